// Host build of kf_gen_last (the real kernel source, see
// tests/test_kernel_emulation.py): the Philox main/tail pass (NOISE=1)
// (with the full and, where nlast <= N/2, the half last stage) against the
// per-entry noise-array path (NOISE=0) fed with the same FP32 Box-Muller
// normals, for the row-parity cases of both storage orders and 2D;
// every buffer is end-aligned against a PROT_NONE guard region, so an access
// past its end stops the run with the buffer and offset named.
#include <hip/hip_runtime.h>
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <thread>
#include <vector>
thread_local emu_dim3 threadIdx;
emu_dim3 blockIdx, blockDim, gridDim;
std::barrier<>* emu_bar;
int emu_perm[1024];
namespace pfamd { alignas(16) unsigned char smem[160 * 1024]; }
#include "pf_fast.hpp"
using namespace pfamd;
// tile shape the launcher picks for FP32 (fast_L / fast_NT in pf_fast_impl.hpp)
constexpr int fast_L32(int N) { int L = 16; while (L > 1 && L * N > 8192) L /= 2; return L; }
constexpr int fast_NT32(int N, int L) { int nt = (L * N + 15) / 16; nt = (nt + 63) / 64 * 64; return nt < 64 ? 64 : nt; }
struct Buf { const char* name; char* lo; char* hi; };
static std::vector<Buf> bufs;
static const size_t GUARD = (size_t)1 << 36;
static void* galloc(const char* name, size_t bytes) {
  const size_t pg = 4096, rounded = (bytes + pg - 1) / pg * pg;
  char* base = (char*)mmap(nullptr, GUARD + rounded + GUARD, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  char* data = base + GUARD; mprotect(data, rounded, PROT_READ | PROT_WRITE);
  char* start = data + (rounded - bytes); bufs.push_back({ name, start, start + bytes }); return start;
}
static const char* cur = "";
static void on_segv(int, siginfo_t* si, void*) {
  char* a = (char*)si->si_addr; char msg[512];
  int n = snprintf(msg, sizeof msg, "SEGV in %s at %p block (%u,%u,%u):", cur, (void*)a, blockIdx.x, blockIdx.y, blockIdx.z);
  for (auto& b : bufs) if (a >= b.lo - GUARD && a < b.hi + GUARD) n += snprintf(msg + n, sizeof msg - n, " %s off %lld size %lld", b.name, (long long)(a - b.lo), (long long)(b.hi - b.lo));
  n += snprintf(msg + n, sizeof msg - n, "\n"); (void)!write(2, msg, n); _exit(3);
}
static void run_block(unsigned lin, int NT, const std::function<void()>& k) {
  const unsigned bx = lin % gridDim.x, t = lin / gridDim.x;
  blockIdx = { bx, t % gridDim.y, t / gridDim.y };
  std::barrier<> bar(NT); emu_bar = &bar;
  std::vector<std::thread> th; for (int i = 0; i < NT; i++) th.emplace_back([&, i] { threadIdx = { (unsigned)i, 0, 0 }; k(); });
  for (auto& x : th) x.join();
}

template <int N>
static int run_case(const char* name, int dim, int transposed, int dct, int N0, int Nmid, int nlast, int nb, int maxblocks)
{
  constexpr int L = fast_L32(N), NT = fast_NT32(N, L);
  GenGeom g{};
  g.dim = dim; g.transposed = transposed; g.N0 = N0; g.H = N0 / 2 + 1; g.dct = dct;
  if (dim == 3) { g.N1 = Nmid; g.N2 = N; g.n0 = N0 / 2; g.n1 = Nmid / 2; g.n2 = nlast; }
  else { g.N1 = N; g.N2 = 1; g.n0 = N0 / 2; g.n1 = nlast; g.n2 = 1; }
  g.P0 = g.N0 / 2 + 1; g.P1 = g.N1 / 2 + 1; g.P2 = dim == 3 ? g.N2 / 2 + 1 : 1;
  g.Hp = (g.H + 15) / 16 * 16;
  const int64_t nent = (int64_t)g.H * g.N1 * (dim == 3 ? g.N2 : 1);
  g.noise_stride = 2 * nent;
  LastAxis a{}; a.H = g.H; a.Hp = g.Hp; a.N0h = N0 / 2; a.Nmid = dim == 3 ? Nmid : 1; a.nlast = nlast;
  if (dim == 3) { a.sA = transposed ? 1 : g.N1; a.sB = transposed ? g.N2 : 1; } else { a.sA = 1; a.sB = 0; }
  const int64_t neig = dct ? (int64_t)g.P0 * g.P1 * g.P2 : nent;
  float* sq = (float*)galloc("sq", neig * 4);
  std::mt19937 rng(5); std::uniform_real_distribution<float> U(0.5f, 2.f);
  for (int64_t i = 0; i < neig; i++) sq[i] = U(rng);
  const uint64_t seed = 0x1234567890abcdefull, first = 3;
  double* noise = (double*)galloc("noise", (size_t)nb * g.noise_stride * 8);
  for (int b = 0; b < nb; b++)
    for (int64_t s = 0; s < nent; s++) {
      float r1, r2; philox_pair<float>(seed, first + b, (uint64_t)s, r1, r2);
      noise[b * g.noise_stride + 2 * s] = r1; noise[b * g.noise_stride + 2 * s + 1] = r2;
    }
  const size_t outn = (size_t)nb * g.Hp * a.Nmid * nlast;
  cplx<float>* B1 = (cplx<float>*)galloc("B1", outn * 8);
  cplx<float>* B2 = (cplx<float>*)galloc("B2", outn * 8);
  cplx<float>* B3 = (cplx<float>*)galloc("B3", outn * 8);
  for (size_t i = 0; i < outn; i++) { B1[i] = { NAN, NAN }; B2[i] = { NAN, NAN }; B3[i] = { NAN, NAN }; }
  const bool half = 2 * nlast <= N;
  cplx<float>* tw = (cplx<float>*)galloc("tw", N * 8);
  for (int j = 0; j < N; j++) tw[j] = { (float)cos(-2 * M_PI * j / N), (float)sin(-2 * M_PI * j / N) };
  blockDim = { (unsigned)NT, 1, 1 };
  gridDim = { (unsigned)(a.N0h / L + 1), (unsigned)a.Nmid, (unsigned)nb };
  const unsigned tot = gridDim.x * gridDim.y * gridDim.z;
  std::vector<unsigned> blocks;
  if ((int)tot <= maxblocks) { for (unsigned i = 0; i < tot; i++) blocks.push_back(i); }
  else {
    for (unsigned i = 0; i < 40; i++) blocks.push_back(i);
    for (unsigned i = tot - 40; i < tot; i++) blocks.push_back(i);
    std::mt19937 r2(9); while ((int)blocks.size() < maxblocks) blocks.push_back(r2() % tot);
  }
  cur = "philox";
  for (unsigned lin : blocks) run_block(lin, NT, [&] { kf_gen_last<float, N, L, NT, 1, false>(g, a, sq, nullptr, seed, first, B1, tw); });
  cur = "philox, lower half of the last stage";
  if (half)
    for (unsigned lin : blocks) run_block(lin, NT, [&] { kf_gen_last<float, N, L, NT, 1, false, true>(g, a, sq, nullptr, seed, first, B3, tw); });
  cur = "noise";
  for (unsigned lin : blocks) run_block(lin, NT, [&] { kf_gen_last<float, N, L, NT, 0, true>(g, a, sq, noise, seed, first, B2, tw); });
  int bad = 0;
  for (int v = 0; v < (half ? 2 : 1); v++) {
    const cplx<float>* X = v ? B3 : B1;
    double mx = 0, md = 0; size_t nan_mis = 0, cnt = 0;
    for (size_t i = 0; i < outn; i++) {
      const bool n1 = std::isnan(X[i].x), n2 = std::isnan(B2[i].x);
      if (n1 != n2) { nan_mis++; continue; }
      if (n1) continue;
      cnt++;
      mx = std::max(mx, (double)std::max(std::fabs(B2[i].x), std::fabs(B2[i].y)));
      md = std::max(md, (double)std::max(std::fabs(X[i].x - B2[i].x), std::fabs(X[i].y - B2[i].y)));
    }
    const bool ok = nan_mis == 0 && cnt > 0 && md <= 2e-5 * mx;
    printf("%-34s %-5s N=%d L=%d NT=%d blocks=%zu/%u written=%zu nan_mismatch=%zu maxdiff/max=%.3g %s\n", name,
           v ? "half" : "full", N, L, NT, blocks.size(), tot, cnt, nan_mis, mx > 0 ? md / mx : -1.0, ok ? "OK" : "FAIL");
    fflush(stdout);
    bad += !ok;
  }
  return bad;
}

int main(int argc, char** argv) {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  int bad = 0;
  bad += run_case<64>("3D transposed r2c", 3, 1, 0, 64, 64, 32, 2, 1 << 30);
  bad += run_case<64>("3D transposed dct", 3, 1, 1, 64, 64, 32, 2, 1 << 30);
  bad += run_case<64>("3D x-fastest r2c", 3, 0, 0, 64, 64, 32, 2, 1 << 30);
  bad += run_case<64>("3D x-fastest dct", 3, 0, 1, 64, 64, 32, 2, 1 << 30);
  bad += run_case<48>("3D transposed dct N=48", 3, 1, 1, 96, 32, 24, 1, 1 << 30);
  bad += run_case<96>("3D x-fastest dct N=96", 3, 0, 1, 32, 48, 48, 1, 1 << 30);
  bad += run_case<64>("2D r2c", 2, 0, 0, 64, 1, 32, 3, 1 << 30);
  bad += run_case<128>("2D dct", 2, 0, 1, 96, 1, 64, 2, 1 << 30);
  bad += run_case<768>("2D dct N=768", 2, 0, 1, 32, 1, 384, 1, 1 << 30);
  bad += run_case<16>("3D transposed r2c N=16", 3, 1, 0, 32, 16, 8, 2, 1 << 30);
  bad += run_case<32>("3D x-fastest r2c nlast=N", 3, 0, 0, 32, 32, 32, 2, 1 << 30);
  bad += run_case<64>("2D r2c nlast=N", 2, 0, 0, 64, 1, 64, 2, 1 << 30);
  // the 512-point tile of the 256^3 / 512^3 ensemble (16 lines, 512 threads)
  // on small x and middle extents
  bad += run_case<512>("3D transposed dct N=512", 3, 1, 1, 32, 16, 256, 2, 1 << 30);
  bad += run_case<512>("3D x-fastest r2c N=512", 3, 0, 0, 32, 16, 256, 1, 1 << 30);
  bad += run_case<512>("2D r2c N=512", 2, 0, 0, 64, 1, 256, 2, 1 << 30);
  printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
