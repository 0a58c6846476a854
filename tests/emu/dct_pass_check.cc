// Host build of the specialised DCT/DST passes (kdf_gen_last, kdf_mid, kdf_x
// of pf_dctdst_fast.hpp; see tests/test_kernel_emulation.py) at the launch
// shapes the library picks: every buffer is end-aligned against a PROT_NONE
// guard region, so an access past its end stops the run with the buffer and
// offset named. Small shapes run every block (one with M = 1024 in z and y
// runs a sample); the 1024^3 shape (2048^3
// embedding, three realizations per launch, as the work-size cap gives at
// that size) runs the first, last, edge and random blocks of each pass.
// Then the three passes fed host noise (NOISE = 0) are checked against the
// same passes fed the FP32 Philox normals in-kernel (NOISE = 1) on a small
// shape, every output compared.
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
#include "pf_dctdst_fast.hpp"
using namespace pfamd;
// tile shape of the FP32 launchers (dct_L / fast_NT in pf_fast_impl.hpp)
constexpr int dct_L32(int M) { int L = 32; while (L > 1 && (long)L * M * 8 > 65536) L /= 2; return L; }
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
// all blocks, or the first and last 8 of the grid, the first and last block
// of every edge-column group and random others
// sampled block counts: the full sample with PF_EMU_FULL=1, else the
// structured picks (first, last and edge blocks) and a few random ones
static unsigned samples(unsigned full, unsigned quick) { return getenv("PF_EMU_FULL") ? full : quick; }
static std::vector<unsigned> pick(unsigned maxblocks) {
  const unsigned tot = gridDim.x * gridDim.y * gridDim.z;
  std::vector<unsigned> v;
  if (tot <= maxblocks) { for (unsigned i = 0; i < tot; i++) v.push_back(i); return v; }
  for (unsigned i = 0; i < 8; i++) { v.push_back(i); v.push_back(tot - 1 - i); }
  const unsigned ex = gridDim.x - 1; // edge column (kx = P0 - 1) of the strided passes
  for (unsigned z = 0; z < gridDim.z; z++) {
    v.push_back(ex + gridDim.x * gridDim.y * z);
    v.push_back(ex + gridDim.x * (gridDim.y - 1 + gridDim.y * z));
  }
  std::mt19937 r(11); while (v.size() < maxblocks) v.push_back(r() % tot);
  return v;
}
static cplx<float>* table(int N) {
  cplx<float>* tw = (cplx<float>*)galloc("tw", (size_t)N * 8);
  for (int j = 0; j < N; j++) tw[j] = { (float)cos(-2 * M_PI * j / N), (float)sin(-2 * M_PI * j / N) };
  return tw;
}
static DctGeom geom(int n0, int n1, int n2, int transposed) {
  DctGeom d{};
  d.dim = 3; d.n0 = n0; d.n1 = n1; d.n2 = n2;
  d.N0 = 2 * n0; d.N1 = 2 * n1; d.N2 = 2 * n2;
  d.P0 = n0 + 1; d.P1 = n1 + 1; d.P2 = n2 + 1;
  d.Ptot = (int64_t)d.P0 * d.P1 * d.P2;
  d.sx = 1;
  d.s1 = transposed ? (int64_t)d.P0 * d.P2 : d.P0;
  d.s2 = transposed ? d.P0 : (int64_t)d.P0 * d.P1;
  d.noise_stride = 8 * d.Ptot;
  d.P0p = (d.P0 + 31) / 32 * 32;
  return d;
}
struct Bufs { float *sq, *G, *H, *out; cplx<float> *twz, *twy, *twx; };
template <int MZ, int MY, int MX>
static void passes(const DctGeom& g, const Bufs& b, const double* noise, int nb, unsigned maxblocks, const char* tag) {
  constexpr int LZ = dct_L32(MZ), NTZ = fast_NT32(MZ, LZ);
  constexpr int LY = dct_L32(MY), NTY = fast_NT32(MY, LY);
  constexpr int LX = dct_L32(MX), NTX = fast_NT32(MX, LX);
  static_assert((size_t)LZ * MZ * 8 <= 65536 && (size_t)LY * MY * 8 <= 65536 && (size_t)LX * MX * 8 <= 65536, "tile within 64 KB");
  if ((g.P0 - 1) % LZ || (g.P0 - 1) % LY) { printf("%s: shape has no specialised passes\n", tag); exit(4); }
  const uint64_t seed = 0x9e3779b97f4a7c15ull, first = 2;
  gridDim = { (unsigned)((g.P0 - 1) / LZ + 1), (unsigned)g.P1, (unsigned)nb }; blockDim = { (unsigned)NTZ, 1, 1 };
  cur = "kdf_gen_last";
  for (unsigned lin : pick(maxblocks))
    run_block(lin, NTZ, [&] {
      if (noise) kdf_gen_last<float, MZ, LZ, NTZ, 0>(g, b.sq, noise, seed, first, b.G, b.twz);
      else kdf_gen_last<float, MZ, LZ, NTZ, 1>(g, b.sq, nullptr, seed, first, b.G, b.twz);
    });
  gridDim = { (unsigned)((g.P0 - 1) / LY + 1), (unsigned)g.n2, (unsigned)nb }; blockDim = { (unsigned)NTY, 1, 1 };
  cur = "kdf_mid";
  for (unsigned lin : pick(maxblocks)) run_block(lin, NTY, [&] { kdf_mid<float, MY, LY, NTY>(g, b.G, b.H, b.twy); });
  const int64_t rows = (int64_t)g.n1 * g.n2;
  gridDim = { (unsigned)((rows + LX - 1) / LX), (unsigned)nb, 1 }; blockDim = { (unsigned)NTX, 1, 1 };
  cur = "kdf_x";
  for (unsigned lin : pick(maxblocks)) run_block(lin, NTX, [&] { kdf_x<float, MX, LX, NTX>(g, b.H, b.out, rows, b.twx, PF_TRANSFORM_SIGN, 0.5f); });
  printf("%-44s n=(%d,%d,%d) nb=%d tiles L=(%d,%d,%d) OK\n", tag, g.n0, g.n1, g.n2, nb, LZ, LY, LX);
  fflush(stdout);
}
static Bufs alloc(const DctGeom& g, int nb, bool zero_sq) {
  Bufs b;
  b.sq = (float*)galloc("sq", (size_t)g.Ptot * 4);
  if (!zero_sq) for (int64_t i = 0; i < g.Ptot; i++) b.sq[i] = 0.5f + 0.001f * (float)(i % 997);
  b.G = (float*)galloc("G", (size_t)nb * 4 * g.P1 * g.P0p * g.n2 * 4);
  b.H = (float*)galloc("Hh", (size_t)nb * 2 * g.n1 * g.n2 * g.P0p * 4);
  b.out = (float*)galloc("out", (size_t)nb * g.n0 * g.n1 * g.n2 * 4);
  b.twz = table(2 * g.n2); b.twy = table(2 * g.n1); b.twx = table(2 * g.n0);
  return b;
}
int main() {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  int bad = 0;
  {  // every block of small shapes, both storage orders
    for (int tr = 0; tr < 2; tr++) {
      DctGeom g = geom(32, 64, 16, tr); Bufs b = alloc(g, 2, false);
      passes<16, 64, 32>(g, b, nullptr, 2, 1u << 30, tr ? "all blocks, transposed" : "all blocks, x-fastest");
    }
    DctGeom g = geom(16, 1024, 1024, 1); Bufs b = alloc(g, 1, false);
    passes<1024, 1024, 16>(g, b, nullptr, 1, samples(160, 24), "M=1024 in z and y, sampled blocks");
  }
  {  // host noise against in-kernel Philox normals (same FP32 Box-Muller values)
    DctGeom g = geom(32, 32, 64, 1);
    const int nb = 2;
    const uint64_t seed = 0x9e3779b97f4a7c15ull, first = 2;
    double* noise = (double*)galloc("noise", (size_t)nb * g.noise_stride * 8);
    for (int r = 0; r < nb; r++)
      for (int64_t s = 0; s < g.noise_stride; s += 2) {
        float a, c; philox_pair<float>(seed, first + r, (uint64_t)(s / 2), a, c);
        noise[r * g.noise_stride + s] = a; noise[r * g.noise_stride + s + 1] = c;
      }
    Bufs b1 = alloc(g, nb, false), b2 = alloc(g, nb, false);
    const size_t n = (size_t)nb * g.n0 * g.n1 * g.n2;
    for (size_t i = 0; i < n; i++) { b1.out[i] = NAN; b2.out[i] = NAN; }
    passes<64, 32, 32>(g, b1, nullptr, nb, 1u << 30, "Philox in-kernel");
    passes<64, 32, 32>(g, b2, noise, nb, 1u << 30, "host noise");
    // sign outputs: compare the Gaussian values instead through the H arrays
    size_t mism = 0, nans = 0;
    for (size_t i = 0; i < n; i++) { nans += std::isnan(b1.out[i]) || std::isnan(b2.out[i]); mism += b1.out[i] != b2.out[i]; }
    const size_t hn = (size_t)nb * 2 * g.n1 * g.n2 * g.P0p;
    double mx = 0, md = 0;
    for (size_t i = 0; i < hn; i++) { mx = std::max(mx, (double)std::fabs(b2.H[i])); md = std::max(md, (double)std::fabs(b1.H[i] - b2.H[i])); }
    const bool ok = nans == 0 && md <= 1e-5 * mx && mism <= n / 1000;
    printf("Philox vs host noise: outputs %zu nan %zu sign mismatches %zu H maxdiff/max %.3g %s\n", n, nans, mism, mx > 0 ? md / mx : -1.0, ok ? "OK" : "FAIL");
    bad += !ok;
  }
  {  // the 1024^3 shape of the DCT/DST generation: sampled blocks
    for (int tr = 1; tr >= 0; tr--) {
      DctGeom g = geom(1024, 1024, 1024, tr); Bufs b = alloc(g, 3, true);
      passes<1024, 1024, 1024>(g, b, nullptr, 3, samples(48, 22), tr ? "1024^3 transposed, sampled blocks" : "1024^3 x-fastest, sampled blocks");
    }
  }
  printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
