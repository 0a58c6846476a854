// Host build of the fused last-axis pass of the covariance product
// (kf_fused_mult of pf_fast.hpp: forward DFT of the n_last leading points of
// every (kx, k_mid) line, times f(lambda)/N_ext, inverse DFT, first n_last
// points back in place) at the launch geometry of the library's launcher
// (fast_L tiles, 8 values per thread, the Nyquist column's own tiles), FP64
// and FP32, 3D in both storage orders and 2D, the DCT-half and the full
// eigenvalue tables, multiply / root / inverse-floor factors. Every buffer
// ends at a PROT_NONE guard region and every block runs; each line is
// compared with a direct O(N^2) evaluation
//   y(j) = sum_k f(lambda(kx, k_mid, k))/N_ext e^{+2 pi i j k/N}
//          sum_{e < n_last} x(e) e^{-2 pi i e k/N},   j < n_last.
#include <hip/hip_runtime.h>
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>
#include <complex>
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
using cd = std::complex<double>;

struct Buf { const char* name; char* lo; char* hi; };
static std::vector<Buf> bufs;
static const size_t GUARD = (size_t)1 << 30;
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
// the launcher's tile rule (pf_fast_impl.hpp fast_L, fused_n)
template <typename T> constexpr int fuse_L(int N) { int L = sizeof(T) == 4 ? 16 : 8; const int lim = sizeof(T) == 4 ? 8192 : 4096; while (L > 1 && L * N > lim) L /= 2; return L; }
template <typename T> constexpr int fuse_NT(int N) { const int L = fuse_L<T>(N); int nt = (L * N / 8 + 63) / 64 * 64; return nt < 64 ? 64 : nt; }

template <typename T, int N>
static int run_case(const char* name, int dim, int transposed, int dct, int N0, int Nmid, int nlast, int op, int nb)
{
  constexpr int L = fuse_L<T>(N), NT = fuse_NT<T>(N);
  GenGeom g{};
  g.dim = dim; g.transposed = transposed; g.N0 = N0; g.H = N0 / 2 + 1; g.dct = dct;
  if (dim == 3) { g.N1 = Nmid; g.N2 = N; } else { g.N1 = N; g.N2 = 1; }
  g.P0 = g.N0 / 2 + 1; g.P1 = g.N1 / 2 + 1; g.P2 = dim == 3 ? g.N2 / 2 + 1 : 1;
  g.Hp = (g.H + 15) / 16 * 16;
  LastAxis a{}; a.H = g.H; a.Hp = g.Hp; a.N0h = N0 / 2; a.Nmid = dim == 3 ? Nmid : 1; a.nlast = nlast;
  if (dim == 3) { a.sA = transposed ? 1 : g.N1; a.sB = transposed ? g.N2 : 1; } else { a.sA = 1; a.sB = 0; }
  const int64_t neig = dct ? (int64_t)g.P0 * g.P1 * g.P2 : (int64_t)g.H * g.N1 * g.N2;
  double* eig = (double*)galloc("eig", neig * 8);
  std::mt19937 rng(7); std::uniform_real_distribution<double> U(0.5, 2.0), V(-1, 1);
  for (int64_t i = 0; i < neig; i++) eig[i] = U(rng);
  double* spec = (double*)galloc("spec", 8);
  spec[0] = 0.9;
  const int64_t plane = (int64_t)a.Hp * a.Nmid;
  const size_t nc = (size_t)nb * plane * nlast;
  cplx<T>* C = (cplx<T>*)galloc("C", nc * sizeof(cplx<T>));
  std::vector<cplx<T>> C0(nc);
  for (size_t i = 0; i < nc; i++) C0[i] = C[i] = { (T)V(rng), (T)V(rng) };
  cplx<T>* tw = (cplx<T>*)galloc("tw", N * sizeof(cplx<T>));
  for (int j = 0; j < N; j++) tw[j] = { (T)cos(-2 * M_PI * j / N), (T)sin(-2 * M_PI * j / N) };
  const double inv_next = 1.0 / ((double)g.N0 * g.N1 * g.N2);
  blockDim = { (unsigned)NT, 1, 1 };
  gridDim = { (unsigned)(a.N0h / L + 1), (unsigned)a.Nmid, (unsigned)nb };
  cur = name;
  const unsigned tot = gridDim.x * gridDim.y * gridDim.z;
  const bool half = 2 * nlast <= N;
  for (unsigned lin = 0; lin < tot; lin++) {
    if (half)
      run_block(lin, NT, [&] { kf_fused_mult<T, N, L, NT, true>(C, g, a, eig, op, inv_next, tw, spec, 0); });
    else
      run_block(lin, NT, [&] { kf_fused_mult<T, N, L, NT, false>(C, g, a, eig, op, inv_next, tw, spec, 0); });
  }
  double md = 0, mx = 0;
  for (int b = 0; b < nb; b++)
    for (int km = 0; km < a.Nmid; km++)
      for (int kx = 0; kx < g.H; kx++) {
        std::vector<cd> X(N), Y(N);
        for (int k = 0; k < N; k++) {
          cd s = 0;
          for (int e = 0; e < nlast; e++) {
            const cplx<T> v = C0[(size_t)b * plane * nlast + (size_t)e * plane + (size_t)km * a.Hp + kx];
            s += cd(v.x, v.y) * std::polar(1.0, -2 * M_PI * (double)e * k / N);
          }
          int k1, k2;
          line_k(g, kx, km, k, k1, k2);
          const double m = eig[eig_index(g, kx, k1, k2)];
          const double f = op == PF_OP_MULTIPLY ? m : op == PF_OP_ROOT ? std::sqrt(m) : 1.0 / std::max(m, spec[0]);
          X[k] = s * f * inv_next;
        }
        for (int j = 0; j < nlast; j++) {
          cd s = 0;
          for (int k = 0; k < N; k++) s += X[k] * std::polar(1.0, 2 * M_PI * (double)j * k / N);
          const cplx<T> got = C[(size_t)b * plane * nlast + (size_t)j * plane + (size_t)km * a.Hp + kx];
          md = std::max(md, std::abs(cd(got.x, got.y) - s));
          mx = std::max(mx, std::abs(s));
        }
      }
  const double tol = sizeof(T) == 8 ? 1e-12 : 2e-5;
  const bool ok = mx > 0 && md <= tol * mx;
  printf("%-36s %s N=%d L=%d NT=%d blocks=%u nlast=%d op=%d maxdiff/max=%.3g %s\n", name, sizeof(T) == 8 ? "fp64" : "fp32", N, L, NT, tot, nlast, op, md / mx, ok ? "OK" : "FAIL");
  fflush(stdout);
  return !ok;
}

int main() {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  int bad = 0;
  bad += run_case<double, 64>("3D transposed dct multiply", 3, 1, 1, 32, 16, 32, PF_OP_MULTIPLY, 2);
  bad += run_case<double, 64>("3D x-fastest dct root", 3, 0, 1, 32, 16, 32, PF_OP_ROOT, 1);
  bad += run_case<double, 48>("3D transposed r2c inverse-floor", 3, 1, 0, 32, 8, 24, PF_OP_INVERSE_FLOOR, 1);
  bad += run_case<double, 32>("3D x-fastest r2c nlast=N", 3, 0, 0, 16, 8, 32, PF_OP_MULTIPLY, 1);
  bad += run_case<double, 256>("3D transposed dct N=256", 3, 1, 1, 16, 4, 128, PF_OP_MULTIPLY, 1);
  bad += run_case<float, 64>("3D transposed dct fp32", 3, 1, 1, 32, 16, 32, PF_OP_MULTIPLY, 1);
  bad += run_case<float, 128>("2D r2c fp32", 2, 0, 0, 64, 1, 64, PF_OP_ROOT, 2);
  bad += run_case<double, 96>("2D dct", 2, 0, 1, 32, 1, 40, PF_OP_INVERSE_FLOOR, 1);
  printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
