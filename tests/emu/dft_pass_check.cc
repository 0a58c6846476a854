// Host build of the DFT field passes of pf_dft.hpp (kdft_gen, kdft_cols,
// kdft_rows, kdft_pad, kdft_mult_1d) at the launch geometry the library
// gives them (tile and thread rule of make_cfg, the four-step split of
// dft_split_1d with a small tile cap so that the 1-D cases take two passes).
// Every buffer ends at a PROT_NONE guard region and every block runs; the
// fields are compared with a direct O(N^2) evaluation of
//   f(x) + i f_spare(x) = sum_k sqrt(lambda_k) (xi_2s + i xi_2s+1) e^{+2 pi i k.x/N}
// (s the storage index of k), and the 1-D products with the circulant sum
//   y(x) = sum_k f(lambda_k)/N e^{+2 pi i k x/N} sum_x' x(x') e^{-2 pi i k x'/N}.
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
#include "pf_dft.hpp"
using namespace pfamd;
using cd = std::complex<double>;
struct dim3 { unsigned x, y, z; dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };

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
static void launch(const char* name, dim3 grid, int nt, const std::function<void()>& k) {
  cur = name;
  gridDim = { grid.x, grid.y, grid.z };
  blockDim = { (unsigned)nt, 1, 1 };
  const unsigned tot = grid.x * grid.y * grid.z;
  for (unsigned lin = 0; lin < tot; lin++) {
    const unsigned bx = lin % grid.x, t = lin / grid.x;
    blockIdx = { bx, t % grid.y, t / grid.y };
    std::barrier<> bar(nt); emu_bar = &bar;
    std::vector<std::thread> th;
    for (int i = 0; i < nt; i++) th.emplace_back([&, i] { threadIdx = { (unsigned)i, 0, 0 }; k(); });
    for (auto& x : th) x.join();
  }
}
static Sched sched(int n) {
  Sched s{}; s.n = n; s.nst = ct_nst(n);
  for (int i = 0; i < s.nst; i++) s.rad[i] = ct_radix(n, i);
  return s;
}
// make_cfg of the library: L lines per tile, one thread per MAXE values
static size_t BUDGET = 65536;
template <typename T> struct Cfg { int L, nt; size_t lds; };
template <typename T> static Cfg<T> cfg(int n, int64_t nlines) {
  const size_t cs = sizeof(cplx<T>);
  int L = sizeof(T) == 4 ? 16 : 8;
  while (L > 1 && ((size_t)L * n + n) * cs > BUDGET) L /= 2;
  while (L > 1 && L / 2 >= nlines) L /= 2;
  const int me = sizeof(T) == 4 ? 32 : 16;
  int nt = (L * n + me - 1) / me; nt = (nt + 63) / 64 * 64; if (nt < 64) nt = 64;
  return { L, nt, ((size_t)L * n + n) * cs };
}
static unsigned nblk(int64_t lines, int L) { return (unsigned)((lines + L - 1) / L); }
template <typename T> static cplx<T>* table(int N) {
  cplx<T>* tw = (cplx<T>*)galloc("tw", (size_t)N * sizeof(cplx<T>));
  for (int j = 0; j < N; j++) tw[j] = { (T)cos(-2 * M_PI * j / N), (T)sin(-2 * M_PI * j / N) };
  return tw;
}
template <typename T> static void cols(const char* name, bool inv, const cplx<T>* in, cplx<T>* out, const ColPass& p, int nb) {
  auto c = cfg<T>(p.E, (int64_t)p.C * p.O);
  const Sched s = sched(p.E);
  const cplx<T>* tw = table<T>(p.E);
  constexpr int ME = sizeof(T) == 4 ? 32 : 16;
  launch(name, dim3(nblk((int64_t)p.C * p.O, c.L), nb), c.nt, [&] {
    if (inv) kdft_cols<T, ME, true>(in, out, p, s, tw, c.L); else kdft_cols<T, ME, false>(in, out, p, s, tw, c.L);
  });
}
template <typename T, bool INV, int OUT> static void rows(const char* name, const cplx<T>* in, cplx<T>* outc, T* outf, const RowPass& r, int E, int nb, int64_t ff, int64_t nf) {
  auto c = cfg<T>(E, r.R);
  const Sched s = sched(E);
  const cplx<T>* tw = table<T>(E);
  constexpr int ME = sizeof(T) == 4 ? 32 : 16;
  launch(name, dim3(nblk(r.R, c.L), nb), c.nt, [&] { kdft_rows<T, T, ME, INV, OUT>(in, outc, outf, r, s, tw, c.L, ff, nf, 0, T(0)); });
}
// the library's four-step split with tile cap `cap`
static void split(int N, int cap, int& A, int& B) {
  for (int a = std::min(N, cap); a >= 1; a--) if (N % a == 0) { if (N / a > cap) break; A = a; B = N / a; return; }
  fprintf(stderr, "no split\n"); exit(2);
}

static int bad = 0;
static void report(const char* tag, double err, double tol) {
  printf("%-44s max rel err %.2e %s\n", tag, err, err <= tol ? "OK" : "FAIL");
  fflush(stdout);
  if (!(err <= tol)) bad++;
}

// storage index of k (x fastest; last two axes swapped when transposed)
static int64_t sidx(int dim, int tr, const int N[3], int k0, int k1, int k2) {
  if (dim == 1) return k0;
  if (dim == 2) return tr ? k1 + (int64_t)N[1] * k0 : k0 + (int64_t)N[0] * k1;
  return tr ? k0 + (int64_t)N[0] * (k2 + (int64_t)N[2] * k1) : k0 + (int64_t)N[0] * (k1 + (int64_t)N[1] * k2);
}

// generation: draws [0, nd) with host noise; fields checked against the
// direct sum for every draw and both parts. f_first = -1 drops the real part
// of draw 0 (an odd first field), nfields cuts the last imaginary part.
template <typename T>
static void gen_case(const char* tag, int dim, const int n[3], int tr, int cap, int nd, int64_t f_first, double tol) {
  int N[3] = { 2 * n[0], dim > 1 ? 2 * n[1] : 1, dim > 2 ? 2 * n[2] : 1 };
  const int64_t Next = (int64_t)N[0] * N[1] * N[2], dom = (int64_t)n[0] * (dim > 1 ? n[1] : 1) * (dim > 2 ? n[2] : 1);
  DftGeom g{};
  g.dim = dim; g.transposed = tr; g.N0 = N[0]; g.N1 = N[1]; g.N2 = N[2]; g.H = N[0] / 2 + 1;
  g.P0 = N[0] / 2 + 1; g.P1 = N[1] / 2 + 1; g.P2 = N[2] / 2 + 1; g.dct = 0; g.noise_stride = 2 * Next;
  if (dim == 1) split(N[0], cap, g.A, g.B); else { g.A = N[0]; g.B = 1; }
  // symmetric spectrum lambda(k) = lambda(-k); half table of kx < H
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.1, 1.0);
  std::vector<double> lam(Next);
  for (int64_t i = 0; i < Next; i++) {
    const int k0 = i % N[0], k1 = (i / N[0]) % N[1], k2 = i / ((int64_t)N[0] * N[1]);
    const int m0 = (N[0] - k0) % N[0], m1 = (N[1] - k1) % N[1], m2 = (N[2] - k2) % N[2];
    const int64_t j = m0 + (int64_t)N[0] * (m1 + (int64_t)N[1] * m2);
    lam[i] = j < i ? lam[j] : U(rng);
  }
  const int64_t half = (int64_t)g.H * N[1] * N[2];
  T* sq = (T*)galloc("sq", half * sizeof(T));
  for (int64_t i = 0; i < half; i++) {
    const int k0 = i % g.H, r = i / g.H;
    sq[i] = (T)std::sqrt(lam[k0 + (int64_t)N[0] * r]);
  }
  std::normal_distribution<double> nd01;
  double* noise = (double*)galloc("noise", (size_t)nd * 2 * Next * sizeof(double));
  for (int64_t i = 0; i < (int64_t)nd * 2 * Next; i++) noise[i] = nd01(rng);
  const int64_t nfields = 2 * nd + f_first; // fields of the call: f_first + 2j (+1), j < nd
  T* out = (T*)galloc("out", (size_t)std::max<int64_t>(nfields, 1) * dom * sizeof(T));
  for (int64_t i = 0; i < nfields * dom; i++) out[i] = (T)NAN;
  // pass 1
  int64_t w1 = dim == 3 ? (int64_t)n[2] * N[1] * N[0] : dim == 2 ? (int64_t)n[1] * N[0] : N[0];
  int64_t w2 = dim == 3 ? (int64_t)n[2] * n[1] * N[0] : 0;
  cplx<T>* W1 = (cplx<T>*)galloc("W1", (size_t)w1 * nd * sizeof(cplx<T>));
  cplx<T>* W2 = w2 ? (cplx<T>*)galloc("W2", (size_t)w2 * nd * sizeof(cplx<T>)) : nullptr;
  ColPass p{};
  if (dim == 3) { p.C = N[0]; p.E = N[2]; p.O = N[1]; p.keep = n[2]; p.out_es = (int64_t)N[1] * N[0]; p.out_os = N[0]; }
  else if (dim == 2) { p.C = N[0]; p.E = N[1]; p.O = 1; p.keep = n[1]; p.out_es = N[0]; }
  else { p.C = g.A; p.E = g.B; p.O = 1; p.keep = g.B; p.in_es = g.A; p.in_bs = N[0]; p.out_es = g.A; p.twN = g.B > 1 ? N[0] : 0; }
  p.out_bs = w1;
  {
    auto c = cfg<T>(p.E, (int64_t)p.C * p.O);
    const Sched s = sched(p.E);
    const cplx<T>* tw = table<T>(p.E);
    constexpr int ME = sizeof(T) == 4 ? 32 : 16;
    launch("kdft_gen", dim3(nblk((int64_t)p.C * p.O, c.L), nd), c.nt, [&] { kdft_gen<T, ME, 0>(g, sq, noise, 0, 0, W1, p, s, tw, c.L); });
  }
  const cplx<T>* rin = W1;
  RowPass r{};
  int E = N[0];
  if (dim == 3) {
    ColPass q{};
    q.C = N[0]; q.E = N[1]; q.O = n[2]; q.keep = n[1]; q.in_es = N[0]; q.in_os = (int64_t)N[1] * N[0]; q.in_bs = w1;
    q.out_es = N[0]; q.out_os = (int64_t)n[1] * N[0]; q.out_bs = w2;
    cols<T>("kdft_cols (y)", true, W1, W2, q, nd);
    rin = W2; r.R = (int64_t)n[1] * n[2]; r.in_bs = w2;
  } else if (dim == 2) { r.R = n[1]; r.in_bs = w1; }
  if (dim >= 2) { r.keep = n[0]; r.in_rs = N[0]; r.out_rs = n[0]; r.out_es = 1; r.limit = dom; }
  else { r.R = g.B; r.keep = g.A; r.in_rs = g.A; r.in_bs = N[0]; r.out_rs = 1; r.out_es = g.B; r.limit = n[0]; E = g.A; }
  r.out_bs = dom;
  rows<T, true, 1>("kdft_rows (x)", rin, nullptr, out, r, E, nd, f_first, nfields);
  // direct sums
  double err = 0, mx = 0;
  for (int d0 = 0; d0 < nd; d0++) {
    const double* z = noise + (int64_t)d0 * 2 * Next;
    for (int64_t o = 0; o < dom; o++) {
      const int x0 = o % n[0], x1 = dim > 1 ? (o / n[0]) % n[1] : 0, x2 = dim > 2 ? o / ((int64_t)n[0] * n[1]) : 0;
      cd acc = 0;
      for (int64_t i = 0; i < Next; i++) {
        const int k0 = i % N[0], k1 = (i / N[0]) % N[1], k2 = i / ((int64_t)N[0] * N[1]);
        const int64_t s = sidx(dim, tr, N, k0, k1, k2);
        const double ph = 2 * M_PI * ((double)k0 * x0 / N[0] + (double)k1 * x1 / N[1] + (double)k2 * x2 / N[2]);
        acc += std::sqrt(lam[i]) * cd(z[2 * s], z[2 * s + 1]) * cd(cos(ph), sin(ph));
      }
      const int64_t fre = f_first + 2 * d0;
      for (int part = 0; part < 2; part++) {
        const int64_t f = fre + part;
        if (f < 0 || f >= nfields) continue;
        const double want = part ? acc.imag() : acc.real();
        const double got = (double)out[f * dom + o];
        mx = std::max(mx, std::abs(want));
        err = std::max(err, std::abs(got - want));
      }
    }
  }
  for (int64_t i = 0; i < nfields * dom; i++) if (std::isnan((double)out[i])) { err = INFINITY; break; }
  char t[128]; snprintf(t, sizeof t, "%s (A=%d B=%d)", tag, g.A, g.B);
  report(t, err / mx, tol);
}

// 1-D product: pad, forward four-step, times lambda / N, inverse four-step
template <typename T>
static void matvec_case(const char* tag, int n0, int cap, int nb, double tol) {
  const int N = 2 * n0, H = N / 2 + 1;
  DftGeom g{}; g.dim = 1; g.N0 = N; g.N1 = g.N2 = 1; g.H = H;
  split(N, cap, g.A, g.B);
  std::mt19937_64 rng(9);
  std::uniform_real_distribution<double> U(0.1, 1.0);
  double* eig = (double*)galloc("eig", H * sizeof(double));
  for (int k = 0; k < H; k++) eig[k] = U(rng);
  T* x = (T*)galloc("x", (size_t)nb * n0 * sizeof(T));
  T* y = (T*)galloc("y", (size_t)nb * n0 * sizeof(T));
  std::normal_distribution<double> nd01;
  for (int i = 0; i < nb * n0; i++) { x[i] = (T)nd01(rng); y[i] = (T)NAN; }
  cplx<T>* W1 = (cplx<T>*)galloc("W1", (size_t)nb * N * sizeof(cplx<T>));
  cplx<T>* W2 = (cplx<T>*)galloc("W2", (size_t)nb * N * sizeof(cplx<T>));
  launch("kdft_pad", dim3(4, 1), 64, [&] { kdft_pad<T, T>(x, W2, n0, N, nb); });
  ColPass p{}; p.C = g.B; p.E = g.A; p.O = 1; p.keep = g.A; p.in_es = g.B; p.in_bs = N; p.out_es = g.B; p.out_bs = N; p.twN = g.B > 1 ? N : 0;
  cols<T>("kdft_cols (forward)", false, W2, W1, p, nb);
  RowPass r{}; r.R = g.A; r.keep = g.B; r.in_rs = g.B; r.in_bs = N; r.out_rs = 1; r.out_es = g.A; r.out_bs = N; r.limit = N;
  rows<T, false, 0>("kdft_rows (forward)", W1, W2, nullptr, r, g.B, nb, 0, 0);
  launch("kdft_mult_1d", dim3(4, 1), 64, [&] { kdft_mult_1d<T>(W2, N, H, eig, 0, 1.0 / N, nb, 0.); });
  ColPass q{}; q.C = g.A; q.E = g.B; q.O = 1; q.keep = g.B; q.in_es = g.A; q.in_bs = N; q.out_es = g.A; q.out_bs = N; q.twN = g.B > 1 ? N : 0;
  cols<T>("kdft_cols (inverse)", true, W2, W1, q, nb);
  RowPass s{}; s.R = g.B; s.keep = g.A; s.in_rs = g.A; s.in_bs = N; s.out_rs = 1; s.out_es = g.B; s.out_bs = n0; s.limit = n0;
  rows<T, true, 2>("kdft_rows (inverse)", W1, nullptr, y, s, g.A, nb, 0, nb);
  double err = 0, mx = 0;
  for (int b = 0; b < nb; b++)
    for (int o = 0; o < n0; o++) {
      cd acc = 0;
      for (int k = 0; k < N; k++) {
        cd X = 0;
        for (int xp = 0; xp < n0; xp++) X += (double)x[b * n0 + xp] * std::polar(1.0, -2 * M_PI * (double)k * xp / N);
        acc += eig[k < H ? k : N - k] / N * X * std::polar(1.0, 2 * M_PI * (double)k * o / N);
      }
      mx = std::max(mx, std::abs(acc.real()));
      const double got = (double)y[b * n0 + o];
      err = std::isnan(got) ? INFINITY : std::max(err, std::abs(got - acc.real()));
    }
  char t[128]; snprintf(t, sizeof t, "%s (A=%d B=%d)", tag, g.A, g.B);
  report(t, err / mx, tol);
}

int main() {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  const int n1[3] = { 12, 1, 1 }, n1b[3] = { 30, 1, 1 };
  gen_case<double>("1D one tile, 2 draws", 1, n1, 0, 4096, 2, 0, 1e-12);
  gen_case<double>("1D four-step, odd first field", 1, n1b, 0, 12, 2, -1, 1e-12);
  gen_case<float>("1D four-step FP32", 1, n1b, 0, 10, 1, 0, 2e-5);
  const int n2[3] = { 5, 3, 1 };
  gen_case<double>("2D x-fastest", 2, n2, 0, 0, 2, 0, 1e-12);
  gen_case<double>("2D transposed, odd first field", 2, n2, 1, 0, 1, -1, 1e-12);
  const int n3[3] = { 3, 2, 4 };
  gen_case<double>("3D x-fastest", 3, n3, 0, 0, 1, 0, 1e-12);
  gen_case<double>("3D transposed", 3, n3, 1, 0, 2, 0, 1e-12);
  gen_case<float>("3D transposed FP32", 3, n3, 1, 0, 1, 0, 2e-5);
  matvec_case<double>("1D product, one tile", 20, 4096, 2, 1e-12);
  matvec_case<double>("1D product, four-step", 36, 12, 2, 1e-12);
  matvec_case<float>("1D product, four-step FP32", 36, 9, 1, 2e-5);
  printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad ? 1 : 0;
}
