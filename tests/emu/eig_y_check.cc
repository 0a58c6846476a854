// Host build of kf_eig_c2r_y (the real kernel source, see
// tests/test_kernel_emulation.py): the real y pass of the 2D R2C
// eigenvalues from the x half spectrum of a point-symmetric array (rows
// 0 .. M of N1 = 2M), against direct O(N1^2) DFTs of the Hermitian-extended
// columns: eigenvalues, square roots, and the classification summed over the
// block partials; two realizations with their own thresholds, every block.
// Every buffer ends at a PROT_NONE guard region, so an access past its end
// stops the run with the buffer and offset named; the pad columns of the
// spectrum hold NaN, so a read of them shows in the results.
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
unsigned long long emu_shfl[1024];
namespace pfamd { alignas(16) unsigned char smem[160 * 1024]; }
#include "pf_fast.hpp"
using namespace pfamd;
// tile shape the launcher picks for FP64 (fast_L / fast_NT in pf_fast_impl.hpp)
constexpr int fast_L64(int N) { int L = 8; while (L > 1 && L * N > 4096) L /= 2; return L; }
constexpr int fast_NT64(int N, int L) { int nt = (L * N + 7) / 8; nt = (nt + 63) / 64 * 64; return nt < 64 ? 64 : nt; }
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

template <typename TS, int M>
static int run_case(int N0, bool keep_eig)
{
  constexpr int L = fast_L64(M), NT = fast_NT64(M, L), N1 = 2 * M;
  const int H = N0 / 2 + 1, Hp = (H + 15) / 16 * 16, N0h = N0 / 2, nf = 2;
  if (N0h % L) { printf("bad case\n"); return 1; }
  const double inv_next = 1.0 / ((double)N0 * N1);
  // point-symmetric real arrays e[j][i] (j <= M), their x half spectra
  std::mt19937_64 rng(N0 * 131 + M);
  std::normal_distribution<double> nd;
  cplx<double>* spec = (cplx<double>*)galloc("spec", (size_t)nf * (M + 1) * Hp * 16);
  std::vector<double> efull((size_t)nf * N1 * N0);
  for (int f = 0; f < nf; f++) {
    double* e = &efull[(size_t)f * N1 * N0];
    for (int j = 0; j <= M; j++)
      for (int i = 0; i < N0; i++) e[(size_t)j * N0 + i] = nd(rng);
    // point symmetry: e[-j][-i] = e[j][i] (rows 0 and M symmetric in i)
    for (int j = 0; j <= M; j++)
      for (int i = 0; i < N0; i++) {
        const int jm = (N1 - j) % N1, im = (N0 - i) % N0;
        if (jm > M || (jm == j && im > i)) continue;
        e[(size_t)jm * N0 + im] = e[(size_t)j * N0 + i];
      }
    for (int j = M + 1; j < N1; j++)
      for (int i = 0; i < N0; i++) e[(size_t)j * N0 + i] = e[(size_t)(N1 - j) * N0 + (N0 - i) % N0];
    for (int j = 0; j <= M; j++)
      for (int k = 0; k < Hp; k++) {
        cplx<double> s = { NAN, NAN };
        if (k < H) {
          s = { 0, 0 };
          for (int i = 0; i < N0; i++) {
            const double a = -2 * M_PI * (double)((int64_t)k * i % N0) / N0;
            s.x += e[(size_t)j * N0 + i] * cos(a);
            s.y += e[(size_t)j * N0 + i] * sin(a);
          }
        }
        spec[((size_t)f * (M + 1) + j) * Hp + k] = s;
      }
  }
  cplx<double>* tw = (cplx<double>*)galloc("tw", N1 * 16);
  for (int j = 0; j < N1; j++) tw[j] = { cos(-2 * M_PI * j / N1), sin(-2 * M_PI * j / N1) };
  const size_t cnt = (size_t)H * N1;
  double* eig = keep_eig ? (double*)galloc("eig", nf * cnt * 8) : nullptr;
  TS* sq = (TS*)galloc("sq", nf * cnt * sizeof(TS));
  const int nb = N0h / L + 1;
  ClassifyPartial* part = (ClassifyPartial*)galloc("partial", (size_t)nf * nb * sizeof(ClassifyPartial));
  double* thrs = (double*)galloc("thrs", nf * 8);
  thrs[0] = 1e-14; thrs[1] = 0.3; // the second one classifies a band of values as zero
  for (size_t i = 0; i < nf * cnt; i++) { sq[i] = NAN; if (eig) eig[i] = NAN; }
  for (int i = 0; i < nf * nb; i++) part[i] = { -1, -1, -1, NAN };
  blockDim = { (unsigned)NT, 1, 1 };
  gridDim = { (unsigned)nb, (unsigned)nf, 1 };
  cur = "kf_eig_c2r_y";
  for (unsigned lin = 0; lin < (unsigned)(nb * nf); lin++)
    run_block(lin, NT, [&] { kf_eig_c2r_y<TS, M, L, NT>(spec, Hp, H, N0h, inv_next, 0., thrs, eig, sq, part, tw); });
  // direct 2D DFT of e (real part), classification and roots
  int bad = 0;
  double md = 0, mx = 0, msq = 0;
  for (int f = 0; f < nf; f++) {
    const double* e = &efull[(size_t)f * N1 * N0];
    ClassifyPartial want = { 0, 0, 0, 1.7976931348623157e308 }, got = { 0, 0, 0, 1.7976931348623157e308 };
    for (int b = 0; b < nb; b++) {
      const ClassifyPartial& q = part[f * nb + b];
      got.small += q.small; got.zero += q.zero; got.negative += q.negative;
      got.smallest = std::min(got.smallest, q.smallest);
    }
    for (int ky = 0; ky < N1; ky++)
      for (int kx = 0; kx < H; kx++) {
        double re = 0, im = 0;
        for (int j = 0; j < N1; j++) {
          // x spectrum of row j, folded to the stored rows j <= M by symmetry
          double sx, sy;
          if (j <= M) { sx = spec[((size_t)f * (M + 1) + j) * Hp + kx].x; sy = spec[((size_t)f * (M + 1) + j) * Hp + kx].y; }
          else { sx = spec[((size_t)f * (M + 1) + (N1 - j)) * Hp + kx].x; sy = -spec[((size_t)f * (M + 1) + (N1 - j)) * Hp + kx].y; }
          const double a = -2 * M_PI * (double)((int64_t)ky * j % N1) / N1;
          re += sx * cos(a) - sy * sin(a);
          im += sx * sin(a) + sy * cos(a);
        }
        double v = re * inv_next;
        mx = std::max(mx, fabs(v));
        classify_one(v, thrs[f], want);
        const size_t o = f * cnt + (size_t)ky * H + kx;
        if (eig) md = std::max(md, fabs(eig[o] - v));
        const double s = sqrt(v);
        msq = std::max(msq, fabs((double)sq[o] - s) / (sizeof(TS) == 4 ? 1e-7 * (1 + s) : 1e-13 * (1 + s)));
        if (std::isnan((double)sq[o])) msq = INFINITY;
        (void)e;
      }
    const bool okc = got.small == want.small && got.zero == want.zero && got.negative == want.negative &&
                     fabs(got.smallest - want.smallest) <= 1e-12 * mx;
    printf("  realization %d: classification small %lld/%lld zero %lld/%lld negative %lld/%lld %s\n", f, got.small,
           want.small, got.zero, want.zero, got.negative, want.negative, okc ? "OK" : "FAIL");
    bad += !okc;
  }
  const bool ok = (!eig || md <= 1e-12 * mx) && msq <= 1.0;
  printf("M=%d N0=%d L=%d NT=%d %s eig %s: max |eig - direct| / max %.3g, roots %s\n", M, N0, L, NT,
         sizeof(TS) == 4 ? "FP32 roots" : "FP64 roots", eig ? "kept" : "not kept", eig ? md / mx : 0.0,
         msq <= 1.0 ? "OK" : "FAIL");
  fflush(stdout);
  return bad + !ok;
}

int main() {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  int bad = 0;
  bad += run_case<float, 64>(64, false);
  bad += run_case<double, 64>(64, true);
  bad += run_case<float, 96>(48, true);
  bad += run_case<float, 16>(32, true);
  bad += run_case<double, 512>(16, true);
  bad += run_case<float, 768>(16, false);
  bad += run_case<float, 1024>(32, true);
  printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
