// Host build of the generic (runtime-size) DCT/DST passes kd_gen_last,
// kd_mid and kd_x of pf_dctdst.hpp at the launch geometry the library gives
// them when their LDS budget is 80 KB: 8-line tiles of M = 1024 with the pad
// line and the X_M slots, ((8 + 1) * 1024 + 8) * 8 = 73,792 bytes. This is
// the tile of the two 1024^3 DCT/DST runs that faulted on the GPU (generic
// passes with an 80 KB budget, and an earlier form of the specialised passes
// with a pad line). Every buffer ends at a PROT_NONE guard region; an access
// past an end or before a start stops the run naming the buffer, offset and
// block. Small shapes run every block; the 1024^3 shape (three realizations
// per launch) runs the first, last and edge blocks and random others.
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
#include "pf_dctdst.hpp"
using namespace pfamd;

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
static std::vector<unsigned> pick(unsigned maxblocks) {
  const unsigned tot = gridDim.x * gridDim.y * gridDim.z;
  std::vector<unsigned> v;
  if (tot <= maxblocks) { for (unsigned i = 0; i < tot; i++) v.push_back(i); return v; }
  for (unsigned i = 0; i < 4; i++) { v.push_back(i); v.push_back(tot - 1 - i); }
  const unsigned ex = gridDim.x - 1; // edge column kx = P0 - 1 (tiles of L rows)
  for (unsigned z = 0; z < gridDim.z; z++) {
    v.push_back(ex + gridDim.x * gridDim.y * z);
    v.push_back(ex + gridDim.x * (gridDim.y - 1 + gridDim.y * z));
  }
  std::mt19937 r(17); while (v.size() < maxblocks) v.push_back(r() % tot);
  return v;
}
static cplx<float>* table(int N) {
  cplx<float>* tw = (cplx<float>*)galloc("tw", (size_t)N * 8);
  for (int j = 0; j < N; j++) tw[j] = { (float)cos(-2 * M_PI * j / N), (float)sin(-2 * M_PI * j / N) };
  return tw;
}
// radix schedule of the library's make_sched (the rule of ct_radix)
static Sched sched(int n) {
  Sched s{}; s.n = n; s.nst = ct_nst(n);
  for (int i = 0; i < s.nst; i++) s.rad[i] = ct_radix(n, i);
  return s;
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
// the library's tile choice for the generic passes (generate_dctdst): the
// largest power of two L <= 16 dividing P0 - 1 (x pass: <= 16, and below the
// number of rows) with ((L + 1) M + L) complex values within the budget; one
// thread per 32 values
static const size_t BUDGET = 80 * 1024;
static int pow2_div(int n, int cap) { int L = 1; while (L * 2 <= cap && n % (L * 2) == 0) L *= 2; return L; }
static int tile_L(int L, int M) { while (L > 1 && ((size_t)(L + 1) * M + L) * 8 > BUDGET) L /= 2; return L; }
static int threads(int L, int M) { return std::min(1024, std::max(64, ((L * M / 32 + 63) / 64) * 64)); }

static void passes(const DctGeom& g, int nb, unsigned maxblocks, const char* tag) {
  float* sq = (float*)galloc("sq", (size_t)g.Ptot * 4);
  // large shapes check bounds only: their root table stays untouched (zero pages)
  if (g.Ptot < ((int64_t)1 << 26))
    for (int64_t i = 0; i < g.Ptot; i++) sq[i] = 0.5f;
  float* G = (float*)galloc("G", (size_t)nb * 4 * g.P1 * g.P0p * g.n2 * 4);
  float* H = (float*)galloc("Hh", (size_t)nb * 2 * g.n1 * g.n2 * g.P0p * 4);
  float* out = (float*)galloc("out", (size_t)nb * g.n0 * g.n1 * g.n2 * 4);
  cplx<float>* twz = table(2 * g.n2); cplx<float>* twy = table(2 * g.n1); cplx<float>* twx = table(2 * g.n0);
  const uint64_t seed = 0x9e3779b97f4a7c15ull, first = 2;
  size_t lds[3];
  int Ls[3];
  {
    const int M = g.n2, L = tile_L(pow2_div(g.P0 - 1, 16), M), nt = threads(L, M);
    Ls[0] = L; lds[0] = ((size_t)(L + 1) * M + L) * 8;
    const Sched s = sched(M);
    gridDim = { (unsigned)((g.P0 - 1) / L + 1), (unsigned)g.P1, (unsigned)nb }; blockDim = { (unsigned)nt, 1, 1 };
    cur = "kd_gen_last";
    for (unsigned lin : pick(maxblocks))
      run_block(lin, nt, [&] { kd_gen_last<float, 0, 0, 0, 1>(g, sq, nullptr, seed, first, G, twz, s, M, L); });
  }
  {
    const int M = g.n1, L = tile_L(pow2_div(g.P0 - 1, 16), M), nt = threads(L, M);
    Ls[1] = L; lds[1] = ((size_t)(L + 1) * M + L) * 8;
    const Sched s = sched(M);
    gridDim = { (unsigned)((g.P0 - 1) / L + 1), (unsigned)g.n2, (unsigned)nb }; blockDim = { (unsigned)nt, 1, 1 };
    cur = "kd_mid";
    for (unsigned lin : pick(maxblocks))
      run_block(lin, nt, [&] { kd_mid<float, 0, 0, 0>(g, G, H, twy, s, M, L); });
  }
  {
    const int M = g.n0;
    const int64_t rows = (int64_t)g.n1 * g.n2;
    int L = tile_L(16, M);
    while (L > 1 && L / 2 >= rows) L /= 2;
    const int nt = threads(L, M);
    Ls[2] = L; lds[2] = ((size_t)(L + 1) * M + L) * 8;
    const Sched s = sched(M);
    gridDim = { (unsigned)((rows + L - 1) / L), (unsigned)nb, 1 }; blockDim = { (unsigned)nt, 1, 1 };
    cur = "kd_x";
    for (unsigned lin : pick(maxblocks))
      run_block(lin, nt, [&] { kd_x<float, 0, 0, 0>(g, H, out, rows, twx, s, M, L, PF_TRANSFORM_SIGN, 0.5f); });
  }
  printf("%-34s n=(%d,%d,%d) nb=%d tiles L=(%d,%d,%d) LDS=(%zu,%zu,%zu) B OK\n", tag, g.n0, g.n1, g.n2, nb,
         Ls[0], Ls[1], Ls[2], lds[0], lds[1], lds[2]);
  fflush(stdout);
}

int main() {
  struct sigaction sa {}; sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO; sigaction(SIGSEGV, &sa, nullptr); sigaction(SIGBUS, &sa, nullptr);
  const bool full = getenv("PF_EMU_FULL") != nullptr;
  for (int tr = 0; tr < 2; tr++)
    passes(geom(32, 64, 16, tr), 2, 1u << 30, tr ? "all blocks, transposed" : "all blocks, x-fastest");
  passes(geom(16, 1024, 1024, 1), 1, full ? 96 : 16, "M=1024 in z and y, sampled");
  for (int tr = 1; tr >= 0; tr--)
    passes(geom(1024, 1024, 1024, tr), 3, full ? 48 : 18, tr ? "1024^3 transposed, sampled" : "1024^3 x-fastest, sampled");
  printf("all OK\n");
  return 0;
}
