// pf_fast.hpp -- compile-time specialised line-FFT kernels of the hot path.
//
// Line length N, lines per tile L and threads NT are template parameters, so
// the radix schedule, the stage loop and every index computation fold to
// constants and shifts. Twiddles come from a small global table (N entries,
// or N0 = 2M for the half-length real transforms) that stays cache resident,
// so the whole LDS budget holds the tile.
//
// Tiles of the strided passes are L consecutive kx of one (k_mid) row, so
// global loads and stores move L*sizeof(complex) contiguous bytes per
// element; the Nyquist column kx = N0/2 forms its own tiles of L
// consecutive rows, which keeps H = N0/2+1 (odd) out of every division.
#pragma once

#include <type_traits>

#include "pf_common.hpp"
#include "pf_fft.hpp"

namespace pfamd {

// --- constexpr radix schedule (same rule as the host's make_sched) --------
constexpr int ct_p2(int n)
{
  int p = 0;
  while (n % 2 == 0) {
    n /= 2;
    p++;
  }
  return p;
}

constexpr int ct_radix(int n, int st)
{
  int p2 = ct_p2(n), m = n;
  while (m % 2 == 0)
    m /= 2;
  int n8 = p2 / 3, rem = p2 % 3;
  int t0 = 0, t1 = 0, nt2 = 0;
  if (rem == 1 && n8 > 0) {
    n8--;
    t0 = 4;
    t1 = 4;
    nt2 = 2;
  } else if (rem == 1) {
    t0 = 2;
    nt2 = 1;
  } else if (rem == 2) {
    t0 = 4;
    nt2 = 1;
  }
  if (st < n8)
    return 8;
  st -= n8;
  if (st < nt2)
    return st == 0 ? t0 : t1;
  st -= nt2;
  while (m % 3 == 0) {
    if (st == 0)
      return 3;
    st--;
    m /= 3;
  }
  while (m % 5 == 0) {
    if (st == 0)
      return 5;
    st--;
    m /= 5;
  }
  while (m % 7 == 0) {
    if (st == 0)
      return 7;
    st--;
    m /= 7;
  }
  return 0;
}

constexpr int ct_nst(int n)
{
  int s = 0;
  while (ct_radix(n, s) != 0)
    s++;
  return s;
}

// Stores of data that no later access of the same kernel reads (the pass
// outputs B1, B2 and the fields): non-temporal, so they do not displace
// what the passes re-read from L2 (twiddles, root eigenvalues).
template <typename T>
__device__ __forceinline__ void store_nt(T* p, T a, T b)
{
  using V = T __attribute__((ext_vector_type(2)));
  const V w = { a, b };
  __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
}
template <typename T>
__device__ __forceinline__ void store_nt(cplx<T>* p, const cplx<T>& v)
{
  store_nt(reinterpret_cast<T*>(p), v.x, v.y);
}

// Source of a transform's first stage / sink of its last stage: the tile in
// LDS (the default), or a callable in(e, l) -> value / out(e, l, value) that
// reads the input from or writes the output to global memory directly.
struct TileLds
{
};

// LDS tile index of (element e, line l). PAD >= 0: row pitch L + PAD.
// PAD < 0: row pitch L with the rows permuted, e -> e ^ ((e >> 3) & 3). A
// first Stockham stage of radix R writes rows j*R + r for eight consecutive j
// at once; for 64-byte rows (L = 8 complex FP32) those rows all start on the
// same bank, and the permutation spreads them over the four bank offsets of
// a 64-byte row. It keeps each aligned group of 8 rows in place as a set, so
// the accesses of the other stages (eight consecutive rows of one aligned
// group per wave) stay contiguous.
template <int L, int PAD>
__device__ __forceinline__ int tix(int e, int l)
{
  if constexpr (PAD >= 0)
    return e * (L + PAD) + l;
  else
    return (e ^ ((e >> 3) & 3)) * L + l;
}

// One Stockham stage; tile layout buf[tix<L, PAD>(e, l)]; lane mapping
// l fastest. With In = TileLds it reads the tile (and waits at a barrier
// before overwriting it), else it takes its inputs from in(e, l); with
// Out = TileLds it writes the tile and ends with a barrier, else it hands
// each output to out(e, l, v) and leaves the LDS tile unused. RO < R (last
// stage, global output only) emits outputs r < RO of each butterfly: with
// NS * R = N those are the elements below RO * NS, and the arithmetic of the
// others is left out.
template <typename T, int N, int L, int NT, int PAD, int R, int NS, bool INV, int TWM,
          typename In = TileLds, typename Out = TileLds, int RO = R>
__device__ __forceinline__ void tstage(cplx<T>* buf, const cplx<T>* tw, int tid,
                                       const In& in = In{}, const Out& out = Out{})
{
  constexpr bool LDS_IN = std::is_same<In, TileLds>::value;
  constexpr bool LDS_OUT = std::is_same<Out, TileLds>::value;
  static_assert(RO == R || (!LDS_OUT && NS * R == N), "pruned outputs: last stage, global sink");
  constexpr int NR = N / R;
  constexpr int TOTAL = L * NR;
  constexpr int IT = (TOTAL + NT - 1) / NT;
  constexpr int TSTEP = (N / (NS * R)) * TWM;
  cplx<T> v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int w = tid + it * NT;
    if (TOTAL % NT == 0 || w < TOTAL) {
      const int l = w % L, j = w / L;
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (LDS_IN)
          v[it][r] = buf[tix<L, PAD>(j + r * NR, l)];
        else
          v[it][r] = in(j + r * NR, l);
      }
      if (NS > 1) {
        // one table load per butterfly: w^r = w^(r/2) * w^(r - r/2) from w
        // (at most three products deep for radix 8) instead of R - 1 loads
        cplx<T> t[R];
        t[1] = tw[k * TSTEP];
        if (INV)
          t[1].y = -t[1].y;
#pragma unroll
        for (int r = 2; r < R; r++)
          t[r] = cmul(t[r / 2], t[r - r / 2]);
#pragma unroll
        for (int r = 1; r < R; r++)
          v[it][r] = cmul(v[it][r], t[r]);
      }
      Butterfly<T, R, INV>::run(v[it]);
    }
  }
  if constexpr (LDS_IN && LDS_OUT)
    __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int w = tid + it * NT;
    if (TOTAL % NT == 0 || w < TOTAL) {
      const int l = w % L, j = w / L;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < RO; r++) {
        if constexpr (LDS_OUT)
          buf[tix<L, PAD>(base + r * NS, l)] = v[it][r];
        else
          out(base + r * NS, l, v[it][r]);
      }
    }
  }
  if constexpr (LDS_OUT)
    __syncthreads();
}

// tid: the thread's index in the block. Callers inside a loop pass a value
// from opaque_tid() so the per-element index math is not hoisted out of the
// loop (where it would stay live in registers).
template <typename T, int N, int L, int NT, int PAD, bool INV, int TWM, int ST = 0, int NS = 1>
__device__ __forceinline__ void tfft(cplx<T>* buf, const cplx<T>* tw, int tid = threadIdx.x)
{
  if constexpr (ST < ct_nst(N)) {
    constexpr int R = ct_radix(N, ST);
    tstage<T, N, L, NT, PAD, R, NS, INV, TWM>(buf, tw, tid);
    tfft<T, N, L, NT, PAD, INV, TWM, ST + 1, NS * R>(buf, tw, tid);
  }
}

// tfft with the first stage reading in(e, l) and the last stage writing
// out(e, l, v) (either may be TileLds). A first stage that reads global
// memory needs no barrier before it writes the tile; a last stage that writes
// global memory leaves the tile without a closing barrier. HALF: the sink
// keeps no output at or above N/2, so a last stage of even radix computes
// and hands over only the lower half of each butterfly.
template <typename T, int N, int L, int NT, int PAD, bool INV, int TWM, bool HALF = false,
          typename In, typename Out, int ST = 0, int NS = 1>
__device__ __forceinline__ void tfft_io(cplx<T>* buf, const cplx<T>* tw, int tid, const In& in,
                                        const Out& out)
{
  if constexpr (ST < ct_nst(N)) {
    constexpr int R = ct_radix(N, ST);
    constexpr bool first = ST == 0, last = ST + 1 == ct_nst(N);
    constexpr bool lds_out = std::is_same<Out, TileLds>::value;
    constexpr int RO = (HALF && last && !lds_out && R % 2 == 0) ? R / 2 : R;
    if constexpr (first && last)
      tstage<T, N, L, NT, PAD, R, NS, INV, TWM, In, Out, RO>(buf, tw, tid, in, out);
    else if constexpr (first)
      tstage<T, N, L, NT, PAD, R, NS, INV, TWM, In, TileLds>(buf, tw, tid, in, TileLds{});
    else if constexpr (last)
      tstage<T, N, L, NT, PAD, R, NS, INV, TWM, TileLds, Out, RO>(buf, tw, tid, TileLds{}, out);
    else
      tstage<T, N, L, NT, PAD, R, NS, INV, TWM>(buf, tw, tid);
    tfft_io<T, N, L, NT, PAD, INV, TWM, HALF, In, Out, ST + 1, NS * R>(buf, tw, tid, in, out);
  }
}

// threadIdx.x through an empty asm statement: each call yields a value the
// compiler cannot prove equal to the previous one.
__device__ __forceinline__ int opaque_tid()
{
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Tile loops with a compile-time trip count: thread t handles elements
// t, t + NT, ... of a tile of TOTAL elements. With the count known, every
// global load of a tile is issued before the first wait (a loop bounded by
// threadIdx.x would wait out one memory latency per element).
template <int TOTAL, int NT>
struct TileLoop
{
  static constexpr int IT = (TOTAL + NT - 1) / NT;
  static __device__ __forceinline__ bool live(int it, int tid = threadIdx.x)
  {
    return TOTAL % NT == 0 || tid + it * NT < TOTAL;
  }
  static __device__ __forceinline__ int idx(int it, int tid = threadIdx.x) { return tid + it * NT; }
};

// Geometry of the half spectrum as the generation / matvec last-axis passes
// see it: line (kx, km) with elements e along the last axis.
struct LastAxis
{
  int H, N0h;        // H = N0/2 + 1, N0h = N0/2
  int Hp;            // row pitch of the work arrays (H rounded up to 128-byte rows)
  int Nmid;          // rows of lines (3D: N1, 2D: 1)
  int nlast;         // pruned output length
  int64_t sA, sB;    // storage index base = H * (e*sA + km*sB)
};

__device__ __forceinline__ bool self_conj(int k, int N) { return k == 0 || 2 * k == N; }

template <typename T>
__device__ __forceinline__ cplx<T> make_entry(T lambda, T r1, T r2, bool selfconj)
{
  if (selfconj)
    return { lambda * r1, T(0) };
  const T ls = lambda * (T)0.70710678118654752440; // lambda / sqrt(2)
  return { ls * r1, ls * r2 };
}

// Spectrum entry of one Box-Muller pair: make_entry(lambda, normals of
// (xa, xb)). FP32 folds the 1/sqrt(2) of the entries off the self-conjugate
// set into the root's argument and lambda into the radius, ahead of the
// cosine / sine products.
template <typename T>
__device__ __forceinline__ cplx<T> gen_entry(uint32_t xa, uint32_t xb, T lambda, bool selfconj)
{
  T r1, r2;
  BoxMuller<T>::pair(xa, xb, r1, r2);
  return make_entry<T>(lambda, r1, r2, selfconj);
}

template <>
__device__ __forceinline__ cplx<float> gen_entry<float>(uint32_t xa, uint32_t xb, float lambda,
                                                        bool selfconj)
{
  // lambda * sqrt(-2 ln u) / sqrt(2) = lambda * sqrt(-ln u); ln u = log2(u) ln 2
  const float k = selfconj ? -2.0f * 0.69314718055994531f : -0.69314718055994531f;
  const float t =
    lambda * __builtin_amdgcn_sqrtf(k * __builtin_amdgcn_logf(BoxMuller<float>::turn(xa)));
  const float ub = BoxMuller<float>::turn(xb);
  return { t * __builtin_amdgcn_cosf(ub), selfconj ? 0.0f : t * __builtin_amdgcn_sinf(ub) };
}

// Philox block `blk` of realization `real`. With 32-bit slot walks every
// block index of a realization is below 2^31 (its high counter word is 0), so
// the rounds with wave-uniform halves run partly on the scalar unit.
template <bool IDX64>
__device__ __forceinline__ void philox_block(uint64_t blk, uint64_t real, uint32_t k0,
                                             uint32_t k1, uint32_t c[4])
{
  if constexpr (IDX64) {
    c[0] = (uint32_t)blk;
    c[1] = (uint32_t)(blk >> 32);
    c[2] = (uint32_t)real;
    c[3] = (uint32_t)(real >> 32);
    philox4x32_10(c, k0, k1);
  } else {
    philox4x32_10_lane0((uint32_t)blk, (uint32_t)real, (uint32_t)(real >> 32), k0, k1, c);
  }
}

// logical (k0, k1, k2) of line (kx, km) element e
__device__ __forceinline__ void line_k(const GenGeom& g, int kx, int km, int e, int& k1, int& k2)
{
  if (g.dim == 3) {
    k1 = km;
    k2 = e;
  } else {
    k1 = e;
    k2 = 0;
  }
}

// Pass G: spectrum generated in registers along the last axis (length N),
// inverse DFT, first nlast outputs -> B1[j][km][kx].
// IDX64: storage indices and eigenvalue offsets in 64 bits (fields with 2^31
// or more half-spectrum entries); otherwise the slot walk runs in 32 bits and
// the eigenvalue loads and B1 stores take 32-bit byte offsets from a
// block-uniform base (the launcher checks that a realization's eigenvalue
// table and B1 slab stay below 4 GiB). HALF: nlast <= N/2, so the last FFT
// stage computes the lower half of its outputs only.
// Body of one tile (bx = kx tile or the Nyquist column, by = k_mid row or
// group, b = realization of the launch).
template <typename T, int N, int L, int NT, int NOISE, bool IDX64, bool HALF = false>
__device__ __forceinline__ void gen_last_tile(const GenGeom& g, const LastAxis& a, const T* sq,
                                              const double* noise, uint64_t seed,
                                              uint64_t first_real, cplx<T>* B1,
                                              const cplx<T>* tw, int bx, int by, int b)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int ntile = a.N0h / L;
  const bool nyq = bx >= ntile;
  int kx0, km0;
  if (!nyq) {
    kx0 = bx * L;
    // DCT-half eigenvalues: rows km and Nmid-km read the same table row, so
    // they take neighbouring logical positions (and one XCD's L2)
    km0 = (g.dct && g.dim == 3) ? ((by & 1) ? a.Nmid - ((by + 1) >> 1) : (by >> 1)) : by;
  } else {
    kx0 = a.N0h;
    km0 = by * L;
    if (km0 >= a.Nmid)
      return;
  }
  const uint64_t real = first_real + (g.fidx ? (uint64_t)g.fidx[b] : (uint64_t)b);
  sq += (int64_t)b * g.sq_stride;
  const double* nz = noise ? noise + (int64_t)b * g.noise_stride : nullptr;
  if (NOISE == 1 && !nyq) {
    // One Philox block yields the entries 2m and 2m+1 of the storage order.
    // Row e of the tile holds entries base(e) + l, l < L (k_mid is fixed,
    // so the storage index and the eigenvalue row are affine in e, and
    // kx < N0/2 needs no folding). An even row base makes the row L/2 whole
    // blocks; an odd one makes it L/2 - 1 whole blocks plus a block whose
    // second entry is column 0 and one whose first entry is column L-1.
    // Main pass: thread t owns block slot q = t % S (S = L/2) of the rows
    // e = t/S + it*ES. ES is even, so a thread's rows share one parity and
    // its two columns 2q - odd and 2q - odd + 1 (column -1 is dropped).
    // Tail pass: column L-1 of the odd rows, one block each. Every lane of
    // the main pass does useful work, and no index test is left in its loop.
    static_assert(L % 2 == 0, "entries are generated in pairs");
    constexpr int S = L / 2, ES = NT / S;
    static_assert(NT % S == 0 && ES % 2 == 0, "a thread's rows share one parity");
    constexpr int ITM = (N + ES - 1) / ES;
    const bool d3 = g.dim == 3;
    const int Pe = d3 ? g.P2 : g.P1, Ne = d3 ? g.N2 : g.N1;
    int64_t eu, es; // eigen row offset = eu + es * e' (e' = e folded for the DCT half)
    if (g.dct) {
      const int f1 = km0 >= g.P1 ? g.N1 - km0 : km0;
      eu = d3 ? (int64_t)g.P0 * f1 : 0;
      es = d3 ? (int64_t)g.P0 * g.P1 : g.P0;
    } else {
      eu = d3 ? (int64_t)g.H * km0 : 0;
      es = d3 ? (int64_t)g.H * g.N1 : g.H;
    }
    // kx = 0 (column 0 of the first tile) is the only self-conjugate column
    const bool scq = (d3 ? self_conj(km0, g.N1) : true) && kx0 == 0;
    const uint32_t sk0 = (uint32_t)seed, sk1 = (uint32_t)(seed >> 32);
    using IX = typename std::conditional<IDX64, int64_t, uint32_t>::type;
    const IX rowstep = (IX)((int64_t)a.H * a.sA); // storage index step per e
    const IX est = (IX)es;
    const T* sqt = sq + eu + kx0;
    // eigenvalue at element offset i from sqt
    auto sq_at = [sqt](auto i) -> T {
      if constexpr (IDX64)
        return sqt[i];
      else
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(sqt) +
                                           (uint32_t)i * (uint32_t)sizeof(T));
    };
    // the eigenvalues at offsets i and i + 1 in one load (element-aligned)
    auto sq_pair = [sqt](auto i, T& x, T& y) {
      const char* p;
      if constexpr (IDX64)
        p = reinterpret_cast<const char*>(sqt + i);
      else
        p = reinterpret_cast<const char*>(sqt) + (uint32_t)i * (uint32_t)sizeof(T);
      T v[2];
      __builtin_memcpy(v, p, sizeof(v));
      x = v[0];
      y = v[1];
    };
    const IX base0 = (IX)((int64_t)a.H * ((int64_t)km0 * a.sB) + kx0);
    const int par0 = (int)(base0 & 1), rs = (int)(rowstep & 1);
    const int tid = (int)threadIdx.x;
    // the tile of kx = 0 holds self-conjugate entries; every other tile runs
    // a copy of the main pass without their tests
    auto main_pass = [&](auto with_sc) {
      constexpr bool SC = decltype(with_sc)::value;
      const int q = tid % S;
      int e = tid / S;
      const int odd = (par0 + rs * e) & 1;
      const int la = 2 * q - odd, lb = la + 1, lac = la < 0 ? 0 : la, ia = la < 0 ? lb : la;
      IX base = base0 + rowstep * (IX)e + (IX)la; // storage index of column la: even
      IX up = est * (IX)e, dn = est * (IX)(Ne - e); // es*e and es*(Ne-e)
#pragma unroll
      for (int it = 0; it < ITM; it++) {
        if (N % ES == 0 || e < N) {
          const IX row = (g.dct && e >= Pe) ? dn : up;
          // columns lac and lac + 1 (= lb, or column 1 when la = -1 and lb = 0)
          T sa, sp;
          sq_pair(row + (IX)lac, sa, sp);
          const T sb = la < 0 ? sa : sp;
          const uint64_t blk = (uint64_t)base >> 1;
          uint32_t c[4];
          philox_block<IDX64>(blk, real, sk0, sk1, c);
          const bool sce = SC && self_conj(e, Ne);
          const cplx<T> va = gen_entry<T>(c[0], c[1], sa, sce && la == 0);
          const cplx<T> vb = gen_entry<T>(c[2], c[3], sb, sce && lb == 0);
          // column -1 is dropped by storing it to column lb first: a store
          // left unconditional keeps va (and its eigenvalue load) out of a
          // branch, where the load's latency would be waited out at once
          buf[e * L + ia] = va;
          buf[e * L + lb] = vb;
        }
        e += ES;
        base += rowstep * (IX)ES;
        up += est * (IX)ES;
        dn -= est * (IX)ES;
      }
    };
    if (scq)
      main_pass(std::true_type{});
    else
      main_pass(std::false_type{});
    {
      // odd rows: every row when the row step is even and base(0) odd, every
      // other row from the first odd one when the row step is odd
      const int e0 = rs ? 1 - par0 : 0, estep = rs ? 2 : 1;
      const int nodd = rs ? (N - e0 + 1) / 2 : (par0 ? N : 0);
      constexpr int ITT = (N + NT - 1) / NT;
#pragma unroll
      for (int it = 0; it < ITT; it++) {
        const int i = tid + it * NT;
        if (i < nodd) {
          const int e = e0 + estep * i;
          const IX row = (g.dct && e >= Pe) ? est * (IX)(Ne - e) : est * (IX)e;
          const T sa = sq_at(row + (IX)(L - 1));
          const uint64_t blk = (uint64_t)(IX)(base0 + rowstep * (IX)e + (IX)(L - 1)) >> 1;
          uint32_t c[4];
          philox_block<IDX64>(blk, real, sk0, sk1, c);
          buf[e * L + L - 1] = gen_entry<T>(c[0], c[1], sa, false);
        }
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < L * N; idx += NT) {
      const int l = idx % L, e = idx / L;
      const int kx = nyq ? kx0 : kx0 + l;
      const int km = nyq ? km0 + l : km0;
      cplx<T> z = { 0, 0 };
      if (km < a.Nmid) {
        int k1, k2;
        line_k(g, kx, km, e, k1, k2);
        const int64_t s = (int64_t)a.H * ((int64_t)e * a.sA + (int64_t)km * a.sB) + kx;
        T r1, r2;
        if (NOISE == 0) {
          r1 = (T)nz[2 * s];
          r2 = (T)nz[2 * s + 1];
        } else {
          philox_pair<T>(seed, real, (uint64_t)s, r1, r2);
        }
        const T lam = sq[eig_index(g, kx, k1, k2)];
        z = make_entry<T>(lam, r1, r2,
                          self_conj(kx, g.N0) && self_conj(k1, g.N1) && self_conj(k2, g.N2));
      }
      buf[e * L + l] = z;
    }
  }
  __syncthreads();
  // the last stage writes its outputs j < nlast straight to B1: element
  // (j, l) of the tile at j*plane + l*lstep from a block-uniform base
  const int64_t plane = (int64_t)a.Hp * a.Nmid;
  cplx<T>* ob = B1 + (int64_t)b * plane * a.nlast + (int64_t)km0 * a.Hp + kx0;
  const int64_t lstep = nyq ? a.Hp : 1;
  const int lmax = nyq ? a.Nmid - km0 : L; // live columns of the tile
  if constexpr (IDX64) {
    auto store = [&](int j, int l, const cplx<T>& v) {
      if (j < a.nlast && l < lmax)
        store_nt(&ob[(int64_t)j * plane + l * lstep], v);
    };
    tfft_io<T, N, L, NT, 0, true, 1, HALF>(buf, tw, (int)threadIdx.x, TileLds{}, store);
  } else {
    const uint32_t pb = (uint32_t)plane * (uint32_t)sizeof(cplx<T>);
    const uint32_t lb = (uint32_t)lstep * (uint32_t)sizeof(cplx<T>);
    char* obb = reinterpret_cast<char*>(ob);
    auto store = [&](int j, int l, const cplx<T>& v) {
      if (j < a.nlast && l < lmax)
        store_nt(reinterpret_cast<cplx<T>*>(obb + ((uint32_t)j * pb + (uint32_t)l * lb)), v);
    };
    tfft_io<T, N, L, NT, 0, true, 1, HALF>(buf, tw, (int)threadIdx.x, TileLds{}, store);
  }
}

template <typename T, int N, int L, int NT, int NOISE, bool IDX64, bool HALF = false>
__global__ void __launch_bounds__(NT) kf_gen_last(GenGeom g, LastAxis a, const T* sq,
                                                  const double* noise, uint64_t seed,
                                                  uint64_t first_real, cplx<T>* B1,
                                                  const cplx<T>* tw)
{
  int bx, by, b;
  xcd_block(bx, by, b);
  gen_last_tile<T, N, L, NT, NOISE, IDX64, HALF>(g, a, sq, noise, seed, first_real, B1, tw, bx,
                                                 by, b);
}

// Pass along the middle axis (3D): lines (kx, z) of in[z][e][kx] (in_rows
// rows per z, n_in leading points, rest zero), DFT, first n_out points ->
// out[z][j][kx] (out_rows rows per z).
// Body of tile (bx, by) of realization b. Loads and stores take 32-bit byte
// offsets from block-uniform bases (the launcher checks that a tile's span
// stays below 4 GiB). HALF: n_out <= N/2, so the last FFT stage computes the
// lower half of its outputs only.
template <typename T, int N, int L, int NT, bool INV, bool HALF = false>
__device__ __forceinline__ void mid_tile(const cplx<T>* in, cplx<T>* out, int Hp, int N0h, int nz,
                                         int in_rows, int n_in, int out_rows, int n_out,
                                         const cplx<T>* tw, int bx, int by, int b)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int ntile = N0h / L;
  const bool nyq = bx >= ntile;
  int kx0, z0;
  if (!nyq) {
    kx0 = bx * L;
    z0 = by;
  } else {
    kx0 = N0h;
    z0 = by * L;
    if (z0 >= nz)
      return;
  }
  in += (int64_t)b * Hp * nz * in_rows;
  out += (int64_t)b * Hp * nz * out_rows;
  // column l of the tile: kx0 + l (or row group z0 + l of the Nyquist column);
  // element e of a line at e*Hp from the column's base. The first stage
  // reads the lines from B1 and the last stage writes rows j < n_out to B2.
  const int lmax = nyq ? nz - z0 : L; // live columns of the tile
  const cplx<T>* ib = in + (int64_t)z0 * in_rows * Hp + kx0;
  cplx<T>* ob = out + (int64_t)z0 * out_rows * Hp + kx0;
  constexpr uint32_t CB = sizeof(cplx<T>);
  const uint32_t eb = (uint32_t)Hp * CB;
  const uint32_t linb = nyq ? (uint32_t)in_rows * eb : CB, loutb = nyq ? (uint32_t)out_rows * eb : CB;
  const char* ibb = reinterpret_cast<const char*>(ib);
  char* obb = reinterpret_cast<char*>(ob);
  auto load = [&](int e, int l) -> cplx<T> {
    cplx<T> v = { 0, 0 };
    if (e < n_in && l < lmax)
      v = *reinterpret_cast<const cplx<T>*>(ibb + ((uint32_t)e * eb + (uint32_t)l * linb));
    return v;
  };
  auto store = [&](int j, int l, const cplx<T>& v) {
    if (j < n_out && l < lmax)
      store_nt(reinterpret_cast<cplx<T>*>(obb + ((uint32_t)j * eb + (uint32_t)l * loutb)), v);
  };
  tfft_io<T, N, L, NT, 0, INV, 1, HALF>(buf, tw, (int)threadIdx.x, load, store);
}

template <typename T, int N, int L, int NT, bool INV, bool HALF = false>
__global__ void __launch_bounds__(NT) kf_mid(const cplx<T>* in, cplx<T>* out, int Hp, int N0h,
                                             int nz, int in_rows, int n_in, int out_rows,
                                             int n_out, const cplx<T>* tw)
{
  // tile rows narrower than a 128-byte line (FP64 tiles of 2 or 4 lines):
  // XCD-aware order, so neighbouring kx tiles, which share lines, run on one
  // XCD and meet in its L2 (whole-line rows keep the hardware order)
  int bx = blockIdx.x, by = blockIdx.y, b = blockIdx.z;
  if constexpr (L * sizeof(cplx<T>) < 128)
    xcd_block(bx, by, b);
  mid_tile<T, N, L, NT, INV, HALF>(in, out, Hp, N0h, nz, in_rows, n_in, out_rows, n_out, tw, bx, by,
                                   b);
}

// c2r along x (see k_c2r_x for the algebra), M = N0/2 point inverse FFT on a
// padded tile of L rows; first n0 outputs with the value transform.
// Body of tile bx of realization b.
template <typename T, typename TO, int M, int L, int NT>
__device__ __forceinline__ void c2r_x_tile(const cplx<T>* in, TO* out, int64_t rows, int Hp,
                                           int n0, const cplx<T>* tw, int transform, T param,
                                           const int64_t* ooff, int bx, int b)
{
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  in += (int64_t)b * rows * Hp;
  out += ooff ? ooff[b] : (int64_t)b * rows * n0; // ooff: output offset of each realization
  const int64_t row0 = (int64_t)bx * L;
  const int lrows = (int)(rows - row0 < L ? rows - row0 : L); // live rows of the tile
  const cplx<T>* ib = in + row0 * Hp;
  // rows are contiguous along k, the tile's line direction: the loads run
  // along k (coalesced) and the tile is transposed through LDS
  using TL = TileLoop<L * M, NT>;
  cplx<T> xk[TL::IT], xm[TL::IT];
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    const int idx = TL::idx(it);
    const int l = idx / M, k = idx % M;
    xk[it] = { 0, 0 };
    xm[it] = { 0, 0 };
    if (TL::live(it) && l < lrows) {
      const uint32_t off = (uint32_t)(l * Hp); // 32-bit offsets within the tile
      xk[it] = ib[off + k];
      xm[it] = ib[off + (M - k)];
    }
  }
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    const int idx = TL::idx(it);
    const int l = idx / M, k = idx % M;
    if (!TL::live(it))
      continue;
    if (k == 0) {
      xk[it].y = 0;
      xm[it].y = 0;
    }
    const cplx<T> s = { xk[it].x + xm[it].x, xk[it].y - xm[it].y };
    const cplx<T> d = { xk[it].x - xm[it].x, xk[it].y + xm[it].y };
    const cplx<T> td = cmul(cconj(tw[k]), d);
    buf[k * SE + l] = { s.x - td.y, s.y + td.x };
  }
  __syncthreads();
  tfft<T, M, L, NT, 1, true, 2>(buf, tw);
  // outputs 2m and 2m+1 of a row are the real and imaginary parts of entry
  // m: one LDS read and, for even n0 on an aligned row, one vector store.
  // Offsets within the tile are 32-bit (L * n0 is far below 2^31), so the
  // stores address a uniform 64-bit base plus a per-lane offset.
  TO* ob = out + row0 * n0;
  const bool vec = n0 % 2 == 0 && (reinterpret_cast<uintptr_t>(out) & (2 * sizeof(TO) - 1)) == 0;
  auto emit = [&](auto kind) {
    constexpr int K = decltype(kind)::value;
    auto one = [&](int it, bool pairs) {
      const int idx = TL::idx(it);
      const int l = idx / M, m = idx % M;
      if (TL::live(it) && l < lrows && 2 * m < n0) {
        const cplx<T> v = buf[m * SE + l];
        const TO lo = (TO)value_transform_k<T, K>(v.x, param);
        const TO hi = (TO)value_transform_k<T, K>(v.y, param);
        const uint32_t off = (uint32_t)(l * n0 + 2 * m);
        if (pairs) {
          store_nt(ob + off, lo, hi);
        } else {
          ob[off] = lo;
          if (2 * m + 1 < n0)
            ob[off + 1] = hi;
        }
      }
    };
    // Box-Cox's power function is long: its loop stays rolled
    if constexpr (K == PF_TRANSFORM_BOXCOX) {
#pragma unroll 1
      for (int it = 0; it < TL::IT; it++)
        one(it, vec);
    } else if (vec) {
#pragma unroll
      for (int it = 0; it < TL::IT; it++)
        one(it, true);
    } else {
#pragma unroll
      for (int it = 0; it < TL::IT; it++)
        one(it, false);
    }
  };
  // the transform is dispatched once per tile, not per output
  switch (transform) {
    case PF_TRANSFORM_LOGNORMAL:
      emit(std::integral_constant<int, PF_TRANSFORM_LOGNORMAL>{});
      break;
    case PF_TRANSFORM_FOLDEDNORMAL:
      emit(std::integral_constant<int, PF_TRANSFORM_FOLDEDNORMAL>{});
      break;
    case PF_TRANSFORM_SIGN:
      emit(std::integral_constant<int, PF_TRANSFORM_SIGN>{});
      break;
    case PF_TRANSFORM_BOXCOX:
      emit(std::integral_constant<int, PF_TRANSFORM_BOXCOX>{});
      break;
    default:
      emit(std::integral_constant<int, PF_TRANSFORM_NONE>{});
  }
}

template <typename T, typename TO, int M, int L, int NT>
__global__ void __launch_bounds__(NT) kf_c2r_x(const cplx<T>* in, TO* out, int64_t rows, int Hp,
                                               int n0, const cplx<T>* tw, int transform, T param,
                                               const int64_t* ooff)
{
  c2r_x_tile<T, TO, M, L, NT>(in, out, rows, Hp, n0, tw, transform, param, ooff, blockIdx.x,
                              blockIdx.y);
}

// Eigenvalues of a 2D R2C embedding from the x half spectrum X of the even
// part of its covariance (rows j = 0 .. M of the N1 = 2M rows, row pitch Hp
// complex; realization f at f * Hp * (M + 1)). Along y every column kx of X
// is Hermitian, X[N1 - j] = conj X[j] with rows 0 and M real, so its forward
// DFT along y is real: the inverse DFT of the conjugated column, computed as
// one M-point complex FFT of the packed column (the packing of c2r_x_tile),
// whose entry m holds outputs 2m and 2m + 1. Each eigenvalue is scaled by
// inv_next, classified and clamped (classify_one) and stored as FP64 (eig,
// optional) and as its square root (sq), both [N1][H] per realization. One
// classification partial per block at partial[f * gridDim.x + tile].
// Tile bx < N0h / L: columns bx*L .. bx*L + L - 1; tile N0h / L: the Nyquist
// column N0h alone. thrs: per-realization thresholds (nullptr: thr0).
template <typename TS, int M, int L, int NT>
__global__ void __launch_bounds__(NT) kf_eig_c2r_y(const cplx<double>* spec, int Hp, int H, int N0h,
                                                   double inv_next, double thr0,
                                                   const double* thrs, double* eig, TS* sq,
                                                   ClassifyPartial* partial,
                                                   const cplx<double>* tw)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem);
  // 4 FP64 columns are half a 128-byte line: neighbouring tiles run on one XCD
  int bx = blockIdx.x, f = blockIdx.y, bz = 0;
  if constexpr (L * sizeof(cplx<double>) < 128)
    xcd_block(bx, f, bz);
  const int kx0 = bx < N0h / L ? bx * L : N0h;
  const int lmax = bx < N0h / L ? L : 1; // live columns of the tile
  const double t = thrs ? thrs[f] : thr0;
  const int64_t cnt = (int64_t)H * (2 * M);
  const char* sb = reinterpret_cast<const char*>(spec + (int64_t)f * Hp * (M + 1) + kx0);
  TS* sqf = sq + f * cnt + kx0;
  double* eigf = eig ? eig + f * cnt + kx0 : nullptr;
  const uint32_t eb = (uint32_t)Hp * (uint32_t)sizeof(cplx<double>);
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  auto store = [&](int m, int l, const cplx<double>& v) {
    if (l < lmax) {
      double a = v.x * inv_next, b = v.y * inv_next;
      classify_one(a, t, c);
      classify_one(b, t, c);
      const uint32_t o0 = (uint32_t)(2 * m) * (uint32_t)H + (uint32_t)l, o1 = o0 + (uint32_t)H;
      if (eigf) {
        eigf[o0] = a;
        eigf[o1] = b;
      }
      sqf[o0] = (TS)sqrt(a);
      sqf[o1] = (TS)sqrt(b);
    }
  };
  {
    // the packed column Z[k] from X[k] and X[M - k]: one thread loads each
    // row pair (k, M - k) once and writes Z[k] and Z[M - k] to the tile
    using TP = TileLoop<L * (M / 2 + 1), NT>;
    cplx<double> xa[TP::IT], xb[TP::IT];
#pragma unroll
    for (int it = 0; it < TP::IT; it++) {
      const int idx = TP::idx(it);
      const int l = idx % L, k = idx / L;
      xa[it] = { 0, 0 };
      xb[it] = { 0, 0 };
      if (TP::live(it) && l < lmax) {
        const uint32_t lo = (uint32_t)l * (uint32_t)sizeof(cplx<double>);
        xa[it] = *reinterpret_cast<const cplx<double>*>(sb + ((uint32_t)k * eb + lo));
        xb[it] = *reinterpret_cast<const cplx<double>*>(sb + ((uint32_t)(M - k) * eb + lo));
      }
    }
#pragma unroll
    for (int it = 0; it < TP::IT; it++) {
      if (!TP::live(it))
        continue;
      const int idx = TP::idx(it);
      const int l = idx % L, k = idx / L;
      cplx<double> xk = xa[it], xm = xb[it];
      if (k == 0) {
        xk.y = 0;
        xm.y = 0;
      }
      {
        // packing of conj X: s = conj X[k] + X[M-k], d = conj X[k] - X[M-k]
        const cplx<double> sm = { xk.x + xm.x, xm.y - xk.y };
        const cplx<double> d = { xk.x - xm.x, -xk.y - xm.y };
        const cplx<double> td = cmul(cconj(tw[k]), d);
        buf[tix<L, 0>(k, l)] = { sm.x - td.y, sm.y + td.x };
      }
      if (k != 0 && 2 * k != M) {
        const cplx<double> sm = { xm.x + xk.x, xk.y - xm.y };
        const cplx<double> d = { xm.x - xk.x, -xm.y - xk.y };
        const cplx<double> td = cmul(cconj(tw[M - k]), d);
        buf[tix<L, 0>(M - k, l)] = { sm.x - td.y, sm.y + td.x };
      }
    }
    __syncthreads();
    tfft_io<double, M, L, NT, 0, true, 2>(buf, tw, (int)threadIdx.x, TileLds{}, store);
  }
  // block sum of the partials through the tile's LDS (free after the barrier)
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    c.small += __shfl_down(c.small, off);
    c.zero += __shfl_down(c.zero, off);
    c.negative += __shfl_down(c.negative, off);
    c.smallest = fmin(c.smallest, __shfl_down(c.smallest, off));
  }
  ClassifyPartial* red = reinterpret_cast<ClassifyPartial*>(smem);
  if (threadIdx.x % 64 == 0)
    red[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    ClassifyPartial r = red[0];
    for (int w = 1; w < NT / 64; w++) {
      r.small += red[w].small;
      r.zero += red[w].zero;
      r.negative += red[w].negative;
      r.smallest = fmin(r.smallest, red[w].smallest);
    }
    partial[(int64_t)f * gridDim.x + bx] = r;
  }
}

// r2c along x of zero-padded real rows (n0 of N0 values), see k_r2c_x.
template <typename T, typename TI, int M, int L, int NT>
__global__ void __launch_bounds__(NT) kf_r2c_x(const TI* in, cplx<T>* out, int64_t rows, int Hp,
                                               int n0, const cplx<T>* tw)
{
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int N0 = 2 * M, SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int b = blockIdx.y;
  constexpr int H = M + 1;
  in += (int64_t)b * rows * n0;
  out += (int64_t)b * rows * Hp;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  using TL = TileLoop<L * M, NT>;
  cplx<T> z[TL::IT];
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    const int idx = TL::idx(it);
    const int l = idx / M, j = idx % M;
    const int64_t r = row0 + l;
    z[it] = { 0, 0 };
    if (TL::live(it) && r < rows) {
      const int x0 = 2 * j, x1 = 2 * j + 1;
      z[it].x = x0 < n0 ? (T)in[r * n0 + x0] : T(0);
      z[it].y = x1 < n0 ? (T)in[r * n0 + x1] : T(0);
    }
  }
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    const int idx = TL::idx(it);
    if (TL::live(it))
      buf[(idx % M) * SE + idx / M] = z[it];
  }
  __syncthreads();
  tfft<T, M, L, NT, 1, false, 2>(buf, tw);
  using TH = TileLoop<L * H, NT>;
#pragma unroll
  for (int it = 0; it < TH::IT; it++) {
    const int idx = TH::idx(it);
    const int l = idx / H, k = idx % H;
    const int64_t r = row0 + l;
    if (TH::live(it) && r < rows) {
      const cplx<T> zk = buf[(k % M) * SE + l];
      const cplx<T> zm = buf[((M - k) % M) * SE + l];
      const cplx<T> s = { (zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5) };
      const cplx<T> d = { (zk.x - zm.x) * T(0.5), (zk.y + zm.y) * T(0.5) };
      const cplx<T> wd = cmul(tw[k], d);
      out[r * Hp + k] = { s.x + wd.y, s.y - wd.x };
    }
  }
}

// Fused last-axis pass of the matvec: forward DFT of n_last leading points,
// times f(k)/N_ext, inverse DFT, first n_last points back in place. The
// forward transform's first stage reads the tile's lines straight from C
// (rows e >= n_last are zero) and the inverse transform's last stage writes
// rows j < n_last straight back (HALF: n_last <= N/2, so that stage computes
// the lower half of its outputs only); the tile passes through LDS in
// between. Element (e, l) of the tile is at e*plane + l*lstep from a
// block-uniform base (32-bit offsets: the launcher checks that a
// realization's span stays below 4 GiB).
template <typename T, int N, int L, int NT, bool HALF>
__global__ void __launch_bounds__(NT) kf_fused_mult(cplx<T>* C, GenGeom g, LastAxis a,
                                                    const double* eig, int op, double inv_next,
                                                    const cplx<T>* tw, const double* spec,
                                                    int nbins)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int ntile = a.N0h / L;
  const bool nyq = (int)blockIdx.x >= ntile;
  int kx0, km0;
  if (!nyq) {
    kx0 = blockIdx.x * L;
    km0 = blockIdx.y;
  } else {
    kx0 = a.N0h;
    km0 = blockIdx.y * L;
    if (km0 >= a.Nmid)
      return;
  }
  const int64_t plane = (int64_t)a.Hp * a.Nmid;
  cplx<T>* cb = C + (int64_t)blockIdx.z * plane * a.nlast + (int64_t)km0 * a.Hp + kx0;
  constexpr uint32_t CB = sizeof(cplx<T>);
  const uint32_t pb = (uint32_t)plane * CB, lb = nyq ? (uint32_t)a.Hp * CB : CB;
  const int lmax = nyq ? a.Nmid - km0 : L; // live columns of the tile
  char* cbb = reinterpret_cast<char*>(cb);
  auto load = [&](int e, int l) -> cplx<T> {
    cplx<T> v = { 0, 0 };
    if (e < a.nlast && l < lmax)
      v = *reinterpret_cast<const cplx<T>*>(cbb + ((uint32_t)e * pb + (uint32_t)l * lb));
    return v;
  };
  auto store = [&](int j, int l, const cplx<T>& v) {
    if (j < a.nlast && l < lmax)
      *reinterpret_cast<cplx<T>*>(cbb + ((uint32_t)j * pb + (uint32_t)l * lb)) = v;
  };
  tfft_io<T, N, L, NT, 0, false, 1>(buf, tw, (int)threadIdx.x, load, TileLds{});
  using TL = TileLoop<L * N, NT>;
  {
    // the eigenvalues (or bin multipliers) of all entries a thread scales are
    // fetched before the first product, with a compile-time trip count, so
    // their loads are in flight together instead of one latency per entry
    double m[TL::IT];
#pragma unroll
    for (int it = 0; it < TL::IT; it++) {
      const int idx = TL::idx(it);
      const int l = idx % L, e = idx / L;
      const int kx = nyq ? kx0 : kx0 + l;
      const int km = nyq ? km0 + l : km0;
      m[it] = 0.;
      if (TL::live(it) && km < a.Nmid) {
        int k1, k2;
        line_k(g, kx, km, e, k1, k2);
        m[it] = op == PF_OP_SPECTRAL ? spectral_factor(g, kx, k1, k2, spec, nbins)
                                     : eig[eig_index(g, kx, k1, k2)];
      }
    }
#pragma unroll
    for (int it = 0; it < TL::IT; it++) {
      const int idx = TL::idx(it);
      const int l = idx % L;
      const int km = nyq ? km0 + l : km0;
      if (TL::live(it) && km < a.Nmid) {
        double f = m[it];
        if (op != PF_OP_SPECTRAL)
          f = op == PF_OP_MULTIPLY ? f
              : op == PF_OP_ROOT   ? sqrt(f)
              : op == PF_OP_INVERSE_FLOOR ? 1. / fmax(f, spec[0])
                                   : 1. / f;
        f *= inv_next;
        buf[idx] = cscale(buf[idx], (T)f);
      }
    }
  }
  __syncthreads();
  tfft_io<T, N, L, NT, 0, true, 1, HALF>(buf, tw, (int)threadIdx.x, TileLds{}, store);
}

} // namespace pfamd
