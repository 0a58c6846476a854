// pf_dctdst_fast.hpp -- compile-time specialised passes of the DCT/DST field
// backend (same algebra as pf_dctdst.hpp: three pruned c2r passes over the
// axes, each an M-point complex FFT of the packed sequence).
//
// M (half line length), L (lines per tile) and NT are template parameters:
// every tile loop has a compile-time trip count, so all global loads of a
// tile are in flight before the first wait, and the Stockham stages fold to
// constants. Tiles of the strided passes are L consecutive kx of one row; for
// 4-byte reals a tile row is only 4L bytes wide, so blocks run in an
// XCD-aware order that keeps the neighbouring tiles of a row (which share its
// cache lines) in one L2.
//
// LDS tile: L lines of M complex values, row pitch L with permuted rows
// (tix<L, -1>), no pad line, so a tile of 8 x 1024 FP32 lines is exactly
// 64 KB. The packed c2r reads only the real parts of X_0 and X_M (FFTW's c2r
// ignores their imaginary parts), so Re X_M is kept in the imaginary slot of
// row 0 and the tile needs no extra row for X_M.
#pragma once

#include "pf_dctdst.hpp"

namespace pfamd {

constexpr int DCT_PAD = -1; // tile layout tix<L, -1>: pitch L, permuted rows

// 16 bytes of consecutive reals at a 16-byte aligned address (rows of the
// work arrays start on 128-byte boundaries and interior tiles at multiples of
// L >= 4 columns): one vector access per lane instead of 16 / sizeof(T).
template <typename T>
struct Vec16
{
  static constexpr int V = 16 / (int)sizeof(T);
  T v[V];
};
template <typename T>
__device__ __forceinline__ Vec16<T> load16(const T* p)
{
  Vec16<T> r;
  __builtin_memcpy(&r, __builtin_assume_aligned(p, 16), 16);
  return r;
}
template <typename T>
__device__ __forceinline__ void store16(T* p, const Vec16<T>& r)
{
  __builtin_memcpy(__builtin_assume_aligned(p, 16), &r, 16);
}

// X_m of line l into the tile (m <= M): row m, except that X_0 and X_M
// share row 0 as (Re X_0, Re X_M).
template <typename T, int M, int L>
__device__ __forceinline__ void put_x(cplx<T>* buf, int m, int l, const cplx<T>& x)
{
  if (m == 0)
    buf[tix<L, DCT_PAD>(0, l)].x = x.x;
  else if (m == M)
    buf[tix<L, DCT_PAD>(0, l)].y = x.x;
  else
    buf[tix<L, DCT_PAD>(m, l)] = x;
}

// Pack X (layout of put_x) into Z in place and run the inverse FFT.
// Afterwards output x of line l is (x odd ? Im : Re) of row x/2.
template <typename T, int M, int L, int NT>
__device__ __forceinline__ void c2r_tile_ct(cplx<T>* buf, const cplx<T>* tw, int tid)
{
  using TP = TileLoop<L * (M / 2 + 1), NT>;
#pragma unroll
  for (int it = 0; it < TP::IT; it++) {
    if (!TP::live(it, tid))
      continue;
    const int idx = TP::idx(it, tid);
    const int l = idx % L, k = idx / L;
    cplx<T> xk, xm;
    if (k == 0) {
      const cplx<T> z = buf[tix<L, DCT_PAD>(0, l)];
      xk = { z.x, T(0) };
      xm = { z.y, T(0) };
    } else {
      xk = buf[tix<L, DCT_PAD>(k, l)];
      xm = buf[tix<L, DCT_PAD>(M - k, l)];
    }
    {
      const cplx<T> sm = { xk.x + xm.x, xk.y - xm.y };
      const cplx<T> d = { xk.x - xm.x, xk.y + xm.y };
      const cplx<T> td = cmul(cconj(tw[k]), d);
      buf[tix<L, DCT_PAD>(k, l)] = { sm.x - td.y, sm.y + td.x };
    }
    if (k != 0 && 2 * k != M) {
      const cplx<T> sm = { xm.x + xk.x, xm.y - xk.y };
      const cplx<T> d = { xm.x - xk.x, xm.y + xk.y };
      const cplx<T> td = cmul(cconj(tw[M - k]), d);
      buf[tix<L, DCT_PAD>(M - k, l)] = { sm.x - td.y, sm.y + td.x };
    }
  }
  __syncthreads();
  tfft<T, M, L, NT, DCT_PAD, true, 2>(buf, tw, tid);
}

// dct_factor(g, t, k0, k1, k2, 1) without branches: 2^{-d/2} 2^{border/2},
// or 0 when an odd axis of type t sits on its border.
template <typename T>
__device__ __forceinline__ T dct_border_factor(const DctGeom& g, int t, int k0, int k1, int k2)
{
  const int b0 = (k0 == 0) | (k0 == g.P0 - 1);
  const int b1 = (k1 == 0) | (k1 == g.P1 - 1);
  const int b2 = g.dim == 3 ? ((k2 == 0) | (k2 == g.P2 - 1)) : 0;
  const int zero = (b0 & t) | (b1 & (t >> 1)) | (b2 & (t >> 2));
  const int border = b0 + b1 + b2;
  T v = g.dim == 3 ? (T)0.35355339059327376220 : (T)0.5; // 1/sqrt(2^d)
  v = (border & 1) ? v * (T)1.41421356237309504880 : v;
  v = (border & 2) ? v * (T)2 : v;
  return (zero & 1) ? T(0) : v;
}

// Pass along the last axis (see kd_gen_last): noise of types (c, c + 2^{d-1})
// times the DCT/DST factors, c2r along the last axis, first n_last outputs ->
// G[c][j][k_mid][kx] (row pitch P0p).
// The noise is drawn without branches, one tile entry (m < M, l) per item:
// Re X_m from type c and Im X_m from type c + 2^{d-1}; the entry of row 0
// takes Re X_0 and Re X_M instead (the c2r uses neither Im X_0 nor Im X_M,
// and the tile keeps Re X_M in the imaginary slot of row 0). Each sample is
// one counter-based normal (philox_normal), scaled by its border factor and
// its root eigenvalue.
template <typename T, int M, int L, int NT, int NOISE>
__global__ void __launch_bounds__(NT, 4) kdf_gen_last(DctGeom g, const T* sq, const double* noise,
                                                   uint64_t seed, uint64_t first_real, T* G,
                                                   const cplx<T>* tw)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int Pm = g.dim == 3 ? g.P1 : 1;
  const int nlast = g.dim == 3 ? g.n2 : g.n1;
  int bx, by, b;
  xcd_block(bx, by, b);
  const bool edge = bx >= (g.P0 - 1) / L;
  int kx0, km0;
  if (!edge) {
    kx0 = bx * L;
    km0 = by;
  } else {
    kx0 = g.P0 - 1;
    km0 = by * L;
    if (km0 >= Pm)
      return;
  }
  const uint64_t real = first_real + b;
  const double* nz = noise ? noise + (int64_t)b * g.noise_stride : nullptr;
  const int ncomb = 1 << (g.dim - 1);
  const int oddlast = 1 << (g.dim - 1);
  const int64_t plane = (int64_t)Pm * g.P0p;
  T* out = G + (int64_t)b * ncomb * plane * nlast;
  // The thread's line l = tid % L is the same for all its items (NT is a
  // multiple of L); its items are the rows m = m_t + it * (NT / L). Lines of
  // an edge tile past the last row (km >= Pm) read row Pm - 1 and are zeroed.
  using TS = TileLoop<L * M, NT>;
  static_assert(NT % L == 0, "items of a thread share one line");
  constexpr int MS = NT / L;
  const int l = (int)threadIdx.x % L, m_t = (int)threadIdx.x / L;
  const int kx = edge ? kx0 : kx0 + l;
  const int km_raw = edge ? km0 + l : km0;
  const bool live = km_raw < Pm;
  const int km = live ? km_raw : Pm - 1;
  // sample index (reference storage order) and root-table offset of row m:
  // base + m * step along the last axis
  const int64_t s_base = (int64_t)kx * g.sx + (g.dim == 3 ? (int64_t)km * g.s1 : 0);
  const int64_t s_step = g.dim == 3 ? g.s2 : g.s1;
  const int64_t r_base = kx + (g.dim == 3 ? (int64_t)g.P0 * km : 0);
  const int64_t r_step = g.dim == 3 ? (int64_t)g.P0 * g.P1 : (int64_t)g.P0;
  const T sM = sq[r_base + M * r_step]; // root of element M (row 0's imaginary slot)
  // border factor of type t for a last-axis element inside (1) or on the
  // border (0 or M) of the axis; the other axes are fixed by the line
  auto fac = [&](int t, int mlast) {
    return dct_border_factor<T>(g, t, kx, g.dim == 3 ? km : mlast, g.dim == 3 ? mlast : 0);
  };
  auto normal = [&](int64_t st) -> T {
    if constexpr (NOISE == 0)
      return (T)nz[st];
    else
      return philox_normal1<T>(seed, real, (uint64_t)st);
  };
  // Philox blocks shared along a tile row: in an interior tile with
  // consecutive kx stored consecutively (sx = 1), the L lanes of one row
  // hold consecutive samples, four to a block. Lanes l < L/2 draw the blocks
  // of the row's real parts (b_re0 + l), lanes l >= L/2 those of its
  // imaginary parts, and every lane takes the words of its sample's block
  // from the lane that drew it (ds_bpermute). A row of L samples spans at
  // most L/2 blocks for L >= 4. FP32 only: FP64's Box-Muller (software
  // log, sincos) twice per item next to the shared words spills registers.
  static_assert(L >= 4, "a tile row spans at most L/2 Philox blocks");
  const bool share = NOISE == 1 && sizeof(T) == 4 && !edge && g.sx == 1;
  const int lane0 = ((int)threadIdx.x & 63) - l; // first lane of the thread's row in its wave
  auto from_lane = [&](uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int src,
                       int64_t st) -> T {
    const int a = src << 2;
    return box_muller_of_sample<T>((uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)c0),
                                   (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)c1),
                                   (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)c2),
                                   (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)c3),
                                   (uint64_t)st);
  };
  // outputs: block-uniform base, 32-bit element offsets (the launcher checks
  // nlast * plane < 2^32)
  const uint32_t lstep = edge ? (uint32_t)g.P0p : 1u;
#pragma unroll 1
  for (int c = 0; c < ncomb; c++) {
    const int t0 = c, t1 = c + oddlast;
    const int tid = opaque_tid();
    const T f0i = fac(t0, 1), f0b = fac(t0, 0), f1i = fac(t1, 1);
    // Re X_m (type c) and Im X_m (type c + 2^{d-1}) of the thread's rows; the
    // entry of row 0 takes Re X_0 and Re X_M
    int64_t s_re = (int64_t)t0 * g.Ptot + s_base + (int64_t)m_t * s_step;
    int64_t s_im = s_re + (int64_t)(t1 - t0) * g.Ptot;
    const int64_t s_xm = (int64_t)t0 * g.Ptot + s_base + (int64_t)M * s_step;
    const T* rp = sq + r_base + (int64_t)m_t * r_step;
    const int64_t ds = (int64_t)MS * s_step, dr = (int64_t)MS * r_step;
#pragma unroll 1
    for (int it = 0; it < TS::IT; it++) {
      if (TS::live(it, tid)) {
        const int m = m_t + it * MS;
        const bool row0 = m == 0;
        const T root_m = *rp;
        const int64_t s_i = row0 ? s_xm : s_im;
        T n_re, n_im;
        if (share) {
          const int64_t b_re0 = (s_re - l) >> 2, b_im0 = (s_i - l) >> 2;
          const int64_t blk = l < L / 2 ? b_re0 + l : b_im0 + (l - L / 2);
          uint32_t cw[4] = { (uint32_t)blk, (uint32_t)((uint64_t)blk >> 32), (uint32_t)real,
                             (uint32_t)(real >> 32) };
          philox4x32_10(cw, (uint32_t)seed, (uint32_t)(seed >> 32));
          n_re = from_lane(cw[0], cw[1], cw[2], cw[3], lane0 + (int)((s_re >> 2) - b_re0), s_re);
          n_im = from_lane(cw[0], cw[1], cw[2], cw[3], lane0 + L / 2 + (int)((s_i >> 2) - b_im0),
                           s_i);
        } else {
          n_re = normal(s_re);
          n_im = normal(s_i);
        }
        const T re = n_re * (row0 ? f0b : f0i) * root_m;
        const T im = n_im * (row0 ? f0b : f1i) * (row0 ? sM : root_m);
        const cplx<T> v = { live ? re : T(0), live ? (row0 ? im : -im) : T(0) };
        buf[tix<L, DCT_PAD>(m, l)] = v;
      }
      s_re += ds;
      s_im += ds;
      rp += dr;
    }
    __syncthreads();
    c2r_tile_ct<T, M, L, NT>(buf, tw, tid);
    T* oc = out + (int64_t)c * plane * nlast + (int64_t)km0 * g.P0p + kx0;
    using TO = TileLoop<L * M, NT>;
#pragma unroll
    for (int it = 0; it < TO::IT; it++) {
      const int idx = TO::idx(it, tid);
      const int l = idx % L, j = idx / L;
      const int km = edge ? km0 + l : km0;
      if (TO::live(it, tid) && j < nlast && km < Pm) {
        const cplx<T> v = buf[tix<L, DCT_PAD>(j >> 1, l)];
        oc[(uint32_t)j * (uint32_t)plane + (uint32_t)l * lstep] = (j & 1) ? v.y : v.x;
      }
    }
    __syncthreads();
  }
}

// 3D pass along y (see kd_mid): X_m = G[tx][z][m][kx] - i G[tx+2][z][m][kx],
// output Hh[tx][z][y][kx] (y < n1).
template <typename T, int M, int L, int NT>
__global__ void __launch_bounds__(NT, 4) kdf_mid(DctGeom g, const T* G, T* Hh, const cplx<T>* tw)
{
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  int bx, by, b;
  xcd_block(bx, by, b);
  const bool edge = bx >= (g.P0 - 1) / L;
  int kx0, z0;
  if (!edge) {
    kx0 = bx * L;
    z0 = by;
  } else {
    kx0 = g.P0 - 1;
    z0 = by * L;
    if (z0 >= g.n2)
      return;
  }
  const int64_t gplane = (int64_t)g.P1 * g.P0p;
  const int64_t gsize = gplane * g.n2;
  const int64_t hplane = (int64_t)g.n1 * g.P0p;
  const int64_t hsize = hplane * g.n2;
  const T* Gb = G + (int64_t)b * 4 * gsize;
  T* Hb = Hh + (int64_t)b * 2 * hsize;
  using TL = TileLoop<L * (M + 1), NT>;
#pragma unroll 1
  for (int tx = 0; tx < 2; tx++) {
    const int tid = opaque_tid();
    // block-uniform bases and 32-bit element offsets (a tile spans at most
    // L z-planes of one G array)
    const T* ge = Gb + (int64_t)tx * gsize + (int64_t)z0 * gplane + kx0;
    const T* go = ge + 2 * gsize;
    const uint32_t zstep = edge ? (uint32_t)gplane : 1u; // step of l
    if (!edge) {
      // interior tile: a row's L columns are contiguous, Vec16 per lane; two
      // batches of loads keep the staging registers within the budget
      constexpr int V = Vec16<T>::V;
      static_assert(L % V == 0, "whole vectors per tile row");
      using TV = TileLoop<(L / V) * (M + 1), NT>;
      constexpr int HV = (TV::IT + 1) / 2;
#pragma unroll
      for (int h = 0; h < TV::IT; h += HV) {
        Vec16<T> ve[HV], vo[HV];
#pragma unroll
        for (int i = 0; i < HV; i++) {
          const int it = h + i;
          const int idx = TV::idx(it, tid);
          const int l0 = (idx % (L / V)) * V, m = idx / (L / V);
          if (it < TV::IT && TV::live(it, tid) && z0 < g.n2) {
            const uint32_t o = (uint32_t)m * (uint32_t)g.P0p + (uint32_t)l0;
            ve[i] = load16(ge + o);
            vo[i] = load16(go + o);
          } else {
#pragma unroll
            for (int u = 0; u < V; u++) {
              ve[i].v[u] = 0;
              vo[i].v[u] = 0;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < HV; i++) {
          const int it = h + i;
          if (it >= TV::IT || !TV::live(it, tid))
            continue;
          const int idx = TV::idx(it, tid);
          const int l0 = (idx % (L / V)) * V, m = idx / (L / V);
#pragma unroll
          for (int u = 0; u < V; u++)
            put_x<T, M, L>(buf, m, l0 + u, { ve[i].v[u], -vo[i].v[u] });
        }
      }
    } else {
    // two batches of loads (all of a batch in flight at once) keep the
    // staging registers within a 4-waves-per-SIMD budget
    constexpr int HB = (TL::IT + 1) / 2;
#pragma unroll
    for (int h = 0; h < TL::IT; h += HB) {
      T ve[HB], vo[HB];
#pragma unroll
      for (int i = 0; i < HB; i++) {
        const int it = h + i;
        const int idx = TL::idx(it, tid);
        const int l = idx % L, m = idx / L;
        const int z = edge ? z0 + l : z0;
        ve[i] = 0;
        vo[i] = 0;
        if (it < TL::IT && TL::live(it, tid) && z < g.n2) {
          const uint32_t o = (uint32_t)m * (uint32_t)g.P0p + (uint32_t)l * zstep;
          ve[i] = ge[o];
          vo[i] = go[o];
        }
      }
#pragma unroll
      for (int i = 0; i < HB; i++) {
        const int it = h + i;
        if (it >= TL::IT || !TL::live(it, tid))
          continue;
        const int idx = TL::idx(it, tid);
        const int l = idx % L, m = idx / L;
        put_x<T, M, L>(buf, m, l, { ve[i], -vo[i] });
      }
    }
    }
    __syncthreads();
    c2r_tile_ct<T, M, L, NT>(buf, tw, tid);
    T* hb = Hb + (int64_t)tx * hsize + (int64_t)z0 * hplane + kx0;
    if (!edge) {
      constexpr int V = Vec16<T>::V;
      using TV = TileLoop<(L / V) * M, NT>;
#pragma unroll
      for (int it = 0; it < TV::IT; it++) {
        const int idx = TV::idx(it, tid);
        const int l0 = (idx % (L / V)) * V, y = idx / (L / V);
        if (TV::live(it, tid) && y < g.n1 && z0 < g.n2) {
          Vec16<T> w;
#pragma unroll
          for (int u = 0; u < V; u++) {
            const cplx<T> v = buf[tix<L, DCT_PAD>(y >> 1, l0 + u)];
            w.v[u] = (y & 1) ? v.y : v.x;
          }
          store16(hb + ((uint32_t)y * (uint32_t)g.P0p + (uint32_t)l0), w);
        }
      }
    } else {
      const uint32_t hstep = (uint32_t)hplane;
      using TO = TileLoop<L * M, NT>;
#pragma unroll
      for (int it = 0; it < TO::IT; it++) {
        const int idx = TO::idx(it, tid);
        const int l = idx % L, y = idx / L;
        const int z = z0 + l;
        if (TO::live(it, tid) && y < g.n1 && z < g.n2) {
          const cplx<T> v = buf[tix<L, DCT_PAD>(y >> 1, l)];
          hb[(uint32_t)y * (uint32_t)g.P0p + (uint32_t)l * hstep] = (y & 1) ? v.y : v.x;
        }
      }
    }
    __syncthreads();
  }
}

// Pass along x (see kd_x): rows r; X_m = Hh[0][r][m] - i Hh[1][r][m]; output
// field[r][x] (x < n0) with the value transform.
template <typename T, int M, int L, int NT>
__global__ void __launch_bounds__(NT) kdf_x(DctGeom g, const T* Hh, T* out, int64_t rows,
                                            const cplx<T>* tw, int transform, T param)
{
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int P0 = M + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  const int b = blockIdx.y;
  const int64_t hsize = rows * g.P0p;
  const T* h0 = Hh + (int64_t)b * 2 * hsize;
  const T* h1 = h0 + hsize;
  out += (int64_t)b * rows * g.n0;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  using TL = TileLoop<L * P0, NT>;
  T va[TL::IT], vb[TL::IT];
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    const int idx = TL::idx(it);
    const int l = idx / P0, m = idx % P0;
    const int64_t r = row0 + l;
    va[it] = 0;
    vb[it] = 0;
    if (TL::live(it) && r < rows) {
      va[it] = h0[r * g.P0p + m];
      vb[it] = h1[r * g.P0p + m];
    }
  }
#pragma unroll
  for (int it = 0; it < TL::IT; it++) {
    if (!TL::live(it))
      continue;
    const int idx = TL::idx(it);
    const int l = idx / P0, m = idx % P0;
    put_x<T, M, L>(buf, m, l, { va[it], -vb[it] });
  }
  __syncthreads();
  c2r_tile_ct<T, M, L, NT>(buf, tw, (int)threadIdx.x);
  using TO = TileLoop<L * M, NT>;
  // partial unroll: the value transform is a runtime switch
#pragma unroll 4
  for (int it = 0; it < TO::IT; it++) {
    const int idx = TO::idx(it);
    const int l = idx / M, x = idx % M;
    const int64_t r = row0 + l;
    if (TO::live(it) && x < g.n0 && r < rows) {
      const cplx<T> v = buf[tix<L, DCT_PAD>(x >> 1, l)];
      out[r * g.n0 + x] = value_transform<T>(transform, (x & 1) ? v.y : v.x, param);
    }
  }
}

} // namespace pfamd
