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
#pragma once

#include "pf_dctdst.hpp"

namespace pfamd {

// Pack X (buf rows m < M, xM = X_M) into Z in place and run the inverse FFT.
template <typename T, int M, int L, int NT>
__device__ __forceinline__ void c2r_tile_ct(cplx<T>* buf, const cplx<T>* xM, const cplx<T>* tw,
                                            int tid)
{
  constexpr int SE = L + 1;
  using TP = TileLoop<L * (M / 2 + 1), NT>;
#pragma unroll
  for (int it = 0; it < TP::IT; it++) {
    if (!TP::live(it, tid))
      continue;
    const int idx = TP::idx(it, tid);
    const int l = idx % L, k = idx / L;
    cplx<T> xk = buf[k * SE + l];
    cplx<T> xm = (k == 0) ? xM[l] : buf[(M - k) * SE + l];
    if (k == 0) {
      xk.y = 0;
      xm.y = 0;
    }
    {
      const cplx<T> sm = { xk.x + xm.x, xk.y - xm.y };
      const cplx<T> d = { xk.x - xm.x, xk.y + xm.y };
      const cplx<T> td = cmul(cconj(tw[k]), d);
      buf[k * SE + l] = { sm.x - td.y, sm.y + td.x };
    }
    if (k != 0 && 2 * k != M) {
      const cplx<T> sm = { xm.x + xk.x, xm.y - xk.y };
      const cplx<T> d = { xm.x - xk.x, xm.y + xk.y };
      const cplx<T> td = cmul(cconj(tw[M - k]), d);
      buf[(M - k) * SE + l] = { sm.x - td.y, sm.y + td.x };
    }
  }
  __syncthreads();
  tfft<T, M, L, NT, 1, true, 2>(buf, tw, tid);
}

// Pass along the last axis (see kd_gen_last): noise of types (c, c + 2^{d-1})
// times the DCT/DST factors, c2r along the last axis, first n_last outputs ->
// G[c][j][k_mid][kx] (row pitch P0p).
template <typename T, int M, int L, int NT, int NOISE>
__global__ void __launch_bounds__(NT, 4) kdf_gen_last(DctGeom g, const T* sq, const double* noise,
                                                   uint64_t seed, uint64_t first_real, T* G,
                                                   const cplx<T>* tw)
{
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + SE * M;
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
#pragma unroll 1
  for (int c = 0; c < ncomb; c++) {
    const int t0 = c, t1 = c + oddlast;
    const int tid = opaque_tid();
    if (NOISE == 1 && !edge && g.sx == 1) {
      // consecutive kx are consecutive samples: one Philox block (4 normals)
      // feeds up to 4 entries; NB blocks cover the L samples of a (type, m)
      // row whatever its alignment
      constexpr int NB = (L + 6) / 4;
      using TG = TileLoop<2 * NB * (M + 1), NT>;
#pragma unroll 2
      for (int it = 0; it < TG::IT; it++) {
        if (!TG::live(it, tid))
          continue;
        const int idx = TG::idx(it, tid);
        const int j = idx % NB, rest = idx / NB;
        const int half = rest & 1, m = rest >> 1;
        const int t = half ? t1 : t0;
        const int k1 = g.dim == 3 ? km0 : m, k2 = g.dim == 3 ? m : 0;
        const int64_t q0 = t * g.Ptot + kx0 + k1 * g.s1 + k2 * g.s2;
        const uint64_t blk = (uint64_t)(q0 >> 2) + j;
        const int64_t qb = (int64_t)(blk << 2);
        if (qb >= q0 + L)
          continue;
        const T* sqrow = sq + (int64_t)g.P0 * (k1 + (int64_t)g.P1 * k2);
        uint32_t cc[4] = { (uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)real,
                           (uint32_t)(real >> 32) };
        philox4x32_10(cc, (uint32_t)seed, (uint32_t)(seed >> 32));
        T nrm[4];
        BoxMuller<T>::pair(cc[0], cc[1], nrm[0], nrm[1]);
        BoxMuller<T>::pair(cc[2], cc[3], nrm[2], nrm[3]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int l = (int)(qb + i - q0);
          if (l < 0 || l >= L)
            continue;
          const int kx = kx0 + l;
          const T v = dct_factor<T>(g, t, kx, k1, k2, sqrow[kx]) * nrm[i];
          cplx<T>* dst = (m < M) ? &buf[m * SE + l] : &xM[l];
          if (half)
            dst->y = -v;
          else
            dst->x = v;
        }
      }
    } else {
      for (int idx = tid; idx < L * (M + 1); idx += NT) {
        const int l = idx % L, m = idx / L;
        const int kx = edge ? kx0 : kx0 + l;
        const int km = edge ? km0 + l : km0;
        cplx<T> x = { 0, 0 };
        if (km < Pm) {
          const int k1 = g.dim == 3 ? km : m, k2 = g.dim == 3 ? m : 0;
          const T sl = sq[kx + (int64_t)g.P0 * (k1 + (int64_t)g.P1 * k2)];
          const int64_t st = kx * g.sx + k1 * g.s1 + k2 * g.s2;
          const T f0 = dct_factor<T>(g, t0, kx, k1, k2, sl);
          const T f1 = dct_factor<T>(g, t1, kx, k1, k2, sl);
          T r0, r1;
          if (NOISE == 0) {
            r0 = (T)nz[t0 * g.Ptot + st];
            r1 = (T)nz[t1 * g.Ptot + st];
          } else {
            r0 = philox_normal<T>(seed, real, (uint64_t)(t0 * g.Ptot + st));
            r1 = philox_normal<T>(seed, real, (uint64_t)(t1 * g.Ptot + st));
          }
          x = { f0 * r0, -(f1 * r1) };
        }
        if (m < M)
          buf[m * SE + l] = x;
        else
          xM[l] = x;
      }
    }
    __syncthreads();
    c2r_tile_ct<T, M, L, NT>(buf, xM, tw, tid);
    using TO = TileLoop<L * M, NT>;
#pragma unroll
    for (int it = 0; it < TO::IT; it++) {
      const int idx = TO::idx(it, tid);
      const int l = idx % L, j = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int km = edge ? km0 + l : km0;
      if (TO::live(it, tid) && j < nlast && km < Pm) {
        const cplx<T> v = buf[(j >> 1) * SE + l];
        out[(int64_t)c * plane * nlast + (int64_t)j * plane + (int64_t)km * g.P0p + kx] =
          (j & 1) ? v.y : v.x;
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
  constexpr int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + SE * M;
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
        const cplx<T> x = { ve[i], -vo[i] };
        if (m < M)
          buf[m * SE + l] = x;
        else
          xM[l] = x;
      }
    }
    __syncthreads();
    c2r_tile_ct<T, M, L, NT>(buf, xM, tw, tid);
    T* hb = Hb + (int64_t)tx * hsize + (int64_t)z0 * hplane + kx0;
    const uint32_t hstep = edge ? (uint32_t)hplane : 1u;
    using TO = TileLoop<L * M, NT>;
#pragma unroll
    for (int it = 0; it < TO::IT; it++) {
      const int idx = TO::idx(it, tid);
      const int l = idx % L, y = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int z = edge ? z0 + l : z0;
      if (TO::live(it, tid) && y < g.n1 && z < g.n2) {
        const cplx<T> v = buf[(y >> 1) * SE + l];
        hb[(uint32_t)y * (uint32_t)g.P0p + (uint32_t)l * hstep] = (y & 1) ? v.y : v.x;
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
  constexpr int SE = L + 1, P0 = M + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + SE * M;
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
    const cplx<T> x = { va[it], -vb[it] };
    if (m < M)
      buf[m * SE + l] = x;
    else
      xM[l] = x;
  }
  __syncthreads();
  c2r_tile_ct<T, M, L, NT>(buf, xM, tw, (int)threadIdx.x);
  using TO = TileLoop<L * M, NT>;
  // partial unroll: the value transform is a runtime switch
#pragma unroll 4
  for (int it = 0; it < TO::IT; it++) {
    const int idx = TO::idx(it);
    const int l = idx / M, x = idx % M;
    const int64_t r = row0 + l;
    if (TO::live(it) && x < g.n0 && r < rows) {
      const cplx<T> v = buf[(x >> 1) * SE + l];
      out[r * g.n0 + x] = value_transform<T>(transform, (x & 1) ? v.y : v.x, param);
    }
  }
}

} // namespace pfamd
