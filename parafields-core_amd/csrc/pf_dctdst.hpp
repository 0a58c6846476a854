// pf_dctdst.hpp -- DCT/DST field backend (DCTDSTFieldBackend,
// dctdstfieldbackend.hh) on the device.
//
// The reference sums 2^d parity types; type t (bit i = odd along axis i)
// holds a_t(k) = 2^{-d/2} sqrt(lambda_k) xi 2^{border/2} (0 on the boundary
// of an odd axis, matrix.hh:342-372 with set, :418-438) and is transformed by
// REDFT00 along even axes and RODFT00 along odd ones. Along one axis
//   REDFT00(a)(x) + RODFT00-with-zero-border(b)(x) = c2r_N(a - i b)(x)
// (c2r with FFTW's semantics, N = 2(P-1)), so the 2^d-term sum factors into
// three pruned passes over the axes:
//   pass last: X_m = a_{t}(..,m) - i a_{t + odd_last}(..,m)  -> 2^{d-1} real arrays
//   pass mid : X_m = G_c(..,m,..) - i G_{c + odd_mid}(..,m,..) -> 2 real arrays (3D)
//   pass x   : X_m = H_0(m,..) - i H_1(m,..)                   -> the field
// Each c2r of length N = 2M runs as an M-point complex inverse FFT of the
// packed sequence Z_k = (X_k + conj X_{M-k}) + i e^{2 pi i k/N} (X_k - conj X_{M-k}).
#pragma once

#include "pf_fast.hpp"

namespace pfamd {

struct DctGeom
{
  int dim;
  int P0, P1, P2;     // N_i/2 + 1
  int N0, N1, N2;
  int n0, n1, n2;
  int64_t Ptot;       // P0*P1*P2
  int64_t sx, s1, s2; // reference storage index of (kx, k1, k2): kx*sx + k1*s1 + k2*s2
  int64_t noise_stride;
  int P0p;            // row pitch (reals) of the G / H work arrays: P0 rounded up to 128 bytes
};

// a_t at logical (k0, k1, k2), without the noise factor: returns the factor
// scale * sqrt(lambda) * 2^{border/2} or 0 (DCTDSTFieldBackend::set).
template <typename T>
__device__ __forceinline__ T dct_factor(const DctGeom& g, int t, int k0, int k1, int k2, T sqlam)
{
  int border = 0;
  const int kk[3] = { k0, k1, k2 };
  const int PP[3] = { g.P0, g.P1, g.P2 };
  for (int i = 0; i < g.dim; i++) {
    if (kk[i] == 0 || kk[i] == PP[i] - 1) {
      if ((t >> i) & 1)
        return T(0);
      border++;
    }
  }
  const T scale = g.dim == 3 ? (T)0.35355339059327376220 : (T)0.5; // 1/sqrt(2^d)
  T v = scale * sqlam;
  if (border == 1)
    v *= (T)1.41421356237309504880;
  else if (border == 2)
    v *= (T)2.0;
  else if (border == 3)
    v *= (T)2.82842712474619009760;
  return v;
}

// In-tile c2r: X_m in buf[m*SE + l] (m < M) and X_M in xM[l]; packs Z in
// place and runs the M-point inverse FFT. Afterwards output x of line l is
// (x odd ? Im : Re) of buf[(x/2)*SE + l].
template <typename T, int MF, int LF, int NTF>
__device__ __forceinline__ void c2r_tile(cplx<T>* buf, const cplx<T>* xM, const cplx<T>* tw, int Mr,
                                         int Lr, const Sched& s)
{
  const int M = MF ? MF : Mr;
  const int L = LF ? LF : Lr;
  const int NT = NTF ? NTF : (int)blockDim.x;
  const int SE = L + 1;
  const int npairs = M / 2 + 1;
  for (int idx = threadIdx.x; idx < L * npairs; idx += NT) {
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
  if constexpr (MF > 0)
    tfft<T, MF, LF, NTF, 1, true, 2>(buf, tw);
  else
    fft_lds<T, sizeof(T) == 4 ? 32 : 16, false, true>(buf, tw, 2, s, L, SE, 1);
}

// Pass along the last axis: lines (kx, k_mid) of the P-grid, X from the noise
// of types (t, t + 2^{d-1}); one combination c of the other parities at a
// time. Output G[c][j][k_mid][kx] (real, j < n_last).
// Tiles: L consecutive kx of one k_mid row; the last column kx = P0-1 is a
// tile of L consecutive rows.
template <typename T, int MF, int LF, int NTF, int NOISE>
__global__ void kd_gen_last(DctGeom g, const T* sq, const double* noise, uint64_t seed,
                            uint64_t first_real, T* G, const cplx<T>* tw, Sched s, int Mr, int Lr)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int M = MF ? MF : Mr;
  const int L = LF ? LF : Lr;
  const int NT = NTF ? NTF : (int)blockDim.x;
  const int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + (int64_t)SE * M;
  const int Pm = g.dim == 3 ? g.P1 : 1;       // rows of lines
  const int nlast = g.dim == 3 ? g.n2 : g.n1;
  const int ntile = (g.P0 - 1) / L;
  const bool edge = (int)blockIdx.x >= ntile;
  int kx0, km0;
  if (!edge) {
    kx0 = blockIdx.x * L;
    km0 = blockIdx.y;
  } else {
    kx0 = g.P0 - 1;
    km0 = blockIdx.y * L;
    if (km0 >= Pm)
      return;
  }
  const int b = blockIdx.z;
  const uint64_t real = first_real + b;
  const double* nz = noise ? noise + (int64_t)b * g.noise_stride : nullptr;
  const int ncomb = 1 << (g.dim - 1);
  const int oddlast = 1 << (g.dim - 1);
  const int64_t plane = (int64_t)Pm * g.P0p;
  T* out = G + (int64_t)b * ncomb * plane * nlast;
  for (int c = 0; c < ncomb; c++) {
    const int t0 = c, t1 = c + oddlast;
    if (NOISE == 1 && !edge && g.sx == 1) {
      // consecutive kx are consecutive samples: one Philox block (4 normals)
      // feeds up to 4 entries of the tile; NB blocks cover the L samples of a
      // (type, m) row whatever its alignment: samples q0 .. q0+L-1 touch at
      // most ceil((L+3)/4) = (L+6)/4 blocks
      const int NB = (L + 6) / 4;
      for (int idx = threadIdx.x; idx < 2 * NB * (M + 1); idx += NT) {
        const int j = idx % NB, rest = idx / NB;
        const int half = rest & 1, m = rest >> 1;
        const int t = half ? t1 : t0;
        const int k1 = g.dim == 3 ? km0 : m, k2 = g.dim == 3 ? m : 0;
        const int64_t q0 = t * g.Ptot + kx0 + k1 * g.s1 + k2 * g.s2;
        const uint64_t blk = (uint64_t)(q0 >> 2) + j;
        const int64_t qb = (int64_t)(blk << 2);
        if (qb >= q0 + L)
          continue;
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
          const T sl = sq[kx + (int64_t)g.P0 * (k1 + (int64_t)g.P1 * k2)];
          const T v = dct_factor<T>(g, t, kx, k1, k2, sl) * nrm[i];
          cplx<T>* dst = (m < M) ? &buf[m * SE + l] : &xM[l];
          if (half)
            dst->y = -v;
          else
            dst->x = v;
        }
      }
      __syncthreads();
    } else
    for (int idx = threadIdx.x; idx < L * (M + 1); idx += NT) {
      const int l = idx % L, m = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int km = edge ? km0 + l : km0;
      cplx<T> x = { 0, 0 };
      if (km < Pm) {
        const int k1 = g.dim == 3 ? km : m, k2 = g.dim == 3 ? m : 0;
        const int64_t ei = kx + (int64_t)g.P0 * (k1 + (int64_t)g.P1 * k2);
        const T sl = sq[ei];
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
    __syncthreads();
    c2r_tile<T, MF, LF, NTF>(buf, xM, tw, M, L, s);
    for (int idx = threadIdx.x; idx < L * nlast; idx += NT) {
      const int l = idx % L, j = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int km = edge ? km0 + l : km0;
      if (km < Pm) {
        const cplx<T> v = buf[(j >> 1) * SE + l];
        out[(int64_t)c * plane * nlast + (int64_t)j * plane + (int64_t)km * g.P0p + kx] =
          (j & 1) ? v.y : v.x;
      }
    }
    __syncthreads();
  }
}

// 3D pass along y: lines (kx, z); X_m = G[tx][z][m][kx] - i G[tx+2][z][m][kx];
// output Hh[tx][z][y][kx] (y < n1).
template <typename T, int MF, int LF, int NTF>
__global__ void kd_mid(DctGeom g, const T* G, T* Hh, const cplx<T>* tw, Sched s, int Mr, int Lr)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int M = MF ? MF : Mr;
  const int L = LF ? LF : Lr;
  const int NT = NTF ? NTF : (int)blockDim.x;
  const int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + (int64_t)SE * M;
  const int ntile = (g.P0 - 1) / L;
  const bool edge = (int)blockIdx.x >= ntile;
  int kx0, z0;
  if (!edge) {
    kx0 = blockIdx.x * L;
    z0 = blockIdx.y;
  } else {
    kx0 = g.P0 - 1;
    z0 = blockIdx.y * L;
    if (z0 >= g.n2)
      return;
  }
  const int b = blockIdx.z;
  const int64_t gplane = (int64_t)g.P1 * g.P0p;     // one z slice of G
  const int64_t gsize = gplane * g.n2;               // one G array
  const int64_t hplane = (int64_t)g.n1 * g.P0p;
  const int64_t hsize = hplane * g.n2;
  const T* Gb = G + (int64_t)b * 4 * gsize;
  T* Hb = Hh + (int64_t)b * 2 * hsize;
  for (int tx = 0; tx < 2; tx++) {
    const T* ge = Gb + (int64_t)tx * gsize;
    const T* go = Gb + (int64_t)(tx + 2) * gsize;
    for (int idx = threadIdx.x; idx < L * (M + 1); idx += NT) {
      const int l = idx % L, m = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int z = edge ? z0 + l : z0;
      cplx<T> x = { 0, 0 };
      if (z < g.n2) {
        const int64_t o = (int64_t)z * gplane + (int64_t)m * g.P0p + kx;
        x = { ge[o], -go[o] };
      }
      if (m < M)
        buf[m * SE + l] = x;
      else
        xM[l] = x;
    }
    __syncthreads();
    c2r_tile<T, MF, LF, NTF>(buf, xM, tw, M, L, s);
    for (int idx = threadIdx.x; idx < L * g.n1; idx += NT) {
      const int l = idx % L, y = idx / L;
      const int kx = edge ? kx0 : kx0 + l;
      const int z = edge ? z0 + l : z0;
      if (z < g.n2) {
        const cplx<T> v = buf[(y >> 1) * SE + l];
        Hb[(int64_t)tx * hsize + (int64_t)z * hplane + (int64_t)y * g.P0p + kx] =
          (y & 1) ? v.y : v.x;
      }
    }
    __syncthreads();
  }
}

// Pass along x: rows r (= y + n1 z); X_m = Hh[0][r][m] - i Hh[1][r][m]
// (m < P0); output field[r][x] (x < n0) with the value transform.
template <typename T, int MF, int LF, int NTF>
__global__ void kd_x(DctGeom g, const T* Hh, T* out, int64_t rows, const cplx<T>* tw, Sched s,
                     int Mr, int Lr, int transform, T param)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int M = MF ? MF : Mr;
  const int L = LF ? LF : Lr;
  const int NT = NTF ? NTF : (int)blockDim.x;
  const int SE = L + 1;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* xM = buf + (int64_t)SE * M;
  const int P0 = M + 1;
  const int b = blockIdx.y;
  const int64_t hsize = rows * g.P0p;
  const T* h0 = Hh + (int64_t)b * 2 * hsize;
  const T* h1 = h0 + hsize;
  out += (int64_t)b * rows * g.n0;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * P0; idx += NT) {
    const int l = idx / P0, m = idx % P0;
    const int64_t r = row0 + l;
    cplx<T> x = { 0, 0 };
    if (r < rows)
      x = { h0[r * g.P0p + m], -h1[r * g.P0p + m] };
    if (m < M)
      buf[m * SE + l] = x;
    else
      xM[l] = x;
  }
  __syncthreads();
  c2r_tile<T, MF, LF, NTF>(buf, xM, tw, M, L, s);
  for (int idx = threadIdx.x; idx < L * g.n0; idx += NT) {
    const int l = idx / g.n0, x = idx % g.n0;
    const int64_t r = row0 + l;
    if (r < rows) {
      const cplx<T> v = buf[(x >> 1) * SE + l];
      out[r * g.n0 + x] = value_transform<T>(transform, (x & 1) ? v.y : v.x, param);
    }
  }
}

} // namespace pfamd
