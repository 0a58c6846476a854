// pf_dft.hpp -- the full complex field backend (DFTFieldBackend,
// dftfieldbackend.hh:19-477) and the 1-D transforms of the DFT matrix backend
// (dftmatrixbackend.hh:19-384), as tiles of lines held in LDS and transformed
// by the Stockham stages of pf_fft.hpp.
//
// Two pass shapes cover every case:
//  * column pass: lines along a strided axis; the L lines of a tile are L
//    consecutive columns (contiguous in memory), so every load and store moves
//    L adjacent complex values;
//  * row pass: lines along the contiguous axis, L rows per tile.
// A d-dimensional inverse transform is column passes over the strided axes
// followed by one row pass. An N-point 1-D transform (N = A B) is one column
// pass and one row pass (four-step): with k = kA + A kB and n = nB + B nA,
//   x[nB + B nA] = sum_kA e^{+2 pi i nA kA / A} e^{+2 pi i nB kA / N}
//                  sum_kB X[kA + A kB] e^{+2 pi i nB kB / B},
// i.e. B-point transforms along kB of the [kB][kA] array (column pass) times
// the twiddle e^{+2 pi i nB kA / N}, then A-point transforms of the rows nB,
// element nA of row nB landing at n = nB + B nA. The forward transform is the
// same with the sign of every exponent flipped and the roles of (nA, nB) and
// (kA, kB) exchanged.
#pragma once

#include "pf_fft.hpp"

namespace pfamd {

// Geometry of the DFT field: full extended grid, noise order and the
// eigenvalue table it reads.
struct DftGeom
{
  int dim, transposed;
  int N0, N1, N2;   // extended sizes
  int H;            // N0 / 2 + 1: x extent of the FP64 half table (DFT matrix)
  int P0, P1, P2;   // DCT half sizes (DCT matrix table)
  int dct;          // 1: the eigenvalue table is the DCT half array
  int A, B;         // 1-D four-step split N0 = A * B
  int64_t noise_stride; // normals per draw on the host (2 N_ext)
};

// Reference storage index of frequency (k0, k1, k2): the order in which
// generateField (matrix.hh:374-404) visits the field entries after
// transposeIfNeeded (dftfieldbackend.hh:140-147): the last two axes swapped
// when transposed.
__device__ __forceinline__ int64_t dft_storage_index(const DftGeom& g, int k0, int k1, int k2)
{
  if (g.dim == 1)
    return k0;
  if (g.dim == 2)
    return g.transposed ? k1 + (int64_t)g.N1 * k0 : k0 + (int64_t)g.N0 * k1;
  return g.transposed ? k0 + (int64_t)g.N0 * (k2 + (int64_t)g.N2 * k1)
                      : k0 + (int64_t)g.N0 * (k1 + (int64_t)g.N1 * k2);
}

// Index of frequency (k0, k1, k2) of the full grid in the eigenvalue table.
// DFT matrix: the table keeps the real parts of kx < H; Re X(k) = Re X(-k)
// for the transform of a real array, so kx >= H reads (-k0, -k1, -k2) mod N.
// DCT matrix: DCTMatrixBackend::eval folds k >= P to N - k per axis
// (dctmatrixbackend.hh:337-345).
__device__ __forceinline__ int64_t dft_eig_index(const DftGeom& g, int k0, int k1, int k2)
{
  if (g.dct) {
    const int a = k0 >= g.P0 ? g.N0 - k0 : k0;
    const int b = k1 >= g.P1 ? g.N1 - k1 : k1;
    const int c = k2 >= g.P2 ? g.N2 - k2 : k2;
    return a + (int64_t)g.P0 * (b + (int64_t)g.P1 * c);
  }
  if (k0 >= g.H) {
    k0 = g.N0 - k0;
    k1 = k1 == 0 ? 0 : g.N1 - k1;
    k2 = k2 == 0 ? 0 : g.N2 - k2;
  }
  return k0 + (int64_t)g.H * (k1 + (int64_t)g.N1 * k2);
}

// DFTFieldBackend::set (dftfieldbackend.hh:236-240): lambda (rand1 + i rand2)
// with the two normals of the entry's storage index s (samples 2s, 2s+1).
template <typename T, int NOISE>
__device__ __forceinline__ cplx<T> dft_entry(const DftGeom& g, const T* sq, int k0, int k1, int k2,
                                             const double* noise, uint64_t seed, uint64_t draw)
{
  const int64_t s = dft_storage_index(g, k0, k1, k2);
  const T lambda = sq[dft_eig_index(g, k0, k1, k2)];
  T r1, r2;
  if (NOISE == 0) {
    r1 = (T)noise[2 * s];
    r2 = (T)noise[2 * s + 1];
  } else {
    philox_pair<T>(seed, draw, (uint64_t)s, r1, r2);
  }
  return { lambda * r1, lambda * r2 };
}

// Strides of a column pass: column c in [0, C), element e in [0, E) (the
// transform axis), slab o in [0, O); grid.y runs over the draws / vectors.
struct ColPass
{
  int C, E, O;
  int keep;                 // outputs kept per line (the first `keep`)
  int64_t in_es, in_os, in_bs;
  int64_t out_es, out_os, out_bs;
  int64_t twN;              // 0: none; else output j of column c times e^{+-2 pi i j c / twN}
};

// e^{sign 2 pi i m / n} in FP64 (m reduced mod n by the caller)
template <typename T>
__device__ __forceinline__ cplx<T> unit_root(int64_t m, int64_t n, bool inv)
{
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  return { (T)c, (T)(inv ? s : -s) };
}

template <typename T, bool INV>
__device__ __forceinline__ void col_store(const ColPass& p, cplx<T>* out, const cplx<T>* buf,
                                          int L, int64_t line0, int64_t nlines)
{
  for (int idx = threadIdx.x; idx < L * p.keep; idx += blockDim.x) {
    const int l = idx % L, j = idx / L;
    const int64_t u = line0 + l;
    if (u >= nlines)
      continue;
    const int64_t c = u % p.C, o = u / p.C;
    cplx<T> v = buf[j * L + l];
    if (p.twN)
      v = cmul(v, unit_root<T>(((int64_t)j * c) % p.twN, p.twN, INV));
    out[o * p.out_os + j * p.out_es + c] = v;
  }
}

// Column pass on an array: lines (c, o) along e, transformed (INV: e^{+i}),
// first `keep` outputs (times the four-step twiddle) written out.
template <typename T, int MAXE, bool INV>
__global__ void kdft_cols(const cplx<T>* in, cplx<T>* out, ColPass p, Sched s, const cplx<T>* tw,
                          int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  in += blockIdx.y * p.in_bs;
  out += blockIdx.y * p.out_bs;
  const int64_t nlines = (int64_t)p.C * p.O;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> v = { 0, 0 };
    if (u < nlines) {
      const int64_t c = u % p.C, o = u / p.C;
      v = in[o * p.in_os + e * p.in_es + c];
    }
    buf[e * L + l] = v;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, INV>(buf, twl, 1, s, L, L, 1);
  col_store<T, INV>(p, out, buf, L, line0, nlines);
}

// First (generating) pass of the DFT field: the spectrum entries of a tile
// are drawn in registers (host noise or Philox, draw first_draw + blockIdx.y)
// and transformed backward along the last axis. 3D: columns kx, slabs ky,
// elements kz; 2D: columns kx, elements ky; 1D: columns kA, elements kB
// (k = kA + A kB) with the four-step twiddle.
template <typename T, int MAXE, int NOISE>
__global__ void kdft_gen(DftGeom g, const T* sq, const double* noise, uint64_t seed,
                         uint64_t first_draw, cplx<T>* out, ColPass p, Sched s, const cplx<T>* tw,
                         int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  const double* nz = noise ? noise + b * g.noise_stride : nullptr;
  const uint64_t draw = first_draw + b;
  out += b * p.out_bs;
  const int64_t nlines = (int64_t)p.C * p.O;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> z = { 0, 0 };
    if (u < nlines) {
      const int c = (int)(u % p.C), o = (int)(u / p.C);
      if (g.dim == 3)
        z = dft_entry<T, NOISE>(g, sq, c, o, e, nz, seed, draw);
      else if (g.dim == 2)
        z = dft_entry<T, NOISE>(g, sq, c, e, 0, nz, seed, draw);
      else
        z = dft_entry<T, NOISE>(g, sq, c + g.A * e, 0, 0, nz, seed, draw);
    }
    buf[e * L + l] = z;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, true>(buf, twl, 1, s, L, L, 1);
  col_store<T, true>(p, out, buf, L, line0, nlines);
}

// Row pass: R rows of E contiguous values (pitch in_rs), transformed; output
// j of row r goes to r * out_rs + j * out_es for j < keep, where that index is
// below `limit`. OUT = 0: complex output `outc`; OUT = 1: real output, the
// real part to field fre and the imaginary part to field fre + 1 of `outf`
// (draw b: fre = f_first + 2 b; fields outside [0, nfields) are skipped),
// after the value transform; OUT = 2: real part only, to vector b of `outf`.
struct RowPass
{
  int64_t R;
  int keep;
  int64_t in_rs, in_bs;
  int64_t out_rs, out_es, out_bs, limit;
};

template <typename T, typename TO, int MAXE, bool INV, int OUT>
__global__ void kdft_rows(const cplx<T>* in, cplx<T>* outc, TO* outf, RowPass p, Sched s,
                          const cplx<T>* tw, int L, int64_t f_first, int64_t nfields,
                          int transform, T param)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  in += b * p.in_bs;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx / N, e = idx % N;
    const int64_t r = row0 + l;
    buf[l * N + e] = r < p.R ? in[r * p.in_rs + e] : cplx<T>{ 0, 0 };
  }
  __syncthreads();
  fft_lds<T, MAXE, true, INV>(buf, twl, 1, s, L, 1, N);
  TO* fre = nullptr;
  TO* fim = nullptr;
  if (OUT == 1) {
    const int64_t f = f_first + 2 * (int64_t)b;
    fre = (f >= 0 && f < nfields) ? outf + f * p.out_bs : nullptr;
    fim = (f + 1 >= 0 && f + 1 < nfields) ? outf + (f + 1) * p.out_bs : nullptr;
  } else if (OUT == 2) {
    fre = outf + b * p.out_bs;
  }
  if (OUT == 0)
    outc += b * p.out_bs;
  // contiguous outputs along j (rows of the 2D/3D field) or along the rows
  // (1-D four-step, n = nB + B nA): the fast loop index follows memory
  const bool jfast = p.out_es == 1;
  for (int idx = threadIdx.x; idx < L * p.keep; idx += blockDim.x) {
    const int l = jfast ? idx / p.keep : idx % L;
    const int j = jfast ? idx % p.keep : idx / L;
    const int64_t r = row0 + l;
    if (r >= p.R)
      continue;
    const int64_t o = r * p.out_rs + (int64_t)j * p.out_es;
    if (o >= p.limit)
      continue;
    const cplx<T> v = buf[l * N + j];
    if (OUT == 0) {
      outc[o] = v;
    } else {
      if (fre)
        fre[o] = (TO)value_transform<T>(transform, v.x, param);
      if (OUT == 1 && fim)
        fim[o] = (TO)value_transform<T>(transform, v.y, param);
    }
  }
}

// Zero-padded complex copy of `batch` real vectors of n values into arrays of
// N entries (pitch N): the input of a forward transform.
template <typename T, typename TI>
__global__ void kdft_pad(const TI* in, cplx<T>* out, int64_t n, int64_t N, int batch)
{
  const int64_t tot = N * batch;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / N, k = i % N;
    out[i] = { k < n ? (T)in[b * n + k] : T(0), T(0) };
  }
}

// 1-D product step in natural frequency order: X[k] *= f(lambda_k) / N with
// f = lambda (multiply), sqrt(lambda) (root), 1 / lambda (inverse of the
// embedding) or 1 / max(lambda, floor) (op 4, the floored inverse of the
// preconditioner); lambda from the FP64 half table (k >= H reads N - k).
template <typename T>
__global__ void kdft_mult_1d(cplx<T>* X, int64_t N, int H, const double* eig, int op,
                             double scale, int batch, double floor)
{
  const int64_t tot = N * batch;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i % N;
    const double lam = eig[k < H ? k : N - k];
    double f = op == 0 ? lam : op == 1 ? sqrt(lam) : op == 4 ? 1. / fmax(lam, floor) : 1. / lam;
    f *= scale;
    X[i] = { (T)((double)X[i].x * f), (T)((double)X[i].y * f) };
  }
}

} // namespace pfamd

