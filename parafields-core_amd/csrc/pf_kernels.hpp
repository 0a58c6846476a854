// pf_kernels.hpp -- the HIP kernels of the engine (gfx950).
//
// Layout in HBM (x fastest everywhere, realization b outermost):
//   eigenvalues   FP64, DCT backend [P2][P1][P0] (P_i = N_i/2+1),
//                 R2C backend [N2][N1][H] (H = N0/2+1, real part)
//   sqrt table    T, same layout as the eigenvalues
//   spectrum B1   complex T [n_last][N_mid][H]     (after the last-axis pass)
//   spectrum B2   complex T [n2][n1][H]            (3D, after the y pass)
//   field         T [n2][n1][n0]
#pragma once

#include "pf_common.hpp"
#include "pf_fft.hpp"

namespace pfamd {

// ---------------------------------------------------------------------------
// Eigenvalues
// ---------------------------------------------------------------------------

// Visits every (row, i) of an array of nrows rows of n entries with the
// threads of all blocks: a block takes one row at a time when rows are at
// least blockDim.x long, else blockDim.x / n whole rows. No 64-bit division
// per entry; the row index is split by the caller once per row.
template <typename F>
__device__ __forceinline__ void for_rows(int64_t nrows, int n, const F& f)
{
  const int bd = blockDim.x, t = threadIdx.x;
  const bool wide = n >= bd;
  const int rpb = wide ? 1 : bd / n;        // rows per block and sweep
  const int rsub = wide ? 0 : t / n;        // this thread's row within them
  const int i0 = wide ? t : t % n, step = wide ? bd : n;
  if (rsub >= rpb)
    return;
  for (int64_t r0 = (int64_t)blockIdx.x * rpb; r0 < nrows; r0 += (int64_t)gridDim.x * rpb) {
    const int64_t row = r0 + rsub;
    if (row >= nrows)
      break;
    for (int i = i0; i < n; i += step)
      f(row, i);
  }
}

// blocks for for_rows over nrows rows of n entries (256 threads)
inline unsigned row_blocks(int64_t nrows, int n)
{
  const int64_t rpb = n >= 256 ? 1 : 256 / n;
  const int64_t b = (nrows + rpb - 1) / rpb;
  return (unsigned)(b < (1 << 20) ? b : (1 << 20));
}

// Covariance on the DCT half grid [P2][P1][P0] (dctmatrixbackend.hh:164-170 via
// matrix.hh:615-638).
// COV: the covariance family at compile time (-1: p.covariance at run time);
// one family per instantiation keeps the other families' code (and
// registers) out of the kernel.
template <int COV, int PER>
__global__ void k_eval_cov_dct(CovParams p, int P0, int P1, int P2, double* out)
{
  for_rows((int64_t)P1 * P2, P0, [&](int64_t row, int i) {
    const int64_t idx[3] = { i, (int64_t)((uint32_t)row % (uint32_t)P1),
                             (int64_t)((uint32_t)row / (uint32_t)P1) };
    out[row * P0 + i] = embedded_cov<COV, PER>(p, idx);
  });
}

// Covariance on the full real grid [N2][N1][N0] (r2cmatrixbackend.hh:150-170;
// the r2c padding entries are never read by the transform, so they are not
// stored).
template <int COV, int PER>
__global__ void k_eval_cov_full(CovParams p, int N0, int N1, int N2, double* out)
{
  for_rows((int64_t)N1 * N2, N0, [&](int64_t row, int i) {
    const int64_t idx[3] = { i, (int64_t)((uint32_t)row % (uint32_t)N1),
                             (int64_t)((uint32_t)row / (uint32_t)N1) };
    out[row * N0 + i] = embedded_cov<COV, PER>(p, idx);
  });
}

// Even part of the 2D embedded covariance, e(i, j) = (c(i, j) + c(-i, -j)) / 2
// with indices mod (N0, N1) (c as in k_eval_cov_full). The real part of the
// embedding's DFT is the DFT of e (the odd part transforms to an imaginary
// array), and e is point-symmetric, so its rows j = 0 .. N1/2 determine it.
// Classical embedding: c(-i, -j) is c(i, j) up to the rounding of the wrapped
// coordinates, except on the Nyquist column 2i = N0, where -i wraps onto i
// and the mirror is c(i, -j), another value under a rotated anisotropy; so
// only that column is averaged. Rows 0 and N1/2 need no mirror: the y pass
// keeps the real part of their x spectra, the DFT of their even part in x.
// Merge and fold embeddings average every entry.
template <int COV, int PER>
__device__ __forceinline__ double cov_even_2d(const CovParams& p, int N0, int N1, int i, int j)
{
  const int64_t a[3] = { i, j, 0 };
  const double v = embedded_cov<COV, PER>(p, a);
  if (PER == 0 && !(2 * i == N0 && j != 0 && 2 * j != N1))
    return v;
  const int64_t b[3] = { i ? N0 - i : 0, j ? N1 - j : 0, 0 };
  return 0.5 * (v + embedded_cov<COV, PER>(p, b));
}

// e on rows 0 .. N1/2 of the 2D embedding: out[j][i], N1/2 + 1 rows of N0.
template <int COV, int PER>
__global__ void k_eval_cov_even2d(CovParams p, int N0, int N1, double* out)
{
  p.dim = 2;
  for_rows(N1 / 2 + 1, N0, [&](int64_t row, int i) {
    out[row * N0 + i] = cov_even_2d<COV, PER>(p, N0, N1, i, (int)row);
  });
}

// Covariance on the 1-D embedding grid (one long row: grid-stride over x).
template <int COV, int PER>
__global__ void k_eval_cov_1d(CovParams p, int64_t N0, double* out)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N0;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t idx[3] = { i, 0, 0 };
    out[i] = embedded_cov<COV, PER>(p, idx);
  }
}

// Cofold (matrix.hh:905-1013) on a dense grid a[n2][n1][n0] of classical
// entries (the DCT half grid or the full R2C grid): max-norm of the wrapped
// coordinates in units of the extensions, and whether the entry lies in the
// border band of the extended domain along some axis.
__device__ inline double cofold_norm(const CovParams& p, const int64_t idx[3], bool& border)
{
  double norm = 0.;
  border = false;
  for (int i = 0; i < p.dim; i++) {
    double c = (double)idx[i] * p.h[i];
    if (c > p.wrap_half[i])
      c -= p.wrap_len[i];
    if (fabs(c) / p.ext[i] > norm)
      norm = fabs(c) / p.ext[i];
    if (p.wrap_half[i] - fabs(c) < p.border[i])
      border = true;
  }
  return norm;
}

// Pass 1 (matrix.hh:913-943): per block, the minimum over entries of norm
// <= 1 and the maximum over border entries of norm > 1; part[2b], part[2b+1].
__global__ void __launch_bounds__(256) k_cofold_bounds(CovParams p, const double* a, int n0,
                                                       int n1, int n2, double* part)
{
  __shared__ double smin[256], smax[256];
  double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
  const int64_t total = (int64_t)n0 * n1 * n2;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * 256) {
    const int64_t idx[3] = { k % n0, (k / n0) % n1, k / ((int64_t)n0 * n1) };
    bool border;
    const double norm = cofold_norm(p, idx, border);
    const double v = a[k];
    if (norm <= 1.) {
      if (v < mn)
        mn = v;
    } else if (border && v > mx)
      mx = v;
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      if (smin[threadIdx.x + s] < smin[threadIdx.x])
        smin[threadIdx.x] = smin[threadIdx.x + s];
      if (smax[threadIdx.x + s] > smax[threadIdx.x])
        smax[threadIdx.x] = smax[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// Pass 2 (matrix.hh:968-1006): entries of norm > 1 below alpha take the
// offset, entries in [alpha, beta) the offset plus the taper's integral from
// alpha, the others stay.
__global__ void __launch_bounds__(256) k_cofold_apply(CovParams p, double* a, int n0, int n1,
                                                      int n2, double alpha, double beta,
                                                      double offset)
{
  const int64_t total = (int64_t)n0 * n1 * n2;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * 256) {
    const int64_t idx[3] = { k % n0, (k / n0) % n1, k / ((int64_t)n0 * n1) };
    bool border;
    if (!(cofold_norm(p, idx, border) > 1.))
      continue;
    const double x = a[k];
    if (x < alpha)
      a[k] = offset;
    else if (x < beta)
      a[k] = cofold_integral(p, alpha, beta, x) + offset;
  }
}

// REDFT00 (FFTW's DCT-I) along one axis of a real FP64 array, computed as the
// DFT of the even extension of length Nfft = 2(P-1); in place. Lines of the
// axis: stride `stride`, P points. LM = true for the contiguous axis.
template <int MAXE, bool LM>
__global__ void k_redft00_axis(double* a, Sched s, const cplx<double>* tw, int P, int64_t stride,
                               int64_t nlines, int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<double>* twl = reinterpret_cast<cplx<double>*>(smem);
  cplx<double>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int64_t line0 = (int64_t)blockIdx.x * L;
  const int se = LM ? 1 : L, sl = LM ? N : 1;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    int l, e;
    if (LM) {
      l = idx / N;
      e = idx % N;
    } else {
      l = idx % L;
      e = idx / L;
    }
    const int64_t u = line0 + l;
    double v = 0.;
    if (u < nlines) {
      const int64_t inner = u % stride, outer = u / stride;
      const int64_t base = inner + outer * stride * P;
      const int src = e < P ? e : N - e;
      v = a[base + src * stride];
    }
    buf[e * se + l * sl] = { v, 0. };
  }
  __syncthreads();
  fft_lds<double, MAXE, LM, false>(buf, twl, 1, s, L, se, sl);
  for (int idx = threadIdx.x; idx < L * P; idx += blockDim.x) {
    int l, e;
    if (LM) {
      l = idx / P;
      e = idx % P;
    } else {
      l = idx % L;
      e = idx / L;
    }
    const int64_t u = line0 + l;
    if (u < nlines) {
      const int64_t inner = u % stride, outer = u / stride;
      const int64_t base = inner + outer * stride * P;
      a[base + e * stride] = buf[e * se + l * sl].x;
    }
  }
}

// Real rows of length N0 -> first H = N0/2+1 complex outputs of the forward
// DFT (full complex FFT of the real row), FP64.
template <int MAXE>
__global__ void k_r2c_rows_full(const double* in, cplx<double>* out, Sched s,
                                const cplx<double>* tw, int64_t rows, int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n, H = N / 2 + 1;
  cplx<double>* twl = reinterpret_cast<cplx<double>*>(smem);
  cplx<double>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int64_t row0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx / N, e = idx % N;
    const int64_t r = row0 + l;
    buf[l * N + e] = { r < rows ? in[r * N + e] : 0., 0. };
  }
  __syncthreads();
  fft_lds<double, MAXE, true, false>(buf, twl, 1, s, L, 1, N);
  for (int idx = threadIdx.x; idx < L * H; idx += blockDim.x) {
    const int l = idx / H, e = idx % H;
    const int64_t r = row0 + l;
    if (r < rows)
      out[r * H + e] = buf[l * N + e];
  }
}

// Complex c2c along a strided axis, in place (lines of n points at
// `stride`), unnormalised. Works for any T.
template <typename T, int MAXE, bool INV>
__global__ void k_c2c_strided(cplx<T>* a, Sched s, const cplx<T>* tw, int64_t stride,
                              int64_t nlines, int L, int64_t bstride)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  a += blockIdx.y * bstride;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> v = { 0, 0 };
    if (u < nlines) {
      const int64_t inner = u % stride, outer = u / stride;
      v = a[inner + outer * stride * N + e * stride];
    }
    buf[e * L + l] = v;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, INV>(buf, twl, 1, s, L, L, 1);
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    if (u < nlines) {
      const int64_t inner = u % stride, outer = u / stride;
      a[inner + outer * stride * N + e * stride] = buf[e * L + l];
    }
  }
}

// Block-wide sum of the classification records; the result is valid in
// thread 0.
__device__ ClassifyPartial block_reduce(ClassifyPartial c)
{
  __shared__ long long s_small[1024 / 64], s_zero[1024 / 64], s_neg[1024 / 64];
  __shared__ double s_min[1024 / 64];
  for (int off = 32; off > 0; off >>= 1) {
    c.small += __shfl_down(c.small, off);
    c.zero += __shfl_down(c.zero, off);
    c.negative += __shfl_down(c.negative, off);
    c.smallest = fmin(c.smallest, __shfl_down(c.smallest, off));
  }
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    s_small[wid] = c.small;
    s_zero[wid] = c.zero;
    s_neg[wid] = c.negative;
    s_min[wid] = c.smallest;
  }
  __syncthreads();
  ClassifyPartial r = { 0, 0, 0, 1.7976931348623157e308 };
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++) {
      r.small += s_small[w];
      r.zero += s_zero[w];
      r.negative += s_neg[w];
      r.smallest = fmin(r.smallest, s_min[w]);
    }
  }
  return r;
}

__device__ void block_reduce_write(ClassifyPartial c, ClassifyPartial* out)
{
  const ClassifyPartial r = block_reduce(c);
  if (threadIdx.x == 0)
    out[blockIdx.x] = r;
}

// Divide by N_ext and classify/clamp every entry (DCT backend; order free),
// matrix.hh:266-291 with dctmatrixbackend.hh:240-272.
__global__ void k_scale_classify_dct(double* a, int64_t n, double inv_next, double thr,
                                     ClassifyPartial* partial)
{
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = a[i] * inv_next;
    classify_one(v, thr, c);
    a[i] = v;
  }
  block_reduce_write(c, partial);
}

// R2C backend: divide the complex spectrum [N2][N1][H] by N_ext.
__global__ void k_scale_complex(cplx<double>* a, int64_t n, double inv_next)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i].x *= inv_next;
    a[i].y *= inv_next;
  }
}

// R2C backend classification over the first `count` reals of the interleaved
// complex buffer in the reference's storage order (r2cmatrixbackend.hh:378-400
// read through matrix.hh:270-291): real i is component i%2 of storage entry
// i/2; 3D transposed storage is (kx, kz, ky).
__global__ void k_classify_r2c(cplx<double>* spec, int64_t count, int H, int N1, int N2,
                               int transposed3d, double thr, ClassifyPartial* partial,
                               int check_all)
{
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  if (check_all) {
    // opt-in: real part of every entry
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
      classify_one(spec[i].x, thr, c);
    block_reduce_write(c, partial);
    return;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i >> 1;
    int64_t logical = s;
    if (transposed3d) {
      const int64_t kx = s % H, kz = (s / H) % N2, ky = s / ((int64_t)H * N2);
      logical = kx + (int64_t)H * (ky + (int64_t)N1 * kz);
    }
    double* p = reinterpret_cast<double*>(&spec[logical]) + (i & 1);
    double v = *p;
    classify_one(v, thr, c);
    *p = v;
  }
  block_reduce_write(c, partial);
}

// DFT matrix eigenvalues of a 1-D embedding: the real parts of the forward
// transform of kx < H, scaled by 1/N, classified and clamped like
// fillTransformedMatrix (matrix.hh:266-291) over the full array: entry kx
// stands for itself and, for 0 < kx < N/2, for its mirror N - kx.
__global__ void kdft_finish_eigs_1d(const cplx<double>* X, int64_t N, int H, double inv_n,
                                    double thr, double* eig, ClassifyPartial* partial)
{
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < H;
       k += (int64_t)gridDim.x * blockDim.x) {
    double v = X[k].x * inv_n;
    const bool twice = k != 0 && 2 * k != N;
    classify_one(v, thr, c);
    if (twice) {
      double w = X[k].x * inv_n;
      classify_one(w, thr, c);
    }
    eig[k] = v;
  }
  block_reduce_write(c, partial);
}

// 2D/3D DFT matrix: classification and clamping of the real parts of the half
// spectrum (natural layout [N2][N1][H]) with the multiplicity of each entry
// in the full array (kx and N0 - kx for 0 < kx < N0/2).
__global__ void kdft_classify(cplx<double>* spec, int64_t count, int H, int N0, double thr,
                              ClassifyPartial* partial)
{
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % H);
    double v = spec[i].x;
    const double v0 = v;
    classify_one(v, thr, c);
    if (kx != 0 && 2 * kx != N0) {
      double w = v0;
      classify_one(w, thr, c);
    }
    spec[i].x = v;
  }
  block_reduce_write(c, partial);
}

// R2C backend, full real-part check, spectrum with row pitch Hp: in one pass
// scale by 1/N_ext, classify and clamp the real part, write the eigenvalue
// (layout [rows][H]) and its square root in the field precision (the work of
// k_scale_complex, k_classify_r2c(check_all), k_real_part and k_sqrt_table).
template <typename T>
__global__ void k_finish_r2c_eigs(const cplx<double>* spec, int Hp, int H, int64_t rows,
                                  double inv_next, double thr, double* eig, T* sq,
                                  ClassifyPartial* partial)
{
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for_rows(rows, H, [&](int64_t r, int kx) {
    double v = spec[r * Hp + kx].x * inv_next;
    classify_one(v, thr, c);
    eig[r * H + kx] = v;
    sq[r * H + kx] = (T)sqrt(v);
  });
  block_reduce_write(c, partial);
}

// ---- fields of one shape in one launch (uneven batches) --------------------
// Field f = blockIdx.y of a group: its own covariance parameters, threshold
// and arrays at stride `stride` (entries) from the group's base.

template <int COV, int PER, int DIM>
__global__ void k_eval_cov_full_group(const CovParams* __restrict__ cps, int N0, int N1, int N2,
                                      double* __restrict__ out, int64_t stride)
{
  // a local copy with the dimension fixed at compile time: every loop over
  // the axes unrolls, every parameter index is static, and the parameters
  // (uniform) stay in scalar registers instead of being loaded per point
  CovParams p = cps[blockIdx.y];
  p.dim = DIM;
  out += (int64_t)blockIdx.y * stride;
  for_rows((int64_t)N1 * N2, N0, [&](int64_t row, int i) {
    const int64_t idx[3] = { i, (int64_t)((uint32_t)row % (uint32_t)N1),
                             (int64_t)((uint32_t)row / (uint32_t)N1) };
    out[row * N0 + i] = embedded_cov<COV, PER>(p, idx);
  });
}

// k_eval_cov_even2d per field (group of 2D fields): field f = blockIdx.y
// with its own parameters, rows at out + f * stride.
template <int COV, int PER>
__global__ void k_eval_cov_even2d_group(const CovParams* __restrict__ cps, int N0, int N1,
                                        double* __restrict__ out, int64_t stride)
{
  CovParams p = cps[blockIdx.y];
  p.dim = 2;
  out += (int64_t)blockIdx.y * stride;
  for_rows(N1 / 2 + 1, N0, [&](int64_t row, int i) {
    out[row * N0 + i] = cov_even_2d<COV, PER>(p, N0, N1, i, (int)row);
  });
}

// k_finish_r2c_eigs per field: spectrum [rows][Hp], eigenvalues and roots
// [rows][H], gridDim.x block partials per field.
template <typename T>
__global__ void k_finish_r2c_eigs_group(const cplx<double>* spec, int Hp, int H, int64_t rows,
                                        double inv_next, const double* thr, double* eig, T* sq,
                                        ClassifyPartial* partial)
{
  const int f = blockIdx.y;
  const int64_t count = (int64_t)H * rows;
  spec += (int64_t)f * Hp * rows;
  if (eig) // nullptr: the FP64 eigenvalues are not kept
    eig += (int64_t)f * count;
  sq += (int64_t)f * count;
  const double t = thr[f];
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for_rows(rows, H, [&](int64_t r, int kx) {
    double v = spec[r * Hp + kx].x * inv_next;
    classify_one(v, t, c);
    if (eig)
      eig[r * H + kx] = v;
    sq[r * H + kx] = (T)sqrt(v);
  });
  block_reduce_write(c, partial + (int64_t)f * gridDim.x);
}

// One block per field: the field's n partials into stats[fidx[f]].
__global__ void k_reduce_partials_group(const ClassifyPartial* partial, int n,
                                        const int64_t* fidx, ClassifyPartial* stats)
{
  const int f = blockIdx.x;
  ClassifyPartial c = { 0, 0, 0, 1.7976931348623157e308 };
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const ClassifyPartial q = partial[(int64_t)f * n + i];
    c.small += q.small;
    c.zero += q.zero;
    c.negative += q.negative;
    c.smallest = fmin(c.smallest, q.smallest);
  }
  const ClassifyPartial r = block_reduce(c);
  if (threadIdx.x == 0)
    stats[fidx[f]] = r;
}

// Sum `n` block partials into one record on the device (capture-safe).
__global__ void k_reduce_partials(const ClassifyPartial* partial, int n, ClassifyPartial* out)
{
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  ClassifyPartial r = { 0, 0, 0, 1.7976931348623157e308 };
  for (int i = 0; i < n; i++) {
    r.small += partial[i].small;
    r.zero += partial[i].zero;
    r.negative += partial[i].negative;
    r.smallest = fmin(r.smallest, partial[i].smallest);
  }
  *out = r;
}

__global__ void k_real_part(const cplx<double>* spec, double* out, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = spec[i].x;
}

template <typename T>
__global__ void k_sqrt_table(const double* eig, T* out, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (T)sqrt(eig[i]);
}

// ---------------------------------------------------------------------------
// Field generation, R2C field backend (matrix.hh:374-421,
// r2cfieldbackend.hh:250-325, :423-441).
// ---------------------------------------------------------------------------
// Spectrum entry: R2CFieldBackend::set, r2cfieldbackend.hh:305-325.
template <typename T, int NOISE>
__device__ __forceinline__ cplx<T> spectrum_entry(const GenGeom& g, const T* sq, int kx, int k1,
                                                  int k2, const double* noise, uint64_t seed,
                                                  uint64_t real)
{
  const int64_t s = storage_index(g, kx, k1, k2);
  const T lambda = sq[eig_index(g, kx, k1, k2)];
  T r1, r2;
  if (NOISE == 0) {
    r1 = (T)noise[2 * s];
    r2 = (T)noise[2 * s + 1];
  } else {
    philox_pair<T>(seed, real, (uint64_t)s, r1, r2);
  }
  const bool allMultiple = ((2 * kx) % g.N0 == 0) && ((2 * k1) % g.N1 == 0) &&
                           ((2 * k2) % g.N2 == 0);
  if (allMultiple)
    return { lambda * r1, T(0) };
  const T ls = lambda * (T)0.70710678118654752440; // lambda / sqrt(2)
  return { ls * r1, ls * r2 };
}

// Pass G: spectrum of line (kx, k_mid) along the last axis, generated in
// registers, inverse DFT, first n_last outputs -> B1[j][k_mid][kx].
template <typename T, int MAXE, int NOISE>
__global__ void k_gen_last(GenGeom g, const T* sq, const double* noise, uint64_t seed,
                           uint64_t first_real, cplx<T>* B1, Sched s, const cplx<T>* tw, int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  const int64_t nlines = (g.dim == 3) ? (int64_t)g.H * g.N1 : g.H;
  const int nlast = (g.dim == 3) ? g.n2 : g.n1;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  const double* nz = noise ? noise + b * g.noise_stride : nullptr;
  const uint64_t real = first_real + b;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> z = { 0, 0 };
    if (u < nlines) {
      const int kx = (int)(u % g.H), km = (int)(u / g.H);
      if (g.dim == 3)
        z = spectrum_entry<T, NOISE>(g, sq, kx, km, e, nz, seed, real);
      else
        z = spectrum_entry<T, NOISE>(g, sq, kx, e, 0, nz, seed, real);
    }
    buf[e * L + l] = z;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, true>(buf, twl, 1, s, L, L, 1);
  const int64_t plane = (int64_t)g.Hp * (g.dim == 3 ? g.N1 : 1);
  cplx<T>* out = B1 + (int64_t)b * plane * nlast;
  for (int idx = threadIdx.x; idx < L * nlast; idx += blockDim.x) {
    const int l = idx % L, j = idx / L;
    const int64_t u = line0 + l;
    if (u < nlines)
      out[j * plane + (u / g.H) * g.Hp + (u % g.H)] = buf[j * L + l];
  }
}

// Pass Y (3D): lines (kx, z) along ky of B1[z][ky][kx]; inverse DFT (or
// forward with FWD); first n_out outputs -> B2[z][y][kx]. The input holds
// n_in <= N1 leading points (rest zero).
template <typename T, int MAXE, bool INV>
__global__ void k_pass_mid(const cplx<T>* in, cplx<T>* out, int H, int Hp, int nz, int in_rows,
                           int n_in, int out_rows, int n_out, Sched s, const cplx<T>* tw, int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  const int64_t nlines = (int64_t)H * nz;
  in += (int64_t)b * Hp * nz * in_rows;
  out += (int64_t)b * Hp * nz * out_rows;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> v = { 0, 0 };
    if (u < nlines && e < n_in) {
      const int64_t kx = u % H, z = u / H;
      v = in[(z * in_rows + e) * Hp + kx];
    }
    buf[e * L + l] = v;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, INV>(buf, twl, 1, s, L, L, 1);
  for (int idx = threadIdx.x; idx < L * n_out; idx += blockDim.x) {
    const int l = idx % L, j = idx / L;
    const int64_t u = line0 + l;
    if (u < nlines) {
      const int64_t kx = u % H, z = u / H;
      out[(z * out_rows + j) * Hp + kx] = buf[j * L + l];
    }
  }
}

// Pass X: 1-D c2r along x of rows of H complex (FFTW c2r semantics: Re of the
// DC and Nyquist entries), as an M = N0/2 point complex inverse FFT of
// Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N0} (X_k - conj X_{M-k});
// out[2j] = Re z_j, out[2j+1] = Im z_j; first n0 values, value transform.
// tw is the N0-point table (the M-point FFT uses every second entry).
template <typename T, typename TO, int MAXE>
__global__ void k_c2r_x(const cplx<T>* in, TO* out, int64_t rows, int Hp, int n0, Sched s,
                        const cplx<T>* tw, int L, int transform, T param)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int M = s.n, N0 = 2 * M;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N0;
  for (int i = threadIdx.x; i < N0; i += blockDim.x)
    twl[i] = tw[i];
  __syncthreads();
  const int b = blockIdx.y;
  in += (int64_t)b * rows * Hp;
  out += (int64_t)b * rows * n0;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
    const int l = idx / M, k = idx % M;
    const int64_t r = row0 + l;
    cplx<T> z = { 0, 0 };
    if (r < rows) {
      cplx<T> xk = in[r * Hp + k];
      cplx<T> xm = in[r * Hp + (M - k)];
      if (k == 0) {
        xk.y = 0;
        xm.y = 0;
      }
      const cplx<T> a = { xk.x + xm.x, xk.y - xm.y };  // X_k + conj X_{M-k}
      const cplx<T> d = { xk.x - xm.x, xk.y + xm.y };  // X_k - conj X_{M-k}
      const cplx<T> t = cconj(twl[k]);                 // e^{+2 pi i k / N0}
      const cplx<T> td = cmul(t, d);
      z = { a.x - td.y, a.y + td.x };
    }
    buf[l * M + k] = z;
  }
  __syncthreads();
  fft_lds<T, MAXE, true, true>(buf, twl, 2, s, L, 1, M);
  for (int idx = threadIdx.x; idx < L * n0; idx += blockDim.x) {
    const int l = idx / n0, x = idx % n0;
    const int64_t r = row0 + l;
    if (r < rows) {
      const cplx<T> v = buf[l * M + (x >> 1)];
      out[r * n0 + x] = (TO)value_transform<T>(transform, (x & 1) ? v.y : v.x, param);
    }
  }
}

// Pass FX (matvec input): zero-padded real rows (n0 of N0 values) -> H
// complex via an M-point FFT of z_j = x_{2j} + i x_{2j+1}:
// X_k = (Z_k + conj Z_{M-k})/2 - i e^{-2 pi i k/N0} (Z_k - conj Z_{M-k})/2.
template <typename T, typename TI, int MAXE>
__global__ void k_r2c_x(const TI* in, cplx<T>* out, int64_t rows, int Hp, int n0, Sched s,
                        const cplx<T>* tw, int L)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int M = s.n, N0 = 2 * M;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N0;
  for (int i = threadIdx.x; i < N0; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  const int H = M + 1;
  in += (int64_t)b * rows * n0;
  out += (int64_t)b * rows * Hp;
  const int64_t row0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
    const int l = idx / M, j = idx % M;
    const int64_t r = row0 + l;
    cplx<T> z = { 0, 0 };
    if (r < rows) {
      const int x0 = 2 * j, x1 = 2 * j + 1;
      z.x = x0 < n0 ? (T)in[r * n0 + x0] : T(0);
      z.y = x1 < n0 ? (T)in[r * n0 + x1] : T(0);
    }
    buf[l * M + j] = z;
  }
  __syncthreads();
  fft_lds<T, MAXE, true, false>(buf, twl, 2, s, L, 1, M);
  for (int idx = threadIdx.x; idx < L * H; idx += blockDim.x) {
    const int l = idx / H, k = idx % H;
    const int64_t r = row0 + l;
    if (r < rows) {
      const cplx<T> zk = buf[l * M + (k % M)];
      const cplx<T> zm = buf[l * M + ((M - k) % M)];
      const cplx<T> a = { (zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5) }; // E_k
      const cplx<T> d = { (zk.x - zm.x) * T(0.5), (zk.y + zm.y) * T(0.5) }; // (Z_k - conj Z_{M-k})/2
      const cplx<T> w = twl[k];                                           // e^{-2 pi i k/N0}
      const cplx<T> wd = cmul(w, d);
      // X_k = a - i * wd
      out[r * Hp + k] = { a.x + wd.y, a.y - wd.x };
    }
  }
}

// Fused pass along the last axis of the matvec: lines (kx, k_mid) of
// C[j][k_mid][kx], n_in leading points (rest zero) -> forward DFT -> times
// f(k)/N_ext -> inverse DFT -> first n_in points back in place.
template <typename T, int MAXE>
__global__ void k_fused_mult(cplx<T>* C, GenGeom g, const double* eig, int op, double inv_next,
                             Sched s, const cplx<T>* tw, int L, const double* spec, int nbins)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = s.n;
  cplx<T>* twl = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* buf = twl + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    twl[i] = tw[i];
  const int b = blockIdx.y;
  const int64_t nlines = (g.dim == 3) ? (int64_t)g.H * g.N1 : g.H;
  const int nlast = (g.dim == 3) ? g.n2 : g.n1;
  const int64_t plane = (int64_t)g.Hp * (g.dim == 3 ? g.N1 : 1);
  C += (int64_t)b * plane * nlast;
  const int64_t line0 = (int64_t)blockIdx.x * L;
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    cplx<T> v = { 0, 0 };
    if (u < nlines && e < nlast)
      v = C[e * plane + (u / g.H) * g.Hp + (u % g.H)];
    buf[e * L + l] = v;
  }
  __syncthreads();
  fft_lds<T, MAXE, false, false>(buf, twl, 1, s, L, L, 1);
  for (int idx = threadIdx.x; idx < L * N; idx += blockDim.x) {
    const int l = idx % L, e = idx / L;
    const int64_t u = line0 + l;
    if (u < nlines) {
      const int kx = (int)(u % g.H), km = (int)(u / g.H);
      const int k1 = (g.dim == 3) ? km : e, k2 = (g.dim == 3) ? e : 0;
      double f;
      if (op == PF_OP_SPECTRAL)
        f = spectral_factor(g, kx, k1, k2, spec, nbins);
      else {
        const double m = eig[eig_index(g, kx, k1, k2)];
        f = op == PF_OP_MULTIPLY ? m
            : op == PF_OP_ROOT   ? sqrt(m)
            : op == PF_OP_INVERSE_FLOOR ? 1. / fmax(m, spec[0])
                                 : 1. / m;
      }
      f *= inv_next;
      buf[e * L + l] = cscale(buf[e * L + l], (T)f);
    }
  }
  __syncthreads();
  fft_lds<T, MAXE, false, true>(buf, twl, 1, s, L, L, 1);
  for (int idx = threadIdx.x; idx < L * nlast; idx += blockDim.x) {
    const int l = idx % L, j = idx / L;
    const int64_t u = line0 + l;
    if (u < nlines)
      C[j * plane + (u / g.H) * g.Hp + (u % g.H)] = buf[j * L + l];
  }
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_value_transform(T* v, int64_t n, int kind, T param)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = value_transform<T>(kind, v[i], param);
}

__global__ void k_philox_normals(uint64_t seed, uint64_t real, uint64_t first, int64_t n,
                                 double* out)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = philox_normal<double>(seed, real, first + (uint64_t)i);
}

__global__ void k_philox_raw(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                             uint32_t k1, int64_t nblocks, uint32_t* out)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nblocks;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t c[4] = { c0 + (uint32_t)i, c1, c2, c3 };
    philox4x32_10(c, k0, k1);
    out[4 * i + 0] = c[0];
    out[4 * i + 1] = c[1];
    out[4 * i + 2] = c[2];
    out[4 * i + 3] = c[3];
  }
}

// dot products: per-block partial sums (FP64)
__global__ void k_dot_partial(const double* a, const double* b, int64_t n, double* partial)
{
  __shared__ double sh[1024 / 64];
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_down(s, off);
  if (threadIdx.x % 64 == 0)
    sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++)
      t += sh[w];
    partial[blockIdx.x] = t;
  }
}

// y = a*x + y
__global__ void k_axpy(double a, const double* x, double* y, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

// y = x + beta*y
__global__ void k_xpby(const double* x, double beta, double* y, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] + beta * y[i];
}

// z = x - y
__global__ void k_sub(const double* x, const double* y, double* z, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = x[i] - y[i];
}

// energy functional partial sums: sum iter_i * (0.5 * Aiter_i - b_i)
__global__ void k_energy_partial(const double* it, const double* ait, const double* b, int64_t n,
                                 double* partial)
{
  __shared__ double sh[1024 / 64];
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += it[i] * (0.5 * ait[i] - b[i]);
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_down(s, off);
  if (threadIdx.x % 64 == 0)
    sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++)
      t += sh[w];
    partial[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// Device-resident (P)CG: the scalars live in device memory and every kernel
// reads them there, so an iteration is enqueued without a host round trip.
// A solve stops updating once `done` is set (the recursive residual met the
// tolerance); the host looks at the scalars every few iterations.
// ---------------------------------------------------------------------------
struct CgScalars
{
  double rzs[2]; // r . z (z = M^{-1} r; r . r for plain CG): iteration k reads slot k % 2 and
                 // the p update of iteration k writes the next value to slot (k + 1) % 2
  double rr;     // r . r
  double bb;     // b . b
  double tol2;   // (rel_tol)^2 * b . b
  int done;      // 1: the recursive residual met the tolerance
  int it;        // iterations taken
};

// Block sums of a . b into partial[blockIdx.x] (deterministic order).
__device__ __forceinline__ double block_sum(double s)
{
  __shared__ double sh[1024 / 64];
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_down(s, off);
  if (threadIdx.x % 64 == 0)
    sh[threadIdx.x / 64] = s;
  __syncthreads();
  double t = 0.;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++)
      t += sh[w];
  return t;
}

// Sum of nb partials into *out by one block (fixed order of the partials
// per thread, then the block tree).
__device__ __forceinline__ double reduce_partials_dev(const double* partial, int nb)
{
  double s = 0.;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    s += partial[i];
  return block_sum(s);
}

// Sum of the nb partials a previous kernel wrote, the same value in every
// thread of every block: each block of the consuming kernel reduces them
// itself in one fixed order (no one-block kernel in between).
__device__ __forceinline__ double all_reduce_partials(const double* partial, int nb)
{
  __shared__ double tot;
  const double t = reduce_partials_dev(partial, nb);
  if (threadIdx.x == 0)
    tot = t;
  __syncthreads();
  const double v = tot;
  __syncthreads(); // block_sum's scratch and tot may be reused after this
  return v;
}

// partial = a . b over the vector (skipped once the solve is done)
__global__ void k_cg_dot(const double* a, const double* b, int64_t n, double* partial,
                         const CgScalars* sc)
{
  if (sc->done)
    return;
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  const double t = block_sum(s);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = t;
}

// x += alpha p, r -= alpha Ap with alpha = rz / pAp (pAp from the nb partials
// of p . Ap, rz from slot par), partial sums of r . r into rr_part
__global__ void k_cg_update_xr(const double* p, const double* Ap, double* x, double* r, int64_t n,
                               const double* pap_part, int nb, double* rr_part,
                               const CgScalars* sc, int par)
{
  if (sc->done)
    return;
  const double alpha = sc->rzs[par] / all_reduce_partials(pap_part, nb);
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    s += ri * ri;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0)
    rr_part[blockIdx.x] = t;
}

// End of an iteration: rr from the partials of r . r. Every block decides
// alike: with rr <= tol^2 b.b the solve is done and p is left as it is;
// otherwise p = z + beta p with beta = rz_new / rz (rz_new from the partials
// of r . z, or rr for plain CG when rz_part is null). Block 0 records rr and
// the iteration, sets done, or stores rz_new in the other slot (the slot the
// blocks read stays untouched during the kernel).
__global__ void k_cg_update_p(const double* z, double* p, int64_t n, const double* rr_part,
                              const double* rz_part, int nb, CgScalars* sc, int par)
{
  if (sc->done)
    return;
  const double rr = all_reduce_partials(rr_part, nb);
  const bool stop = rr <= sc->tol2;
  const double rz_new = stop ? 0. : rz_part ? all_reduce_partials(rz_part, nb) : rr;
  const double rz = sc->rzs[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rr = rr;
    sc->it += 1;
    if (stop)
      sc->done = 1;
    else
      sc->rzs[par ^ 1] = rz_new;
  }
  if (stop)
    return;
  const double beta = rz_new / rz;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

} // namespace pfamd
