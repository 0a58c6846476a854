// pf_mlp.hpp -- kernels of the small MLP behind the learned spectral
// preconditioner (no counterpart in the reference): FP32 dense layers with
// exact GELU, mean-squared-error loss, backward pass and Adam, plus the
// radial binning of embedding eigenvalues that produces its targets.
#pragma once

#include "pf_common.hpp"

namespace pfamd {

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x)
{
  const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// Z = X W^T + b (X: B x a, W: b x a row-major); A = gelu(Z) if act else Z.
__global__ void k_mlp_linear_fwd(const float* X, const float* W, const float* bias, float* Z,
                                 float* A, int B, int a, int b, int act)
{
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * b)
    return;
  const int s = (int)(t / b), j = (int)(t % b);
  float z = bias[j];
  for (int i = 0; i < a; i++)
    z += W[(int64_t)j * a + i] * X[(int64_t)s * a + i];
  Z[t] = z;
  A[t] = act ? gelu(z) : z;
}

// D = 2 (P - Y) / n, per-block partial sums of (P - Y)^2 / n.
__global__ void k_mlp_mse(const float* P, const float* Y, float* D, int64_t n, double* partial)
{
  __shared__ double sh[1024 / 64];
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float e = P[i] - Y[i];
    D[i] = 2.f * e / (float)n;
    s += (double)e * e / (double)n;
  }
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

// D *= gelu'(Z) elementwise
__global__ void k_mlp_gelu_back(float* D, const float* Z, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    D[i] *= gelu_grad(Z[i]);
}

// Weight and bias gradients of one layer, gW[j][i] = sum_s Dz[s][j] X[s][i]
// and gb[j] = sum_s Dz[s][j], in two steps: block (split q, output tile) sums
// the samples s = q, q + nsplit, ... of its outputs t (t < a b: weight j i,
// t = a b + j: bias j) into part[q][t]; k_mlp_grad_sum adds the nsplit
// partials of each output in one fixed order.
__global__ void k_mlp_grad_part(const float* Dz, const float* X, float* part, int B, int a, int b)
{
  const int t = blockIdx.y * blockDim.x + threadIdx.x;
  const int q = blockIdx.x, nsplit = gridDim.x;
  const int nout = a * b + b;
  if (t >= nout)
    return;
  const bool w = t < a * b;
  const int j = w ? t / a : t - a * b, i = w ? t % a : 0;
  float acc = 0.f;
  for (int s = q; s < B; s += nsplit)
    acc += Dz[(int64_t)s * b + j] * (w ? X[(int64_t)s * a + i] : 1.f);
  part[(int64_t)q * nout + t] = acc;
}

__global__ void k_mlp_grad_sum(const float* part, int nsplit, float* gW, float* gb, int a, int b)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nout = a * b + b;
  if (t >= nout)
    return;
  float acc = 0.f;
  for (int q = 0; q < nsplit; q++)
    acc += part[(int64_t)q * nout + t];
  if (t < a * b)
    gW[t] = acc;
  else
    gb[t - a * b] = acc;
}

// Din[s][i] = sum_j Dz[s][j] W[j][i]
__global__ void k_mlp_back_in(const float* Dz, const float* W, float* Din, int B, int a, int b)
{
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * a)
    return;
  const int s = (int)(t / a), i = (int)(t % a);
  float acc = 0.f;
  for (int j = 0; j < b; j++)
    acc += Dz[(int64_t)s * b + j] * W[(int64_t)j * a + i];
  Din[t] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                           float b1, float b2, float eps, float c1, float c2)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Per-bin sums of max(lambda, floor) and counts over the DCT-half eigenvalue
// array [P2][P1][P0] (every folded spectral position once). Each block keeps
// its bins in LDS (dynamic shared memory of nbins sums and nbins counts) and
// adds them to the global bins once at its end.
__global__ void k_radial_sums(const double* eig, int P0, int P1, int P2, int dim, int nbins,
                              double floor_v, double* sums, unsigned long long* counts)
{
  extern __shared__ __align__(16) unsigned char smem[];
  double* lsum = reinterpret_cast<double*>(smem);
  unsigned long long* lcnt = reinterpret_cast<unsigned long long*>(lsum + nbins);
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
    lsum[i] = 0.;
    lcnt[i] = 0ull;
  }
  __syncthreads();
  const int64_t total = (int64_t)P0 * P1 * P2;
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < total;
       f += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(f % P0), b = (int)((f / P0) % P1), c = (int)(f / ((int64_t)P0 * P1));
    const int bin = radial_bin(a, b, c, P0 - 1, P1 - 1, P2 > 1 ? P2 - 1 : 1, dim, nbins);
    const double v = eig[f] > floor_v ? eig[f] : floor_v;
    atomicAdd(&lsum[bin], v);
    atomicAdd(&lcnt[bin], 1ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += blockDim.x)
    if (lcnt[i] > 0) {
      atomicAdd(&sums[i], lsum[i]);
      atomicAdd(&counts[i], lcnt[i]);
    }
}

} // namespace pfamd
