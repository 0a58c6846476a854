// pf_common.hpp -- spectrum geometry shared by the generic and the
// specialised kernels: how a half-spectrum position maps to the reference's
// noise order and to the eigenvalue table.
#pragma once

#include "pf_device.hpp"

namespace pfamd {

// XCD-aware block order. The hardware hands linear block b (x fastest) to
// XCD b % 8 (MI355X: 8 XCDs of 32 CUs, one L2 each). This returns the
// logical block (bx, by, bz) so that the blocks one XCD receives form one
// contiguous run of logical indices: tiles that share cache lines (the
// neighbouring kx tiles of a row) then meet in the same L2.
// Logical index of linear block b of n (a bijection on [0, n)).
__device__ __forceinline__ unsigned xcd_linear(unsigned b, unsigned n)
{
  const unsigned x = b & 7u, q = n >> 3, r = n & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ void xcd_block(int& bx, int& by, int& bz)
{
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned n = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned lam = xcd_linear(b, n);
  bx = (int)(lam % gx);
  const unsigned t = lam / gx;
  by = (int)(t % gy);
  bz = (int)(t / gy);
}

struct GenGeom
{
  int dim, transposed;
  int N0, N1, N2, H;
  int n0, n1, n2;
  int P0, P1, P2;       // DCT half sizes (eigen layout of the DCT backend)
  int dct;              // 1: eigen table is the DCT half array (fold k >= P)
  int64_t noise_stride; // samples per realization
  int Hp;               // row pitch (complex entries) of the spectrum work arrays, H rounded up
                        // to 128-byte rows
  // fields of one shape generated together (uneven batches): realization b
  // draws Philox realization first + fidx[b] and reads its own root table at
  // sq + b * sq_stride; nullptr / 0 for the realizations of one plan
  const int64_t* fidx = nullptr;
  int64_t sq_stride = 0;
};

// Reference storage index of half-spectrum entry (kx, k1, k2): the order in
// which generateField draws the noise.
__device__ __forceinline__ int64_t storage_index(const GenGeom& g, int kx, int k1, int k2)
{
  if (g.dim == 3 && g.transposed)
    return kx + (int64_t)g.H * (k2 + (int64_t)g.N2 * k1);
  return kx + (int64_t)g.H * (k1 + (int64_t)g.N1 * k2);
}

__device__ __forceinline__ int64_t eig_index(const GenGeom& g, int kx, int k1, int k2)
{
  if (g.dct) {
    const int a = kx >= g.P0 ? g.N0 - kx : kx;
    const int b = k1 >= g.P1 ? g.N1 - k1 : k1;
    const int c = k2 >= g.P2 ? g.N2 - k2 : k2;
    return a + (int64_t)g.P0 * (b + (int64_t)g.P1 * c);
  }
  return kx + (int64_t)g.H * (k1 + (int64_t)g.N1 * k2);
}

// Eigenvalue classification (matrix.hh:266-291): counts of small, zero and
// negative eigenvalues against the threshold, and the smallest value; the
// value is clamped to 0 when negative.
struct ClassifyPartial
{
  long long small, zero, negative;
  double smallest;
};

__device__ __forceinline__ void classify_one(double& v, double thr, ClassifyPartial& c)
{
  if (v < c.smallest)
    c.smallest = v;
  if (v < 1e-6) {
    if (v < thr) {
      if (v > -thr)
        c.zero++;
      else
        c.negative++;
    } else
      c.small++;
  }
  if (v < 0.)
    v = 0.;
}

// Operator index of the learned spectral preconditioner in the fused matvec
// pass (after PF_OP_MULTIPLY / PF_OP_ROOT / PF_OP_INVERSE_EMBED).
constexpr int PF_OP_SPECTRAL = 3;
// Inverse of the embedding with the eigenvalues floored at table[0]: the
// embedding-inverse preconditioner of (P)CG, 1 / max(lambda, nugget / N_ext),
// defined where the clamp left zero eigenvalues.
constexpr int PF_OP_INVERSE_FLOOR = 4;

// Radial bin of a (folded) spectral position: r = |k / (N/2)| / sqrt(dim).
__device__ __forceinline__ int radial_bin(int a, int b, int c, int h0, int h1, int h2, int dim,
                                          int nbins)
{
  const double x = (double)a / h0, y = (double)b / h1, z = dim == 3 ? (double)c / h2 : 0.;
  const double r = sqrt((x * x + y * y + z * z) / dim);
  int bin = (int)(r * nbins);
  return bin < nbins ? bin : nbins - 1;
}

// Multiplier of a spectral position by the radial table (folded k).
__device__ __forceinline__ double spectral_factor(const GenGeom& g, int kx, int k1, int k2,
                                                  const double* table, int nbins)
{
  const int a = 2 * kx > g.N0 ? g.N0 - kx : kx;
  const int b = 2 * k1 > g.N1 ? g.N1 - k1 : k1;
  const int c = 2 * k2 > g.N2 ? g.N2 - k2 : k2;
  return table[radial_bin(a, b, c, g.N0 / 2, g.N1 > 1 ? g.N1 / 2 : 1, g.N2 > 1 ? g.N2 / 2 : 1,
                          g.dim, nbins)];
}

} // namespace pfamd
