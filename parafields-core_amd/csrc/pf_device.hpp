// pf_device.hpp -- device-side building blocks of the MI355X circulant-
// embedding engine: complex arithmetic, the Philox4x32-10 normal stream,
// covariance families and coordinate geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/parafields_amd.h"

namespace pfamd {

template <typename T>
struct cplx
{
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b)
{
  return { a.x + b.x, a.y + b.y };
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b)
{
  return { a.x - b.x, a.y - b.y };
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
  return { a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x };
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a)
{
  return { a.x, -a.y };
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s)
{
  return { a.x * s, a.y * s };
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). The normal stream of a
// realization: sample s lives in block s/4 (counter = (block, realization)),
// key = 64-bit seed; uniforms u = (x + 0.5) 2^-32; Box-Muller on (u0,u1) and
// (u2,u3): n_a = sqrt(-2 ln u_a) cos(2 pi u_b), n_b = ... sin(2 pi u_b).
// ---------------------------------------------------------------------------
// a ^ b ^ k in one gfx950 three-input bit operation (register-only ALU
// instruction; truth table 0x96 is the parity of the three inputs, symmetric
// in their order). k must be wave-uniform: every caller passes a key derived
// from a kernel argument, held in an SGPR.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t k)
{
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
  return r;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32x32->64 product per multiplier (v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = xor3(hi1, c[1], k0);
    const uint32_t n2 = xor3(hi0, c[3], k1);
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// philox4x32_10 of the counter (c0, 0, u2, u3) where only c0 varies across the
// wave (a block index below 2^32; u2, u3 = the realization, wave-uniform).
// The same ten rounds, with the wave-uniform halves of the first three
// rounds written out so the compiler keeps them in scalar registers: round 1
// multiplies the uniform word u2, round 2 the uniform output of round 1, so
// two of the twenty 32x32->64 products and one XOR leave the vector unit.
// Outputs equal philox4x32_10 on { c0, 0, u2, u3 } bit for bit.
__device__ __forceinline__ void philox4x32_10_lane0(uint32_t c0, uint32_t u2, uint32_t u3,
                                                    uint32_t k0, uint32_t k1, uint32_t out[4])
{
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  // round 1 (keys k0, k1): uniform product M1 * u2
  const uint64_t q1 = (uint64_t)M1 * u2;
  const uint32_t a0 = (uint32_t)(q1 >> 32) ^ k0; // ^ c1 = 0
  const uint32_t a1 = (uint32_t)q1;
  const uint64_t p0 = (uint64_t)M0 * c0;
  const uint32_t a2 = (uint32_t)(p0 >> 32) ^ (u3 ^ k1);
  const uint32_t a3 = (uint32_t)p0;
  // round 2 (keys + W): uniform product M0 * a0
  const uint64_t q0 = (uint64_t)M0 * a0;
  const uint64_t p1 = (uint64_t)M1 * a2;
  const uint32_t b0 = (uint32_t)(p1 >> 32) ^ (a1 ^ (k0 + W0));
  const uint32_t b1 = (uint32_t)p1;
  const uint32_t b2 = a3 ^ ((uint32_t)(q0 >> 32) ^ (k1 + W1));
  const uint32_t b3 = (uint32_t)q0;
  // round 3 (keys + 2W): b3 is uniform
  const uint64_t r0 = (uint64_t)M0 * b0;
  const uint64_t r1 = (uint64_t)M1 * b2;
  uint32_t c[4];
  c[0] = xor3((uint32_t)(r1 >> 32), b1, k0 + 2 * W0);
  c[1] = (uint32_t)r1;
  c[2] = (uint32_t)(r0 >> 32) ^ (b3 ^ (k1 + 2 * W1));
  c[3] = (uint32_t)r0;
  // rounds 4..10
  uint32_t kk0 = k0 + 2 * W0, kk1 = k1 + 2 * W1;
#pragma unroll
  for (int r = 3; r < 10; r++) {
    kk0 += W0;
    kk1 += W1;
    const uint64_t s0 = (uint64_t)M0 * c[0];
    const uint64_t s1 = (uint64_t)M1 * c[2];
    const uint32_t n0 = xor3((uint32_t)(s1 >> 32), c[1], kk0);
    const uint32_t n2 = xor3((uint32_t)(s0 >> 32), c[3], kk1);
    c[0] = n0;
    c[1] = (uint32_t)s1;
    c[2] = n2;
    c[3] = (uint32_t)s0;
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = c[3];
}

// Pair of normals (xi_1, xi_2) = samples (2e, 2e+1) of the stream: one
// Box-Muller pair of block e/2.
template <typename T>
struct BoxMuller;

template <>
struct BoxMuller<double>
{
  static __device__ __forceinline__ void pair(uint32_t xa, uint32_t xb, double& na, double& nb)
  {
    const double ua = ((double)xa + 0.5) * (1.0 / 4294967296.0);
    const double ub = ((double)xb + 0.5) * (1.0 / 4294967296.0);
    const double r = sqrt(-2.0 * log(ua));
    double s, c;
    sincospi(2.0 * ub, &s, &c);
    na = r * c;
    nb = r * s;
  }
};

template <>
struct BoxMuller<float>
{
  // u in (0, 1]: float of the 32-bit integer scaled by 2^-32 plus half an ulp
  static __device__ __forceinline__ float turn(uint32_t x)
  {
    return fmaf((float)x, 2.3283064365386963e-10f, 1.1641532182693481e-10f);
  }
  // sqrt(-2 ln u): ln(u) = log2(u) * ln 2 with the hardware log2; the
  // argument of the root is in [0, 44.4] (never denormal), so the hardware
  // square root (1 ulp) needs none of the IEEE wrapper's scaling steps
  static __device__ __forceinline__ float radius(uint32_t xa)
  {
    return __builtin_amdgcn_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(turn(xa)));
  }
  static __device__ __forceinline__ void pair(uint32_t xa, uint32_t xb, float& na, float& nb)
  {
    const float r = radius(xa);
    // hardware sin/cos take their argument in revolutions: sin(2 pi ub)
    const float ub = turn(xb);
    na = r * __builtin_amdgcn_cosf(ub);
    nb = r * __builtin_amdgcn_sinf(ub);
  }
};

template <typename T>
__device__ __forceinline__ void philox_pair(uint64_t seed, uint64_t real, uint64_t entry, T& xi1,
                                            T& xi2)
{
  const uint64_t blk = entry >> 1;
  uint32_t c[4] = { (uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)real, (uint32_t)(real >> 32) };
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  if (entry & 1)
    BoxMuller<T>::pair(c[2], c[3], xi1, xi2);
  else
    BoxMuller<T>::pair(c[0], c[1], xi1, xi2);
}

// Single normal of sample index s.
template <typename T>
__device__ __forceinline__ T philox_normal(uint64_t seed, uint64_t real, uint64_t s)
{
  T a, b;
  philox_pair<T>(seed, real, s >> 1, a, b);
  return (s & 1) ? b : a;
}

// philox_normal for a sample drawn alone: the same value, computing only the
// half of the Box-Muller pair that sample s takes. FP32: cos(2 pi u) is
// taken as sin(2 pi (u + 1/4)), one hardware sine (the sum rounds the turn
// by at most 6e-8).
template <typename T>
__device__ __forceinline__ T philox_normal1(uint64_t seed, uint64_t real, uint64_t s)
{
  return philox_normal<T>(seed, real, s);
}

// The normal of sample s from the four words c of its Philox block (block
// s / 4): Box-Muller of the word pair (s / 2) mod 2, its cosine (s even) or
// sine (s odd) part.
// (The words are passed as values: a select between elements of an array
// would keep the array in private memory.)
template <typename T>
__device__ __forceinline__ T box_muller_of_sample(uint32_t c0, uint32_t c1, uint32_t c2,
                                                  uint32_t c3, uint64_t s)
{
  const bool hi = (s >> 1) & 1;
  T a, b;
  BoxMuller<T>::pair(hi ? c2 : c0, hi ? c3 : c1, a, b);
  return (s & 1) ? b : a;
}

template <>
__device__ __forceinline__ float box_muller_of_sample<float>(uint32_t c0, uint32_t c1, uint32_t c2,
                                                             uint32_t c3, uint64_t s)
{
  const bool hi = (s >> 1) & 1;
  const float r = BoxMuller<float>::radius(hi ? c2 : c0);
  const float u = BoxMuller<float>::turn(hi ? c3 : c1);
  return r * __builtin_amdgcn_sinf((s & 1) ? u : u + 0.25f);
}

template <>
__device__ __forceinline__ float philox_normal1<float>(uint64_t seed, uint64_t real, uint64_t s)
{
  const uint64_t blk = s >> 2;
  uint32_t c[4] = { (uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)real, (uint32_t)(real >> 32) };
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return box_muller_of_sample<float>(c[0], c[1], c[2], c[3], s);
}

// ---------------------------------------------------------------------------
// Covariance model on the device (FP64), covariance.hh:18-823.
// ---------------------------------------------------------------------------
struct CovParams
{
  int dim;
  int covariance;
  int aniso;         // effective geometry kind
  double variance;
  double m[9];       // 1/l, 1/l_i or A^{-1} (row-major)
  double nu, sqrt2nu, matern_c; // Matern: 2^{1-nu}/Gamma(nu)
  double gamma, alpha, beta;
  double nugget;
  // embedding grid geometry (matrix.hh:615-638)
  double h[3], wrap_half[3], wrap_len[3];
  int factor;
  // merge embedding (matrix.hh:652-775): taper kind, combination, recursion
  // count, taper interval in units of the extensions, extensions
  int period;         // enum pf_periodization: 0 classical, 1 merge, 2 fold
  int sigmoid;        // enum pf_sigmoid
  int radial;         // radial taper (combine radial/radialPlus and sqrt(dim) < factor / 2)
  int radial_plus;    // the remainder takes the covariance at the centre of the extended domain
  int recursions;
  double sig_start, sig_end, ext[3];
  double fold_beta;   // fold: the taper reaches zero at this norm (matrix.hh:831-833)
  double border[3];   // cofold: ext_i / extended cells_i, the border band (matrix.hh:936-941)
};

// Modified Bessel K_nu(x) (Temme series for x < 2, Steed's CF2 otherwise,
// forward recurrence in nu) -- the same method as the oracle's bessel_k.
__device__ inline double bessel_k(double nu, double x)
{
  const double EPS = 1e-16, PI = 3.14159265358979323846;
  int nl = (int)(nu + 0.5);
  double xmu = nu - nl, xmu2 = xmu * xmu;
  double xi = 1.0 / x, xi2 = 2.0 * xi;
  double rkmu, rk1;
  if (x < 2.0) {
    double x2 = 0.5 * x, pimu = PI * xmu;
    double fact = (fabs(pimu) < EPS ? 1.0 : pimu / sin(pimu));
    double d = -log(x2), e = xmu * d;
    double fact2 = (fabs(e) < EPS ? 1.0 : sinh(e) / e);
    double gampl = 1.0 / tgamma(1.0 + xmu), gammi = 1.0 / tgamma(1.0 - xmu);
    double gam1 = (fabs(xmu) < 1e-2)
                    ? -(0.5772156649015329 - 0.0420026350340952 * xmu2 -
                        0.0421977345555443 * xmu2 * xmu2 + 0.0072189432466630 * xmu2 * xmu2 * xmu2)
                    : (gammi - gampl) / (2.0 * xmu);
    double gam2 = 0.5 * (gammi + gampl);
    double ff = fact * (gam1 * cosh(e) + gam2 * fact2 * d);
    double sum = ff;
    e = exp(e);
    double p = 0.5 * e / gampl, q = 0.5 / (e * gammi), c = 1.0;
    d = x2 * x2;
    double sum1 = p;
    for (int i = 1; i <= 10000; i++) {
      ff = (i * ff + p + q) / (i * i - xmu2);
      c *= (d / i);
      p /= (i - xmu);
      q /= (i + xmu);
      double del = c * ff;
      sum += del;
      double del1 = c * (p - i * ff);
      sum1 += del1;
      if (fabs(del) < fabs(sum) * EPS)
        break;
    }
    rkmu = sum;
    rk1 = sum1 * xi2;
  } else {
    double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, delh = d;
    double q1 = 0.0, q2 = 1.0, a1 = 0.25 - xmu2, q = a1, c = a1, a = -a1;
    double s = 1.0 + q * delh;
    for (int i = 2; i <= 10000; i++) {
      a -= 2 * (i - 1);
      c = -a * c / i;
      double qnew = (q1 - b * q2) / a;
      q1 = q2;
      q2 = qnew;
      q += c * qnew;
      b += 2.0;
      d = 1.0 / (b + a * d);
      delh = (b * d - 1.0) * delh;
      h += delh;
      double dels = q * delh;
      s += dels;
      if (fabs(dels / s) < EPS)
        break;
    }
    h = a1 * h;
    rkmu = sqrt(PI / (2.0 * x)) * exp(-x) / s;
    rk1 = rkmu * (xmu + x + 0.5 - h) * xi;
  }
  for (int i = 1; i <= nl; i++) {
    double rktemp = (xmu + i) * xi2 * rk1 + rkmu;
    rkmu = rk1;
    rk1 = rktemp;
  }
  return rkmu;
}

// Covariance at the physical offset coord: geometry transform, covariance
// functor (+ nugget at the origin).
// COV >= 0: the family is fixed at compile time (p.covariance is not read).
template <int COV = -1>
__device__ inline double cov_at(const CovParams& p, const double coord[3])
{
#pragma clang fp contract(off)
  double t[3] = { 0., 0., 0. };
  if (p.aniso == PF_ANISO_NONE) {
    for (int i = 0; i < p.dim; i++)
      t[i] = p.m[0] * coord[i];
  } else if (p.aniso == PF_ANISO_AXIPARALLEL) {
    for (int i = 0; i < p.dim; i++)
      t[i] = p.m[i] * coord[i];
  } else {
    for (int i = 0; i < p.dim; i++) {
      t[i] = 0.;
      for (int j = 0; j < p.dim; j++)
        t[i] += p.m[i * p.dim + j] * coord[j];
    }
  }
  double sum = 0.;
  for (int i = 0; i < p.dim; i++)
    sum += t[i] * t[i];
  const double hh = sqrt(sum);
  const double var = p.variance;
  double v = 0.;
  switch (COV >= 0 ? COV : p.covariance) {
    case PF_COV_SPHERICAL:
      if (hh > 1.)
        v = 0.;
      else if (p.dim == 3)
        v = var * (1. - 1.5 * hh + 0.5 * pow(hh, 3.));
      else if (p.dim == 2)
        v = var * (1. - 2. / 3.14159265358979323846 * (hh * sqrt(1. - pow(hh, 2.)) + asin(hh)));
      else
        v = var * (1. - hh);
      break;
    case PF_COV_EXPONENTIAL:
    case PF_COV_CUSTOM:
      v = var * exp(-hh);
      break;
    case PF_COV_GAMMA_EXPONENTIAL:
      v = var * exp(-pow(hh, p.gamma));
      break;
    case PF_COV_GAUSSIAN:
      v = var * exp(-hh * hh);
      break;
    case PF_COV_SEPARABLE_EXPONENTIAL: {
      double s = 0.;
      for (int i = 0; i < p.dim; i++)
        s += fabs(t[i]);
      v = var * exp(-s);
      break;
    }
    case PF_COV_MATERN:
      if (hh < 1e-10)
        v = var;
      else
        v = var * p.matern_c * pow(p.sqrt2nu * hh, p.nu) * bessel_k(p.nu, p.sqrt2nu * hh);
      break;
    case PF_COV_MATERN32:
      v = var * (1. + sqrt(3.) * hh) * exp(-sqrt(3.) * hh);
      break;
    case PF_COV_MATERN52:
      v = var * (1. + sqrt(5.) * hh + 5. / 3. * hh * hh) * exp(-sqrt(5.) * hh);
      break;
    case PF_COV_DAMPED_OSCILLATION:
      v = (p.dim == 3) ? var * exp(-hh) * cos(hh / sqrt(3.)) : var * exp(-hh) * cos(hh);
      break;
    case PF_COV_CAUCHY:
      v = var * pow(1. + pow(hh, 2.), -3.);
      break;
    case PF_COV_GENERALIZED_CAUCHY:
      v = var * pow(1. + pow(hh, p.alpha), -p.beta);
      break;
    case PF_COV_CUBIC:
      v = (hh > 1.) ? 0.
                    : var * (1. - 7. * pow(hh, 2.) + 8.75 * pow(hh, 3.) - 3.5 * pow(hh, 5.) +
                             0.75 * pow(hh, 7.));
      break;
    case PF_COV_WHITE_NOISE:
      v = var;
      for (int i = 0; i < p.dim; i++)
        if (fabs(t[i]) > 1e-10)
          v = 0.;
      break;
    default:
      v = 0.;
  }
  if (p.nugget != 0.) {
    bool origin = true;
    for (int i = 0; i < p.dim; i++)
      if (t[i] != 0.)
        origin = false;
    if (origin)
      v += p.nugget;
  }
  return v;
}

// Sigmoid tapers of the merge embedding (covariance.hh:147-290): 1 for |x|
// below `one`, 0 beyond `zero`. Clamping: linear in between; smooth:
// theta(x1) / (theta(x1) + theta(x2)) with theta(y) = exp(-1/y); smoothstep:
// the Hermite interpolation matching `rec` derivatives (the regularized
// incomplete beta function I_t(rec+1, rec+1) of the clamped value t), by the
// reference's recursion.
__host__ __device__ inline double sigmoid_clamp(double one, double zero, double x)
{
#pragma clang fp contract(off)
  const double ax = fabs(x);
  return ax < one ? 1. : (ax > zero ? 0. : (zero - ax) / (zero - one));
}

__host__ __device__ inline double sigmoid_taper(const CovParams& p, double one, double zero, double x)
{
#pragma clang fp contract(off)
  if (p.sigmoid == PF_SIGMOID_SMOOTH) {
    const double ax = fabs(x);
    if (ax < one)
      return 1.;
    if (ax > zero)
      return 0.;
    const double x1 = (zero - ax) / (zero - one), x2 = (ax - one) / (zero - one);
    const double t1 = x1 > 0. ? exp(-1. / x1) : 0., t2 = x2 > 0. ? exp(-1. / x2) : 0.;
    return t1 / (t1 + t2);
  }
  const double x2 = 2. * sigmoid_clamp(one, zero, x) - 1.;
  double norm = 1., add = 1., out = 1.;
  for (int i = 1; i <= p.recursions; i++) {
    out *= 2. * i;
    add *= 1 - x2 * x2;
    out += add;
    out /= 2. * i + 1.;
    norm *= 2. * i / (2. * i + 1.);
  }
  out *= x2 / norm;
  out += 1.;
  out /= 2.;
  return out;
}

// Entry of the merge embedding at grid position idx (no wrap),
// matrix.hh:652-775: the 2^dim mirror images x_i -> ext_i * factor - x_i each
// contribute the covariance times a taper (radial in the extension-scaled
// norm, or the product of per-axis tapers); what the tapers leave takes the
// covariance at the centre of the extended domain (radialPlus) or nothing.
template <int COV = -1>
__device__ inline double merged_cov(const CovParams& p, const int64_t idx[3])
{
#pragma clang fp contract(off)
  const double eps = 1e-10;
  const double sqrt_dim = sqrt((double)p.dim);
  const double factor = (double)p.factor;
  double value = 0., remainder = 1.;
  for (int j = 0; j < (1 << p.dim); j++) {
    double mirror[3] = { 0., 0., 0. };
    for (int i = 0; i < p.dim; i++) {
      const double x = (double)idx[i] * p.h[i];
      mirror[i] = ((j >> i) & 1) ? p.ext[i] * factor - x : x;
    }
    double damp = 1.;
    if (p.radial) {
      double r = 0.;
      for (int i = 0; i < p.dim; i++)
        r += (mirror[i] * mirror[i]) / (p.ext[i] * p.ext[i]);
      r = sqrt(r);
      damp = sigmoid_taper(p, p.sig_start * sqrt_dim - eps,
                           factor - (factor - p.sig_end) * sqrt_dim + eps, r);
    } else {
      for (int i = 0; i < p.dim; i++)
        damp *= sigmoid_taper(p, p.ext[i] * p.sig_start - eps, p.ext[i] * p.sig_end + eps,
                              mirror[i]);
    }
    if (damp < eps)
      continue;
    remainder -= damp;
    value += cov_at<COV>(p, mirror) * damp;
  }
  if (p.radial_plus) {
    double c[3] = { 0., 0., 0. };
    for (int i = 0; i < p.dim; i++)
      c[i] = p.ext[i] * factor / 2.;
    // the centre value without a nugget (the offset is never the origin)
    value += cov_at<COV>(p, c) * remainder;
  }
  return value;
}

// Integral of f over [lo, hi]: composite 5-point Gauss-Legendre over 16
// panels (nodes 0, +-sqrt(5 -+ 2 sqrt(10/7)) / 3, weights 128/225,
// (322 +- 13 sqrt(70)) / 900). The tapers below are smooth on the intervals
// it is used on, and the rule integrates them to round-off for the recursion
// counts in use.
template <typename F>
__host__ __device__ inline double gauss_legendre_16x5(F&& f, double lo, double hi)
{
  const double r10 = sqrt(10. / 7.), r70 = sqrt(70.);
  const double x1 = sqrt(5. - 2. * r10) / 3., x2 = sqrt(5. + 2. * r10) / 3.;
  const double w0 = 128. / 225., w1 = (322. + 13. * r70) / 900., w2 = (322. - 13. * r70) / 900.;
  const int panels = 16;
  const double half = 0.5 * (hi - lo) / panels;
  double acc = 0.;
  for (int k = 0; k < panels; k++) {
    const double c = lo + (2 * k + 1) * half;
    acc += w0 * f(c);
    acc += w1 * (f(c - half * x1) + f(c + half * x1));
    acc += w2 * (f(c - half * x2) + f(c + half * x2));
  }
  return acc * half;
}

// Integral of the taper sigmoid(1, fold_beta, x) over x in [1, norm]. The
// taper is zero from fold_beta on, so the integral stops there. The reference
// calls gsl_integration_qng with absolute tolerance 1e-6 and relative
// tolerance 1e-3 on [1, norm] (matrix.hh:866-868).
__device__ inline double fold_integral(const CovParams& p, double norm)
{
  const double b = p.fold_beta;
  const double hi = norm < b ? norm : b;
  if (!(hi > 1.))
    return 0.;
  return gauss_legendre_16x5([&](double x) { return sigmoid_taper(p, 1., b, x); }, 1., hi);
}

// Cofold (matrix.hh:946-961): integral of the codomain taper
// sigmoid(0, beta - alpha, beta - t) over t in [alpha, x], alpha <= x <= beta.
__host__ __device__ inline double cofold_integral(const CovParams& p, double alpha, double beta,
                                                  double x)
{
  if (!(x > alpha))
    return 0.;
  const double w = beta - alpha;
  return gauss_legendre_16x5([&](double t) { return sigmoid_taper(p, 0., w, beta - t); }, alpha,
                             x);
}

// Covariance at the embedding grid position idx (integer cell offsets):
// coordinates idx*h wrapped by the classical rule (matrix.hh:615-638), then
// folded, or the merge embedding. PER: 0 classical, 1 merge, 2 fold, -1 as
// p.period says (a kernel compiled for one periodization keeps the others'
// code and registers out).
template <int COV = -1, int PER = -1>
__device__ inline double embedded_cov(const CovParams& p, const int64_t idx[3])
{
#pragma clang fp contract(off)
  if constexpr (PER == 1)
    return merged_cov<COV>(p, idx);
  if constexpr (PER == -1) {
    if (p.period == 1)
      return merged_cov<COV>(p, idx);
  }
  double coord[3] = { 0., 0., 0. };
  for (int i = 0; i < p.dim; i++) {
    coord[i] = (double)idx[i] * p.h[i];
    if (coord[i] > p.wrap_half[i])
      coord[i] -= p.wrap_len[i];
  }
  // fold (matrix.hh:850-881): coordinates whose extension-scaled norm
  // exceeds 1 move towards the origin, to norm 1 + the taper's integral
  if (PER == 2 || (PER == -1 && p.period == 2)) {
    double norm = 0.;
    for (int i = 0; i < p.dim; i++)
      norm += (coord[i] * coord[i]) / (p.ext[i] * p.ext[i]);
    norm = sqrt(norm / (double)p.dim);
    if (norm > 1.) {
      const double val = fold_integral(p, norm) + 1.;
      for (int i = 0; i < p.dim; i++)
        coord[i] *= val / norm;
    }
  }
  return cov_at<COV>(p, coord);
}

// Value transforms of mutators.hh:35-163 (sign threshold is a parameter,
// 0 in the reference).
// value_transform with the kind fixed at compile time
template <typename T, int K>
__device__ __forceinline__ T value_transform_k(T v, T param)
{
  if constexpr (K == PF_TRANSFORM_LOGNORMAL)
    return exp(v);
  else if constexpr (K == PF_TRANSFORM_FOLDEDNORMAL)
    return fabs(v);
  else if constexpr (K == PF_TRANSFORM_SIGN)
    return (v < param) ? T(-1) : T(1);
  else if constexpr (K == PF_TRANSFORM_BOXCOX)
    return pow(v * param + T(1), -param);
  else
    return v;
}

template <typename T>
__device__ __forceinline__ T value_transform(int kind, T v, T param)
{
  switch (kind) {
    case PF_TRANSFORM_LOGNORMAL:
      return exp(v);
    case PF_TRANSFORM_FOLDEDNORMAL:
      return fabs(v);
    case PF_TRANSFORM_SIGN:
      return (v < param) ? T(-1) : T(1);
    case PF_TRANSFORM_BOXCOX:
      return pow(v * param + T(1), -param);
    default:
      return v;
  }
}

} // namespace pfamd
