// pf_oracle.cc -- FP64 CPU restatement of parafields-core's circulant-embedding
// engine. TEST INFRASTRUCTURE ONLY: it is the checker that the HIP library is
// compared against (tests, the smoke check and the CPU-baseline leg of the
// benchmark). The product never links, loads or calls it.
//
// Every function follows the reference loop it restates (file:line relative
// to the reference repository root, include/parafields/...). Arithmetic is
// RF = double throughout, as in a reference build with RF = double. The
// transforms that FFTW performs in the reference are computed with this
// file's own mixed-radix FFT using FFTW's documented definitions:
//   r2c   : X_k = sum_x x_x e^{-2 pi i k x / N}, k < N/2+1   (unnormalised)
//   c2r   : c2c along all but the x axis, then a 1-D c2r along x that reads
//           only Re of the DC and Nyquist entries                (unnormalised)
//   REDFT00 (DCT-I), RODFT00 (DST-I): FFTW's definitions, unnormalised.
//
// Parity status of each function (what pins it) is listed in the tests that
// exercise it (tests/test_oracle.py).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../include/parafields_amd.h"

namespace {

using cd = std::complex<double>;
using Index = long long;

int g_threads = 1;

// ----------------------------------------------------------------------------
// Threading helper: split [0, n) over g_threads workers.
// ----------------------------------------------------------------------------
void parallel_for(Index n, const std::function<void(Index, Index)>& body)
{
  int nt = std::max(1, std::min<int>(g_threads, (int)std::min<Index>(n, 1 << 20)));
  if (nt == 1 || n < 64) {
    body(0, n);
    return;
  }
  std::vector<std::thread> pool;
  Index chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    Index b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e)
      break;
    pool.emplace_back([&body, b, e] { body(b, e); });
  }
  for (auto& th : pool)
    th.join();
}

// ----------------------------------------------------------------------------
// Mixed-radix recursive FFT (decimation in time), unnormalised.
// sign = -1: forward (e^{-i}), +1: backward (e^{+i}).
// ----------------------------------------------------------------------------
struct FFTPlan
{
  int n;
  std::vector<cd> tw; // e^{-2 pi i j / n}, j < n
  explicit FFTPlan(int n_)
    : n(n_)
    , tw(n_)
  {
    for (int j = 0; j < n; j++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * j / n;
      tw[j] = cd((double)cosl(a), (double)sinl(a));
    }
  }
};

int smallest_factor(int n)
{
  if (n % 4 == 0)
    return 4;
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0)
      return p;
  return n;
}

void fft_rec(const cd* in, Index stride, cd* out, int n, int sign, const FFTPlan& plan,
             int twstep)
{
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = smallest_factor(n);
  const int m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + q * stride, stride * p, out + q * m, m, sign, plan, twstep * p);
  cd tbuf[16], rbuf[16];
  std::vector<cd> tv, rv;
  cd* t = tbuf;
  cd* r = rbuf;
  if (p > 16) {
    tv.resize(p);
    rv.resize(p);
    t = tv.data();
    r = rv.data();
  }
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      cd w = plan.tw[(Index)q * k * twstep % plan.n];
      if (sign > 0)
        w = std::conj(w);
      t[q] = w * out[q * m + k];
    }
    for (int s = 0; s < p; s++) {
      cd acc = 0;
      for (int q = 0; q < p; q++) {
        cd w = plan.tw[(Index)((q * s) % p) * m * twstep % plan.n];
        if (sign > 0)
          w = std::conj(w);
        acc += t[q] * w;
      }
      r[s] = acc;
    }
    for (int s = 0; s < p; s++)
      out[k + m * s] = r[s];
  }
}

void fft1d(cd* data, int n, int sign)
{
  static thread_local std::vector<std::pair<int, FFTPlan*>> cache;
  FFTPlan* plan = nullptr;
  for (auto& e : cache)
    if (e.first == n)
      plan = e.second;
  if (!plan) {
    plan = new FFTPlan(n);
    cache.push_back({ n, plan });
  }
  std::vector<cd> in(data, data + n);
  fft_rec(in.data(), 1, data, n, sign, *plan, 1);
}

// Apply a 1-D complex transform along axis `ax` of a complex array with
// dims d[0..2] (x fastest). Only lines are transformed; the array is in place.
void c2c_axis(cd* a, const Index d[3], int ax, int sign)
{
  Index stride = 1;
  for (int i = 0; i < ax; i++)
    stride *= d[i];
  const int n = (int)d[ax];
  const Index total = d[0] * d[1] * d[2];
  const Index lines = total / n;
  parallel_for(lines, [&](Index b, Index e) {
    std::vector<cd> line(n);
    for (Index l = b; l < e; l++) {
      Index inner = l % stride, outer = l / stride;
      Index base = inner + outer * stride * n;
      for (int j = 0; j < n; j++)
        line[j] = a[base + j * stride];
      fft1d(line.data(), n, sign);
      for (int j = 0; j < n; j++)
        a[base + j * stride] = line[j];
    }
  });
}

// 1-D real-to-complex along x of real rows of length n0 (input pitch ip,
// output pitch h = n0/2+1 complex), `rows` rows.
void r2c_x(const double* in, Index ip, cd* out, int n0, Index rows)
{
  const int h = n0 / 2 + 1;
  parallel_for(rows, [&](Index b, Index e) {
    std::vector<cd> line(n0);
    for (Index r = b; r < e; r++) {
      for (int j = 0; j < n0; j++)
        line[j] = in[r * ip + j];
      fft1d(line.data(), n0, -1);
      for (int k = 0; k < h; k++)
        out[r * h + k] = line[k];
    }
  });
}

// 1-D complex-to-real along x (FFTW c2r semantics: Im of the DC and Nyquist
// entries are ignored, the rest is Hermitian-extended), unnormalised.
void c2r_x(const cd* in, double* out, Index op, int n0, Index rows)
{
  const int h = n0 / 2 + 1;
  parallel_for(rows, [&](Index b, Index e) {
    std::vector<cd> line(n0);
    for (Index r = b; r < e; r++) {
      const cd* z = in + r * h;
      line[0] = cd(z[0].real(), 0.);
      for (int k = 1; k < h; k++) {
        line[k] = z[k];
        if (n0 - k != k)
          line[n0 - k] = std::conj(z[k]);
      }
      if (n0 % 2 == 0)
        line[n0 / 2] = cd(z[n0 / 2].real(), 0.);
      fft1d(line.data(), n0, +1);
      for (int j = 0; j < n0; j++)
        out[r * op + j] = line[j].real();
    }
  });
}

// FFTW REDFT00 along an axis of a real array (n = P points):
// Y_k = X_0 + (-1)^k X_{n-1} + 2 sum_{j=1}^{n-2} X_j cos(pi j k / (n-1)),
// computed as the DFT of the even extension of length 2(n-1).
void redft00_axis(double* a, const Index d[3], int ax)
{
  Index stride = 1;
  for (int i = 0; i < ax; i++)
    stride *= d[i];
  const int n = (int)d[ax];
  const Index total = d[0] * d[1] * d[2];
  const Index lines = total / n;
  if (n == 1)
    return;
  const int N = 2 * (n - 1);
  parallel_for(lines, [&](Index b, Index e) {
    std::vector<cd> line(N);
    for (Index l = b; l < e; l++) {
      Index inner = l % stride, outer = l / stride;
      Index base = inner + outer * stride * n;
      for (int j = 0; j < n; j++)
        line[j] = a[base + j * stride];
      for (int j = 1; j < n - 1; j++)
        line[N - j] = line[j];
      fft1d(line.data(), N, -1);
      for (int k = 0; k < n; k++)
        a[base + k * stride] = line[k].real();
    }
  });
}

// FFTW RODFT00 along an axis of length n (interior points only):
// Y_k = 2 sum_{j=0}^{n-1} X_j sin(pi (j+1)(k+1) / (n+1)),
// computed from the odd extension of length 2(n+1).
void rodft00_line(const double* x, double* y, int n)
{
  const int N = 2 * (n + 1);
  std::vector<cd> line(N, 0.);
  for (int j = 0; j < n; j++) {
    line[j + 1] = x[j];
    line[N - 1 - j] = -x[j];
  }
  fft1d(line.data(), N, -1);
  // DFT of odd sequence: -2i sum x_j sin(...)  =>  Y_k = -Im(F_{k+1})
  for (int k = 0; k < n; k++)
    y[k] = -line[k + 1].imag();
}

// ----------------------------------------------------------------------------
// Configuration-derived sizes, restating RandomFieldTraits::update,
// fieldtraits.hh:217-314 (commSize = 1).
// ----------------------------------------------------------------------------
struct Geo
{
  int dim;
  Index n[3], N[3], H;
  double ext[3], h[3];
  int factor;
  bool transposed;
  Index Next, domain;
};

Geo make_geo(const pf_desc& d)
{
  if (d.dim < 1 || d.dim > 3)
    throw std::runtime_error("dim must be 1, 2 or 3");
  if (d.dim == 1 && (d.matrix_backend != PF_MATRIX_DFT || d.field_backend != PF_FIELD_DFT))
    throw std::runtime_error("1-D fields use DFTMatrixBackend and DFTFieldBackend");
  Geo g;
  g.dim = d.dim;
  g.factor = d.periodic ? 1 : d.embedding_factor;
  if (g.factor < 1)
    throw std::runtime_error("embedding.factor must be >= 1");
  if (d.periodization < PF_PERIOD_CLASSICAL || d.periodization > PF_PERIOD_COFOLD)
    throw std::runtime_error(
      "stochastic.periodization must be \"classical\", \"merge\", \"fold\", or \"cofold\"");
  if (d.sigmoid != PF_SIGMOID_SMOOTHSTEP && d.sigmoid != PF_SIGMOID_SMOOTH)
    throw std::runtime_error("embedding.sigmoid must be \"smooth\" or \"smoothstep\"");
  if (d.sigmoid_combine < PF_COMBINE_RADIAL_PLUS || d.sigmoid_combine > PF_COMBINE_TENSOR)
    throw std::runtime_error("sigmoidCombine must be tensor, radial or radialPlus");
  g.Next = 1;
  g.domain = 1;
  for (int i = 0; i < 3; i++) {
    if (i < d.dim) {
      g.n[i] = d.cells[i];
      g.ext[i] = d.extensions[i];
      g.h[i] = d.extensions[i] / d.cells[i];
      g.N[i] = (Index)g.factor * d.cells[i];
    } else {
      g.n[i] = g.N[i] = 1;
      g.ext[i] = g.h[i] = 1.;
    }
    g.Next *= g.N[i];
    g.domain *= g.n[i];
  }
  g.H = g.N[0] / 2 + 1;
  // fieldtraits.hh:229-267: default dim > 1; 2D with an R2C backend -> false.
  if (d.dim == 1)
    g.transposed = false; // fieldtraits.hh:229-233
  else if (d.transposed >= 0)
    g.transposed = d.transposed != 0;
  else if (d.dim == 2)
    g.transposed = !(d.matrix_backend == PF_MATRIX_R2C || d.field_backend == PF_FIELD_R2C);
  else
    g.transposed = true;
  return g;
}

// ----------------------------------------------------------------------------
// Covariance functors, covariance.hh:296-823, and geometry, :18-138.
// ----------------------------------------------------------------------------
struct Geometry
{
  int dim, kind;
  double m[9];
  explicit Geometry(const pf_desc& d)
  {
    dim = d.dim;
    kind = d.anisotropy;
    if (kind == PF_ANISO_NONE || d.covariance == PF_COV_CUSTOM) {
      // ScaledIdentityMatrix, covariance.hh:18-44 (custom-iso always uses it,
      // matrix.hh:261-264)
      kind = PF_ANISO_NONE;
      m[0] = 1. / d.corr_length[0];
    } else if (kind == PF_ANISO_AXIPARALLEL) {
      // DiagonalMatrix, covariance.hh:49-74
      for (int i = 0; i < dim; i++)
        m[i] = 1. / d.corr_length[i];
    } else if (kind == PF_ANISO_GEOMETRIC) {
      // GeneralMatrix, covariance.hh:79-138: explicit cofactor inverse
      const double* c = d.corr_length;
      if (dim == 1) {
        m[0] = 1. / c[0];
      } else if (dim == 3) {
        m[0] = c[4] * c[8] - c[5] * c[7];
        m[1] = c[2] * c[7] - c[1] * c[8];
        m[2] = c[1] * c[5] - c[2] * c[4];
        m[3] = c[5] * c[6] - c[3] * c[8];
        m[4] = c[0] * c[8] - c[2] * c[6];
        m[5] = c[2] * c[3] - c[0] * c[5];
        m[6] = c[3] * c[7] - c[4] * c[6];
        m[7] = c[1] * c[6] - c[0] * c[7];
        m[8] = c[0] * c[4] - c[1] * c[3];
        const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) -
                           c[1] * (c[3] * c[8] - c[5] * c[6]) +
                           c[2] * (c[3] * c[7] - c[4] * c[6]);
        for (int i = 0; i < 9; i++)
          m[i] /= det;
      } else {
        m[0] = c[3];
        m[1] = -c[1];
        m[2] = -c[2];
        m[3] = c[0];
        const double det = c[0] * c[3] - c[1] * c[2];
        for (int i = 0; i < 4; i++)
          m[i] /= det;
      }
    } else
      throw std::runtime_error("stochastic.anisotropy must be none, axiparallel or geometric");
  }
  void transform(const double* x, double* y) const
  {
    if (kind == PF_ANISO_NONE)
      for (int i = 0; i < dim; i++)
        y[i] = m[0] * x[i];
    else if (kind == PF_ANISO_AXIPARALLEL)
      for (int i = 0; i < dim; i++)
        y[i] = m[i] * x[i];
    else
      for (int i = 0; i < dim; i++) {
        y[i] = 0.;
        for (int j = 0; j < dim; j++)
          y[i] += m[i * dim + j] * x[j];
      }
  }
};

// Modified Bessel function K_nu(x), x > 0, nu >= 0: Temme's series for
// |x| <= 2 and Steed's continued fraction CF2 for x > 2 (the method of
// Numerical Recipes' bessik / Temme 1975), followed by forward recurrence.
// Used for the general Matern family, which the reference evaluates with
// GSL's gsl_sf_bessel_Knu (covariance.hh:539-555).
double bessel_k(double nu, double x)
{
  const double EPS = 1e-16, FPMIN = 1e-300, PI = 3.14159265358979323846;
  const int MAXIT = 100000;
  int nl = (int)(nu + 0.5);
  double xmu = nu - nl, xmu2 = xmu * xmu;
  double xi = 1.0 / x, xi2 = 2.0 * xi;
  double rkmu, rk1;
  if (x < 2.0) {
    double x2 = 0.5 * x, pimu = PI * xmu;
    double fact = (std::fabs(pimu) < EPS ? 1.0 : pimu / std::sin(pimu));
    double d = -std::log(x2), e = xmu * d;
    double fact2 = (std::fabs(e) < EPS ? 1.0 : std::sinh(e) / e);
    // gamma1 = (1/Gamma(1-xmu) - 1/Gamma(1+xmu)) / (2 xmu), gamma2 = (..+..)/2
    double gampl = 1.0 / std::tgamma(1.0 + xmu), gammi = 1.0 / std::tgamma(1.0 - xmu);
    // gamma1 -> -(c2 + c4 mu^2 + c6 mu^4 + c8 mu^6) near mu = 0, with c_k the
    // Taylor coefficients of 1/Gamma(z) (Abramowitz & Stegun 6.1.34)
    double gam1 = (std::fabs(xmu) < 1e-2)
                    ? -(0.5772156649015329 - 0.0420026350340952 * xmu2 -
                        0.0421977345555443 * xmu2 * xmu2 + 0.0072189432466630 * xmu2 * xmu2 * xmu2)
                    : (gammi - gampl) / (2.0 * xmu);
    double gam2 = 0.5 * (gammi + gampl);
    double ff = fact * (gam1 * std::cosh(e) + gam2 * fact2 * d);
    double sum = ff;
    e = std::exp(e);
    double p = 0.5 * e / gampl, q = 0.5 / (e * gammi), c = 1.0;
    d = x2 * x2;
    double sum1 = p;
    int i;
    for (i = 1; i <= MAXIT; i++) {
      ff = (i * ff + p + q) / (i * i - xmu2);
      c *= (d / i);
      p /= (i - xmu);
      q /= (i + xmu);
      double del = c * ff;
      sum += del;
      double del1 = c * (p - i * ff);
      sum1 += del1;
      if (std::fabs(del) < std::fabs(sum) * EPS)
        break;
    }
    rkmu = sum;
    rk1 = sum1 * xi2;
  } else {
    double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, delh = d;
    double q1 = 0.0, q2 = 1.0, a1 = 0.25 - xmu2, q = a1, c = a1, a = -a1;
    double s = 1.0 + q * delh;
    int i;
    for (i = 2; i <= MAXIT; i++) {
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
      if (std::fabs(dels / s) < EPS)
        break;
    }
    h = a1 * h;
    rkmu = std::sqrt(PI / (2.0 * x)) * std::exp(-x) / s;
    rk1 = rkmu * (xmu + x + 0.5 - h) * xi;
  }
  (void)FPMIN;
  for (int i = 1; i <= nl; i++) {
    double rktemp = (xmu + i) * xi2 * rk1 + rkmu;
    rkmu = rk1;
    rk1 = rktemp;
  }
  return rkmu;
}

struct Covariance
{
  pf_desc d;
  explicit Covariance(const pf_desc& d_)
    : d(d_)
  {
    if (d.covariance == PF_COV_GAMMA_EXPONENTIAL && (d.exp_gamma < 0. || d.exp_gamma > 2.))
      throw std::runtime_error("exponent gamma has to be between 0 and 2");
    if (d.covariance == PF_COV_GENERALIZED_CAUCHY &&
        (d.cauchy_alpha <= 0. || d.cauchy_alpha > 2. || d.cauchy_beta <= 0.))
      throw std::runtime_error("generalized Cauchy parameters out of range");
    if (d.covariance == PF_COV_MATERN && d.matern_nu < 0.)
      throw std::runtime_error("matern nu has to be positive");
    if (d.covariance == PF_COV_CUBIC && d.dim == 3)
      throw std::runtime_error("cubic covariance only applicable in 1D or 2D");
    if (d.covariance < 0 || d.covariance > PF_COV_CUSTOM)
      throw std::runtime_error("covariance structure not known");
  }
  double operator()(const double* x) const
  {
    const int dim = d.dim;
    const double var = d.variance;
    double sum = 0.;
    for (int i = 0; i < dim; i++)
      sum += x[i] * x[i];
    const double h = std::sqrt(sum);
    double v = 0.;
    switch (d.covariance) {
      case PF_COV_SPHERICAL:
        if (h > 1.)
          v = 0.;
        else if (dim == 3)
          v = var * (1. - 1.5 * h + 0.5 * std::pow(h, 3));
        else if (dim == 2)
          v = var * (1. - 2. / 3.14159265358979323846 *
                            (h * std::sqrt(1. - std::pow(h, 2)) + std::asin(h)));
        else
          v = var * (1. - h); // covariance.hh:331-335
        break;
      case PF_COV_EXPONENTIAL:
      case PF_COV_CUSTOM: // custom-iso nugget model: exponential + nugget
        v = var * std::exp(-h);
        break;
      case PF_COV_GAMMA_EXPONENTIAL:
        v = var * std::exp(-std::pow(h, d.exp_gamma));
        break;
      case PF_COV_GAUSSIAN:
        v = var * std::exp(-h * h);
        break;
      case PF_COV_SEPARABLE_EXPONENTIAL: {
        double s = 0.;
        for (int i = 0; i < dim; i++)
          s += std::fabs(x[i]);
        v = var * std::exp(-s);
        break;
      }
      case PF_COV_MATERN: {
        const double nu = d.matern_nu;
        if (h < 1e-10)
          v = var;
        else {
          const double s2n = std::sqrt(2. * nu);
          v = var * std::pow(2., 1. - nu) / std::tgamma(nu) * std::pow(s2n * h, nu) *
              bessel_k(nu, s2n * h);
        }
        break;
      }
      case PF_COV_MATERN32:
        v = var * (1. + std::sqrt(3.) * h) * std::exp(-std::sqrt(3.) * h);
        break;
      case PF_COV_MATERN52:
        v = var * (1. + std::sqrt(5.) * h + 5. / 3. * h * h) * std::exp(-std::sqrt(5.) * h);
        break;
      case PF_COV_DAMPED_OSCILLATION:
        if (dim == 3)
          v = var * std::exp(-h) * std::cos(h / std::sqrt(3.));
        else
          v = var * std::exp(-h) * std::cos(h);
        break;
      case PF_COV_CAUCHY:
        v = var * std::pow(1. + std::pow(h, 2), -3);
        break;
      case PF_COV_GENERALIZED_CAUCHY:
        v = var * std::pow(1. + std::pow(h, d.cauchy_alpha), -d.cauchy_beta);
        break;
      case PF_COV_CUBIC:
        if (h > 1.)
          v = 0.;
        else
          v = var * (1. - 7. * std::pow(h, 2) + 8.75 * std::pow(h, 3) - 3.5 * std::pow(h, 5) +
                     0.75 * std::pow(h, 7));
        break;
      case PF_COV_WHITE_NOISE: {
        v = var;
        for (int i = 0; i < dim; i++)
          if (std::fabs(x[i]) > 1e-10)
            v = 0.;
        break;
      }
      default:
        throw std::runtime_error("covariance structure not known");
    }
    return v;
  }
};

// Covariance at the physical offset coord (geometry transform, family,
// nugget at the origin).
double cov_at(const Geo& g, const Geometry& geo, const Covariance& cov, const pf_desc& d,
              const double coord[3])
{
  double tc[3] = { 0, 0, 0 };
  geo.transform(coord, tc);
  double v = cov(tc);
  if (d.nugget != 0.) {
    bool origin = true;
    for (int i = 0; i < g.dim; i++)
      if (tc[i] != 0.)
        origin = false;
    if (origin)
      v += d.nugget;
  }
  return v;
}

// Sigmoids of the merge embedding, covariance.hh:147-290: value 1 for |x|
// below oneEdge, 0 beyond zeroEdge.
// ClampingSigmoid (covariance.hh:204-234): linear in between.
double sigmoid_clamp(double one, double zero, double x)
{
  if (std::fabs(x) < one)
    return 1.;
  if (std::fabs(x) > zero)
    return 0.;
  return (zero - std::fabs(x)) / (zero - one);
}

// SmoothSigmoid (covariance.hh:147-196): theta(x1) / (theta(x1) + theta(x2)),
// theta(y) = exp(-1/y) for y > 0.
double sigmoid_smooth(double one, double zero, double x)
{
  if (std::fabs(x) < one)
    return 1.;
  if (std::fabs(x) > zero)
    return 0.;
  const double x1 = (zero - std::fabs(x)) / (zero - one);
  const double x2 = (std::fabs(x) - one) / (zero - one);
  auto theta = [](double y) { return y > 0. ? std::exp(-1. / y) : 0.; };
  return theta(x1) / (theta(x1) + theta(x2));
}

// SmoothstepSigmoid (covariance.hh:246-290): Hermite interpolation matching
// `rec` derivatives, built by the reference's recursion over
// (1 - x2^2)^i with x2 = 2 clamp - 1.
double sigmoid_smoothstep(double one, double zero, double x, unsigned rec)
{
  const double x2 = 2. * sigmoid_clamp(one, zero, x) - 1.;
  double norm = 1., add = 1., out = 1.;
  for (unsigned i = 1; i <= rec; i++) {
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

// embedding.mergeRecursions: the reference's value 99 (the default) picks a
// count per covariance family, matrix.hh:679-705.
unsigned merge_recursions(const pf_desc& d)
{
  if (d.merge_recursions > 0)
    return (unsigned)d.merge_recursions;
  if (d.merge_recursions < 0)
    return 0;
  const double half_dim = (double)d.dim / 2.;
  switch (d.covariance) {
    case PF_COV_MATERN:
      return (unsigned)std::floor(d.matern_nu * 2. + half_dim);
    case PF_COV_GAMMA_EXPONENTIAL:
      return (unsigned)std::floor(d.exp_gamma + half_dim);
    case PF_COV_GENERALIZED_CAUCHY:
      return (unsigned)std::floor(d.cauchy_alpha + half_dim);
    case PF_COV_EXPONENTIAL:
      return 2;
    case PF_COV_MATERN32:
      return 4;
    case PF_COV_MATERN52:
      return 6;
    case PF_COV_CAUCHY:
      return 3;
    default:
      throw std::runtime_error(
        "magic recursions value 99 not implemented for choice of covariance");
  }
}

// Entry of the merge embedding at integer position idx (no wrap):
// Matrix::computeMatrixEntriesWithMerge, matrix.hh:652-775. The 2^dim mirror
// images x_i -> ext_i * factor - x_i of the position each contribute the
// covariance times a sigmoid taper; what the tapers leave (remainder) takes
// the covariance at the centre of the extended domain (radialPlus) or 0.
double merged_cov(const Geo& g, const Geometry& geo, const Covariance& cov, const pf_desc& d,
                  const Index idx[3])
{
  const int dim = g.dim;
  const unsigned factor = (unsigned)g.factor;
  const unsigned rec = merge_recursions(d);
  const int combine = d.sigmoid_combine;
  const bool radial = combine == PF_COMBINE_RADIAL || combine == PF_COMBINE_RADIAL_PLUS;
  const double start = d.sigmoid_start != 0. ? d.sigmoid_start : 1.;
  const double end = d.sigmoid_end != 0. ? d.sigmoid_end : (double)factor - 1.;
  auto sig = [&](double one, double zero, double x) {
    return d.sigmoid == PF_SIGMOID_SMOOTH ? sigmoid_smooth(one, zero, x)
                                          : sigmoid_smoothstep(one, zero, x, rec);
  };
  double constValue = 0.;
  if (combine == PF_COMBINE_RADIAL_PLUS) {
    double c[3] = { 0, 0, 0 }, tc[3] = { 0, 0, 0 };
    for (int i = 0; i < dim; i++)
      c[i] = g.ext[i] * (double)factor / 2.;
    geo.transform(c, tc);
    constValue = cov(tc);
  }
  double trueCoord[3] = { 0, 0, 0 };
  for (int i = 0; i < dim; i++)
    trueCoord[i] = (double)idx[i] * g.h[i];
  double value = 0., remainder = 1.;
  const double eps = 1e-10;
  const double sqrtDim = std::sqrt((double)dim);
  for (unsigned j = 0; j < (1u << dim); j++) {
    double mirror[3] = { 0, 0, 0 };
    for (int i = 0; i < dim; i++)
      mirror[i] = (j & (1u << i)) ? g.ext[i] * (double)factor - trueCoord[i] : trueCoord[i];
    double damp = 1.;
    if (radial && sqrtDim < factor / 2.) {
      double radius = 0.;
      for (int i = 0; i < dim; i++)
        radius += (mirror[i] * mirror[i]) / (g.ext[i] * g.ext[i]);
      radius = std::sqrt(radius);
      damp = sig(start * sqrtDim - eps, factor - (factor - end) * sqrtDim + eps, radius);
    } else {
      for (int i = 0; i < dim; i++)
        damp *= sig(g.ext[i] * start - eps, g.ext[i] * end + eps, mirror[i]);
    }
    if (damp < eps)
      continue;
    remainder -= damp;
    value += cov_at(g, geo, cov, d, mirror) * damp;
  }
  return value + constValue * remainder;
}

// Adaptive Simpson quadrature of f on [a, b] to absolute tolerance eps.
template <typename F>
double simpson_adapt(F& f, double a, double b, double fa, double fm, double fb, double whole,
                     double eps, int depth)
{
  const double m = 0.5 * (a + b), lm = 0.5 * (a + m), rm = 0.5 * (m + b);
  const double flm = f(lm), frm = f(rm);
  const double left = (m - a) / 6. * (fa + 4. * flm + fm);
  const double right = (b - m) / 6. * (fm + 4. * frm + fb);
  if (depth <= 0 || std::fabs(left + right - whole) <= 15. * eps)
    return left + right + (left + right - whole) / 15.;
  return simpson_adapt(f, a, m, fa, flm, fm, left, 0.5 * eps, depth - 1) +
         simpson_adapt(f, m, b, fm, frm, fb, right, 0.5 * eps, depth - 1);
}

// Integral of the taper sigmoid(one, zero, x) over x in [one, hi]; the taper
// is zero from `zero` on. Smoothstep: exact antiderivative. With
// u = 1 - 2 (x - one) / (zero - one) and w = 1 - u^2 the taper is
// (1 + u S(u) / n) / 2, S = sum_j a_j w^j from the recursion
// out_i = (2i out_{i-1} + w^i) / (2i + 1); Q(u) = sum_j a_j
// (-w^(j+1) / (2 (j + 1))) has Q' = u S and follows the same recursion, and
// the integral is (zero - one) / 2 * (G(1) - G(u)), G(u) = u / 2 + Q(u) / (2 n),
// G(1) = 1/2. Smooth: adaptive Simpson to 1e-14.
double taper_integral(int sigmoid, double one, double zero, double hi, unsigned rec)
{
  hi = std::min(hi, zero);
  if (!(hi > one))
    return 0.;
  if (sigmoid == PF_SIGMOID_SMOOTHSTEP) {
    const double u = 1. - 2. * (hi - one) / (zero - one), w = 1. - u * u;
    double n = 1., q = -0.5 * w, wp = w;
    for (unsigned i = 1; i <= rec; i++) {
      wp *= w;
      q = (2. * i * q - wp / (2. * (i + 1.))) / (2. * i + 1.);
      n *= 2. * i / (2. * i + 1.);
    }
    return 0.5 * (zero - one) * (0.5 - (0.5 * u + 0.5 * q / n));
  }
  auto f = [&](double x) { return sigmoid_smooth(one, zero, x); };
  const double fa = f(one), fm = f(0.5 * (one + hi)), fb = f(hi);
  return simpson_adapt(f, one, hi, fa, fm, fb, (hi - one) / 6. * (fa + 4. * fm + fb), 1e-14, 40);
}

// The fold embedding's integral of sigmoid(1, beta, x) over [1, norm]
// (gsl_integration_qng in the reference, absolute tolerance 1e-6, relative
// 1e-3; matrix.hh:866-868).
double fold_integral(int sigmoid, double beta, double norm, unsigned rec)
{
  return taper_integral(sigmoid, 1., beta, norm, rec);
}

// Matrix::modifyMatrixEntriesWithCofold, matrix.hh:905-1013, on the classical
// entries of a dense grid with C[i] points per axis, reached through at(f):
// alpha = largest entry of norm > 1 on the border band of the extended
// domain, beta = min(cofoldMaxFactor * alpha, smallest entry of max-norm
// <= 1); entries of norm > 1 below alpha take offset = beta - I(beta), those
// in [alpha, beta) I(x) + offset, with I(x) the integral of
// sigmoid(0, beta - alpha, beta - t) over t in [alpha, x], i.e.
// F(beta - alpha) - F(beta - x), F(y) the taper's integral over [0, y].
template <typename At>
void cofold(const pf_desc& d, const Geo& g, const Index C[3], At&& at)
{
  const Index total = C[0] * C[1] * C[2];
  auto norm_of = [&](Index f, bool& border) {
    const Index idx[3] = { f % C[0], (f / C[0]) % C[1], f / (C[0] * C[1]) };
    double norm = 0.;
    border = false;
    for (int i = 0; i < g.dim; i++) {
      double c = (double)idx[i] * g.h[i];
      if (c > 0.5 * g.ext[i] * (double)g.factor)
        c -= g.ext[i] * (double)g.factor;
      norm = std::max(norm, std::fabs(c) / g.ext[i]);
      if (0.5 * g.ext[i] * (double)g.factor - std::fabs(c) < g.ext[i] / (double)g.N[i])
        border = true;
    }
    return norm;
  };
  double min_constrained = std::numeric_limits<double>::max();
  double max_border = -std::numeric_limits<double>::max();
  for (Index f = 0; f < total; f++) {
    bool border;
    const double norm = norm_of(f, border);
    if (norm <= 1.)
      min_constrained = std::min(min_constrained, at(f));
    else if (border)
      max_border = std::max(max_border, at(f));
  }
  const double max_factor = d.fold_max_factor != 0. ? d.fold_max_factor : 5.;
  const unsigned rec =
    d.merge_recursions > 0 ? (unsigned)d.merge_recursions : (d.merge_recursions < 0 ? 0u : 1u);
  const double alpha = max_border, beta = std::min(max_factor * max_border, min_constrained);
  const double w = beta - alpha;
  auto F = [&](double y) { return taper_integral(d.sigmoid, 0., w, y, rec); };
  const double offset = w > 0. ? beta - F(w) : beta;
  for (Index f = 0; f < total; f++) {
    bool border;
    if (!(norm_of(f, border) > 1.))
      continue;
    const double x = at(f);
    if (x < alpha)
      at(f) = offset;
    else if (x < beta)
      at(f) = (F(w) - F(beta - x)) + offset;
  }
}

// Entry value of the embedding grid at integer position idx:
// Matrix::computeMatrixEntriesWithMirroring, matrix.hh:615-638 (classical),
// computeMatrixEntriesWithFold, matrix.hh:790-888 (fold: coordinates whose
// extension-scaled norm exceeds 1 move towards the origin, to norm
// 1 + the taper's integral), or the merge embedding.
double embedded_cov(const Geo& g, const Geometry& geo, const Covariance& cov, const pf_desc& d,
                    const Index idx[3])
{
  if (d.periodization == PF_PERIOD_MERGE)
    return merged_cov(g, geo, cov, d, idx);
  double coord[3] = { 0, 0, 0 };
  for (int i = 0; i < g.dim; i++) {
    coord[i] = (double)idx[i] * g.h[i];
    if (coord[i] > 0.5 * g.ext[i] * (double)g.factor)
      coord[i] -= g.ext[i] * (double)g.factor;
  }
  if (d.periodization == PF_PERIOD_FOLD) {
    const double max_factor = d.fold_max_factor != 0. ? d.fold_max_factor : 1.;
    const double beta = max_factor * 0.5 * (double)g.factor / std::sqrt((double)g.dim);
    double norm = 0.;
    for (int i = 0; i < g.dim; i++)
      norm += (coord[i] * coord[i]) / (g.ext[i] * g.ext[i]);
    norm = std::sqrt(norm / (double)g.dim);
    if (norm > 1.) {
      const double val = fold_integral(d.sigmoid, beta, norm, merge_recursions(d)) + 1.;
      for (int i = 0; i < g.dim; i++)
        coord[i] *= val / norm;
    }
  }
  return cov_at(g, geo, cov, d, coord);
}

// ----------------------------------------------------------------------------
// Eigenvalues. Result kept in a logical (natural x-fastest) layout:
//   DCT: P0*P1*P2 reals, P_i = N_i/2+1.
//   R2C: H*N1*N2 reals (Re part after the reference's clamping).
// ----------------------------------------------------------------------------
struct Eigs
{
  std::vector<double> v;
  pf_eig_stats st;
};

// Classification loop, matrix.hh:266-291, over `count` reals visited in the
// reference's storage order through `at(i)` (a reference to the value).
template<typename At>
void classify(Index count, double threshold, At at, pf_eig_stats& st)
{
  st.small = st.zero = st.negative = 0;
  st.smallest = std::numeric_limits<double>::max();
  st.entries = count;
  for (Index i = 0; i < count; i++) {
    double& value = at(i);
    if (value < st.smallest)
      st.smallest = value;
    if (value < 1e-6) {
      if (value < threshold) {
        if (value > -threshold)
          st.zero++;
        else
          st.negative++;
      } else
        st.small++;
    }
    if (value < 0.)
      value = 0.;
  }
}

// DCTMatrixBackend (dctmatrixbackend.hh:240-272) filled through
// computeMatrixEntriesWithMirroring; `custom` (if non-null) supplies the
// entries in x-fastest order instead.
Eigs eigs_dct(const pf_desc& d, const Geo& g, const double* custom)
{
  Geometry geo(d);
  Covariance cov(d);
  Index P[3];
  for (int i = 0; i < 3; i++)
    P[i] = (i < g.dim) ? g.N[i] / 2 + 1 : 1;
  const Index total = P[0] * P[1] * P[2];
  Eigs e;
  e.v.resize(total);
  parallel_for(total, [&](Index b, Index en) {
    for (Index f = b; f < en; f++) {
      Index idx[3] = { f % P[0], (f / P[0]) % P[1], f / (P[0] * P[1]) };
      e.v[f] = custom ? custom[f] : embedded_cov(g, geo, cov, d, idx);
    }
  });
  if (!custom && d.periodization == PF_PERIOD_COFOLD)
    cofold(d, g, P, [&](Index f) -> double& { return e.v[f]; });
  for (int ax = 0; ax < g.dim; ax++)
    redft00_axis(e.v.data(), P, ax);
  for (auto& x : e.v)
    x /= (double)g.Next;
  // classification is order independent for the DCT array
  classify(total, d.threshold, [&](Index i) -> double& { return e.v[i]; }, e.st);
  return e;
}


// R2CMatrixBackend (r2cmatrixbackend.hh:254-285, finalize :412-421) with the
// reference's classification over the interleaved buffer: after the forward
// transform localMatrixSize() is the complex count (r2cmatrixbackend.hh:150-156)
// but get/set index the buffer as reals (:378-400), so matrix.hh:270-291
// checks and clamps Re and Im of the first half of the complex entries only.
Eigs eigs_r2c(const pf_desc& d, const Geo& g, const double* custom)
{
  Geometry geo(d);
  Covariance cov(d);
  const Index H = g.H, N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const Index rows = N1 * N2;
  std::vector<double> real(2 * H * rows);
  parallel_for(rows, [&](Index b, Index en) {
    for (Index r = b; r < en; r++) {
      Index idx[3] = { 0, r % N1, r / N1 };
      for (Index x = 0; x < 2 * H; x++) {
        idx[0] = x;
        real[r * 2 * H + x] = custom ? custom[r * 2 * H + x] : embedded_cov(g, geo, cov, d, idx);
      }
    }
  });
  if (!custom && d.periodization == PF_PERIOD_COFOLD) {
    const Index C[3] = { N0, N1, N2 };
    cofold(d, g, C, [&](Index f) -> double& { return real[(f / N0) * 2 * H + f % N0]; });
  }
  std::vector<cd> spec(H * rows);
  r2c_x(real.data(), 2 * H, spec.data(), (int)N0, rows);
  Index dims[3] = { H, N1, N2 };
  for (int ax = 1; ax < g.dim; ax++)
    c2c_axis(spec.data(), dims, ax, -1);
  for (auto& c : spec)
    c /= (double)g.Next;
  const bool tr = g.transposed && g.dim == 3;
  // storage index (FFTW_MPI_TRANSPOSED_OUT in 3D: kx fastest, then kz, then ky)
  // -> logical natural index kx + H (ky + N1 kz)
  auto storage_to_logical = [&](Index s) -> Index {
    if (!tr)
      return s;
    Index kx = s % H, kz = (s / H) % N2, ky = s / (H * N2);
    return kx + H * (ky + N1 * kz);
  };
  const Index count = H * rows;
  Eigs e;
  if (d.r2c_check_all)
    classify(
      count, d.threshold,
      [&](Index i) -> double& { return reinterpret_cast<double*>(&spec[i])[0]; }, e.st);
  else
    classify(
      count, d.threshold,
      [&](Index i) -> double& {
        double* p = reinterpret_cast<double*>(&spec[storage_to_logical(i / 2)]);
        return p[i % 2];
      },
      e.st);
  e.v.resize(count);
  for (Index c = 0; c < count; c++)
    e.v[c] = spec[c].real();
  return e;
}

// Storage index -> logical (k0,k1,k2) of the full grid of the DFT backends:
// transposeIfNeeded swaps the two last extents (dftmatrixbackend.hh:150-158,
// dftfieldbackend.hh:140-147).
void dft_storage_to_k(const Geo& g, Index s, Index k[3])
{
  const Index N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  if (g.dim == 2 && g.transposed) {
    k[1] = s % N1;
    k[0] = s / N1;
    k[2] = 0;
  } else if (g.dim == 3 && g.transposed) {
    k[0] = s % N0;
    k[2] = (s / N0) % N2;
    k[1] = s / (N0 * N2);
  } else {
    k[0] = s % N0;
    k[1] = (s / N0) % N1;
    k[2] = s / (N0 * N1);
  }
}

// DFTMatrixBackend (dftmatrixbackend.hh:209-273): the covariance on every
// entry of the extended grid (computeMatrixEntries, real part, imaginary part
// 0), the complex forward transform divided by N_ext, then
// fillTransformedMatrix's classification (matrix.hh:266-291) over
// localMatrixSize() = N_ext entries of get(index) = the real part, in the
// storage order. `custom`: N0 x N1 x N2 values, x fastest.
Eigs eigs_dft(const pf_desc& d, const Geo& g, const double* custom)
{
  Geometry geo(d);
  Covariance cov(d);
  const Index N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const Index total = N0 * N1 * N2;
  std::vector<double> real(total);
  parallel_for(total, [&](Index b, Index en) {
    for (Index f = b; f < en; f++) {
      Index idx[3] = { f % N0, (f / N0) % N1, f / (N0 * N1) };
      real[f] = custom ? custom[f] : embedded_cov(g, geo, cov, d, idx);
    }
  });
  if (!custom && d.periodization == PF_PERIOD_COFOLD) {
    const Index C[3] = { N0, N1, N2 };
    cofold(d, g, C, [&](Index f) -> double& { return real[f]; });
  }
  std::vector<cd> spec(total);
  for (Index f = 0; f < total; f++)
    spec[f] = cd(real[f], 0.);
  Index dims[3] = { N0, N1, N2 };
  for (int ax = 0; ax < g.dim; ax++)
    c2c_axis(spec.data(), dims, ax, -1);
  Eigs e;
  e.v.resize(total);
  for (Index f = 0; f < total; f++)
    e.v[f] = spec[f].real() / (double)g.Next;
  classify(
    total, d.threshold,
    [&](Index s) -> double& {
      Index k[3];
      dft_storage_to_k(g, s, k);
      return e.v[k[0] + N0 * (k[1] + N1 * k[2])];
    },
    e.st);
  return e;
}

Eigs compute_eigs(const pf_desc& d, const Geo& g, const double* custom)
{
  if (d.matrix_backend == PF_MATRIX_DCT)
    return eigs_dct(d, g, custom);
  if (d.matrix_backend == PF_MATRIX_R2C)
    return eigs_r2c(d, g, custom);
  if (d.matrix_backend == PF_MATRIX_DFT)
    return eigs_dft(d, g, custom);
  throw std::runtime_error("unknown matrix backend");
}

// Eigenvalue at logical spectral position (k0,k1,k2) of the R2C half spectrum
// (DCTMatrixBackend::eval folds k >= P to N-k, dctmatrixbackend.hh:337-345;
// R2CMatrixBackend::eval reads the same position, r2cmatrixbackend.hh:337-342).
double eig_at(const pf_desc& d, const Geo& g, const Eigs& e, Index k0, Index k1, Index k2)
{
  // DFTMatrixBackend::eval(indices) reads the full grid (dftmatrixbackend.hh:309-318)
  if (d.matrix_backend == PF_MATRIX_DFT)
    return e.v[k0 + g.N[0] * (k1 + g.N[1] * k2)];
  if (d.matrix_backend == PF_MATRIX_DCT) {
    Index P0 = g.N[0] / 2 + 1, P1 = g.N[1] / 2 + 1, P2 = g.N[2] / 2 + 1;
    if (g.dim < 3)
      P2 = 1;
    Index a = k0 >= P0 ? g.N[0] - k0 : k0;
    Index b = k1 >= P1 ? g.N[1] - k1 : k1;
    Index c = (g.dim == 3 && k2 >= P2) ? g.N[2] - k2 : k2;
    return e.v[a + P0 * (b + P1 * c)];
  }
  return e.v[k0 + g.H * (k1 + g.N[1] * k2)];
}

// Storage index -> logical (k0,k1,k2) of the R2C half spectrum: the order in
// which generateField visits entries (matrix.hh:374-404; 3D transposed
// layout after transposeIfNeeded, r2cfieldbackend.hh:190-199).
void spec_storage_to_k(const Geo& g, Index s, Index k[3])
{
  const Index H = g.H;
  if (g.dim == 3 && g.transposed) {
    k[0] = s % H;
    k[2] = (s / H) % g.N[2];
    k[1] = s / (H * g.N[2]);
  } else {
    k[0] = s % H;
    k[1] = (s / H) % g.N[1];
    k[2] = s / (H * g.N[1]);
  }
}

// Backward transform of the R2C field backend (r2cfieldbackend.hh:250-274)
// followed by extendedFieldToField (:423-441).
void r2c_backward_extract(const Geo& g, std::vector<cd>& spec, double* field)
{
  Index dims[3] = { g.H, g.N[1], g.N[2] };
  for (int ax = g.dim - 1; ax >= 1; ax--)
    c2c_axis(spec.data(), dims, ax, +1);
  const Index rows = g.N[1] * g.N[2];
  std::vector<double> real(g.N[0] * rows);
  c2r_x(spec.data(), real.data(), g.N[0], (int)g.N[0], rows);
  for (Index z = 0; z < g.n[2]; z++)
    for (Index y = 0; y < g.n[1]; y++)
      for (Index x = 0; x < g.n[0]; x++)
        field[x + g.n[0] * (y + g.n[1] * z)] = real[x + g.N[0] * (y + g.N[1] * z)];
}

// Matrix::generateField, general version (matrix.hh:374-421) with
// R2CFieldBackend::set (r2cfieldbackend.hh:305-325).
void generate_r2c(const pf_desc& d, const Geo& g, const Eigs& e, const double* noise,
                  double* field)
{
  static const double sqrtTwo = std::sqrt(2.);
  const Index count = g.H * g.N[1] * g.N[2];
  std::vector<cd> spec(count);
  for (Index s = 0; s < count; s++) {
    Index k[3];
    spec_storage_to_k(g, s, k);
    const double lambda = std::sqrt(eig_at(d, g, e, k[0], k[1], k[2]));
    const double rand1 = noise[2 * s];
    const double rand2 = noise[2 * s + 1];
    bool allMultiple = true;
    for (int i = 0; i < g.dim; i++)
      if ((2 * k[i]) % g.N[i] != 0)
        allMultiple = false;
    const Index c = k[0] + g.H * (k[1] + g.N[1] * k[2]);
    if (allMultiple)
      spec[c] = cd(lambda * rand1, 0.);
    else
      spec[c] = cd(lambda / sqrtTwo * rand1, lambda / sqrtTwo * rand2);
  }
  r2c_backward_extract(g, spec, field);
}

// Storage permutation of the DCT/DST arrays: transposeIfNeeded swaps the two
// last dims (dctdstfieldbackend.hh:868-882).
void dct_storage_to_k(const Geo& g, const Index P[3], Index s, Index k[3])
{
  int perm[3] = { 0, 1, 2 };
  if (g.transposed)
    std::swap(perm[g.dim - 2], perm[g.dim - 1]);
  Index r = s;
  for (int j = 0; j < 3; j++) {
    Index sz = P[perm[j]];
    k[perm[j]] = r % sz;
    r /= sz;
  }
}

// Field extraction walk of metaForLoopFromExtended (dctdstfieldbackend.hh:905-930):
// x runs over n0 entries then skips 1, each y row skips P0 after n1 rows.
Index dct_extract_index(const Geo& g, const Index P[3], Index x, Index y, Index z)
{
  if (g.dim == 2)
    return x + y * (g.n[0] + 1);
  return x + y * (g.n[0] + 1) + z * (g.n[1] * (g.n[0] + 1) + P[0]);
}

// Matrix::generateField, DCT/DST version (matrix.hh:342-372) with
// DCTDSTFieldBackend::set (dctdstfieldbackend.hh:418-438), the r2r backward
// transforms (:326-404; REDFT00 on even axes, RODFT00 on the interior of odd
// axes with the zero boundary reinserted, :1041-1140) and the additive
// extraction (:631-648).
void generate_dctdst(const pf_desc& d, const Geo& g, const Eigs& e, const double* noise,
                     double* field)
{
  if (d.matrix_backend != PF_MATRIX_DCT)
    throw std::runtime_error("DCTDSTFieldBackend requires DCTMatrixBackend");
  Index P[3];
  for (int i = 0; i < 3; i++)
    P[i] = (i < g.dim) ? g.N[i] / 2 + 1 : 1;
  const Index count = P[0] * P[1] * P[2];
  const double scale = 1. / std::sqrt((double)(1 << g.dim));
  for (Index i = 0; i < g.domain; i++)
    field[i] = 0.;
  std::vector<double> a(count);
  for (unsigned type = 0; type < (1u << g.dim); type++) {
    bool odd[3] = { false, false, false };
    for (int i = 0; i < g.dim; i++)
      odd[i] = (type >> i) & 1;
    for (Index s = 0; s < count; s++) {
      Index k[3];
      dct_storage_to_k(g, P, s, k);
      const Index c = k[0] + P[0] * (k[1] + P[1] * k[2]);
      const double lambda = std::sqrt(e.v[c]);
      const double rand = noise[type * count + s];
      unsigned border = 0;
      bool zero = false;
      for (int i = 0; i < g.dim; i++)
        if (k[i] == 0 || k[i] == P[i] - 1) {
          if (odd[i]) {
            zero = true;
            break;
          }
          border++;
        }
      if (zero) {
        a[c] = 0.;
        continue;
      }
      double v = scale * lambda * rand;
      if (border > 0)
        v *= std::sqrt((double)(1 << border));
      a[c] = v;
    }
    for (int ax = 0; ax < g.dim; ax++) {
      if (!odd[ax]) {
        redft00_axis(a.data(), P, ax);
        continue;
      }
      Index stride = 1;
      for (int i = 0; i < ax; i++)
        stride *= P[i];
      const int n = (int)P[ax];
      const Index lines = count / n;
      parallel_for(lines, [&](Index b, Index en) {
        std::vector<double> in(n), out(n);
        for (Index l = b; l < en; l++) {
          Index inner = l % stride, outer = l / stride;
          Index base = inner + outer * stride * n;
          for (int j = 0; j < n - 2; j++)
            in[j] = a[base + (j + 1) * stride];
          rodft00_line(in.data(), out.data(), n - 2);
          a[base] = 0.;
          a[base + (n - 1) * stride] = 0.;
          for (int j = 0; j < n - 2; j++)
            a[base + (j + 1) * stride] = out[j];
        }
      });
    }
    for (Index z = 0; z < g.n[2]; z++)
      for (Index y = 0; y < g.n[1]; y++)
        for (Index x = 0; x < g.n[0]; x++)
          field[x + g.n[0] * (y + g.n[1] * z)] += a[dct_extract_index(g, P, x, y, z)];
  }
}

// Matrix::multiplyExtended / multiplyRootExtended / multiplyInverseExtended,
// general version (matrix.hh:1443-1646): zero-pad (r2cfieldbackend.hh:352-368),
// r2c / N_ext (:210-239), mult (:337-341), c2r, restrict.
// Learned spectral preconditioner table (exp of the bin log-multipliers),
// used by matvec op 3: bin of the folded position |k/(N/2)|/sqrt(dim).
thread_local std::vector<double> g_spec_table;

// Multiplier of an eigenvalue for the products: lambda (operator*), sqrt
// (multiplyRoot), 1 / lambda (multiplyInverseExtended), or op 4: the PCG
// embedding-inverse preconditioner 1 / max(lambda, g_inv_floor).
double g_inv_floor = 0.;
double op_factor(int op, double m)
{
  if (op == PF_OP_MULTIPLY)
    return m;
  if (op == PF_OP_ROOT)
    return std::sqrt(m);
  if (op == 4)
    return 1. / std::max(m, g_inv_floor);
  return 1. / m;
}

// Backward complex transform of the full grid (dftfieldbackend.hh:166-196)
// followed by extendedFieldToField of component `comp` (:280-318): the first
// n_i entries per axis of the real (0) or imaginary (1) part.
void dft_backward_extract(const Geo& g, std::vector<cd>& spec, double* re, double* im)
{
  Index dims[3] = { g.N[0], g.N[1], g.N[2] };
  for (int ax = 0; ax < g.dim; ax++)
    c2c_axis(spec.data(), dims, ax, +1);
  for (Index z = 0; z < g.n[2]; z++)
    for (Index y = 0; y < g.n[1]; y++)
      for (Index x = 0; x < g.n[0]; x++) {
        const cd v = spec[x + g.N[0] * (y + g.N[1] * z)];
        const Index o = x + g.n[0] * (y + g.n[1] * z);
        if (re)
          re[o] = v.real();
        if (im)
          im[o] = v.imag();
      }
}

// Matrix::generateField, general version (matrix.hh:374-421) with
// DFTFieldBackend::set (dftfieldbackend.hh:236-240): entry `index` of the
// storage order gets lambda (rand1 + i rand2) with lambda = sqrt(eval(index))
// (the matrix and field layouts agree for the DFT and DCT matrices); the
// backward transform gives the field (real part) and the spare field
// (imaginary part, hasSpareField, :232).
void generate_dft(const pf_desc& d, const Geo& g, const Eigs& e, const double* noise,
                  double* field, double* spare)
{
  const Index count = g.N[0] * g.N[1] * g.N[2];
  std::vector<cd> spec(count);
  for (Index s = 0; s < count; s++) {
    Index k[3];
    dft_storage_to_k(g, s, k);
    const double lambda = std::sqrt(eig_at(d, g, e, k[0], k[1], k[2]));
    spec[k[0] + g.N[0] * (k[1] + g.N[1] * k[2])] =
      cd(lambda * noise[2 * s], lambda * noise[2 * s + 1]);
  }
  dft_backward_extract(g, spec, field, spare);
}

// Products with the DFT field backend (matrix.hh:1443-1646):
// fieldToExtendedField (dftfieldbackend.hh:244-278: real part, zero
// padding), forward complex transform / N_ext (:198-226), mult by f(lambda)
// on both parts (:242), backward transform, real part of the first n_i.
void matvec_dft(const pf_desc& d, const Geo& g, const Eigs& e, int op, const double* x, double* y)
{
  const Index N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  std::vector<cd> spec(N0 * N1 * N2, cd(0., 0.));
  for (Index z = 0; z < g.n[2]; z++)
    for (Index yy = 0; yy < g.n[1]; yy++)
      for (Index xx = 0; xx < g.n[0]; xx++)
        spec[xx + N0 * (yy + N1 * z)] = cd(x[xx + g.n[0] * (yy + g.n[1] * z)], 0.);
  Index dims[3] = { N0, N1, N2 };
  for (int ax = 0; ax < g.dim; ax++)
    c2c_axis(spec.data(), dims, ax, -1);
  for (auto& c : spec)
    c /= (double)g.Next;
  for (Index k2 = 0; k2 < N2; k2++)
    for (Index k1 = 0; k1 < N1; k1++)
      for (Index k0 = 0; k0 < N0; k0++) {
        const double m = eig_at(d, g, e, k0, k1, k2);
        const double f = op_factor(op, m);
        spec[k0 + N0 * (k1 + N1 * k2)] *= f;
      }
  dft_backward_extract(g, spec, y, nullptr);
}

int radial_bin(const Geo& g, Index k0, Index k1, Index k2, int nbins)
{
  Index a = 2 * k0 > g.N[0] ? g.N[0] - k0 : k0;
  Index b = 2 * k1 > g.N[1] ? g.N[1] - k1 : k1;
  Index c = 2 * k2 > g.N[2] ? g.N[2] - k2 : k2;
  const double x = (double)a / (g.N[0] / 2), y = (double)b / std::max<Index>(g.N[1] / 2, 1),
               z = g.dim == 3 ? (double)c / std::max<Index>(g.N[2] / 2, 1) : 0.;
  const double r = std::sqrt((x * x + y * y + z * z) / g.dim);
  int bin = (int)(r * nbins);
  return bin < nbins ? bin : nbins - 1;
}

void matvec_r2c(const pf_desc& d, const Geo& g, const Eigs& e, int op, const double* x,
                double* y)
{
  const Index rows = g.N[1] * g.N[2];
  std::vector<double> real(g.N[0] * rows, 0.);
  for (Index z = 0; z < g.n[2]; z++)
    for (Index yy = 0; yy < g.n[1]; yy++)
      for (Index xx = 0; xx < g.n[0]; xx++)
        real[xx + g.N[0] * (yy + g.N[1] * z)] = x[xx + g.n[0] * (yy + g.n[1] * z)];
  std::vector<cd> spec(g.H * rows);
  r2c_x(real.data(), g.N[0], spec.data(), (int)g.N[0], rows);
  Index dims[3] = { g.H, g.N[1], g.N[2] };
  for (int ax = 1; ax < g.dim; ax++)
    c2c_axis(spec.data(), dims, ax, -1);
  for (auto& c : spec)
    c /= (double)g.Next;
  for (Index k2 = 0; k2 < g.N[2]; k2++)
    for (Index k1 = 0; k1 < g.N[1]; k1++)
      for (Index k0 = 0; k0 < g.H; k0++) {
        double f;
        if (op == 3)
          f = g_spec_table[radial_bin(g, k0, k1, k2, (int)g_spec_table.size())];
        else {
          const double m = eig_at(d, g, e, k0, k1, k2);
          f = op_factor(op, m);
        }
        spec[k0 + g.H * (k1 + g.N[1] * k2)] *= f;
      }
  r2c_backward_extract(g, spec, y);
}

struct Engine
{
  pf_desc d;
  Geo g;
  Eigs e;
  int status;
};

Engine* get_engine(const pf_desc* d, const double* custom = nullptr)
{
  static thread_local Engine* cached = nullptr;
  if (!custom && cached && std::memcmp(&cached->d, d, sizeof(pf_desc)) == 0)
    return cached;
  Engine* en = new Engine;
  en->d = *d;
  en->g = make_geo(*d);
  if (d->field_backend == PF_FIELD_R2C && d->dim == 2 && en->g.transposed)
    throw std::runtime_error("R2CFieldBackend supports transposed output only for dim > 2");
  en->e = compute_eigs(*d, en->g, custom);
  en->status = (en->e.st.negative > 0 && !d->approximate) ? PF_ENEG : PF_OK;
  if (custom)
    return en; // caller owns
  delete cached;
  cached = en;
  return en;
}

// r2r along one axis of the P-grid array for one parity: REDFT00 on all P
// entries of an even axis, RODFT00 on the P-2 interior entries of an odd axis
// with the boundary entries kept at zero (dctdstfieldbackend.hh:236-404, the
// zero strip/reinsert of :960-1171). Both are unnormalised and their own
// inverses up to 2(P-1) and 2(P-1) respectively.
void r2r_axis(double* a, const Index P[3], int ax, bool odd)
{
  if (!odd) {
    redft00_axis(a, P, ax);
    return;
  }
  Index stride = 1;
  for (int i = 0; i < ax; i++)
    stride *= P[i];
  const int n = (int)P[ax];
  const Index lines = P[0] * P[1] * P[2] / n;
  parallel_for(lines, [&](Index b, Index en) {
    std::vector<double> in(n), out(n);
    for (Index l = b; l < en; l++) {
      Index inner = l % stride, outer = l / stride;
      Index base = inner + outer * stride * n;
      for (int j = 0; j < n - 2; j++)
        in[j] = a[base + (j + 1) * stride];
      rodft00_line(in.data(), out.data(), n - 2);
      a[base] = 0.;
      a[base + (n - 1) * stride] = 0.;
      for (int j = 0; j < n - 2; j++)
        a[base + (j + 1) * stride] = out[j];
    }
  });
}

// Matrix::multiplyExtended / multiplyRootExtended / multiplyInverseExtended,
// DCT/DST version (matrix.hh:1443-1470, :1514-1541, :1587-1614):
// fieldToExtendedField (dctdstfieldbackend.hh:489-500: zeroed P-grid array,
// the first n_i entries of every axis from the field), then for every parity
// type t: setComponent (:440-458: value / 2^{d - border}, zero on the boundary
// of an odd axis), forwardTransform (:236-316: r2r per axis, / N_ext),
// mult by f(lambda) (:461-466), backwardTransform (:326-404) and the
// extraction (:631-648, :905-955), which adds into the output for every type
// (the non-additive walk adds too, :923). The forward and backward transforms
// both leave the array in the storage order (transposed or not) in which the
// matrix backend stores its eigenvalues, so index k of one meets index k of
// the other; here both stay in the logical order. `y` is accumulated into:
// Matrix::operator* passes the input itself as the output (matrix.hh:142-151),
// so the reference returns input + C x / N_ext there.
void matvec_dctdst(const pf_desc& d, const Geo& g, const Eigs& e, int op, const double* x,
                   double* y)
{
  Index P[3];
  for (int i = 0; i < 3; i++)
    P[i] = (i < g.dim) ? g.N[i] / 2 + 1 : 1;
  const Index count = P[0] * P[1] * P[2];
  std::vector<double> ext(count, 0.), a(count);
  for (Index z = 0; z < g.n[2]; z++)
    for (Index yy = 0; yy < g.n[1]; yy++)
      for (Index xx = 0; xx < g.n[0]; xx++)
        ext[xx + P[0] * (yy + P[1] * z)] = x[xx + g.n[0] * (yy + g.n[1] * z)];
  for (unsigned type = 0; type < (1u << g.dim); type++) {
    bool odd[3] = { false, false, false };
    for (int i = 0; i < g.dim; i++)
      odd[i] = (type >> i) & 1;
    for (Index c = 0; c < count; c++) {
      const Index k[3] = { c % P[0], (c / P[0]) % P[1], c / (P[0] * P[1]) };
      unsigned border = 0;
      bool zero = false;
      for (int i = 0; i < g.dim; i++)
        if (k[i] == 0 || k[i] == P[i] - 1) {
          if (odd[i])
            zero = true;
          border++;
        }
      a[c] = zero ? 0. : ext[c] / (double)(1u << (g.dim - border));
    }
    for (int ax = 0; ax < g.dim; ax++)
      r2r_axis(a.data(), P, ax, odd[ax]);
    for (Index c = 0; c < count; c++) {
      const double m = e.v[c];
      double f;
      if (op == 3) {
        const Index k0 = c % P[0], k1 = (c / P[0]) % P[1], k2 = c / (P[0] * P[1]);
        f = g_spec_table[radial_bin(g, k0, k1, k2, (int)g_spec_table.size())];
      } else
        f = op_factor(op, m);
      a[c] = a[c] / (double)g.Next * f;
    }
    for (int ax = 0; ax < g.dim; ax++)
      r2r_axis(a.data(), P, ax, odd[ax]);
    for (Index z = 0; z < g.n[2]; z++)
      for (Index yy = 0; yy < g.n[1]; yy++)
        for (Index xx = 0; xx < g.n[0]; xx++)
          y[xx + g.n[0] * (yy + g.n[1] * z)] += a[xx + P[0] * (yy + P[1] * z)];
  }
}

// The product M x (no accumulation) for either field backend.
void matvec(const Engine& en, int op, const double* x, double* y)
{
  if (en.d.field_backend == PF_FIELD_DCTDST) {
    std::vector<double> acc(en.g.domain, 0.);
    matvec_dctdst(en.d, en.g, en.e, op, x, acc.data());
    std::copy(acc.begin(), acc.end(), y);
    return;
  }
  if (en.d.field_backend == PF_FIELD_DFT) {
    matvec_dft(en.d, en.g, en.e, op, x, y);
    return;
  }
  matvec_r2c(en.d, en.g, en.e, op, x, y);
}

double dot(const std::vector<double>& a, const std::vector<double>& b)
{
  double s = 0.;
  for (size_t i = 0; i < a.size(); i++)
    s += a[i] * b[i];
  return s;
}

// Matrix::innerCG, matrix.hh:1315-1432, restated line by line (commSize 1).
int inner_cg(const Engine& en, std::vector<double>& iter, const std::vector<double>& solution,
             bool precondition, unsigned cgIterations)
{
  const size_t n = iter.size();
  std::vector<double> tempSolution = solution;
  std::vector<double> matrixTimesSolution(n), matrixTimesIter(n), residual(n), precResidual(n),
    direction(n), matrixTimesDirection(n);
  double scalarProd, scalarProd2, alphaDenominator, alpha, beta;
  matvec(en, PF_OP_MULTIPLY, tempSolution.data(), matrixTimesSolution.data()); // unused product
  matvec(en, PF_OP_MULTIPLY, iter.data(), matrixTimesIter.data());
  for (size_t i = 0; i < n; i++)
    residual[i] = solution[i] - matrixTimesIter[i];
  if (precondition)
    matvec(en, PF_OP_INVERSE_EMBED, residual.data(), precResidual.data());
  else
    precResidual = residual;
  direction = precResidual;
  bool converged = false;
  scalarProd = dot(precResidual, residual);
  scalarProd2 = dot(residual, residual);
  if (std::sqrt(std::abs(scalarProd2)) < 1e-6)
    converged = true;
  double firstValue = 0.;
  for (size_t i = 0; i < n; i++)
    firstValue += iter[i] * (0.5 * matrixTimesIter[i] - solution[i]);
  unsigned count = 0;
  while (!converged && count < cgIterations) {
    matvec(en, PF_OP_MULTIPLY, direction.data(), matrixTimesDirection.data());
    alphaDenominator = dot(direction, matrixTimesDirection);
    alpha = scalarProd / alphaDenominator;
    double oldValue = 0.;
    for (size_t i = 0; i < n; i++)
      oldValue += iter[i] * (0.5 * matrixTimesIter[i] - solution[i]);
    for (size_t i = 0; i < n; i++) {
      iter[i] += alpha * direction[i];
      matrixTimesIter[i] += alpha * matrixTimesDirection[i];
    }
    double value = 0.;
    for (size_t i = 0; i < n; i++)
      value += iter[i] * (0.5 * matrixTimesIter[i] - solution[i]);
    for (size_t i = 0; i < n; i++)
      residual[i] = solution[i] - matrixTimesIter[i];
    if (precondition)
      matvec(en, PF_OP_INVERSE_EMBED, residual.data(), precResidual.data());
    else
      precResidual = residual;
    beta = 1. / scalarProd;
    scalarProd = dot(precResidual, residual);
    beta *= scalarProd;
    for (size_t i = 0; i < n; i++)
      direction[i] = precResidual[i] + beta * direction[i];
    if (value != firstValue) {
      if (std::abs(value - oldValue) / std::abs(value - firstValue) < 1e-16)
        converged = true;
    }
    count++;
  }
  return (int)count;
}

// (P)CG with a relative-residual stop: recursive residual, replaced by the
// true residual b - A x every 50 iterations and whenever the recursion claims
// convergence (the stop is decided on the true residual). `prec` applies
// M^{-1}; x starts at 0.
int pcg(const Engine& en, const std::vector<double>& b, std::vector<double>& x,
        const std::function<void(const std::vector<double>&, std::vector<double>&)>* prec,
        int maxit, double rel_tol, double* relres)
{
  const size_t n = b.size();
  x.assign(n, 0.);
  std::vector<double> r = b, z(n), p(n), Ap(n), Ax(n);
  const double bnorm = std::sqrt(dot(b, b));
  if (bnorm == 0.) {
    *relres = 0.;
    return 0;
  }
  if (prec)
    (*prec)(r, z);
  else
    z = r;
  p = z;
  double rz = dot(r, z);
  int it = 0;
  double rr = 1.;
  auto true_residual = [&]() {
    matvec(en, PF_OP_MULTIPLY, x.data(), Ax.data());
    for (size_t i = 0; i < n; i++)
      r[i] = b[i] - Ax[i];
    return std::sqrt(dot(r, r)) / bnorm;
  };
  while (it < maxit) {
    matvec(en, PF_OP_MULTIPLY, p.data(), Ap.data());
    const double alpha = rz / dot(p, Ap);
    for (size_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    it++;
    rr = std::sqrt(dot(r, r)) / bnorm;
    if (rr <= rel_tol || it % 50 == 0) {
      rr = true_residual();
      if (rr <= rel_tol)
        break;
    }
    if (prec)
      (*prec)(r, z);
    else
      z = r;
    const double rzn = dot(r, z);
    const double beta = rzn / rz;
    rz = rzn;
    for (size_t i = 0; i < n; i++)
      p[i] = z[i] + beta * p[i];
  }
  if (rr > rel_tol || it == 0)
    rr = true_residual();
  *relres = rr;
  return it;
}

// FP64 MLP (exact GELU hidden layers, linear output), forward + backward of
// the mean squared error; parameter layout W (rows = outputs) then b, layer
// by layer. Reference for the device training kernels.
double gelu(double x) { return 0.5 * x * (1. + std::erf(x / std::sqrt(2.))); }
double gelu_grad(double x)
{
  return 0.5 * (1. + std::erf(x / std::sqrt(2.))) +
         x * std::exp(-0.5 * x * x) / std::sqrt(2. * 3.14159265358979323846);
}

double mlp_forward_backward(const std::vector<int>& dims, const double* prm, const double* X,
                            const double* Y, int B, double* grads, double* out)
{
  const int L = (int)dims.size() - 1;
  std::vector<Index> woff(L), boff(L);
  Index off = 0;
  for (int l = 0; l < L; l++) {
    woff[l] = off;
    off += (Index)dims[l] * dims[l + 1];
    boff[l] = off;
    off += dims[l + 1];
  }
  std::vector<std::vector<double>> Z(L), A(L);
  for (int l = 0; l < L; l++) {
    const int a = dims[l], b = dims[l + 1];
    const double* in = l == 0 ? X : A[l - 1].data();
    Z[l].resize((size_t)B * b);
    A[l].resize((size_t)B * b);
    for (int s = 0; s < B; s++)
      for (int j = 0; j < b; j++) {
        double z = prm[boff[l] + j];
        for (int i = 0; i < a; i++)
          z += prm[woff[l] + (Index)j * a + i] * in[(Index)s * a + i];
        Z[l][(Index)s * b + j] = z;
        A[l][(Index)s * b + j] = l + 1 < L ? gelu(z) : z;
      }
  }
  const int no = dims.back();
  const double n = (double)B * no;
  double loss = 0.;
  std::vector<double> D((size_t)B * no);
  for (Index i = 0; i < (Index)B * no; i++) {
    const double e = A[L - 1][i] - Y[i];
    loss += e * e / n;
    D[i] = 2. * e / n;
    if (out)
      out[i] = A[L - 1][i];
  }
  if (grads) {
    for (int l = L - 1; l >= 0; l--) {
      const int a = dims[l], b = dims[l + 1];
      const double* in = l == 0 ? X : A[l - 1].data();
      if (l + 1 < L)
        for (Index i = 0; i < (Index)B * b; i++)
          D[i] *= gelu_grad(Z[l][i]);
      for (int j = 0; j < b; j++) {
        double gb = 0.;
        for (int s = 0; s < B; s++)
          gb += D[(Index)s * b + j];
        grads[boff[l] + j] = gb;
        for (int i = 0; i < a; i++) {
          double gw = 0.;
          for (int s = 0; s < B; s++)
            gw += D[(Index)s * b + j] * in[(Index)s * a + i];
          grads[woff[l] + (Index)j * a + i] = gw;
        }
      }
      if (l > 0) {
        std::vector<double> Dn((size_t)B * a, 0.);
        for (int s = 0; s < B; s++)
          for (int i = 0; i < a; i++) {
            double acc = 0.;
            for (int j = 0; j < b; j++)
              acc += D[(Index)s * b + j] * prm[woff[l] + (Index)j * a + i];
            Dn[(Index)s * a + i] = acc;
          }
        D.swap(Dn);
      }
    }
  }
  return loss;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11 "Parallel random numbers: as easy as
// 1, 2, 3"), the counter-based generator of the device throughput path.
// ---------------------------------------------------------------------------
void philox4x32_10(uint32_t c[4], const uint32_t key[2])
{
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// Normal sample `s` of realization `real`: block s/4 of the Philox stream
// keyed by the 64-bit seed; uniforms u = (x + 0.5) 2^-32; Box-Muller on the
// pairs (u0,u1) and (u2,u3): r = sqrt(-2 ln u_a), t = 2 pi u_b,
// n_a = r cos t, n_b = r sin t.
void philox_normals(uint64_t seed, uint64_t real, uint64_t first, Index n, double* out)
{
  const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
  const double TWO_PI = 6.283185307179586476925286766559;
  // samples s and s + 1 (s even) are the cosine and sine normals of one
  // Box-Muller pair; the four samples of block s / 4 share one Philox call
  parallel_for(n, [&](Index b, Index e) {
    uint64_t have = ~(uint64_t)0;
    uint32_t c[4] = { 0, 0, 0, 0 };
    for (Index i = b; i < e; i++) {
      const uint64_t s = first + (uint64_t)i;
      const uint64_t blk = s >> 2;
      if (blk != have) {
        c[0] = (uint32_t)blk;
        c[1] = (uint32_t)(blk >> 32);
        c[2] = (uint32_t)real;
        c[3] = (uint32_t)(real >> 32);
        philox4x32_10(c, key);
        have = blk;
      }
      const int lane = (int)(s & 3);
      const int pa = lane & 2;
      const double ua = ((double)c[pa] + 0.5) * (1.0 / 4294967296.0);
      const double ub = ((double)c[pa + 1] + 0.5) * (1.0 / 4294967296.0);
      const double r = std::sqrt(-2.0 * std::log(ua));
      const double t = TWO_PI * ub;
      out[i] = (lane & 1) ? r * std::sin(t) : r * std::cos(t);
    }
  });
}

// ValueTransform, mutators.hh:35-163.
void value_transform(int kind, double param, double* v, Index n)
{
  for (Index i = 0; i < n; i++) {
    double& x = v[i];
    switch (kind) {
      case PF_TRANSFORM_NONE:
        break;
      case PF_TRANSFORM_LOGNORMAL:
        x = std::exp(x);
        break;
      case PF_TRANSFORM_FOLDEDNORMAL:
        x = std::abs(x);
        break;
      case PF_TRANSFORM_SIGN:
        x = (x < param) ? -1. : 1.;
        break;
      case PF_TRANSFORM_BOXCOX:
        x = std::pow(x * param + 1, -param);
        break;
      default:
        throw std::runtime_error("transform type not known");
    }
  }
}

thread_local std::string g_err;

} // namespace

#define PFO_TRY(...)                                                                               \
  try {                                                                                            \
    __VA_ARGS__;                                                                                   \
  } catch (const std::exception& ex) {                                                             \
    g_err = ex.what();                                                                             \
    return PF_EINVAL;                                                                              \
  }

extern "C" {

const char* pfo_last_error(void) { return g_err.c_str(); }

int pfo_set_threads(int n)
{
  g_threads = n < 1 ? 1 : n;
  return PF_OK;
}

int pfo_sizes(const pf_desc* d, pf_sizes* s)
{
  PFO_TRY({
    Geo g = make_geo(*d);
    s->dim = g.dim;
    for (int i = 0; i < 3; i++) {
      s->cells[i] = (int)g.n[i];
      s->extended[i] = (int)g.N[i];
    }
    s->transposed = g.transposed;
    s->domain_size = g.domain;
    s->extended_size = g.Next;
    Index P = 1;
    for (int i = 0; i < g.dim; i++)
      P *= g.N[i] / 2 + 1;
    const Index half = g.H * g.N[1] * g.N[2];
    s->eig_entries = d->matrix_backend == PF_MATRIX_DCT ? P
                     : d->matrix_backend == PF_MATRIX_DFT ? g.Next
                                                          : half;
    s->noise_per_field = d->field_backend == PF_FIELD_DCTDST ? (P << g.dim)
                         : d->field_backend == PF_FIELD_DFT  ? 2 * g.Next
                                                             : 2 * half;
  });
  return PF_OK;
}

// Eigenvalues (fillTransformedMatrix). `out` (optional) receives them in the
// reference's storage layout. Returns PF_ENEG when negatives exist and
// approximate is off (stats filled either way).
int pfo_eigs(const pf_desc* d, const double* custom, double* out, pf_eig_stats* st)
{
  PFO_TRY({
    Engine* en = get_engine(d, custom);
    if (st)
      *st = en->e.st;
    if (out) {
      const Geo& g = en->g;
      if (d->matrix_backend == PF_MATRIX_DFT) {
        for (Index s = 0; s < g.Next; s++) {
          Index k[3];
          dft_storage_to_k(g, s, k);
          out[s] = en->e.v[k[0] + g.N[0] * (k[1] + g.N[1] * k[2])];
        }
      } else if (d->matrix_backend == PF_MATRIX_DCT) {
        Index P[3];
        for (int i = 0; i < 3; i++)
          P[i] = (i < g.dim) ? g.N[i] / 2 + 1 : 1;
        const Index count = P[0] * P[1] * P[2];
        for (Index s = 0; s < count; s++) {
          Index k[3];
          dct_storage_to_k(g, P, s, k);
          out[s] = en->e.v[k[0] + P[0] * (k[1] + P[1] * k[2])];
        }
      } else {
        const Index count = g.H * g.N[1] * g.N[2];
        for (Index s = 0; s < count; s++) {
          Index k[3];
          spec_storage_to_k(g, s, k);
          out[s] = en->e.v[k[0] + g.H * (k[1] + g.N[1] * k[2])];
        }
      }
    }
    int status = en->status;
    if (custom)
      delete en;
    return status;
  });
}

// One realization: noise consumed in generateField's order, untransformed.
// DFT field: one draw (2 N_ext normals), `field` receives 2 domain values:
// the field, then the spare field (matrix.hh:411-420).
int pfo_generate(const pf_desc* d, const double* noise, double* field)
{
  PFO_TRY({
    Engine* en = get_engine(d);
    if (en->status != PF_OK) {
      g_err = "negative eigenvalues in covariance matrix";
      return en->status;
    }
    if (d->field_backend == PF_FIELD_DCTDST)
      generate_dctdst(en->d, en->g, en->e, noise, field);
    else if (d->field_backend == PF_FIELD_DFT)
      generate_dft(en->d, en->g, en->e, noise, field, field + en->g.domain);
    else
      generate_r2c(en->d, en->g, en->e, noise, field);
  });
  return PF_OK;
}

int pfo_matvec(const pf_desc* d, int op, const double* x, double* y)
{
  PFO_TRY({
    Engine* en = get_engine(d);
    if (en->status != PF_OK) {
      g_err = "negative eigenvalues in covariance matrix";
      return en->status;
    }
    matvec(*en, op, x, y);
  });
  return PF_OK;
}

// Matrix::multiplyInverse, matrix.hh:189-206.
int pfo_multiply_inverse(const pf_desc* d, const double* b, double* x, int* iterations)
{
  PFO_TRY({
    Engine* en = get_engine(d);
    if (en->status != PF_OK) {
      g_err = "negative eigenvalues in covariance matrix";
      return en->status;
    }
    const Index n = en->g.domain;
    std::vector<double> input(b, b + n), output(b, b + n);
    bool fieldZero = true;
    for (Index i = 0; i < n; i++)
      if (std::abs(input[i]) > 1e-10)
        fieldZero = false;
    int it = 0;
    if (!fieldZero) {
      std::vector<double> tmp(n);
      matvec(*en, PF_OP_INVERSE_EMBED, output.data(), tmp.data());
      output = tmp;
      it = inner_cg(*en, output, input, true, (unsigned)d->cg_iterations);
    }
    std::copy(output.begin(), output.end(), x);
    if (iterations)
      *iterations = it;
  });
  return PF_OK;
}

// Relative-residual (P)CG: mode PF_CG_PLAIN or PF_CG_EMBED_PCG.
int pfo_pcg(const pf_desc* d, int mode, int maxit, double rel_tol, const double* b, double* x,
            int* iterations, double* relres)
{
  PFO_TRY({
    Engine* en = get_engine(d);
    if (en->status != PF_OK) {
      g_err = "negative eigenvalues in covariance matrix";
      return en->status;
    }
    const Index n = en->g.domain;
    g_inv_floor = d->nugget / (double)en->g.Next;
    std::vector<double> bb(b, b + n), xx;
    std::function<void(const std::vector<double>&, std::vector<double>&)> prec =
      [&](const std::vector<double>& r, std::vector<double>& z) {
        z.resize(r.size());
        // embedding inverse floored at nugget / N_ext when there is a nugget
        const int op = mode == PF_CG_LEARNED_PCG ? 3 : d->nugget > 0. ? 4 : PF_OP_INVERSE_EMBED;
        matvec(*en, op, r.data(), z.data());
      };
    int it = pcg(*en, bb, xx, mode == PF_CG_PLAIN ? nullptr : &prec, maxit, rel_tol, relres);
    std::copy(xx.begin(), xx.end(), x);
    if (iterations)
      *iterations = it;
  });
  return PF_OK;
}

int pfo_value_transform(int kind, double param, double* v, long long n)
{
  PFO_TRY(value_transform(kind, param, v, n));
  return PF_OK;
}

// CppRNGBackend (cpprngbackend.hh:57-100): std::default_random_engine seeded
// with seed (+ rank 0, randomfield.hh:525-526), std::normal_distribution<double>.
int pfo_noise_cpp(unsigned seed, long long n, double* out)
{
  std::default_random_engine generator;
  std::normal_distribution<double> normalDist(0., 1.);
  generator.seed(seed);
  for (long long i = 0; i < n; i++)
    out[i] = normalDist(generator);
  return PF_OK;
}

int pfo_noise_philox(uint64_t seed, uint64_t realization, uint64_t first, long long n,
                     double* out)
{
  philox_normals(seed, realization, first, n, out);
  return PF_OK;
}

int pfo_philox_raw(const uint32_t ctr[4], const uint32_t key[2], long long nblocks,
                   uint32_t* out)
{
  for (long long i = 0; i < nblocks; i++) {
    uint32_t c[4] = { ctr[0] + (uint32_t)i, ctr[1], ctr[2], ctr[3] };
    philox4x32_10(c, key);
    for (int j = 0; j < 4; j++)
      out[4 * i + j] = c[j];
  }
  return PF_OK;
}

// Covariance functor (with geometry) at `n` points of raw coordinates.
int pfo_cov_eval(const pf_desc* d, const double* x, long long n, double* out)
{
  PFO_TRY({
    Geometry geo(*d);
    Covariance cov(*d);
    for (long long i = 0; i < n; i++) {
      double t[3] = { 0, 0, 0 };
      geo.transform(x + i * d->dim, t);
      out[i] = cov(t);
    }
  });
  return PF_OK;
}

double pfo_bessel_k(double nu, double x) { return bessel_k(nu, x); }

// Embedding grid before the forward transform (classical or merge), natural
// x-fastest layout: the DCT half grid (N_i/2+1 per axis) or the full
// extended grid (R2C matrix).
int pfo_cov_grid(const pf_desc* d, double* out)
{
  PFO_TRY({
    Geo g = make_geo(*d);
    Geometry geo(*d);
    Covariance cov(*d);
    Index C[3];
    for (int i = 0; i < 3; i++)
      C[i] = i >= g.dim ? 1 : (d->matrix_backend == PF_MATRIX_DCT ? g.N[i] / 2 + 1 : g.N[i]);
    const Index total = C[0] * C[1] * C[2];
    parallel_for(total, [&](Index b, Index en) {
      for (Index f = b; f < en; f++) {
        Index idx[3] = { f % C[0], (f / C[0]) % C[1], f / (C[0] * C[1]) };
        out[f] = embedded_cov(g, geo, cov, *d, idx);
      }
    });
    if (d->periodization == PF_PERIOD_COFOLD)
      cofold(*d, g, C, [&](Index f) -> double& { return out[f]; });
  });
  return PF_OK;
}

// Sigmoids of the merge embedding, for the tests: kind 0 clamping, 1 smooth,
// 2 smoothstep with `rec` recursions.
double pfo_sigmoid(int kind, double one, double zero, double x, int rec)
{
  if (kind == 0)
    return sigmoid_clamp(one, zero, x);
  if (kind == 1)
    return sigmoid_smooth(one, zero, x);
  return sigmoid_smoothstep(one, zero, x, (unsigned)rec);
}
// Fold taper integral, for the tests (sigmoid: enum pf_sigmoid).
double pfo_fold_integral(int sigmoid, double beta, double norm, int rec)
{
  return fold_integral(sigmoid, beta, norm, (unsigned)rec);
}

// Integral of sigmoid(one, zero, x) over [one, hi], for the tests.
double pfo_taper_integral(int sigmoid, double one, double zero, double hi, int rec)
{
  return taper_integral(sigmoid, one, zero, hi, (unsigned)rec);
}
// Learned-PCG: preconditioner R F^{-1} diag(exp(L_bin)) F E.
int pfo_pcg_learned(const pf_desc* d, const double* log_mult, int nbins, int maxit, double rel_tol,
                    const double* b, double* x, int* iterations, double* relres)
{
  g_spec_table.resize(nbins);
  for (int i = 0; i < nbins; i++)
    g_spec_table[i] = std::exp(log_mult[i]);
  return pfo_pcg(d, PF_CG_LEARNED_PCG, maxit, rel_tol, b, x, iterations, relres);
}

// Radial targets -log(mean_bin max(lambda, floor)) over the DCT-half array.
int pfo_spectral_targets(const pf_desc* d, int nbins, double floor_value, double* out)
{
  PFO_TRY({
    Engine* en = get_engine(d);
    if (d->matrix_backend != PF_MATRIX_DCT)
      throw std::runtime_error("spectral targets need the DCT matrix backend");
    const Geo& g = en->g;
    Index P[3];
    for (int i = 0; i < 3; i++)
      P[i] = (i < g.dim) ? g.N[i] / 2 + 1 : 1;
    std::vector<double> sums(nbins, 0.);
    std::vector<long long> counts(nbins, 0);
    for (Index c = 0; c < P[2]; c++)
      for (Index b2 = 0; b2 < P[1]; b2++)
        for (Index a = 0; a < P[0]; a++) {
          const int bin = radial_bin(g, a, b2, c, nbins);
          const double v = en->e.v[a + P[0] * (b2 + P[1] * c)];
          sums[bin] += v > floor_value ? v : floor_value;
          counts[bin]++;
        }
    double last = 0.;
    for (int b = 0; b < nbins; b++) {
      if (counts[b] > 0)
        last = -std::log(sums[b] / counts[b]);
      out[b] = counts[b] > 0 ? last : NAN;
    }
    for (int b = nbins - 1; b >= 0; b--)
      if (std::isnan(out[b]))
        out[b] = (b + 1 < nbins) ? out[b + 1] : last;
  });
  return PF_OK;
}

// MLP loss, gradients (optional) and outputs (optional), FP64.
double pfo_mlp(const int* dims, int ndims, const double* params, const double* x, const double* y,
               int batch, double* grads, double* out)
{
  std::vector<int> dv(dims, dims + ndims);
  return mlp_forward_backward(dv, params, x, y, batch, grads, out);
}


} // extern "C"
