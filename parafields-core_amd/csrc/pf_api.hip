// pf_api.hip -- plan objects and the extern "C" ABI of libparafields_amd.so.
//
// A plan is the device-side state of one reference `Matrix`
// (include/parafields/matrix.hh:55-462): the configuration, the lazily built
// FP64 embedding eigenvalues, the per-precision sqrt table and the work
// buffers of the field transforms. All work of a plan runs on its own HIP
// stream; calls on a plan are serialised by a mutex; every call returns after
// its device work has completed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "pf_dctdst.hpp"
#include "pf_dft.hpp"
#include "pf_fast.hpp"
#include "pf_fast_launch.hpp"
#include "pf_kernels.hpp"
#include "pf_mlp.hpp"

using namespace pfamd;

namespace {

thread_local std::string g_err;

struct PfError : std::runtime_error
{
  int code;
  PfError(int c, const std::string& m)
    : std::runtime_error(m)
    , code(c)
  {
  }
};

#define HIPCHECK(x)                                                                                \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw PfError(PF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));                      \
  } while (0)

struct DevBuf
{
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b)
  {
    if (b <= bytes)
      return;
    release();
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      bytes = 0;
      throw PfError(PF_ENOMEM, "device allocation of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
  }
  void release()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename X>
  X* as() const
  {
    return reinterpret_cast<X*>(p);
  }
  ~DevBuf() { release(); }
};

// Radix schedule of an n-point Stockham transform. order 0: radix 8 first,
// then 4 / 2, then 3, 5, 7; order 1: the same radices in reverse order;
// order 2: radix 4 (and one 2) instead of 8. Every order computes the same
// transform; the autotuner picks one per line length for the generic kernels.
Sched make_sched(int n, int order = 0)
{
  Sched s;
  s.n = n;
  s.nst = 0;
  int m = n, p2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    p2++;
  }
  std::vector<int> r;
  if (order == 2) {
    for (int i = 0; i < p2 / 2; i++)
      r.push_back(4);
    if (p2 % 2)
      r.push_back(2);
  } else {
    int n8 = p2 / 3, rem = p2 % 3;
    if (rem == 1 && n8 > 0) {
      n8--;
      r.push_back(4);
      r.push_back(4);
    } else if (rem == 1)
      r.push_back(2);
    else if (rem == 2)
      r.push_back(4);
    for (int i = 0; i < n8; i++)
      r.insert(r.begin(), 8);
  }
  for (int q : { 3, 5, 7 })
    while (m % q == 0) {
      m /= q;
      r.push_back(q);
    }
  if (order == 1)
    std::reverse(r.begin(), r.end());
  if (m != 1)
    throw PfError(PF_EINVAL, "transform length " + std::to_string(n) +
                               " has a prime factor above 7 (supported: 2, 3, 5, 7)");
  if ((int)r.size() > kMaxStages)
    throw PfError(PF_EINVAL, "transform length too long");
  for (size_t i = 0; i < r.size(); i++)
    s.rad[i] = r[i];
  s.nst = (int)r.size();
  return s;
}

template <typename T>
constexpr int maxe()
{
  return sizeof(T) == 4 ? 32 : 16;
}

struct Cfg
{
  int L, threads;
  size_t lds;
};

// Tile of L lines of length n (+ a twiddle table of ntab entries) within 64 KB
// of LDS, one thread per MAXE elements.
size_t lds_budget()
{
  static size_t budget = 0;
  if (budget == 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    budget = std::min<size_t>((size_t)v, 65536);
    // PF_LDS_BUDGET can only lower the 64 KB cap
    if (const char* e = getenv("PF_LDS_BUDGET"))
      budget = std::min<size_t>(budget, (size_t)atol(e));
  }
  return budget;
}

// LDS budget of the specialised kernels; PF_NO_FAST=1 routes every pass
// through the generic kernels (used by the tests to check both).
size_t fast_budget() { return getenv("PF_NO_FAST") ? 0 : lds_budget(); }

// LDS budget of the specialised DCT/DST passes (pf_dctdst_fast.hpp), whose
// tiles are at most 64 KB. Two 1024^3 runs of the DCT/DST generation with
// tiles above 64 KB of LDS (an earlier form of the specialised kernels with a
// pad line, 73.8 KB, then the generic ones with 80 KB tiles), each after
// other workloads in the same process, ended in GPU memory faults whose cause
// is not identified; no kernel here takes more than 64 KB of LDS since.
size_t dct_fast_budget() { return fast_budget(); }

// Tiles of the generic DCT/DST kernels: at most 64 KB (see above).
size_t dct_generic_budget() { return lds_budget(); }

template <typename T>
Cfg make_cfg(int n, int64_t nlines, int ntab, int max_lines = 0)
{
  const size_t cs = sizeof(cplx<T>);
  int L = sizeof(T) == 4 ? 16 : 8;
  while (max_lines > 0 && L > max_lines)
    L /= 2;
  while (L > 1 && ((size_t)L * n + ntab) * cs > lds_budget())
    L /= 2;
  if (((size_t)L * n + ntab) * cs > lds_budget())
    throw PfError(PF_EINVAL, "line length " + std::to_string(n) + " exceeds the LDS budget");
  while (L > 1 && L / 2 >= nlines)
    L /= 2;
  // small problems: fewer lines per tile until the launch has ~2 workgroups
  // per CU (a 71-line pass on 16-line tiles is 5 workgroups)
  while (L > 1 && (nlines + L - 1) / L < 512)
    L /= 2;
  const int me = maxe<T>();
  int threads = (L * n + me - 1) / me;
  threads = ((threads + 63) / 64) * 64;
  if (threads < 64)
    threads = 64;
  if (threads > 1024)
    throw PfError(PF_EINVAL, "line length " + std::to_string(n) + " too long for one workgroup");
  Cfg c;
  c.L = L;
  c.threads = threads;
  c.lds = (size_t)(ntab + (size_t)L * n) * cs;
  return c;
}

inline unsigned nblocks(int64_t lines, int L)
{
  return (unsigned)((lines + L - 1) / L);
}

struct Geo
{
  int dim, factor, transposed;
  int n[3], N[3], H, P[3];
  int Hp;  // row pitch of the complex work arrays: H rounded up to 16 entries (128-byte rows)
  double h[3], ext[3];
  int64_t Next, domain, eig_entries, noise_per_field, half;
};

struct Timing
{
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev;
};

// Launch choices an autotuner can change without recompiling: the
// analogue of the FFTW planner choices the reference keeps in wisdom files
// (fftw.useWisdom, dftmatrixbackend.hh:52-70). Zero / -1 keep the defaults.
struct Tuning
{
  int ldiv = 0;     // y and x passes of the R2C generation: tiles of L / ldiv lines
  int chunks = 0;   // chunks of the two-stream generation pipeline (0: 8)
  int overlap = -1; // two-stream pipeline: 0 off, 1 on, -1 default (PF_OVERLAP)
  int lines = 0;    // generic kernels: at most this many lines per tile
  int radix = 0;    // generic kernels: radix order of make_sched
  int sub = 0;      // DCT/DST and DFT fields: realizations per sub-batch (0: the work cap)
};

struct Plan
{
  pf_desc d;
  Geo g;
  Tuning tune;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  bool capturing = false;
  hipStream_t stream2 = nullptr;     // second stream of the pipelined generation
  hipEvent_t ev_join = nullptr;
  std::array<hipEvent_t, 8> ev_chunk{};
  std::mutex mu;
  bool eig_valid = false;
  int eig_status = PF_OK;
  pf_eig_stats st{};
  DevBuf eig;       // FP64 eigenvalues (logical layout)
  DevBuf sq32, sq64; // sqrt tables
  DevBuf w1, w2, w3, noise, stage_in, stage_out, partial, scratch;
  DevBuf spec;       // learned radial multipliers exp(L_b) (FP64)
  int spec_nbins = 0;
  DevBuf floor;      // PF_OP_INVERSE_FLOOR: the eigenvalue floor (one FP64)
  double floor_value = 0.;
  std::map<std::pair<int, int>, std::unique_ptr<DevBuf>> tw; // (precision bytes, n) -> table
  bool profiling = false;
  // set during pf_generate_noalloc: a lookup that would allocate (a missing
  // twiddle table) fails instead, and no profiling events are created
  bool noalloc = false;
  std::vector<std::pair<std::string, double>> last_timings;
  std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> pending;

  ~Plan()
  {
    for (auto& p : pending) {
      (void)hipEventDestroy(std::get<1>(p));
      (void)hipEventDestroy(std::get<2>(p));
    }
    if (stream2) {
      (void)hipStreamSynchronize(stream2);
      (void)hipStreamDestroy(stream2);
      (void)hipEventDestroy(ev_join);
      for (auto e : ev_chunk)
        (void)hipEventDestroy(e);
    }
    if (stream && own_stream)
      (void)hipStreamDestroy(stream);
  }

  template <typename T>
  const cplx<T>* table(int n)
  {
    auto key = std::make_pair((int)sizeof(T), n);
    auto it = tw.find(key);
    if (it != tw.end())
      return it->second->template as<cplx<T>>();
    if (noalloc)
      throw PfError(PF_ESTATE, "plan not prepared for generation without allocation: "
                               "call pf_eigs first");
    std::vector<cplx<T>> h(n);
    for (int j = 0; j < n; j++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * j / n;
      h[j] = { (T)cosl(a), (T)sinl(a) };
    }
    auto b = std::make_unique<DevBuf>();
    b->ensure(n * sizeof(cplx<T>));
    HIPCHECK(hipMemcpyAsync(b->p, h.data(), n * sizeof(cplx<T>), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    const cplx<T>* ptr = b->template as<cplx<T>>();
    tw[key] = std::move(b);
    return ptr;
  }

  // optional per-launch timing: events on the stream the launch goes to
  void mark_begin(const char* name, hipEvent_t& a, hipEvent_t& b, hipStream_t s)
  {
    if (!profiling || noalloc)
      return;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, s));
    pending.emplace_back(name, a, b);
  }
  void mark_end(hipEvent_t b, hipStream_t s)
  {
    if (!profiling || noalloc)
      return;
    HIPCHECK(hipEventRecord(b, s));
  }
  void collect_timings()
  {
    last_timings.clear();
    for (auto& p : pending) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, std::get<1>(p), std::get<2>(p)));
      last_timings.emplace_back(std::get<0>(p), (double)ms);
      (void)hipEventDestroy(std::get<1>(p));
      (void)hipEventDestroy(std::get<2>(p));
    }
    pending.clear();
  }
  void sync()
  {
    HIPCHECK(hipStreamSynchronize(stream));
    collect_timings();
  }
};

struct Scope
{
  Plan* p;
  hipStream_t s;
  hipEvent_t a{}, b{};
  Scope(Plan* p_, const char* name, hipStream_t s_ = nullptr)
    : p(p_)
    , s(s_ ? s_ : p_->stream)
  {
    p->mark_begin(name, a, b, s);
  }
  ~Scope() noexcept(false) { p->mark_end(b, s); }
};

// Launch configuration of a generic kernel under the plan's tuning.
template <typename T>
Cfg pcfg(const Plan& P, int n, int64_t nlines, int ntab)
{
  return make_cfg<T>(n, nlines, ntab, P.tune.lines);
}
Sched psched(const Plan& P, int n) { return make_sched(n, P.tune.radix); }

void check_launch()
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw PfError(PF_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

// embedding.mergeRecursions / embedding.foldRecursions (smoothstep taper of
// the merge and fold embeddings): the reference's value 99, its default,
// picks a count per covariance family (matrix.hh:679-705, :805-829, the same
// rule) and throws for the others.
int merge_recursions(const pf_desc& d)
{
  if (d.merge_recursions > 0)
    return d.merge_recursions;
  if (d.merge_recursions < 0)
    return 0;
  const double half_dim = (double)d.dim / 2.;
  switch (d.covariance) {
    case PF_COV_MATERN:
      return (int)std::floor(d.matern_nu * 2. + half_dim);
    case PF_COV_GAMMA_EXPONENTIAL:
      return (int)std::floor(d.exp_gamma + half_dim);
    case PF_COV_GENERALIZED_CAUCHY:
      return (int)std::floor(d.cauchy_alpha + half_dim);
    case PF_COV_EXPONENTIAL:
      return 2;
    case PF_COV_MATERN32:
      return 4;
    case PF_COV_MATERN52:
      return 6;
    case PF_COV_CAUCHY:
      return 3;
    default:
      throw PfError(PF_EINVAL,
                    "magic recursions value 99 not implemented for choice of covariance");
  }
}

void dft_split_1d(int64_t N, size_t cs, int& A, int& B);

Geo make_geo(const pf_desc& d)
{
  Geo g{};
  if (d.dim < 1 || d.dim > 3)
    throw PfError(PF_EINVAL, "dim must be 1, 2 or 3");
  g.dim = d.dim;
  g.factor = d.periodic ? 1 : d.embedding_factor;
  if (g.factor < 1)
    throw PfError(PF_EINVAL, "embedding.factor must be >= 1");
  g.Next = 1;
  g.domain = 1;
  for (int i = 0; i < 3; i++) {
    if (i < d.dim) {
      if (d.cells[i] < 1)
        throw PfError(PF_EINVAL, "grid.cells must be positive");
      g.n[i] = d.cells[i];
      g.N[i] = g.factor * d.cells[i];
      g.ext[i] = d.extensions[i];
      g.h[i] = d.extensions[i] / d.cells[i];
    } else {
      g.n[i] = g.N[i] = 1;
      g.ext[i] = g.h[i] = 1.;
    }
    g.P[i] = (i < d.dim) ? g.N[i] / 2 + 1 : 1;
    g.Next *= g.N[i];
    g.domain *= g.n[i];
  }
  if (g.N[0] % 2 != 0)
    throw PfError(PF_EINVAL, "the extended x size must be even");
  g.H = g.N[0] / 2 + 1;
  g.Hp = (g.H + 15) / 16 * 16;
  if (d.dim == 1)
    g.transposed = 0; // fieldtraits.hh:229-233: transposed output needs two dims
  else if (d.transposed >= 0)
    g.transposed = d.transposed != 0;
  else if (d.dim == 2)
    g.transposed = !(d.matrix_backend == PF_MATRIX_R2C || d.field_backend == PF_FIELD_R2C);
  else
    g.transposed = 1;
  if (d.field_backend == PF_FIELD_R2C && d.dim == 2 && g.transposed)
    throw PfError(PF_EINVAL, "R2CFieldBackend supports transposed output only for dim > 2");
  if (d.matrix_backend != PF_MATRIX_DCT && d.matrix_backend != PF_MATRIX_R2C &&
      d.matrix_backend != PF_MATRIX_DFT)
    throw PfError(PF_EINVAL, "unknown matrix backend");
  if (d.field_backend != PF_FIELD_R2C && d.field_backend != PF_FIELD_DCTDST &&
      d.field_backend != PF_FIELD_DFT)
    throw PfError(PF_EINVAL, "unknown field backend");
  // 1-D fields: the reference's default DFTMatrixBackend + DFTFieldBackend
  // (matrix.hh:1660-1736)
  if (d.dim == 1 && (d.matrix_backend != PF_MATRIX_DFT || d.field_backend != PF_FIELD_DFT))
    throw PfError(PF_EINVAL, "1-D fields use DFTMatrixBackend and DFTFieldBackend");
  if (d.field_backend == PF_FIELD_DFT && d.matrix_backend == PF_MATRIX_R2C)
    throw PfError(PF_EINVAL, "DFTFieldBackend needs DFTMatrixBackend or DCTMatrixBackend");
  if (d.field_backend == PF_FIELD_DCTDST && d.matrix_backend != PF_MATRIX_DCT)
    throw PfError(PF_EINVAL, "DCTDSTFieldBackend requires DCTMatrixBackend");
  // the reference's DCT/DST extraction walk (dctdstfieldbackend.hh:905-930)
  // addresses the first n_i entries per axis only for embedding factor 2
  if (d.field_backend == PF_FIELD_DCTDST && g.factor != 2)
    throw PfError(PF_EINVAL, "DCTDSTFieldBackend needs embedding.factor 2");
  if (d.covariance < 0 || d.covariance > PF_COV_CUSTOM)
    throw PfError(PF_EINVAL, "covariance structure not known");
  if (d.covariance == PF_COV_CUBIC && d.dim == 3)
    throw PfError(PF_EINVAL, "cubic covariance only applicable in 1D or 2D");
  if (d.covariance == PF_COV_GAMMA_EXPONENTIAL && (d.exp_gamma < 0. || d.exp_gamma > 2.))
    throw PfError(PF_EINVAL, "exponent gamma has to be between 0 and 2");
  if (d.covariance == PF_COV_GENERALIZED_CAUCHY &&
      (d.cauchy_alpha <= 0. || d.cauchy_alpha > 2. || d.cauchy_beta <= 0.))
    throw PfError(PF_EINVAL, "generalized Cauchy parameters out of range");
  if (d.covariance == PF_COV_MATERN && d.matern_nu < 0.)
    throw PfError(PF_EINVAL, "matern nu has to be positive");
  if (d.anisotropy < 0 || d.anisotropy > 2)
    throw PfError(PF_EINVAL, "stochastic.anisotropy must be none, axiparallel or geometric");
  if (d.precision != PF_FP64 && d.precision != PF_FP32)
    throw PfError(PF_EINVAL, "precision must be PF_FP64 or PF_FP32");
  if (d.periodization < PF_PERIOD_CLASSICAL || d.periodization > PF_PERIOD_COFOLD)
    throw PfError(PF_EINVAL, "stochastic.periodization must be \"classical\", \"merge\", \"fold\", "
                             "or \"cofold\"");
  if (d.sigmoid != PF_SIGMOID_SMOOTHSTEP && d.sigmoid != PF_SIGMOID_SMOOTH)
    throw PfError(PF_EINVAL, "embedding.sigmoid must be \"smooth\" or \"smoothstep\"");
  if (d.sigmoid_combine < PF_COMBINE_RADIAL_PLUS || d.sigmoid_combine > PF_COMBINE_TENSOR)
    throw PfError(PF_EINVAL, "sigmoidCombine must be tensor, radial or radialPlus");
  if (d.periodization == PF_PERIOD_MERGE || d.periodization == PF_PERIOD_FOLD)
    (void)merge_recursions(d); // rejects families without a default count
  if (!(d.fold_max_factor >= 0.))
    throw PfError(PF_EINVAL, "embedding.foldMaxFactor must be positive");
  const int64_t half = (int64_t)g.H * g.N[1] * g.N[2];
  const int64_t P = (int64_t)g.P[0] * g.P[1] * g.P[2];
  g.half = half;
  g.eig_entries = d.matrix_backend == PF_MATRIX_DCT ? P : half;
  // DFT field: the 2 N_ext normals of one draw, which fills two fields
  g.noise_per_field = d.field_backend == PF_FIELD_DCTDST ? (P << g.dim)
                      : d.field_backend == PF_FIELD_DFT  ? 2 * g.Next
                                                         : 2 * half;
  // every transform length must factor into 2, 3, 5, 7
  for (int i = 0; i < d.dim; i++)
    (void)make_sched(g.N[i]);
  (void)make_sched(g.N[0] / 2);
  if (d.dim == 1) {
    // the FP64 eigenvalue and product transforms must split into two tiles
    int A, B;
    dft_split_1d(g.N[0], sizeof(cplx<double>), A, B);
  }
  return g;
}

CovParams make_cov(const pf_desc& d, const Geo& g)
{
  CovParams p{};
  p.dim = d.dim;
  p.covariance = d.covariance;
  p.variance = d.variance;
  p.aniso = d.anisotropy;
  if (p.aniso == PF_ANISO_NONE || d.covariance == PF_COV_CUSTOM) {
    p.aniso = PF_ANISO_NONE;
    p.m[0] = 1. / d.corr_length[0];
  } else if (p.aniso == PF_ANISO_AXIPARALLEL) {
    for (int i = 0; i < d.dim; i++)
      p.m[i] = 1. / d.corr_length[i];
  } else if (d.dim == 1) {
    p.m[0] = 1. / d.corr_length[0]; // GeneralMatrix in 1D, covariance.hh:124-125
  } else {
    const double* c = d.corr_length;
    if (d.dim == 3) {
      p.m[0] = c[4] * c[8] - c[5] * c[7];
      p.m[1] = c[2] * c[7] - c[1] * c[8];
      p.m[2] = c[1] * c[5] - c[2] * c[4];
      p.m[3] = c[5] * c[6] - c[3] * c[8];
      p.m[4] = c[0] * c[8] - c[2] * c[6];
      p.m[5] = c[2] * c[3] - c[0] * c[5];
      p.m[6] = c[3] * c[7] - c[4] * c[6];
      p.m[7] = c[1] * c[6] - c[0] * c[7];
      p.m[8] = c[0] * c[4] - c[1] * c[3];
      const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                         c[2] * (c[3] * c[7] - c[4] * c[6]);
      for (int i = 0; i < 9; i++)
        p.m[i] /= det;
    } else {
      p.m[0] = c[3];
      p.m[1] = -c[1];
      p.m[2] = -c[2];
      p.m[3] = c[0];
      const double det = c[0] * c[3] - c[1] * c[2];
      for (int i = 0; i < 4; i++)
        p.m[i] /= det;
    }
  }
  p.nu = d.matern_nu;
  p.sqrt2nu = std::sqrt(2. * d.matern_nu);
  p.matern_c = d.covariance == PF_COV_MATERN ? std::pow(2., 1. - d.matern_nu) / std::tgamma(d.matern_nu)
                                              : 0.;
  p.gamma = d.exp_gamma;
  p.alpha = d.cauchy_alpha;
  p.beta = d.cauchy_beta;
  p.nugget = d.nugget;
  for (int i = 0; i < 3; i++) {
    p.h[i] = g.h[i];
    p.wrap_half[i] = 0.5 * g.ext[i] * (double)g.factor;
    p.wrap_len[i] = g.ext[i] * (double)g.factor;
  }
  p.factor = g.factor;
  p.period = d.periodization;
  if (p.period == PF_PERIOD_FOLD) {
    // matrix.hh:800-834: embedding.foldMaxFactor (default 1) and
    // embedding.foldRecursions (same family rule as mergeRecursions)
    p.sigmoid = d.sigmoid;
    p.recursions = merge_recursions(d);
    const double max_factor = d.fold_max_factor != 0. ? d.fold_max_factor : 1.;
    p.fold_beta = max_factor * 0.5 * (double)g.factor / std::sqrt((double)d.dim);
    for (int i = 0; i < 3; i++)
      p.ext[i] = g.ext[i];
  }
  if (p.period == PF_PERIOD_COFOLD) {
    // matrix.hh:945-951: embedding.cofoldRecursions (default 1); the border
    // band is one cell of the extended grid (matrix.hh:936-941)
    p.sigmoid = d.sigmoid;
    p.recursions = d.merge_recursions > 0 ? d.merge_recursions : (d.merge_recursions < 0 ? 0 : 1);
    for (int i = 0; i < 3; i++) {
      p.ext[i] = g.ext[i];
      p.border[i] = g.ext[i] / (double)g.N[i];
    }
  }
  if (p.period == PF_PERIOD_MERGE) {
    p.sigmoid = d.sigmoid;
    p.radial = d.sigmoid_combine != PF_COMBINE_TENSOR &&
               std::sqrt((double)d.dim) < (double)g.factor / 2.;
    p.radial_plus = d.sigmoid_combine == PF_COMBINE_RADIAL_PLUS;
    p.recursions = merge_recursions(d);
    p.sig_start = d.sigmoid_start != 0. ? d.sigmoid_start : 1.;
    p.sig_end = d.sigmoid_end != 0. ? d.sigmoid_end : (double)g.factor - 1.;
    for (int i = 0; i < 3; i++)
      p.ext[i] = g.ext[i];
  }
  return p;
}

GenGeom gen_geom(const Plan& P)
{
  const Geo& g = P.g;
  GenGeom gg;
  gg.dim = g.dim;
  gg.transposed = g.transposed;
  gg.N0 = g.N[0];
  gg.N1 = g.N[1];
  gg.N2 = g.N[2];
  gg.H = g.H;
  gg.n0 = g.n[0];
  gg.n1 = g.n[1];
  gg.n2 = g.n[2];
  gg.P0 = g.P[0];
  gg.P1 = g.P[1];
  gg.P2 = g.P[2];
  gg.dct = P.d.matrix_backend == PF_MATRIX_DCT;
  gg.noise_stride = g.noise_per_field;
  gg.Hp = g.Hp;
  return gg;
}

// f(std::integral_constant<int, family>) for the covariance family `cov`
// (-1 for families without an instantiation of their own).
// the covariance kernels' periodization as a compile-time constant
template <typename F>
void with_period(int per, F&& f)
{
  if (per == PF_PERIOD_MERGE)
    f(std::integral_constant<int, 1>{});
  else if (per == PF_PERIOD_FOLD)
    f(std::integral_constant<int, 2>{});
  else
    f(std::integral_constant<int, 0>{});
}

template <typename F>
void with_cov(int cov, F&& f)
{
  switch (cov) {
#define PF_COV_CASE(c)                                                                             \
  case c:                                                                                          \
    f(std::integral_constant<int, c>{});                                                           \
    break;
    PF_COV_CASE(PF_COV_EXPONENTIAL)
    PF_COV_CASE(PF_COV_GAUSSIAN)
    PF_COV_CASE(PF_COV_SPHERICAL)
    PF_COV_CASE(PF_COV_SEPARABLE_EXPONENTIAL)
    PF_COV_CASE(PF_COV_MATERN)
    PF_COV_CASE(PF_COV_MATERN32)
    PF_COV_CASE(PF_COV_MATERN52)
    PF_COV_CASE(PF_COV_DAMPED_OSCILLATION)
    PF_COV_CASE(PF_COV_GAMMA_EXPONENTIAL)
    PF_COV_CASE(PF_COV_CAUCHY)
    PF_COV_CASE(PF_COV_GENERALIZED_CAUCHY)
    PF_COV_CASE(PF_COV_CUBIC)
    PF_COV_CASE(PF_COV_WHITE_NOISE)
#undef PF_COV_CASE
    default:
      f(std::integral_constant<int, -1>{});
  }
}

unsigned grid1d(int64_t n, int threads = 256)
{
  int64_t b = (n + threads - 1) / threads;
  if (b > 65536)
    b = 65536;
  if (b < 1)
    b = 1;
  return (unsigned)b;
}

// Cofold (Matrix::modifyMatrixEntriesWithCofold, matrix.hh:905-1013) on the
// classical entries of the dense grid a[n2][n1][n0]: block minima / maxima on
// the device, their reduction and the offset on the host, then the remap of
// the entries beyond the domain on the device. alpha is the largest border
// entry, beta = min(cofoldMaxFactor * alpha, smallest entry on the domain).
void apply_cofold(Plan& P, const CovParams& cp, double* a, int n0, int n1, int n2)
{
  const int blocks = 512;
  DevBuf part;
  part.ensure(2 * blocks * sizeof(double));
  k_cofold_bounds<<<blocks, 256, 0, P.stream>>>(cp, a, n0, n1, n2, part.as<double>());
  check_launch();
  std::vector<double> h(2 * blocks);
  HIPCHECK(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          P.stream));
  HIPCHECK(hipStreamSynchronize(P.stream));
  part.release();
  double min_constrained = 1.7976931348623157e308, max_border = -1.7976931348623157e308;
  for (int b = 0; b < blocks; b++) {
    if (h[2 * b] < min_constrained)
      min_constrained = h[2 * b];
    if (h[2 * b + 1] > max_border)
      max_border = h[2 * b + 1];
  }
  const double max_factor = P.d.fold_max_factor != 0. ? P.d.fold_max_factor : 5.;
  const double alpha = max_border, beta = std::min(max_factor * max_border, min_constrained);
  const double offset = beta - cofold_integral(cp, alpha, beta, beta);
  k_cofold_apply<<<blocks, 256, 0, P.stream>>>(cp, a, n0, n1, n2, alpha, beta, offset);
  check_launch();
}

// ---------------------------------------------------------------------------
// Eigenvalues: Matrix::fillTransformedMatrix, matrix.hh:223-318.
// ---------------------------------------------------------------------------
pf_eig_stats reduce_partials(Plan& P, int nb)
{
  std::vector<ClassifyPartial> h(nb);
  HIPCHECK(hipMemcpyAsync(h.data(), P.partial.p, nb * sizeof(ClassifyPartial),
                          hipMemcpyDeviceToHost, P.stream));
  HIPCHECK(hipStreamSynchronize(P.stream));
  pf_eig_stats s{};
  s.smallest = 1.7976931348623157e308;
  for (auto& c : h) {
    s.small += c.small;
    s.zero += c.zero;
    s.negative += c.negative;
    s.smallest = std::min(s.smallest, c.smallest);
  }
  return s;
}

// Whether the 2D R2C eigenvalues can come from the even part of the
// covariance: specialised kernels for the r2c along x of rows 0 .. N1/2 and
// for the real y pass (kf_eig_c2r_y) of N1/2 points.
bool even2d_ok(const Geo& g)
{
  if (g.dim != 2 || g.N[0] % 2 != 0 || g.N[1] % 2 != 0 || fast_budget() == 0)
    return false;
  FastCtx q{ nullptr, fast_budget() };
  q.query = true;
  const int N0 = g.N[0], N1 = g.N[1];
  return fast_r2c<double>(q, N0 / 2, nullptr, nullptr, N1 / 2 + 1, g.Hp, N0, nullptr, 1) &&
         fast_eig_c2r_y<float>(q, N1 / 2, nullptr, g.Hp, (int)g.H, N0 / 2, 0., 0., nullptr,
                               nullptr, nullptr, nullptr, nullptr, 1, nullptr);
}

// Classification partials per realization of kf_eig_c2r_y for this shape.
int even2d_blocks(const Geo& g)
{
  FastCtx q{ nullptr, fast_budget() };
  q.query = true;
  int nb = 0;
  fast_eig_c2r_y<float>(q, g.N[1] / 2, nullptr, g.Hp, (int)g.H, g.N[0] / 2, 0., 0., nullptr,
                        nullptr, nullptr, nullptr, nullptr, 1, &nb);
  return nb;
}

// R2C matrix, 2D, full real-part check, from the even part e of the
// embedded covariance (k_eval_cov_even2d): e on rows 0 .. N1/2, r2c of those
// rows along x, then one real pass along y that scales, classifies and
// writes the eigenvalues and the square-root table (kf_eig_c2r_y). The
// eigenvalues are the real part of the embedding's DFT, as in the full
// transform; the passes move about a third of its bytes.
bool compute_eigs_even2d(Plan& P, const double* custom_host, ClassifyPartial* dev_stats,
                         const CovParams& cp, double inv_next)
{
  const Geo& g = P.g;
  const pf_desc& d = P.d;
  const int64_t H = g.H, N0 = g.N[0], N1 = g.N[1], hr = N1 / 2 + 1;
  const int64_t count = H * N1;
  P.scratch.ensure((size_t)N0 * hr * sizeof(double) + (size_t)g.Hp * hr * sizeof(cplx<double>));
  double* real = P.scratch.as<double>();
  cplx<double>* spec = reinterpret_cast<cplx<double>*>(real + N0 * hr);
  if (custom_host) {
    // custom values arrive as the padded 2H x N1 real array: its even part
    std::vector<double> even(N0 * hr);
    for (int64_t j = 0; j < hr; j++) {
      const int64_t jm = j ? N1 - j : 0;
      for (int64_t i = 0; i < N0; i++) {
        const int64_t im = i ? N0 - i : 0;
        even[j * N0 + i] = 0.5 * (custom_host[j * 2 * H + i] + custom_host[jm * 2 * H + im]);
      }
    }
    HIPCHECK(hipMemcpyAsync(real, even.data(), even.size() * sizeof(double),
                            hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
  } else {
    Scope sc(&P, "eig_cov");
    with_cov(cp.covariance, [&](auto c) {
      constexpr int C = decltype(c)::value;
      with_period(cp.period, [&](auto q) {
        k_eval_cov_even2d<C, decltype(q)::value><<<row_blocks(hr, (int)N0), 256, 0, P.stream>>>(
          cp, (int)N0, (int)N1, real);
      });
    });
    check_launch();
  }
  const FastCtx fc{ P.stream, fast_budget() };
  {
    Scope sc(&P, "eig_r2c_x");
    if (!fast_r2c<double>(fc, (int)(N0 / 2), real, spec, hr, g.Hp, (int)N0,
                          P.table<double>((int)N0), 1))
      throw PfError(PF_EHIP, "even-part eigenvalue r2c launch rejected");
    check_launch();
  }
  P.eig.ensure(count * sizeof(double));
  const int64_t ne = g.eig_entries;
  int nb = even2d_blocks(g);
  P.partial.ensure((size_t)std::max(nb, 1024) * sizeof(ClassifyPartial));
  {
    Scope sc(&P, "eig_classify");
    const cplx<double>* tw = P.table<double>((int)N1);
    bool ok;
    if (d.precision == PF_FP32) {
      P.sq32.ensure(ne * sizeof(float));
      ok = fast_eig_c2r_y<float>(fc, (int)(N1 / 2), spec, g.Hp, (int)H, (int)(N0 / 2), inv_next,
                                 d.threshold, nullptr, P.eig.as<double>(), P.sq32.as<float>(),
                                 P.partial.as<ClassifyPartial>(), tw, 1, &nb);
    } else {
      P.sq64.ensure(ne * sizeof(double));
      ok = fast_eig_c2r_y<double>(fc, (int)(N1 / 2), spec, g.Hp, (int)H, (int)(N0 / 2), inv_next,
                                  d.threshold, nullptr, P.eig.as<double>(), P.sq64.as<double>(),
                                  P.partial.as<ClassifyPartial>(), tw, 1, &nb);
    }
    if (!ok)
      throw PfError(PF_EHIP, "even-part eigenvalue y-pass launch rejected");
    check_launch();
  }
  if (dev_stats)
    k_reduce_partials<<<1, 64, 0, P.stream>>>(P.partial.as<ClassifyPartial>(), nb, dev_stats);
  else {
    P.st = reduce_partials(P, nb);
    P.st.entries = count;
    P.scratch.release();
  }
  return true;
}

// R2C matrix, 2D, full real-part check (r2c_check_all): covariance on the
// padded grid, compile-time r2c along x and forward DFT along y (in place,
// row pitch Hp), then one fused pass for scaling, classification, the real
// part and the square-root table. Returns false (nothing changed but the
// scratch contents) when a line length has no specialised kernel. Shapes the
// even-part kernels cover take compute_eigs_even2d (except cofold, whose
// offset pass works on the full grid).
bool compute_eigs_r2c2d_fast(Plan& P, const double* custom_host, ClassifyPartial* dev_stats,
                             const CovParams& cp, double inv_next, int nbp)
{
  const Geo& g = P.g;
  const pf_desc& d = P.d;
  if (d.periodization != PF_PERIOD_COFOLD && even2d_ok(g))
    return compute_eigs_even2d(P, custom_host, dev_stats, cp, inv_next);
  const int64_t H = g.H, N0 = g.N[0], N1 = g.N[1], rows = N1;
  const int64_t count = H * rows;
  P.scratch.ensure((size_t)N0 * rows * sizeof(double) +
                   (size_t)g.Hp * rows * sizeof(cplx<double>));
  double* real = P.scratch.as<double>();
  cplx<double>* spec = reinterpret_cast<cplx<double>*>(real + N0 * rows);
  if (custom_host) {
    std::vector<double> compact(N0 * rows);
    for (int64_t r = 0; r < rows; r++)
      std::memcpy(&compact[r * N0], custom_host + r * 2 * H, N0 * sizeof(double));
    HIPCHECK(hipMemcpyAsync(real, compact.data(), compact.size() * sizeof(double),
                            hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
  } else {
    Scope sc(&P, "eig_cov");
    with_cov(cp.covariance, [&](auto c) {
      constexpr int C = decltype(c)::value;
      const dim3 grid = row_blocks((int64_t)g.N[1] * g.N[2], g.N[0]);
      with_period(cp.period, [&](auto q) {
        k_eval_cov_full<C, decltype(q)::value><<<grid, 256, 0, P.stream>>>(cp, g.N[0], g.N[1], g.N[2], real);
      });
    });
    check_launch();
    if (P.d.periodization == PF_PERIOD_COFOLD)
      apply_cofold(P, cp, real, g.N[0], g.N[1], g.N[2]);
  }
  const FastCtx fc{ P.stream, fast_budget() };
  {
    Scope sc(&P, "eig_r2c_x");
    if (!fast_r2c<double>(fc, (int)(N0 / 2), real, spec, rows, g.Hp, (int)N0,
                          P.table<double>((int)N0), 1))
      return false;
    check_launch();
  }
  {
    Scope sc(&P, "eig_c2c");
    if (!fast_mid<double>(fc, (int)N1, false, spec, spec, g.Hp, (int)(N0 / 2), 1, (int)N1,
                          (int)N1, (int)N1, (int)N1, P.table<double>((int)N1), 1))
      return false;
    check_launch();
  }
  P.eig.ensure(count * sizeof(double));
  const int64_t ne = g.eig_entries;
  const unsigned nb = std::min<unsigned>(nbp, grid1d(count));
  {
    Scope sc(&P, "eig_classify");
    if (d.precision == PF_FP32) {
      P.sq32.ensure(ne * sizeof(float));
      k_finish_r2c_eigs<float><<<nb, 256, 0, P.stream>>>(spec, g.Hp, (int)H, rows, inv_next,
                                                         d.threshold, P.eig.as<double>(),
                                                         P.sq32.as<float>(),
                                                         P.partial.as<ClassifyPartial>());
    } else {
      P.sq64.ensure(ne * sizeof(double));
      k_finish_r2c_eigs<double><<<nb, 256, 0, P.stream>>>(spec, g.Hp, (int)H, rows, inv_next,
                                                          d.threshold, P.eig.as<double>(),
                                                          P.sq64.as<double>(),
                                                          P.partial.as<ClassifyPartial>());
    }
    check_launch();
  }
  if (dev_stats)
    k_reduce_partials<<<1, 64, 0, P.stream>>>(P.partial.as<ClassifyPartial>(), nb, dev_stats);
  else {
    P.st = reduce_partials(P, (int)nb);
    P.st.entries = count;
    P.scratch.release();
  }
  return true;
}

// dev_stats != nullptr: capture-safe mode for batches (no host syncs, no
// frees); the classification lands in *dev_stats and the caller finalises.
void compute_eigs_dft_1d(Plan& P, const double* custom_host, const CovParams& cp);

void compute_eigs(Plan& P, const double* custom_host, ClassifyPartial* dev_stats = nullptr)
{
  const Geo& g = P.g;
  const pf_desc& d = P.d;
  CovParams cp = make_cov(d, g);
  const double inv_next = 1.0 / (double)g.Next;
  const int nbp = 1024;
  P.partial.ensure(nbp * sizeof(ClassifyPartial));
  if (g.dim == 1) {
    compute_eigs_dft_1d(P, custom_host, cp);
  } else if (d.matrix_backend == PF_MATRIX_DCT) {
    const int64_t total = (int64_t)g.P[0] * g.P[1] * g.P[2];
    P.eig.ensure(total * sizeof(double));
    double* a = P.eig.as<double>();
    if (custom_host)
      HIPCHECK(hipMemcpyAsync(a, custom_host, total * sizeof(double), hipMemcpyHostToDevice,
                              P.stream));
    else {
      Scope sc(&P, "eig_cov");
      with_cov(cp.covariance, [&](auto c) {
        constexpr int C = decltype(c)::value;
        const dim3 grid = row_blocks((int64_t)g.P[1] * g.P[2], g.P[0]);
        with_period(cp.period, [&](auto q) {
          k_eval_cov_dct<C, decltype(q)::value><<<grid, 256, 0, P.stream>>>(cp, g.P[0], g.P[1], g.P[2], a);
        });
      });
      check_launch();
      if (P.d.periodization == PF_PERIOD_COFOLD)
        apply_cofold(P, cp, a, g.P[0], g.P[1], g.P[2]);
    }
    int64_t stride = 1;
    for (int ax = 0; ax < g.dim; ax++) {
      const int Pn = g.P[ax], Nf = g.N[ax];
      const int64_t nlines = total / Pn;
      if (Pn > 1) {
        Sched s = make_sched(Nf);
        Cfg c = make_cfg<double>(Nf, nlines, Nf);
        Scope sc(&P, "eig_redft00");
        if (ax == 0)
          k_redft00_axis<16, true><<<nblocks(nlines, c.L), c.threads, c.lds, P.stream>>>(
            a, s, P.table<double>(Nf), Pn, stride, nlines, c.L);
        else
          k_redft00_axis<16, false><<<nblocks(nlines, c.L), c.threads, c.lds, P.stream>>>(
            a, s, P.table<double>(Nf), Pn, stride, nlines, c.L);
        check_launch();
      }
      stride *= Pn;
    }
    const unsigned nb = std::min<unsigned>(nbp, grid1d(total));
    {
      Scope sc(&P, "eig_classify");
      k_scale_classify_dct<<<nb, 256, 0, P.stream>>>(a, total, inv_next, d.threshold,
                                                    P.partial.as<ClassifyPartial>());
      check_launch();
    }
    if (dev_stats)
      k_reduce_partials<<<1, 64, 0, P.stream>>>(P.partial.as<ClassifyPartial>(), nb, dev_stats);
    else
      P.st = reduce_partials(P, nb);
    P.st.entries = total;
  } else if (g.dim == 2 && d.matrix_backend == PF_MATRIX_R2C && d.r2c_check_all &&
             fast_budget() > 0 &&
             compute_eigs_r2c2d_fast(P, custom_host, dev_stats, cp, inv_next, nbp)) {
    // eigenvalues, classification and square roots done by the specialised
    // 2D path (full real-part check)
    if (dev_stats)
      return;
    P.sync();
    P.eig_valid = true;
    P.eig_status = (P.st.negative > 0 && !d.approximate) ? PF_ENEG : PF_OK;
    return;
  } else {
    const int64_t H = g.H, N0 = g.N[0], rows = (int64_t)g.N[1] * g.N[2];
    const int64_t count = H * rows;
    P.eig.ensure(count * sizeof(double));
    P.scratch.ensure(std::max<size_t>(N0 * rows * sizeof(double), 0) +
                     count * sizeof(cplx<double>));
    double* real = P.scratch.as<double>();
    cplx<double>* spec = reinterpret_cast<cplx<double>*>(real + N0 * rows);
    if (custom_host) {
      // custom values arrive as the padded 2H x N1 x N2 real array (R2C) or
      // the N0 x N1 x N2 array (DFT)
      const int64_t pitch = d.matrix_backend == PF_MATRIX_DFT ? N0 : 2 * H;
      std::vector<double> compact(N0 * rows);
      for (int64_t r = 0; r < rows; r++)
        std::memcpy(&compact[r * N0], custom_host + r * pitch, N0 * sizeof(double));
      HIPCHECK(hipMemcpyAsync(real, compact.data(), compact.size() * sizeof(double),
                              hipMemcpyHostToDevice, P.stream));
      HIPCHECK(hipStreamSynchronize(P.stream));
    } else {
      Scope sc(&P, "eig_cov");
      with_cov(cp.covariance, [&](auto c) {
      constexpr int C = decltype(c)::value;
      const dim3 grid = row_blocks((int64_t)g.N[1] * g.N[2], g.N[0]);
      with_period(cp.period, [&](auto q) {
        k_eval_cov_full<C, decltype(q)::value><<<grid, 256, 0, P.stream>>>(cp, g.N[0], g.N[1], g.N[2], real);
      });
    });
      check_launch();
      if (P.d.periodization == PF_PERIOD_COFOLD)
        apply_cofold(P, cp, real, g.N[0], g.N[1], g.N[2]);
    }
    {
      Sched s = make_sched((int)N0);
      Cfg c = make_cfg<double>((int)N0, rows, (int)N0);
      Scope sc(&P, "eig_r2c_x");
      k_r2c_rows_full<16><<<nblocks(rows, c.L), c.threads, c.lds, P.stream>>>(
        real, spec, s, P.table<double>((int)N0), rows, c.L);
      check_launch();
    }
    int64_t stride = H;
    for (int ax = 1; ax < g.dim; ax++) {
      const int Nf = g.N[ax];
      const int64_t nlines = count / Nf;
      Sched s = make_sched(Nf);
      Cfg c = make_cfg<double>(Nf, nlines, Nf);
      Scope sc(&P, "eig_c2c");
      k_c2c_strided<double, 16, false><<<dim3(nblocks(nlines, c.L), 1), c.threads, c.lds,
                                         P.stream>>>(spec, s, P.table<double>(Nf), stride, nlines,
                                                     c.L, 0);
      check_launch();
      stride *= Nf;
    }
    k_scale_complex<<<grid1d(count), 256, 0, P.stream>>>(spec, count, inv_next);
    check_launch();
    const unsigned nb = std::min<unsigned>(nbp, grid1d(count));
    if (d.matrix_backend == PF_MATRIX_DFT)
      kdft_classify<<<nb, 256, 0, P.stream>>>(spec, count, g.H, g.N[0], d.threshold,
                                              P.partial.as<ClassifyPartial>());
    else
      k_classify_r2c<<<nb, 256, 0, P.stream>>>(spec, count, g.H, g.N[1], g.N[2],
                                               (g.dim == 3 && g.transposed) ? 1 : 0, d.threshold,
                                               P.partial.as<ClassifyPartial>(), d.r2c_check_all);
    check_launch();
    if (dev_stats)
      k_reduce_partials<<<1, 64, 0, P.stream>>>(P.partial.as<ClassifyPartial>(), nb, dev_stats);
    else
      P.st = reduce_partials(P, nb);
    P.st.entries = d.matrix_backend == PF_MATRIX_DFT ? g.Next : count;
    k_real_part<<<grid1d(count), 256, 0, P.stream>>>(spec, P.eig.as<double>(), count);
    check_launch();
    if (!dev_stats)
      P.scratch.release();
  }
  // sqrt tables for generation
  const int64_t ne = g.eig_entries;
  if (d.precision == PF_FP32) {
    P.sq32.ensure(ne * sizeof(float));
    k_sqrt_table<float><<<grid1d(ne), 256, 0, P.stream>>>(P.eig.as<double>(), P.sq32.as<float>(), ne);
  } else {
    P.sq64.ensure(ne * sizeof(double));
    k_sqrt_table<double>
      <<<grid1d(ne), 256, 0, P.stream>>>(P.eig.as<double>(), P.sq64.as<double>(), ne);
  }
  check_launch();
  if (dev_stats)
    return;
  P.sync();
  P.eig_valid = true;
  P.eig_status = (P.st.negative > 0 && !d.approximate) ? PF_ENEG : PF_OK;
}

// Twiddle tables of the R2C generation passes (last axis, middle axis, x),
// built with the eigenvalues so that a prepared plan generates without
// allocating device memory (pf_generate_noalloc).
void prepare_generation(Plan& P)
{
  if (P.d.field_backend != PF_FIELD_R2C)
    return;
  const Geo& g = P.g;
  const int ns[3] = { g.dim == 3 ? g.N[2] : g.N[1], g.dim == 3 ? g.N[1] : 0, g.N[0] };
  for (int n : ns) {
    if (n <= 0)
      continue;
    if (P.d.precision == PF_FP32)
      P.table<float>(n);
    else
      P.table<double>(n);
  }
}

void ensure_eigs(Plan& P)
{
  if (!P.eig_valid)
    compute_eigs(P, nullptr);
  if (P.eig_status != PF_OK)
    throw PfError(PF_ENEG, "negative eigenvalues in covariance matrix");
}

// ---------------------------------------------------------------------------
// Generation (R2C field backend)
// ---------------------------------------------------------------------------
LastAxis last_axis(const Geo& g)
{
  LastAxis a;
  a.H = g.H;
  a.Hp = g.Hp;
  a.N0h = g.N[0] / 2;
  a.Nmid = g.dim == 3 ? g.N[1] : 1;
  a.nlast = g.dim == 3 ? g.n[2] : g.n[1];
  if (g.dim == 3) {
    // reference storage index of (kx, k1 = km, k2 = e)
    a.sA = g.transposed ? 1 : g.N[1];
    a.sB = g.transposed ? g.N[2] : 1;
  } else {
    a.sA = 1;
    a.sB = 0;
  }
  return a;
}

// One chunk of realizations [b0, b0 + nb): pass G on stream sa, then the
// middle and x passes on stream sb (sb waits on an event after pass G).
template <typename T>
void generate_r2c_chunk(Plan& P, hipStream_t sa, hipStream_t sb, hipEvent_t ev,
                        const double* dnoise, uint64_t seed, uint64_t first, int b0, int nb,
                        int transform, double param, T* out, cplx<T>* W1, cplx<T>* W2)
{
  const Geo& g = P.g;
  GenGeom gg = gen_geom(P);
  LastAxis la = last_axis(g);
  const T* sq = sizeof(T) == 4 ? (const T*)P.sq32.p : (const T*)P.sq64.p;
  const int dim = g.dim;
  const int Nlast = dim == 3 ? g.N[2] : g.N[1];
  const int nlast = la.nlast;
  const int64_t nlinesG = dim == 3 ? (int64_t)g.H * g.N[1] : g.H;
  const int64_t s1 = (int64_t)g.Hp * (dim == 3 ? g.N[1] : 1) * nlast;  // B1 per realization
  const int64_t s2 = dim == 3 ? (int64_t)g.Hp * g.n[1] * g.n[2] : 0;   // B2 per realization
  cplx<T>* B1 = W1 + (int64_t)b0 * s1;
  cplx<T>* B2 = dim == 3 ? W2 + (int64_t)b0 * s2 : nullptr;
  const double* nz = dnoise ? dnoise + (int64_t)b0 * g.noise_per_field : nullptr;
  out += (int64_t)b0 * g.domain;
  constexpr int ME = maxe<T>();
  {
    FastCtx fc{ sa, fast_budget() };
    Scope sc(&P, "gen_last", sa);
    const cplx<T>* tw = P.table<T>(Nlast);
    if (!fast_gen_last<T>(fc, Nlast, gg, la, sq, nz, seed, first + b0, B1, tw, nb)) {
      Sched s = psched(P, Nlast);
      Cfg c = pcfg<T>(P, Nlast, nlinesG, Nlast);
      dim3 grid(nblocks(nlinesG, c.L), nb);
      if (nz)
        k_gen_last<T, ME, 0><<<grid, c.threads, c.lds, sa>>>(gg, sq, nz, seed, first + b0, B1, s,
                                                            tw, c.L);
      else
        k_gen_last<T, ME, 1><<<grid, c.threads, c.lds, sa>>>(gg, sq, nullptr, seed, first + b0,
                                                            B1, s, tw, c.L);
    }
    check_launch();
  }
  if (sb != sa) {
    HIPCHECK(hipEventRecord(ev, sa));
    HIPCHECK(hipStreamWaitEvent(sb, ev, 0));
  }
  // half-width tiles for the y and x passes of the generation: the plan's
  // tuning, else PF_TILE_DIV=2
  const char* td = getenv("PF_TILE_DIV");
  const int ldiv = P.tune.ldiv > 0 ? P.tune.ldiv : (td && td[0] == '2') ? 2 : 1;
  FastCtx fb{ sb, fast_budget(), ldiv };
  const cplx<T>* xin = B1;
  if (dim == 3) {
    Scope sc(&P, "gen_mid", sb);
    const cplx<T>* tw = P.table<T>(g.N[1]);
    if (!fast_mid<T>(fb, g.N[1], true, B1, B2, g.Hp, g.N[0] / 2, g.n[2], g.N[1], g.N[1], g.n[1],
                     g.n[1], tw, nb)) {
      const int64_t nl = (int64_t)g.H * g.n[2];
      Sched s = psched(P, g.N[1]);
      Cfg c = pcfg<T>(P, g.N[1], nl, g.N[1]);
      dim3 grid(nblocks(nl, c.L), nb);
      k_pass_mid<T, ME, true><<<grid, c.threads, c.lds, sb>>>(
        B1, B2, g.H, g.Hp, g.n[2], g.N[1], g.N[1], g.n[1], g.n[1], s, tw, c.L);
    }
    check_launch();
    xin = B2;
  }
  {
    const int M = g.N[0] / 2;
    const int64_t rows = (int64_t)g.n[1] * g.n[2];
    Scope sc(&P, "gen_c2r_x", sb);
    const cplx<T>* tw = P.table<T>(g.N[0]);
    if (!fast_c2r<T>(fb, M, xin, out, rows, g.Hp, g.n[0], tw, transform, (T)param, nb)) {
      Sched s = psched(P, M);
      Cfg c = pcfg<T>(P, M, rows, g.N[0]);
      dim3 grid(nblocks(rows, c.L), nb);
      k_c2r_x<T, T, ME><<<grid, c.threads, c.lds, sb>>>(xin, out, rows, g.Hp, g.n[0], s, tw, c.L,
                                                       transform, (T)param);
    }
    check_launch();
  }
}

// Generation of `batch` realizations. Large batches are cut into chunks and
// pipelined over two streams: the VALU-bound pass G of chunk k+1 runs while
// the HBM-bound middle and x passes of chunk k run (PF_OVERLAP=0 disables).
// With profiling on, each launch is bracketed by events on its own stream,
// so its time includes any sharing of the device with the other stream.
template <typename T>
void generate_r2c(Plan& P, const double* dnoise, uint64_t seed, uint64_t first, int batch,
                  int transform, double param, T* out)
{
  const Geo& g = P.g;
  const int dim = g.dim;
  const int nlast = dim == 3 ? g.n[2] : g.n[1];
  const int64_t nlinesG = dim == 3 ? (int64_t)g.H * g.N[1] : g.H;
  P.w1.ensure((size_t)g.Hp * (dim == 3 ? g.N[1] : 1) * nlast * sizeof(cplx<T>) * batch);
  if (dim == 3)
    P.w2.ensure((size_t)g.Hp * g.n[1] * g.n[2] * sizeof(cplx<T>) * batch);
  const char* ov = getenv("PF_OVERLAP");
  const bool want = P.tune.overlap >= 0 ? P.tune.overlap != 0 : !(ov && ov[0] == '0');
  const bool overlap = batch >= 4 && want && !P.capturing;
  if (!overlap) {
    generate_r2c_chunk<T>(P, P.stream, P.stream, nullptr, dnoise, seed, first, 0, batch,
                          transform, param, out, P.w1.as<cplx<T>>(), P.w2.as<cplx<T>>());
    return;
  }
  if (!P.stream2) {
    HIPCHECK(hipStreamCreateWithFlags(&P.stream2, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&P.ev_join, hipEventDisableTiming));
    for (auto& e : P.ev_chunk)
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  // stream2 starts after everything already queued on the plan stream
  HIPCHECK(hipEventRecord(P.ev_join, P.stream));
  HIPCHECK(hipStreamWaitEvent(P.stream2, P.ev_join, 0));
  const int nchunks = std::min(batch, P.tune.chunks > 0 ? std::min(P.tune.chunks, 8) : 8);
  const int per = (batch + nchunks - 1) / nchunks;
  int k = 0;
  for (int b0 = 0; b0 < batch; b0 += per, k++) {
    const int nb = std::min(per, batch - b0);
    generate_r2c_chunk<T>(P, P.stream, P.stream2, P.ev_chunk[k % P.ev_chunk.size()], dnoise,
                          seed, first, b0, nb, transform, param, out, P.w1.as<cplx<T>>(),
                          P.w2.as<cplx<T>>());
  }
  HIPCHECK(hipEventRecord(P.ev_join, P.stream2));
  HIPCHECK(hipStreamWaitEvent(P.stream, P.ev_join, 0));
}

// ---------------------------------------------------------------------------
// Generation, DCT/DST field backend (matrix.hh:342-372)
// ---------------------------------------------------------------------------
DctGeom dct_geom(const Plan& P)
{
  const Geo& g = P.g;
  DctGeom d;
  d.dim = g.dim;
  d.P0 = g.P[0];
  d.P1 = g.P[1];
  d.P2 = g.P[2];
  d.N0 = g.N[0];
  d.N1 = g.N[1];
  d.N2 = g.N[2];
  d.n0 = g.n[0];
  d.n1 = g.n[1];
  d.n2 = g.n[2];
  d.Ptot = (int64_t)g.P[0] * g.P[1] * g.P[2];
  // storage order of the reference arrays: transposeIfNeeded swaps the two
  // last dims (dctdstfieldbackend.hh:868-882)
  if (g.dim == 3) {
    d.sx = 1;
    d.s1 = g.transposed ? (int64_t)g.P[0] * g.P[2] : g.P[0];
    d.s2 = g.transposed ? g.P[0] : (int64_t)g.P[0] * g.P[1];
  } else {
    d.sx = g.transposed ? g.P[1] : 1;
    d.s1 = g.transposed ? 1 : g.P[0];
    d.s2 = 0;
  }
  d.noise_stride = g.noise_per_field;
  d.P0p = (g.P[0] + 31) / 32 * 32;
  return d;
}

// largest power of two <= cap dividing n
int pow2_divisor(int n, int cap)
{
  int L = 1;
  while (L * 2 <= cap && n % (L * 2) == 0)
    L *= 2;
  return L;
}

size_t work_cap()
{
  if (const char* e = getenv("PF_MAX_WORK_BYTES"))
    return (size_t)atoll(e);
  return (size_t)96 << 30;
}

template <typename T>
void generate_dctdst(Plan& P, const double* dnoise, uint64_t seed, uint64_t first, int batch,
                     int transform, double param, T* out)
{
  const Geo& g = P.g;
  DctGeom dg = dct_geom(P);
  const T* sq = sizeof(T) == 4 ? (const T*)P.sq32.p : (const T*)P.sq64.p;
  const int dim = g.dim;
  const int Pm = dim == 3 ? g.P[1] : 1;
  const int nlast = dim == 3 ? g.n[2] : g.n[1];
  const int Nlast = dim == 3 ? g.N[2] : g.N[1];
  const int ncomb = 1 << (dim - 1);
  const size_t gsz = (size_t)ncomb * Pm * dg.P0p * nlast * sizeof(T);
  const size_t hsz = dim == 3 ? (size_t)2 * g.n[1] * g.n[2] * dg.P0p * sizeof(T) : 0;
  int sub = (int)std::max<size_t>(1, work_cap() / (gsz + hsz));
  if (P.tune.sub > 0)
    sub = std::min(sub, P.tune.sub);
  sub = std::min(sub, batch);
  P.w1.ensure(gsz * sub);
  if (dim == 3)
    P.w2.ensure(hsz * sub);
  const size_t cs = sizeof(cplx<T>);
  const int maxL = sizeof(T) == 4 ? 16 : 8;
  for (int b0 = 0; b0 < batch; b0 += sub) {
    const int nb = std::min(sub, batch - b0);
    const double* nz = dnoise ? dnoise + (int64_t)b0 * g.noise_per_field : nullptr;
    const FastCtx fc{ P.stream, dct_fast_budget() };
    T* G = P.w1.as<T>();
    T* Hh = dim == 3 ? P.w2.as<T>() : nullptr;
    T* o = out + (int64_t)b0 * g.domain;
    const int64_t xrows = (int64_t)g.n[1] * g.n[2];
    // the specialised passes run when all of them apply (the G and H layouts
    // of both kinds agree, but one kind throughout keeps the choice simple)
    FastCtx q = fc;
    q.query = true;
    const bool fast =
      fast_dct_gen_last<T>(q, Nlast / 2, dg, sq, nz, seed, first + b0, G, nullptr, nb) &&
      (dim != 3 || fast_dct_mid<T>(q, g.N[1] / 2, dg, G, Hh, nullptr, nb)) &&
      fast_dct_x<T>(q, g.N[0] / 2, dg, dim == 3 ? Hh : G, o, xrows, nullptr, transform, (T)param,
                    nb);
    if (fast) {
      {
        Scope sc(&P, "dct_gen_last_fast");
        fast_dct_gen_last<T>(fc, Nlast / 2, dg, sq, nz, seed, first + b0, G, P.table<T>(Nlast),
                             nb);
        check_launch();
      }
      if (dim == 3) {
        Scope sc(&P, "dct_mid_fast");
        fast_dct_mid<T>(fc, g.N[1] / 2, dg, G, Hh, P.table<T>(g.N[1]), nb);
        check_launch();
      }
      {
        Scope sc(&P, "dct_x_fast");
        fast_dct_x<T>(fc, g.N[0] / 2, dg, dim == 3 ? Hh : G, o, xrows, P.table<T>(g.N[0]),
                      transform, (T)param, nb);
        check_launch();
      }
      continue;
    }
    // generic kernels (any size with prime factors <= 7): all three passes
    {
      const int M = Nlast / 2;
      int L = pow2_divisor(g.P[0] - 1, maxL);
      while (L > 1 && ((size_t)(L + 1) * M + L) * cs > dct_generic_budget())
        L /= 2;
      const size_t lds = ((size_t)(L + 1) * M + L) * cs;
      const int nt = std::min(1024, std::max(64, ((L * M / maxe<T>() + 63) / 64) * 64));
      dim3 grid((g.P[0] - 1) / L + 1, Pm, nb);
      Scope sc(&P, "dct_gen_last");
      if (nz)
        kd_gen_last<T, 0, 0, 0, 0><<<grid, nt, lds, P.stream>>>(
          dg, sq, nz, seed, first + b0, P.w1.as<T>(), P.table<T>(Nlast), psched(P, M), M, L);
      else
        kd_gen_last<T, 0, 0, 0, 1><<<grid, nt, lds, P.stream>>>(
          dg, sq, nullptr, seed, first + b0, P.w1.as<T>(), P.table<T>(Nlast), psched(P, M), M, L);
      check_launch();
    }
    const T* hin = P.w1.as<T>();
    if (dim == 3) {
      const int M = g.N[1] / 2;
      int L = pow2_divisor(g.P[0] - 1, maxL);
      while (L > 1 && ((size_t)(L + 1) * M + L) * cs > dct_generic_budget())
        L /= 2;
      const size_t lds = ((size_t)(L + 1) * M + L) * cs;
      const int nt = std::min(1024, std::max(64, ((L * M / maxe<T>() + 63) / 64) * 64));
      dim3 grid((g.P[0] - 1) / L + 1, g.n[2], nb);
      Scope sc(&P, "dct_mid");
      kd_mid<T, 0, 0, 0><<<grid, nt, lds, P.stream>>>(dg, P.w1.as<T>(), P.w2.as<T>(),
                                                      P.table<T>(g.N[1]), psched(P, M), M, L);
      check_launch();
      hin = P.w2.as<T>();
    }
    {
      const int M = g.N[0] / 2;
      const int64_t rows = (int64_t)g.n[1] * g.n[2];
      int L = maxL;
      while (L > 1 && ((size_t)(L + 1) * M + L) * cs > dct_generic_budget())
        L /= 2;
      while (L > 1 && L / 2 >= rows)
        L /= 2;
      const size_t lds = ((size_t)(L + 1) * M + L) * cs;
      const int nt = std::min(1024, std::max(64, ((L * M / maxe<T>() + 63) / 64) * 64));
      dim3 grid((unsigned)((rows + L - 1) / L), nb);
      Scope sc(&P, "dct_x");
      kd_x<T, 0, 0, 0><<<grid, nt, lds, P.stream>>>(dg, hin, out + (int64_t)b0 * g.domain, rows,
                                                    P.table<T>(g.N[0]), psched(P, M), M, L,
                                                    transform, (T)param);
      check_launch();
    }
  }
}

// ---------------------------------------------------------------------------
// DFT field backend (dftfieldbackend.hh) and the 1-D DFT matrix transforms
// ---------------------------------------------------------------------------

// Longest line a generic tile holds (one line plus its twiddle table).
int dft_max_line(size_t cs) { return (int)(lds_budget() / (2 * cs)); }

// Four-step split N = A * B of a 1-D transform: A the largest divisor of N
// that one tile holds, B = N / A (B = 1 when N fits one tile).
void dft_split_1d(int64_t N, size_t cs, int& A, int& B)
{
  const int cap = dft_max_line(cs);
  for (int64_t a = std::min<int64_t>(N, cap); a >= 1; a--)
    if (N % a == 0) {
      if (N / a > cap)
        break;
      A = (int)a;
      B = (int)(N / a);
      return;
    }
  throw PfError(PF_EINVAL, "1-D extended size " + std::to_string(N) +
                             " has no split into two transforms of at most " +
                             std::to_string(cap) + " points");
}

DftGeom dft_geom(const Plan& P, size_t cs)
{
  const Geo& g = P.g;
  DftGeom d{};
  d.dim = g.dim;
  d.transposed = g.transposed;
  d.N0 = g.N[0];
  d.N1 = g.N[1];
  d.N2 = g.N[2];
  d.H = g.H;
  d.P0 = g.P[0];
  d.P1 = g.P[1];
  d.P2 = g.P[2];
  d.dct = P.d.matrix_backend == PF_MATRIX_DCT ? 1 : 0;
  d.noise_stride = g.noise_per_field;
  if (g.dim == 1)
    dft_split_1d(g.N[0], cs, d.A, d.B);
  else {
    d.A = g.N[0];
    d.B = 1;
  }
  return d;
}

template <typename T>
void launch_cols(Plan& P, const char* name, bool inv, const cplx<T>* in, cplx<T>* out,
                 const ColPass& p, int nb)
{
  Sched s = psched(P, p.E);
  Cfg c = pcfg<T>(P, p.E, (int64_t)p.C * p.O, p.E);
  Scope sc(&P, name);
  const dim3 grid(nblocks((int64_t)p.C * p.O, c.L), nb);
  if (inv)
    kdft_cols<T, maxe<T>(), true><<<grid, c.threads, c.lds, P.stream>>>(in, out, p, s,
                                                                          P.table<T>(p.E), c.L);
  else
    kdft_cols<T, maxe<T>(), false><<<grid, c.threads, c.lds, P.stream>>>(in, out, p, s,
                                                                           P.table<T>(p.E), c.L);
  check_launch();
}

template <typename T, typename TO, bool INV, int OUT>
void launch_rows(Plan& P, const char* name, const cplx<T>* in, cplx<T>* outc, TO* outf,
                 const RowPass& r, int E, int nb, int64_t f_first, int64_t nfields, int transform,
                 double param)
{
  Sched s = make_sched(E);
  Cfg c = pcfg<T>(P, E, r.R, E);
  Scope sc(&P, name);
  kdft_rows<T, TO, maxe<T>(), INV, OUT><<<dim3(nblocks(r.R, c.L), nb), c.threads, c.lds,
                                          P.stream>>>(in, outc, outf, r, s, P.table<T>(E), c.L,
                                                      f_first, nfields, transform, (T)param);
  check_launch();
}

// Forward 1-D transform of `nb` complex arrays of N values in natural order
// (in -> X through the work array W): column pass over the [nA][nB] view
// with the twiddle e^{-2 pi i nB kA / N}, then the B-point rows kA written
// to k = kA + A kB.
template <typename T>
void dft_forward_1d(Plan& P, const DftGeom& dg, const cplx<T>* in, cplx<T>* W, cplx<T>* X, int nb)
{
  const int64_t N = dg.N0;
  ColPass p{};
  p.C = dg.B;
  p.E = dg.A;
  p.O = 1;
  p.keep = dg.A;
  p.in_es = dg.B;
  p.in_bs = N;
  p.out_es = dg.B;
  p.out_bs = N;
  p.twN = dg.B > 1 ? N : 0;
  launch_cols<T>(P, "dft_fwd_cols", false, in, W, p, nb);
  RowPass r{};
  r.R = dg.A;
  r.keep = dg.B;
  r.in_rs = dg.B;
  r.in_bs = N;
  r.out_rs = 1;
  r.out_es = dg.A;
  r.out_bs = N;
  r.limit = N;
  launch_rows<T, T, false, 0>(P, "dft_fwd_rows", W, X, (T*)nullptr, r, dg.B, nb, 0, 0,
                              PF_TRANSFORM_NONE, 0.);
}

// Inverse 1-D transform's row pass geometry: rows nB of A values, element nA
// of row nB to n = nB + B nA, the first n_keep outputs.
RowPass dft_inverse_rows_1d(const DftGeom& dg, int64_t n_keep, int64_t out_bs)
{
  RowPass r{};
  r.R = dg.B;
  r.keep = dg.A;
  r.in_rs = dg.A;
  r.in_bs = dg.N0;
  r.out_rs = 1;
  r.out_es = dg.B;
  r.out_bs = out_bs;
  r.limit = n_keep;
  return r;
}

ColPass dft_inverse_cols_1d(const DftGeom& dg)
{
  ColPass p{};
  p.C = dg.A;
  p.E = dg.B;
  p.O = 1;
  p.keep = dg.B;
  p.in_es = dg.A;
  p.in_bs = dg.N0;
  p.out_es = dg.A;
  p.out_bs = dg.N0;
  p.twN = dg.B > 1 ? dg.N0 : 0;
  return p;
}

// DFT matrix eigenvalues of a 1-D embedding (dftmatrixbackend.hh:209-236):
// covariance on the N-point grid, forward transform, real parts / N; the
// classification of fillTransformedMatrix (matrix.hh:266-291) over the full
// array. `custom_host`: N covariance values instead of the covariance family.
void compute_eigs_dft_1d(Plan& P, const double* custom_host, const CovParams& cp)
{
  const Geo& g = P.g;
  const pf_desc& d = P.d;
  const int64_t N = g.N[0];
  DftGeom dg = dft_geom(P, sizeof(cplx<double>));
  P.eig.ensure((size_t)g.H * sizeof(double));
  P.scratch.ensure((size_t)N * (sizeof(double) + 3 * sizeof(cplx<double>)));
  double* real = P.scratch.as<double>();
  cplx<double>* Z = reinterpret_cast<cplx<double>*>(real + N);
  cplx<double>* W = Z + N;
  cplx<double>* X = W + N;
  if (custom_host) {
    HIPCHECK(hipMemcpyAsync(real, custom_host, N * sizeof(double), hipMemcpyHostToDevice,
                            P.stream));
  } else {
    Scope sc(&P, "eig_cov");
    with_cov(cp.covariance, [&](auto c) {
      constexpr int C = decltype(c)::value;
      with_period(cp.period, [&](auto q) {
        k_eval_cov_1d<C, decltype(q)::value><<<grid1d(N), 256, 0, P.stream>>>(cp, N, real);
      });
    });
    check_launch();
    if (d.periodization == PF_PERIOD_COFOLD)
      apply_cofold(P, cp, real, g.N[0], 1, 1);
  }
  kdft_pad<double, double><<<grid1d(N), 256, 0, P.stream>>>(real, Z, N, N, 1);
  check_launch();
  dft_forward_1d<double>(P, dg, Z, W, X, 1);
  const unsigned nb = std::min<unsigned>(1024, grid1d(g.H));
  kdft_finish_eigs_1d<<<nb, 256, 0, P.stream>>>(X, N, g.H, 1.0 / (double)N, d.threshold,
                                                 P.eig.as<double>(),
                                                 P.partial.as<ClassifyPartial>());
  check_launch();
  P.st = reduce_partials(P, nb);
  P.st.entries = N;
  P.scratch.release();
}

// Matrix::generateField with the DFT field backend (matrix.hh:374-421,
// dftfieldbackend.hh:236-240, :280-318): draw j fills the spectrum with
// lambda (rand1 + i rand2), the backward transform gives field 2j as its real
// part and field 2j + 1 (the reference's spare field, handed out by the next
// generateField call) as its imaginary part. Output field i of this call is
// field first + i: draw (first + i) / 2, part (first + i) % 2. With host noise
// the draws are 0 .. ceil(batch / 2) - 1 of the noise array.
template <typename T>
void generate_dft(Plan& P, const double* dnoise, uint64_t seed, uint64_t first, int batch,
                  int transform, double param, T* out)
{
  const Geo& g = P.g;
  const size_t cs = sizeof(cplx<T>);
  DftGeom dg = dft_geom(P, cs);
  const T* sq = sizeof(T) == 4 ? (const T*)P.sq32.p : (const T*)P.sq64.p;
  const uint64_t d_begin = dnoise ? 0 : first / 2;
  const int64_t f_first = dnoise ? 0 : 2 * (int64_t)d_begin - (int64_t)first;
  const int ndraws = (int)((batch - f_first + 1) / 2);
  // per-draw work arrays
  int64_t w1 = 0, w2 = 0;
  if (g.dim == 3) {
    w1 = (int64_t)g.n[2] * g.N[1] * g.N[0];
    w2 = (int64_t)g.n[2] * g.n[1] * g.N[0];
  } else if (g.dim == 2) {
    w1 = (int64_t)g.n[1] * g.N[0];
  } else {
    w1 = g.N[0];
  }
  int sub = (int)std::max<int64_t>(1, (int64_t)(work_cap() / ((w1 + w2) * cs)));
  if (P.tune.sub > 0)
    sub = std::min(sub, P.tune.sub);
  sub = std::min(sub, ndraws);
  P.w1.ensure((size_t)w1 * cs * sub);
  if (w2)
    P.w2.ensure((size_t)w2 * cs * sub);
  cplx<T>* W1 = P.w1.as<cplx<T>>();
  cplx<T>* W2 = w2 ? P.w2.as<cplx<T>>() : nullptr;
  for (int j0 = 0; j0 < ndraws; j0 += sub) {
    const int nb = std::min(sub, ndraws - j0);
    const double* nz = dnoise ? dnoise + (int64_t)j0 * g.noise_per_field : nullptr;
    // pass 1: spectrum drawn in registers, backward along the last axis
    ColPass p{};
    if (g.dim == 3) {
      p.C = g.N[0];
      p.E = g.N[2];
      p.O = g.N[1];
      p.keep = g.n[2];
      p.out_es = (int64_t)g.N[1] * g.N[0];
      p.out_os = g.N[0];
    } else if (g.dim == 2) {
      p.C = g.N[0];
      p.E = g.N[1];
      p.O = 1;
      p.keep = g.n[1];
      p.out_es = g.N[0];
    } else {
      p = dft_inverse_cols_1d(dg);
    }
    p.out_bs = w1;
    {
      Sched s = psched(P, p.E);
      Cfg c = pcfg<T>(P, p.E, (int64_t)p.C * p.O, p.E);
      Scope sc(&P, "dft_gen");
      const dim3 grid(nblocks((int64_t)p.C * p.O, c.L), nb);
      if (nz)
        kdft_gen<T, maxe<T>(), 0><<<grid, c.threads, c.lds, P.stream>>>(
          dg, sq, nz, seed, d_begin + j0, W1, p, s, P.table<T>(p.E), c.L);
      else
        kdft_gen<T, maxe<T>(), 1><<<grid, c.threads, c.lds, P.stream>>>(
          dg, sq, nullptr, seed, d_begin + j0, W1, p, s, P.table<T>(p.E), c.L);
      check_launch();
    }
    const cplx<T>* rows_in = W1;
    RowPass r{};
    int E = g.N[0];
    if (g.dim == 3) {
      // pass 2: y lines of W1[z][ky][kx], first n1 outputs -> W2[z][y][kx]
      ColPass q{};
      q.C = g.N[0];
      q.E = g.N[1];
      q.O = g.n[2];
      q.keep = g.n[1];
      q.in_es = g.N[0];
      q.in_os = (int64_t)g.N[1] * g.N[0];
      q.in_bs = w1;
      q.out_es = g.N[0];
      q.out_os = (int64_t)g.n[1] * g.N[0];
      q.out_bs = w2;
      launch_cols<T>(P, "dft_mid", true, W1, W2, q, nb);
      rows_in = W2;
      r.R = (int64_t)g.n[1] * g.n[2];
      r.in_bs = w2;
    } else if (g.dim == 2) {
      r.R = g.n[1];
      r.in_bs = w1;
    }
    if (g.dim >= 2) {
      r.keep = g.n[0];
      r.in_rs = g.N[0];
      r.out_rs = g.n[0];
      r.out_es = 1;
      r.limit = g.domain;
    } else {
      r = dft_inverse_rows_1d(dg, g.n[0], g.domain);
      E = dg.A;
    }
    r.out_bs = g.domain;
    launch_rows<T, T, true, 1>(P, "dft_x", rows_in, nullptr, out, r, E, nb,
                               f_first + 2 * (int64_t)j0, batch, transform, param);
  }
}

// Products of the 1-D DFT field (matrix.hh:1443-1646 with
// dftfieldbackend.hh:166-196, :244-318): zero-padded forward transform,
// times f(lambda) / N, backward transform, the real part of the first n0
// values.
template <typename T>
void matvec_dft_1d(Plan& P, int op, const T* x, T* y, int batch)
{
  const Geo& g = P.g;
  if (op != PF_OP_MULTIPLY && op != PF_OP_ROOT && op != PF_OP_INVERSE_EMBED &&
      op != PF_OP_INVERSE_FLOOR)
    throw PfError(PF_EINVAL, "operator not available for 1-D fields");
  const int64_t N = g.N[0];
  DftGeom dg = dft_geom(P, sizeof(cplx<T>));
  P.w1.ensure((size_t)N * batch * sizeof(cplx<T>));
  P.w2.ensure((size_t)N * batch * sizeof(cplx<T>));
  cplx<T>* W1 = P.w1.as<cplx<T>>();
  cplx<T>* W2 = P.w2.as<cplx<T>>();
  kdft_pad<T, T><<<grid1d(N * batch), 256, 0, P.stream>>>(x, W2, g.n[0], N, batch);
  check_launch();
  // forward: W2 -> (W1) -> W2, natural frequency order
  dft_forward_1d<T>(P, dg, W2, W1, W2, batch);
  {
    Scope sc(&P, "dft_mult");
    kdft_mult_1d<T><<<grid1d(N * batch), 256, 0, P.stream>>>(W2, N, g.H, P.eig.as<double>(), op,
                                                             1.0 / (double)N, batch,
                                                             P.floor_value);
    check_launch();
  }
  launch_cols<T>(P, "dft_inv_cols", true, W2, W1, dft_inverse_cols_1d(dg), batch);
  launch_rows<T, T, true, 2>(P, "dft_inv_rows", W1, nullptr, y,
                             dft_inverse_rows_1d(dg, g.n[0], g.n[0]), dg.A, batch, 0, batch,
                             PF_TRANSFORM_NONE, 0.);
}

// ---------------------------------------------------------------------------
// Matvec (R2C field backend): multiplyExtended & co, matrix.hh:1443-1646
// ---------------------------------------------------------------------------
template <typename T>
void matvec_r2c(Plan& P, int op, const T* x, T* y, int batch)
{
  const Geo& g = P.g;
  GenGeom gg = gen_geom(P);
  LastAxis la = last_axis(g);
  FastCtx fc{ P.stream, fast_budget() };
  const int dim = g.dim;
  constexpr int ME = maxe<T>();
  const int M = g.N[0] / 2;
  const int64_t rows = (int64_t)g.n[1] * g.n[2];
  const size_t c1 = (size_t)rows * g.Hp * sizeof(cplx<T>);
  const size_t c2 = dim == 3 ? (size_t)g.Hp * g.N[1] * g.n[2] * sizeof(cplx<T>) : 0;
  P.w1.ensure(c1 * batch);
  if (dim == 3)
    P.w2.ensure(c2 * batch);
  {
    Scope sc(&P, "mv_r2c_x");
    const cplx<T>* tw = P.table<T>(g.N[0]);
    if (!fast_r2c<T>(fc, M, x, P.w1.as<cplx<T>>(), rows, g.Hp, g.n[0], tw, batch)) {
      Sched s = psched(P, M);
      Cfg c = pcfg<T>(P, M, rows, g.N[0]);
      k_r2c_x<T, T, ME><<<dim3(nblocks(rows, c.L), batch), c.threads, c.lds, P.stream>>>(
        x, P.w1.as<cplx<T>>(), rows, g.Hp, g.n[0], s, tw, c.L);
    }
    check_launch();
  }
  cplx<T>* C = P.w1.as<cplx<T>>();
  if (dim == 3) {
    Scope sc(&P, "mv_fwd_mid");
    const cplx<T>* tw = P.table<T>(g.N[1]);
    if (!fast_mid<T>(fc, g.N[1], false, P.w1.as<cplx<T>>(), P.w2.as<cplx<T>>(), g.Hp, g.N[0] / 2,
                     g.n[2], g.n[1], g.n[1], g.N[1], g.N[1], tw, batch)) {
      const int64_t nl = (int64_t)g.H * g.n[2];
      Sched s = psched(P, g.N[1]);
      Cfg c = pcfg<T>(P, g.N[1], nl, g.N[1]);
      k_pass_mid<T, ME, false><<<dim3(nblocks(nl, c.L), batch), c.threads, c.lds, P.stream>>>(
        P.w1.as<cplx<T>>(), P.w2.as<cplx<T>>(), g.H, g.Hp, g.n[2], g.n[1], g.n[1], g.N[1], g.N[1],
        s, tw, c.L);
    }
    check_launch();
    C = P.w2.as<cplx<T>>();
  }
  {
    const int Nlast = dim == 3 ? g.N[2] : g.N[1];
    const int64_t nlines = dim == 3 ? (int64_t)g.H * g.N[1] : g.H;
    Scope sc(&P, "mv_fused_mult");
    const cplx<T>* tw = P.table<T>(Nlast);
    const double* tab = op == PF_OP_INVERSE_FLOOR ? P.floor.as<double>() : P.spec.as<double>();
    if (!fast_fused<T>(fc, Nlast, C, gg, la, P.eig.as<double>(), op, 1.0 / (double)g.Next, tw,
                       batch, tab, P.spec_nbins)) {
      Sched s = psched(P, Nlast);
      Cfg c = pcfg<T>(P, Nlast, nlines, Nlast);
      k_fused_mult<T, ME><<<dim3(nblocks(nlines, c.L), batch), c.threads, c.lds, P.stream>>>(
        C, gg, P.eig.as<double>(), op, 1.0 / (double)g.Next, s, tw, c.L, tab,
        P.spec_nbins);
    }
    check_launch();
  }
  if (dim == 3) {
    Scope sc(&P, "mv_inv_mid");
    const cplx<T>* tw = P.table<T>(g.N[1]);
    if (!fast_mid<T>(fc, g.N[1], true, P.w2.as<cplx<T>>(), P.w1.as<cplx<T>>(), g.Hp, g.N[0] / 2,
                     g.n[2], g.N[1], g.N[1], g.n[1], g.n[1], tw, batch)) {
      const int64_t nl = (int64_t)g.H * g.n[2];
      Sched s = psched(P, g.N[1]);
      Cfg c = pcfg<T>(P, g.N[1], nl, g.N[1]);
      k_pass_mid<T, ME, true><<<dim3(nblocks(nl, c.L), batch), c.threads, c.lds, P.stream>>>(
        P.w2.as<cplx<T>>(), P.w1.as<cplx<T>>(), g.H, g.Hp, g.n[2], g.N[1], g.N[1], g.n[1], g.n[1],
        s, tw, c.L);
    }
    check_launch();
  }
  {
    Scope sc(&P, "mv_c2r_x");
    const cplx<T>* tw = P.table<T>(g.N[0]);
    if (!fast_c2r<T>(fc, M, P.w1.as<cplx<T>>(), y, rows, g.Hp, g.n[0], tw, PF_TRANSFORM_NONE, T(0),
                     batch)) {
      Sched s = psched(P, M);
      Cfg c = pcfg<T>(P, M, rows, g.N[0]);
      k_c2r_x<T, T, ME><<<dim3(nblocks(rows, c.L), batch), c.threads, c.lds, P.stream>>>(
        P.w1.as<cplx<T>>(), y, rows, g.Hp, g.n[0], s, tw, c.L, PF_TRANSFORM_NONE, T(0));
    }
    check_launch();
  }
}

// The DCT/DST field backend takes the same products: its route in the
// reference (matrix.hh:1456-1470, per parity type setComponent, r2r forward /
// N_ext, times f(lambda), r2r backward, summed into the output) adds up to
// R f(C_ext) E x / N_ext with the DCT eigenvalues, which matvec_r2c applies
// through the half-spectrum passes (the fused pass folds kx, ky, kz onto the
// DCT-half table). The reference's accumulation into its output
// (dctdstfieldbackend.hh:923) is the caller's: the C++ Matrix adds the input
// back for operator* and multiplyRoot.

// ---------------------------------------------------------------------------
// Conjugate gradients (FP64 vectors and dot products on the device)
// ---------------------------------------------------------------------------
struct Vec
{
  DevBuf b;
  double* p(int64_t n)
  {
    b.ensure(n * sizeof(double));
    return b.as<double>();
  }
};

double dot(Plan& P, const double* a, const double* b, int64_t n)
{
  const unsigned nb = std::min<unsigned>(1024, grid1d(n));
  P.partial.ensure(1024 * sizeof(ClassifyPartial));
  k_dot_partial<<<nb, 256, 0, P.stream>>>(a, b, n, P.partial.as<double>());
  check_launch();
  std::vector<double> h(nb);
  HIPCHECK(hipMemcpyAsync(h.data(), P.partial.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                          P.stream));
  HIPCHECK(hipStreamSynchronize(P.stream));
  double s = 0.;
  for (double v : h)
    s += v;
  return s;
}

double energy(Plan& P, const double* it, const double* ait, const double* b, int64_t n)
{
  const unsigned nb = std::min<unsigned>(1024, grid1d(n));
  P.partial.ensure(1024 * sizeof(ClassifyPartial));
  k_energy_partial<<<nb, 256, 0, P.stream>>>(it, ait, b, n, P.partial.as<double>());
  check_launch();
  std::vector<double> h(nb);
  HIPCHECK(hipMemcpyAsync(h.data(), P.partial.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                          P.stream));
  HIPCHECK(hipStreamSynchronize(P.stream));
  double s = 0.;
  for (double v : h)
    s += v;
  return s;
}

void matvec64(Plan& P, int op, const double* x, double* y)
{
  if (op == PF_OP_SPECTRAL && P.spec_nbins <= 0)
    throw PfError(PF_EINVAL, "no learned multipliers loaded");
  if (P.g.dim == 1)
    matvec_dft_1d<double>(P, op, x, y, 1);
  else
    matvec_r2c<double>(P, op, x, y, 1);
}

// Matrix::innerCG, matrix.hh:1315-1432, on device vectors.
int inner_cg(Plan& P, double* iter, const double* solution, int64_t n, unsigned cgIterations)
{
  Vec vMS, vMI, vR, vPR, vD, vMD;
  double* matrixTimesSolution = vMS.p(n);
  double* matrixTimesIter = vMI.p(n);
  double* residual = vR.p(n);
  double* precResidual = vPR.p(n);
  double* direction = vD.p(n);
  double* matrixTimesDirection = vMD.p(n);
  const unsigned gb = grid1d(n);
  matvec64(P, PF_OP_MULTIPLY, solution, matrixTimesSolution); // unused product (matrix.hh:1329)
  matvec64(P, PF_OP_MULTIPLY, iter, matrixTimesIter);
  k_sub<<<gb, 256, 0, P.stream>>>(solution, matrixTimesIter, residual, n);
  matvec64(P, PF_OP_INVERSE_EMBED, residual, precResidual);
  HIPCHECK(hipMemcpyAsync(direction, precResidual, n * sizeof(double), hipMemcpyDeviceToDevice,
                          P.stream));
  bool converged = false;
  double scalarProd = dot(P, precResidual, residual, n);
  double scalarProd2 = dot(P, residual, residual, n);
  if (std::sqrt(std::abs(scalarProd2)) < 1e-6)
    converged = true;
  const double firstValue = energy(P, iter, matrixTimesIter, solution, n);
  unsigned count = 0;
  while (!converged && count < cgIterations) {
    matvec64(P, PF_OP_MULTIPLY, direction, matrixTimesDirection);
    const double alphaDenominator = dot(P, direction, matrixTimesDirection, n);
    const double alpha = scalarProd / alphaDenominator;
    const double oldValue = energy(P, iter, matrixTimesIter, solution, n);
    k_axpy<<<gb, 256, 0, P.stream>>>(alpha, direction, iter, n);
    k_axpy<<<gb, 256, 0, P.stream>>>(alpha, matrixTimesDirection, matrixTimesIter, n);
    const double value = energy(P, iter, matrixTimesIter, solution, n);
    k_sub<<<gb, 256, 0, P.stream>>>(solution, matrixTimesIter, residual, n);
    matvec64(P, PF_OP_INVERSE_EMBED, residual, precResidual);
    double beta = 1. / scalarProd;
    scalarProd = dot(P, precResidual, residual, n);
    beta *= scalarProd;
    k_xpby<<<gb, 256, 0, P.stream>>>(precResidual, beta, direction, n);
    check_launch();
    if (value != firstValue) {
      if (std::abs(value - oldValue) / std::abs(value - firstValue) < 1e-16)
        converged = true;
    }
    count++;
  }
  return (int)count;
}

// (P)CG with a relative-residual stop, device resident: alpha, beta and the
// dot products stay in device memory (CgScalars), the x/r update is one fused
// pass that also sums r . r, and the host reads the scalars every
// kCheck iterations instead of once per dot product. Once the recursive
// residual meets the tolerance the device stops updating (done); the host
// then recomputes the FP64 true residual b - A x and either stops or goes on
// from it (residual replacement: the direction is kept). Every kRefresh
// iterations the residual is replaced by the true residual as well.
// Preconditioner: none, the embedding inverse, or the learned radial
// spectral multiplier (PF_OP_SPECTRAL).
int pcg(Plan& P, const double* b, double* x, int64_t n, int mode, int maxit, double rel_tol,
        double* relres)
{
  constexpr int kCheck = 10, kRefresh = 50;
  Vec vR, vZ, vP, vAp, vAx;
  double* r = vR.p(n);
  double* z = mode == PF_CG_PLAIN ? r : vZ.p(n);
  double* p = vP.p(n);
  double* Ap = vAp.p(n);
  double* Ax = vAx.p(n);
  const unsigned gb = std::min<unsigned>(1024, grid1d(n));
  DevBuf dsc, dpart;
  dsc.ensure(sizeof(CgScalars));
  dpart.ensure(3 * 1024 * sizeof(double));
  CgScalars* sc = dsc.as<CgScalars>();
  // block partials of p . Ap, r . r and r . z
  double* pap_part = dpart.as<double>();
  double* rr_part = pap_part + 1024;
  double* rz_part = rr_part + 1024;
  HIPCHECK(hipMemsetAsync(x, 0, n * sizeof(double), P.stream));
  HIPCHECK(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, P.stream));
  const double bb = dot(P, b, b, n);
  if (bb == 0.) {
    *relres = 0.;
    return 0;
  }
  // embedding-inverse preconditioner: 1 / max(lambda, nugget / N_ext), so
  // that eigenvalues the clamp set to zero do not make it infinite
  const bool floored = mode == PF_CG_EMBED_PCG && P.d.nugget > 0.;
  if (floored) {
    P.floor_value = P.d.nugget / (double)P.g.Next;
    P.floor.ensure(sizeof(double));
    HIPCHECK(hipMemcpyAsync(P.floor.p, &P.floor_value, sizeof(double), hipMemcpyHostToDevice,
                            P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
  }
  auto prec = [&](const double* rr, double* zz) {
    if (mode == PF_CG_EMBED_PCG)
      matvec64(P, floored ? PF_OP_INVERSE_FLOOR : PF_OP_INVERSE_EMBED, rr, zz);
    else if (mode == PF_CG_LEARNED_PCG)
      matvec64(P, PF_OP_SPECTRAL, rr, zz);
  };
  auto put = [&](const CgScalars& h) {
    HIPCHECK(hipMemcpyAsync(sc, &h, sizeof h, hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
  };
  auto true_residual = [&]() {
    matvec64(P, PF_OP_MULTIPLY, x, Ax);
    k_sub<<<grid1d(n), 256, 0, P.stream>>>(b, Ax, r, n);
    check_launch();
    return dot(P, r, r, n);
  };
  // iteration i (the device's count) reads r . z from slot i % 2 and its p
  // update writes the next value to the other slot (CgScalars::rzs)
  int par = 0;
  {
    // start: z = M^{-1} r, p = z, rz = r . z
    if (mode != PF_CG_PLAIN)
      prec(r, z);
    HIPCHECK(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, P.stream));
    CgScalars h{};
    h.rr = bb;
    h.rzs[par] = mode == PF_CG_PLAIN ? bb : dot(P, r, z, n);
    h.bb = bb;
    h.tol2 = rel_tol * rel_tol * bb;
    h.done = h.rr <= h.tol2 ? 1 : 0;
    put(h);
  }
  int it = 0, since_refresh = 0;
  double rr = 1.;
  while (true) {
    // the plan stream is non-blocking: read the scalars in its order
    CgScalars h{};
    HIPCHECK(hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
    it = h.it;
    // the slot of the next iteration; an iteration that set done did not
    // write it, and its own r . z is in the other slot
    par = it & 1;
    if (!std::isfinite(h.rr) || !std::isfinite(h.rzs[h.done ? par ^ 1 : par])) {
      // the recurrence broke down (an infinite or undefined preconditioner)
      rr = std::numeric_limits<double>::quiet_NaN();
      break;
    }
    if (h.done || it >= maxit || since_refresh >= kRefresh) {
      const double rr2 = true_residual(); // r = b - A x
      rr = std::sqrt(rr2 / bb);
      if (rr <= rel_tol || it >= maxit)
        break;
      h.rr = rr2;
      // z and r . z of the replaced residual
      if (mode != PF_CG_PLAIN)
        prec(r, z);
      const double rzn = mode == PF_CG_PLAIN ? rr2 : dot(P, r, z, n);
      if (h.done) {
        // the iteration that set `done` stopped before its direction update:
        // finish it from the true residual, p = z + (r . z / rz) p
        k_xpby<<<gb, 256, 0, P.stream>>>(z, rzn / h.rzs[par ^ 1], p, n);
        check_launch();
        h.done = 0;
      }
      // the next step size uses the replaced residual's r . z
      h.rzs[par] = rzn;
      put(h);
      since_refresh = 0;
      continue;
    }
    const int steps = std::min(kCheck, maxit - it);
    for (int k = 0; k < steps; k++) {
      // once an iteration sets done, the rest of this round of kernels skips
      // (all test sc->done) and the next round starts from the device's count
      matvec64(P, PF_OP_MULTIPLY, p, Ap);
      k_cg_dot<<<gb, 256, 0, P.stream>>>(p, Ap, n, pap_part, sc);
      k_cg_update_xr<<<gb, 256, 0, P.stream>>>(p, Ap, x, r, n, pap_part, (int)gb, rr_part, sc,
                                               par);
      if (mode != PF_CG_PLAIN) {
        prec(r, z);
        k_cg_dot<<<gb, 256, 0, P.stream>>>(r, z, n, rz_part, sc);
      }
      k_cg_update_p<<<gb, 256, 0, P.stream>>>(z, p, n, rr_part,
                                              mode == PF_CG_PLAIN ? nullptr : rz_part, (int)gb,
                                              sc, par);
      check_launch();
      par ^= 1;
    }
    since_refresh += steps;
  }
  *relres = rr;
  return it;
}

// ---------------------------------------------------------------------------
// MLP of the learned spectral preconditioner
// ---------------------------------------------------------------------------
struct Mlp
{
  std::vector<int> dims;
  std::vector<int64_t> woff, boff;
  int64_t nparams = 0;
  DevBuf params, grads, m, v, x, y, d0, d1, partial;
  // weight gradients: block partials over kGradSplit interleaved sample sets
  static constexpr int kGradSplit = 32;
  DevBuf gpart;
  std::vector<DevBuf> Z, A;
  int t = 0;
  hipStream_t stream = nullptr;
  int cap = 0;
  ~Mlp()
  {
    if (stream)
      (void)hipStreamDestroy(stream);
  }
  int nl() const { return (int)dims.size() - 1; }
  void ensure_batch(int B)
  {
    if (B <= cap)
      return;
    Z.resize(nl());
    A.resize(nl());
    int wmax = 0;
    for (int l = 0; l < nl(); l++) {
      Z[l].ensure((size_t)B * dims[l + 1] * sizeof(float));
      A[l].ensure((size_t)B * dims[l + 1] * sizeof(float));
    }
    for (int d : dims)
      wmax = std::max(wmax, d);
    x.ensure((size_t)B * dims[0] * sizeof(float));
    y.ensure((size_t)B * dims.back() * sizeof(float));
    d0.ensure((size_t)B * wmax * sizeof(float));
    d1.ensure((size_t)B * wmax * sizeof(float));
    partial.ensure(1024 * sizeof(double));
    int64_t gmax = 0;
    for (int l = 0; l < nl(); l++)
      gmax = std::max<int64_t>(gmax, (int64_t)dims[l] * dims[l + 1] + dims[l + 1]);
    gpart.ensure((size_t)kGradSplit * gmax * sizeof(float));
    cap = B;
  }
  void forward(int B)
  {
    const float* in = x.as<float>();
    for (int l = 0; l < nl(); l++) {
      const int a = dims[l], b = dims[l + 1];
      const int64_t tot = (int64_t)B * b;
      k_mlp_linear_fwd<<<(unsigned)((tot + 255) / 256), 256, 0, stream>>>(
        in, params.as<float>() + woff[l], params.as<float>() + boff[l], Z[l].as<float>(),
        A[l].as<float>(), B, a, b, l + 1 < nl() ? 1 : 0);
      in = A[l].as<float>();
    }
    check_launch();
  }
  double backward(int B)
  {
    const int L = nl();
    const int64_t nout = (int64_t)B * dims.back();
    const unsigned nb = std::min<unsigned>(1024, grid1d(nout));
    float* D = d0.as<float>();
    float* Dn = d1.as<float>();
    k_mlp_mse<<<nb, 256, 0, stream>>>(A[L - 1].as<float>(), y.as<float>(), D, nout,
                                      partial.as<double>());
    for (int l = L - 1; l >= 0; l--) {
      const int a = dims[l], b = dims[l + 1];
      if (l + 1 < L)
        k_mlp_gelu_back<<<grid1d((int64_t)B * b), 256, 0, stream>>>(D, Z[l].as<float>(),
                                                                    (int64_t)B * b);
      const float* in = l == 0 ? x.as<float>() : A[l - 1].as<float>();
      const int nout = a * b + b, nsplit = std::max(1, std::min(kGradSplit, B / 16));
      k_mlp_grad_part<<<dim3((unsigned)nsplit, (unsigned)((nout + 255) / 256)), 256, 0, stream>>>(
        D, in, gpart.as<float>(), B, a, b);
      k_mlp_grad_sum<<<(unsigned)((nout + 255) / 256), 256, 0, stream>>>(
        gpart.as<float>(), nsplit, grads.as<float>() + woff[l], grads.as<float>() + boff[l], a, b);
      if (l > 0) {
        k_mlp_back_in<<<(unsigned)(((int64_t)B * a + 255) / 256), 256, 0, stream>>>(
          D, params.as<float>() + woff[l], Dn, B, a, b);
        std::swap(D, Dn);
      }
    }
    check_launch();
    std::vector<double> h(nb);
    HIPCHECK(hipMemcpyAsync(h.data(), partial.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                            stream));
    HIPCHECK(hipStreamSynchronize(stream));
    double loss = 0.;
    for (double q : h)
      loss += q;
    return loss;
  }
};


// ---------------------------------------------------------------------------
// Uneven batches: one stream, optional hipGraph capture of the whole batch
// ---------------------------------------------------------------------------
// A field the grouped path can take: 2D, R2C matrix with the full real-part
// check, R2C field, FP32 realizations, and a specialised kernel for every
// pass of its shape.
bool groupable(const Plan& P)
{
  const pf_desc& d = P.d;
  const Geo& g = P.g;
  if (g.dim != 2 || d.matrix_backend != PF_MATRIX_R2C || !d.r2c_check_all ||
      d.field_backend != PF_FIELD_R2C || d.precision != PF_FP32 || fast_budget() == 0)
    return false;
  FastCtx q{ nullptr, fast_budget() };
  q.query = true;
  const int N0 = g.N[0], N1 = g.N[1];
  const GenGeom gg = gen_geom(P);
  const LastAxis la = last_axis(g);
  return (even2d_ok(g) ||
          (fast_r2c<double>(q, N0 / 2, nullptr, nullptr, N1, g.Hp, N0, nullptr, 1) &&
           fast_mid<double>(q, N1, false, nullptr, nullptr, g.Hp, N0 / 2, 1, N1, N1, N1, N1,
                            nullptr, 1))) &&
         fast_gen_last<float>(q, N1, gg, la, nullptr, nullptr, 0, 0, nullptr, nullptr, 1) &&
         fast_c2r<float>(q, N0 / 2, nullptr, nullptr, g.n[1], g.Hp, g.n[0], nullptr, 0, 0.f, 1);
}

// Fields of one shape that run together: every kernel of their eigenvalues
// and realizations is one launch over the group (field f = launch index f),
// with per-field covariance parameters, thresholds, Philox realizations,
// root tables and output offsets. 2D R2C fields (full real-part check, FP32
// realizations) with specialised kernels for every pass are grouped.
struct FieldGroup
{
  std::vector<size_t> fields; // batch indices
  DevBuf cov, fidx, ooff, thr; // per field: CovParams, batch index, output offset, threshold
  DevBuf scratch, sq, partial, w1;
};

// A unit of work of the batch: one group, or one field through its own plan.
struct BatchUnit
{
  int group; // index into Batch::groups, or -1
  size_t field;
  double cost;
};

struct Batch
{
  std::vector<std::unique_ptr<Plan>> plans;
  std::vector<int64_t> offsets;
  std::vector<double> cost;           // algorithmic bytes per field (bin packing weight)
  std::vector<std::unique_ptr<FieldGroup>> groups;
  std::vector<BatchUnit> units;
  hipStream_t stream = nullptr;       // origin stream
  std::vector<hipStream_t> branch;    // extra streams of the forked graph branches
  std::vector<hipEvent_t> ev;         // fork/join events (one per branch)
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> fev;        // per-field timing of the eager pass (2 per field)
  bool field_timing = false, timed = false;
  int branches = 1;
  DevBuf stats, stage;
  bool prepared = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t key_seed = 0, key_first = 0;
  int key_transform = -1, key_branches = 0;
  double key_param = 0.;
  void* key_out = nullptr;
  ~Batch()
  {
    if (exec)
      (void)hipGraphExecDestroy(exec);
    if (graph)
      (void)hipGraphDestroy(graph);
    plans.clear();
    for (auto s : branch)
      (void)hipStreamDestroy(s);
    for (auto e : ev)
      (void)hipEventDestroy(e);
    for (auto e : fev)
      (void)hipEventDestroy(e);
    if (ev_fork)
      (void)hipEventDestroy(ev_fork);
    if (stream)
      (void)hipStreamDestroy(stream);
  }
  void ensure_branches(int k)
  {
    while ((int)branch.size() < k - 1) {
      hipStream_t s;
      HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      branch.push_back(s);
      hipEvent_t e;
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
    }
    if (!ev_fork)
      HIPCHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  }
  void enqueue_field(size_t i, uint64_t seed, uint64_t first, int transform, double param,
                     void* out)
  {
    Plan& P = *plans[i];
    compute_eigs(P, nullptr, stats.as<ClassifyPartial>() + i);
    if (P.d.precision == PF_FP32) {
      float* o = (float*)out + offsets[i];
      if (P.d.field_backend == PF_FIELD_DCTDST)
        generate_dctdst<float>(P, nullptr, seed, first + i, 1, transform, param, o);
      else
        generate_r2c<float>(P, nullptr, seed, first + i, 1, transform, param, o);
    } else {
      double* o = (double*)out + offsets[i];
      if (P.d.field_backend == PF_FIELD_DCTDST)
        generate_dctdst<double>(P, nullptr, seed, first + i, 1, transform, param, o);
      else
        generate_r2c<double>(P, nullptr, seed, first + i, 1, transform, param, o);
    }
  }
  // One group on stream s: FP64 eigenvalues of all its fields (covariance,
  // r2c along x, y pass, scale/classify/root, per-field reduction into
  // stats), then one FP32 realization of each (pass G along y, c2r along x
  // into the field's place in the output). The eigenvalues take the path of
  // a single plan of the shape: from the even part of the covariance where
  // its kernels exist (compute_eigs_even2d), else from the full grid.
  void enqueue_group(FieldGroup& G, hipStream_t s, uint64_t seed, uint64_t first, int transform,
                     double param, void* out)
  {
    Plan& P = *plans[G.fields[0]]; // shape: geometry and twiddle tables
    P.stream = s;
    const Geo& g = P.g;
    const int nf = (int)G.fields.size();
    const bool even = even2d_ok(g);
    const int64_t H = g.H, N0 = g.N[0], N1 = g.N[1], rows = N1, count = H * rows;
    const int64_t hr = even ? N1 / 2 + 1 : rows; // covariance rows evaluated
    const double inv_next = 1.0 / (double)g.Next;
    G.scratch.ensure((size_t)nf * (N0 * hr * sizeof(double) + g.Hp * hr * sizeof(cplx<double>)));
    double* real = G.scratch.as<double>();
    cplx<double>* spec = reinterpret_cast<cplx<double>*>(real + nf * N0 * hr);
    with_cov(P.d.covariance, [&](auto c) {
      // the fields of a group share one periodization (part of the group key)
      constexpr int C = decltype(c)::value;
      const dim3 grid(row_blocks(hr, (int)N0), nf);
      with_period(P.d.periodization, [&](auto q) {
        // groups are 2D (groupable)
        if (even)
          k_eval_cov_even2d_group<C, decltype(q)::value><<<grid, 256, 0, s>>>(
            G.cov.as<CovParams>(), (int)N0, (int)N1, real, N0 * hr);
        else
          k_eval_cov_full_group<C, decltype(q)::value, 2><<<grid, 256, 0, s>>>(
            G.cov.as<CovParams>(), (int)N0, (int)N1, 1, real, N0 * rows);
      });
    });
    check_launch();
    const FastCtx fc{ s, fast_budget() };
    G.sq.ensure((size_t)nf * count * sizeof(float)); // the FP64 eigenvalues are not kept
    if (!fast_r2c<double>(fc, (int)(N0 / 2), real, spec, hr, g.Hp, (int)N0,
                          P.table<double>((int)N0), nf))
      throw PfError(PF_EHIP, "grouped eigenvalue launch rejected");
    int nb;
    if (even) {
      nb = even2d_blocks(g);
      G.partial.ensure((size_t)nf * nb * sizeof(ClassifyPartial));
      if (!fast_eig_c2r_y<float>(fc, (int)(N1 / 2), spec, g.Hp, (int)H, (int)(N0 / 2), inv_next,
                                 0., G.thr.as<double>(), nullptr, G.sq.as<float>(),
                                 G.partial.as<ClassifyPartial>(), P.table<double>((int)N1), nf,
                                 &nb))
        throw PfError(PF_EHIP, "grouped eigenvalue launch rejected");
      check_launch();
    } else {
      if (!fast_mid<double>(fc, (int)N1, false, spec, spec, g.Hp, (int)(N0 / 2), 1, (int)N1,
                            (int)N1, (int)N1, (int)N1, P.table<double>((int)N1), nf))
        throw PfError(PF_EHIP, "grouped eigenvalue launch rejected");
      check_launch();
      nb = (int)std::min<unsigned>(1024, grid1d(count));
      G.partial.ensure((size_t)nf * nb * sizeof(ClassifyPartial));
      k_finish_r2c_eigs_group<float><<<dim3(nb, nf), 256, 0, s>>>(
        spec, g.Hp, (int)H, rows, inv_next, G.thr.as<double>(), nullptr, G.sq.as<float>(),
        G.partial.as<ClassifyPartial>());
      check_launch();
    }
    k_reduce_partials_group<<<nf, 256, 0, s>>>(G.partial.as<ClassifyPartial>(), nb,
                                               G.fidx.as<int64_t>(), stats.as<ClassifyPartial>());
    check_launch();
    GenGeom gg = gen_geom(P);
    gg.fidx = G.fidx.as<int64_t>();
    gg.sq_stride = count;
    const LastAxis la = last_axis(g);
    G.w1.ensure((size_t)nf * g.Hp * la.nlast * sizeof(cplx<float>));
    cplx<float>* B1 = G.w1.as<cplx<float>>();
    FastCtx fx = fc;
    fx.ooff = G.ooff.as<int64_t>();
    if (!fast_gen_last<float>(fc, (int)N1, gg, la, G.sq.as<float>(), nullptr, seed, first, B1,
                              P.table<float>((int)N1), nf) ||
        !fast_c2r<float>(fx, (int)(N0 / 2), B1, (float*)out, g.n[1], g.Hp, g.n[0],
                         P.table<float>((int)N0), transform, (float)param, nf))
      throw PfError(PF_EHIP, "grouped generation launch rejected");
    check_launch();
  }
  void enqueue_unit(const BatchUnit& u, hipStream_t s, uint64_t seed, uint64_t first,
                    int transform, double param, void* out)
  {
    if (u.group >= 0) {
      enqueue_group(*groups[u.group], s, seed, first, transform, param, out);
    } else {
      plans[u.field]->stream = s;
      enqueue_field(u.field, seed, first, transform, param, out);
    }
  }
  // Every unit: FP64 eigenvalues (classification reduced on the device),
  // then one realization per field. No host synchronisation, so it can be
  // captured. With k > 1 branches the units are bin-packed by algorithmic
  // bytes (largest first onto the least loaded branch) and each branch runs
  // on its own stream forked from and joined to the origin stream by events.
  // time_units (eager, serial only): an event pair around every unit.
  void enqueue(uint64_t seed, uint64_t first, int transform, double param, void* out, int k,
               bool time_units = false)
  {
    const size_t n = units.size();
    // never fork a branch that would receive no unit: an empty forked
    // branch joins the capture without any node of its own
    k = std::min<int>(k, (int)n);
    if (k <= 1) {
      if (time_units) {
        while (fev.size() < 2 * n) {
          hipEvent_t e;
          HIPCHECK(hipEventCreate(&e));
          fev.push_back(e);
        }
      }
      for (size_t i = 0; i < n; i++) {
        if (time_units)
          HIPCHECK(hipEventRecord(fev[2 * i], stream));
        enqueue_unit(units[i], stream, seed, first, transform, param, out);
        if (time_units)
          HIPCHECK(hipEventRecord(fev[2 * i + 1], stream));
      }
      timed = time_units;
      return;
    }
    ensure_branches(k);
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++)
      order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return units[a].cost > units[b].cost; });
    std::vector<double> load(k, 0.);
    std::vector<std::vector<size_t>> bins(k);
    for (size_t i : order) {
      const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[j] += units[i].cost;
      bins[j].push_back(i);
    }
    HIPCHECK(hipEventRecord(ev_fork, stream));
    for (int j = 1; j < k; j++)
      HIPCHECK(hipStreamWaitEvent(branch[j - 1], ev_fork, 0));
    for (int j = 0; j < k; j++) {
      hipStream_t s = j == 0 ? stream : branch[j - 1];
      for (size_t i : bins[j])
        enqueue_unit(units[i], s, seed, first, transform, param, out);
    }
    for (int j = 1; j < k; j++) {
      HIPCHECK(hipEventRecord(ev[j - 1], branch[j - 1]));
      HIPCHECK(hipStreamWaitEvent(stream, ev[j - 1], 0));
    }
  }
};

#define PF_GUARD(...)                                                                              \
  try {                                                                                            \
    __VA_ARGS__                                                                                    \
  } catch (const PfError& e) {                                                                     \
    g_err = e.what();                                                                              \
    return e.code;                                                                                 \
  } catch (const std::exception& e) {                                                              \
    g_err = e.what();                                                                              \
    return PF_EINVAL;                                                                              \
  }

size_t elem_size(const Plan& P) { return P.d.precision == PF_FP32 ? 4 : 8; }

// ---------------------------------------------------------------------------
// Generation without allocation (pf_generate_noalloc): the work arrays of the
// passes live in one static device buffer per GPU. On the host `scratch` is
// a stand-in for the device copies, never an address: host code takes each
// device's address from hipGetSymbolAddress (once per device, with that
// device current) and passes that pointer to the kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PF_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;

void* scratch_ptr(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end())
    return it->second;
  void* ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr)
    throw PfError(PF_EHIP, "hipGetSymbolAddress returned no address for the scratch buffer");
  g_scratch_ptr[device] = ptr;
  return ptr;
}

// Bytes of B1 (+ B2 in 3D) of one realization, each rounded up to 256 bytes.
size_t noalloc_bytes_per_realization(const Plan& P)
{
  const Geo& g = P.g;
  const size_t cs = 2 * elem_size(P);
  const int nlast = g.dim == 3 ? g.n[2] : g.n[1];
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t b = up((size_t)g.Hp * (g.dim == 3 ? g.N[1] : 1) * nlast * cs);
  if (g.dim == 3)
    b += up((size_t)g.Hp * g.n[1] * g.n[2] * cs);
  return b;
}

// Realizations [0, batch) in groups that fit the buffer, on the plan stream,
// without the two-stream pipeline (which would create a stream and events).
template <typename T>
void generate_r2c_noalloc(Plan& P, uint64_t seed, uint64_t first, int batch, int transform,
                          double param, T* out, void* buf, int per_group)
{
  const Geo& g = P.g;
  const int nlast = g.dim == 3 ? g.n[2] : g.n[1];
  const size_t s1 = (size_t)g.Hp * (g.dim == 3 ? g.N[1] : 1) * nlast * sizeof(cplx<T>);
  const size_t s1g = (s1 * per_group + 255) / 256 * 256;
  cplx<T>* W1 = reinterpret_cast<cplx<T>*>(buf);
  cplx<T>* W2 = g.dim == 3 ? reinterpret_cast<cplx<T>*>(static_cast<char*>(buf) + s1g) : nullptr;
  for (int b0 = 0; b0 < batch; b0 += per_group) {
    const int nb = std::min(per_group, batch - b0);
    generate_r2c_chunk<T>(P, P.stream, P.stream, nullptr, nullptr, seed, first + b0, 0, nb,
                          transform, param, out + (int64_t)b0 * g.domain, W1, W2);
  }
}

// Generation dispatch by field backend and precision (device output).
void run_generate(Plan& P, const double* dn, uint64_t seed, uint64_t first, int batch,
                  int transform, double param, void* dout)
{
  const int fb = P.d.field_backend;
  if (P.d.precision == PF_FP32) {
    if (fb == PF_FIELD_DFT)
      generate_dft<float>(P, dn, seed, first, batch, transform, param, (float*)dout);
    else if (fb == PF_FIELD_DCTDST)
      generate_dctdst<float>(P, dn, seed, first, batch, transform, param, (float*)dout);
    else
      generate_r2c<float>(P, dn, seed, first, batch, transform, param, (float*)dout);
  } else {
    if (fb == PF_FIELD_DFT)
      generate_dft<double>(P, dn, seed, first, batch, transform, param, (double*)dout);
    else if (fb == PF_FIELD_DCTDST)
      generate_dctdst<double>(P, dn, seed, first, batch, transform, param, (double*)dout);
    else
      generate_r2c<double>(P, dn, seed, first, batch, transform, param, (double*)dout);
  }
}

// Whether every pass of the plan's generation has a specialised kernel (then
// the generic kernels' tile lines and radix order do not matter).
bool all_specialised(const Plan& P)
{
  if (fast_budget() == 0)
    return false;
  const Geo& g = P.g;
  if (P.d.field_backend == PF_FIELD_DFT)
    return false;
  if (P.d.field_backend == PF_FIELD_DCTDST) {
    DctGeom dg = dct_geom(P);
    FastCtx q{ nullptr, dct_fast_budget() };
    q.query = true;
    const int Nlast = g.dim == 3 ? g.N[2] : g.N[1];
    const int64_t xrows = (int64_t)g.n[1] * g.n[2];
    if (P.d.precision == PF_FP32)
      return fast_dct_gen_last<float>(q, Nlast / 2, dg, nullptr, nullptr, 0, 0, nullptr, nullptr, 1) &&
             (g.dim != 3 || fast_dct_mid<float>(q, g.N[1] / 2, dg, nullptr, nullptr, nullptr, 1)) &&
             fast_dct_x<float>(q, g.N[0] / 2, dg, nullptr, nullptr, xrows, nullptr, 0, 0.f, 1);
    return fast_dct_gen_last<double>(q, Nlast / 2, dg, nullptr, nullptr, 0, 0, nullptr, nullptr, 1) &&
           (g.dim != 3 || fast_dct_mid<double>(q, g.N[1] / 2, dg, nullptr, nullptr, nullptr, 1)) &&
           fast_dct_x<double>(q, g.N[0] / 2, dg, nullptr, nullptr, xrows, nullptr, 0, 0., 1);
  }
  FastCtx q{ nullptr, fast_budget() };
  q.query = true;
  const GenGeom gg = gen_geom(P);
  const LastAxis la = last_axis(g);
  const int Nlast = g.dim == 3 ? g.N[2] : g.N[1];
  const int64_t rows = (int64_t)g.n[1] * g.n[2];
  auto all = [&](auto t) {
    using T = decltype(t);
    return fast_gen_last<T>(q, Nlast, gg, la, nullptr, nullptr, 0, 0, nullptr, nullptr, 1) &&
           (g.dim != 3 || fast_mid<T>(q, g.N[1], true, nullptr, nullptr, g.Hp, g.N[0] / 2, g.n[2],
                                      g.N[1], g.N[1], g.n[1], g.n[1], nullptr, 1)) &&
           fast_c2r<T>(q, g.N[0] / 2, nullptr, nullptr, rows, g.Hp, g.n[0], nullptr, 0, T(0), 1);
  };
  return P.d.precision == PF_FP32 ? all(0.f) : all(0.);
}

// Autotuner cache key: device, field geometry and batch.
std::string tune_key(const Plan& P, int batch)
{
  hipDeviceProp_t prop{};
  HIPCHECK(hipGetDeviceProperties(&prop, P.device));
  const Geo& g = P.g;
  char k[512];
  std::snprintf(k, sizeof k, "%s:%dcu|%s|m%d|f%d|d%d|n%dx%dx%d|N%dx%dx%d|t%d|b%d", prop.gcnArchName,
                prop.multiProcessorCount, P.d.precision == PF_FP32 ? "fp32" : "fp64",
                P.d.matrix_backend, P.d.field_backend, g.dim, g.n[0], g.n[1], g.n[2], g.N[0],
                g.N[1], g.N[2], g.transposed, batch);
  return k;
}

bool tune_lookup(const std::string& path, const std::string& key, Tuning& t, double& ms)
{
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f)
    return false;
  char line[1024];
  bool found = false;
  while (std::fgets(line, sizeof line, f)) {
    char k[600];
    Tuning x;
    double m = 0.;
    if (std::sscanf(line, "%599s %d %d %d %d %d %d %lf", k, &x.ldiv, &x.chunks, &x.overlap,
                    &x.lines, &x.radix, &x.sub, &m) == 8 &&
        key == k) {
      t = x;
      ms = m;
      found = true; // the last entry of a key wins
    }
  }
  std::fclose(f);
  return found;
}

void tune_store(const std::string& path, const std::string& key, const Tuning& t, double ms)
{
  FILE* f = std::fopen(path.c_str(), "a");
  if (!f)
    throw PfError(PF_EINVAL, "cannot write the tuning cache " + path);
  std::fprintf(f, "%s %d %d %d %d %d %d %.6f\n", key.c_str(), t.ldiv, t.chunks, t.overlap,
               t.lines, t.radix, t.sub, ms);
  std::fclose(f);
}

} // namespace

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_plan_create(const pf_desc* desc, void** plan)
{
  PF_GUARD({
    if (!desc || !plan)
      throw PfError(PF_EINVAL, "null argument");
    auto P = std::make_unique<Plan>();
    P->d = *desc;
    P->g = make_geo(*desc);
    HIPCHECK(hipGetDevice(&P->device));
    HIPCHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    if (getenv("PF_PROFILE"))
      P->profiling = true;
    *plan = P.release();
    return PF_OK;
  })
}

int pf_plan_destroy(void* plan)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    if (P) {
      (void)hipStreamSynchronize(P->stream);
      delete P;
    }
    return PF_OK;
  })
}

int pf_plan_sizes(void* plan, pf_sizes* s)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    const Geo& g = P->g;
    s->dim = g.dim;
    for (int i = 0; i < 3; i++) {
      s->cells[i] = g.n[i];
      s->extended[i] = g.N[i];
    }
    s->transposed = g.transposed;
    s->domain_size = g.domain;
    s->extended_size = g.Next;
    // DFT matrix: the reference's array holds every entry of the grid
    s->eig_entries = P->d.matrix_backend == PF_MATRIX_DFT ? g.Next : g.eig_entries;
    s->noise_per_field = g.noise_per_field;
    return PF_OK;
  })
}

int pf_device_sync(void* plan)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    P->sync();
    return PF_OK;
  })
}

int pf_eigs(void* plan, pf_eig_stats* stats)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (!P->eig_valid)
      compute_eigs(*P, nullptr);
    prepare_generation(*P);
    if (stats)
      *stats = P->st;
    if (P->eig_status != PF_OK)
      g_err = "negative eigenvalues in covariance matrix";
    return P->eig_status;
  })
}

int pf_eigs_custom(void* plan, const double* values, pf_eig_stats* stats)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    compute_eigs(*P, values);
    prepare_generation(*P);
    if (stats)
      *stats = P->st;
    if (P->eig_status != PF_OK)
      g_err = "negative eigenvalues in covariance matrix";
    return P->eig_status;
  })
}

int pf_eigs_copy(void* plan, double* out, int64_t n)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (!P->eig_valid)
      compute_eigs(*P, nullptr);
    const Geo& g = P->g;
    const bool dft = P->d.matrix_backend == PF_MATRIX_DFT;
    if (n > (dft ? g.Next : g.eig_entries))
      throw PfError(PF_EINVAL, "n exceeds the number of eigenvalues");
    std::vector<double> logical(g.eig_entries);
    HIPCHECK(hipMemcpy(logical.data(), P->eig.p, g.eig_entries * sizeof(double),
                       hipMemcpyDeviceToHost));
    // reference storage order (3D transposed: x, z, y)
    const bool tr = g.transposed;
    if (dft) {
      // every entry of the grid in the storage order of DFTMatrixBackend
      // (last two axes swapped when transposed, dftmatrixbackend.hh:150-158);
      // the table holds kx < H, the rest is Re X(-k)
      const int64_t N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
      for (int64_t s = 0; s < n; s++) {
        int64_t k0, k1, k2;
        if (g.dim == 2 && tr) {
          k1 = s % N1;
          k0 = s / N1;
          k2 = 0;
        } else if (g.dim == 3 && tr) {
          k0 = s % N0;
          k2 = (s / N0) % N2;
          k1 = s / (N0 * N2);
        } else {
          k0 = s % N0;
          k1 = (s / N0) % N1;
          k2 = s / (N0 * N1);
        }
        if (k0 >= g.H) {
          k0 = N0 - k0;
          k1 = (N1 - k1) % N1;
          k2 = (N2 - k2) % N2;
        }
        out[s] = logical[k0 + g.H * (k1 + N1 * k2)];
      }
    } else if (P->d.matrix_backend == PF_MATRIX_DCT) {
      int perm[3] = { 0, 1, 2 };
      if (tr)
        std::swap(perm[g.dim - 2], perm[g.dim - 1]);
      for (int64_t s = 0; s < n; s++) {
        int64_t k[3], r = s;
        for (int j = 0; j < 3; j++) {
          k[perm[j]] = r % g.P[perm[j]];
          r /= g.P[perm[j]];
        }
        out[s] = logical[k[0] + g.P[0] * (k[1] + (int64_t)g.P[1] * k[2])];
      }
    } else {
      for (int64_t s = 0; s < n; s++) {
        int64_t kx = s % g.H, k1, k2;
        if (g.dim == 3 && tr) {
          k2 = (s / g.H) % g.N[2];
          k1 = s / ((int64_t)g.H * g.N[2]);
        } else {
          k1 = (s / g.H) % g.N[1];
          k2 = s / ((int64_t)g.H * g.N[1]);
        }
        out[s] = logical[kx + g.H * (k1 + (int64_t)g.N[1] * k2)];
      }
    }
    return PF_OK;
  })
}

int pf_generate(void* plan, const double* noise, uint64_t seed, uint64_t first_realization,
                int32_t batch, int32_t transform, double transform_param, void* out,
                int32_t out_mem)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (batch < 1)
      throw PfError(PF_EINVAL, "batch must be >= 1");
    if (transform < 0 || transform > PF_TRANSFORM_BOXCOX)
      throw PfError(PF_EINVAL, "transform type not known");
    ensure_eigs(*P);
    const Geo& g = P->g;
    const bool dftf = P->d.field_backend == PF_FIELD_DFT;
    const double* dn = nullptr;
    if (noise) {
      // DFT field: one draw of noise_per_field normals fills two fields
      const int64_t draws = dftf ? (batch + 1) / 2 : batch;
      const size_t nb = (size_t)draws * g.noise_per_field * sizeof(double);
      P->noise.ensure(nb);
      HIPCHECK(hipMemcpyAsync(P->noise.p, noise, nb, hipMemcpyHostToDevice, P->stream));
      dn = P->noise.as<double>();
    }
    const size_t es = elem_size(*P);
    const size_t outb = (size_t)batch * g.domain * es;
    void* dout = out;
    if (out_mem == PF_MEM_HOST) {
      P->stage_out.ensure(outb);
      dout = P->stage_out.p;
    }
    run_generate(*P, dn, seed, first_realization, batch, transform, transform_param, dout);
    if (out_mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(out, dout, outb, hipMemcpyDeviceToHost, P->stream));
    P->sync();
    return PF_OK;
  })
}

static void to_c(const Tuning& t, pf_tuning* o)
{
  o->tile_div = t.ldiv;
  o->chunks = t.chunks;
  o->overlap = t.overlap;
  o->lines = t.lines;
  o->radix_order = t.radix;
  o->sub_batch = t.sub;
}

int pf_get_tuning(void* plan, pf_tuning* t)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    to_c(P->tune, t);
    return PF_OK;
  })
}

int pf_set_tuning(void* plan, const pf_tuning* t)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (t->tile_div < 0 || t->tile_div > 2 || t->chunks < 0 || t->chunks > 8 ||
        t->overlap < -1 || t->overlap > 1 || t->lines < 0 || t->radix_order < 0 ||
        t->radix_order > 2 || t->sub_batch < 0)
      throw PfError(PF_EINVAL, "tuning value out of range");
    Tuning x;
    x.ldiv = t->tile_div;
    x.chunks = t->chunks;
    x.overlap = t->overlap;
    x.lines = t->lines;
    x.radix = t->radix_order;
    x.sub = t->sub_batch;
    P->tune = x;
    return PF_OK;
  })
}

// The autotuner: times pf_generate of `batch` device-Philox realizations
// into `out` (device memory) under each candidate and keeps the fastest in
// the plan. Candidates, tuned one group after the other: the tile width of
// the y and x passes and the two-stream pipeline (off, or on with 2, 4, 8
// chunks) for the R2C field, the realizations per sub-batch for the DCT/DST
// and DFT fields, then the generic kernels' tile lines and radix order when
// some pass runs on them. cache_path (optional): a text file of earlier
// results, one line per device and field shape; a hit applies the stored
// choice without timing, a miss appends the new result.
int pf_autotune(void* plan, int32_t batch, uint64_t seed, int32_t transform, double param,
                void* out, const char* cache_path, pf_tune_result* res)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    if (!P || !out || batch < 1)
      throw PfError(PF_EINVAL, "pf_autotune: null plan or output, or batch < 1");
    std::lock_guard<std::mutex> lk(P->mu);
    ensure_eigs(*P);
    const std::string key = tune_key(*P, batch);
    pf_tune_result r{};
    double cached_ms = 0.;
    Tuning hit;
    if (cache_path && tune_lookup(cache_path, key, hit, cached_ms)) {
      P->tune = hit;
      to_c(hit, &r.best);
      r.ms = cached_ms;
      r.from_cache = 1;
      if (res)
        *res = r;
      return PF_OK;
    }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    auto time_it = [&](const Tuning& t) {
      P->tune = t;
      run_generate(*P, nullptr, seed, 0, batch, transform, param, out); // allocates, warms
      float best = 1e30f;
      for (int rep = 0; rep < 2; rep++) {
        HIPCHECK(hipEventRecord(a, P->stream));
        run_generate(*P, nullptr, seed, 0, batch, transform, param, out);
        HIPCHECK(hipEventRecord(b, P->stream));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        best = std::min(best, ms);
      }
      r.candidates++;
      return (double)best;
    };
    Tuning best;
    double best_ms = time_it(best);
    r.default_ms = best_ms;
    auto consider = [&](const Tuning& t) {
      const double ms = time_it(t);
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    };
    const int fb = P->d.field_backend;
    if (fb == PF_FIELD_R2C) {
      Tuning base = best;
      for (int ldiv : { 1, 2 })
        for (int ov : { 0, 1 })
          for (int ch : { 2, 4, 8 }) {
            if (ov == 0 && ch != 8)
              continue;
            if (ov == 1 && batch < 4)
              continue;
            Tuning t = base;
            t.ldiv = ldiv;
            t.overlap = ov;
            t.chunks = ov ? ch : 0;
            consider(t);
          }
    } else {
      Tuning base = best;
      for (int sub : { 1, 2, 4 })
        if (sub < batch) {
          Tuning t = base;
          t.sub = sub;
          consider(t);
        }
    }
    if (!all_specialised(*P)) {
      Tuning base = best;
      for (int lines : { 0, 4, 2, 1 })
        for (int radix : { 0, 1, 2 }) {
          if (lines == 0 && radix == 0)
            continue;
          Tuning t = base;
          t.lines = lines;
          t.radix = radix;
          consider(t);
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    P->tune = best;
    to_c(best, &r.best);
    r.ms = best_ms;
    if (cache_path)
      tune_store(cache_path, key, best, best_ms);
    if (res)
      *res = r;
    return PF_OK;
  })
}

int pf_generate_noalloc(void* plan, uint64_t seed, uint64_t first_realization, int32_t batch,
                        int32_t transform, double transform_param, void* out)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    if (!P || !out)
      throw PfError(PF_EINVAL, "pf_generate_noalloc: null plan or output");
    std::lock_guard<std::mutex> lk(P->mu);
    if (batch < 1)
      throw PfError(PF_EINVAL, "batch must be >= 1");
    if (transform < 0 || transform > PF_TRANSFORM_BOXCOX)
      throw PfError(PF_EINVAL, "transform type not known");
    if (P->d.field_backend != PF_FIELD_R2C)
      throw PfError(PF_EINVAL, "pf_generate_noalloc supports the R2C field backend only");
    if (!P->eig_valid)
      throw PfError(PF_ESTATE, "pf_generate_noalloc needs a prepared plan: call pf_eigs first "
                               "(it computes the eigenvalues and transform tables, which "
                               "allocates)");
    if (P->eig_status != PF_OK)
      throw PfError(PF_ENEG, "negative eigenvalues in covariance matrix");
    const size_t per = noalloc_bytes_per_realization(*P);
    if (per > (size_t)PF_NOALLOC_SCRATCH_BYTES)
      throw PfError(PF_EINVAL, "one realization needs " + std::to_string(per) +
                                 " bytes of work arrays, above the " +
                                 std::to_string((size_t)PF_NOALLOC_SCRATCH_BYTES) +
                                 "-byte buffer of pf_generate_noalloc");
    // the output must be device memory (no staging without allocation)
    hipPointerAttribute_t attr{};
    const hipError_t pe = hipPointerGetAttributes(&attr, out);
    (void)hipGetLastError(); // a host pointer leaves an error that the launch check would see
    if (pe != hipSuccess || attr.type != hipMemoryTypeDevice)
      throw PfError(PF_EINVAL, "pf_generate_noalloc writes device memory only");
    const int per_group = (int)std::min<size_t>((size_t)batch, PF_NOALLOC_SCRATCH_BYTES / per);
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != P->device)
      HIPCHECK(hipSetDevice(P->device));
    struct Restore
    {
      Plan* p;
      int dev, cur;
      ~Restore()
      {
        p->noalloc = false;
        if (dev != cur)
          (void)hipSetDevice(cur);
      }
    } restore{ P, P->device, cur };
    void* buf = scratch_ptr(P->device);
    P->noalloc = true;
    if (P->d.precision == PF_FP32)
      generate_r2c_noalloc<float>(*P, seed, first_realization, batch, transform, transform_param,
                                  static_cast<float*>(out), buf, per_group);
    else
      generate_r2c_noalloc<double>(*P, seed, first_realization, batch, transform,
                                   transform_param, static_cast<double*>(out), buf, per_group);
    HIPCHECK(hipStreamSynchronize(P->stream));
    return PF_OK;
  })
}

int pf_matvec(void* plan, int32_t op, const void* x, void* y, int32_t batch, int32_t mem)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (op < 0 || op > PF_OP_INVERSE_EMBED)
      throw PfError(PF_EINVAL, "unknown matvec operator");
    ensure_eigs(*P);
    const Geo& g = P->g;
    const size_t es = elem_size(*P);
    const size_t nb = (size_t)batch * g.domain * es;
    const void* dx = x;
    void* dy = y;
    if (mem == PF_MEM_HOST) {
      P->stage_in.ensure(nb);
      P->stage_out.ensure(nb);
      HIPCHECK(hipMemcpyAsync(P->stage_in.p, x, nb, hipMemcpyHostToDevice, P->stream));
      dx = P->stage_in.p;
      dy = P->stage_out.p;
    }
    if (g.dim == 1) {
      if (P->d.precision == PF_FP32)
        matvec_dft_1d<float>(*P, op, (const float*)dx, (float*)dy, batch);
      else
        matvec_dft_1d<double>(*P, op, (const double*)dx, (double*)dy, batch);
    } else if (P->d.precision == PF_FP32)
      matvec_r2c<float>(*P, op, (const float*)dx, (float*)dy, batch);
    else
      matvec_r2c<double>(*P, op, (const double*)dx, (double*)dy, batch);
    if (mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(y, dy, nb, hipMemcpyDeviceToHost, P->stream));
    P->sync();
    return PF_OK;
  })
}

int pf_cg(void* plan, const double* b, double* x, int32_t mem, const pf_cg_opts* opts,
          pf_cg_stats* stats)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    ensure_eigs(*P);
    const int64_t n = P->g.domain;
    const int mode = opts ? opts->mode : PF_CG_REFERENCE;
    const int maxit = (opts && opts->max_iterations > 0) ? opts->max_iterations
                                                         : P->d.cg_iterations;
    const double tol = opts ? opts->rel_tol : 1e-8;
    if (mode == PF_CG_LEARNED_PCG) {
      if (!opts->bin_log_multipliers || opts->nbins < 1)
        throw PfError(PF_EINVAL, "learned PCG needs bin_log_multipliers");
      std::vector<double> t(opts->nbins);
      for (int i = 0; i < opts->nbins; i++)
        t[i] = std::exp((double)opts->bin_log_multipliers[i]);
      P->spec.ensure(t.size() * sizeof(double));
      HIPCHECK(hipMemcpy(P->spec.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
      P->spec_nbins = opts->nbins;
    }
    // the CG runs in FP64 whatever the plan's field precision
    if (P->d.precision != PF_FP64)
      throw PfError(PF_EINVAL, "pf_cg needs a PF_FP64 plan (FP64 matvecs and dot products)");
    Vec vb, vx;
    double* db = vb.p(n);
    double* dx = vx.p(n);
    HIPCHECK(hipMemcpyAsync(db, b, n * sizeof(double),
                            mem == PF_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                            P->stream));
    auto t0 = std::chrono::steady_clock::now();
    int it = 0;
    double rr = 0.;
    if (mode == PF_CG_REFERENCE) {
      // Matrix::multiplyInverse, matrix.hh:189-206
      std::vector<double> hb(n);
      HIPCHECK(hipMemcpyAsync(hb.data(), db, n * sizeof(double), hipMemcpyDeviceToHost,
                              P->stream));
      HIPCHECK(hipStreamSynchronize(P->stream));
      bool fieldZero = true;
      for (int64_t i = 0; i < n; i++)
        if (std::abs(hb[i]) > 1e-10)
          fieldZero = false;
      if (fieldZero)
        HIPCHECK(hipMemcpyAsync(dx, db, n * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
      else {
        matvec64(*P, PF_OP_INVERSE_EMBED, db, dx);
        it = inner_cg(*P, dx, db, n, (unsigned)maxit);
      }
      Vec vax, vr;
      double* ax = vax.p(n);
      double* r = vr.p(n);
      matvec64(*P, PF_OP_MULTIPLY, dx, ax);
      k_sub<<<grid1d(n), 256, 0, P->stream>>>(db, ax, r, n);
      const double bn = std::sqrt(dot(*P, db, db, n));
      rr = bn > 0 ? std::sqrt(dot(*P, r, r, n)) / bn : 0.;
    } else {
      it = pcg(*P, db, dx, n, mode, maxit, tol, &rr);
    }
    P->sync();
    auto t1 = std::chrono::steady_clock::now();
    HIPCHECK(hipMemcpyAsync(x, dx, n * sizeof(double),
                            mem == PF_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                            P->stream));
    P->sync();
    if (stats) {
      stats->iterations = it;
      stats->rel_residual = rr;
      stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return PF_OK;
  })
}

int pf_value_transform(void* plan, int32_t transform, double param, void* data, int64_t n,
                       int32_t mem)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (transform < 0 || transform > PF_TRANSFORM_BOXCOX)
      throw PfError(PF_EINVAL, "transform type not known");
    const size_t es = elem_size(*P);
    void* d = data;
    if (mem == PF_MEM_HOST) {
      P->stage_in.ensure(n * es);
      HIPCHECK(hipMemcpyAsync(P->stage_in.p, data, n * es, hipMemcpyHostToDevice, P->stream));
      d = P->stage_in.p;
    }
    if (es == 4)
      k_value_transform<float><<<grid1d(n), 256, 0, P->stream>>>((float*)d, n, transform,
                                                                 (float)param);
    else
      k_value_transform<double><<<grid1d(n), 256, 0, P->stream>>>((double*)d, n, transform,
                                                                  param);
    check_launch();
    if (mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(data, d, n * es, hipMemcpyDeviceToHost, P->stream));
    P->sync();
    return PF_OK;
  })
}

int pf_philox_normals(void* plan, uint64_t seed, uint64_t realization, uint64_t first, int64_t n,
                      double* out, int32_t mem)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    double* d = out;
    if (mem == PF_MEM_HOST) {
      P->stage_out.ensure(n * sizeof(double));
      d = P->stage_out.as<double>();
    }
    k_philox_normals<<<grid1d(n), 256, 0, P->stream>>>(seed, realization, first, n, d);
    check_launch();
    if (mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    P->sync();
    return PF_OK;
  })
}

int pf_philox_raw(void* plan, const uint32_t ctr[4], const uint32_t key[2], int64_t nblocks,
                  uint32_t* out, int32_t mem)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    uint32_t* d = out;
    const size_t nb = (size_t)nblocks * 4 * sizeof(uint32_t);
    if (mem == PF_MEM_HOST) {
      P->stage_out.ensure(nb);
      d = P->stage_out.as<uint32_t>();
    }
    k_philox_raw<<<grid1d(nblocks), 256, 0, P->stream>>>(ctr[0], ctr[1], ctr[2], ctr[3], key[0],
                                                         key[1], nblocks, d);
    check_launch();
    if (mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(out, d, nb, hipMemcpyDeviceToHost, P->stream));
    P->sync();
    return PF_OK;
  })
}


// ---- learned spectral preconditioner --------------------------------------

int pf_spectral_targets(void* plan, int32_t nbins, double floor_value, double* log_targets)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    if (P->d.matrix_backend != PF_MATRIX_DCT)
      throw PfError(PF_EINVAL, "spectral targets need the DCT matrix backend");
    if (nbins < 1)
      throw PfError(PF_EINVAL, "nbins must be positive");
    if (!P->eig_valid)
      compute_eigs(*P, nullptr);
    const Geo& g = P->g;
    DevBuf sums, counts;
    sums.ensure(nbins * sizeof(double));
    counts.ensure(nbins * sizeof(unsigned long long));
    HIPCHECK(hipMemsetAsync(sums.p, 0, nbins * sizeof(double), P->stream));
    HIPCHECK(hipMemsetAsync(counts.p, 0, nbins * sizeof(unsigned long long), P->stream));
    const int64_t total = (int64_t)g.P[0] * g.P[1] * g.P[2];
    const size_t lds = (size_t)nbins * (sizeof(double) + sizeof(unsigned long long));
    if (lds > 65536)
      throw PfError(PF_EINVAL, "nbins too large for the per-block bins");
    k_radial_sums<<<std::min<unsigned>(1024, grid1d(total)), 256, lds, P->stream>>>(
      P->eig.as<double>(), g.P[0], g.P[1], g.P[2], g.dim, nbins, floor_value, sums.as<double>(),
      counts.as<unsigned long long>());
    check_launch();
    std::vector<double> hs(nbins);
    std::vector<unsigned long long> hc(nbins);
    HIPCHECK(hipMemcpyAsync(hs.data(), sums.p, nbins * sizeof(double), hipMemcpyDeviceToHost,
                            P->stream));
    HIPCHECK(hipMemcpyAsync(hc.data(), counts.p, nbins * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, P->stream));
    P->sync();
    double last = 0.;
    for (int b = 0; b < nbins; b++) {
      // empty bins (possible near r = 0) take the nearest filled neighbour
      if (hc[b] > 0)
        last = -std::log(hs[b] / (double)hc[b]);
      log_targets[b] = hc[b] > 0 ? last : NAN;
    }
    for (int b = nbins - 1; b >= 0; b--)
      if (std::isnan(log_targets[b]))
        log_targets[b] = (b + 1 < nbins) ? log_targets[b + 1] : last;
    return PF_OK;
  })
}

int pf_mlp_create(const int32_t* dims, int32_t ndims, uint64_t seed, void** mlp)
{
  PF_GUARD({
    if (ndims < 2)
      throw PfError(PF_EINVAL, "an MLP needs at least input and output sizes");
    auto M = std::make_unique<Mlp>();
    M->dims.assign(dims, dims + ndims);
    int64_t off = 0;
    for (int l = 0; l + 1 < ndims; l++) {
      M->woff.push_back(off);
      off += (int64_t)dims[l] * dims[l + 1];
      M->boff.push_back(off);
      off += dims[l + 1];
    }
    M->nparams = off;
    HIPCHECK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    M->params.ensure(off * sizeof(float));
    M->grads.ensure(off * sizeof(float));
    M->m.ensure(off * sizeof(float));
    M->v.ensure(off * sizeof(float));
    // LeCun-normal weights, zero biases
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l + 1 < ndims; l++) {
      std::normal_distribution<double> nd(0., 1. / std::sqrt((double)dims[l]));
      for (int64_t i = 0; i < (int64_t)dims[l] * dims[l + 1]; i++)
        h[M->woff[l] + i] = (float)nd(rng);
    }
    HIPCHECK(hipMemcpy(M->params.p, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(M->m.p, 0, off * sizeof(float)));
    HIPCHECK(hipMemset(M->v.p, 0, off * sizeof(float)));
    *mlp = M.release();
    return PF_OK;
  })
}

int pf_mlp_destroy(void* mlp)
{
  PF_GUARD({
    delete static_cast<Mlp*>(mlp);
    return PF_OK;
  })
}

int pf_mlp_nparams(void* mlp, int64_t* n)
{
  PF_GUARD({
    *n = static_cast<Mlp*>(mlp)->nparams;
    return PF_OK;
  })
}

int pf_mlp_params(void* mlp, float* values, int64_t n, int32_t set)
{
  PF_GUARD({
    Mlp* M = static_cast<Mlp*>(mlp);
    if (n != M->nparams)
      throw PfError(PF_EINVAL, "parameter count mismatch");
    if (set)
      HIPCHECK(hipMemcpy(M->params.p, values, n * sizeof(float), hipMemcpyHostToDevice));
    else
      HIPCHECK(hipMemcpy(values, M->params.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return PF_OK;
  })
}

int pf_mlp_forward(void* mlp, const float* x, int32_t batch, float* y)
{
  PF_GUARD({
    Mlp* M = static_cast<Mlp*>(mlp);
    M->ensure_batch(batch);
    HIPCHECK(hipMemcpyAsync(M->x.p, x, (size_t)batch * M->dims[0] * sizeof(float),
                            hipMemcpyHostToDevice, M->stream));
    M->forward(batch);
    HIPCHECK(hipMemcpyAsync(y, M->A[M->nl() - 1].p, (size_t)batch * M->dims.back() * sizeof(float),
                            hipMemcpyDeviceToHost, M->stream));
    HIPCHECK(hipStreamSynchronize(M->stream));
    return PF_OK;
  })
}

int pf_mlp_gradients(void* mlp, const float* x, const float* y, int32_t batch, float* grads,
                     double* loss)
{
  PF_GUARD({
    Mlp* M = static_cast<Mlp*>(mlp);
    M->ensure_batch(batch);
    HIPCHECK(hipMemcpyAsync(M->x.p, x, (size_t)batch * M->dims[0] * sizeof(float),
                            hipMemcpyHostToDevice, M->stream));
    HIPCHECK(hipMemcpyAsync(M->y.p, y, (size_t)batch * M->dims.back() * sizeof(float),
                            hipMemcpyHostToDevice, M->stream));
    M->forward(batch);
    const double l = M->backward(batch);
    HIPCHECK(hipMemcpy(grads, M->grads.p, M->nparams * sizeof(float), hipMemcpyDeviceToHost));
    if (loss)
      *loss = l;
    return PF_OK;
  })
}

// Full-batch Adam on the mean-squared error for `steps` steps.
int pf_mlp_train(void* mlp, const float* x, const float* y, int32_t batch, int32_t steps, float lr,
                 double* final_loss)
{
  PF_GUARD({
    Mlp* M = static_cast<Mlp*>(mlp);
    M->ensure_batch(batch);
    HIPCHECK(hipMemcpyAsync(M->x.p, x, (size_t)batch * M->dims[0] * sizeof(float),
                            hipMemcpyHostToDevice, M->stream));
    HIPCHECK(hipMemcpyAsync(M->y.p, y, (size_t)batch * M->dims.back() * sizeof(float),
                            hipMemcpyHostToDevice, M->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double l = 0.;
    for (int s = 0; s < steps; s++) {
      M->forward(batch);
      l = M->backward(batch);
      M->t++;
      const float c1 = 1.f - std::pow(b1, (float)M->t), c2 = 1.f - std::pow(b2, (float)M->t);
      k_mlp_adam<<<grid1d(M->nparams), 256, 0, M->stream>>>(
        M->params.as<float>(), M->grads.as<float>(), M->m.as<float>(), M->v.as<float>(),
        M->nparams, lr, b1, b2, eps, c1, c2);
      check_launch();
    }
    HIPCHECK(hipStreamSynchronize(M->stream));
    if (final_loss)
      *final_loss = l;
    return PF_OK;
  })
}


int pf_batch_create(const pf_desc* descs, int32_t n, void** batch)
{
  PF_GUARD({
    if (!descs || n < 1 || !batch)
      throw PfError(PF_EINVAL, "a batch needs at least one descriptor");
    auto B = std::make_unique<Batch>();
    HIPCHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    int64_t off = 0;
    for (int i = 0; i < n; i++) {
      if (descs[i].precision != descs[0].precision)
        throw PfError(PF_EINVAL, "all fields of a batch must use one precision");
      if (descs[i].periodization == PF_PERIOD_COFOLD)
        throw PfError(PF_EINVAL, "cofold fields are not supported in a batch (the cofold offset "
                                 "is reduced on the host between its two passes)");
      if (descs[i].matrix_backend == PF_MATRIX_DFT || descs[i].field_backend == PF_FIELD_DFT)
        throw PfError(PF_EINVAL, "DFT backends are not supported in a batch (a draw gives two "
                                 "fields; generate them with pf_generate)");
      auto P = std::make_unique<Plan>();
      P->d = descs[i];
      P->g = make_geo(descs[i]);
      HIPCHECK(hipGetDevice(&P->device));
      P->stream = B->stream;
      P->own_stream = false;
      P->capturing = true;
      B->offsets.push_back(off);
      off += P->g.domain;
      // algorithmic bytes of one field: FP64 eigen passes + spectrum + output
      const double es = P->d.precision == PF_FP32 ? 4. : 8.;
      B->cost.push_back(48. * P->g.half + 4. * es * P->g.half + es * P->g.domain);
      B->plans.push_back(std::move(P));
    }
    B->offsets.push_back(off);
    B->stats.ensure((size_t)n * sizeof(ClassifyPartial));
    // group the fields of one shape that have specialised kernels for every
    // pass (PF_BATCH_GROUP=0: every field through its own plan)
    const char* ge = getenv("PF_BATCH_GROUP");
    const bool grouping = !(ge && ge[0] == '0');
    std::map<std::array<int64_t, 6>, int> key_group;
    for (int i = 0; i < n; i++) {
      const Plan& P = *B->plans[i];
      if (!grouping || !groupable(P)) {
        B->units.push_back({ -1, (size_t)i, B->cost[i] });
        continue;
      }
      const std::array<int64_t, 6> key = { P.g.n[0], P.g.n[1], P.g.N[0], P.g.N[1],
                                           (int64_t)P.g.transposed,
                                           (int64_t)P.d.covariance * 4 + P.d.periodization };
      auto it = key_group.find(key);
      if (it == key_group.end()) {
        it = key_group.emplace(key, (int)B->groups.size()).first;
        B->groups.push_back(std::make_unique<FieldGroup>());
        B->units.push_back({ it->second, 0, 0. });
      }
      B->groups[it->second]->fields.push_back(i);
    }
    for (auto& u : B->units) {
      if (u.group < 0)
        continue;
      FieldGroup& G = *B->groups[u.group];
      const size_t nf = G.fields.size();
      u.field = G.fields[0];
      std::vector<CovParams> cps(nf);
      std::vector<int64_t> fidx(nf), ooff(nf);
      std::vector<double> thr(nf);
      for (size_t f = 0; f < nf; f++) {
        const Plan& P = *B->plans[G.fields[f]];
        cps[f] = make_cov(P.d, P.g);
        fidx[f] = (int64_t)G.fields[f];
        ooff[f] = B->offsets[G.fields[f]];
        thr[f] = P.d.threshold;
        u.cost += B->cost[G.fields[f]];
      }
      G.cov.ensure(nf * sizeof(CovParams));
      G.fidx.ensure(nf * sizeof(int64_t));
      G.ooff.ensure(nf * sizeof(int64_t));
      G.thr.ensure(nf * sizeof(double));
      HIPCHECK(hipMemcpy(G.cov.p, cps.data(), nf * sizeof(CovParams), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(G.fidx.p, fidx.data(), nf * sizeof(int64_t), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(G.ooff.p, ooff.data(), nf * sizeof(int64_t), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(G.thr.p, thr.data(), nf * sizeof(double), hipMemcpyHostToDevice));
    }
    *batch = B.release();
    return PF_OK;
  })
}

int pf_batch_destroy(void* batch)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    if (B) {
      (void)hipStreamSynchronize(B->stream);
      delete B;
    }
    return PF_OK;
  })
}

int pf_batch_offsets(void* batch, int64_t* offsets)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    for (size_t i = 0; i < B->offsets.size(); i++)
      offsets[i] = B->offsets[i];
    return PF_OK;
  })
}

int pf_batch_generate(void* batch, uint64_t seed, uint64_t first_realization, int32_t transform,
                      double transform_param, void* out, int32_t out_mem, int32_t use_graph,
                      pf_eig_stats* stats)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    if (transform < 0 || transform > PF_TRANSFORM_BOXCOX)
      throw PfError(PF_EINVAL, "transform type not known");
    const size_t es = B->plans[0]->d.precision == PF_FP32 ? 4 : 8;
    const size_t outb = (size_t)B->offsets.back() * es;
    void* dout = out;
    if (out_mem == PF_MEM_HOST) {
      B->stage.ensure(outb);
      dout = B->stage.p;
    }
    const int k = std::max(1, std::min<int>(B->branches, (int)B->plans.size()));
    B->timed = false;
    if (!B->prepared) {
      // eager pass: allocates every buffer and twiddle table the capture needs
      B->enqueue(seed, first_realization, transform, transform_param, dout, 1);
      HIPCHECK(hipStreamSynchronize(B->stream));
      B->prepared = true;
      if (!use_graph && B->field_timing) // the timed pass runs without allocations
        B->enqueue(seed, first_realization, transform, transform_param, dout, 1, true);
    } else if (!use_graph) {
      B->enqueue(seed, first_realization, transform, transform_param, dout, 1, B->field_timing);
    }
    if (use_graph) {
      const bool same = B->exec && B->key_seed == seed && B->key_first == first_realization &&
                        B->key_transform == transform && B->key_param == transform_param &&
                        B->key_out == dout && B->key_branches == k;
      if (!same) {
        if (B->exec)
          (void)hipGraphExecDestroy(B->exec);
        if (B->graph)
          (void)hipGraphDestroy(B->graph);
        B->exec = nullptr;
        B->graph = nullptr;
        B->ensure_branches(k);
        HIPCHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeRelaxed));
        B->enqueue(seed, first_realization, transform, transform_param, dout, k);
        HIPCHECK(hipStreamEndCapture(B->stream, &B->graph));
        HIPCHECK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->key_seed = seed;
        B->key_first = first_realization;
        B->key_transform = transform;
        B->key_param = transform_param;
        B->key_out = dout;
        B->key_branches = k;
      }
      HIPCHECK(hipGraphLaunch(B->exec, B->stream));
    }
    if (out_mem == PF_MEM_HOST)
      HIPCHECK(hipMemcpyAsync(out, dout, outb, hipMemcpyDeviceToHost, B->stream));
    HIPCHECK(hipStreamSynchronize(B->stream));
    const size_t n = B->plans.size();
    std::vector<ClassifyPartial> h(n);
    HIPCHECK(hipMemcpy(h.data(), B->stats.p, n * sizeof(ClassifyPartial), hipMemcpyDeviceToHost));
    int status = PF_OK;
    for (size_t i = 0; i < n; i++) {
      Plan& P = *B->plans[i];
      pf_eig_stats st{};
      st.small = h[i].small;
      st.zero = h[i].zero;
      st.negative = h[i].negative;
      st.smallest = h[i].smallest;
      st.entries = P.d.matrix_backend == PF_MATRIX_DCT
                     ? (int64_t)P.g.P[0] * P.g.P[1] * P.g.P[2]
                     : P.g.half;
      if (stats)
        stats[i] = st;
      if (st.negative > 0 && !P.d.approximate)
        status = PF_ENEG;
    }
    if (status != PF_OK)
      g_err = "negative eigenvalues in covariance matrix";
    return status;
  })
}

int pf_batch_set_field_timing(void* batch, int32_t on)
{
  PF_GUARD({
    static_cast<Batch*>(batch)->field_timing = on != 0;
    return PF_OK;
  })
}

int pf_batch_field_times(void* batch, double* ms)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    if (!B->timed)
      throw PfError(PF_ESTATE, "the last pf_batch_generate was not a timed eager pass");
    for (size_t i = 0; i < B->units.size(); i++) {
      float t = 0.f;
      HIPCHECK(hipEventElapsedTime(&t, B->fev[2 * i], B->fev[2 * i + 1]));
      const BatchUnit& u = B->units[i];
      if (u.group < 0) {
        ms[u.field] = t;
      } else {
        // a group's fields share its time equally
        const auto& f = B->groups[u.group]->fields;
        for (size_t j : f)
          ms[j] = t / (double)f.size();
      }
    }
    return PF_OK;
  })
}

int pf_batch_set_branches(void* batch, int32_t branches)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    if (branches < 1 || branches > 64)
      throw PfError(PF_EINVAL, "branches must be in [1, 64]");
    B->branches = branches;
    return PF_OK;
  })
}

// Autotuner of the batch packing: times one graph replay per candidate
// branch count (1, 2, 4, 8, 16) on the real batch and keeps the fastest.
int pf_batch_autotune(void* batch, uint64_t seed, uint64_t first_realization, int32_t transform,
                      double transform_param, void* out, int32_t out_mem, int32_t* best,
                      double* best_ms)
{
  PF_GUARD({
    Batch* B = static_cast<Batch*>(batch);
    int bestk = 1;
    double bestt = 1e300;
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    for (int k : { 1, 2, 4, 8, 16 }) {
      if (k > 1 && k > (int)B->plans.size())
        break; // more branches than fields: nothing left to try
      B->branches = k;
      int st = pf_batch_generate(batch, seed, first_realization, transform, transform_param, out,
                                 out_mem, 1, nullptr); // builds the graph for k
      if (st != PF_OK && st != PF_ENEG)
        return st;
      HIPCHECK(hipEventRecord(a, B->stream));
      HIPCHECK(hipGraphLaunch(B->exec, B->stream));
      HIPCHECK(hipEventRecord(b, B->stream));
      HIPCHECK(hipEventSynchronize(b));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, a, b));
      if (ms < bestt) {
        bestt = ms;
        bestk = k;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    B->branches = bestk;
    if (best)
      *best = bestk;
    if (best_ms)
      *best_ms = bestt;
    return PF_OK;
  })
}

// Per-launch device times (ms) of the last call when profiling is enabled
// (environment PF_PROFILE or pf_set_profiling): names and durations.
int pf_set_profiling(void* plan, int32_t on)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    P->profiling = on != 0;
    return PF_OK;
  })
}

int pf_last_timings(void* plan, char* names, int32_t names_len, double* ms, int32_t max_n,
                    int32_t* n_out)
{
  PF_GUARD({
    Plan* P = static_cast<Plan*>(plan);
    std::lock_guard<std::mutex> lk(P->mu);
    std::string all;
    int n = 0;
    for (auto& t : P->last_timings) {
      if (n < max_n)
        ms[n] = t.second;
      all += t.first;
      all += ';';
      n++;
    }
    if (names && names_len > 0) {
      std::strncpy(names, all.c_str(), names_len - 1);
      names[names_len - 1] = 0;
    }
    *n_out = n;
    return PF_OK;
  })
}

} // extern "C"
