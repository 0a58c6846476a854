// randomfield.hh -- C++ host API of the MI355X engine, a drop-in for the
// hot path of parafields-core's RandomField / Matrix
// (include/parafields/randomfield.hh, matrix.hh). Same method names,
// argument meaning and exceptions; the arithmetic runs on the GPU through
// the C ABI of parafields_amd.h.
//
//   parafields_amd::ParameterTree config;           // same keys as the reference
//   config["grid.cells"] = "256 256"; ...
//   parafields_amd::RandomField<parafields_amd::GridTraits<double, double, 2>> f(config);
//   f.generate(42);                                  // CppRNGBackend noise order
//   f.bulkEvaluate(values, sizes);
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../parafields_amd.h"
#include "io.hh"
#include "parametertree.hh"
#include "trend.hh"

namespace parafields_amd {

// exceptions.hh:8-15
class NegativeEigenvalueError : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

inline void check(int status)
{
  if (status == PF_OK)
    return;
  const std::string msg = pf_last_error();
  if (status == PF_ENEG)
    throw NegativeEigenvalueError(msg);
  throw std::runtime_error(msg);
}

// Backend tags: the template arguments of the reference's Matrix
// (matrix.hh:55-60).
struct DCTMatrixBackend
{
  static constexpr int id = PF_MATRIX_DCT;
};
struct R2CMatrixBackend
{
  static constexpr int id = PF_MATRIX_R2C;
};
struct R2CFieldBackend
{
  static constexpr int id = PF_FIELD_R2C;
};
struct DCTDSTFieldBackend
{
  static constexpr int id = PF_FIELD_DCTDST;
};
// DFTMatrixBackend / DFTFieldBackend (dftmatrixbackend.hh, dftfieldbackend.hh):
// the 1-D defaults (matrix.hh:1660-1736)
struct DFTMatrixBackend
{
  static constexpr int id = PF_MATRIX_DFT;
};
struct DFTFieldBackend
{
  static constexpr int id = PF_FIELD_DFT;
};

// Default backends per dimension (DefaultIsoMatrix / DefaultAnisoMatrix,
// matrix.hh:1696-1736): DCT matrix + R2C field (iso) and R2C matrix + R2C
// field (aniso) for dim > 1; the DFT matrix and field for 1-D.
template <int dim>
struct DefaultBackends
{
  using IsoMB = DCTMatrixBackend;
  using IsoFB = R2CFieldBackend;
  using AnisoMB = R2CMatrixBackend;
  using AnisoFB = R2CFieldBackend;
};
template <>
struct DefaultBackends<1>
{
  using IsoMB = DFTMatrixBackend;
  using IsoFB = DFTFieldBackend;
  using AnisoMB = DFTMatrixBackend;
  using AnisoFB = DFTFieldBackend;
};

// test/traits.hh analogue
template <typename DF, typename RF, unsigned int dimension>
struct GridTraits
{
  enum
  {
    dim = dimension
  };
  using RangeField = RF;
  using DomainField = DF;
  using Domain = std::array<DF, dimension>;
};

// Configuration of one field (RandomFieldTraits, fieldtraits.hh:169-209),
// reduced to the descriptor the engine consumes.
// Config: this library's ParameterTree or Dune::ParameterTree (the same
// get<T>(key), get<T>(key, default) and hasKey).
template <typename RF, int dim, typename Config>
pf_desc make_desc(const Config& c, int matrix_backend, int field_backend)
{
  pf_desc d{};
  d.dim = dim;
  const auto cells = c.template get<std::array<int, dim>>("grid.cells");
  const auto ext = c.template get<std::array<double, dim>>("grid.extensions");
  for (int i = 0; i < dim; i++) {
    d.cells[i] = cells[i];
    d.extensions[i] = ext[i];
  }
  d.embedding_factor = c.template get<int>("embedding.factor", 2);
  d.periodic = c.template get<bool>("randomField.periodic", false);
  d.variance = c.template get<double>("stochastic.variance");
  const std::string cov = c.template get<std::string>("stochastic.covariance");
  static const char* names[] = { "exponential", "gaussian", "spherical", "separableExponential",
                                 "matern", "matern32", "matern52", "dampedOscillation",
                                 "gammaExponential", "cauchy", "generalizedCauchy", "cubic",
                                 "whiteNoise", "custom-iso" };
  d.covariance = -1;
  for (int i = 0; i < 14; i++)
    if (cov == names[i])
      d.covariance = i;
  if (cov == "custom-aniso")
    d.covariance = PF_COV_CUSTOM;
  if (d.covariance < 0)
    throw std::runtime_error("covariance structure " + cov + " not known");
  const std::string an = c.template get<std::string>("stochastic.anisotropy", "none");
  if (an == "none")
    d.anisotropy = PF_ANISO_NONE;
  else if (an == "axiparallel")
    d.anisotropy = PF_ANISO_AXIPARALLEL;
  else if (an == "geometric")
    d.anisotropy = PF_ANISO_GEOMETRIC;
  else
    throw std::runtime_error(
      "stochastic.anisotropy must be \"none\", \"axiparallel\" or \"geometric\"");
  const auto cl = c.template get<std::vector<double>>("stochastic.corrLength");
  for (size_t i = 0; i < cl.size() && i < 9; i++)
    d.corr_length[i] = cl[i];
  d.matern_nu = c.template get<double>("stochastic.maternNu", 1.);
  d.exp_gamma = c.template get<double>("stochastic.expGamma", 1.);
  d.cauchy_alpha = c.template get<double>("stochastic.cauchyAlpha", 1.);
  d.cauchy_beta = c.template get<double>("stochastic.cauchyBeta", 1.);
  d.matrix_backend = matrix_backend;
  d.field_backend = field_backend;
  d.transposed = c.hasKey("fftw.transposed") ? (int)c.template get<bool>("fftw.transposed") : -1;
  d.approximate = c.template get<bool>("embedding.approximate", false);
  d.threshold = c.template get<double>("embedding.threshold", 1e-14);
  d.precision = std::is_same<RF, float>::value ? PF_FP32 : PF_FP64;
  d.cg_iterations = c.template get<int>("randomField.cgIterations", 100);
  d.nugget = 0.;
  d.r2c_check_all = 0;
  // embedding.periodization and the merge / fold / cofold taper
  // (matrix.hh:515-580, :652-1013)
  const std::string per = c.template get<std::string>("embedding.periodization", "classical");
  if (per == "classical")
    d.periodization = PF_PERIOD_CLASSICAL;
  else if (per == "merge")
    d.periodization = PF_PERIOD_MERGE;
  else if (per == "fold")
    d.periodization = PF_PERIOD_FOLD;
  else if (per == "cofold")
    d.periodization = PF_PERIOD_COFOLD;
  else
    throw std::runtime_error(
      "stochastic.periodization must be \"classical\", \"merge\", \"fold\", or \"cofold\"");
  const std::string sg = c.template get<std::string>("embedding.sigmoid", "smoothstep");
  if (sg == "smoothstep")
    d.sigmoid = PF_SIGMOID_SMOOTHSTEP;
  else if (sg == "smooth")
    d.sigmoid = PF_SIGMOID_SMOOTH;
  else
    throw std::runtime_error("embedding.sigmoid must be \"smooth\" or \"smoothstep\"");
  const std::string sc = c.template get<std::string>("stochastic.sigmoidCombine", "radialPlus");
  if (sc == "radialPlus")
    d.sigmoid_combine = PF_COMBINE_RADIAL_PLUS;
  else if (sc == "radial")
    d.sigmoid_combine = PF_COMBINE_RADIAL;
  else if (sc == "tensor")
    d.sigmoid_combine = PF_COMBINE_TENSOR;
  else
    throw std::runtime_error("sigmoidCombine must be tensor, radial or radialPlus");
  if (per == "cofold") {
    const int rec = (int)c.template get<double>("embedding.cofoldRecursions", 1.);
    d.merge_recursions = rec == 0 ? -1 : rec;
  } else {
    const int rec =
      c.template get<int>(per == "fold" ? "embedding.foldRecursions" : "embedding.mergeRecursions", 99);
    d.merge_recursions = rec == 99 ? 0 : (rec == 0 ? -1 : rec);
  }
  d.sigmoid_start = c.template get<double>("embedding.sigmoidStart", 1.);
  d.sigmoid_end = c.hasKey("embedding.sigmoidEnd") ? c.template get<double>("embedding.sigmoidEnd") : 0.;
  // pf_desc.fold_max_factor = 0 selects the default (1 for fold, 5 for
  // cofold), so an explicit 0 cannot be expressed and is rejected
  const char* mf = per == "cofold" ? "embedding.cofoldMaxFactor" : "embedding.foldMaxFactor";
  d.fold_max_factor = c.template get<double>(mf, per == "cofold" ? 5. : 1.);
  if (c.hasKey(mf) && d.fold_max_factor == 0.)
    throw std::runtime_error(std::string(mf) + " = 0 is not supported by this engine");
  return d;
}

// The field data (StochasticPart, stochastic.hh:54): n cells, x fastest;
// empty until the first generate, as in the reference.
template <typename RF>
struct StochasticPart
{
  std::vector<RF> dataVector;
  bool evalValid = false;
};

// CppRNGBackend (cpprngbackend.hh:57-100): libstdc++ default_random_engine
// + normal_distribution<RF>.
template <typename RF>
class CppRNGBackend
{
  std::default_random_engine generator;
  std::normal_distribution<RF> normalDist{ 0., 1. };

public:
  void seed(unsigned int s) { generator.seed(s); }
  RF sample() { return normalDist(generator); }
};

// Matrix<Traits, MatrixBackend, FieldBackend> (matrix.hh:55-462) over one
// engine plan. RF = double -> FP64 device arithmetic; RF = float -> FP32
// realizations (eigenvalues stay FP64). multiplyInverse always runs FP64
// (a second FP64 plan is created on first use for RF = float).
template <typename RF, int dim, typename MB, typename FB>
class Matrix
{
  pf_desc desc_;
  void* plan_ = nullptr;
  mutable void* plan64_ = nullptr;
  pf_sizes sizes_{};
  mutable std::vector<double> custom_; // host covariance values of a callable
  mutable std::vector<RF> spare_;      // DFT field: the second field of the last draw

public:
  using StochasticPartType = StochasticPart<RF>;

  explicit Matrix(const pf_desc& d)
    : desc_(d)
  {
    desc_.matrix_backend = MB::id;
    desc_.field_backend = FB::id;
    check(pf_plan_create(&desc_, &plan_));
    check(pf_plan_sizes(plan_, &sizes_));
  }
  Matrix(const Matrix&) = delete;
  Matrix& operator=(const Matrix&) = delete;
  ~Matrix()
  {
    if (plan_)
      pf_plan_destroy(plan_);
    if (plan64_)
      pf_plan_destroy(plan64_);
  }

  void* plan() const { return plan_; }
  const pf_sizes& sizes() const { return sizes_; }

  // Matrix::fillTransformedMatrix(std::string), matrix.hh:223-318
  pf_eig_stats fillTransformedMatrix() const
  {
    pf_eig_stats st{};
    check(pf_eigs(plan_, &st));
    return st;
  }

  // fillTransformedMatrix(Covariance&&): a user callable
  // RF(RF variance, const std::array<RF,dim>& x) evaluated on the host over
  // the embedding grid with the scaled-identity geometry (matrix.hh:261-264).
  template <typename Covariance>
  pf_eig_stats fillTransformedMatrix(Covariance&& covariance) const
  {
    if (desc_.periodization != PF_PERIOD_CLASSICAL)
      throw std::runtime_error("a custom covariance is evaluated with the classical embedding only");
    const int f = desc_.periodic ? 1 : desc_.embedding_factor;
    std::array<long, 3> N{ 1, 1, 1 }, cnt{ 1, 1, 1 };
    for (int i = 0; i < dim; i++)
      N[i] = (long)f * desc_.cells[i];
    const bool dct = MB::id == PF_MATRIX_DCT;
    for (int i = 0; i < dim; i++)
      cnt[i] = dct ? N[i] / 2 + 1 : N[i];
    if (MB::id == PF_MATRIX_R2C)
      cnt[0] = 2 * (N[0] / 2 + 1); // padded r2c rows (the DFT matrix: every entry)
    std::vector<double> vals((size_t)cnt[0] * cnt[1] * cnt[2]);
    const double inv = 1. / desc_.corr_length[0];
    std::array<RF, dim> x;
    for (long k = 0; k < cnt[2]; k++)
      for (long j = 0; j < cnt[1]; j++)
        for (long i = 0; i < cnt[0]; i++) {
          const long idx[3] = { i, j, k };
          for (int a = 0; a < dim; a++) {
            double h = desc_.extensions[a] / desc_.cells[a];
            double cc = idx[a] * h;
            if (cc > 0.5 * desc_.extensions[a] * f)
              cc -= desc_.extensions[a] * f;
            x[a] = (RF)(inv * cc);
          }
          vals[(size_t)i + cnt[0] * ((size_t)j + cnt[1] * k)] =
            covariance((RF)desc_.variance, x);
        }
    pf_eig_stats st{};
    check(pf_eigs_custom(plan_, vals.data(), &st));
    custom_ = std::move(vals);
    return st;
  }

  // Matrix::generateField, matrix.hh:330-421: the RNG's sample() sequence is
  // consumed in the reference's order.
  template <typename RNG>
  void generateField(RNG& rng, StochasticPartType& sp) const
  {
    if (FB::id == PF_FIELD_DFT) {
      // one draw gives two fields: the first call returns the real part and
      // keeps the imaginary part, the next call returns it without drawing
      // (spareField, matrix.hh:336, :411-420)
      if (!spare_.empty()) {
        sp.dataVector = std::move(spare_);
        spare_.clear();
        sp.evalValid = false;
        return;
      }
      std::vector<double> noise((size_t)sizes_.noise_per_field);
      for (auto& v : noise)
        v = (double)rng.sample();
      std::vector<RF> both(2 * (size_t)sizes_.domain_size);
      check(pf_generate(plan_, noise.data(), 0, 0, 2, PF_TRANSFORM_NONE, 0., both.data(),
                        PF_MEM_HOST));
      const auto mid = both.begin() + (std::ptrdiff_t)sizes_.domain_size;
      sp.dataVector.assign(both.begin(), mid);
      spare_.assign(mid, both.end());
      sp.evalValid = false;
      return;
    }
    std::vector<double> noise((size_t)sizes_.noise_per_field);
    for (auto& v : noise)
      v = (double)rng.sample();
    std::vector<RF> field((size_t)sizes_.domain_size);
    check(pf_generate(plan_, noise.data(), 0, 0, 1, PF_TRANSFORM_NONE, 0., field.data(),
                      PF_MEM_HOST));
    if (FB::id == PF_FIELD_DCTDST && sp.dataVector.size() == field.size()) {
      // the reference's DCT/DST extraction adds onto an existing field
      // (dctdstfieldbackend.hh:923)
      for (size_t i = 0; i < field.size(); i++)
        sp.dataVector[i] += field[i];
    } else
      sp.dataVector = std::move(field);
    sp.evalValid = false;
  }

  // Throughput path: `batch` realizations from the device Philox stream.
  void generateBatch(uint64_t seed, uint64_t first, int batch, int transform, double param,
                     RF* out, int mem) const
  {
    check(pf_generate(plan_, nullptr, seed, first, batch, transform, param, out, mem));
  }

  // Matrix::generateUncorrelatedField, matrix.hh:433-445
  void generateUncorrelatedField(unsigned int seed, StochasticPartType& sp) const
  {
    std::default_random_engine generator(seed);
    std::normal_distribution<RF> normalDist(0., 1.);
    sp.dataVector.resize((size_t)sizes_.domain_size);
    for (auto& v : sp.dataVector)
      v = normalDist(generator);
    sp.evalValid = false;
  }

  // Matrix::setVarianceAsField, matrix.hh:456-462
  void setVarianceAsField(StochasticPartType& sp) const
  {
    sp.dataVector.assign((size_t)sizes_.domain_size, (RF)desc_.variance);
    sp.evalValid = false;
  }

  // Matrix::operator*, matrix.hh:142-151 (C x / N_ext)
  StochasticPartType operator*(const StochasticPartType& in) const { return apply(in, PF_OP_MULTIPLY); }
  // Matrix::multiplyRoot, matrix.hh:167-176
  StochasticPartType multiplyRoot(const StochasticPartType& in) const { return apply(in, PF_OP_ROOT); }

  // Matrix::multiplyInverse, matrix.hh:189-206 (zero check, embedding-inverse
  // initial guess, preconditioned innerCG)
  StochasticPartType multiplyInverse(const StochasticPartType& in) const
  {
    void* p = plan_;
    if (desc_.precision != PF_FP64) {
      if (!plan64_) {
        pf_desc d = desc_;
        d.precision = PF_FP64;
        check(pf_plan_create(&d, &plan64_));
        if (!custom_.empty()) {
          pf_eig_stats st{};
          check(pf_eigs_custom(plan64_, custom_.data(), &st));
        }
      }
      p = plan64_;
    }
    std::vector<double> b(in.dataVector.begin(), in.dataVector.end()), x(b.size());
    pf_cg_opts o{};
    o.mode = PF_CG_REFERENCE;
    pf_cg_stats st{};
    check(pf_cg(p, b.data(), x.data(), PF_MEM_HOST, &o, &st));
    StochasticPartType out;
    out.dataVector.assign(x.begin(), x.end());
    return out;
  }

private:
  StochasticPartType apply(const StochasticPartType& in, int op) const
  {
    StochasticPartType out;
    out.dataVector.resize(in.dataVector.size());
    check(pf_matvec(plan_, op, in.dataVector.data(), out.dataVector.data(), 1, PF_MEM_HOST));
    return out;
  }
};

// ValueTransform, mutators.hh:126-163
template <typename RF>
class ValueTransform
{
  int kind_;
  RF lambda_;

public:
  explicit ValueTransform(const ParameterTree& c)
  {
    const std::string t = c.get<std::string>("randomField.transform", "none");
    lambda_ = 0;
    if (t == "none")
      kind_ = PF_TRANSFORM_NONE;
    else if (t == "logNormal")
      kind_ = PF_TRANSFORM_LOGNORMAL;
    else if (t == "foldedNormal")
      kind_ = PF_TRANSFORM_FOLDEDNORMAL;
    else if (t == "sign")
      kind_ = PF_TRANSFORM_SIGN;
    else if (t == "boxCox") {
      kind_ = PF_TRANSFORM_BOXCOX;
      lambda_ = c.get<RF>("transform.boxCoxLambda");
    } else
      throw std::runtime_error{ "transform type not known" };
  }
  int kind() const { return kind_; }
  RF param() const { return lambda_; }
  void apply(RF& v) const
  {
    switch (kind_) {
      case PF_TRANSFORM_LOGNORMAL:
        v = std::exp(v);
        break;
      case PF_TRANSFORM_FOLDEDNORMAL:
        v = std::abs(v);
        break;
      case PF_TRANSFORM_SIGN:
        v = v < 0. ? RF(-1) : RF(1);
        break;
      case PF_TRANSFORM_BOXCOX:
        v = std::pow(v * lambda_ + 1, -lambda_);
        break;
      default:
        break;
    }
  }
};

// Refinement of cell data (StochasticPart::refine / coarsen,
// stochastic.hh:597-760): refine copies every cell into its 2^d children,
// coarsen averages the 2^d children of every coarse cell. `cells` is the
// grid after the operation. An empty vector stays empty.
template <typename RF, int dim>
std::vector<RF> refine_cells(const std::vector<RF>& in, const std::array<int, dim>& cells)
{
  if (in.empty())
    return in;
  std::array<size_t, 3> n{ 1, 1, 1 };
  for (int i = 0; i < dim; i++)
    n[i] = (size_t)cells[i];
  std::vector<RF> out(n[0] * n[1] * n[2]);
  for (size_t z = 0; z < n[2]; z++)
    for (size_t y = 0; y < n[1]; y++)
      for (size_t x = 0; x < n[0]; x++) {
        const size_t ox = x / 2, oy = dim > 1 ? y / 2 : 0, oz = dim > 2 ? z / 2 : 0;
        out[x + n[0] * (y + n[1] * z)] = in.at(ox + (n[0] / 2) * (oy + (n[1] / (dim > 1 ? 2 : 1)) * oz));
      }
  return out;
}

template <typename RF, int dim>
std::vector<RF> coarsen_cells(const std::vector<RF>& in, const std::array<int, dim>& cells)
{
  if (in.empty())
    return in;
  std::array<size_t, 3> n{ 1, 1, 1 }, f{ 1, 1, 1 };
  for (int i = 0; i < dim; i++) {
    n[i] = (size_t)cells[i];
    f[i] = 2 * n[i];
  }
  std::vector<RF> out(n[0] * n[1] * n[2]);
  for (size_t z = 0; z < n[2]; z++)
    for (size_t y = 0; y < n[1]; y++)
      for (size_t x = 0; x < n[0]; x++) {
        RF sum = 0.;
        for (size_t c = 0; c < ((size_t)1 << dim); c++) {
          const size_t fx = 2 * x + (c & 1), fy = dim > 1 ? 2 * y + ((c >> 1) & 1) : 0,
                       fz = dim > 2 ? 2 * z + ((c >> 2) & 1) : 0;
          sum += in.at(fx + f[0] * (fy + f[1] * fz));
        }
        out[x + n[0] * (y + n[1] * z)] = sum / (RF)((size_t)1 << dim);
      }
  return out;
}

// RandomField (randomfield.hh:42-1389), the hot-path surface: generation,
// products with the covariance matrix, its root and its inverse (with the
// optional cached inverse products of randomField.cacheInvMatvec and
// randomField.cacheInvRootMatvec), evaluation with the value transform,
// refinement, arithmetic and the on-disk format. Iso/aniso dispatch as in
// randomfield.hh:153-164.
template <typename GT, typename IsoMB = typename DefaultBackends<GT::dim>::IsoMB,
          typename IsoFB = typename DefaultBackends<GT::dim>::IsoFB,
          typename AnisoMB = typename DefaultBackends<GT::dim>::AnisoMB,
          typename AnisoFB = typename DefaultBackends<GT::dim>::AnisoFB>
class RandomField
{
public:
  using RF = typename GT::RangeField;
  enum
  {
    dim = GT::dim
  };
  using IsoMatrix = Matrix<RF, dim, IsoMB, IsoFB>;
  using AnisoMatrix = Matrix<RF, dim, AnisoMB, AnisoFB>;

private:
  ParameterTree config_;
  ValueTransform<RF> transform_;
  std::shared_ptr<IsoMatrix> iso_;
  std::shared_ptr<AnisoMatrix> aniso_;
  StochasticPart<RF> sp_;
  std::array<int, dim> cells_;
  std::array<double, dim> ext_;
  TrendPart<RF, dim> trend_;
  // cached products (randomfield.hh:100-107, :220-226): invMatvec_ holds
  // C^{-1} sp_ and invRootMatvec_ C^{-1/2} sp_ when their flags are set
  bool cacheInvMatvec_ = false, cacheInvRootMatvec_ = false;
  StochasticPart<RF> invMatvec_, invRootMatvec_;
  bool invMatvecValid_ = false, invRootMatvecValid_ = false;

  static ParameterTree read_field_file(const std::string& fileName)
  {
    ParameterTree c;
    c.readINIFile(fileName + ".field");
    return c;
  }

  void make_matrix()
  {
    const std::string an = config_.get<std::string>("stochastic.anisotropy", "none");
    const std::string cov = config_.get<std::string>("stochastic.covariance");
    auto with_cells = [&](pf_desc d) {
      for (int i = 0; i < dim; i++)
        d.cells[i] = cells_[i];
      return d;
    };
    iso_.reset();
    aniso_.reset();
    if (cov != "custom-aniso" && (an == "none" || an == "axiparallel"))
      iso_ = std::make_shared<IsoMatrix>(with_cells(make_desc<RF, dim>(config_, IsoMB::id, IsoFB::id)));
    else
      aniso_ = std::make_shared<AnisoMatrix>(
        with_cells(make_desc<RF, dim>(config_, AnisoMB::id, AnisoFB::id)));
  }
  StochasticPart<RF> times(const StochasticPart<RF>& x) const { return aniso_ ? (*aniso_) * x : (*iso_) * x; }
  StochasticPart<RF> root(const StochasticPart<RF>& x) const
  {
    return aniso_ ? aniso_->multiplyRoot(x) : iso_->multiplyRoot(x);
  }
  StochasticPart<RF> inverse(const StochasticPart<RF>& x) const
  {
    return aniso_ ? aniso_->multiplyInverse(x) : iso_->multiplyInverse(x);
  }
  static void scale(StochasticPart<RF>& x, RF alpha)
  {
    for (auto& v : x.dataVector)
      v *= alpha;
  }
  static void axpy_part(StochasticPart<RF>& x, const StochasticPart<RF>& o, RF alpha)
  {
    if (o.dataVector.empty())
      return;
    if (x.dataVector.empty())
      x.dataVector.assign(o.dataVector.size(), RF(0));
    for (size_t i = 0; i < x.dataVector.size(); i++)
      x.dataVector[i] += o.dataVector.at(i) * alpha;
    x.evalValid = false;
  }
  void read_config()
  {
    cells_ = config_.get<std::array<int, dim>>("grid.cells");
    ext_ = config_.get<std::array<double, dim>>("grid.extensions");
    cacheInvMatvec_ = config_.get<bool>("randomField.cacheInvMatvec", false);
    cacheInvRootMatvec_ = config_.get<bool>("randomField.cacheInvRootMatvec", false);
  }

public:
  explicit RandomField(const ParameterTree& config)
    : config_(config)
    , transform_(config)
  {
    read_config();
    trend_ = TrendPart<RF, dim>(config_, ext_);
    make_matrix();
  }

  // randomfield.hh:192-227: the configuration from <fileName>.field, the
  // trend coefficients from <fileName>.trend and the field values from
  // <fileName>.stoch.h5 (dataset /stochastic); the caches start from the
  // field read in, not valid
  explicit RandomField(const std::string& fileName)
    : config_(read_field_file(fileName))
    , transform_(config_)
  {
    read_config();
    trend_ = TrendPart<RF, dim>(config_, ext_, fileName);
    std::array<unsigned int, dim> c;
    for (int i = 0; i < dim; i++)
      c[i] = (unsigned int)cells_[i];
    sp_.dataVector = readStochasticHDF5<RF, dim>(c, "/stochastic", fileName + ".stoch.h5");
    make_matrix();
    invMatvec_ = sp_;
    invRootMatvec_ = sp_;
  }

  // randomfield.hh:706-713 with StochasticPart::writeToFile
  // (stochastic.hh:273-293) and TrendPart::writeToFile (trend.hh:1515-1527)
  void writeToFile(const std::string& fileName) const
  {
    std::array<unsigned int, dim> c;
    std::array<RF, dim> e;
    for (int i = 0; i < dim; i++) {
      c[i] = (unsigned int)cells_[i];
      e[i] = (RF)ext_[i];
    }
    std::vector<RF> data = sp_.dataVector;
    if (data.empty())
      data.assign((size_t)domain_size(), RF(0));
    writeStochasticHDF5<RF, dim>(c, data, "/stochastic", fileName + ".stoch.h5");
    writeToXDMF<RF, dim>(c, e, fileName);
    {
      std::ofstream t(fileName + ".trend", std::ofstream::trunc);
      trend_.writeToFile(t);
    }
    std::ofstream f(fileName + ".field", std::ofstream::trunc);
    ParameterTree out = config_;
    std::ostringstream cs;
    for (int i = 0; i < dim; i++)
      cs << (i ? " " : "") << cells_[i];
    out["grid.cells"] = cs.str(); // the current resolution
    out.report(f);
  }

  // randomfield.hh:491-499: time-based seed
  void generate(bool allowNonWorldComm = false)
  {
    unsigned int seed = (unsigned int)clock();
    seed += static_cast<int>(reinterpret_cast<uintptr_t>(&sp_));
    generate(seed, allowNonWorldComm);
  }

  // randomfield.hh:512-529 (rank 0: seed + 0)
  void generate(unsigned int seed, bool allowNonWorldComm = false)
  {
    CppRNGBackend<RF> rng;
    rng.seed(seed);
    generateWithRNG(rng, allowNonWorldComm);
  }

  // randomfield.hh:541-558: the field, then the trend coefficients from the
  // same RNG; the cached products become invalid
  template <typename RNG>
  void generateWithRNG(RNG& rng, bool = false)
  {
    if (aniso_)
      aniso_->generateField(rng, sp_);
    else
      iso_->generateField(rng, sp_);
    trend_.generate(rng);
    invMatvecValid_ = false;
    invRootMatvecValid_ = false;
  }

  // randomfield.hh:589-603
  void generateUncorrelated(unsigned int seed, bool = false)
  {
    if (aniso_)
      aniso_->generateUncorrelatedField(seed, sp_);
    else
      iso_->generateUncorrelatedField(seed, sp_);
    trend_.generateUncorrelated(seed);
    invMatvecValid_ = false;
    invRootMatvecValid_ = false;
  }

  template <typename Covariance>
  void fillMatrix(Covariance&& cov)
  {
    if (aniso_)
      aniso_->fillTransformedMatrix(std::forward<Covariance>(cov));
    else
      iso_->fillTransformedMatrix(std::forward<Covariance>(cov));
  }

  // randomfield.hh:1181-1205: caches the input as C^{-1} (result) and its
  // root product as C^{-1/2} (result)
  void timesMatrix()
  {
    if (cacheInvMatvec_) {
      invMatvec_ = sp_;
      invMatvecValid_ = true;
    }
    if (cacheInvRootMatvec_) {
      invRootMatvec_ = root(sp_);
      invRootMatvecValid_ = true;
    }
    sp_ = times(sp_);
    trend_.timesMatrix();
  }

  // randomfield.hh:1214-1243: a valid cached C^{-1} x replaces the solve
  void timesInverseMatrix()
  {
    if (cacheInvMatvec_ && invMatvecValid_) {
      if (cacheInvRootMatvec_) {
        invRootMatvec_ = root(invMatvec_);
        invRootMatvecValid_ = true;
      }
      sp_ = invMatvec_;
      invMatvecValid_ = false;
    } else {
      sp_ = inverse(sp_);
      if (cacheInvMatvec_)
        invMatvecValid_ = false;
      if (cacheInvRootMatvec_)
        invRootMatvecValid_ = false;
    }
    trend_.timesInverseMatrix();
  }

  // randomfield.hh:1255-1272
  void timesMatrixRoot()
  {
    if (cacheInvMatvec_ && cacheInvRootMatvec_) {
      invMatvec_ = invRootMatvec_;
      invMatvecValid_ = invRootMatvecValid_;
    }
    if (cacheInvRootMatvec_) {
      invRootMatvec_ = sp_;
      invRootMatvecValid_ = true;
    }
    sp_ = root(sp_);
    trend_.timesMatrixRoot();
  }

  // randomfield.hh:1282-1310: C^{-1/2} x = root(inverse(x)), or the cached
  // C^{-1/2} x when valid
  void timesInvMatRoot()
  {
    if (cacheInvRootMatvec_ && invRootMatvecValid_) {
      sp_ = invRootMatvec_;
      invRootMatvecValid_ = false;
      if (cacheInvMatvec_) {
        invRootMatvec_ = invMatvec_;
        invRootMatvecValid_ = invMatvecValid_;
        invMatvecValid_ = false;
      }
    } else {
      sp_ = inverse(sp_);
      if (cacheInvRootMatvec_) {
        invRootMatvec_ = sp_;
        invRootMatvecValid_ = true;
      }
      sp_ = root(sp_);
      if (cacheInvMatvec_)
        invMatvecValid_ = false;
    }
    trend_.timesInvMatRoot();
  }

  // randomfield.hh:854-869
  void zero()
  {
    trend_.zero();
    for (auto& v : sp_.dataVector)
      v = RF(0);
    if (cacheInvMatvec_) {
      for (auto& v : invMatvec_.dataVector)
        v = RF(0);
      invMatvecValid_ = true;
    }
    if (cacheInvRootMatvec_) {
      for (auto& v : invRootMatvec_.dataVector)
        v = RF(0);
      invRootMatvecValid_ = true;
    }
  }

  // randomfield.hh:880-886: the grid doubles its cells in every axis and the
  // matrix is rebuilt; refine() then moves the field to the new grid
  void refineMatrix()
  {
    for (auto& c : cells_)
      c *= 2;
    make_matrix();
  }
  void coarsenMatrix()
  {
    for (auto& c : cells_) {
      if (c % 2)
        throw std::runtime_error("coarsening needs an even number of cells");
      c /= 2;
    }
    make_matrix();
  }

  // randomfield.hh:896-940: with a valid cached C^{-1} x the refined field is
  // C (refined cache) 2^d (smoother than copying cells), likewise from a valid
  // C^{-1/2} x through the root; otherwise cells are copied
  void refine() { regrid(true); }
  // randomfield.hh:959-1004: the same with averaging of cells, scale 2^-d
  void coarsen() { regrid(false); }

  // arithmetic (randomfield.hh:1023-1150): the caches follow when both
  // operands hold them, else they become invalid
  RandomField& operator+=(const RandomField& o) { return axpy(o, RF(1)); }
  RandomField& operator-=(const RandomField& o) { return axpy(o, RF(-1)); }
  RandomField& operator*=(RF alpha)
  {
    trend_ *= alpha;
    scale(sp_, alpha);
    if (cacheInvMatvec_)
      scale(invMatvec_, alpha);
    if (cacheInvRootMatvec_)
      scale(invRootMatvec_, alpha);
    return *this;
  }
  RandomField& axpy(const RandomField& o, RF alpha)
  {
    trend_.axpy(o.trend_, alpha);
    axpy_part(sp_, o.sp_, alpha);
    if (cacheInvMatvec_) {
      if (o.cacheInvMatvec_) {
        axpy_part(invMatvec_, o.invMatvec_, alpha);
        invMatvecValid_ = invMatvecValid_ && o.invMatvecValid_;
      } else
        invMatvecValid_ = false;
    }
    if (cacheInvRootMatvec_) {
      if (o.cacheInvRootMatvec_) {
        axpy_part(invRootMatvec_, o.invRootMatvec_, alpha);
        invRootMatvecValid_ = invRootMatvecValid_ && o.invRootMatvecValid_;
      } else
        invRootMatvecValid_ = false;
    }
    return *this;
  }
  RandomField& axpy(RF alpha, const RandomField& o) { return axpy(o, alpha); }

  // scalar product and norms (randomfield.hh:1163-1343)
  RF operator*(const RandomField& o) const
  {
    RF out = 0.;
    if (sp_.dataVector.size() == o.sp_.dataVector.size())
      for (size_t i = 0; i < sp_.dataVector.size(); i++)
        out += sp_.dataVector[i] * o.sp_.dataVector[i];
    return out + trend_ * o.trend_;
  }
  RF oneNorm() const
  {
    RF out = 0.;
    for (RF v : sp_.dataVector)
      out += std::abs(v);
    return out + trend_.oneNorm();
  }
  RF twoNorm() const { return std::sqrt(*this * *this); }
  RF infNorm() const
  {
    RF out = 0.;
    for (RF v : sp_.dataVector)
      out = std::max(out, std::abs(v));
    return std::max(out, trend_.infNorm());
  }

  // randomfield.hh:641-651: stochastic value (cell lookup) plus trend, then
  // the value transform
  void evaluate(const std::array<double, dim>& location, RF& output) const
  {
    size_t idx = 0, stride = 1;
    for (int i = 0; i < dim; i++) {
      long c = (long)std::floor(location[i] / (ext_[i] / cells_[i]));
      c = std::min<long>(std::max<long>(c, 0), cells_[i] - 1);
      idx += (size_t)c * stride;
      stride *= cells_[i];
    }
    output = sp_.dataVector.at(idx) + trend_.evaluate(location);
    transform_.apply(output);
  }

  // randomfield.hh:664-695: all cells, x fastest, each at its centre
  void bulkEvaluate(std::vector<RF>& output, std::array<unsigned int, dim>& sizes) const
  {
    for (int i = 0; i < dim; i++)
      sizes[i] = cells_[i];
    output = sp_.dataVector;
    const bool trend = trend_.size() > 0;
    std::array<double, dim> x{};
    for (size_t i = 0; i < output.size(); i++) {
      if (trend) {
        size_t r = i;
        for (int a = 0; a < dim; a++) {
          x[a] = ((double)(r % cells_[a]) + 0.5) * (ext_[a] / cells_[a]);
          r /= cells_[a];
        }
        output[i] += trend_.evaluate(x);
      }
      transform_.apply(output[i]);
    }
  }

  const TrendPart<RF, dim>& trendPart() const { return trend_; }
  TrendPart<RF, dim>& trendPart() { return trend_; }

  const std::vector<RF>& dataVector() const { return sp_.dataVector; }
  std::vector<RF>& dataVector() { return sp_.dataVector; }
  const std::vector<RF>& invMatvecData() const { return invMatvec_.dataVector; }
  const std::vector<RF>& invRootMatvecData() const { return invRootMatvec_.dataVector; }
  bool invMatvecValid() const { return invMatvecValid_; }
  bool invRootMatvecValid() const { return invRootMatvecValid_; }
  const std::array<int, dim>& cells() const { return cells_; }
  bool usesAnisoMatrix() const { return (bool)aniso_; }
  void* plan() const { return aniso_ ? aniso_->plan() : iso_->plan(); }

private:
  int64_t domain_size() const
  {
    int64_t n = 1;
    for (int i = 0; i < dim; i++)
      n *= cells_[i];
    return n;
  }

  // the field's data is on the grid before the last refineMatrix /
  // coarsenMatrix; `up` moves it to the current cells
  void regrid(bool up)
  {
    const size_t want = (size_t)domain_size();
    auto move = [&](StochasticPart<RF>& x) {
      if (!x.dataVector.empty() && x.dataVector.size() != want)
        x.dataVector = up ? refine_cells<RF, dim>(x.dataVector, cells_)
                          : coarsen_cells<RF, dim>(x.dataVector, cells_);
    };
    const RF s = (RF)std::pow(2., up ? (double)dim : -(double)dim);
    if (cacheInvMatvec_ && invMatvecValid_) {
      move(invMatvec_);
      sp_ = times(invMatvec_);
      scale(sp_, s);
      scale(invMatvec_, s);
      if (cacheInvRootMatvec_) {
        invRootMatvec_ = root(invMatvec_);
        scale(invRootMatvec_, s);
        invRootMatvecValid_ = true;
      }
    } else if (cacheInvRootMatvec_ && invRootMatvecValid_) {
      move(invRootMatvec_);
      sp_ = root(invRootMatvec_);
      scale(sp_, s);
      scale(invRootMatvec_, s);
      if (cacheInvMatvec_) {
        invMatvec_ = sp_;
        invMatvecValid_ = false;
      }
    } else {
      move(sp_);
      if (cacheInvMatvec_)
        move(invMatvec_);
      if (cacheInvRootMatvec_)
        move(invRootMatvec_);
    }
  }
};

} // namespace parafields_amd
