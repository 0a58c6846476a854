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
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../parafields_amd.h"
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
    if (!dct)
      cnt[0] = 2 * (N[0] / 2 + 1); // padded r2c rows
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

// RandomField (randomfield.hh:42-1389), the hot-path surface: generation,
// products with the covariance matrix, its root and its inverse, and
// evaluation with the value transform. Iso/aniso dispatch as in
// randomfield.hh:153-164.
template <typename GT, typename IsoMB = DCTMatrixBackend, typename IsoFB = R2CFieldBackend,
          typename AnisoMB = R2CMatrixBackend, typename AnisoFB = R2CFieldBackend>
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

public:
  explicit RandomField(const ParameterTree& config)
    : config_(config)
    , transform_(config)
  {
    cells_ = config.get<std::array<int, dim>>("grid.cells");
    ext_ = config.get<std::array<double, dim>>("grid.extensions");
    trend_ = TrendPart<RF, dim>(config, ext_);
    const std::string an = config.get<std::string>("stochastic.anisotropy", "none");
    const std::string cov = config.get<std::string>("stochastic.covariance");
    if (cov != "custom-aniso" && (an == "none" || an == "axiparallel"))
      iso_ = std::make_shared<IsoMatrix>(make_desc<RF, dim>(config, IsoMB::id, IsoFB::id));
    else
      aniso_ = std::make_shared<AnisoMatrix>(make_desc<RF, dim>(config, AnisoMB::id, AnisoFB::id));
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
  // same RNG
  template <typename RNG>
  void generateWithRNG(RNG& rng, bool = false)
  {
    if (aniso_)
      aniso_->generateField(rng, sp_);
    else
      iso_->generateField(rng, sp_);
    trend_.generate(rng);
  }

  // randomfield.hh:589-603
  void generateUncorrelated(unsigned int seed, bool = false)
  {
    if (aniso_)
      aniso_->generateUncorrelatedField(seed, sp_);
    else
      iso_->generateUncorrelatedField(seed, sp_);
    trend_.generateUncorrelated(seed);
  }

  template <typename Covariance>
  void fillMatrix(Covariance&& cov)
  {
    if (aniso_)
      aniso_->fillTransformedMatrix(std::forward<Covariance>(cov));
    else
      iso_->fillTransformedMatrix(std::forward<Covariance>(cov));
  }

  // randomfield.hh:1181-1309: the stochastic part through the matrix, the
  // trend coefficients scaled by their variances
  void timesMatrix()
  {
    sp_ = aniso_ ? (*aniso_) * sp_ : (*iso_) * sp_;
    trend_.timesMatrix();
  }
  void timesInverseMatrix()
  {
    sp_ = aniso_ ? aniso_->multiplyInverse(sp_) : iso_->multiplyInverse(sp_);
    trend_.timesInverseMatrix();
  }
  void timesMatrixRoot()
  {
    sp_ = aniso_ ? aniso_->multiplyRoot(sp_) : iso_->multiplyRoot(sp_);
    trend_.timesMatrixRoot();
  }
  void timesInvMatRoot()
  {
    sp_ = aniso_ ? aniso_->multiplyInverse(sp_) : iso_->multiplyInverse(sp_);
    sp_ = aniso_ ? aniso_->multiplyRoot(sp_) : iso_->multiplyRoot(sp_);
    trend_.timesInvMatRoot();
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

  const std::vector<RF>& dataVector() const { return sp_.dataVector; }
  std::vector<RF>& dataVector() { return sp_.dataVector; }
  bool usesAnisoMatrix() const { return (bool)aniso_; }
  void* plan() const { return aniso_ ? aniso_->plan() : iso_->plan(); }
};

} // namespace parafields_amd
