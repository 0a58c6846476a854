// gpumatrix.hh -- the reference-side drop-in: a `Matrix` for
// parafields::RandomField (include/parafields/randomfield.hh:42-47) whose
// arithmetic runs on the MI355X engine through the C ABI of parafields_amd.h.
//
// In a parafields-core build, include this header after the reference's own
// headers (it names parafields::StochasticPart) and pass the aliases at the
// end of this file as the IsoMatrix / AnisoMatrix template arguments:
//
//   #include <parafields/randomfield.hh>
//   #include <parafields_amd/gpumatrix.hh>
//   using Field = parafields::RandomField<GridTraits, parafields_amd::GpuIsoMatrix,
//                                         parafields_amd::GpuAnisoMatrix>;
//
// GpuMatrix<Traits, MB, FB> provides the surface the reference's Matrix
// (matrix.hh:55-462) offers its callers: the constructor from
// std::shared_ptr<Traits> (:98), update (:121), operator* (:142),
// multiplyRoot (:167), multiplyInverse (:189), fillTransformedMatrix (:223),
// generateField (:330), generateUncorrelatedField (:433), setVarianceAsField
// (:456) and the MatrixBackendType / FieldBackendType / StochasticPartType
// typedefs (:63-66) that RandomFieldTraits befriends and compares
// (fieldtraits.hh:102-107, :245-257). Like the reference's Matrix it reads
// the traits' private members (it is the traits' IsoMatrixType /
// AnisoMatrixType, their friend) and writes StochasticPart::dataVector /
// evalValid (stochastic.hh:37-38).
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../parafields_amd.h"
#include "randomfield.hh"

namespace parafields {
template <typename Traits>
class StochasticPart;
}

namespace parafields_amd {

// Backend tags standing in for the reference's backend classes in the
// Matrix typedefs (matrix.hh:63-66). They are distinct class types, so the
// reference's `friend typename ...::MatrixBackendType;` declarations and its
// std::is_same checks against R2CMatrixBackend<Traits> / R2CFieldBackend<Traits>
// (fieldtraits.hh:245-257) compile and see "not the reference's R2C".
template <typename Traits>
struct GpuDCTMatrixBackend
{
  static constexpr int id = PF_MATRIX_DCT;
};
template <typename Traits>
struct GpuR2CMatrixBackend
{
  static constexpr int id = PF_MATRIX_R2C;
};
template <typename Traits>
struct GpuR2CFieldBackend
{
  static constexpr int id = PF_FIELD_R2C;
};
template <typename Traits>
struct GpuDCTDSTFieldBackend
{
  static constexpr int id = PF_FIELD_DCTDST;
};
template <typename Traits>
struct GpuDFTMatrixBackend
{
  static constexpr int id = PF_MATRIX_DFT;
};
template <typename Traits>
struct GpuDFTFieldBackend
{
  static constexpr int id = PF_FIELD_DFT;
};
// The defaults of DefaultIsoMatrix / DefaultAnisoMatrix, which depend on the
// dimension (matrix.hh:1660-1736): DCT (iso) or R2C (aniso) matrix with the
// R2C field for dim > 1, the DFT matrix and field for 1-D.
template <typename Traits>
struct GpuIsoMatrixBackend
{
  static constexpr int id = Traits::dim == 1 ? PF_MATRIX_DFT : PF_MATRIX_DCT;
};
template <typename Traits>
struct GpuAnisoMatrixBackend
{
  static constexpr int id = Traits::dim == 1 ? PF_MATRIX_DFT : PF_MATRIX_R2C;
};
template <typename Traits>
struct GpuDefaultFieldBackend
{
  static constexpr int id = Traits::dim == 1 ? PF_FIELD_DFT : PF_FIELD_R2C;
};

template <typename Traits, template <typename> class MB, template <typename> class FB>
class GpuMatrix
{
public:
  using MatrixBackendType = MB<Traits>;
  using FieldBackendType = FB<Traits>;
  using StochasticPartType = parafields::StochasticPart<Traits>;

private:
  using RF = typename Traits::RF;
  enum
  {
    dim = Traits::dim
  };
  static constexpr bool dctdst = FieldBackendType::id == PF_FIELD_DCTDST;

  const std::shared_ptr<Traits> traits;
  pf_desc desc_{};
  void* plan_ = nullptr;
  mutable void* plan64_ = nullptr; // FP64 plan of multiplyInverse for RF = float
  pf_sizes sizes_{};
  mutable bool filled_ = false;
  mutable std::vector<double> custom_; // host covariance values of a callable
  // DFT field: the second field of the last draw (spareField, matrix.hh:90);
  // like the reference's, it survives update()
  mutable std::vector<RF> spare_;

  void release()
  {
    if (plan_)
      pf_plan_destroy(plan_);
    if (plan64_)
      pf_plan_destroy(plan64_);
    plan_ = plan64_ = nullptr;
  }

  // The eigenvalues are computed lazily on first use, as the reference's
  // matrixBackend.valid() check does (matrix.hh:333-334).
  void ensure_filled() const
  {
    if (!filled_) {
      pf_eig_stats st{};
      check(pf_eigs(plan_, &st));
      filled_ = true;
    }
  }

  std::vector<RF>& data(StochasticPartType& sp) const { return sp.dataVector; }
  const std::vector<RF>& data(const StochasticPartType& sp) const { return sp.dataVector; }

  // output = input (+ for the DCT/DST field: the reference's extraction adds
  // onto the output, which operator* and multiplyRoot pass as the input
  // itself, matrix.hh:144-146 with dctdstfieldbackend.hh:923) + op(input)
  StochasticPartType apply(const StochasticPartType& input, int op) const
  {
    ensure_filled();
    StochasticPartType output(input);
    const std::vector<RF>& x = data(input);
    std::vector<RF> y(x.size());
    check(pf_matvec(plan_, op, x.data(), y.data(), 1, PF_MEM_HOST));
    std::vector<RF>& o = data(output);
    for (size_t i = 0; i < y.size(); i++)
      o[i] = dctdst ? o[i] + y[i] : y[i];
    output.evalValid = false;
    return output;
  }

public:
  explicit GpuMatrix(const std::shared_ptr<Traits>& traits_)
    : traits(traits_)
  {
    update();
  }
  GpuMatrix(const GpuMatrix&) = delete;
  GpuMatrix& operator=(const GpuMatrix&) = delete;
  ~GpuMatrix() { release(); }

  // Matrix::update (matrix.hh:121-133): the traits may have been refined;
  // the plan is rebuilt from their current cells and extensions.
  void update()
  {
    release();
    desc_ = make_desc<RF, dim>((*traits).config, MatrixBackendType::id, FieldBackendType::id);
    for (int i = 0; i < dim; i++) {
      desc_.cells[i] = (int)(*traits).cells[i];
      desc_.extensions[i] = (double)(*traits).extensions[i];
    }
    desc_.embedding_factor = (int)(*traits).embeddingFactor;
    check(pf_plan_create(&desc_, &plan_));
    check(pf_plan_sizes(plan_, &sizes_));
    filled_ = false;
    custom_.clear(); // a custom covariance is filled again on the new grid
  }

  // Matrix::operator*, matrix.hh:142-151: C x / N_ext
  StochasticPartType operator*(const StochasticPartType& input) const
  {
    return apply(input, PF_OP_MULTIPLY);
  }

  // Matrix::multiplyRoot, matrix.hh:167-176
  StochasticPartType multiplyRoot(const StochasticPartType& input) const
  {
    return apply(input, PF_OP_ROOT);
  }

  // Matrix::multiplyInverse, matrix.hh:189-206: the input itself when every
  // entry is at most 1e-10 in magnitude, else the embedding inverse as the
  // initial guess of the preconditioned innerCG, FP64 on the device.
  StochasticPartType multiplyInverse(const StochasticPartType& input) const
  {
    ensure_filled();
    StochasticPartType output(input);
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
    const std::vector<RF>& in = data(input);
    std::vector<double> b(in.begin(), in.end()), x(b.size());
    pf_cg_opts o{};
    o.mode = PF_CG_REFERENCE;
    pf_cg_stats st{};
    check(pf_cg(p, b.data(), x.data(), PF_MEM_HOST, &o, &st));
    std::vector<RF>& out = data(output);
    for (size_t i = 0; i < x.size(); i++)
      out[i] = (RF)x[i];
    output.evalValid = false;
    return output;
  }

  // Matrix::fillTransformedMatrix, matrix.hh:223-318. A std::string names a
  // built-in covariance (it must be the configured stochastic.covariance,
  // which the plan evaluates on the device); any other argument is a callable
  // RF(RF variance, const std::array<RF, dim>& x) of the scaled coordinate,
  // evaluated on the host over the classical embedding grid (custom-iso
  // semantics, matrix.hh:261-264) and transformed on the device.
  template <typename Covariance>
  void fillTransformedMatrix(Covariance&& covariance) const
  {
    if constexpr (std::is_convertible<Covariance, std::string>::value) {
      const std::string name(covariance);
      if (name != (*traits).config.template get<std::string>("stochastic.covariance"))
        throw std::runtime_error("covariance structure " + name +
                                 " differs from the configured one");
      ensure_filled();
    } else {
      if (desc_.periodization != PF_PERIOD_CLASSICAL)
        throw std::runtime_error(
          "a custom covariance is evaluated with the classical embedding only");
      const int f = desc_.periodic ? 1 : desc_.embedding_factor;
      std::array<long, 3> N{ 1, 1, 1 }, cnt{ 1, 1, 1 };
      for (int i = 0; i < dim; i++)
        N[i] = (long)f * desc_.cells[i];
      const bool dct = MatrixBackendType::id == PF_MATRIX_DCT;
      const bool r2c = MatrixBackendType::id == PF_MATRIX_R2C;
      for (int i = 0; i < dim; i++)
        cnt[i] = dct ? N[i] / 2 + 1 : N[i];
      if (r2c)
        cnt[0] = 2 * (N[0] / 2 + 1); // padded r2c rows (the DFT matrix: every entry)
      std::vector<double> vals((size_t)cnt[0] * cnt[1] * cnt[2]);
      const double inv = 1. / desc_.corr_length[0];
      std::array<RF, dim> x;
      for (long k = 0; k < cnt[2]; k++)
        for (long j = 0; j < cnt[1]; j++)
          for (long i = 0; i < cnt[0]; i++) {
            const long idx[3] = { i, j, k };
            for (int a = 0; a < dim; a++) {
              const double h = desc_.extensions[a] / desc_.cells[a];
              double c = idx[a] * h;
              if (c > 0.5 * desc_.extensions[a] * f)
                c -= desc_.extensions[a] * f;
              x[a] = (RF)(inv * c);
            }
            vals[(size_t)i + cnt[0] * ((size_t)j + cnt[1] * k)] =
              (double)covariance((RF)desc_.variance, x);
          }
      pf_eig_stats st{};
      check(pf_eigs_custom(plan_, vals.data(), &st));
      custom_ = std::move(vals);
      filled_ = true;
    }
  }

  // Matrix::generateField, matrix.hh:330-421: the RNG's sample() sequence in
  // the reference's order (two samples per half-spectrum entry for the R2C
  // field, one per entry and parity type for the DCT/DST field).
  template <typename RNG>
  void generateField(RNG& rngBackend, StochasticPartType& stochasticPart) const
  {
    ensure_filled();
    if (FieldBackendType::id == PF_FIELD_DFT) {
      // one draw, two fields: the real part now, the imaginary part on the
      // next call without drawing (matrix.hh:336, :411-420)
      std::vector<RF>& out = data(stochasticPart);
      if (!spare_.empty()) {
        out = std::move(spare_);
        spare_.clear();
        stochasticPart.evalValid = false;
        return;
      }
      std::vector<double> noise((size_t)sizes_.noise_per_field);
      for (auto& v : noise)
        v = (double)rngBackend.sample();
      std::vector<RF> both(2 * (size_t)sizes_.domain_size);
      check(pf_generate(plan_, noise.data(), 0, 0, 2, PF_TRANSFORM_NONE, 0., both.data(),
                        PF_MEM_HOST));
      const auto mid = both.begin() + (std::ptrdiff_t)sizes_.domain_size;
      out.assign(both.begin(), mid);
      spare_.assign(mid, both.end());
      stochasticPart.evalValid = false;
      return;
    }
    std::vector<double> noise((size_t)sizes_.noise_per_field);
    for (auto& v : noise)
      v = (double)rngBackend.sample();
    std::vector<RF> field((size_t)sizes_.domain_size);
    check(pf_generate(plan_, noise.data(), 0, 0, 1, PF_TRANSFORM_NONE, 0., field.data(),
                      PF_MEM_HOST));
    std::vector<RF>& out = data(stochasticPart);
    if (dctdst) {
      // the DCT/DST extraction resizes and adds (dctdstfieldbackend.hh:640, :923)
      out.resize(field.size());
      for (size_t i = 0; i < field.size(); i++)
        out[i] += field[i];
    } else
      out = std::move(field);
    stochasticPart.evalValid = false;
  }

  // Matrix::generateUncorrelatedField, matrix.hh:433-445 (one process: rank 0)
  void generateUncorrelatedField(unsigned int seed, StochasticPartType& stochasticPart) const
  {
    std::default_random_engine generator(seed);
    std::normal_distribution<RF> normalDist(0., 1.);
    std::vector<RF>& out = data(stochasticPart);
    out.resize((size_t)sizes_.domain_size);
    for (auto& v : out)
      v = normalDist(generator);
    stochasticPart.evalValid = false;
  }

  // Matrix::setVarianceAsField, matrix.hh:456-462
  void setVarianceAsField(StochasticPartType& stochasticPart) const
  {
    std::vector<RF>& out = data(stochasticPart);
    out.assign((size_t)sizes_.domain_size, (RF)desc_.variance);
    stochasticPart.evalValid = false;
  }

  void* plan() const { return plan_; }
};

// The reference's default matrices (matrix.hh:1660-1736): DCT matrix with the
// R2C field for isotropic and axiparallel covariances, the R2C matrix for
// geometric anisotropy (the DFT matrix and field in 1-D), and the DCT/DST
// and DFT fields as drop-ins for either.
template <typename Traits>
using GpuIsoMatrix = GpuMatrix<Traits, GpuIsoMatrixBackend, GpuDefaultFieldBackend>;
template <typename Traits>
using GpuAnisoMatrix = GpuMatrix<Traits, GpuAnisoMatrixBackend, GpuDefaultFieldBackend>;
template <typename Traits>
using GpuDFTMatrix = GpuMatrix<Traits, GpuDFTMatrixBackend, GpuDFTFieldBackend>;
template <typename Traits>
using GpuDCTDSTMatrix = GpuMatrix<Traits, GpuDCTMatrixBackend, GpuDCTDSTFieldBackend>;

} // namespace parafields_amd
