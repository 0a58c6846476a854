// Drop-in check of include/parafields_amd/gpumatrix.hh: GpuMatrix plugged
// into stand-ins for the reference's RandomFieldTraits and StochasticPart
// with the same shape as the reference's classes (private members, the
// friend declarations of fieldtraits.hh:102-107 and stochastic.hh:37-38, the
// backend is_same checks of fieldtraits.hh:245-257). Compiling and linking
// it instantiates every member of the adapter.
//
//   gpumatrix_check compile       exits after construction-free checks (no GPU)
//   gpumatrix_check run <dir>     runs every method on the GPU and writes the
//                                 results as raw doubles into <dir>
#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "parafields_amd/parametertree.hh"

namespace parafields {

template <typename Traits>
class StochasticPart;
template <typename Traits>
struct R2CFieldBackend
{};
template <typename Traits>
struct R2CMatrixBackend
{};

template <typename DF, typename RF, unsigned int dimension>
struct GridTraits
{
  enum
  {
    dim = dimension
  };
  using RangeField = RF;
  using DomainField = DF;
};

// RandomFieldTraits (fieldtraits.hh:60-160): the members a Matrix reads
template <typename GT, template <typename> class IsoMatrix, template <typename> class AnisoMatrix>
class RandomFieldTraits
{
  using ThisType = RandomFieldTraits<GT, IsoMatrix, AnisoMatrix>;

public:
  enum
  {
    dim = GT::dim
  };
  using RF = typename GT::RangeField;
  using IsoMatrixType = IsoMatrix<ThisType>;
  using AnisoMatrixType = AnisoMatrix<ThisType>;
  using Index = unsigned int;
  using Indices = std::array<Index, dim>;

private:
  friend StochasticPart<ThisType>;
  friend IsoMatrix<ThisType>;
  friend AnisoMatrix<ThisType>;
  friend typename IsoMatrix<ThisType>::MatrixBackendType;
  friend typename IsoMatrix<ThisType>::FieldBackendType;
  friend typename AnisoMatrix<ThisType>::MatrixBackendType;
  friend typename AnisoMatrix<ThisType>::FieldBackendType;

  const parafields_amd::ParameterTree& config;
  const std::array<RF, dim> extensions;
  unsigned int embeddingFactor;
  Indices cells;
  Index domainSize = 1;

public:
  explicit RandomFieldTraits(const parafields_amd::ParameterTree& c)
    : config(c)
    , extensions(c.get<std::array<RF, dim>>("grid.extensions"))
    , embeddingFactor(c.get<unsigned int>("embedding.factor", 2))
    , cells(c.get<Indices>("grid.cells"))
  {
    for (int i = 0; i < dim; i++)
      domainSize *= cells[i];
  }
  // fieldtraits.hh:245-257: the dispatch asks whether an R2C backend is used
  static constexpr bool isoFieldIsReferenceR2C =
    std::is_same<typename IsoMatrix<ThisType>::FieldBackendType, R2CFieldBackend<ThisType>>::value;
  static constexpr bool anisoMatrixIsReferenceR2C =
    std::is_same<typename AnisoMatrix<ThisType>::MatrixBackendType,
                 R2CMatrixBackend<ThisType>>::value;
  void refine()
  {
    for (auto& c : cells)
      c *= 2;
    domainSize <<= dim;
  }
};

// StochasticPart (stochastic.hh:30-120): dataVector and evalValid are
// private, written by the matrices (friends)
template <typename Traits>
class StochasticPart
{
  using RF = typename Traits::RF;
  friend typename Traits::IsoMatrixType;
  friend typename Traits::AnisoMatrixType;
  std::shared_ptr<Traits> traits;
  unsigned int localDomainSize;
  mutable std::vector<RF> dataVector;
  mutable bool evalValid = false;

public:
  explicit StochasticPart(const std::shared_ptr<Traits>& t)
    : traits(t)
    , localDomainSize((*t).domainSize)
    , dataVector((*t).domainSize, RF(0))
  {}
  const std::vector<RF>& values() const { return dataVector; }
  void set(const std::vector<RF>& v) { dataVector = v; }
  bool valid() const { return evalValid; }
};

} // namespace parafields

#include "parafields_amd/gpumatrix.hh"

using namespace parafields_amd;

template <unsigned dim>
using Traits = parafields::RandomFieldTraits<parafields::GridTraits<double, double, dim>,
                                             GpuIsoMatrix, GpuAnisoMatrix>;
template <unsigned dim>
using TraitsDCTDST = parafields::RandomFieldTraits<parafields::GridTraits<double, double, dim>,
                                                   GpuDCTDSTMatrix, GpuAnisoMatrix>;
using Traits2f = parafields::RandomFieldTraits<parafields::GridTraits<double, float, 2>,
                                               GpuIsoMatrix, GpuAnisoMatrix>;

static_assert(!Traits<2>::isoFieldIsReferenceR2C, "tag types are not the reference's backends");
static_assert(Traits<1>::IsoMatrixType::MatrixBackendType::id == PF_MATRIX_DFT &&
                Traits<1>::IsoMatrixType::FieldBackendType::id == PF_FIELD_DFT &&
                Traits<2>::IsoMatrixType::MatrixBackendType::id == PF_MATRIX_DCT &&
                Traits<3>::AnisoMatrixType::MatrixBackendType::id == PF_MATRIX_R2C,
              "default backends per dimension (matrix.hh:1660-1736)");
static_assert(!Traits<3>::anisoMatrixIsReferenceR2C, "tag types are not the reference's backends");
static_assert(std::is_same<Traits<2>::IsoMatrixType::StochasticPartType,
                           parafields::StochasticPart<Traits<2>>>::value,
              "StochasticPartType is the reference's container");

static std::string outdir;

template <typename RF>
static void dump(const char* name, const std::vector<RF>& v)
{
  const std::string p = outdir + "/" + name + ".bin";
  FILE* f = std::fopen(p.c_str(), "wb");
  for (RF x : v) {
    const double d = (double)x;
    std::fwrite(&d, sizeof d, 1, f);
  }
  std::fclose(f);
}

// CppRNGBackend (cpprngbackend.hh:57-100)
template <typename RF>
struct Rng
{
  std::default_random_engine g;
  std::normal_distribution<RF> n{ 0., 1. };
  explicit Rng(unsigned s) { g.seed(s); }
  RF sample() { return n(g); }
};

static ParameterTree config(const char* cells, const char* ext, const char* cov,
                            const char* aniso, const char* corr)
{
  ParameterTree c;
  c["grid.cells"] = cells;
  c["grid.extensions"] = ext;
  c["stochastic.variance"] = "1.5";
  c["stochastic.covariance"] = cov;
  c["stochastic.anisotropy"] = aniso;
  c["stochastic.corrLength"] = corr;
  c["embedding.approximate"] = "true";
  return c;
}

int main(int argc, char** argv)
{
  if (argc < 2)
    return 2;
  if (std::strcmp(argv[1], "compile") == 0) {
    std::printf("gpumatrix_check: adapter instantiated\nall OK\n");
    return 0;
  }
  if (argc < 3)
    return 2;
  outdir = argv[2];
  try {
    // 2D isotropic (DCT matrix + R2C field)
    {
      ParameterTree c = config("40 24", "1 0.6", "exponential", "none", "0.1");
      auto t = std::make_shared<Traits<2>>(c);
      GpuIsoMatrix<Traits<2>> m(t);
      m.fillTransformedMatrix(std::string("exponential"));
      parafields::StochasticPart<Traits<2>> sp(t);
      Rng<double> rng(42);
      m.generateField(rng, sp);
      dump("iso_gen", sp.values());
      dump("iso_mult", (m * sp).values());
      dump("iso_root", m.multiplyRoot(sp).values());
      dump("iso_inv", m.multiplyInverse(sp).values());
      m.generateUncorrelatedField(7, sp);
      dump("iso_uncorr", sp.values());
      m.setVarianceAsField(sp);
      dump("iso_var", sp.values());
      // refinement: the traits double the cells, update rebuilds the plan
      t->refine();
      m.update();
      parafields::StochasticPart<Traits<2>> sp2(t);
      Rng<double> rng2(43);
      m.generateField(rng2, sp2);
      dump("iso_refined_gen", sp2.values());
    }
    // 3D geometric anisotropy (R2C matrix + R2C field)
    {
      ParameterTree c = config("12 10 8", "1 1 1", "exponential", "geometric",
                               "0.1 0.02 0 0.01 0.08 0 0 0 0.05");
      auto t = std::make_shared<Traits<3>>(c);
      GpuAnisoMatrix<Traits<3>> m(t);
      parafields::StochasticPart<Traits<3>> sp(t);
      Rng<double> rng(5);
      m.generateField(rng, sp);
      dump("aniso_gen", sp.values());
      dump("aniso_mult", (m * sp).values());
    }
    // DCT/DST field: generate adds onto the existing field; operator* and
    // multiplyRoot return input + product (the reference's accumulation)
    {
      ParameterTree c = config("16 12 10", "1 1 1", "exponential", "axiparallel", "0.2 0.1 0.15");
      auto t = std::make_shared<TraitsDCTDST<3>>(c);
      GpuDCTDSTMatrix<TraitsDCTDST<3>> m(t);
      parafields::StochasticPart<TraitsDCTDST<3>> sp(t);
      Rng<double> rng(9);
      m.generateField(rng, sp);
      dump("dct_gen1", sp.values());
      m.generateField(rng, sp);
      dump("dct_gen2", sp.values());
      dump("dct_mult", (m * sp).values());
      dump("dct_root", m.multiplyRoot(sp).values());
      dump("dct_inv", m.multiplyInverse(sp).values());
    }
    // 1-D: the DFT matrix and field; one draw gives two fields, the second
    // generateField returns the spare field without drawing
    {
      ParameterTree c = config("500", "1", "exponential", "none", "0.05");
      auto t = std::make_shared<Traits<1>>(c);
      GpuIsoMatrix<Traits<1>> m(t);
      parafields::StochasticPart<Traits<1>> sp(t);
      Rng<double> rng(11);
      m.generateField(rng, sp);
      dump("dft1_gen1", sp.values());
      m.generateField(rng, sp);
      dump("dft1_gen2", sp.values());
      m.generateField(rng, sp); // a new draw from the same stream
      dump("dft1_gen3", sp.values());
      dump("dft1_mult", (m * sp).values());
      dump("dft1_inv", m.multiplyInverse(sp).values());
    }
    // FP32 realizations with a custom covariance callable
    {
      ParameterTree c = config("32 32", "1 1", "custom-iso", "none", "0.1");
      auto t = std::make_shared<Traits2f>(c);
      GpuIsoMatrix<Traits2f> m(t);
      m.fillTransformedMatrix([](float var, const std::array<float, 2>& x) {
        return var * std::exp(-std::sqrt(x[0] * x[0] + x[1] * x[1]));
      });
      parafields::StochasticPart<Traits2f> sp(t);
      Rng<float> rng(3);
      m.generateField(rng, sp);
      dump("custom_f32_gen", sp.values());
      dump("custom_f32_inv", m.multiplyInverse(sp).values());
    }
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  std::printf("all OK\n");
  return 0;
}
