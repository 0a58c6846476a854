// Host test of the configuration reader (make_desc in
// include/parafields_amd/randomfield.hh), no GPU: the embedding keys of the
// classical, merge, fold and cofold periodizations and their defaults
// (matrix.hh:515-580, :652-1013).
#include <cstdio>
#include <stdexcept>

#include "parafields_amd/randomfield.hh"

using namespace parafields_amd;

static int bad = 0;
static void expect(bool ok, const char* what)
{
  if (!ok) {
    std::printf("FAIL %s\n", what);
    bad++;
  }
}

static ParameterTree base(const char* per)
{
  ParameterTree c;
  c["grid.cells"] = "16 16";
  c["grid.extensions"] = "1 1";
  c["stochastic.covariance"] = "exponential";
  c["stochastic.variance"] = "1";
  c["stochastic.corrLength"] = "0.1";
  if (per)
    c["embedding.periodization"] = per;
  return c;
}

int main()
{
  const pf_desc d0 = make_desc<double, 2>(base(nullptr), PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(d0.periodization == PF_PERIOD_CLASSICAL, "classical default");
  // fold: foldRecursions 99 (default) -> 0 (family rule), foldMaxFactor 1
  pf_desc f = make_desc<double, 2>(base("fold"), PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(f.periodization == PF_PERIOD_FOLD && f.merge_recursions == 0 && f.fold_max_factor == 1.,
         "fold defaults");
  ParameterTree cf = base("fold");
  cf["embedding.foldRecursions"] = "3";
  cf["embedding.foldMaxFactor"] = "1.5";
  cf["embedding.mergeRecursions"] = "7"; // not read for fold
  f = make_desc<double, 2>(cf, PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(f.merge_recursions == 3 && f.fold_max_factor == 1.5, "fold keys");
  // cofold: cofoldRecursions 1, cofoldMaxFactor 5; 0 recursions -> -1
  pf_desc k = make_desc<double, 2>(base("cofold"), PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(k.periodization == PF_PERIOD_COFOLD && k.merge_recursions == 1 &&
           k.fold_max_factor == 5.,
         "cofold defaults");
  ParameterTree ck = base("cofold");
  ck["embedding.cofoldRecursions"] = "0";
  ck["embedding.cofoldMaxFactor"] = "2.5";
  ck["embedding.sigmoid"] = "smooth";
  k = make_desc<double, 2>(ck, PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(k.merge_recursions == -1 && k.fold_max_factor == 2.5 && k.sigmoid == PF_SIGMOID_SMOOTH,
         "cofold keys");
  // merge keeps mergeRecursions
  ParameterTree cm = base("merge");
  cm["embedding.mergeRecursions"] = "2";
  const pf_desc m = make_desc<double, 2>(cm, PF_MATRIX_DCT, PF_FIELD_R2C);
  expect(m.periodization == PF_PERIOD_MERGE && m.merge_recursions == 2, "merge keys");
  bool threw = false;
  try {
    (void)make_desc<double, 2>(base("spiral"), PF_MATRIX_DCT, PF_FIELD_R2C);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "unknown periodization rejected");
  if (bad == 0)
    std::printf("all OK\n");
  return bad ? 1 : 0;
}
