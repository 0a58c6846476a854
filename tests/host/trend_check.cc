// Host test of the trend part (include/parafields_amd/trend.hh), no GPU:
// component order and parameter counts, shapes at points whose values are
// known in closed form, the draw order and scaling of generate (a scripted
// RNG returning 1, 2, 3, ...), the variance scalings of the matrix products,
// and generateUncorrelated against std::default_random_engine itself.
#include <cmath>
#include <cstdio>
#include <random>

#include "parafields_amd/trend.hh"

using namespace parafields_amd;

static int bad = 0;
static void expect(bool ok, const char* what)
{
  if (!ok) {
    std::printf("FAIL %s\n", what);
    bad++;
  }
}
static bool near(double a, double b) { return std::fabs(a - b) <= 1e-14 * (1 + std::fabs(b)); }

struct Counting
{
  double v = 0;
  double sample() { return v += 1; }
};

int main()
{
  ParameterTree c;
  c["mean.mean"] = "0.5";
  c["mean.variance"] = "4";
  c["slope.mean"] = "1 -2";
  c["slope.variance"] = "1 9";
  c["disk0.mean"] = "0.25 0.75 0.1 2";
  c["disk0.variance"] = "0 0 0 0";
  c["disk1.mean"] = "0.5 0.5 0.2 1";
  c["disk1.variance"] = "1 1 1 1";
  c["block0.mean"] = "0.5 0.5 0.2 0.4 3";
  c["block0.variance"] = "0 0 0 0 0";
  using T = TrendPart<double, 2>;
  T t(c, { 1., 1. });
  expect(t.size() == 5, "five components");
  expect(t.component(0).kind == T::Kind::Mean && t.component(1).kind == T::Kind::Slope &&
           t.component(2).kind == T::Kind::Disk && t.component(3).kind == T::Kind::Disk &&
           t.component(4).kind == T::Kind::Block,
         "order mean, slope, disk0, disk1, block0");
  const std::array<double, 2> ext{ 1., 1. };
  // shapes (shifts zero)
  expect(near(t.component(0).evaluate({ 0.1, 0.9 }, ext), 0.5), "mean is constant");
  expect(near(t.component(1).evaluate({ 0.5, 0.5 }, ext), 0.), "slope is zero at the centre");
  expect(near(t.component(1).evaluate({ 1.0, 0.5 }, ext), 0.5), "slope along x");
  expect(near(t.component(1).evaluate({ 0.5, 0.0 }, ext), 1.0), "slope along y");
  expect(near(t.component(3).evaluate({ 0.5, 0.5 }, ext), 1.), "disk is 1 at its centre");
  expect(near(t.component(3).evaluate({ 0.7, 0.5 }, ext), std::exp(-1.)),
         "disk is exp(-value) at its radius");
  expect(near(t.component(2).evaluate({ 0.25, 0.85 }, ext), std::exp(-2.)),
         "second disk at its radius");
  expect(near(t.component(4).evaluate({ 0.55, 0.65 }, ext), 3.), "block inside");
  expect(near(t.component(4).evaluate({ 0.65, 0.5 }, ext), 0.), "block outside in x");
  expect(near(t.component(4).evaluate({ 0.5, 0.75 }, ext), 0.), "block outside in y");
  expect(near(t.evaluate({ 0.5, 0.5 }), 0.5 + 0. + std::exp(-(0.0625 + 0.0625) / 0.01 * 2.) + 1. + 3.),
         "sum of the components");
  // generate: sample() * sqrt(variance), components in order
  Counting rng;
  t.generate(rng);
  const double want[5][5] = { { 1 * 2. }, { 2 * 1., 3 * 3. }, { 0, 0, 0, 0 }, { 8, 9, 10, 11 },
                              { 0, 0, 0, 0, 0 } };
  for (size_t k = 0; k < 5; k++)
    for (size_t i = 0; i < t.component(k).shift.size(); i++)
      expect(near(t.component(k).shift[i], want[k][i]), "generate draw order and scaling");
  expect(rng.v == 16., "16 samples drawn");
  expect(near(t.component(0).evaluate({ 0.3, 0.3 }, ext), 2.5), "mean plus shift");
  // matrix products scale the shifts
  t.timesMatrix();
  expect(near(t.component(0).shift[0], 8.) && near(t.component(1).shift[1], 81.), "timesMatrix");
  t.timesInverseMatrix();
  expect(near(t.component(0).shift[0], 2.) && near(t.component(1).shift[1], 9.),
         "timesInverseMatrix");
  t.timesMatrixRoot();
  expect(near(t.component(0).shift[0], 4.) && near(t.component(1).shift[1], 27.),
         "timesMatrixRoot");
  t.timesInvMatRoot();
  expect(near(t.component(0).shift[0], 2.) && near(t.component(1).shift[1], 9.), "timesInvMatRoot");
  // generateUncorrelated: a fresh default_random_engine(seed) per component
  t.generateUncorrelated(5u);
  for (size_t k = 0; k < 5; k++) {
    std::default_random_engine g(5u);
    std::normal_distribution<double> n(0., 1.);
    for (size_t i = 0; i < t.component(k).shift.size(); i++)
      expect(t.component(k).shift[i] == n(g), "generateUncorrelated stream");
  }
  // parameter counts are checked
  ParameterTree wrong;
  wrong["disk0.mean"] = "0.5 0.5 0.1";
  wrong["disk0.variance"] = "0 0 0";
  bool threw = false;
  try {
    T w(wrong, { 1., 1. });
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "a disk needs dim + 2 parameters");
  std::printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
