// Host test of the on-disk format (include/parafields_amd/io.hh, the trend
// file of trend.hh, ParameterTree::report), no GPU: the HDF5 field file read
// back bit for bit in FP64 and FP32, a grid mismatch rejected, the XDMF text,
// trend coefficients through a .trend file, and a configuration through
// report() and readINI. Writes <dir>/f3.stoch.h5 etc. for h5dump.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>
#include <string>

#include "parafields_amd/io.hh"
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

struct Counting
{
  double v = 0.1;
  double sample() { return v += 0.37; }
};

static std::string slurp(const std::string& f)
{
  std::ifstream in(f);
  std::stringstream s;
  s << in.rdbuf();
  return s.str();
}

int main(int argc, char** argv)
{
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::mt19937_64 rng(5);
  std::normal_distribution<double> nd;
  // 3D FP64 field, 5 x 4 x 3 cells (x fastest)
  {
    std::array<unsigned int, 3> cells{ 5, 4, 3 };
    std::vector<double> v(60);
    for (auto& x : v)
      x = nd(rng);
    writeStochasticHDF5<double, 3>(cells, v, "/stochastic", dir + "/f3.stoch.h5");
    const auto r = readStochasticHDF5<double, 3>(cells, "/stochastic", dir + "/f3.stoch.h5");
    expect(r == v, "FP64 3D round trip");
    bool threw = false;
    try {
      readStochasticHDF5<double, 3>({ 4, 5, 3 }, "/stochastic", dir + "/f3.stoch.h5");
    } catch (const std::exception&) {
      threw = true;
    }
    expect(threw, "grid mismatch rejected");
    threw = false;
    try {
      readStochasticHDF5<double, 3>(cells, "/stochastic", dir + "/missing.stoch.h5");
    } catch (const std::exception&) {
      threw = true;
    }
    expect(threw, "missing file rejected");
    std::array<double, 3> ext{ 1.0, 2.0, 0.5 };
    writeToXDMF<double, 3>(cells, ext, dir + "/f3");
    const std::string x = slurp(dir + "/f3.xdmf");
    expect(x.find("NumberOfElements=\"3 4 5 \"") != std::string::npos, "XDMF topology dims");
    expect(x.find("0.2 0.5 0.166667 ") != std::string::npos, "XDMF cell sizes");
    expect(x.find(dir + "/f3.stoch.h5:/stochastic") != std::string::npos, "XDMF dataset");
  }
  // known values, x fastest, for h5dump: v[i] = i + 0.25 on 4 x 3 cells
  {
    std::vector<double> v(12);
    for (size_t i = 0; i < v.size(); i++)
      v[i] = (double)i + 0.25;
    writeStochasticHDF5<double, 2>({ 4, 3 }, v, "/stochastic", dir + "/lin.stoch.h5");
  }
  // 2D FP32 field stored as F64 (values exactly representable)
  {
    std::array<unsigned int, 2> cells{ 7, 3 };
    std::vector<float> v(21);
    for (auto& x : v)
      x = (float)nd(rng);
    writeStochasticHDF5<float, 2>(cells, v, "/stochastic", dir + "/f2.stoch.h5");
    const auto r = readStochasticHDF5<float, 2>(cells, "/stochastic", dir + "/f2.stoch.h5");
    expect(r == v, "FP32 2D round trip");
    const auto r64 = readStochasticHDF5<double, 2>(cells, "/stochastic", dir + "/f2.stoch.h5");
    bool same = true;
    for (size_t i = 0; i < v.size(); i++)
      same = same && r64[i] == (double)v[i];
    expect(same, "FP32 values stored as exact F64");
    std::array<double, 2> ext{ 1.0, 1.0 };
    writeToXDMF<double, 2>(cells, ext, dir + "/f2");
    const std::string x = slurp(dir + "/f2.xdmf");
    // 2D pads the spacing with the first cell size (three entries)
    expect(x.find("0.142857 0.142857 0.333333 ") != std::string::npos, "XDMF 2D padding");
  }
  // trend coefficients through the .trend file
  {
    ParameterTree c;
    c["mean.mean"] = "0.5";
    c["mean.variance"] = "4";
    c["slope.mean"] = "1 -2";
    c["slope.variance"] = "1 9";
    c["disk0.mean"] = "0.25 0.75 0.1 2";
    c["disk0.variance"] = "0.01 0.01 0.001 1";
    c["block0.mean"] = "0.5 0.5 0.2 0.3 1";
    c["block0.variance"] = "0 0 0 0 2";
    TrendPart<double, 2> t(c, { 1.0, 1.0 });
    Counting r;
    t.generate(r);
    {
      std::ofstream f(dir + "/t.trend");
      t.writeToFile(f);
    }
    TrendPart<double, 2> u(c, { 1.0, 1.0 }, dir + "/t");
    // the file holds the coefficients mean + shift at 17 digits; the shifts
    // come back as coefficient - mean, within rounding of that subtraction
    bool same = u.size() == t.size();
    for (size_t k = 0; same && k < t.size(); k++)
      for (size_t i = 0; i < t.component(k).shift.size(); i++) {
        const double c = t.component(k).mean[i] + t.component(k).shift[i];
        same = same && std::abs(u.component(k).shift[i] - t.component(k).shift[i]) <=
                         4e-16 * (std::abs(c) + std::abs(t.component(k).mean[i]));
      }
    expect(same, "trend shifts through the file");
    const std::string tx = slurp(dir + "/t.trend");
    expect(tx.rfind("mean =", 0) == 0 && tx.find("\nslope =") != std::string::npos &&
             tx.find("\ndisk0 =") != std::string::npos &&
             tx.find("\nblock0 =") != std::string::npos,
           "trend file lines in component order");
    // arithmetic on the shifts
    TrendPart<double, 2> w = t;
    w.axpy(t, 2.0);
    expect(std::abs(w * t - 3.0 * (t * t)) <= 1e-12 * (t * t), "trend axpy and scalar product");
    w.zero();
    expect(w.oneNorm() == 0. && w.infNorm() == 0., "trend zero");
  }
  // configuration through report() and readINI
  {
    ParameterTree c;
    c["grid.cells"] = "64 32";
    c["grid.extensions"] = "1 0.5";
    c["stochastic.corrLength"] = "0.1 0.05";
    c["stochastic.covariance"] = "matern32";
    c["embedding.factor"] = "2";
    c["top"] = "x y";
    std::stringstream s;
    c.report(s);
    ParameterTree d;
    d.readINI(s);
    const auto a = d.get<std::array<int, 2>>("grid.cells");
    expect(a[0] == 64 && a[1] == 32, "report/readINI array");
    expect(d.get<std::string>("stochastic.covariance") == "matern32", "report/readINI string");
    expect(d.get<std::string>("top") == "x y", "report/readINI unsectioned key");
    expect(s.str().find("[ stochastic ]") != std::string::npos, "report sections");
  }
  std::printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
