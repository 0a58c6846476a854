// Host test of the image trend component (include/parafields_amd/trend.hh
// with pngreader.hh), no GPU: the component's value at a set of points for a
// grayscale and an RGB PNG, its place after the other components in
// generate's draw order and in the .trend file, and the rejected
// configurations.
//
//   image_check <gray.png> <rgb.png> <out dir>
// writes <out dir>/values.txt: one line per point "x y gray rgb".
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

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
  double v = 0.;
  double sample() { return v += 1.; }
};

static ParameterTree config(const std::string& png)
{
  ParameterTree c;
  c["mean.mean"] = "0.5";
  c["mean.variance"] = "0";
  c["image.mean"] = "2";
  c["image.variance"] = "0.25";
  c["image.filename"] = png;
  return c;
}

int main(int argc, char** argv)
{
  if (argc < 4)
    return 2;
  const std::string dir = argv[3];
  try {
    const std::array<double, 2> ext{ 2.0, 1.0 };
    TrendPart<double, 2> g(config(argv[1]), ext), c(config(argv[2]), ext);
    expect(g.size() == 2 && g.component(1).name == "image", "image component last");
    std::ofstream out(dir + "/values.txt");
    out.precision(17);
    for (int j = 0; j < 6; j++)
      for (int i = 0; i < 8; i++) {
        const std::array<double, 2> x{ (i + 0.5) * ext[0] / 8, (j + 0.5) * ext[1] / 6 };
        out << x[0] << " " << x[1] << " " << g.evaluate(x) << " " << c.evaluate(x) << "\n";
      }
    // generate: the image coefficient is drawn after the mean's
    Counting r;
    g.generate(r);
    expect(g.component(0).shift[0] == 0. && g.component(1).shift[0] == 2. * 0.5,
           "generate draws the image shift last");
    {
      std::ofstream f(dir + "/t.trend");
      g.writeToFile(f);
    }
    TrendPart<double, 2> back(config(argv[1]), ext, dir + "/t");
    expect(back.component(1).shift[0] == g.component(1).shift[0], "image shift through the file");
    std::ifstream tf(dir + "/t.trend");
    std::stringstream ts;
    ts << tf.rdbuf();
    expect(ts.str().find("\nimage = 3") != std::string::npos, "image line of the trend file");
    // rejected configurations
    bool threw = false;
    try {
      ParameterTree k = config(argv[1]);
      k["image.mean"] = "1 2";
      k["image.variance"] = "1 1";
      TrendPart<double, 2> t(k, ext);
    } catch (const std::exception&) {
      threw = true;
    }
    expect(threw, "two image parameters rejected");
    threw = false;
    try {
      TrendPart<double, 3> t(config(argv[1]), { 1., 1., 1. });
    } catch (const std::exception&) {
      threw = true;
    }
    expect(threw, "image component in 3D rejected");
    threw = false;
    try {
      TrendPart<double, 2> t(config(dir + "/missing.png"), ext);
    } catch (const std::exception&) {
      threw = true;
    }
    expect(threw, "missing PNG rejected");
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  std::printf(bad ? "FAILURES %d\n" : "all OK\n", bad);
  return bad;
}
