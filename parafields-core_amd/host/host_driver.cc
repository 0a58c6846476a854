// host_driver.cc -- exercises the C++ drop-in API (parafields_amd/randomfield.hh)
// end to end on the GPU.
//   pf_host_driver smoke           reference smoke matrix (test/fieldtest.cc:26-57)
//   pf_host_driver field <prefix>  2D 256x256 exponential l=0.05, generate(42) in
//                                  double, then timesMatrix, timesMatrixRoot and
//                                  timesInverseMatrix; raw doubles to <prefix>.*.bin
//   pf_host_driver trend <prefix>  2D 64x48 exponential l=0.1 with mean, slope, disk
//                                  and block trend components: generate(42), the
//                                  stochastic part, the trend coefficients and
//                                  bulkEvaluate to <prefix>.*.bin
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "parafields_amd/randomfield.hh"

using namespace parafields_amd;

template <typename RF, int dim>
int smoke_one(const char* cov)
{
  ParameterTree config;
  config["grid.cells"] = dim == 2 ? "16 16" : "8 8 8";
  config["grid.extensions"] = dim == 2 ? "1 1" : "1 1 1";
  config["stochastic.variance"] = "1";
  config["stochastic.corrLength"] = "0.05";
  config["stochastic.covariance"] = cov;
  RandomField<GridTraits<RF, RF, dim>> field(config);
  field.generate();
  for (RF v : field.dataVector())
    if (!std::isfinite((double)v))
      return 1;
  return 0;
}

template <typename V>
void write(const std::string& path, const V& v)
{
  std::ofstream o(path, std::ios::binary);
  o.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(v[0]));
}

int main(int argc, char** argv)
{
  if (argc < 2) {
    std::cerr << "usage: pf_host_driver smoke | field <prefix>\n";
    return 2;
  }
  try {
    if (!std::strcmp(argv[1], "smoke")) {
      int bad = 0;
      for (const char* cov : { "exponential", "gaussian", "spherical" }) {
        bad += smoke_one<float, 2>(cov) + smoke_one<double, 2>(cov);
        bad += smoke_one<float, 3>(cov) + smoke_one<double, 3>(cov);
      }
      std::cout << (bad ? "smoke: FAILED" : "smoke: ok") << std::endl;
      return bad ? 1 : 0;
    }
    if (!std::strcmp(argv[1], "field") && argc > 2) {
      const std::string pre = argv[2];
      ParameterTree config;
      std::istringstream ini("[grid]\ncells = 256 256\nextensions = 1 1\n"
                             "[stochastic]\nvariance = 1.0\ncovariance = exponential\n"
                             "corrLength = 0.05\n");
      config.readINI(ini);
      RandomField<GridTraits<double, double, 2>> field(config);
      field.generate(42u);
      write(pre + ".gen.bin", field.dataVector());
      field.timesMatrix();
      write(pre + ".mul.bin", field.dataVector());
      field.timesMatrixRoot();
      write(pre + ".root.bin", field.dataVector());
      field.generate(43u);
      write(pre + ".gen43.bin", field.dataVector());
      field.timesInverseMatrix();
      write(pre + ".inv.bin", field.dataVector());
      std::vector<double> vals;
      std::array<unsigned, 2> sizes;
      field.bulkEvaluate(vals, sizes);
      std::cout << "field: ok " << sizes[0] << "x" << sizes[1] << std::endl;
      return 0;
    }
    if (!std::strcmp(argv[1], "trend") && argc > 2) {
      const std::string pre = argv[2];
      ParameterTree config;
      std::istringstream ini("[grid]\ncells = 64 48\nextensions = 1 1\n"
                             "[stochastic]\nvariance = 1.0\ncovariance = exponential\n"
                             "corrLength = 0.1\n"
                             "[mean]\nmean = 0.5\nvariance = 0.25\n"
                             "[slope]\nmean = 1 -2\nvariance = 0.5 2\n"
                             "[disk0]\nmean = 0.3 0.6 0.1 1\nvariance = 0.01 0.01 0.001 0.5\n"
                             "[block0]\nmean = 0.5 0.5 0.2 0.3 2\nvariance = 0.01 0.01 0 0 1\n");
      config.readINI(ini);
      RandomField<GridTraits<double, double, 2>> field(config);
      field.generate(42u);
      write(pre + ".stoch.bin", field.dataVector());
      std::vector<double> shifts, means;
      for (size_t k = 0; k < field.trendPart().size(); k++)
        for (size_t i = 0; i < field.trendPart().component(k).shift.size(); i++) {
          shifts.push_back(field.trendPart().component(k).shift[i]);
          means.push_back(field.trendPart().component(k).mean[i]);
        }
      write(pre + ".shift.bin", shifts);
      std::vector<double> vals;
      std::array<unsigned, 2> sizes;
      field.bulkEvaluate(vals, sizes);
      write(pre + ".bulk.bin", vals);
      // evaluate() at every cell centre agrees with bulkEvaluate
      double worst = 0.;
      for (unsigned j = 0; j < sizes[1]; j++)
        for (unsigned i = 0; i < sizes[0]; i++) {
          double v;
          field.evaluate({ (i + 0.5) / sizes[0], (j + 0.5) / sizes[1] }, v);
          worst = std::max(worst, std::fabs(v - vals[i + sizes[0] * j]));
        }
      std::cout << "trend: ok " << shifts.size() << " coefficients, evaluate/bulk " << worst
                << std::endl;
      return worst < 1e-12 ? 0 : 1;
    }
  } catch (const NegativeEigenvalueError& e) {
    std::cerr << "NegativeEigenvalueError: " << e.what() << std::endl;
    return 3;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 4;
  }
  return 2;
}
