// trend.hh -- deterministic trend components of a random field (host side).
//
// Counterpart of TrendComponent / TrendPart (trend.hh:29-640, :851-1505 of the
// reference): mean, slope, disk and block components read from the field's
// configuration, each with coefficients mean + shift. `generate` draws the
// shifts from the field's RNG after the stochastic part (trend.hh:324-342,
// :1148-1155, called from randomfield.hh:550-554), `generateUncorrelated`
// from a fresh std::default_random_engine(seed) per component
// (trend.hh:355-385), and the matrix products scale the shifts by the
// variances (trend.hh:394-441). The image component (trend.hh:679-836) scales
// 1 - pixel / 255 of a PNG file (pngreader.hh) by its one coefficient.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <ostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <memory>

#include "parametertree.hh"
#include "pngreader.hh"

namespace parafields_amd {

template <typename RF, int dim>
class TrendPart
{
public:
  enum class Kind { Mean, Slope, Disk, Block, Image };

  struct Component
  {
    Kind kind;
    std::string name; // mean, slope, disk<k>, block<k>, image
    std::vector<RF> mean, variance, shift;
    std::shared_ptr<const PNGReader> image; // Kind::Image

    // value at location x (trend.hh:585-620)
    RF evaluate(const std::array<double, dim>& x, const std::array<double, dim>& ext) const
    {
      auto c = [&](size_t i) { return mean[i] + shift[i]; };
      switch (kind) {
        case Kind::Mean:
          return c(0);
        case Kind::Slope: {
          RF out = 0.;
          for (int i = 0; i < dim; i++)
            out += c(i) * (RF)(x[i] - ext[i] / 2.);
          return out;
        }
        case Kind::Disk: {
          RF d2 = 0.;
          for (int i = 0; i < dim; i++)
            d2 += std::pow((RF)x[i] - c(i), 2);
          return std::exp(-d2 / std::pow(c(dim), 2) * c(dim + 1));
        }
        case Kind::Block:
          for (int i = 0; i < dim; i++)
            if (std::abs((RF)x[i] - c(i)) > 0.5 * c(dim + i))
              return 0.;
          return c(2 * dim);
        case Kind::Image:
          // ImageComponent::evaluate, trend.hh:798-806
          return c(0) * (RF)image->evaluate(x, ext);
      }
      return 0.;
    }
  };

  TrendPart() = default;

  // Components in the reference's order: mean, slope, disk0, disk1, ...,
  // block0, block1, ... (trend.hh:895-1035); keys <name>.mean and
  // <name>.variance. With a file name, the coefficients come from
  // <fileName>.trend (one line "<name> = c_0 c_1 ..." per component, as
  // writeToFile writes them) and the shifts are coefficient - mean
  // (trend.hh:111-150); without one every shift is 0.
  TrendPart(const ParameterTree& config, const std::array<double, dim>& ext,
            const std::string& fileName = "")
    : ext_(ext)
  {
    ParameterTree stored;
    if (!fileName.empty())
      stored.readINIFile(fileName + ".trend");
    const ParameterTree* file = fileName.empty() ? nullptr : &stored;
    add(config, file, "mean", Kind::Mean, 1);
    add(config, file, "slope", Kind::Slope, dim);
    for (int k = 0; add(config, file, "disk" + std::to_string(k), Kind::Disk, dim + 2); k++) {
    }
    for (int k = 0; add(config, file, "block" + std::to_string(k), Kind::Block, 2 * dim + 1);
         k++) {
    }
    // the image component comes last (trend.hh:1015-1032); it needs 2D and
    // one parameter (trend.hh:722-735)
    if (config.hasKey("image.mean")) {
      if (dim != 2)
        throw std::runtime_error("image trend components require dim == 2");
      add(config, file, "image", Kind::Image, 1);
      comps_.back().image =
        std::make_shared<const PNGReader>(config.get<std::string>("image.filename"));
    }
  }

  // TrendComponent::writeToFile / TrendPart::writeToFile (trend.hh:656-675,
  // :1515-1527): "<name> = mean_i + shift_i ..." per component
  void writeToFile(std::ostream& out) const
  {
    // 17 significant digits: the coefficients read back bit for bit (the
    // reference's stream default of 6 digits would round them)
    const auto prec = out.precision(17);
    for (const auto& c : comps_) {
      out << c.name << " =";
      for (size_t i = 0; i < c.shift.size(); i++)
        out << " " << c.mean[i] + c.shift[i];
      out << std::endl;
    }
    out.precision(prec);
  }

  // arithmetic on the shifts (trend.hh:450-645, :1294-1420)
  TrendPart& operator+=(const TrendPart& o) { return axpy(o, RF(1)); }
  TrendPart& operator-=(const TrendPart& o) { return axpy(o, RF(-1)); }
  TrendPart& operator*=(RF alpha)
  {
    for (auto& c : comps_)
      for (auto& s : c.shift)
        s *= alpha;
    return *this;
  }
  TrendPart& axpy(const TrendPart& o, RF alpha)
  {
    for (size_t k = 0; k < comps_.size(); k++)
      for (size_t i = 0; i < comps_[k].shift.size(); i++)
        comps_[k].shift[i] += o.comps_.at(k).shift.at(i) * alpha;
    return *this;
  }
  void zero()
  {
    for (auto& c : comps_)
      for (auto& s : c.shift)
        s = RF(0);
  }
  RF operator*(const TrendPart& o) const
  {
    RF out = 0.;
    for (size_t k = 0; k < comps_.size(); k++)
      for (size_t i = 0; i < comps_[k].shift.size(); i++)
        out += comps_[k].shift[i] * o.comps_.at(k).shift.at(i);
    return out;
  }
  RF oneNorm() const
  {
    RF out = 0.;
    for (const auto& c : comps_)
      for (RF s : c.shift)
        out += std::abs(s);
    return out;
  }
  RF infNorm() const
  {
    RF out = 0.;
    for (const auto& c : comps_)
      for (RF s : c.shift)
        out = std::max(out, std::abs(s));
    return out;
  }

  size_t size() const { return comps_.size(); }
  const Component& component(size_t i) const { return comps_.at(i); }
  Component& component(size_t i) { return comps_.at(i); }

  // shift_i = sample() * sqrt(variance_i), component by component
  // (trend.hh:324-342, :1148-1155)
  template <typename RNG>
  void generate(RNG& rng)
  {
    for (auto& c : comps_)
      for (size_t i = 0; i < c.shift.size(); i++)
        c.shift[i] = rng.sample() * std::sqrt(c.variance[i]);
  }

  // white-noise coefficients: each component draws from its own
  // std::default_random_engine(seed) with std::normal_distribution
  // (trend.hh:355-385)
  void generateUncorrelated(unsigned int seed)
  {
    for (auto& c : comps_) {
      std::default_random_engine generator(seed);
      std::normal_distribution<RF> normal(0., 1.);
      for (auto& s : c.shift)
        s = normal(generator);
    }
  }

  // trend.hh:394-441
  void timesMatrix() { scale([](RF v) { return v; }); }
  void timesInverseMatrix() { scale([](RF v) { return 1. / v; }); }
  void timesMatrixRoot() { scale([](RF v) { return std::sqrt(v); }); }
  void timesInvMatRoot() { scale([](RF v) { return 1. / std::sqrt(v); }); }

  // sum of the component values at x (trend.hh:1489-1505)
  RF evaluate(const std::array<double, dim>& x) const
  {
    RF out = 0.;
    for (const auto& c : comps_)
      out += c.evaluate(x, ext_);
    return out;
  }

private:
  std::array<double, dim> ext_{};
  std::vector<Component> comps_;

  bool add(const ParameterTree& config, const ParameterTree* file, const std::string& name,
           Kind kind, size_t n)
  {
    if (!config.hasKey(name + ".mean"))
      return false;
    Component c;
    c.kind = kind;
    c.name = name;
    c.mean = config.get<std::vector<RF>>(name + ".mean");
    c.variance = config.get<std::vector<RF>>(name + ".variance");
    if (c.mean.size() != c.variance.size())
      throw std::runtime_error("trend component size does not match");
    if (c.mean.size() != n)
      throw std::runtime_error(kind == Kind::Image
                                 ? std::string("Image component must only contain one parameter")
                                 : "trend component " + name + " must contain " +
                                     std::to_string(n) + " parameters");
    c.shift.assign(n, RF(0));
    if (file) {
      const auto coef = file->get<std::vector<RF>>(name);
      if (coef.size() != n)
        throw std::runtime_error("trend component size does not match");
      for (size_t i = 0; i < n; i++)
        c.shift[i] = coef[i] - c.mean[i];
    }
    comps_.push_back(c);
    return true;
  }

  template <typename F>
  void scale(F f)
  {
    for (auto& c : comps_)
      for (size_t i = 0; i < c.shift.size(); i++)
        c.shift[i] *= f(c.variance[i]);
  }
};

} // namespace parafields_amd
