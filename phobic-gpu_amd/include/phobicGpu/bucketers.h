// bucketers.h -- relative bucket functions used to build the fulcrum table
// (same interface and formulas as the reference's include/phobicGpu/bucketers.h:10-50:
// the formulas define the key -> bucket mapping, so they must be identical).
#pragma once
#include <cmath>

namespace phobicgpu {

class Bucketer {
 public:
  virtual void init(double lambda, double expectedTableSize) {}
  virtual double getBucketRel(double relInput) const = 0;
  virtual ~Bucketer() {}
};

// y(x) = x (bucketers.h:18-24)
class UniformBucketer : public Bucketer {
 public:
  double getBucketRel(double x) const override { return x; }
};

// PHOBIC's optimised bucket function (bucketers.h:26-39):
// c = 0.2 lambda / sqrt(P), y(x) = (x + (1-x) ln(1-x)) (1-c) + c x.
class OptBucketer : public Bucketer {
 public:
  double c = 0;
  void init(double lambda, double expectedTableSize) override { c = 0.2 * lambda / std::sqrt(expectedTableSize); }
  double getBucketRel(double x) const override { return (x + (1.0 - x) * std::log(1.0 - x)) * (1.0 - c) + c * x; }
};

// Two-slope skew (bucketers.h:41-50).
class SkewBucketer : public Bucketer {
 public:
  double getBucketRel(double x) const override {
    return x < 0.6 ? x / 0.6 * 0.3 : 0.3 + (x - 0.6) / 0.4 * 0.7;
  }
};

}  // namespace phobicgpu
