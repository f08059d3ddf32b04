// PanNDE::MultiVariate: a source f(x, y, z, t) with an optional support test checkAt
// (reference modules/PanNDE/include/MultiVariate.hpp:77-141; same signatures; the code is
// this library's own).  The default checkAt accepts every point.
#pragma once

#include <memory>

#include "Array.hpp"

namespace PanNDE {

class MultiVariate {
 public:
  virtual double evalAt(std::shared_ptr<PanNDE::Array<double>> arguments) = 0;
  virtual double evalAt(double arguments[]) = 0;
  virtual bool checkAt(std::shared_ptr<PanNDE::Array<double>> arguments) { return true; }
  virtual bool checkAt(double arguments[]) { return true; }
  virtual int argumentCount() = 0;
  virtual std::shared_ptr<PanNDE::Array<double>> sampleArgument(int arg_index) = 0;
  virtual ~MultiVariate() {}
};

}  // namespace PanNDE
