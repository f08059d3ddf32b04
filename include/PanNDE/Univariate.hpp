// PanNDE::Univariate: a function of time (reference modules/PanNDE/include/Univariate.hpp:69-81;
// same signature; the code is this library's own).
#pragma once

namespace PanNDE {

class Univariate {
 public:
  virtual double evalAt(double t) = 0;
  virtual ~Univariate() {}
};

}  // namespace PanNDE
