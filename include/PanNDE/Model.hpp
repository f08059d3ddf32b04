// PanNDE::Model: a time-stepping solver (reference modules/PanNDE/include/Model.hpp:77-106;
// same signatures; the code is this library's own).  solve(n) advances n steps and returns the
// simulation time; getStates() returns the live solution bundle.
#pragma once

#include <memory>

#include "Field.hpp"
#include "Mesh.hpp"

namespace PanNDE {

class Model {
 public:
  virtual double solve(int nsteps = 1) = 0;
  virtual std::shared_ptr<PanNDE::FieldBundle> getStates() = 0;
  virtual ~Model() {}
};

}  // namespace PanNDE
