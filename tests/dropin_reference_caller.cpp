// Compile-only check: a caller written against the reference's own containers
// (HostData::HexMesh, HostField, HostFieldBundle, HostUnivariate, Stubs::HostNormalXDZ)
// swaps HostSolver::HostRSGTDUT / HostUTModel for the DeviceSolver models unchanged.
// Built by tests/test_dropin.py with the reference's include directories only.
#include <memory>

#include "HexMesh.hpp"
#include "HostArray.hpp"
#include "HostField.hpp"
#include "HostNormalXDZ.hpp"
#include "HostUnivariate.hpp"
#include "DeviceSolver/DeviceRSGTDUT.hpp"

std::shared_ptr<PanNDE::Model> build(double dt, std::shared_ptr<HostData::HostFieldBundle> domain,
                                     std::shared_ptr<HostData::HostFieldBundle> solution,
                                     std::shared_ptr<PanNDE::Array<double>> t,
                                     std::shared_ptr<PanNDE::Array<double>> v, bool moments) {
  double center[3] = {0, 0, 0};
  auto sig = HostData::HostUnivariate::makeShared(t, v);
  std::shared_ptr<PanNDE::MultiVariate> xd = Stubs::HostNormalXDZ::makeShared(center, 1e-3, 1e-4, sig);
  if (!moments) {
    std::vector<std::shared_ptr<PanNDE::MultiVariate>> xducers = {nullptr, nullptr, xd};
    auto xds = (std::shared_ptr<PanNDE::MultiVariate>(*)[3])xducers.data();
    return std::make_shared<DeviceSolver::DeviceRSGTDUT>(dt, domain, solution, xds, 1, nullptr);
  }
  auto ut = DeviceSolver::DeviceUTModel::makeShared(dt, domain, solution, nullptr);
  std::shared_ptr<PanNDE::MultiVariate> row[6] = {xd, xd, xd, nullptr, nullptr, nullptr};
  ut->applyMomentExcitation(row);
  return ut;
}
