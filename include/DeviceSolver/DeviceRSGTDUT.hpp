// DeviceRSGTDUT.hpp -- the MI355X DeviceSolver as a PanNDE::Model.
//
// Drop-in for HostSolver::HostRSGTDUT (modules/HostSolver/include/HostRSGTDUT.hpp:87-169)
// and HostSolver::HostUTModel (modules/HostSolver/include/HostUTModel.hpp:84-193): same
// constructor arguments, same solve/getStates contract, the same fields looked up by the
// same standard names.  The arithmetic runs in libpannde_rsg.so (include/pannde_rsg.h);
// this header only reads the caller's mesh, fields and sources through the PanNDE
// interfaces and marshals them across the C ABI.
//
// It includes the interfaces by bare name ("Model.hpp", ...), as the reference solver
// does, so it compiles against the reference's modules/PanNDE/include or against this
// repository's include/PanNDE.
//
// Differences a caller can see:
//  * getStates() copies the device state into the solution fields it returns; between
//    calls the host fields are not refreshed (the reference updates them in place every
//    step).  Fields written by the caller are uploaded with pushStates().
//  * The communicator may be nullptr and is never used (the whole mesh is on one GPU).
//  * Errors from the device library are thrown as std::runtime_error (mesh spacing errors
//    as std::logic_error with the reference's "DX"/"DY"/"DZ" text).
//  * Sources: a NormalXDZ source (below) is evaluated on the device from its signal
//    table.  Any other MultiVariate is evaluated on the host, per step, at every cell
//    centre (force) or node (moment) where its checkAt accepts the point, in the
//    reference's order; the default checkAt (true everywhere) makes that every point.
//    evalAt is assumed to be zero wherever checkAt is false.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Field.hpp"
#include "Mesh.hpp"
#include "Model.hpp"
#include "MultiVariate.hpp"
#include "StandardNames.hpp"
#include "Univariate.hpp"
#include "../pannde_rsg.h"

namespace PanNDE {
class Communicator;
}

namespace DeviceSolver {

inline void check(int rc) {
  if (rc == RSG_OK) return;
  char msg[512];
  rsg_last_error(msg, sizeof msg);
  if (rc == RSG_ERR_MESH) throw std::logic_error(msg);
  throw std::runtime_error(std::string("DeviceSolver: ") + msg);
}

// A time series with HostData::HostUnivariate semantics (HostUnivariate.hpp:68-142):
// keys ascend, a repeated key keeps its first value, constant after the last key,
// linear interpolation below it.
class TableSignal : public PanNDE::Univariate {
 public:
  TableSignal(std::vector<double> times, std::vector<double> values) {
    if (times.size() != values.size()) throw std::runtime_error("mismatched array lengths");
    for (size_t k = 0; k < times.size(); k++) table.emplace(times[k], values[k]);
    for (auto& kv : table) {
      t.push_back(kv.first);
      v.push_back(kv.second);
    }
  }
  double evalAt(double at) override { return value(at); }
  double value(double at) const {
    double out = 0.0;
    check(rsg_signal_eval((int32_t)t.size(), t.data(), v.data(), 1, &at, &out));
    return out;
  }
  const std::vector<double>& times() const { return t; }
  const std::vector<double>& values() const { return v; }

 private:
  std::map<double, double> table;
  std::vector<double> t, v;
};

// A circular normal source with the semantics of Stubs::HostNormalXDZ
// (modules/Stubs/include/HostNormalXDZ.hpp:117-145): the window |z-zc| <= 0.75*ds and
// distance < radius, times a TableSignal.  The device evaluates it from the table.
class NormalXDZ : public PanNDE::MultiVariate {
 public:
  NormalXDZ(const double c[3], double radius, double ds, std::shared_ptr<TableSignal> signal)
      : radius_(radius), ds_(ds), signal_(std::move(signal)) {
    for (int k = 0; k < 3; k++) center_[k] = c[k];
  }
  double evalAt(std::shared_ptr<PanNDE::Array<double>> a) override { return evalAt(a->data()); }
  double evalAt(double a[]) override {
    const double scale = signal_->evalAt(a[3]);
    return scale * (double)window(a);
  }
  bool checkAt(std::shared_ptr<PanNDE::Array<double>> a) override { return checkAt(a->data()); }
  bool checkAt(double a[]) override { return window(a); }
  int argumentCount() override { return 4; }
  std::shared_ptr<PanNDE::Array<double>> sampleArgument(int) override { return nullptr; }

  const double* center() const { return center_; }
  double radius() const { return radius_; }
  double ds() const { return ds_; }
  const TableSignal& signal() const { return *signal_; }

 private:
  bool window(const double a[]) const {
    double res = 0.0;
    for (int k = 0; k < 3; k++) res += (a[k] - center_[k]) * (a[k] - center_[k]);
    const double dist = std::sqrt(res);
    return (std::fabs(a[2] - center_[2]) <= 0.75 * ds_) && (dist < radius_) && (dist >= 0.);
  }
  double center_[3];
  double radius_, ds_;
  std::shared_ptr<TableSignal> signal_;
};

// Shared body of the two models.
class DeviceModel : public PanNDE::Model {
 public:
  virtual ~DeviceModel() {
    if (h_) rsg_destroy(h_);
  }
  DeviceModel(const DeviceModel&) = delete;
  DeviceModel& operator=(const DeviceModel&) = delete;

  // Model::solve (HostRSGTDUT.hpp:162-169): nsteps of velocity, stress, time += dt.
  double solve(int nsteps = 1) override {
    if (nsteps <= 0) return time_;
    commit();
    if (opaque_.empty()) {
      check(rsg_step(h_, nsteps, nullptr, &time_));
      return time_;
    }
    const size_t nch = kinds_.size();
    amps_.assign((size_t)nsteps * nch, 0.0);
    double t = time_;
    for (int k = 0; k < nsteps; k++) {
      double* row = amps_.data() + (size_t)k * nch;
      for (size_t c = 0; c < nch; c++) {
        const Src& s = srcs_[c];
        double args[4] = {s.point[0], s.point[1], s.point[2], t};
        row[c] = s.table ? 0.125 * (s.table->signal().value(t) * 1.0) : 0.125 * (s.mv->evalAt(args));
      }
      t += dt_;
    }
    check(rsg_step(h_, nsteps, amps_.data(), &time_));
    return time_;
  }

  // Model::getStates (HostRSGTDUT.hpp:153): the caller's solution bundle, refreshed from
  // the device.
  std::shared_ptr<PanNDE::FieldBundle> getStates() override {
    pullStates();
    return solution_;
  }

  // Copies the device state into the solution fields.
  void pullStates() {
    std::vector<double> V((size_t)3 * nn_), S((size_t)6 * nc_);
    check(rsg_get_state(h_, V.data(), S.data()));
    for (int k = 0; k < 3; k++) {
      auto f = solution_->field(names_.V[k]);
      for (int32_t i = 0; i < nn_; i++) f->at(i) = V[(size_t)k * nn_ + i];
    }
    for (int k = 0; k < 6; k++) {
      auto f = solution_->field(names_.SI[k]);
      for (int32_t i = 0; i < nc_; i++) f->at(i) = S[(size_t)k * nc_ + i];
    }
  }

  // Uploads the solution fields (after the caller changed them).
  void pushStates() {
    std::vector<double> V((size_t)3 * nn_), S((size_t)6 * nc_);
    for (int k = 0; k < 3; k++) {
      auto f = solution_->field(names_.V[k]);
      for (int32_t i = 0; i < nn_; i++) V[(size_t)k * nn_ + i] = f->at(i);
    }
    for (int k = 0; k < 6; k++) {
      auto f = solution_->field(names_.SI[k]);
      for (int32_t i = 0; i < nc_; i++) S[(size_t)k * nc_ + i] = f->at(i);
    }
    check(rsg_set_state(h_, V.data(), S.data()));
  }

  // Sensor nodes recorded after every step on the device; traces()[step][sensor][3].
  void setSensors(const std::vector<int32_t>& nodes) {
    check(rsg_set_sensors(h_, (int32_t)nodes.size(), nodes.data()));
    nsens_ = nodes.size();
    trace_base_ = steps();
  }
  std::vector<double> traces() {
    const int64_t n = steps() - trace_base_;
    std::vector<double> out((size_t)n * nsens_ * 3);
    if (n > 0) check(rsg_get_traces(h_, 0, n, out.data()));
    return out;
  }
  int64_t steps() {
    int64_t s = 0;
    check(rsg_get_time(h_, nullptr, &s));
    return s;
  }
  void synchronize() { check(rsg_synchronize(h_)); }
  rsg_handle handle() { return h_; }

 protected:
  DeviceModel() = default;

  void setup(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
             std::shared_ptr<PanNDE::FieldBundle> initial_condition, int moment_form, int precision) {
    dt_ = dt;
    domain_ = domain;
    solution_ = initial_condition;
    mesh_ = domain->mesh();
    nn_ = mesh_->nodeCount();
    nc_ = mesh_->cellCount();
    x_.resize(nn_);
    y_.resize(nn_);
    z_.resize(nn_);
    for (int32_t n = 0; n < nn_; n++) {
      double p[3];
      mesh_->nodeCoordinate(n, p);
      x_[n] = p[0];
      y_[n] = p[1];
      z_[n] = p[2];
    }
    cells_.resize((size_t)nc_ * 8);
    for (int32_t c = 0; c < nc_; c++) mesh_->cell(c, cells_.data() + (size_t)c * 8);
    rsg_mesh_desc md{nn_, nc_, x_.data(), y_.data(), z_.data(), cells_.data()};
    rsg_model_desc mo;
    mo.dt = dt;
    mo.moment_form = moment_form;
    keep_.clear();
    mo.density = fieldDesc(domain->field(mat_.density));
    static const int kI[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
    static const int kJ[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};
    for (int k = 0; k < 21; k++) mo.cij[k] = fieldDesc(domain->field(mat_.CIJ[kI[k]][kJ[k]]));
    check(rsg_create(&md, &mo, precision, &h_));
    keep_.clear();
    pushStates();
  }

  void addForceRow(std::shared_ptr<PanNDE::MultiVariate> row[3]) {
    for (int d = 0; d < 3; d++)
      if (row[d]) force_.push_back({d, row[d]});
    dirty_ = true;
  }
  void addMomentRow(std::shared_ptr<PanNDE::MultiVariate> row[6]) {
    for (int k = 0; k < 6; k++)
      if (row[k]) moment_.push_back({k, row[k]});
    dirty_ = true;
  }

 private:
  struct Row {
    int comp;
    std::shared_ptr<PanNDE::MultiVariate> mv;
  };
  struct Src {  // one device channel
    const NormalXDZ* table;  // table-driven source, or
    PanNDE::MultiVariate* mv;
    double point[3];  // the evaluation point of an opaque single-point channel
  };

  rsg_field_desc fieldDesc(std::shared_ptr<PanNDE::Field> f) {
    if (!f) throw std::runtime_error("DeviceSolver: missing material field");
    rsg_field_desc d;
    d.type = (int32_t)f->type();
    const int32_t n = f->size();
    std::vector<double> v((size_t)n);
    for (int32_t i = 0; i < n; i++) v[i] = f->at(i);
    keep_.push_back(std::move(v));
    d.data = keep_.back().data();
    return d;
  }

  // HostRSGTDUT::cellcenter (HostRSGTDUT.hpp:419-428).
  void cellCenter(int32_t c, double cc[3]) const {
    cc[0] = cc[1] = cc[2] = 0.;
    for (int kb = 0; kb < 8; kb++) {
      const int32_t v = cells_[(size_t)c * 8 + kb];
      cc[0] += 0.125 * x_[v];
      cc[1] += 0.125 * y_[v];
      cc[2] += 0.125 * z_[v];
    }
  }

  // Builds the device channels: one per table-driven source row entry, one per support
  // point of every other source (their amplitudes are evaluated here, per step).
  void commit() {
    if (!dirty_) return;
    kinds_.clear();
    srcs_.clear();
    opaque_.clear();
    std::vector<std::vector<int32_t>> supports;
    std::vector<const double*> sig_t, sig_v;
    std::vector<int32_t> sig_n;
    rsg_mesh_desc md{nn_, nc_, x_.data(), y_.data(), z_.data(), cells_.data()};
    for (int pass = 0; pass < 2; pass++) {
      const std::vector<Row>& rows = pass ? moment_ : force_;
      const int kind = pass ? RSG_MOMENT : RSG_FORCE;
      for (const Row& r : rows) {
        if (auto* nx = dynamic_cast<const NormalXDZ*>(r.mv.get())) {
          std::vector<int32_t> ids((size_t)std::max(nn_, nc_));
          int32_t cnt = 0;
          check(rsg_disc_support(&md, pass, nx->center(), nx->radius(), nx->ds(), ids.data(), &cnt));
          ids.resize(cnt);
          supports.push_back(std::move(ids));
          kinds_.push_back({kind, r.comp});
          srcs_.push_back({nx, nullptr, {0, 0, 0}});
          sig_n.push_back((int32_t)nx->signal().times().size());
          sig_t.push_back(nx->signal().times().data());
          sig_v.push_back(nx->signal().values().data());
          continue;
        }
        // opaque source: one single-point channel per accepted point, point order
        const int32_t npts = pass ? nn_ : nc_;
        for (int32_t p = 0; p < npts; p++) {
          double a[4];
          if (pass) {
            a[0] = x_[p];
            a[1] = y_[p];
            a[2] = z_[p];
          } else {
            cellCenter(p, a);
          }
          a[3] = 0.0;
          if (!r.mv->checkAt(a)) continue;
          supports.push_back({p});
          kinds_.push_back({kind, r.comp});
          srcs_.push_back({nullptr, r.mv.get(), {a[0], a[1], a[2]}});
          sig_n.push_back(0);
          sig_t.push_back(nullptr);
          sig_v.push_back(nullptr);
          opaque_.push_back(srcs_.size() - 1);
        }
      }
    }
    // Single-point channels of one row must keep the row's position in the term order:
    // the library orders a point's terms by channel index, and channels were appended row
    // by row, so rows stay in the reference's order and points never share a channel.
    std::vector<rsg_channel_desc> ch(kinds_.size());
    for (size_t k = 0; k < ch.size(); k++) {
      ch[k].kind = kinds_[k][0];
      ch[k].component = kinds_[k][1];
      ch[k].n_support = (int32_t)supports[k].size();
      ch[k].support = supports[k].data();
      ch[k].n_signal = sig_n[k];
      ch[k].signal_t = sig_t[k];
      ch[k].signal_v = sig_v[k];
    }
    check(rsg_set_channels(h_, (int32_t)ch.size(), ch.data()));
    dirty_ = false;
  }

  PanNDE::ElasticMaterialNames mat_;
  PanNDE::ElasticStateNames names_;
  rsg_handle h_ = nullptr;
  double dt_ = 0, time_ = 0;
  int32_t nn_ = 0, nc_ = 0;
  std::shared_ptr<PanNDE::FieldBundle> domain_, solution_;
  std::shared_ptr<PanNDE::Mesh> mesh_;
  std::vector<double> x_, y_, z_;
  std::vector<int32_t> cells_;
  std::vector<std::vector<double>> keep_;
  std::vector<Row> force_, moment_;
  std::vector<std::array<int, 2>> kinds_;
  std::vector<Src> srcs_;
  std::vector<size_t> opaque_;
  std::vector<double> amps_;
  bool dirty_ = true;
  size_t nsens_ = 0;
  int64_t trace_base_ = 0;
};

// Drop-in for HostSolver::HostRSGTDUT (HostRSGTDUT.hpp:100-146).  `precision` (RSG_FP64
// by default: the reference's arithmetic, bit for bit) is an addition at the end.
class DeviceRSGTDUT : public DeviceModel {
 public:
  static std::shared_ptr<DeviceRSGTDUT> makeShared(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
                                                   std::shared_ptr<PanNDE::FieldBundle> initial_condition,
                                                   std::shared_ptr<PanNDE::MultiVariate> transducers[][3],
                                                   int Nxducers = 0,
                                                   std::shared_ptr<PanNDE::Communicator> comm = nullptr,
                                                   int precision = RSG_FP64) {
    return std::make_shared<DeviceRSGTDUT>(dt, domain, initial_condition, transducers, Nxducers, comm,
                                           precision);
  }
  DeviceRSGTDUT(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
                std::shared_ptr<PanNDE::FieldBundle> initial_condition,
                std::shared_ptr<PanNDE::MultiVariate> transducer[3],
                std::shared_ptr<PanNDE::Communicator> comm = nullptr, int precision = RSG_FP64) {
    (void)comm;
    setup(dt, domain, initial_condition, 0, precision);
    addForceRow(transducer);
  }
  DeviceRSGTDUT(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
                std::shared_ptr<PanNDE::FieldBundle> initial_condition,
                std::shared_ptr<PanNDE::MultiVariate> transducers[][3], int Nxducers = 0,
                std::shared_ptr<PanNDE::Communicator> comm = nullptr, int precision = RSG_FP64) {
    (void)comm;
    setup(dt, domain, initial_condition, 0, precision);
    for (int k = 0; k < Nxducers; k++) addForceRow(transducers[k]);
  }
};

// Drop-in for HostSolver::HostUTModel (HostUTModel.hpp:100-170).
class DeviceUTModel : public DeviceModel {
 public:
  static std::shared_ptr<DeviceUTModel> makeShared(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
                                                   std::shared_ptr<PanNDE::FieldBundle> initial_condition,
                                                   std::shared_ptr<PanNDE::Communicator> comm = nullptr,
                                                   int precision = RSG_FP64) {
    return std::make_shared<DeviceUTModel>(dt, domain, initial_condition, comm, precision);
  }
  DeviceUTModel(double dt, std::shared_ptr<PanNDE::FieldBundle> domain,
                std::shared_ptr<PanNDE::FieldBundle> initial_condition,
                std::shared_ptr<PanNDE::Communicator> comm = nullptr, int precision = RSG_FP64) {
    (void)comm;
    setup(dt, domain, initial_condition, 1, precision);
  }
  void applyTransducer(std::shared_ptr<PanNDE::MultiVariate> xducer[3]) { addForceRow(xducer); }
  void applyMomentExcitation(std::shared_ptr<PanNDE::MultiVariate> moment_tensor[6]) {
    addMomentRow(moment_tensor);
  }
  void applyMomentExcitation(std::shared_ptr<PanNDE::MultiVariate> moment_tensor[3][3]) {
    std::shared_ptr<PanNDE::MultiVariate> row[6] = {moment_tensor[0][0], moment_tensor[1][1],
                                                    moment_tensor[2][2], moment_tensor[1][2],
                                                    moment_tensor[0][2], moment_tensor[0][1]};
    addMomentRow(row);
  }
};

}  // namespace DeviceSolver
