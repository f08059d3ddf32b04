// rsg_driver.cpp -- a TransducerModel / WaterColumnModel style driver of the DeviceSolver.
//
// It reads a job in the RSGCASE1 binary format (mesh, density and C11..C66 fields with
// their NODE/CELL/CONSTANT type, dt, the force and moment discs with their signal
// tables, sensor nodes, step count, optional initial state), builds the mesh and fields
// in host containers, constructs DeviceSolver::DeviceRSGTDUT (no moment sources) or
// DeviceSolver::DeviceUTModel exactly as applications/source/CaseRunners/
// TransducerModel.cpp:99-105 and WaterColumnModel.cpp:127-135 construct the host models,
// steps with the drivers' loop (time = model->solve() until the step count is reached,
// TransducerModel.cpp:68-73), and writes an RSGOUT01 file: step count, time, V, S (read
// back through getStates()) and the per-step sensor velocities.
//
// Usage: rsg_driver CASE OUT [--fp32] [--opaque] [--chunk K] [--time-from W]
//   --opaque  hides the sources behind a plain PanNDE::MultiVariate wrapper, so the
//             model takes its general host-evaluated source path
//   --chunk K calls solve(K) instead of solve() per iteration
//   --time-from W  the printed seconds cover steps W.. only (the first launches load the
//             kernels' code objects); every step still runs and is written out
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "DeviceSolver/DeviceRSGTDUT.hpp"
#include "DeviceSolver/HostContainers.hpp"

namespace {

struct Reader {
  FILE* f;
  template <class T>
  T get() {
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("case file truncated");
    return v;
  }
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> out(n);
    if (n && fread(out.data(), sizeof(T), n, f) != n) throw std::runtime_error("case file truncated");
    return out;
  }
};

struct Disc {
  int32_t comp;
  double center[3], radius, ds;
  std::vector<double> t, v;
};

// Any MultiVariate: forwards evalAt/checkAt, hides the concrete type.
class Opaque : public PanNDE::MultiVariate {
 public:
  explicit Opaque(std::shared_ptr<PanNDE::MultiVariate> in) : in_(std::move(in)) {}
  double evalAt(std::shared_ptr<PanNDE::Array<double>> a) override { return in_->evalAt(a); }
  double evalAt(double a[]) override { return in_->evalAt(a); }
  bool checkAt(std::shared_ptr<PanNDE::Array<double>> a) override { return in_->checkAt(a); }
  bool checkAt(double a[]) override { return in_->checkAt(a); }
  int argumentCount() override { return 4; }
  std::shared_ptr<PanNDE::Array<double>> sampleArgument(int) override { return nullptr; }

 private:
  std::shared_ptr<PanNDE::MultiVariate> in_;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: rsg_driver CASE OUT [--fp32] [--opaque] [--chunk K] [--time-from W]\n");
    return 2;
  }
  int precision = RSG_FP64, chunk = 1, time_from = 0;
  bool opaque = false;
  for (int k = 3; k < argc; k++) {
    std::string a = argv[k];
    if (a == "--fp32") precision = RSG_FP32;
    if (a == "--opaque") opaque = true;
    if (a == "--chunk" && k + 1 < argc) chunk = std::max(1, atoi(argv[++k]));
    if (a == "--time-from" && k + 1 < argc) time_from = std::max(0, atoi(argv[++k]));
  }
  try {
    FILE* f = fopen(argv[1], "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    Reader r{f};
    char magic[8];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "RSGCASE1", 8) != 0)
      throw std::runtime_error("not an RSGCASE1 file");
    const int32_t nn = r.get<int32_t>(), nc = r.get<int32_t>();
    auto x = r.vec<double>(nn), y = r.vec<double>(nn), z = r.vec<double>(nn);
    auto cells = r.vec<int32_t>((size_t)nc * 8);
    auto mesh = std::make_shared<DeviceSolver::HexMesh>(x, y, z, cells);
    auto domain = std::make_shared<DeviceSolver::SimpleFieldBundle>(mesh);
    auto readField = [&](const std::string& name) {
      const int32_t type = r.get<int32_t>();
      const size_t len = type == 0 ? nn : (type == 1 ? nc : 1);
      auto v = r.vec<double>(len);
      domain->emplaceField(name, (PanNDE::Field::FieldType)type);
      auto fld = domain->field(name);
      for (size_t i = 0; i < len; i++) fld->at((int)i) = v[i];
    };
    PanNDE::ElasticMaterialNames mat;
    readField(mat.density);
    static const int kI[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
    static const int kJ[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};
    for (int k = 0; k < 21; k++) readField(mat.CIJ[kI[k]][kJ[k]]);
    const double dt = r.get<double>();
    const int32_t moment_form = r.get<int32_t>();
    std::vector<Disc> discs[2];
    for (int pass = 0; pass < 2; pass++) {
      const int32_t n = r.get<int32_t>();
      for (int32_t i = 0; i < n; i++) {
        Disc d;
        d.comp = r.get<int32_t>();
        for (int k = 0; k < 3; k++) d.center[k] = r.get<double>();
        d.radius = r.get<double>();
        d.ds = r.get<double>();
        const int32_t ns = r.get<int32_t>();
        d.t = r.vec<double>(ns);
        d.v = r.vec<double>(ns);
        discs[pass].push_back(d);
      }
    }
    const int32_t nsens = r.get<int32_t>();
    auto sens = r.vec<int32_t>(nsens);
    const int32_t nsteps = r.get<int32_t>();
    const int32_t has_init = r.get<int32_t>();
    std::vector<double> V0, S0;
    if (has_init) {
      V0 = r.vec<double>((size_t)3 * nn);
      S0 = r.vec<double>((size_t)6 * nc);
    }
    fclose(f);

    // zero initial condition, as TransducerModel.cpp:150-159 emplaces it
    PanNDE::ElasticStateNames st;
    auto solution = std::make_shared<DeviceSolver::SimpleFieldBundle>(mesh);
    for (int kd = 0; kd < 3; kd++) solution->emplaceField(st.V[kd], PanNDE::Field::NODE);
    for (int ki = 0; ki < 3; ki++)
      for (int kj = ki; kj < 3; kj++) solution->emplaceField(st.S[ki][kj], PanNDE::Field::CELL);
    if (has_init) {
      for (int k = 0; k < 3; k++)
        for (int32_t i = 0; i < nn; i++) solution->field(st.V[k])->at(i) = V0[(size_t)k * nn + i];
      for (int k = 0; k < 6; k++)
        for (int32_t i = 0; i < nc; i++) solution->field(st.SI[k])->at(i) = S0[(size_t)k * nc + i];
    }

    auto makeSource = [&](const Disc& d) -> std::shared_ptr<PanNDE::MultiVariate> {
      auto sig = std::make_shared<DeviceSolver::TableSignal>(d.t, d.v);
      auto src = std::make_shared<DeviceSolver::NormalXDZ>(d.center, d.radius, d.ds, sig);
      if (opaque) return std::make_shared<Opaque>(src);
      return src;
    };
    std::shared_ptr<DeviceSolver::DeviceModel> model;
    if (!moment_form) {
      if (!discs[1].empty()) throw std::runtime_error("moment sources need the HostUTModel form");
      std::vector<std::shared_ptr<PanNDE::MultiVariate>> flat(3 * std::max<size_t>(1, discs[0].size()));
      for (size_t i = 0; i < discs[0].size(); i++) flat[3 * i + discs[0][i].comp] = makeSource(discs[0][i]);
      auto xds = (std::shared_ptr<PanNDE::MultiVariate>(*)[3])flat.data();
      model = std::make_shared<DeviceSolver::DeviceRSGTDUT>(dt, domain, solution, xds, (int)discs[0].size(),
                                                            nullptr, precision);
    } else {
      auto ut = std::make_shared<DeviceSolver::DeviceUTModel>(dt, domain, solution, nullptr, precision);
      for (auto& d : discs[0]) {
        std::shared_ptr<PanNDE::MultiVariate> row[3] = {nullptr, nullptr, nullptr};
        row[d.comp] = makeSource(d);
        ut->applyTransducer(row);
      }
      for (auto& d : discs[1]) {
        std::shared_ptr<PanNDE::MultiVariate> row[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        row[d.comp] = makeSource(d);
        ut->applyMomentExcitation(row);
      }
      model = ut;
    }
    model->setSensors(sens);

    auto t0 = std::chrono::steady_clock::now();
    double time = 0;
    int32_t done = 0, timed_from = 0;
    while (done < nsteps) {  // the drivers' loop: time = model->solve()
      if (time_from > 0 && done == std::min(time_from, nsteps)) {  // timer starts here
        model->synchronize();
        t0 = std::chrono::steady_clock::now();
        timed_from = done;
      }
      const int32_t k = std::min(chunk, (done < time_from ? std::min(time_from, nsteps) : nsteps) - done);
      time = model->solve(k);
      done += k;
    }
    model->synchronize();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto states = model->getStates();
    auto tr = model->traces();

    FILE* o = fopen(argv[2], "wb");
    if (!o) throw std::runtime_error("cannot write output");
    int32_t h[4] = {nn, nc, nsteps, nsens};
    fwrite("RSGOUT01", 1, 8, o);
    fwrite(h, sizeof(int32_t), 4, o);
    fwrite(&time, sizeof(double), 1, o);
    for (int k = 0; k < 3; k++) fwrite(states->field(st.V[k])->data(), sizeof(double), nn, o);
    for (int k = 0; k < 6; k++) fwrite(states->field(st.SI[k])->data(), sizeof(double), nc, o);
    fwrite(tr.data(), sizeof(double), tr.size(), o);
    fclose(o);
    printf("DRIVER steps=%d seconds=%.6f timed_steps=%d cells=%d nodes=%d precision=%s sources=%s\n", nsteps, secs,
           nsteps - timed_from, nc, nn, precision == RSG_FP64 ? "fp64" : "fp32", opaque ? "opaque" : "table");
  } catch (const std::exception& e) {
    fprintf(stderr, "rsg_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
