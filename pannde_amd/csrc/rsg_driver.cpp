// rsg_driver.cpp -- TransducerModel / WaterColumnModel style drivers of the DeviceSolver.
//
// Two ways in:
//
//   rsg_driver -f job.vtu -o out [--water-column] [--fp32] [--opaque] [--no-fields] [--ascii]
//              [--tables-dir DIR] [--partitioned] [--dump-case FILE]
//     The reference's case runners on the DeviceSolver.  Reads a VTK XML job (vtk_job.hpp):
//     the mesh, density and C11..C66 (PointData first, then CellData), dt and write_times, and
//     either the transducers XD<k>/{XCenter,YCenter,ZCenter,Radius,SignalTimes,SignalValues}
//     driving the z force (TransducerModel.cpp:160-204; Boss.hpp:502-536) or, with
//     --water-column, the water columns WC<k>/{...,SignalSigmaXX,SignalSigmaYY,SignalSigmaZZ}
//     applied as XX, YY, ZZ moment excitations of the HostUTModel form, column by column
//     (WaterColumnModel.cpp:85-180; Boss.hpp:555-650).  The window depth is the z-size of cell
//     0.  It builds DeviceSolver::DeviceRSGTDUT / DeviceUTModel as the host models are built
//     (TransducerModel.cpp:99-105, WaterColumnModel.cpp:127-135), starts from zero, and for
//     every write time calls time = model->solve() while time < write_time
//     (TransducerModel.cpp:63-83), then writes what the reference writes (VTKGateway.hpp:218-283):
//     <out>_<k>.vtu (Vx, Vy, Vz point data, Sxx..Sxy cell data, FieldData time), and for the
//     transducer model node_coords.vtt (the nodes any transducer's window accepts,
//     TransducerModel.cpp:106-130) and velocities_<k>.vtt (their velocities at the write time)
//     in the tables directory.  --partitioned writes the layout the reference writes on several
//     processes: .pvtu / .pvtt summaries, each with one piece <name>_0.vtu / .vtt.
//     --dump-case FILE writes the job as read (RSGCASE1, below) and exits without the GPU.
//
//   rsg_driver CASE OUT [--fp32] [--opaque] [--chunk K] [--time-from W]
//     Reads a job in the RSGCASE1 binary format (mesh, density and C11..C66 fields with their
//     NODE/CELL/CONSTANT type, dt, the force and moment discs with their signal tables, sensor
//     nodes, step count, optional initial state), steps with the drivers' loop (time =
//     model->solve() until the step count is reached) and writes an RSGOUT01 file: step count,
//     time, V, S (read back through getStates()) and the per-step sensor velocities.
//     --chunk K calls solve(K) instead of solve() per iteration; --time-from W: the printed
//     seconds cover steps W.. only (the first launches load the kernels' code objects); every
//     step still runs and is written out.
//
// --opaque hides the sources behind a plain PanNDE::MultiVariate wrapper, so the model takes its
// general host-evaluated source path (one channel per support point, evalAt every step).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "DeviceSolver/DeviceRSGTDUT.hpp"
#include "DeviceSolver/HostContainers.hpp"
#include "vtk_job.hpp"

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

struct FieldIn {
  int32_t type;  // 0 NODE, 1 CELL, 2 CONSTANT
  std::vector<double> v;
};

// A job as the model needs it: mesh, the 22 material fields in ElasticMaterialNames order
// (density, then C11..C66 row by row over the upper triangle), sources, sensors, steps.
struct Job {
  int32_t nn = 0, nc = 0;
  std::vector<double> x, y, z;
  std::vector<int32_t> cells;
  FieldIn fields[22];
  double dt = 0.0;
  int32_t moment_form = 0;
  std::vector<Disc> discs[2];  // force, moment
  std::vector<int32_t> sensors;
  int32_t nsteps = 0;
  std::vector<double> V0, S0;  // optional initial state
  std::vector<double> write_times;
};

const int kI[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
const int kJ[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};

Job read_case_file(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  Reader r{f};
  char magic[8];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "RSGCASE1", 8) != 0) {
    fclose(f);
    throw std::runtime_error("not an RSGCASE1 file");
  }
  Job j;
  j.nn = r.get<int32_t>();
  j.nc = r.get<int32_t>();
  j.x = r.vec<double>(j.nn);
  j.y = r.vec<double>(j.nn);
  j.z = r.vec<double>(j.nn);
  j.cells = r.vec<int32_t>((size_t)j.nc * 8);
  for (auto& fld : j.fields) {
    fld.type = r.get<int32_t>();
    fld.v = r.vec<double>(fld.type == 0 ? j.nn : (fld.type == 1 ? j.nc : 1));
  }
  j.dt = r.get<double>();
  j.moment_form = r.get<int32_t>();
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
      j.discs[pass].push_back(d);
    }
  }
  const int32_t nsens = r.get<int32_t>();
  j.sensors = r.vec<int32_t>(nsens);
  j.nsteps = r.get<int32_t>();
  if (r.get<int32_t>()) {
    j.V0 = r.vec<double>((size_t)3 * j.nn);
    j.S0 = r.vec<double>((size_t)6 * j.nc);
  }
  fclose(f);
  return j;
}

void write_case_file(const char* path, const Job& j) {
  FILE* f = fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + path);
  auto put = [&](const void* p, size_t bytes) {
    if (bytes && fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write failed");
  };
  auto i32 = [&](int32_t v) { put(&v, 4); };
  auto f64 = [&](double v) { put(&v, 8); };
  put("RSGCASE1", 8);
  i32(j.nn);
  i32(j.nc);
  for (auto* a : {&j.x, &j.y, &j.z}) put(a->data(), a->size() * 8);
  put(j.cells.data(), j.cells.size() * 4);
  for (auto& fld : j.fields) {
    i32(fld.type);
    put(fld.v.data(), fld.v.size() * 8);
  }
  f64(j.dt);
  i32(j.moment_form);
  for (auto& list : j.discs) {
    i32((int32_t)list.size());
    for (auto& d : list) {
      i32(d.comp);
      for (double c : d.center) f64(c);
      f64(d.radius);
      f64(d.ds);
      i32((int32_t)d.t.size());
      put(d.t.data(), d.t.size() * 8);
      put(d.v.data(), d.v.size() * 8);
    }
  }
  i32((int32_t)j.sensors.size());
  put(j.sensors.data(), j.sensors.size() * 4);
  i32(j.nsteps);
  i32(j.V0.empty() ? 0 : 1);
  if (!j.V0.empty()) {
    put(j.V0.data(), j.V0.size() * 8);
    put(j.S0.data(), j.S0.size() * 8);
  }
  fclose(f);
}

// Cumulative steps at each write time: while (time < t_k) time += dt, in FP64 as
// Model::solve accumulates it (TransducerModel.cpp:63-83).
std::vector<int64_t> schedule(double dt, const std::vector<double>& write_times) {
  std::vector<int64_t> out;
  double t = 0.0;
  int64_t n = 0;
  for (double tw : write_times) {
    while (t < tw) {
      t += dt;
      n++;
    }
    out.push_back(n);
  }
  return out;
}

Job read_vtu_job(const std::string& path, bool water_column) {
  const vtkjob::Vtu v = vtkjob::read_vtu(path);
  if (v.ncells == 0) throw std::runtime_error("the job has no cells");
  Job j;
  j.nn = v.npts;
  j.nc = v.ncells;
  j.x = v.x;
  j.y = v.y;
  j.z = v.z;
  j.cells = v.cells;
  PanNDE::ElasticMaterialNames mat;
  j.fields[0].type = v.field(mat.density, j.fields[0].v);
  for (int k = 0; k < 21; k++) j.fields[k + 1].type = v.field(mat.CIJ[kI[k]][kJ[k]], j.fields[k + 1].v);
  for (auto& fld : j.fields)
    if (fld.v.size() != (size_t)(fld.type == 0 ? j.nn : j.nc))
      throw std::runtime_error("a material field does not match the mesh");
  // window depth: the z-size of cell 0 (Boss.hpp:530-536, :639-645)
  const double ds = std::fabs(j.z[j.cells[0]] - j.z[j.cells[6]]);
  auto disc = [&](const std::string& p, int comp, const std::string& sig) {
    Disc d;
    d.comp = comp;
    d.center[0] = v.value(p + "XCenter");
    d.center[1] = v.value(p + "YCenter");
    d.center[2] = v.value(p + "ZCenter");
    d.radius = v.value(p + "Radius");
    d.ds = ds;
    d.t = v.array(p + "SignalTimes");
    d.v = v.array(p + sig);
    if (d.t.size() != d.v.size()) throw std::runtime_error(p + ": signal times and values differ in length");
    return d;
  };
  if (!water_column) {
    const int n = (int)v.value("NTransducers");
    for (int k = 0; k < n; k++) j.discs[0].push_back(disc("XD" + std::to_string(k) + "/", 2, "SignalValues"));
  } else {
    j.moment_form = 1;
    const int n = (int)v.value("NWaterColumns");
    static const char* kSig[3] = {"SignalSigmaXX", "SignalSigmaYY", "SignalSigmaZZ"};
    for (int k = 0; k < n; k++)
      for (int c = 0; c < 3; c++) j.discs[1].push_back(disc("WC" + std::to_string(k) + "/", c, kSig[c]));
  }
  j.dt = v.value("dt");
  j.write_times = v.array("write_times");
  const auto steps = schedule(j.dt, j.write_times);
  if (!steps.empty() && steps.back() > INT32_MAX) throw std::runtime_error("too many steps");
  j.nsteps = steps.empty() ? 0 : (int32_t)steps.back();
  if (!water_column) {
    // the recorded nodes: those any transducer's window accepts, in node order
    std::vector<std::shared_ptr<DeviceSolver::NormalXDZ>> win;
    for (auto& d : j.discs[0])
      win.push_back(std::make_shared<DeviceSolver::NormalXDZ>(
          d.center, d.radius, d.ds, std::make_shared<DeviceSolver::TableSignal>(d.t, d.v)));
    for (int32_t n = 0; n < j.nn; n++) {
      double p[3] = {j.x[n], j.y[n], j.z[n]};
      for (auto& w : win)
        if (w->checkAt(p)) {
          j.sensors.push_back(n);
          break;
        }
    }
  }
  return j;
}

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

struct Built {
  std::shared_ptr<DeviceSolver::HexMesh> mesh;
  std::shared_ptr<DeviceSolver::SimpleFieldBundle> solution;
  std::shared_ptr<DeviceSolver::DeviceModel> model;
};

// The domain and solution bundles and the model, constructed as the case runners construct
// the host models; zero initial condition as TransducerModel.cpp:150-159 emplaces it.
Built build_model(const Job& j, int precision, bool opaque) {
  Built b;
  b.mesh = std::make_shared<DeviceSolver::HexMesh>(j.x, j.y, j.z, j.cells);
  auto domain = std::make_shared<DeviceSolver::SimpleFieldBundle>(b.mesh);
  PanNDE::ElasticMaterialNames mat;
  for (int k = 0; k < 22; k++) {
    const std::string name = k == 0 ? mat.density : mat.CIJ[kI[k - 1]][kJ[k - 1]];
    domain->emplaceField(name, (PanNDE::Field::FieldType)j.fields[k].type);
    auto fld = domain->field(name);
    for (size_t i = 0; i < j.fields[k].v.size(); i++) fld->at((int)i) = j.fields[k].v[i];
  }
  PanNDE::ElasticStateNames st;
  b.solution = std::make_shared<DeviceSolver::SimpleFieldBundle>(b.mesh);
  for (int kd = 0; kd < 3; kd++) b.solution->emplaceField(st.V[kd], PanNDE::Field::NODE);
  for (int ki = 0; ki < 3; ki++)
    for (int kj = ki; kj < 3; kj++) b.solution->emplaceField(st.S[ki][kj], PanNDE::Field::CELL);
  if (!j.V0.empty()) {
    for (int k = 0; k < 3; k++)
      for (int32_t i = 0; i < j.nn; i++) b.solution->field(st.V[k])->at(i) = j.V0[(size_t)k * j.nn + i];
    for (int k = 0; k < 6; k++)
      for (int32_t i = 0; i < j.nc; i++) b.solution->field(st.SI[k])->at(i) = j.S0[(size_t)k * j.nc + i];
  }
  auto makeSource = [&](const Disc& d) -> std::shared_ptr<PanNDE::MultiVariate> {
    auto sig = std::make_shared<DeviceSolver::TableSignal>(d.t, d.v);
    auto src = std::make_shared<DeviceSolver::NormalXDZ>(d.center, d.radius, d.ds, sig);
    if (opaque) return std::make_shared<Opaque>(src);
    return src;
  };
  if (!j.moment_form) {
    if (!j.discs[1].empty()) throw std::runtime_error("moment sources need the HostUTModel form");
    std::vector<std::shared_ptr<PanNDE::MultiVariate>> flat(3 * std::max<size_t>(1, j.discs[0].size()));
    for (size_t i = 0; i < j.discs[0].size(); i++) flat[3 * i + j.discs[0][i].comp] = makeSource(j.discs[0][i]);
    auto xds = (std::shared_ptr<PanNDE::MultiVariate>(*)[3])flat.data();
    b.model = std::make_shared<DeviceSolver::DeviceRSGTDUT>(j.dt, domain, b.solution, xds, (int)j.discs[0].size(),
                                                            nullptr, precision);
  } else {
    auto ut = std::make_shared<DeviceSolver::DeviceUTModel>(j.dt, domain, b.solution, nullptr, precision);
    for (auto& d : j.discs[0]) {
      std::shared_ptr<PanNDE::MultiVariate> row[3] = {nullptr, nullptr, nullptr};
      row[d.comp] = makeSource(d);
      ut->applyTransducer(row);
    }
    for (auto& d : j.discs[1]) {
      std::shared_ptr<PanNDE::MultiVariate> row[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
      row[d.comp] = makeSource(d);
      ut->applyMomentExcitation(row);
    }
    b.model = ut;
  }
  return b;
}

int run_case_file(int argc, char** argv) {
  int precision = RSG_FP64, chunk = 1, time_from = 0;
  bool opaque = false;
  for (int k = 3; k < argc; k++) {
    std::string a = argv[k];
    if (a == "--fp32") precision = RSG_FP32;
    if (a == "--opaque") opaque = true;
    if (a == "--chunk" && k + 1 < argc) chunk = std::max(1, atoi(argv[++k]));
    if (a == "--time-from" && k + 1 < argc) time_from = std::max(0, atoi(argv[++k]));
  }
  const Job j = read_case_file(argv[1]);
  Built b = build_model(j, precision, opaque);
  auto& model = b.model;
  model->setSensors(j.sensors);

  auto t0 = std::chrono::steady_clock::now();
  double time = 0;
  int32_t done = 0, timed_from = 0;
  const int32_t nsteps = j.nsteps;
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

  PanNDE::ElasticStateNames st;
  FILE* o = fopen(argv[2], "wb");
  if (!o) throw std::runtime_error("cannot write output");
  const int32_t nsens = (int32_t)j.sensors.size();
  int32_t h[4] = {j.nn, j.nc, nsteps, nsens};
  fwrite("RSGOUT01", 1, 8, o);
  fwrite(h, sizeof(int32_t), 4, o);
  fwrite(&time, sizeof(double), 1, o);
  for (int k = 0; k < 3; k++) fwrite(states->field(st.V[k])->data(), sizeof(double), j.nn, o);
  for (int k = 0; k < 6; k++) fwrite(states->field(st.SI[k])->data(), sizeof(double), j.nc, o);
  fwrite(tr.data(), sizeof(double), tr.size(), o);
  fclose(o);
  printf("DRIVER steps=%d seconds=%.6f timed_steps=%d cells=%d nodes=%d precision=%s sources=%s\n", nsteps, secs,
         nsteps - timed_from, j.nc, j.nn, precision == RSG_FP64 ? "fp64" : "fp32", opaque ? "opaque" : "table");
  return 0;
}

int run_job(int argc, char** argv) {
  std::string job, out, tables_dir, dump;
  int precision = RSG_FP64;
  bool water = false, opaque = false, fields = true, ascii = false, partitioned = false;
  for (int k = 1; k < argc; k++) {
    const std::string a = argv[k];
    const bool more = k + 1 < argc;
    if (a == "-f" && more) job = argv[++k];
    else if (a == "-o" && more) out = argv[++k];
    else if (a == "--tables-dir" && more) tables_dir = argv[++k];
    else if (a == "--dump-case" && more) dump = argv[++k];
    else if (a == "--water-column") water = true;
    else if (a == "--fp32") precision = RSG_FP32;
    else if (a == "--opaque") opaque = true;
    else if (a == "--no-fields") fields = false;
    else if (a == "--ascii") ascii = true;
    else if (a == "--partitioned") partitioned = true;
    else throw std::runtime_error("unknown or incomplete option " + a);
  }
  if (job.empty() || (out.empty() && dump.empty())) throw std::runtime_error("-f job.vtu and -o out are required");
  Job j = read_vtu_job(job, water);
  printf("job %s: %d nodes, %d cells, %zu %s, dt %.6e, %zu write times, %d steps\n", job.c_str(), j.nn, j.nc,
         water ? j.discs[1].size() / 3 : j.discs[0].size(), water ? "water columns" : "transducers", j.dt,
         j.write_times.size(), j.nsteps);
  if (!dump.empty()) {  // the job as read, for a comparison with other readers; no GPU
    write_case_file(dump.c_str(), j);
    return 0;
  }
  const std::string t_ext = partitioned ? ".pvtt" : ".vtt", g_ext = partitioned ? ".pvtu" : ".vtu";
  const std::string tdir = tables_dir.empty() ? "" : (tables_dir.back() == '/' ? tables_dir : tables_dir + "/");
  auto write_table = [&](const std::string& path, const std::vector<vtkjob::Column>& cols) {
    partitioned ? vtkjob::write_pvtt(path, cols) : vtkjob::write_vtt(path, cols);
  };
  const std::vector<int32_t>& sens = j.sensors;
  if (!water) {
    std::vector<vtkjob::Column> c = {{"node_x", {}}, {"node_y", {}}, {"node_z", {}}};
    for (int32_t n : sens) {
      c[0].v.push_back(j.x[n]);
      c[1].v.push_back(j.y[n]);
      c[2].v.push_back(j.z[n]);
    }
    write_table(tdir + "node_coords" + t_ext, c);
  }
  Built b = build_model(j, precision, opaque);
  PanNDE::ElasticStateNames st;
  const auto t0 = std::chrono::steady_clock::now();
  double time = 0.0;
  int64_t done = 0;
  for (size_t kw = 0; kw < j.write_times.size(); kw++) {
    while (time < j.write_times[kw]) {  // TransducerModel.cpp:68-73
      time = b.model->solve();
      done++;
    }
    auto states = b.model->getStates();
    if (!water) {
      std::vector<vtkjob::Column> c = {{"Vx", {}}, {"Vy", {}}, {"Vz", {}}};
      for (int k = 0; k < 3; k++) {
        auto f = states->field(st.V[k]);
        for (int32_t n : sens) c[k].v.push_back(f->at(n));
      }
      write_table(tdir + "velocities_" + std::to_string(kw) + t_ext, c);
    }
    if (fields || water) {
      vtkjob::GridOut g{&j.x, &j.y, &j.z, &j.cells, {}, {}, time};
      for (int k = 0; k < 3; k++) {
        auto f = states->field(st.V[k]);
        g.point_data.push_back({st.V[k], std::vector<double>(f->data(), f->data() + j.nn)});
      }
      for (int k = 0; k < 6; k++) {
        auto f = states->field(st.SI[k]);
        g.cell_data.push_back({st.SI[k], std::vector<double>(f->data(), f->data() + j.nc)});
      }
      const std::string name = out + "_" + std::to_string(kw) + g_ext;
      partitioned ? vtkjob::write_pvtu(name, g, ascii) : vtkjob::write_vtu(name, g, ascii);
    }
    printf("t=%.6e saved write index %zu (%lld steps)\n", time, kw, (long long)done);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("done: %lld steps in %.3f s\n", (long long)done, secs);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const bool job_mode = argc >= 2 && std::strcmp(argv[1], "-f") == 0;
  if (!job_mode && argc < 3) {
    fprintf(stderr,
            "usage: rsg_driver -f job.vtu -o out [--water-column] [--fp32] [--opaque] [--no-fields] [--ascii]\n"
            "                  [--tables-dir DIR] [--partitioned] [--dump-case FILE]\n"
            "       rsg_driver CASE OUT [--fp32] [--opaque] [--chunk K] [--time-from W]\n");
    return 2;
  }
  try {
    return job_mode ? run_job(argc, argv) : run_case_file(argc, argv);
  } catch (const std::exception& e) {
    fprintf(stderr, "rsg_driver: %s\n", e.what());
    return 1;
  }
}
