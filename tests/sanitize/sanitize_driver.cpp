// sanitize_driver.cpp -- runs the host paths of libpannde_rsg (rsg_host.cpp) over the HIP
// stand-ins of hip_host_stub.cpp, for AddressSanitizer / UndefinedBehaviorSanitizer and
// ThreadSanitizer builds on a machine without a GPU.  TEST INFRASTRUCTURE ONLY.
//
// Covers: signal tables and disc supports; rsg_create on a holed, two-material box large
// enough for the multi-threaded set-up (node spacings, locality orders, class keys, LDS
// plans) in both precisions and several tunings, plus the error paths (DX, bad node id,
// bad tuning); channel and sensor tables, stepping bookkeeping across amplitude and trace
// buffers, state round trips, energy; batches with join, leave, graph and direct launches.
// Every device allocation must be released at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pannde_rsg.h"

extern "C" long stub_live_allocations();

namespace {

int failures = 0;

#define EXPECT(cond)                                                     \
  do {                                                                   \
    if (!(cond)) {                                                       \
      char msg_[512];                                                    \
      rsg_last_error(msg_, sizeof(msg_));                                \
      std::fprintf(stderr, "FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, msg_); \
      failures++;                                                        \
    }                                                                    \
  } while (0)

struct Box {
  int nx, ny, nz;
  double h;
  std::vector<double> x, y, z;
  std::vector<int32_t> cells;
  std::vector<double> rho, c[21];
  int32_t nn() const { return (int32_t)x.size(); }
  int32_t nc() const { return (int32_t)(cells.size() / 8); }
};

// nx*ny*nz cells of size h with a square hole through z (cells removed, orphan nodes dropped)
// and two materials (lower and upper half in z).
Box make_box(int nx, int ny, int nz, double h) {
  Box b{nx, ny, nz, h, {}, {}, {}, {}, {}, {}};
  auto hole = [&](int i, int j) { return i >= nx / 3 && i < nx / 3 + 4 && j >= ny / 2 && j < ny / 2 + 3; };
  std::vector<int32_t> id((size_t)(nx + 1) * (ny + 1) * (nz + 1), -1);
  auto nid = [&](int i, int j, int k) { return ((size_t)i * (ny + 1) + j) * (nz + 1) + k; };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        if (hole(i, j)) continue;
        const int v[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (auto& o : v) {
          int32_t& n = id[nid(i + o[0], j + o[1], k + o[2])];
          if (n < 0) {
            n = (int32_t)b.x.size();
            b.x.push_back((i + o[0]) * h);
            b.y.push_back((j + o[1]) * h);
            b.z.push_back((k + o[2]) * h);
          }
          b.cells.push_back(n);
        }
        const bool upper = k >= nz / 2;
        b.rho.push_back(upper ? 4430.0 : 2700.0);
        const double lam = upper ? 90e9 : 51e9, mu = upper ? 44e9 : 26e9;
        double m[6][6] = {};
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 3; q++) m[p][q] = lam + (p == q ? 2 * mu : 0);
        for (int p = 3; p < 6; p++) m[p][p] = mu;
        int t = 0;
        for (int p = 0; p < 6; p++)
          for (int q = p; q < 6; q++) b.c[t++].push_back(m[p][q]);
      }
  return b;
}

rsg_mesh_desc mesh_of(const Box& b) { return rsg_mesh_desc{b.nn(), b.nc(), b.x.data(), b.y.data(), b.z.data(), b.cells.data()}; }

rsg_model_desc model_of(const Box& b, double dt, int moment_form) {
  rsg_model_desc m;
  m.dt = dt;
  m.density = rsg_field_desc{b.rho.data(), RSG_FIELD_CELL};
  for (int k = 0; k < 21; k++) m.cij[k] = rsg_field_desc{b.c[k].data(), RSG_FIELD_CELL};
  m.moment_form = moment_form;
  return m;
}

void signal(std::vector<double>& t, std::vector<double>& v, double dt) {
  for (int i = 0; i < 40; i++) {
    t.push_back(i * dt);
    v.push_back(std::sin(0.3 * i));
  }
}

rsg_handle make_solver(const Box& b, int precision, const rsg_tuning* tun, int moment_form, int nsens) {
  const double dt = 1e-8;
  rsg_mesh_desc md = mesh_of(b);
  rsg_model_desc mo = model_of(b, dt, moment_form);
  rsg_handle h = nullptr;
  EXPECT(rsg_create_ex(&md, &mo, precision, tun, &h) == RSG_OK);
  if (!h) return nullptr;
  const double center[3] = {b.nx * b.h / 2, b.ny * b.h / 2, b.nz * b.h};
  std::vector<int32_t> sup(std::max(b.nc(), b.nn()));
  int32_t cnt = 0;
  std::vector<double> t, v;
  signal(t, v, dt);
  std::vector<rsg_channel_desc> ch;
  EXPECT(rsg_disc_support(&md, 0, center, 6 * b.h, b.h, sup.data(), &cnt) == RSG_OK && cnt > 0);
  std::vector<int32_t> csup(sup.begin(), sup.begin() + cnt);
  ch.push_back(rsg_channel_desc{RSG_FORCE, 2, cnt, csup.data(), (int32_t)t.size(), t.data(), v.data()});
  std::vector<int32_t> nsup;
  if (moment_form) {
    EXPECT(rsg_disc_support(&md, 1, center, 5 * b.h, b.h, sup.data(), &cnt) == RSG_OK && cnt > 0);
    nsup.assign(sup.begin(), sup.begin() + cnt);
    for (int comp = 0; comp < 3; comp++)
      ch.push_back(rsg_channel_desc{RSG_MOMENT, comp, cnt, nsup.data(), (int32_t)t.size(), t.data(), v.data()});
  }
  EXPECT(rsg_set_channels(h, (int32_t)ch.size(), ch.data()) == RSG_OK);
  std::vector<int32_t> sens;
  for (int s = 0; s < nsens; s++) sens.push_back((int32_t)((size_t)s * 7919 % b.nn()));
  EXPECT(rsg_set_sensors(h, nsens, sens.data()) == RSG_OK);
  return h;
}

void single_instance(const Box& b) {
  const rsg_tuning none{};
  rsg_tuning t[4] = {none, none, none, none};
  t[1].coeff_classes = -1;
  t[2].lds_velocity = -1;
  t[2].lds_stress = -1;
  t[3].per_thread_velocity = 4;
  t[3].per_thread_stress = 2;
  t[3].block_velocity = 512;
  for (int precision : {RSG_FP64, RSG_FP32})
    for (int k = 0; k < 4; k++) {
      rsg_handle h = make_solver(b, precision, &t[k], k % 2, 5);
      if (!h) continue;
      double time = 0;
      EXPECT(rsg_step(h, 300, nullptr, &time) == RSG_OK);  // two amplitude stages
      EXPECT(std::fabs(time - 300 * 1e-8) < 1e-14);
      std::vector<double> V((size_t)3 * b.nn()), S((size_t)6 * b.nc()), tr(300 * 5 * 3);
      EXPECT(rsg_get_state(h, V.data(), S.data()) == RSG_OK);
      EXPECT(rsg_set_state(h, V.data(), S.data()) == RSG_OK);
      EXPECT(rsg_get_traces(h, 0, 300, tr.data()) == RSG_OK);
      EXPECT(rsg_get_traces(h, 0, 301, tr.data()) == RSG_ERR_ARG);
      double ke = -1, se = -1;
      EXPECT(rsg_energy(h, &ke, &se) == RSG_OK);
      EXPECT(rsg_energy_noalloc(h, &ke, &se) == RSG_OK);
      // caller-supplied amplitudes
      std::vector<double> amp(20 * (k % 2 ? 4 : 1), 0.5);
      EXPECT(rsg_step(h, 20, amp.data(), nullptr) == RSG_OK);
      rsg_tuning got;
      EXPECT(rsg_get_tuning(h, &got) == RSG_OK);
      int64_t bytes = 0;
      EXPECT(rsg_device_bytes(h, &bytes) == RSG_OK && bytes > 0);
      EXPECT(rsg_destroy(h) == RSG_OK);
    }
}

void error_paths(const Box& b) {
  rsg_mesh_desc md = mesh_of(b);
  rsg_model_desc mo = model_of(b, 1e-8, 0);
  rsg_handle h = nullptr;
  // a node id out of range
  std::vector<int32_t> bad(b.cells);
  bad[8 * 10 + 3] = b.nn();
  rsg_mesh_desc mb = md;
  mb.cell_nodes = bad.data();
  EXPECT(rsg_create(&mb, &mo, RSG_FP64, &h) == RSG_ERR_ARG && h == nullptr);
  // a cell whose vertices share one x: a zero node spacing (RightHex::nodePrep "DX")
  {
    const double cx[8] = {0, 0, 0, 0, 0, 0, 0, 0}, cy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const int32_t cn[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    const double one = 1.0;
    rsg_mesh_desc m1{8, 1, cx, cy, cz, cn};
    rsg_model_desc o1;
    o1.dt = 1e-8;
    o1.density = rsg_field_desc{&one, RSG_FIELD_CONSTANT};
    for (int k = 0; k < 21; k++) o1.cij[k] = rsg_field_desc{&one, RSG_FIELD_CONSTANT};
    o1.moment_form = 0;
    char msg[256];
    EXPECT(rsg_create(&m1, &o1, RSG_FP64, &h) == RSG_ERR_MESH && h == nullptr);
    rsg_last_error(msg, sizeof(msg));
    EXPECT(std::strncmp(msg, "DX", 2) == 0);
  }
  rsg_tuning t{};
  t.block_stress = 96;
  EXPECT(rsg_create_ex(&md, &mo, RSG_FP64, &t, &h) == RSG_ERR_ARG && h == nullptr);
  // signal tables: a repeated key keeps its first value; a time before the first key fails
  const double ts[4] = {0.0, 1.0, 1.0, 2.0}, vs[4] = {0.0, 10.0, 99.0, 20.0};
  const double at[3] = {0.5, 1.0, 3.0};
  double out[3];
  EXPECT(rsg_signal_eval(4, ts, vs, 3, at, out) == RSG_OK && out[0] == 5.0 && out[1] == 10.0 && out[2] == 20.0);
  const double early = -1.0;
  EXPECT(rsg_signal_eval(4, ts, vs, 1, &early, out) == RSG_ERR_ARG);
}

void batches(const Box& b, const Box& small) {
  for (int precision : {RSG_FP64, RSG_FP32})
    for (int use_graph : {1, 0}) {
      rsg_batch bt = nullptr;
      EXPECT(rsg_batch_create(precision, 16, 0, use_graph, &bt) == RSG_OK);
      EXPECT(rsg_batch_set_option(bt, RSG_BATCH_LANES, 3) == RSG_OK);
      EXPECT(rsg_batch_set_option(bt, RSG_BATCH_GROUP_BYTES, 1 << 20) == RSG_OK);
      rsg_tuning t{};
      t.per_thread_velocity = 2;
      // the last two differ only in LDS staging (a staged member keeps no connectivity on
      // the device, so the two must not share a launch)
      rsg_tuning unstaged{};
      unstaged.lds_velocity = -1;
      unstaged.lds_stress = -1;
      rsg_handle h[4] = {make_solver(b, precision, nullptr, 0, 4), make_solver(small, precision, &t, 1, 3),
                         make_solver(small, precision, nullptr, 0, 0), make_solver(small, precision, &unstaged, 0, 2)};
      int32_t slot[4];
      const int64_t steps[4] = {70, 33, 100, 40};
      for (int i = 0; i < 2; i++) EXPECT(rsg_batch_add(bt, h[i], steps[i], &slot[i]) == RSG_OK);
      EXPECT(rsg_step(h[0], 1, nullptr, nullptr) == RSG_ERR_STATE);
      int32_t left = -1;
      EXPECT(rsg_batch_run(bt, 2, &left) == RSG_OK);
      for (int i = 2; i < 4; i++) EXPECT(rsg_batch_add(bt, h[i], steps[i], &slot[i]) == RSG_OK);  // join later
      while (left > 0) EXPECT(rsg_batch_run(bt, 1, &left) == RSG_OK);
      EXPECT(rsg_batch_run(bt, 3, &left) == RSG_OK && left == 0);
      EXPECT(rsg_batch_synchronize(bt) == RSG_OK);
      for (int i = 0; i < 4; i++) {
        int64_t done = 0, total = 0;
        EXPECT(rsg_batch_slot_status(bt, slot[i], &done, &total) == RSG_OK && done == steps[i] && total == steps[i]);
        EXPECT(rsg_batch_remove(bt, slot[i]) == RSG_OK);
      }
      EXPECT(rsg_batch_destroy(bt) == RSG_OK);
      for (auto x : h)
        if (x) EXPECT(rsg_destroy(x) == RSG_OK);
    }
  // one lane, one unbounded group budget: a staged and an unstaged member of the same
  // coefficient form (per-entity arrays) and workgroup shape still take separate launches
  for (int precision : {RSG_FP64, RSG_FP32}) {
    rsg_batch bt = nullptr;
    EXPECT(rsg_batch_create(precision, 8, 0, 1, &bt) == RSG_OK);
    EXPECT(rsg_batch_set_option(bt, RSG_BATCH_LANES, 1) == RSG_OK);
    EXPECT(rsg_batch_set_option(bt, RSG_BATCH_GROUP_BYTES, 0) == RSG_OK);
    rsg_tuning staged{}, unstaged{};
    staged.coeff_classes = unstaged.coeff_classes = -1;
    unstaged.lds_velocity = -1;
    unstaged.lds_stress = -1;
    rsg_handle h[2] = {make_solver(small, precision, &staged, 0, 1), make_solver(small, precision, &unstaged, 0, 1)};
    int32_t slot[2];
    for (int i = 0; i < 2; i++) EXPECT(rsg_batch_add(bt, h[i], 20, &slot[i]) == RSG_OK);
    int32_t left = -1;
    for (int q = 0; q < 8 && left != 0; q++) EXPECT(rsg_batch_run(bt, 1, &left) == RSG_OK);
    EXPECT(left == 0);
    for (int i = 0; i < 2; i++) EXPECT(rsg_batch_remove(bt, slot[i]) == RSG_OK);
    EXPECT(rsg_batch_destroy(bt) == RSG_OK);
    for (auto x : h) EXPECT(rsg_destroy(x) == RSG_OK);
  }
}

}  // namespace

// Cells that duplicate others, appended at the end (each of their vertex slots is claimed
// twice, on the set-up's threads): set-up, source tables with moments (whose special cells
// then come from every cell's vertices) and a few steps, in both precisions.
void duplicated_cells(Box b) {
  const int32_t nc = b.nc();
  for (int32_t k : {0, nc / 3, nc / 2, nc - 1}) {
    for (int v = 0; v < 8; v++) b.cells.push_back(b.cells[(size_t)k * 8 + v]);
    b.rho.push_back(b.rho[k]);
    for (int t = 0; t < 21; t++) b.c[t].push_back(b.c[t][k]);
  }
  for (int precision : {RSG_FP64, RSG_FP32}) {
    rsg_handle h = make_solver(b, precision, nullptr, 1, 3);
    if (!h) continue;
    EXPECT(rsg_step(h, 5, nullptr, nullptr) == RSG_OK);
    EXPECT(rsg_destroy(h) == RSG_OK);
  }
}

int main() {
  const Box big = make_box(64, 60, 40, 1e-4);    // > 128 Ki cells: several set-up threads
  const Box small = make_box(14, 12, 8, 1e-4);
  std::printf("box: %d cells, %d nodes\n", big.nc(), big.nn());
  single_instance(big);
  error_paths(small);
  batches(big, small);
  duplicated_cells(big);
  const long live = stub_live_allocations();
  if (live != 0) {
    std::fprintf(stderr, "FAILED: %ld device allocations still live at exit\n", live);
    failures++;
  }
  std::printf("sanitize driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
