// rsg_host.cpp -- host half of libpannde_rsg.so: the C ABI declared in
// include/pannde_rsg.h.
//
// Set-up (rsg_create) reproduces, in FP64 and in the reference's operation order, what
// HostRSGTDUT / HostUTModel compute before their first step:
//   * node-to-cell links, HexMesh::buildLinks (HexMesh.hpp:300-313): slot = vertex ^ 6;
//   * spacings, RightHex::cellPrep / nodePrep (RightHex.hpp:253-309), including the
//     mirrored boundary bound and the logic_error on a non-positive spacing;
//   * density at nodes and C_IJ at cells with HostField::atNode / atCell
//     (HostField.hpp:121-154): the /8 rule that counts missing cells as zero;
//   * dt/rho_n, the factor updateVsAtNode recomputes every step (HostRSGTDUT.hpp:343).
// It then reorders nodes and cells for locality (z, y, x lexicographic on quantised
// coordinates, so a wavefront's gathers hit a few contiguous runs) and uploads
// structure-of-arrays buffers.  Source channels become per-entity term lists in the
// reference's summation order (evalXds: connected-cell slot, then transducer,
// HostRSGTDUT.hpp:387-402; evalMoments: vertex, then excitation, HostUTModel.hpp:296-311).
#include <hip/hip_runtime_api.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/pannde_rsg.h"
#include "rsg_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// Set-up phase timing on stderr (wall time, and the process's CPU time over all threads),
// compiled in only with -DRSG_CREATE_TIMING.
#ifdef RSG_CREATE_TIMING
inline std::chrono::steady_clock::time_point& create_clock() {
  static thread_local std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  return t;
}
inline double process_cpu_ms() {
  timespec ts;
  clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}
inline double& create_cpu() {
  static thread_local double c = process_cpu_ms();
  return c;
}
#define CREATE_PHASE(name)                                                                        \
  do {                                                                                           \
    const auto t_now_ = std::chrono::steady_clock::now();                                        \
    const double c_now_ = process_cpu_ms();                                                      \
    fprintf(stderr, "rsg_create %-12s %8.1f ms wall %8.1f ms cpu\n", name,                       \
            std::chrono::duration<double, std::milli>(t_now_ - create_clock()).count(),          \
            c_now_ - create_cpu());                                                              \
    create_clock() = t_now_;                                                                     \
    create_cpu() = c_now_;                                                                       \
  } while (0)
#else
#define CREATE_PHASE(name) \
  do {                     \
  } while (0)
#endif

// A roctx range over an entry point's host work (rocprofv3 --marker-trace shows them next to
// the kernels): set-up, stepping, batch blocks, state and energy transfers.
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(RSG_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

// HostUnivariate::evalAt / interpolate (HostUnivariate.hpp:109-142) over m strictly
// ascending keys: the last value from the last key on, (Dv/Dt)*(t-t_prev)+v_prev below it;
// false below the first key (undefined in the reference).
bool signal_at(const double* t, const double* v, size_t m, double at, double* out) {
  if (!(at < t[m - 1])) {
    *out = v[m - 1];
    return true;
  }
  const size_t i = (size_t)(std::upper_bound(t, t + m, at) - t);
  if (i == 0) return false;
  const double Dt = t[i] - t[i - 1];
  const double Dv = v[i] - v[i - 1];
  *out = (Dv / Dt) * (at - t[i - 1]) + v[i - 1];
  return true;
}

// HostUnivariate (HostUnivariate.hpp:68-142): keys ascend as in a std::map, a repeated
// key keeps the value emplaced first; evaluation as signal_at.
struct Signal {
  std::vector<double> t, v;
  void build(int32_t n, const double* ts, const double* vs) {
    std::vector<int32_t> idx(n);
    for (int32_t i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return ts[a] < ts[b]; });
    t.clear();
    v.clear();
    for (int32_t i : idx) {
      if (!t.empty() && !(t.back() < ts[i])) continue;  // equal key: first one stays
      t.push_back(ts[i]);
      v.push_back(vs[i]);
    }
  }
  bool eval(double at, double* out) const { return signal_at(t.data(), v.data(), t.size(), at, out); }
};

struct Channel {
  int32_t kind, comp;
  std::vector<int32_t> support;
  bool has_signal = false;
  Signal sig;
};

// Device memory of the handles comes from a stream-ordered pool of the library's own, one per
// device.  hipFree waits for every stream of the device before it frees, so a process that
// destroys handles while batches keep other streams busy (a server) would wait about one K-step
// block per buffer; hipFreeAsync on the handle's own stream, idle by then, returns the memory
// to the pool without that wait, and later handles reuse it.  The pool keeps what is freed
// (its release threshold is unbounded) until the process ends.
hipError_t device_pool(hipMemPool_t* out) {
  static std::mutex mu;
  static std::vector<std::pair<int, hipMemPool_t>> pools;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  for (auto& dp : pools)
    if (dp.first == dev) {
      *out = dp.second;
      return hipSuccess;
    }
  hipMemPoolProps props;
  memset(&props, 0, sizeof(props));
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = dev;
  hipMemPool_t pool = nullptr;
  if ((e = hipMemPoolCreate(&pool, &props)) != hipSuccess) return e;
  uint64_t keep = UINT64_MAX;
  if ((e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) != hipSuccess) return e;
  pools.emplace_back(dev, pool);
  *out = pool;
  return hipSuccess;
}

// `bytes` of device memory from the pool, ordered on stream s (of the current device): work
// queued on s may use it at once, the host and other streams once s has been synchronized.
hipError_t device_alloc(void** p, size_t bytes, hipStream_t s) {
  hipMemPool_t pool = nullptr;
  hipError_t e = device_pool(&pool);
  if (e == hipSuccess) e = hipMallocFromPoolAsync(p, bytes, pool, s);
  return e;
}

// Returns pool memory once the work queued on s so far has run; the caller guarantees that no
// other stream still uses it.
void device_free(void* p, hipStream_t s) { (void)hipFreeAsync(p, s); }

// A copy ordered on s (a non-blocking stream) that has finished when this returns.  hipMemcpy
// runs on the null stream and waits for the work queued on every stream, so while batches keep
// other streams busy a handle's uploads would wait behind them.
hipError_t copy_on(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

// A device buffer owned by its handle, from the device pool on the handle's stream: released
// on reallocation, and by the handle before it destroys that stream.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int alloc(size_t count, hipStream_t stream) {
    release();
    n = count;
    if (count == 0) return RSG_OK;
    hipError_t e = device_alloc((void**)&p, count * sizeof(T), stream);
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
      return fail(RSG_ERR_NOMEM, "device allocation (%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    s = stream;
    return RSG_OK;
  }
  void release() {
    if (p) device_free(p, s);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

constexpr int kBlockSteps = 256;   // steps whose amplitudes are staged per upload
constexpr int kTraceCapSteps = 4096;

}  // namespace

struct rsg_instance {
  int32_t precision = RSG_FP64;
  int32_t device = 0;  // the GPU the handle's memory and stream live on
  int32_t nn = 0, nc = 0;
  int32_t moment_form = 0;
  double dt = 0, time = 0;
  int64_t steps = 0;
  int block_v = 256, block_s = 256;
  int cpt_v = 1, cpt_s = 1;       // entities per thread
  bool in_batch = false;          // stepped by a batch, not by rsg_step
  hipStream_t batch_stream = nullptr;  // the batch's stream while in_batch
  int xcd = 1;                    // XCD-contiguous workgroup remap
  int lds_v = 1, lds_s = 1;       // stage gathers through LDS when the block sets fit
  int c_blocked = 1;              // C as [n/64][21][64] instead of [21][n]
  int classes = 1;                // coefficient class tables when the rows repeat
  int tile[3] = {8, 8, 4};        // locality brick (x, y, z) in cells

  // permutations: internal index -> caller index, and the inverse
  std::vector<int32_t> node_of, cell_of, node_at, cell_at;
  // caller-numbered connectivity (kept for source term lists)
  std::vector<int32_t> cells;       // [nc][8]
  std::vector<int32_t> node_cells;  // [nn][8]
  // FP64 host copies in internal order of the values that get NaN-boxed
  std::vector<double> idt_h, dcx_h;

  // device state and coefficients (T = double or float)
  void* V[3] = {nullptr, nullptr, nullptr};
  void* S[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void* idt = nullptr;
  void* dn[3] = {nullptr, nullptr, nullptr};
  void* C = nullptr;
  int64_t ldc = 0;
  void* dc[3] = {nullptr, nullptr, nullptr};
  int32_t* ncell = nullptr;
  int32_t* cnode = nullptr;
  // coefficient class tables (null keys: per-entity arrays above)
  uint32_t* nkey = nullptr;
  void* nmat = nullptr;   // [materials] dt/rho
  void* naxis = nullptr;  // [3][kAxisClasses] node spacings
  int32_t n_nmat = -1;
  uint32_t* ckey = nullptr;
  void* cmat = nullptr;   // [materials][kMatRow] C
  void* caxis = nullptr;  // [3][kAxisClasses] cell spacings
  int32_t n_cmat = -1;
  std::vector<uint32_t> nkey_h, ckey_h;  // keys without the special flag, internal order
  // LDS staging plans (null lconn: direct gathers)
  int32_t* v_uoff = nullptr;
  int32_t* v_ulist = nullptr;
  int32_t v_chunk = 0;  // velocity gather lists in chunk words
  uint16_t* v_lconn = nullptr;
  uint32_t* v_lkey = nullptr;           // [n][4] 10-bit links + key (keys with staged sets <= 1022)
  std::vector<uint16_t> v_lconn_h;      // host links for packing v_lkey
  int32_t v_maxu = 0;
  int32_t* s_uoff = nullptr;
  int32_t* s_ulist = nullptr;
  uint16_t* s_lconn = nullptr;
  uint32_t* s_lkey = nullptr;
  std::vector<uint16_t> s_lconn_h;
  int32_t s_maxu = 0;
  // entities whose key on the device carries the special flag (internal ids, ascending): the
  // next table build restores their plain keys unless they stay special
  std::vector<int32_t> v_flagged, s_flagged;
  bool incidence_exact = true;  // node_cells holds every cell of every node (no slot claimed twice)
  double* node_m = nullptr;    // node masses rho_n*dxn*dyn*dzn (energy)
  double* cell_vol = nullptr;  // cell volumes dxc*dyc*dzc (energy)

  // source / sensor tables
  std::vector<Channel> channels;
  std::vector<int32_t> sensors;
  bool tables_dirty = true;
  int32_t n_sp_node = 0, n_sp_cell = 0;
  DevBuf<char> sp_node_val;   // T[nsp]
  DevBuf<int32_t> sp_node_off;
  DevBuf<int32_t> sp_node_sens;
  DevBuf<int32_t> node_terms;
  DevBuf<char> sp_cell_val;   // T[nsp]
  DevBuf<int32_t> sp_node_key, sp_cell_key;  // full key of each special entity (class tables)
  DevBuf<int32_t> sp_cell_off;
  DevBuf<int32_t> cell_terms;
  DevBuf<double> amp[2];
  double* stage[2] = {nullptr, nullptr};  // pinned host staging
  size_t stage_len[2] = {0, 0};            // doubles each staging buffer holds
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_busy[2] = {false, false};
  int cur = 0;
  // amplitudes of the library's signal tables uploaded a block of kBlockSteps steps ahead:
  // amp[win_buf] holds those of steps [win_first, win_first + win_len), so calls of a few
  // steps each (a driver's solve() per step) read them without a copy of their own
  int64_t win_first = -1;
  int32_t win_len = 0;
  int win_buf = 0;
  DevBuf<double> trace;
  int64_t trace_dev_steps = 0;       // rows currently held on the device
  std::vector<double> trace_host;    // rows flushed to the host
  DevBuf<double> partial;  // energy: [2][kEnergyMaxBlocks] partials, then the two totals

  // timing
  int32_t timing = 0;  // 0: off; n > 0: every n-th step's two launches are timed
  std::vector<hipEvent_t> tev;
  double t_vel_ms = 0, t_str_ms = 0;
  int64_t t_launches = 0;

  hipStream_t stream = nullptr;
  std::vector<void*> allocations;
  int64_t dev_bytes = 0;

  size_t elt() const { return precision == RSG_FP64 ? 8 : 4; }

  std::mutex alloc_mu;  // set-up allocates from two threads

  // host sources of the copies queued on `stream` since it was last synchronized; a set-up
  // queues all its uploads and waits once (finish_uploads)
  std::mutex hold_mu;
  std::vector<std::shared_ptr<const void>> held;

  // device memory from the pool on the handle's stream, released with the handle
  int dalloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = device_alloc(p, bytes, stream);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(RSG_ERR_NOMEM, "device allocation (%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lock(alloc_mu);
    allocations.push_back(*p);
    dev_bytes += (int64_t)bytes;
    return RSG_OK;
  }

  ~rsg_instance() {
    // every buffer is freed on the handle's stream, in order behind its own work; a handle
    // in a batch cannot be destroyed, and leaving the batch waited for the lane's work on it
    for (void* p : allocations) device_free(p, stream);
    sp_node_val.release();
    sp_node_off.release();
    sp_node_sens.release();
    node_terms.release();
    sp_cell_val.release();
    sp_node_key.release();
    sp_cell_key.release();
    sp_cell_off.release();
    cell_terms.release();
    amp[0].release();
    amp[1].release();
    trace.release();
    partial.release();
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; i++) {
      if (stage[i]) (void)hipHostFree(stage[i]);
      if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
    }
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int field_len(const rsg_field_desc& f, int32_t nn, int32_t nc) {
  return f.type == RSG_FIELD_NODE ? nn : (f.type == RSG_FIELD_CELL ? nc : 1);
}

int check_field(const rsg_field_desc& f, const char* name) {
  if (!f.data) return fail(RSG_ERR_ARG, "field %s: null data", name);
  if (f.type < RSG_FIELD_NODE || f.type > RSG_FIELD_CONSTANT)
    return fail(RSG_ERR_ARG, "field %s: type %d is not NODE/CELL/CONSTANT", name, f.type);
  return RSG_OK;
}

// HostField::atNode (HostField.hpp:142-154).
__attribute__((always_inline)) inline double field_at_node(const rsg_field_desc& f, const int32_t* ncells, int32_t n) {
  if (f.type == RSG_FIELD_NODE) return f.data[n];
  double v = 0.0;
  for (int k = 0; k < 8; k++) {
    const int32_t c = ncells[(size_t)n * 8 + k];
    v += (-1 == c) ? 0. : (f.type == RSG_FIELD_CONSTANT ? f.data[0] : f.data[c]);
  }
  return v / 8.0;
}

// HostField::atCell (HostField.hpp:121-133).
__attribute__((always_inline)) inline double field_at_cell(const rsg_field_desc& f, const int32_t* cells, int32_t c) {
  if (f.type == RSG_FIELD_NODE) {
    double v = 0.0;
    for (int k = 0; k < 8; k++) v += f.data[cells[(size_t)c * 8 + k]];
    return v / 8.0;
  }
  return f.type == RSG_FIELD_CONSTANT ? f.data[0] : f.data[c];
}

// Threads of one set-up: 8, or PANNDE_RSG_SETUP_THREADS (1..64), read once per process.  A
// process that sets up many handles at once (a server's preparation threads) may spend less
// CPU in total with fewer threads per set-up.  Results do not depend on it.
int setup_threads() {
  static const int t = [] {
    const char* e = std::getenv("PANNDE_RSG_SETUP_THREADS");
    const int v = e ? std::atoi(e) : 0;
    return v >= 1 && v <= 64 ? v : 8;
  }();
  return t;
}

// Set-up loops over n entities run in up to setup_threads() contiguous chunks of at least
// 64 Ki entities.
int chunk_count(int64_t n) {
  constexpr int64_t kMinChunk = 1 << 16;
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<int64_t>(
      1, std::min<int64_t>(std::min<int64_t>(setup_threads(), hw), (n + kMinChunk - 1) / kMinChunk));
}

// Runs fn(w) for w in [0, t), one thread each (the caller takes w = 0); an exception of any
// task (std::bad_alloc) is rethrown in the caller once all have ended.
template <typename F>
void parallel_tasks(int t, F&& fn) {
  std::vector<std::exception_ptr> errs((size_t)t);
  std::vector<std::thread> workers;
  for (int w = 1; w < t; w++)
    workers.emplace_back([&fn, &errs, w]() {
      try {
        fn(w);
      } catch (...) {
        errs[w] = std::current_exception();
      }
    });
  try {
    fn(0);
  } catch (...) {
    errs[0] = std::current_exception();
  }
  for (auto& w : workers) w.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
}

// Runs fn(chunk, begin, end) for chunk_count(n) contiguous chunks of [0, n), one thread each.
// The chunks must write disjoint data.
template <typename F>
void parallel_chunks(int64_t n, F&& fn) {
  const int t = chunk_count(n);
  const int64_t step = (n + t - 1) / t;
  parallel_tasks(t, [&fn, n, step](int w) { fn(w, std::min(n, w * step), std::min(n, (w + 1) * step)); });
}

template <typename F>
void parallel_ranges(int64_t n, F&& fn) {
  parallel_chunks(n, [&fn](int, int64_t b, int64_t e) { fn(b, e); });
}

// A coordinate quantised to the finest spacing: floor((p - lo) / h + 0.5) clamped to
// [0, 2^21 - 1]; a NaN coordinate quantises to 0.  Monotonic in p.
inline uint32_t quantise(double p, double lo, double h) {
  double s = std::floor((p - lo) / h + 0.5);
  if (!(s >= 0)) s = 0;
  if (s > 2097151.0) s = 2097151.0;
  return (uint32_t)s;
}

// Locality order: the points quantised, (qx, qy, qz), sorted by bricks of tile[0] x tile[1] x
// tile[2] (brick z, brick y, brick x, then z, y, x inside the brick), equal keys in index order.
// tile = {huge, 1, 1} is plain (z, y, x) lexicographic order: x-runs are contiguous, so each
// wavefront's gathers touch a few short runs.
std::vector<int32_t> locality_order(int32_t n, const double* px, const double* py, const double* pz,
                                    const double h[3], const double lo[3], const int tile[3]) {
  const double* P[3] = {px, py, pz};
  const int nt = chunk_count(n);
  // Extents: quantisation is monotonic, so the largest quantised value along an axis is that
  // of the largest coordinate (NaN coordinates, which quantise to 0, never compare larger).
  std::vector<std::array<double, 3>> max_of((size_t)nt, {-INFINITY, -INFINITY, -INFINITY});
  parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
    for (int d = 0; d < 3; d++) {
      double v = -INFINITY;
      for (int64_t i = b; i < e; i++)
        if (P[d][i] > v) v = P[d][i];
      max_of[w][d] = v;
    }
  });
  uint64_t ext[3] = {1, 1, 1};
  for (int d = 0; d < 3; d++) {
    double v = -INFINITY;
    for (auto& m : max_of)
      if (m[d] > v) v = m[d];
    ext[d] = (uint64_t)quantise(v, lo[d], h[d]) + 1;
  }
  // brick (z, y, x), then (z, y, x) inside the brick; both fit 64 bits for extents < 2^21
  const uint64_t t[3] = {std::min<uint64_t>((uint64_t)tile[0], ext[0]),
                         std::min<uint64_t>((uint64_t)tile[1], ext[1]),
                         std::min<uint64_t>((uint64_t)tile[2], ext[2])};
  const uint64_t nb[3] = {(ext[0] + t[0] - 1) / t[0], (ext[1] + t[1] - 1) / t[1], (ext[2] + t[2] - 1) / t[2]};
  const uint64_t vol = t[0] * t[1] * t[2];
  // quantised values are below 2^21, so 32-bit division suffices; power-of-two tiles (the
  // default) divide by shifting
  const bool pow2 = ((t[0] & (t[0] - 1)) | (t[1] & (t[1] - 1)) | (t[2] & (t[2] - 1))) == 0;
  int sh[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++)
    while ((1ull << sh[d]) < t[d]) sh[d]++;
  const uint32_t t32[3] = {(uint32_t)t[0], (uint32_t)t[1], (uint32_t)t[2]};
  auto key_of = [&](int64_t i, uint64_t& brick, uint64_t& inner) {
    const uint32_t qx = quantise(px[i], lo[0], h[0]), qy = quantise(py[i], lo[1], h[1]),
                   qz = quantise(pz[i], lo[2], h[2]);
    uint32_t b[3], r[3];
    if (pow2) {
      b[0] = qx >> sh[0], b[1] = qy >> sh[1], b[2] = qz >> sh[2];
      r[0] = qx & (t32[0] - 1), r[1] = qy & (t32[1] - 1), r[2] = qz & (t32[2] - 1);
    } else {
      b[0] = qx / t32[0], b[1] = qy / t32[1], b[2] = qz / t32[2];
      r[0] = qx - b[0] * t32[0], r[1] = qy - b[1] * t32[1], r[2] = qz - b[2] * t32[2];
    }
    brick = ((uint64_t)b[2] * nb[1] + b[1]) * nb[0] + b[0];
    inner = ((uint64_t)r[2] * t[1] + r[1]) * t[0] + r[0];
  };
  std::vector<int32_t> order((size_t)n);
  const long double span = (long double)(nb[0] * t[0]) * (long double)(nb[1] * t[1]) * (long double)(nb[2] * t[2]);
  const int64_t step = ((int64_t)n + nt - 1) / nt;
  // One stable LSD radix pass over the chunks on the digit (word >> shift) & (bins - 1): every
  // chunk counts its digits, chunk-major prefix sums give each chunk its output range per
  // digit, and each chunk scatters its words in order, so a pass produces what a serial stable
  // pass produces.
  auto radix_pass = [&](auto& src, auto& dst, auto&& digit, size_t bins, std::vector<int32_t>& count, auto&& move) {
    count.assign(bins * nt, 0);
    parallel_tasks(nt, [&](int w) {
      int32_t* c = &count[bins * w];
      const int64_t b = std::min<int64_t>(n, w * step), e = std::min<int64_t>(n, (w + 1) * step);
      for (int64_t i = b; i < e; i++) c[digit(src[i])]++;
    });
    int32_t sum = 0;
    for (size_t dgt = 0; dgt < bins; dgt++)
      for (int w = 0; w < nt; w++) {
        int32_t& c = count[bins * w + dgt];
        const int32_t v = c;
        c = sum;
        sum += v;
      }
    parallel_tasks(nt, [&](int w) {
      int32_t* c = &count[bins * w];
      const int64_t b = std::min<int64_t>(n, w * step), e = std::min<int64_t>(n, (w + 1) * step);
      for (int64_t i = b; i < e; i++) move(i, c[digit(src[i])]++);
    });
    src.swap(dst);
  };
  std::vector<int32_t> count;
  if (span <= 4294967296.0L) {
    // keys below 2^32: one word per point, (key << 32) | index, sorted on 8-bit digits of the
    // key (few output streams per chunk, so the scatter stays in cache)
    std::vector<uint64_t> a((size_t)n), a2((size_t)n);
    std::vector<uint64_t> kmax_of((size_t)nt, 0);
    parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
      uint64_t m = 0;
      for (int64_t i = b; i < e; i++) {
        uint64_t brick, inner;
        key_of(i, brick, inner);
        const uint64_t k = brick * vol + inner;
        a[i] = (k << 32) | (uint64_t)i;
        m = std::max(m, k);
      }
      kmax_of[w] = m;
    });
    const uint64_t kmax = *std::max_element(kmax_of.begin(), kmax_of.end());
    for (int shift = 0; shift < 32 && (kmax >> shift) != 0; shift += 8) {
      const int sh = 32 + shift;
      radix_pass(a, a2, [sh](uint64_t x) { return (size_t)((x >> sh) & 255); }, 256, count,
                 [&](int64_t i, int32_t dst) { a2[dst] = a[i]; });
    }
    parallel_ranges(n, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) order[i] = (int32_t)(uint32_t)a[i];
    });
    return order;
  }
  if (span < 9.2e18L) {
    // one 64-bit key per point, brick * (brick volume) + inner < span, and the index beside it
    std::vector<uint64_t> k((size_t)n), k2((size_t)n);
    std::vector<int32_t> o2((size_t)n);
    std::vector<uint64_t> kmax_of((size_t)nt, 0);
    parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
      uint64_t m = 0;
      for (int64_t i = b; i < e; i++) {
        uint64_t brick, inner;
        key_of(i, brick, inner);
        k[i] = brick * vol + inner;
        m = std::max(m, k[i]);
        order[i] = (int32_t)i;
      }
      kmax_of[w] = m;
    });
    const uint64_t kmax = *std::max_element(kmax_of.begin(), kmax_of.end());
    for (int shift = 0; shift < 64 && (kmax >> shift) != 0; shift += 11) {
      radix_pass(k, k2, [shift](uint64_t x) { return (size_t)((x >> shift) & 2047); }, 2048, count,
                 [&](int64_t i, int32_t dst) {
                   k2[dst] = k[i];
                   o2[dst] = order[i];
                 });
      order.swap(o2);
    }
    return order;
  }
  std::vector<std::pair<std::pair<uint64_t, uint64_t>, int32_t>> key((size_t)n);
  for (int32_t i = 0; i < n; i++) {
    uint64_t brick, inner;
    key_of(i, brick, inner);
    key[i] = {{brick, inner}, i};
  }
  std::sort(key.begin(), key.end());
  for (int32_t i = 0; i < n; i++) order[i] = key[i].second;
  return order;
}

// LDS staging plan for a gather kernel launched with `block` threads per workgroup:
// workgroup b handles entities [b*block, (b+1)*block); its unique gather set (sorted
// internal ids) is ulist[u_off[b]..u_off[b+1]) and each entity's 8 links become 16-bit
// indices into that set (0xFFFF for a missing link).  With `chunked` the set is cut into runs
// of at most 4 consecutive ids, each one list word start << 2 | (len - 1) and 4 LDS slots, and
// the links and max_u count slots (a quarter of the list bytes; padding slots stay unused).
struct LdsPlan {
  std::vector<int32_t> u_off, ulist;
  std::vector<uint16_t> lconn;
  int32_t max_u = 0;
};

bool build_lds_plan(const std::vector<int32_t>& conn, int32_t n, int block, LdsPlan& p,
                    bool chunked = false) {
  const int32_t nb = (n + block - 1) / block;
  p.u_off.assign((size_t)nb + 1, 0);
  p.ulist.clear();
  p.lconn.assign((size_t)n * 8, 0xFFFF);
  p.max_u = 0;
  // contiguous ranges of blocks on their own threads, each with its own tag array: the block
  // that last saw an entity (high 32 bits) and the entity's index in that block's set (low
  // bits), so finding and numbering a link touches one word.  A range's tag array covers only
  // the window of ids its links reach (in locality order a narrow one).  The per-range lists
  // are joined in block order.
  struct Range {
    std::vector<int32_t> ulist;
    int32_t max_u = 0;
    bool ok = true;
  };
  std::vector<Range> ranges(chunk_count(n));
  const int64_t nr = (int64_t)ranges.size(), step = (nb + nr - 1) / nr;
  parallel_tasks((int)nr, [&](int w) {
    Range& r = ranges[w];
    const int32_t b0 = (int32_t)std::min<int64_t>(nb, w * step), b1 = (int32_t)std::min<int64_t>(nb, (w + 1) * step);
    const size_t e0 = (size_t)b0 * block * 8, e1 = (size_t)std::min<int64_t>(n, (int64_t)b1 * block) * 8;
    int32_t w_lo = INT32_MAX, w_hi = -1;
    for (size_t i = e0; i < e1; i++) {
      const int32_t g = conn[i];
      if (g >= 0) {
        w_lo = std::min(w_lo, g);
        w_hi = std::max(w_hi, g);
      }
    }
    if (w_hi < 0) w_lo = w_hi = 0;  // no links: every set is empty
    std::vector<uint64_t> tag((size_t)(w_hi - w_lo) + 1, ~0ull);
    std::vector<int32_t> ids;
    ids.reserve((size_t)block * 8);
    r.ulist.reserve((size_t)(b1 - b0) * block * 2);
    for (int32_t b = b0; b < b1; b++) {
      const int32_t lo = b * block, hi = std::min(n, lo + block);
      const uint64_t stamp = (uint64_t)(uint32_t)b << 32;
      ids.clear();
      int32_t s_lo = INT32_MAX, s_hi = -1;
      for (size_t i = (size_t)lo * 8; i < (size_t)hi * 8; i++) {
        const int32_t g = conn[i];
        if (g >= 0 && (tag[g - w_lo] & ~0xFFFFFFFFull) != stamp) {
          tag[g - w_lo] = stamp;
          ids.push_back(g);
          s_lo = std::min(s_lo, g);
          s_hi = std::max(s_hi, g);
        }
      }
      // the set in increasing order: read off the tags when it is dense in its span, else sorted
      if (!ids.empty() && (int64_t)s_hi - s_lo < 16 * (int64_t)ids.size()) {
        size_t k = 0;
        for (int32_t g = s_lo; g <= s_hi; g++)
          if ((tag[g - w_lo] & ~0xFFFFFFFFull) == stamp) ids[k++] = g;
      } else {
        std::sort(ids.begin(), ids.end());
      }
      int32_t slots = 0, words = 0;
      if (chunked) {
        for (size_t k = 0; k < ids.size();) {
          if (ids[k] >= (1 << 30)) {  // start << 2 must fit 32 bits
            r.ok = false;
            return;
          }
          int32_t len = 1;
          while (len < 4 && k + len < ids.size() && ids[k + len] == ids[k] + len) len++;
          for (int32_t o = 0; o < len; o++) tag[ids[k + o] - w_lo] = stamp | (uint64_t)(slots + o);
          r.ulist.push_back((ids[k] << 2) | (len - 1));
          slots += 4;
          words++;
          k += len;
        }
      } else {
        for (size_t k = 0; k < ids.size(); k++) tag[ids[k] - w_lo] = stamp | k;
        r.ulist.insert(r.ulist.end(), ids.begin(), ids.end());
        slots = words = (int32_t)ids.size();
      }
      if (slots >= 0xFFFF) {
        r.ok = false;
        return;
      }
      for (size_t i = (size_t)lo * 8; i < (size_t)hi * 8; i++)
        if (conn[i] >= 0) p.lconn[i] = (uint16_t)(tag[conn[i] - w_lo] & 0xFFFFFFFFull);
      p.u_off[b + 1] = words;  // list lengths for now, offsets below
      r.max_u = std::max(r.max_u, slots);
    }
  });
  size_t total = 0;
  for (auto& r : ranges) {
    if (!r.ok) return false;
    total += r.ulist.size();
    p.max_u = std::max(p.max_u, r.max_u);
  }
  for (int32_t b = 0; b < nb; b++) p.u_off[b + 1] += p.u_off[b];
  p.ulist.reserve(total);
  for (auto& r : ranges) p.ulist.insert(p.ulist.end(), r.ulist.begin(), r.ulist.end());
  return true;
}

// Queues a host-to-device copy of src (moved in and held until finish_uploads) on the
// handle's stream.  A synchronous copy per array would wait, each time, for whatever the
// device's busy hardware queues hold ahead of it.
template <typename T>
int queue_upload(rsg_instance* h, void* dst, std::vector<T>&& src) {
  if (src.empty()) return RSG_OK;
  auto keep = std::make_shared<std::vector<T>>(std::move(src));
  HIP_TRY(hipMemcpyAsync(dst, keep->data(), keep->size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
  std::lock_guard<std::mutex> lock(h->hold_mu);
  h->held.push_back(std::move(keep));
  return RSG_OK;
}

// Waits for the handle's queued copies and releases their sources.
int finish_uploads(rsg_instance* h) {
  HIP_TRY(hipStreamSynchronize(h->stream));
  std::lock_guard<std::mutex> lock(h->hold_mu);
  h->held.clear();
  return RSG_OK;
}

// Allocates a device array for src (moved in) and queues its upload.
template <typename T>
int upload(rsg_instance* h, void** dst, std::vector<T>&& src) {
  int rc = h->dalloc(dst, src.size() * sizeof(T));
  if (rc) return rc;
  return queue_upload(h, *dst, std::move(src));
}

// Converts FP64 values to the handle's precision and uploads them.
int upload_real(rsg_instance* h, void** dst, std::vector<double> src) {
  if (h->precision == RSG_FP64) return upload(h, dst, std::move(src));
  std::vector<float> f(src.size());
  for (size_t i = 0; i < src.size(); i++) f[i] = (float)src[i];
  return upload(h, dst, std::move(f));
}

// Assigns rows of W values (in the handle's precision) to classes of equal rows, in
// order of first appearance; add() returns -1 once more than max_classes are needed.
template <typename T, int W>
struct Classifier {
  size_t max_classes;
  std::vector<T> table;
  std::unordered_map<uint64_t, std::vector<int32_t>> index;
  int32_t last = -1;
  explicit Classifier(size_t m) : max_classes(m) {}
  int32_t add(const T (&row)[W]) {
    if (last >= 0 && memcmp(row, &table[(size_t)last * W], sizeof(row)) == 0) return last;
    // the row's bytes hashed 8 at a time (multiply and fold); equal rows are found by comparing
    // the rows themselves, so the hash only picks the bucket
    uint64_t hsh = 0x9E3779B97F4A7C15ull;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(row);
    for (size_t i = 0; i < sizeof(row); i += 8) {
      uint64_t wd = 0;
      memcpy(&wd, b + i, std::min<size_t>(8, sizeof(row) - i));
      hsh = (hsh ^ wd) * 0xFF51AFD7ED558CCDull;
      hsh ^= hsh >> 29;
    }
    auto& list = index[hsh];
    for (int32_t c : list)
      if (memcmp(row, &table[(size_t)c * W], sizeof(row)) == 0) return last = c;
    const size_t n = table.size() / W;
    if (n >= max_classes) return -1;
    table.insert(table.end(), row, row + W);
    list.push_back((int32_t)n);
    return last = (int32_t)n;
  }
};

// Material rows of n entities in the order given (row(c, m) fills the W values of entity c):
// row_of[c] receives the entity's class and rows the class rows, numbered by first appearance
// in that order; chunks are classified on their own threads and merged in chunk order.  False
// when there are more than kMaxMaterials distinct rows.
template <typename T, int W, typename F>
bool classify_rows(int32_t n, F&& row, std::vector<int32_t>& row_of, std::vector<T>& rows) {
  struct Chunk {
    Classifier<T, W> mats{(size_t)rsg::kMaxMaterials};
    bool overflow = false;
    std::vector<int32_t> map;
  };
  std::vector<Chunk> chunks(chunk_count(n));
  row_of.resize(n);
  parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
    Chunk& ch = chunks[w];
    T m[W];
    for (int64_t c = b; c < e; c++) {
      row((int32_t)c, m);
      const int32_t k = ch.mats.add(m);
      if (k < 0) {
        ch.overflow = true;
        return;
      }
      row_of[c] = k;
    }
  });
  Classifier<T, W> all((size_t)rsg::kMaxMaterials);
  for (auto& ch : chunks) {
    if (ch.overflow) return false;
    const size_t nm = ch.mats.table.size() / W;
    ch.map.resize(nm);
    for (size_t k = 0; k < nm; k++) {
      T r[W];
      memcpy(r, &ch.mats.table[k * W], sizeof(r));
      const int32_t g = all.add(r);
      if (g < 0) return false;
      ch.map[k] = g;
    }
  }
  parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
    for (int64_t c = b; c < e; c++) row_of[c] = chunks[w].map[row_of[c]];
  });
  rows = std::move(all.table);
  return true;
}

// Keys of n entities from their material row m (W values) and spacings sp (3 values),
// both in the handle's precision: material classes and per-axis spacing classes in order
// of first appearance.  On success uploads the material table (rows of ROW values) and
// the [3][kAxisClasses] spacing table and sets *materials; leaves it at -1 when the mesh
// has more than kMaxMaterials materials or kAxisClasses spacings along an axis.
// Contiguous chunks are classified on their own threads, then their classes are merged in
// chunk order, which numbers every class by its first appearance exactly as one pass
// would; entity() must be safe to call from several threads.
template <typename T, int W, int ROW, typename F>
int make_keys(rsg_instance* h, int32_t n, F&& entity, std::vector<uint32_t>& keys, void** mat_table,
              void** axis_table, int32_t* materials) {
  *materials = -1;
  keys.resize(n);
  struct Chunk {
    Classifier<T, W> mats{(size_t)rsg::kMaxMaterials};
    std::vector<T> axv[3];  // spacing classes in order of first appearance in the chunk
    bool overflow = false;
    std::vector<int32_t> mat_map, ax_map[3];  // chunk class -> merged class
  };
  std::vector<Chunk> chunks(chunk_count(n));
  parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
    Chunk& ch = chunks[w];
    std::unordered_map<uint64_t, int> ax[3];
    uint64_t last_bits[3] = {0, 0, 0};
    int last_class[3] = {-1, -1, -1};  // the previous entity's spacing classes
    T m[W], sp[3];
    for (int32_t i = (int32_t)b; i < (int32_t)e; i++) {
      entity(i, m, sp);
      const int32_t mc = ch.mats.add(m);
      if (mc < 0) {
        ch.overflow = true;
        return;
      }
      int ic[3];
      for (int d = 0; d < 3; d++) {
        uint64_t bits = 0;
        memcpy(&bits, &sp[d], sizeof(T));
        if (last_class[d] >= 0 && bits == last_bits[d]) {
          ic[d] = last_class[d];
          continue;
        }
        auto it = ax[d].find(bits);
        if (it == ax[d].end()) {
          if ((int)ch.axv[d].size() >= rsg::kAxisClasses) {
            ch.overflow = true;
            return;
          }
          it = ax[d].emplace(bits, (int)ch.axv[d].size()).first;
          ch.axv[d].push_back(sp[d]);
        }
        ic[d] = it->second;
        last_bits[d] = bits;
        last_class[d] = ic[d];
      }
      keys[i] = rsg::make_key(mc, ic[0], ic[1], ic[2]);
    }
  });
  // merge in chunk order
  Classifier<T, W> mats(rsg::kMaxMaterials);
  std::unordered_map<uint64_t, int> ax[3];
  std::vector<T> axv[3];
  for (auto& ch : chunks) {
    if (ch.overflow) {
      keys.clear();
      return RSG_OK;
    }
    const size_t nm = ch.mats.table.size() / W;
    ch.mat_map.resize(nm);
    for (size_t c = 0; c < nm; c++) {
      T row[W];
      memcpy(row, &ch.mats.table[c * W], sizeof(row));
      const int32_t g = mats.add(row);
      if (g < 0) {
        keys.clear();
        return RSG_OK;
      }
      ch.mat_map[c] = g;
    }
    for (int d = 0; d < 3; d++) {
      ch.ax_map[d].resize(ch.axv[d].size());
      for (size_t k = 0; k < ch.axv[d].size(); k++) {
        uint64_t bits = 0;
        memcpy(&bits, &ch.axv[d][k], sizeof(T));
        auto it = ax[d].find(bits);
        if (it == ax[d].end()) {
          if ((int)axv[d].size() >= rsg::kAxisClasses) {
            keys.clear();
            return RSG_OK;
          }
          it = ax[d].emplace(bits, (int)axv[d].size()).first;
          axv[d].push_back(ch.axv[d][k]);
        }
        ch.ax_map[d][k] = it->second;
      }
    }
  }
  parallel_chunks(n, [&](int w, int64_t b, int64_t e) {
    const Chunk& ch = chunks[w];
    for (int64_t i = b; i < e; i++) {
      const uint32_t k = keys[i];
      keys[i] = rsg::make_key(ch.mat_map[k >> rsg::kMatShift], ch.ax_map[0][(k >> 14) & 0x7F],
                              ch.ax_map[1][(k >> 7) & 0x7F], ch.ax_map[2][k & 0x7F]);
    }
  });
  const size_t nm = mats.table.size() / W;
  std::vector<T> mt(nm * ROW, T(0)), at((size_t)3 * rsg::kAxisClasses, T(0));
  for (size_t c = 0; c < nm; c++)
    for (int j = 0; j < W; j++) mt[c * ROW + j] = mats.table[c * W + j];
  for (int d = 0; d < 3; d++)
    for (size_t k = 0; k < axv[d].size(); k++) at[(size_t)d * rsg::kAxisClasses + k] = axv[d][k];
  int rc;
  if ((rc = upload(h, mat_table, std::move(mt)))) return rc;
  if ((rc = upload(h, axis_table, std::move(at)))) return rc;
  *materials = (int32_t)nm;
  return RSG_OK;
}

// Uploads src in the handle's precision with the entries listed in boxes = {(index,
// special id)} replaced by NaN boxes written directly in the target format.
int copy_boxed(rsg_instance* h, void* dst, const std::vector<double>& src,
               const std::vector<std::pair<int32_t, int32_t>>& boxes) {
  if (h->precision == RSG_FP64) {
    std::vector<double> d(src);
    for (auto& b : boxes) d[b.first] = rsg::box_f64(b.second);
    return queue_upload(h, dst, std::move(d));
  } else {
    std::vector<float> f(src.size());
    for (size_t i = 0; i < src.size(); i++) f[i] = (float)src[i];
    for (auto& b : boxes) f[b.first] = rsg::box_f32(b.second);
    return queue_upload(h, dst, std::move(f));
  }
  return RSG_OK;
}

// Packs 8 16-bit local links (0xFFFF = none) and the key of every entity into the 16-byte
// records the packed kernels read: 10 bits per link, 3 links per word, the key in word 3.
inline void pack_links(const uint16_t* l8, uint32_t* w3) {
  w3[0] = w3[1] = w3[2] = 0;
  for (int k = 0; k < 8; k++) {
    const uint32_t v = l8[k] == 0xFFFF ? rsg::kNoLink : (uint32_t)l8[k];
    w3[k / 3] |= v << (10 * (k % 3));
  }
}

int upload_packed(rsg_instance* h, uint32_t* dst, const std::vector<uint16_t>& lconn,
                  const std::vector<uint32_t>& keys) {
  const size_t n = keys.size();
  std::vector<uint32_t> rec(n * 4);
  for (size_t i = 0; i < n; i++) {
    pack_links(&lconn[i * 8], &rec[i * 4]);
    rec[i * 4 + 3] = keys[i];
  }
  return queue_upload(h, dst, std::move(rec));
}

// Rewrites on the device the keys whose special flag changes: those of the entities flagged by
// the last table build (`flagged`, which then lists the entities flagged now) and those of the
// special entities now (`boxes`: internal id and special index, ids ascending).  `packed`: the
// 16-byte records (links from lconn, the key in word 3); else the 4-byte keys.  Set-up uploads
// every entity's plain key, so a build touches only the few special entities.
int patch_keys(rsg_instance* h, uint32_t* dst, bool packed, const std::vector<uint16_t>& lconn,
               const std::vector<uint32_t>& base, const std::vector<std::pair<int32_t, int32_t>>& boxes,
               std::vector<int32_t>& flagged) {
  std::vector<int32_t> ids(flagged);
  for (auto& b : boxes) ids.push_back(b.first);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  const size_t W = packed ? 4 : 1;
  auto rec = std::make_shared<std::vector<uint32_t>>(ids.size() * W);
  size_t q = 0;
  for (size_t a = 0; a < ids.size(); a++) {
    const int32_t i = ids[a];
    while (q < boxes.size() && boxes[q].first < i) q++;
    const bool special = q < boxes.size() && boxes[q].first == i;
    const uint32_t key = special ? rsg::kSpecialKey | (uint32_t)boxes[q].second : base[i];
    if (packed) {
      pack_links(&lconn[(size_t)i * 8], &(*rec)[a * 4]);
      (*rec)[a * 4 + 3] = key;
    } else {
      (*rec)[a] = key;
    }
  }
  flagged.clear();
  for (auto& b : boxes) flagged.push_back(b.first);
  for (size_t a = 0; a < ids.size();) {  // one copy per run of consecutive ids
    size_t e = a + 1;
    while (e < ids.size() && ids[e] == ids[e - 1] + 1) e++;
    HIP_TRY(hipMemcpyAsync(dst + (size_t)ids[a] * W, rec->data() + a * W, (e - a) * W * sizeof(uint32_t),
                           hipMemcpyHostToDevice, h->stream));
    a = e;
  }
  std::lock_guard<std::mutex> lock(h->hold_mu);
  h->held.push_back(std::move(rec));
  return RSG_OK;
}

// Builds the special-node / special-cell tables from the channels and sensors, and
// NaN-boxes the idt / dcx entries of the special entities on the device.
int build_tables(rsg_instance* h) {
  const int32_t nn = h->nn;
  // (entity, channel) for every support entry, ordered by entity then channel: the force
  // channels covering a cell and the moment channels covering a node (caller numbering) are a
  // run of the list, in channel order
  std::vector<std::pair<int32_t, int32_t>> cell_k, node_k;
  for (size_t k = 0; k < h->channels.size(); k++) {
    const Channel& ch = h->channels[k];
    auto& list = ch.kind == RSG_FORCE ? cell_k : node_k;
    for (int32_t e : ch.support) list.push_back({e, (int32_t)k});
  }
  std::sort(cell_k.begin(), cell_k.end());
  std::sort(node_k.begin(), node_k.end());
  using Run = std::pair<std::vector<std::pair<int32_t, int32_t>>::const_iterator,
                        std::vector<std::pair<int32_t, int32_t>>::const_iterator>;
  auto run_of = [](const std::vector<std::pair<int32_t, int32_t>>& list, int32_t e) -> Run {
    auto lo = std::lower_bound(list.begin(), list.end(), std::make_pair(e, INT32_MIN));
    auto hi = lo;
    while (hi != list.end() && hi->first == e) ++hi;
    return {lo, hi};
  };
  // a node's sensor slot: the last sensor entry that names it
  std::vector<std::pair<int32_t, int32_t>> sens;
  for (size_t s = 0; s < h->sensors.size(); s++) sens.push_back({h->sensors[s], (int32_t)s});
  std::sort(sens.begin(), sens.end());
  // special nodes: a vertex of a cell some force channel covers, or a sensor (internal ids,
  // ascending)
  std::vector<int32_t> sp_nodes;
  for (size_t a = 0; a < cell_k.size(); a++)
    if (a == 0 || cell_k[a].first != cell_k[a - 1].first)
      for (int v = 0; v < 8; v++) sp_nodes.push_back(h->node_at[h->cells[(size_t)cell_k[a].first * 8 + v]]);
  for (int32_t s : h->sensors) sp_nodes.push_back(h->node_at[s]);
  std::sort(sp_nodes.begin(), sp_nodes.end());
  sp_nodes.erase(std::unique(sp_nodes.begin(), sp_nodes.end()), sp_nodes.end());

  std::vector<double> sp_idt;
  std::vector<int32_t> sp_off, sp_sens, terms;
  const std::vector<double>& idt = h->idt_h;
  std::vector<std::pair<int32_t, int32_t>> node_boxes, cell_boxes;
  for (int32_t i : sp_nodes) {
    const int32_t n = h->node_of[i];
    const int32_t sp = (int32_t)sp_idt.size();
    if (h->precision == RSG_FP32 && sp > rsg::kMaxSpecialF32)
      return fail(RSG_ERR_ARG, "too many source/sensor nodes for the FP32 form");
    sp_idt.push_back(idt[i]);
    const Run sr = run_of(sens, n);
    sp_sens.push_back(sr.first == sr.second ? -1 : (sr.second - 1)->second);
    for (int d = 0; d < 3; d++) {
      sp_off.push_back((int32_t)terms.size());
      for (int kb = 0; kb < 8; kb++) {  // evalXds: slot, then transducer
        const int32_t c = h->node_cells[(size_t)n * 8 + kb];
        if (c == -1) continue;
        const Run r = run_of(cell_k, c);
        for (auto it = r.first; it != r.second; ++it)
          if (h->channels[it->second].comp == d) terms.push_back(it->second);
      }
    }
    sp_off.push_back((int32_t)terms.size());
    node_boxes.push_back({i, sp});
  }
  // special cells: a vertex lies in a moment channel's support (internal ids, ascending)
  std::vector<int32_t> sp_cells;
  if (!node_k.empty()) {
    if (h->incidence_exact) {  // the cells around each support node
      for (size_t a = 0; a < node_k.size(); a++)
        if (a == 0 || node_k[a].first != node_k[a - 1].first)
          for (int kb = 0; kb < 8; kb++) {
            const int32_t c = h->node_cells[(size_t)node_k[a].first * 8 + kb];
            if (c >= 0) sp_cells.push_back(h->cell_at[c]);
          }
      std::sort(sp_cells.begin(), sp_cells.end());
      sp_cells.erase(std::unique(sp_cells.begin(), sp_cells.end()), sp_cells.end());
    } else {  // a node slot two cells claim: every cell's vertices
      std::vector<char> in_support(nn, 0);
      for (auto& e : node_k) in_support[e.first] = 1;
      for (int32_t i = 0; i < h->nc; i++) {
        const int32_t c = h->cell_of[i];
        bool touched = false;
        for (int v = 0; v < 8 && !touched; v++) touched = in_support[h->cells[(size_t)c * 8 + v]] != 0;
        if (touched) sp_cells.push_back(i);
      }
    }
  }
  std::vector<double> sp_dcx;
  std::vector<int32_t> csp_off, cterms;
  const std::vector<double>& dcx = h->dcx_h;
  for (int32_t i : sp_cells) {
    const int32_t c = h->cell_of[i];
    const int32_t sp = (int32_t)sp_dcx.size();
    if (h->precision == RSG_FP32 && sp > rsg::kMaxSpecialF32)
      return fail(RSG_ERR_ARG, "too many source cells for the FP32 form");
    sp_dcx.push_back(dcx[i]);
    for (int ks = 0; ks < 6; ks++) {
      csp_off.push_back((int32_t)cterms.size());
      for (int kb = 0; kb < 8; kb++) {  // evalMoments: vertex, then excitation
        const Run r = run_of(node_k, h->cells[(size_t)c * 8 + kb]);
        for (auto it = r.first; it != r.second; ++it)
          if (h->channels[it->second].comp == ks) cterms.push_back(it->second);
      }
    }
    csp_off.push_back((int32_t)cterms.size());
    cell_boxes.push_back({i, sp});
  }
  h->n_sp_node = (int32_t)sp_idt.size();
  h->n_sp_cell = (int32_t)sp_dcx.size();
  // class-table form: keys carry the special flag; the special tables carry the class
  std::vector<int32_t> sp_nkey, sp_ckey;
  if (h->n_nmat >= 0)
    for (auto& b : node_boxes) sp_nkey.push_back((int32_t)h->nkey_h[b.first]);
  if (h->n_cmat >= 0)
    for (auto& b : cell_boxes) sp_ckey.push_back((int32_t)h->ckey_h[b.first]);

  auto put_real = [&](DevBuf<char>& buf, const std::vector<double>& v) -> int {
    int rc = buf.alloc(std::max<size_t>(1, v.size()) * h->elt(), h->stream);
    if (rc) return rc;
    if (v.empty()) return RSG_OK;
    if (h->precision == RSG_FP64) return queue_upload(h, buf.p, std::vector<double>(v));
    return queue_upload(h, buf.p, std::vector<float>(v.begin(), v.end()));
  };
  auto put_int = [&](DevBuf<int32_t>& buf, const std::vector<int32_t>& v) -> int {
    int rc = buf.alloc(std::max<size_t>(1, v.size()), h->stream);
    if (rc) return rc;
    return queue_upload(h, buf.p, std::vector<int32_t>(v));
  };
  int rc;
  if ((rc = put_real(h->sp_node_val, sp_idt))) return rc;
  if ((rc = put_int(h->sp_node_off, sp_off))) return rc;
  if ((rc = put_int(h->sp_node_sens, sp_sens))) return rc;
  if ((rc = put_int(h->node_terms, terms))) return rc;
  if ((rc = put_real(h->sp_cell_val, sp_dcx))) return rc;
  if ((rc = put_int(h->sp_cell_off, csp_off))) return rc;
  if ((rc = put_int(h->cell_terms, cterms))) return rc;
  if ((rc = put_int(h->sp_node_key, sp_nkey))) return rc;
  if ((rc = put_int(h->sp_cell_key, sp_ckey))) return rc;
  if (h->v_lkey || h->nkey) {
    if ((rc = patch_keys(h, h->v_lkey ? h->v_lkey : h->nkey, h->v_lkey != nullptr, h->v_lconn_h, h->nkey_h,
                         node_boxes, h->v_flagged)))
      return rc;
  } else if ((rc = copy_boxed(h, h->idt, idt, node_boxes))) {
    return rc;
  }
  if (h->s_lkey || h->ckey) {
    if ((rc = patch_keys(h, h->s_lkey ? h->s_lkey : h->ckey, h->s_lkey != nullptr, h->s_lconn_h, h->ckey_h,
                         cell_boxes, h->s_flagged)))
      return rc;
  } else if ((rc = copy_boxed(h, h->dc[0], dcx, cell_boxes))) {
    return rc;
  }
  // amplitude staging for kBlockSteps steps (reallocated: no amplitudes are held ahead)
  h->win_first = -1;
  const size_t nch = std::max<size_t>(1, h->channels.size());
  for (int b = 0; b < 2; b++) {
    if (h->stage_ev[b]) HIP_TRY(hipEventSynchronize(h->stage_ev[b]));
    h->stage_busy[b] = false;
    if ((rc = h->amp[b].alloc(nch * kBlockSteps, h->stream))) return rc;
    // the pinned staging comes with the first rsg_step that uploads amplitudes (a batched
    // handle never does: hipHostFree, like hipFree, waits for every stream of the device)
    if (h->stage[b] && h->stage_len[b] < nch * kBlockSteps) {
      (void)hipHostFree(h->stage[b]);
      h->stage[b] = nullptr;
    }
    if (!h->stage_ev[b]) HIP_TRY(hipEventCreateWithFlags(&h->stage_ev[b], hipEventDisableTiming));
  }
  if ((rc = finish_uploads(h))) return rc;
  h->tables_dirty = false;
  return RSG_OK;
}

int flush_traces(rsg_instance* h) {
  const size_t row = h->sensors.size() * 3;
  if (h->trace_dev_steps == 0 || row == 0) return RSG_OK;
  const size_t old = h->trace_host.size();
  h->trace_host.resize(old + (size_t)h->trace_dev_steps * row);
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->batch_stream) HIP_TRY(hipStreamSynchronize(h->batch_stream));
  HIP_TRY(copy_on(h->trace_host.data() + old, h->trace.p, (size_t)h->trace_dev_steps * row * sizeof(double),
                  hipMemcpyDeviceToHost, h->stream));
  h->trace_dev_steps = 0;
  return RSG_OK;
}

template <typename T>
void fill_args(rsg_instance* h, rsg::VelocityArgs<T>& va, rsg::StressArgs<T>& sa) {
  va.V0 = (T*)h->V[0];
  va.V1 = (T*)h->V[1];
  va.V2 = (T*)h->V[2];
  va.S0 = (const T*)h->S[0];
  va.S1 = (const T*)h->S[1];
  va.S2 = (const T*)h->S[2];
  va.S3 = (const T*)h->S[3];
  va.S4 = (const T*)h->S[4];
  va.S5 = (const T*)h->S[5];
  va.idt = (const T*)h->idt;
  va.dnx = (const T*)h->dn[0];
  va.dny = (const T*)h->dn[1];
  va.dnz = (const T*)h->dn[2];
  va.ncell = h->ncell;
  va.n = h->nn;
  va.xcd = h->xcd;
  va.cpt = h->cpt_v;
  va.u_off = h->v_uoff;
  va.ulist = h->v_ulist;
  va.uchunk = h->v_chunk;
  va.lconn = h->v_lconn;
  va.max_u = h->v_maxu;
  va.sp_idt = (const T*)h->sp_node_val.p;
  va.sp_off = h->sp_node_off.p;
  va.sp_sens = h->sp_node_sens.p;
  va.term = h->node_terms.p;
  va.key = h->nkey;
  va.lkey = h->v_lkey;
  va.cls = (const T*)h->nmat;
  va.axis = (const T*)h->naxis;
  va.sp_key = (const uint32_t*)h->sp_node_key.p;
  sa.S0 = (T*)h->S[0];
  sa.S1 = (T*)h->S[1];
  sa.S2 = (T*)h->S[2];
  sa.S3 = (T*)h->S[3];
  sa.S4 = (T*)h->S[4];
  sa.S5 = (T*)h->S[5];
  sa.V0 = (const T*)h->V[0];
  sa.V1 = (const T*)h->V[1];
  sa.V2 = (const T*)h->V[2];
  sa.C = (const T*)h->C;
  sa.ldc = h->ldc;
  sa.dcx = (const T*)h->dc[0];
  sa.dcy = (const T*)h->dc[1];
  sa.dcz = (const T*)h->dc[2];
  sa.cnode = h->cnode;
  sa.n = h->nc;
  sa.xcd = h->xcd;
  sa.cpt = h->cpt_s;
  sa.u_off = h->s_uoff;
  sa.ulist = h->s_ulist;
  sa.lconn = h->s_lconn;
  sa.max_u = h->s_maxu;
  sa.dt = (T)h->dt;
  sa.sp_dcx = (const T*)h->sp_cell_val.p;
  sa.sp_off = h->sp_cell_off.p;
  sa.term = h->cell_terms.p;
  sa.key = h->ckey;
  sa.lkey = h->s_lkey;
  sa.cls = (const T*)h->cmat;
  sa.axis = (const T*)h->caxis;
  sa.sp_key = (const uint32_t*)h->sp_cell_key.p;
}

int ensure_events(rsg_instance* h, size_t n) {
  while (h->tev.size() < n) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    h->tev.push_back(e);
  }
  return RSG_OK;
}

template <typename T>
int run_block(rsg_instance* h, int kb, const double* amp_dev) {
  rsg::VelocityArgs<T> va;
  rsg::StressArgs<T> sa;
  fill_args<T>(h, va, sa);
  const size_t nch = std::max<size_t>(1, h->channels.size());
  const size_t row = h->sensors.size() * 3;
  // timed steps get three events: before the velocity launch, between the two, after the
  // stress launch (events between every launch slow the stream down, so a stride samples)
  const int stride = h->timing;
  const int nt = stride > 0 ? (kb + stride - 1) / stride : 0;
  if (nt) {
    int rc = ensure_events(h, (size_t)3 * nt);
    if (rc) return rc;
  }
  for (int k = 0; k < kb; k++) {
    const bool timed = stride > 0 && k % stride == 0;
    hipEvent_t* ev = timed ? &h->tev[(size_t)3 * (k / stride)] : nullptr;
    va.amp = amp_dev + (size_t)k * nch;
    sa.amp = va.amp;
    va.trace = h->trace.p ? h->trace.p + (size_t)(h->trace_dev_steps + k) * row : nullptr;
    hipError_t e;
    if (timed) HIP_TRY(hipEventRecord(ev[0], h->stream));
    if (sizeof(T) == 8)
      e = rsg::launch_velocity_f64(*(rsg::VelocityArgs<double>*)&va, h->block_v, h->stream);
    else
      e = rsg::launch_velocity_f32(*(rsg::VelocityArgs<float>*)&va, h->block_v, h->stream);
    if (e != hipSuccess) return fail(RSG_ERR_DEVICE, "velocity launch: %s", hipGetErrorString(e));
    if (timed) HIP_TRY(hipEventRecord(ev[1], h->stream));
    if (sizeof(T) == 8)
      e = rsg::launch_stress_f64(*(rsg::StressArgs<double>*)&sa, h->block_s, h->stream);
    else
      e = rsg::launch_stress_f32(*(rsg::StressArgs<float>*)&sa, h->block_s, h->stream);
    if (e != hipSuccess) return fail(RSG_ERR_DEVICE, "stress launch: %s", hipGetErrorString(e));
    if (timed) HIP_TRY(hipEventRecord(ev[2], h->stream));
  }
  if (nt) {
    HIP_TRY(hipEventSynchronize(h->tev[(size_t)3 * nt - 1]));
    for (int q = 0; q < nt; q++) {
      float a = 0, b = 0;
      HIP_TRY(hipEventElapsedTime(&a, h->tev[(size_t)3 * q], h->tev[(size_t)3 * q + 1]));
      HIP_TRY(hipEventElapsedTime(&b, h->tev[(size_t)3 * q + 1], h->tev[(size_t)3 * q + 2]));
      h->t_vel_ms += a;
      h->t_str_ms += b;
    }
    h->t_launches += nt;
  }
  h->trace_dev_steps += row ? kb : 0;
  return RSG_OK;
}

// Energy of the handle's state with partials and totals in `work` (2 * kEnergyMaxBlocks + 2
// doubles of device memory): both kernels on the handle's stream, then the two totals copied
// out.  Waits for the batch stream first when the handle is in a batch.
template <typename T>
void fill_energy_args(rsg_instance* h, rsg::EnergyArgs<T>& a, double* work) {
  rsg::VelocityArgs<T> va;
  fill_args<T>(h, va, a.s);
  a.V0 = (const T*)h->V[0];
  a.V1 = (const T*)h->V[1];
  a.V2 = (const T*)h->V[2];
  a.node_m = h->node_m;
  a.nn = h->nn;
  a.cell_vol = h->cell_vol;
  a.inv_scale = h->precision == RSG_FP64 ? 1.0 : 1.0 / h->dt;
  a.partial = work;
}

int energy_on_stream(rsg_instance* h, double* work, double* kinetic, double* strain) {
  Range range_("rsg_energy");
  if (h->batch_stream) HIP_TRY(hipStreamSynchronize(h->batch_stream));
  if (h->tables_dirty) {
    int rc = build_tables(h);  // the special-cell keys the kernel resolves
    if (rc) return rc;
  }
  const int nb = rsg::kEnergyMaxBlocks;
  double* out = work + 2 * nb;
  hipError_t e;
  if (h->precision == RSG_FP64) {
    rsg::EnergyArgs<double> a;
    fill_energy_args<double>(h, a, work);
    e = rsg::launch_energy_f64(a, nb, out, h->stream);
  } else {
    rsg::EnergyArgs<float> a;
    fill_energy_args<float>(h, a, work);
    e = rsg::launch_energy_f32(a, nb, out, h->stream);
  }
  if (e != hipSuccess) return fail(RSG_ERR_DEVICE, "energy launch: %s", hipGetErrorString(e));
  double tot[2] = {0.0, 0.0};
  HIP_TRY(hipMemcpyAsync(tot, out, sizeof(tot), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (kinetic) *kinetic = tot[0];
  if (strain) *strain = tot[1];
  return RSG_OK;
}

int check_mesh(const rsg_mesh_desc* m) {
  if (!m) return fail(RSG_ERR_ARG, "mesh is null");
  if (m->n_nodes <= 0 || m->n_cells <= 0)
    return fail(RSG_ERR_ARG, "mesh needs nodes and cells (got %d, %d)", m->n_nodes, m->n_cells);
  if (!m->x || !m->y || !m->z || !m->cell_nodes) return fail(RSG_ERR_ARG, "mesh array is null");
  for (int64_t i = 0; i < (int64_t)m->n_cells * 8; i++) {
    const int32_t v = m->cell_nodes[i];
    if (v < 0 || v >= m->n_nodes)
      return fail(RSG_ERR_ARG, "invalid node id: Node count: %d, provided index: %d", m->n_nodes, v);
  }
  return RSG_OK;
}

}  // namespace

extern "C" {

int rsg_last_error(char* buf, size_t len) {
  if (buf && len) {
    size_t k = std::min(len - 1, g_last_error.size());
    memcpy(buf, g_last_error.data(), k);
    buf[k] = 0;
  }
  return (int)g_last_error.size();
}

int rsg_create(const rsg_mesh_desc* mesh, const rsg_model_desc* model, int32_t precision,
               rsg_handle* out) {
  return rsg_create_ex(mesh, model, precision, nullptr, out);
}

int rsg_get_tuning(rsg_handle h, rsg_tuning* t) {
  if (!h || !t) return fail(RSG_ERR_ARG, "null argument");
  t->block_velocity = h->block_v;
  t->block_stress = h->block_s;
  t->per_thread_velocity = h->cpt_v;
  t->per_thread_stress = h->cpt_s;
  t->lds_velocity = h->lds_v ? 1 : -1;
  t->lds_stress = h->lds_s ? 1 : -1;
  for (int d = 0; d < 3; d++) t->tile[d] = h->tile[d];
  t->c_layout = h->c_blocked ? 1 : -1;
  t->xcd_remap = h->xcd ? 1 : -1;
  t->coeff_classes = h->classes ? 1 : -1;
  t->node_classes = h->n_nmat;
  t->cell_classes = h->n_cmat;
  t->links_velocity = h->v_lkey ? 2 : h->v_lconn ? 1 : 0;
  t->links_stress = h->s_lkey ? 2 : h->s_lconn ? 1 : 0;
  return RSG_OK;
}

int rsg_create_ex(const rsg_mesh_desc* mesh, const rsg_model_desc* model, int32_t precision,
                  const rsg_tuning* tuning, rsg_handle* out) {
  Range range_("rsg_create");
  if (!out) return fail(RSG_ERR_ARG, "out is null");
  *out = nullptr;
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!model) return fail(RSG_ERR_ARG, "model is null");
  if (precision != RSG_FP64 && precision != RSG_FP32)
    return fail(RSG_ERR_ARG, "precision %d is neither RSG_FP64 nor RSG_FP32", precision);
  if ((rc = check_field(model->density, "density"))) return rc;
  for (int k = 0; k < 21; k++)
    if ((rc = check_field(model->cij[k], "C_IJ"))) return rc;
  if (!std::isfinite(model->dt)) return fail(RSG_ERR_ARG, "dt is not finite");

  std::unique_ptr<rsg_instance> h(new (std::nothrow) rsg_instance());
  if (!h) return fail(RSG_ERR_NOMEM, "host allocation failed");
  const int32_t nn = mesh->n_nodes, nc = mesh->n_cells;
  h->precision = precision;
  h->nn = nn;
  h->nc = nc;
  h->dt = model->dt;
  h->moment_form = model->moment_form ? 1 : 0;
  if (tuning) {
    const rsg_tuning& t = *tuning;
    if (t.block_velocity) h->block_v = t.block_velocity;
    if (t.block_stress) h->block_s = t.block_stress;
    if (t.per_thread_velocity) h->cpt_v = t.per_thread_velocity;
    if (t.per_thread_stress) h->cpt_s = t.per_thread_stress;
    if (t.lds_velocity) h->lds_v = t.lds_velocity > 0;
    if (t.lds_stress) h->lds_s = t.lds_stress > 0;
    if (t.c_layout) h->c_blocked = t.c_layout > 0;
    if (t.xcd_remap) h->xcd = t.xcd_remap > 0;
    if (t.coeff_classes) h->classes = t.coeff_classes > 0;
    for (int d = 0; d < 3; d++) {
      if (t.tile[d] < 0) return fail(RSG_ERR_ARG, "tile[%d] = %d is negative", d, t.tile[d]);
      if (t.tile[d]) h->tile[d] = t.tile[d];
    }
  }
  // two cells per thread for the stress kernel on meshes of 3 Mi cells or more (their
  // streaming operands are requested one cell ahead: -9 % stress time on a 3.4 M-cell plate),
  // one below (measured faster on 32^3..128^3 instances), unless the caller chose
  if (!tuning || tuning->per_thread_stress == 0) h->cpt_s = nc >= (3 << 20) ? 2 : 1;
  // 512-node velocity workgroups on meshes of 3 Mi nodes or more (-5 % velocity time in FP32
  // and -4 % in FP64 on the 3.5 M-node plate; 256 again if their staged sets do not fit in
  // LDS or in packed links), unless the caller chose the velocity workgroup or its nodes
  // per thread
  bool auto_block_v = false;
  if ((!tuning || (tuning->block_velocity == 0 && tuning->per_thread_velocity == 0)) && nn >= (3 << 20)) {
    h->block_v = 512;
    auto_block_v = true;
  }
  for (int blk : {h->block_v, h->block_s})
    if (blk < 64 || blk > 512 || blk % 64 != 0)
      return fail(RSG_ERR_ARG, "workgroup size %d: must be a multiple of 64 in [64, 512]", blk);
  for (int c : {h->cpt_v, h->cpt_s})
    if (c != 1 && c != 2 && c != 4) return fail(RSG_ERR_ARG, "entities per thread %d: must be 1, 2 or 4", c);

  CREATE_PHASE("start");
  // links (HexMesh::buildLinks): slot = vertex ^ 6, the later cell winning a slot two cells
  // claim, as in the reference's loop
  h->cells.resize((size_t)nc * 8);
  h->node_cells.resize((size_t)nn * 8);
  parallel_ranges(nn, [&](int64_t b, int64_t e) {
    std::fill(h->node_cells.begin() + b * 8, h->node_cells.begin() + e * 8, -1);
  });
  // on threads over ranges of cells: a slot keeps the largest cell that claims it (the later
  // one in the reference's loop), an atomic maximum, so the result is the serial loop's
  std::atomic<bool> exact{true};
  parallel_ranges(nc, [&](int64_t b, int64_t e) {
    std::copy(mesh->cell_nodes + b * 8, mesh->cell_nodes + e * 8, h->cells.begin() + b * 8);
    for (int64_t k = b; k < e; k++)
      for (int v = 0; v < 8; v++) {
        int32_t* slot = &h->node_cells[(size_t)h->cells[(size_t)k * 8 + v] * 8 + (v ^ 6)];
        int32_t seen = __atomic_load_n(slot, __ATOMIC_RELAXED);
        if (seen >= 0) exact.store(false, std::memory_order_relaxed);
        while (seen < (int32_t)k &&
               !__atomic_compare_exchange_n(slot, &seen, (int32_t)k, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        if (seen >= 0) exact.store(false, std::memory_order_relaxed);
      }
  });
  h->incidence_exact = exact.load();

  const double* X[3] = {mesh->x, mesh->y, mesh->z};
  CREATE_PHASE("links");
  // RightHex::cellPrep, and each cell's coordinate bounds for nodePrep below
  std::vector<double> dcell[3];
  for (int d = 0; d < 3; d++) dcell[d].resize(nc);
  std::vector<double> cbox((size_t)nc * 6);  // per cell: min x, y, z, max x, y, z
  parallel_ranges(nc, [&](int64_t c0, int64_t c1) {
    for (int32_t c = (int32_t)c0; c < (int32_t)c1; c++) {
      const int32_t* v = &h->cells[(size_t)c * 8];
      for (int d = 0; d < 3; d++) {
        dcell[d][c] = X[d][v[6]] - X[d][v[0]];
        double lo = INFINITY, hi = -INFINITY;
        for (int k = 0; k < 8; k++) {
          const double p = X[d][v[k]];
          lo = (p < lo) ? p : lo;
          hi = (hi < p) ? p : hi;
        }
        cbox[(size_t)c * 6 + d] = lo;
        cbox[(size_t)c * 6 + 3 + d] = hi;
      }
    }
  });
  CREATE_PHASE("cellprep");
  // RightHex::nodePrep: the bounds over the node's cells' vertices, scanned from the node's own
  // coordinate in slot order.  A strict comparison keeps the first value of the extreme in scan
  // order, and each cell's bounds are the first extreme of its vertices in vertex order, so
  // taking the cells' bounds in slot order selects the same value as the vertex-by-vertex scan.
  // The first node whose spacing fails is reported.
  std::vector<double> dnode[3];
  for (int d = 0; d < 3; d++) dnode[d].resize(nn);
  std::mutex bad_mu;
  int64_t bad_node = -1;
  int bad_axis = 0;
  parallel_ranges(nn, [&](int64_t n0, int64_t n1) {
  for (int32_t n = (int32_t)n0; n < (int32_t)n1; n++) {
    double pn[3], lo[3], hi[3];
    for (int d = 0; d < 3; d++) pn[d] = lo[d] = hi[d] = X[d][n];
    for (int kb = 0; kb < 8; kb++) {
      const int32_t c = h->node_cells[(size_t)n * 8 + kb];
      if (c == -1) continue;
      const double* b = &cbox[(size_t)c * 6];
      for (int d = 0; d < 3; d++) {
        lo[d] = (b[d] < lo[d]) ? b[d] : lo[d];
        hi[d] = (hi[d] < b[3 + d]) ? b[3 + d] : hi[d];
      }
    }
    for (int d = 0; d < 3; d++) {
      if (lo[d] == pn[d]) lo[d] = 2. * pn[d] - hi[d];
      if (hi[d] == pn[d]) hi[d] = 2. * pn[d] - lo[d];
    }
    for (int d = 0; d < 3; d++) {
      dnode[d][n] = 0.5 * (hi[d] - lo[d]);
      if (0 >= dnode[d][n]) {
        std::lock_guard<std::mutex> lock(bad_mu);
        if (bad_node < 0 || n < bad_node) {
          bad_node = n;
          bad_axis = d;
        }
        return;
      }
    }
  }
  });
  std::vector<double>().swap(cbox);
  if (bad_node >= 0) {
    static const char* kAxis[3] = {"DX", "DY", "DZ"};
    return fail(RSG_ERR_MESH, "%s (node %d)", kAxis[bad_axis], (int)bad_node);
  }

  CREATE_PHASE("nodeprep");
  // locality order
  double lo[3], hmin[3];
  {
    // per chunk: the smallest positive cell spacing along each axis
    std::vector<std::array<double, 3>> hmin_of((size_t)chunk_count(nc), {INFINITY, INFINITY, INFINITY});
    parallel_chunks(nc, [&](int w, int64_t b, int64_t e) {
      for (int d = 0; d < 3; d++) {
        double m = INFINITY;
        for (int64_t c = b; c < e; c++) {
          const double s = std::fabs(dcell[d][c]);
          if (s > 0 && s < m) m = s;
        }
        hmin_of[w][d] = m;
      }
    });
    for (int d = 0; d < 3; d++) {
      lo[d] = *std::min_element(X[d], X[d] + nn);
      hmin[d] = INFINITY;
      for (auto& m : hmin_of) hmin[d] = std::min(hmin[d], m[d]);
      if (!std::isfinite(hmin[d])) hmin[d] = 1.0;
    }
  }
  {
    std::vector<double> cx(nc), cy(nc), cz(nc);
    parallel_ranges(nc, [&](int64_t c0, int64_t c1) {
      for (int32_t c = (int32_t)c0; c < (int32_t)c1; c++) {
        double s[3] = {0, 0, 0};
        for (int v = 0; v < 8; v++)
          for (int d = 0; d < 3; d++) s[d] += X[d][h->cells[(size_t)c * 8 + v]];
        cx[c] = s[0] / 8;
        cy[c] = s[1] / 8;
        cz[c] = s[2] / 8;
      }
    });
    double clo[3] = {lo[0] + 0.5 * hmin[0], lo[1] + 0.5 * hmin[1], lo[2] + 0.5 * hmin[2]};
    bool cells_ok = true, nodes_ok = true;
    std::thread cell_sort([&]() {  // the two orders are independent
      try {
        h->cell_of = locality_order(nc, cx.data(), cy.data(), cz.data(), hmin, clo, h->tile);
      } catch (const std::bad_alloc&) {
        cells_ok = false;
      }
    });
    try {
      h->node_of = locality_order(nn, X[0], X[1], X[2], hmin, lo, h->tile);
    } catch (const std::bad_alloc&) {
      nodes_ok = false;
    }
    cell_sort.join();
    if (!cells_ok || !nodes_ok) return fail(RSG_ERR_NOMEM, "host allocation failed");
  }
  h->cell_at.resize(nc);
  h->node_at.resize(nn);
  parallel_ranges(nc, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) h->cell_at[h->cell_of[i]] = (int32_t)i;
  });
  parallel_ranges(nn, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) h->node_at[h->node_of[i]] = (int32_t)i;
  });

  CREATE_PHASE("order");
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  size_t lds_limit = 0;
  {
    int dev = 0, bytes = 0;
    HIP_TRY(hipGetDevice(&dev));
    h->device = dev;
    HIP_TRY(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    lds_limit = (size_t)std::min(bytes, 64 * 1024);
  }

  CREATE_PHASE("stream");
  // node arrays and cell arrays are independent: built on two threads
  auto node_part = [&]() -> int {
    int rc;
    std::vector<double> idt(nn), w(nn), dn[3];
    for (int d = 0; d < 3; d++) dn[d].resize(nn);
    std::vector<int32_t> conn((size_t)nn * 8);
    std::mutex bad_mu;
    int64_t bad = -1;  // first internal node whose dt/density is not finite
    parallel_ranges(nn, [&](int64_t i0, int64_t i1) {
      for (int32_t i = (int32_t)i0; i < (int32_t)i1; i++) {
        const int32_t n = h->node_of[i];
        const double rho = field_at_node(model->density, h->node_cells.data(), n);
        idt[i] = h->dt / (rho);
        if (!(idt[i] == idt[i]) || std::isinf(idt[i])) {
          std::lock_guard<std::mutex> lock(bad_mu);
          if (bad < 0 || i < bad) bad = i;
          return;
        }
        w[i] = rho * dnode[0][n] * dnode[1][n] * dnode[2][n];  // the node's share of the mass
        for (int d = 0; d < 3; d++)
          dn[d][i] = precision == RSG_FP64 ? dnode[d][n] : 0.25 / dnode[d][n];
        for (int k = 0; k < 8; k++) {
          const int32_t c = h->node_cells[(size_t)n * 8 + k];
          conn[(size_t)i * 8 + k] = c < 0 ? -1 : h->cell_at[c];
        }
      }
    });
    if (bad >= 0) {
      const int32_t n = h->node_of[bad];
      return fail(RSG_ERR_ARG, "dt/density is not finite at node %d (density %g)", n,
                  field_at_node(model->density, h->node_cells.data(), n));
    }
    h->idt_h = idt;
    if (h->classes) {  // material: dt/rho; spacings as the per-node arrays hold them
      auto entity = [&](int32_t i, auto& m, auto& sp) {
        using T = std::remove_reference_t<decltype(m[0])>;
        m[0] = (T)idt[i];
        for (int d = 0; d < 3; d++) sp[d] = (T)dn[d][i];
      };
      rc = precision == RSG_FP64
               ? make_keys<double, 1, 1>(h.get(), nn, entity, h->nkey_h, &h->nmat, &h->naxis, &h->n_nmat)
               : make_keys<float, 1, 1>(h.get(), nn, entity, h->nkey_h, &h->nmat, &h->naxis, &h->n_nmat);
      if (rc) return rc;
    }
    if (h->n_nmat < 0) {
      if ((rc = upload_real(h.get(), &h->idt, idt))) return rc;
      for (int d = 0; d < 3; d++)
        if ((rc = upload_real(h.get(), &h->dn[d], dn[d]))) return rc;
    }
    void* p;
    LdsPlan plan;
    bool packed = false;
    // Gather lists in chunk words (a quarter of the list bytes; measured 2-5 % faster velocity
    // launches) unless their padded sets no longer fit LDS or 10-bit links where plain lists do.
    bool chunk_v = true;
    auto plan_fits = [&]() {
      chunk_v = true;
      bool ok = build_lds_plan(conn, nn, h->block_v * h->cpt_v, plan, true) &&
                rsg::velocity_lds_bytes(plan.max_u, h->elt()) <= lds_limit;
      if (!(ok && (h->n_nmat < 0 || plan.max_u <= rsg::kMaxPackedSet))) {
        LdsPlan plain;
        if (build_lds_plan(conn, nn, h->block_v * h->cpt_v, plain, false) &&
            rsg::velocity_lds_bytes(plain.max_u, h->elt()) <= lds_limit) {
          plan = std::move(plain);
          chunk_v = false;
          ok = true;
        }
      }
      return ok;
    };
    bool fits = h->lds_v && plan_fits();
    auto packs = [&]() { return h->n_nmat >= 0 && plan.max_u <= rsg::kMaxPackedSet; };
    if (h->lds_v && auto_block_v && (!fits || (h->n_nmat >= 0 && !packs()))) {
      // the 512-node sets are too large for LDS or for 10-bit links: 256-node workgroups if
      // their sets are not
      const bool fits512 = fits, packs512 = fits && packs(), chunk512 = chunk_v;
      LdsPlan plan512 = std::move(plan);
      h->block_v = 256;
      fits = plan_fits();
      if (!((fits && !fits512) || (fits && packs() && !packs512))) {
        h->block_v = 512;
        plan = std::move(plan512);
        fits = fits512;
        chunk_v = chunk512;
      }
    }
    if (fits) {
      if ((rc = upload(h.get(), &p, std::move(plan.u_off)))) return rc;
      h->v_uoff = (int32_t*)p;
      if ((rc = upload(h.get(), &p, std::move(plan.ulist)))) return rc;
      h->v_ulist = (int32_t*)p;
      h->v_maxu = plan.max_u;
      h->v_chunk = chunk_v;
      // with class keys and 10-bit local indices the links and the key share one record
      packed = h->n_nmat >= 0 && plan.max_u <= rsg::kMaxPackedSet;
      if (packed) {
        if ((rc = h->dalloc((void**)&h->v_lkey, (size_t)nn * 16))) return rc;
        h->v_lconn_h = std::move(plan.lconn);
      } else {
        if ((rc = upload(h.get(), &p, std::move(plan.lconn)))) return rc;
        h->v_lconn = (uint16_t*)p;
      }
    }
    if (h->n_nmat >= 0 && !packed)
      if ((rc = h->dalloc((void**)&h->nkey, (size_t)nn * 4))) return rc;
    // every node's plain key (the table builds flag the special ones)
    if (h->v_lkey) {
      if ((rc = upload_packed(h.get(), h->v_lkey, h->v_lconn_h, h->nkey_h))) return rc;
    } else if (h->nkey) {
      if ((rc = queue_upload(h.get(), h->nkey, std::vector<uint32_t>(h->nkey_h)))) return rc;
    }
    if ((rc = upload(h.get(), &p, std::move(w)))) return rc;
    h->node_m = (double*)p;
    if (!fits) {  // a staged kernel reads its links from the plan, not the connectivity
      if ((rc = upload(h.get(), &p, std::move(conn)))) return rc;
      h->ncell = (int32_t*)p;
    }
    return RSG_OK;
  };
  auto cell_part = [&]() -> int {
    int rc;
    std::vector<double> dc[3], w(nc);
    for (int d = 0; d < 3; d++) dc[d].resize(nc);
    std::vector<int32_t> conn((size_t)nc * 8);
    const int64_t ncp = ((int64_t)nc + 63) / 64 * 64;
    h->ldc = h->c_blocked ? 0 : ncp;
    parallel_ranges(nc, [&](int64_t i0, int64_t i1) {
      for (int32_t i = (int32_t)i0; i < (int32_t)i1; i++) {
        const int32_t c = h->cell_of[i];
        for (int d = 0; d < 3; d++)
          dc[d][i] = precision == RSG_FP64 ? dcell[d][c] : 0.25 / dcell[d][c];
        w[i] = dcell[0][c] * dcell[1][c] * dcell[2][c];
        for (int k = 0; k < 8; k++) conn[(size_t)i * 8 + k] = h->node_at[h->cells[(size_t)c * 8 + k]];
      }
    });
    h->dcx_h = dc[0];
    // C entry j of internal cell i as the kernels read it: FP64 C_IJ, FP32 dt*C_IJ
    auto c_entry = [&](int32_t i, int j) {
      const double cij = field_at_cell(model->cij[j], h->cells.data(), h->cell_of[i]);
      return precision == RSG_FP64 ? cij : h->dt * cij;
    };
    if (h->classes) {  // material: the 21 C; spacings as the per-cell arrays hold them
      // The rows are classified in the caller's cell order first, where each of the 21 fields
      // is read sequentially; an internal cell then copies its row from the small class
      // table.  make_keys numbers the classes by first appearance in internal order, as it
      // would from the fields themselves.
      std::vector<int32_t> row_of;
      std::vector<double> rows_d;
      std::vector<float> rows_f;
      auto caller_row = [&](int32_t c, auto& m) {
        using T = std::remove_reference_t<decltype(m[0])>;
        for (int j = 0; j < 21; j++) {
          const double cij = field_at_cell(model->cij[j], h->cells.data(), c);
          m[j] = (T)(precision == RSG_FP64 ? cij : h->dt * cij);
        }
      };
      const bool few = precision == RSG_FP64 ? classify_rows<double, 21>(nc, caller_row, row_of, rows_d)
                                             : classify_rows<float, 21>(nc, caller_row, row_of, rows_f);
      if (few) {  // else more materials than the tables hold: per-cell arrays
        auto entity = [&](int32_t i, auto& m, auto& sp) {
          using T = std::remove_reference_t<decltype(m[0])>;
          const T* r;
          if constexpr (std::is_same<T, double>::value)
            r = rows_d.data() + (size_t)row_of[h->cell_of[i]] * 21;
          else
            r = rows_f.data() + (size_t)row_of[h->cell_of[i]] * 21;
          for (int j = 0; j < 21; j++) m[j] = r[j];
          for (int d = 0; d < 3; d++) sp[d] = (T)dc[d][i];
        };
        rc = precision == RSG_FP64
                 ? make_keys<double, 21, rsg::kMatRow>(h.get(), nc, entity, h->ckey_h, &h->cmat, &h->caxis, &h->n_cmat)
                 : make_keys<float, 21, rsg::kMatRow>(h.get(), nc, entity, h->ckey_h, &h->cmat, &h->caxis, &h->n_cmat);
        if (rc) return rc;
      }
    }
    if (h->n_cmat < 0) {
      std::vector<double> C((size_t)21 * ncp, 0.0);
      for (int32_t i = 0; i < nc; i++)
        for (int j = 0; j < 21; j++) {
          const size_t idx = h->c_blocked ? ((size_t)(i >> 6) * 21 + j) * 64 + (i & 63) : (size_t)j * ncp + i;
          C[idx] = c_entry(i, j);
        }
      for (int d = 0; d < 3; d++)
        if ((rc = upload_real(h.get(), &h->dc[d], dc[d]))) return rc;
      if ((rc = upload_real(h.get(), &h->C, std::move(C)))) return rc;
    }
    void* p;
    LdsPlan plan;
    bool packed = false;
    if (h->lds_s && build_lds_plan(conn, nc, h->block_s * h->cpt_s, plan) &&
        (size_t)3 * plan.max_u * h->elt() <= lds_limit) {
      if ((rc = upload(h.get(), &p, std::move(plan.u_off)))) return rc;
      h->s_uoff = (int32_t*)p;
      if ((rc = upload(h.get(), &p, std::move(plan.ulist)))) return rc;
      h->s_ulist = (int32_t*)p;
      h->s_maxu = plan.max_u;
      packed = h->n_cmat >= 0 && plan.max_u <= rsg::kMaxPackedSet;
      if (packed) {
        if ((rc = h->dalloc((void**)&h->s_lkey, (size_t)nc * 16))) return rc;
        h->s_lconn_h = std::move(plan.lconn);
      } else {
        if ((rc = upload(h.get(), &p, std::move(plan.lconn)))) return rc;
        h->s_lconn = (uint16_t*)p;
      }
    }
    if (h->n_cmat >= 0 && !packed)
      if ((rc = h->dalloc((void**)&h->ckey, (size_t)nc * 4))) return rc;
    // every cell's plain key (the table builds flag the special ones)
    if (h->s_lkey) {
      if ((rc = upload_packed(h.get(), h->s_lkey, h->s_lconn_h, h->ckey_h))) return rc;
    } else if (h->ckey) {
      if ((rc = queue_upload(h.get(), h->ckey, std::vector<uint32_t>(h->ckey_h)))) return rc;
    }
    if ((rc = upload(h.get(), &p, std::move(w)))) return rc;
    h->cell_vol = (double*)p;
    if (!(h->s_lconn || h->s_lkey)) {  // a staged kernel reads its links from the plan
      if ((rc = upload(h.get(), &p, std::move(conn)))) return rc;
      h->cnode = (int32_t*)p;
    }
    return RSG_OK;
  };
  {
    int rc_cells = RSG_OK;
    std::string err_cells;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::thread worker([&]() {
      try {
        // a new thread starts on device 0: work on the caller's device
        const hipError_t e = hipSetDevice(dev);
        rc_cells = e != hipSuccess ? fail(RSG_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e)) : cell_part();
      } catch (const std::bad_alloc&) {
        rc_cells = fail(RSG_ERR_NOMEM, "host allocation failed");
      }
      if (rc_cells) err_cells = g_last_error;
    });
    int rc_nodes;
    try {
      rc_nodes = node_part();
    } catch (const std::bad_alloc&) {
      rc_nodes = fail(RSG_ERR_NOMEM, "host allocation failed");
    }
    worker.join();
    if (rc_nodes) return rc_nodes;
    if (rc_cells) {
      g_last_error = err_cells;
      return rc_cells;
    }
  }
  CREATE_PHASE("arrays");
  // zero state (TransducerModel.cpp:150-159)
  for (int d = 0; d < 3; d++) {
    if ((rc = h->dalloc(&h->V[d], (size_t)nn * h->elt()))) return rc;
    HIP_TRY(hipMemsetAsync(h->V[d], 0, (size_t)nn * h->elt(), h->stream));
  }
  for (int s = 0; s < 6; s++) {
    if ((rc = h->dalloc(&h->S[s], (size_t)nc * h->elt()))) return rc;
    HIP_TRY(hipMemsetAsync(h->S[s], 0, (size_t)nc * h->elt(), h->stream));
  }
  if ((rc = h->partial.alloc(2 * rsg::kEnergyMaxBlocks + 2, h->stream))) return rc;
  if ((rc = finish_uploads(h.get()))) return rc;  // this handle's stream only (others may be running)
  CREATE_PHASE("state");
  *out = h.release();
  return RSG_OK;
}

int rsg_set_channels(rsg_handle h, int32_t n, const rsg_channel_desc* ch) {
  Range range_("rsg_set_channels");
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (h->in_batch) return fail(RSG_ERR_STATE, "the handle is in a batch; remove it first");
  if (n < 0 || (n > 0 && !ch)) return fail(RSG_ERR_ARG, "bad channel list");
  std::vector<Channel> list((size_t)n);
  for (int32_t k = 0; k < n; k++) {
    const rsg_channel_desc& d = ch[k];
    Channel& c = list[k];
    c.kind = d.kind;
    c.comp = d.component;
    if (d.kind == RSG_FORCE) {
      if (d.component < 0 || d.component > 2) return fail(RSG_ERR_ARG, "channel %d: force direction %d", k, d.component);
    } else if (d.kind == RSG_MOMENT) {
      if (!h->moment_form)
        return fail(RSG_ERR_ARG, "channel %d: stress sources need the HostUTModel form (moment_form=1)", k);
      if (d.component < 0 || d.component > 5) return fail(RSG_ERR_ARG, "channel %d: Voigt index %d", k, d.component);
    } else {
      return fail(RSG_ERR_ARG, "channel %d: kind %d", k, d.kind);
    }
    const int32_t lim = d.kind == RSG_FORCE ? h->nc : h->nn;
    if (d.n_support < 0 || (d.n_support > 0 && !d.support)) return fail(RSG_ERR_ARG, "channel %d: support", k);
    c.support.assign(d.support, d.support + d.n_support);
    std::sort(c.support.begin(), c.support.end());
    c.support.erase(std::unique(c.support.begin(), c.support.end()), c.support.end());
    for (int32_t id : c.support)
      if (id < 0 || id >= lim) return fail(RSG_ERR_ARG, "channel %d: support id %d out of range", k, id);
    if (d.n_signal > 0) {
      if (!d.signal_t || !d.signal_v) return fail(RSG_ERR_ARG, "channel %d: signal arrays", k);
      c.has_signal = true;
      c.sig.build(d.n_signal, d.signal_t, d.signal_v);
    }
  }
  if (h->stream) HIP_TRY(hipStreamSynchronize(h->stream));
  h->channels.swap(list);
  h->tables_dirty = true;
  return build_tables(h);
}

int rsg_set_sensors(rsg_handle h, int32_t n, const int32_t* nodes) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (h->in_batch) return fail(RSG_ERR_STATE, "the handle is in a batch; remove it first");
  if (n < 0 || (n > 0 && !nodes)) return fail(RSG_ERR_ARG, "bad sensor list");
  for (int32_t s = 0; s < n; s++)
    if (nodes[s] < 0 || nodes[s] >= h->nn) return fail(RSG_ERR_ARG, "sensor node %d out of range", nodes[s]);
  std::vector<char> seen(h->nn, 0);
  for (int32_t s = 0; s < n; s++) {
    if (seen[nodes[s]]) return fail(RSG_ERR_ARG, "sensor node %d listed twice", nodes[s]);
    seen[nodes[s]] = 1;
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->sensors.assign(nodes, nodes + n);
  h->trace_host.clear();
  h->trace_dev_steps = 0;
  int rc = h->trace.alloc((size_t)std::max(1, n) * 3 * kTraceCapSteps, h->stream);
  if (rc) return rc;
  h->tables_dirty = true;
  return build_tables(h);
}

int rsg_step(rsg_handle h, int32_t nsteps, const double* amplitudes, double* time_out) {
  Range range_("rsg_step");
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (h->in_batch) return fail(RSG_ERR_STATE, "the handle is stepped by a batch; remove it first");
  if (nsteps < 0) return fail(RSG_ERR_ARG, "nsteps < 0");
  int rc;
  if (h->tables_dirty && (rc = build_tables(h))) return rc;
  const size_t nch = h->channels.size();
  if (!amplitudes)
    for (size_t k = 0; k < nch; k++)
      if (!h->channels[k].has_signal)
        return fail(RSG_ERR_ARG, "channel %zu has no signal table and no amplitudes were given", k);
  const size_t row = h->sensors.size() * 3;
  for (int32_t done = 0; done < nsteps;) {
    int kb = std::min(kBlockSteps, nsteps - done);
    const double* amp = nch ? nullptr : h->amp[0].p;  // no channels: a buffer no term reads
    if (nch && !amplitudes && h->win_first >= 0 && h->steps >= h->win_first &&
        h->steps < h->win_first + h->win_len) {
      // the next steps' amplitudes are on the device already (the signal tables are the
      // handle's and the times follow from the same time += dt as below)
      const int64_t off = h->steps - h->win_first;
      kb = (int)std::min<int64_t>(kb, h->win_len - off);
      amp = h->amp[h->win_buf].p + off * (int64_t)nch;
    }
    if (row && h->trace_dev_steps + kb > kTraceCapSteps && (rc = flush_traces(h))) return rc;
    if (nch && !amp) {
      const int b = h->cur;
      // a table-driven upload covers a whole block ahead; caller amplitudes cover this call
      const int up = amplitudes ? kb : kBlockSteps;
      if (h->stage_busy[b]) HIP_TRY(hipEventSynchronize(h->stage_ev[b]));
      if (!h->stage[b]) {
        HIP_TRY(hipHostMalloc((void**)&h->stage[b], nch * kBlockSteps * sizeof(double), hipHostMallocDefault));
        h->stage_len[b] = nch * kBlockSteps;
      }
      double* st = h->stage[b];
      double t = h->time;
      for (int k = 0; k < up; k++) {
        for (size_t c = 0; c < nch; c++) {
          double a;
          if (amplitudes) {
            a = amplitudes[(size_t)(done + k) * nch + c];
          } else {
            double s;
            if (!h->channels[c].sig.eval(t, &s))
              return fail(RSG_ERR_ARG, "channel %zu: time %g precedes its first signal time", c, t);
            a = 0.125 * (s * 1.0);  // HostNormalXDZ::evalAt inside its window, times 0.125
          }
          st[(size_t)k * nch + c] = a;
        }
        t += h->dt;
      }
      HIP_TRY(hipMemcpyAsync(h->amp[b].p, st, (size_t)up * nch * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
      HIP_TRY(hipEventRecord(h->stage_ev[b], h->stream));
      h->stage_busy[b] = true;
      h->cur ^= 1;
      amp = h->amp[b].p;
      if (amplitudes) {
        h->win_first = -1;
      } else {
        h->win_first = h->steps;
        h->win_len = up;
        h->win_buf = b;
      }
    }
    rc = h->precision == RSG_FP64 ? run_block<double>(h, kb, amp) : run_block<float>(h, kb, amp);
    if (rc) return rc;
    for (int k = 0; k < kb; k++) h->time += h->dt;  // HostRSGTDUT::solve: time += dt
    h->steps += kb;
    done += kb;
  }
  if (time_out) *time_out = h->time;
  return RSG_OK;
}

int rsg_get_state(rsg_handle h, double* V, double* S) {
  Range range_("rsg_get_state");
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->batch_stream) HIP_TRY(hipStreamSynchronize(h->batch_stream));
  const size_t e = h->elt();
  std::vector<char> buf((size_t)std::max(h->nn, h->nc) * e);
  auto fetch = [&](void* src, int32_t n, const std::vector<int32_t>& of, double* dst) -> int {
    HIP_TRY(copy_on(buf.data(), src, (size_t)n * e, hipMemcpyDeviceToHost, h->stream));
    if (e == 8) {
      const double* b = (const double*)buf.data();
      for (int32_t i = 0; i < n; i++) dst[of[i]] = b[i];
    } else {
      const float* b = (const float*)buf.data();
      for (int32_t i = 0; i < n; i++) dst[of[i]] = (double)b[i];
    }
    return RSG_OK;
  };
  int rc;
  if (V)
    for (int d = 0; d < 3; d++)
      if ((rc = fetch(h->V[d], h->nn, h->node_of, V + (size_t)d * h->nn))) return rc;
  if (S)
    for (int s = 0; s < 6; s++)
      if ((rc = fetch(h->S[s], h->nc, h->cell_of, S + (size_t)s * h->nc))) return rc;
  return RSG_OK;
}

int rsg_set_state(rsg_handle h, const double* V, const double* S) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (h->in_batch) return fail(RSG_ERR_STATE, "the handle is in a batch; remove it first");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t e = h->elt();
  std::vector<char> buf((size_t)std::max(h->nn, h->nc) * e);
  auto put = [&](void* dst, int32_t n, const std::vector<int32_t>& of, const double* src) -> int {
    if (e == 8) {
      double* b = (double*)buf.data();
      for (int32_t i = 0; i < n; i++) b[i] = src[of[i]];
    } else {
      float* b = (float*)buf.data();
      for (int32_t i = 0; i < n; i++) b[i] = (float)src[of[i]];
    }
    HIP_TRY(copy_on(dst, buf.data(), (size_t)n * e, hipMemcpyHostToDevice, h->stream));
    return RSG_OK;
  };
  int rc;
  if (V)
    for (int d = 0; d < 3; d++)
      if ((rc = put(h->V[d], h->nn, h->node_of, V + (size_t)d * h->nn))) return rc;
  if (S)
    for (int s = 0; s < 6; s++)
      if ((rc = put(h->S[s], h->nc, h->cell_of, S + (size_t)s * h->nc))) return rc;
  return RSG_OK;
}

int rsg_get_traces(rsg_handle h, int64_t first, int64_t n, double* out) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  int rc = flush_traces(h);
  if (rc) return rc;
  const size_t row = h->sensors.size() * 3;
  const int64_t have = row ? (int64_t)(h->trace_host.size() / row) : 0;
  if (first < 0 || n < 0 || first + n > have)
    return fail(RSG_ERR_ARG, "trace steps [%lld, %lld) outside the %lld recorded", (long long)first,
                (long long)(first + n), (long long)have);
  if (n && !out) return fail(RSG_ERR_ARG, "out is null");
  if (n) memcpy(out, h->trace_host.data() + (size_t)first * row, (size_t)n * row * sizeof(double));
  return RSG_OK;
}

int rsg_synchronize(rsg_handle h) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->batch_stream) HIP_TRY(hipStreamSynchronize(h->batch_stream));
  return RSG_OK;
}

int rsg_get_time(rsg_handle h, double* time, int64_t* steps) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (time) *time = h->time;
  if (steps) *steps = h->steps;
  return RSG_OK;
}

int rsg_energy(rsg_handle h, double* kinetic, double* strain) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  int rc = energy_on_stream(h, h->partial.p, kinetic, strain);
  return rc;
}

int rsg_energy_noalloc(rsg_handle h, double* kinetic, double* strain) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (h->nn > rsg::kEnergyMaxEntities || h->nc > rsg::kEnergyMaxEntities)
    return fail(RSG_ERR_ARG, "%d nodes / %d cells: the allocation-free energy takes at most %d of each", h->nn,
                h->nc, rsg::kEnergyMaxEntities);
  if (h->tables_dirty)  // rebuilding them would allocate
    return fail(RSG_ERR_STATE, "source tables are not built: call rsg_step or rsg_energy once first");
  double* work = nullptr;
  const hipError_t e = rsg::energy_scratch(h->device, &work);
  if (e != hipSuccess) return fail(RSG_ERR_DEVICE, "energy work area: %s", hipGetErrorString(e));
  return energy_on_stream(h, work, kinetic, strain);
}

int rsg_set_timing(rsg_handle h, int32_t enabled) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (enabled < 0) return fail(RSG_ERR_ARG, "timing stride %d < 0", enabled);
  h->timing = enabled;
  return RSG_OK;
}

int rsg_get_kernel_times(rsg_handle h, double* vms, double* sms, int64_t* launches) {
  if (!h) return fail(RSG_ERR_ARG, "handle is null");
  if (vms) *vms = h->t_vel_ms;
  if (sms) *sms = h->t_str_ms;
  if (launches) *launches = h->t_launches;
  h->t_vel_ms = h->t_str_ms = 0;
  h->t_launches = 0;
  return RSG_OK;
}

int rsg_device_bytes(rsg_handle h, int64_t* bytes) {
  if (!h || !bytes) return fail(RSG_ERR_ARG, "null argument");
  *bytes = h->dev_bytes + (int64_t)(h->sp_node_val.bytes() + h->sp_node_off.bytes() +
                                    h->sp_node_sens.bytes() + h->node_terms.bytes() +
                                    h->sp_cell_val.bytes() + h->sp_cell_off.bytes() +
                                    h->cell_terms.bytes() + h->amp[0].bytes() + h->amp[1].bytes() +
                                    h->trace.bytes() + h->partial.bytes());
  return RSG_OK;
}

int rsg_destroy(rsg_handle h) {
  if (h && h->in_batch) return fail(RSG_ERR_STATE, "the handle is in a batch; remove it first");
  delete h;
  return RSG_OK;
}

int rsg_disc_support(const rsg_mesh_desc* mesh, int32_t at, const double center[3], double radius,
                     double ds, int32_t* ids, int32_t* count) {
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!center || !ids || !count || (at != 0 && at != 1)) return fail(RSG_ERR_ARG, "bad argument");
  auto sqr = [](double a) { return a * a; };
  auto check = [&](const double p[3]) {  // HostNormalXDZ::checkAt
    double res = 0.0;
    for (int k = 0; k < 3; k++) res += sqr(p[k] - center[k]);
    const double dist = std::sqrt(res);
    const double distz = std::fabs(p[2] - center[2]);
    return (distz <= 0.75 * ds) * (dist < radius) * (dist >= 0.);
  };
  int32_t m = 0;
  if (at == 0) {
    for (int32_t c = 0; c < mesh->n_cells; c++) {
      double cc[3] = {0., 0., 0.};  // HostRSGTDUT::cellcenter
      for (int kb = 0; kb < 8; kb++) {
        const int32_t v = mesh->cell_nodes[(size_t)c * 8 + kb];
        cc[0] += 0.125 * mesh->x[v];
        cc[1] += 0.125 * mesh->y[v];
        cc[2] += 0.125 * mesh->z[v];
      }
      if (check(cc)) ids[m++] = c;
    }
  } else {
    for (int32_t n = 0; n < mesh->n_nodes; n++) {
      const double p[3] = {mesh->x[n], mesh->y[n], mesh->z[n]};
      if (check(p)) ids[m++] = n;
    }
  }
  *count = m;
  return RSG_OK;
}

int rsg_signal_eval(int32_t n, const double* t, const double* v, int32_t n_eval, const double* at,
                    double* out) {
  if (n <= 0 || !t || !v || n_eval < 0 || (n_eval && (!at || !out)))
    return fail(RSG_ERR_ARG, "bad argument");
  // keys already strictly ascending (a table read back from a map): evaluated in place, with
  // no copy or sort, so per-step callers pay only the search
  bool ascending = true;
  for (int32_t i = 1; i < n && ascending; i++) ascending = t[i - 1] < t[i];
  Signal s;
  if (!ascending) s.build(n, t, v);
  const double* kt = ascending ? t : s.t.data();
  const double* kv = ascending ? v : s.v.data();
  const size_t m = ascending ? (size_t)n : s.t.size();
  for (int32_t i = 0; i < n_eval; i++)
    if (!signal_at(kt, kv, m, at[i], &out[i])) return fail(RSG_ERR_ARG, "time %g precedes the first key", at[i]);
  return RSG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches

struct rsg_batch_s {
  int32_t precision = RSG_FP64;
  int K = 50;
  int grid = 0;         // workgroups of the persistent launches
  bool dense = false;   // one workgroup per work item
  bool use_graph = true;
  // instances are scheduled in groups whose state arrays (V and S) fit this many bytes;
  // each group takes all K steps of a block before the next group of its lane starts
  // (0: one group per lane).  Set at creation: 128 MiB for FP32, 96 MiB for FP64 (measured
  // optima with two lanes on 64 instances of 32^3..128^3 cells)
  int64_t group_bytes = (int64_t)96 << 20;
  // work-table order across XCDs: -1 identity, 0 one contiguous eighth per XCD, n > 0 runs
  // of n consecutive items per XCD in turn
  int32_t xcd_run = 64;
  // an instance joins the lane with the fewest state bytes among the first `lanes`; lanes
  // are independent schedulers, each on a stream of its own, so one lane's launches fill
  // the tails of the other's and neither waits for the other at block ends
  static constexpr int kMaxLanes = 4;
  int32_t lanes = 2;
  struct Slot {
    rsg_instance* h = nullptr;
    int64_t total = 0, done = 0;
    int32_t lane = 0;
    bool attached = false;
    hipEvent_t fin = nullptr;   // recorded after the block in which the instance finishes
    bool fin_recorded = false;
  };
  std::vector<Slot> slots;
  // launch layout of a lane's block: per group, table word offsets, lengths and grids
  struct Launch {
    int64_t off_v, off_s;
    int32_t n_v, n_s;          // table lengths (padded)
    int32_t grid_v, grid_s;
    int32_t km_v, km_s;        // coefficient forms (0 arrays, 1 keys, 3 keys packed with links)
    bool lds_v, lds_s;         // LDS staging
    int32_t max_u_v, max_u_s;  // largest staged set of a member
    int32_t block_v, block_s, cpt_v, cpt_s;  // the members' workgroup sizes and per-thread counts
    bool operator==(const Launch& o) const {
      return off_v == o.off_v && off_s == o.off_s && n_v == o.n_v && n_s == o.n_s && grid_v == o.grid_v &&
             grid_s == o.grid_s && km_v == o.km_v && km_s == o.km_s && lds_v == o.lds_v && lds_s == o.lds_s &&
             max_u_v == o.max_u_v && max_u_s == o.max_u_s && block_v == o.block_v && block_s == o.block_s &&
             cpt_v == o.cpt_v && cpt_s == o.cpt_s;
    }
  };
  struct Lane {
    hipStream_t stream = nullptr;
    // control block on the device: descriptors (one per slot) | velocity tables | stress
    // tables | amplitudes.  Each group's table holds its (instance, block) items in launch
    // order (XCD order applied), padded with (-1, 0) to a multiple of kItemQuantum items so
    // small changes of membership keep the layout; tables are 256-byte aligned.  The host
    // copy is double-buffered so the next block is prepared while the last one uploads.
    size_t cap_inst = 0, cap_wv = 0, cap_ws = 0, cap_amp = 0;  // cap_w*: int32 words
    size_t off_wv = 0, off_ws = 0, off_amp = 0, ctl_bytes = 0;
    char* ctl_dev = nullptr;
    char* ctl_host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
    int cur = 0;
    std::vector<Launch> layout, graph_layout;
    hipGraph_t graph = nullptr;  // the lane's K-step block, replayed while its layout holds
    hipGraphExec_t exec = nullptr;
    // graphs replaced while their last launch may still run on the stream: each is destroyed
    // once an event recorded behind that launch has completed
    struct Retired {
      hipGraphExec_t exec;
      hipGraph_t graph;
      hipEvent_t done;
    };
    std::vector<Retired> retired;

    // Destroys the graph now: only when nothing of it is queued (the stream is idle or the
    // graph was never launched).
    void drop_graph() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
      graph_layout.clear();
    }
    // Sets the graph aside until the work queued so far on the stream has finished.
    int retire_graph() {
      if (exec || graph) {
        hipEvent_t e = nullptr;
        hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventRecord(e, stream);
        if (err != hipSuccess) {
          if (e) (void)hipEventDestroy(e);
          return fail(RSG_ERR_DEVICE, "batch graph retirement: %s", hipGetErrorString(err));
        }
        retired.push_back({exec, graph, e});
      }
      exec = nullptr;
      graph = nullptr;
      graph_layout.clear();
      return RSG_OK;
    }
    // Destroys the retired graphs whose launches have finished (all of them with wait).
    void reap(bool wait) {
      size_t keep = 0;
      for (size_t i = 0; i < retired.size(); i++) {
        Retired& r = retired[i];
        const bool finished = wait ? hipEventSynchronize(r.done) == hipSuccess : hipEventQuery(r.done) == hipSuccess;
        if (finished || wait) {
          if (r.exec) (void)hipGraphExecDestroy(r.exec);
          if (r.graph) (void)hipGraphDestroy(r.graph);
          (void)hipEventDestroy(r.done);
        } else {
          retired[keep++] = r;
        }
      }
      retired.resize(keep);
    }
    void release() {
      if (stream) (void)hipStreamSynchronize(stream);
      reap(true);
      drop_graph();
      if (ctl_dev) (void)hipFree(ctl_dev);
      ctl_dev = nullptr;
      for (int i = 0; i < 2; i++) {
        if (ctl_host[i]) (void)hipHostFree(ctl_host[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        ctl_host[i] = nullptr;
        ev[i] = nullptr;
      }
      if (stream) (void)hipStreamDestroy(stream);
      stream = nullptr;
    }
  };
  Lane lane[kMaxLanes];

  size_t desc_size() const {
    return precision == RSG_FP64 ? sizeof(rsg::BatchDesc<double>) : sizeof(rsg::BatchDesc<float>);
  }
  ~rsg_batch_s() {
    for (auto& ln : lane)
      if (ln.stream) (void)hipStreamSynchronize(ln.stream);
    for (auto& sl : slots) {
      if (sl.attached && sl.h) {
        sl.h->in_batch = false;
        sl.h->batch_stream = nullptr;
      }
      if (sl.fin) (void)hipEventDestroy(sl.fin);
    }
    for (auto& ln : lane) ln.release();
  }
};

namespace {

constexpr int32_t kItemQuantum = 256;

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
size_t padded_items(size_t n) { return (n + kItemQuantum - 1) / kItemQuantum * kItemQuantum; }
size_t table_words(size_t n) { return (2 * padded_items(n) + 63) / 64 * 64; }  // 64-word aligned

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup p -> XCD p % 8).  The
// item launched at position p of an n-item table: run < 0 keeps table order; run == 0 gives
// XCD x the x-th contiguous eighth of the table (the first n % 8 eighths one item longer);
// run > 0 gives each XCD runs of `run` consecutive items in turn (run r of XCD x is items
// [(8r + x) run, (8r + x + 1) run)), with the last n mod 8 run items in place.  Each
// mapping is a permutation of [0, n).
int32_t launch_order(int32_t p, int32_t n, int32_t run) {
  if (run < 0) return p;
  const int32_t x = p & 7, j = p >> 3;
  if (run == 0) {
    const int32_t q = n >> 3, r = n & 7;
    return x * q + std::min(x, r) + j;
  }
  const int32_t n8 = n - n % (8 * run);
  if (p >= n8) return p;
  const int32_t g = j / run;
  return (g * 8 + x) * run + (j - g * run);
}

int64_t state_bytes(const rsg_batch_s* b, const rsg_instance* h) {
  return (6 * (int64_t)h->nc + 3 * (int64_t)h->nn) * (b->precision == RSG_FP64 ? 8 : 4);
}

bool active(const rsg_batch_s::Slot& sl) { return sl.attached && sl.done < sl.total; }

// Grows a lane's control block to hold its instances (amplitude area for K steps).
int batch_reserve(rsg_batch_s* b, int l) {
  rsg_batch_s::Lane& ln = b->lane[l];
  size_t ninst = b->slots.size(), nwv = 0, nws = 0, namp = 0, nmem = 0;
  for (auto& sl : b->slots) {
    if (!sl.attached || sl.lane != l) continue;
    rsg_instance* h = sl.h;
    nwv += (size_t)(h->nn + h->block_v * h->cpt_v - 1) / (h->block_v * h->cpt_v);
    nws += (size_t)(h->nc + h->block_s * h->cpt_s - 1) / (h->block_s * h->cpt_s);
    namp += std::max<size_t>(1, h->channels.size()) * (size_t)b->K;
    nmem++;
  }
  // every group table wastes at most 2 * (kItemQuantum - 1) + 63 words; there are at most
  // as many groups as members
  const size_t slack = (size_t)(2 + 2 * kItemQuantum + 64);
  const size_t wv = 2 * nwv + slack * std::max<size_t>(1, nmem), ws = 2 * nws + slack * std::max<size_t>(1, nmem);
  if (ninst <= ln.cap_inst && wv <= ln.cap_wv && ws <= ln.cap_ws && namp <= ln.cap_amp && ln.ctl_dev)
    return RSG_OK;
  HIP_TRY(hipStreamSynchronize(ln.stream));
  ln.cap_inst = std::max(ln.cap_inst, ninst + ninst / 2 + 4);
  ln.cap_wv = std::max(ln.cap_wv, wv + wv / 2 + 4 * slack);
  ln.cap_ws = std::max(ln.cap_ws, ws + ws / 2 + 4 * slack);
  ln.cap_amp = std::max(ln.cap_amp, namp + namp / 2 + 64);
  ln.off_wv = align256(ln.cap_inst * b->desc_size());
  ln.off_ws = ln.off_wv + align256(ln.cap_wv * sizeof(int32_t));
  ln.off_amp = ln.off_ws + align256(ln.cap_ws * sizeof(int32_t));
  ln.ctl_bytes = ln.off_amp + align256(ln.cap_amp * sizeof(double));
  if (ln.ctl_dev) (void)hipFree(ln.ctl_dev);
  ln.ctl_dev = nullptr;
  HIP_TRY(hipMalloc((void**)&ln.ctl_dev, ln.ctl_bytes));
  for (int i = 0; i < 2; i++) {
    if (ln.ctl_host[i]) (void)hipHostFree(ln.ctl_host[i]);
    ln.ctl_host[i] = nullptr;
    HIP_TRY(hipHostMalloc((void**)&ln.ctl_host[i], ln.ctl_bytes, hipHostMallocDefault));
    ln.busy[i] = false;
  }
  ln.drop_graph();  // the graph's kernels point into the control block
  return RSG_OK;
}

// Enqueues one K-step block of a lane: every group in turn takes its K steps (velocity
// then stress per step).  The launches form a chain of kernel nodes in a graph built
// explicitly (not by stream capture, so other threads' HIP calls cannot disturb it) and
// replayed while the lane's launch layout stays the same.
template <typename T>
int batch_launch_block(rsg_batch_s* b, rsg_batch_s::Lane& ln) {
  const auto* desc = (const rsg::BatchDesc<T>*)ln.ctl_dev;
  const int32_t* wv = (const int32_t*)(ln.ctl_dev + ln.off_wv);
  const int32_t* ws = (const int32_t*)(ln.ctl_dev + ln.off_ws);
  auto emit = [&](hipGraph_t graph, hipGraphNode_t* node) -> hipError_t {
    for (const auto& g : ln.layout)
      for (int k = 0; k < b->K; k++) {
        hipError_t e;
        if (sizeof(T) == 8) {
          e = rsg::launch_batch_f64((const rsg::BatchDesc<double>*)desc, wv + g.off_v, g.n_v, k, g.grid_v, g.block_v,
                                    g.lds_v, g.max_u_v, g.cpt_v, g.km_v, true, ln.stream, graph, node);
          if (e != hipSuccess) return e;
          e = rsg::launch_batch_f64((const rsg::BatchDesc<double>*)desc, ws + g.off_s, g.n_s, k, g.grid_s, g.block_s,
                                    g.lds_s, g.max_u_s, g.cpt_s, g.km_s, false, ln.stream, graph, node);
        } else {
          e = rsg::launch_batch_f32((const rsg::BatchDesc<float>*)desc, wv + g.off_v, g.n_v, k, g.grid_v, g.block_v,
                                    g.lds_v, g.max_u_v, g.cpt_v, g.km_v, true, ln.stream, graph, node);
          if (e != hipSuccess) return e;
          e = rsg::launch_batch_f32((const rsg::BatchDesc<float>*)desc, ws + g.off_s, g.n_s, k, g.grid_s, g.block_s,
                                    g.lds_s, g.max_u_s, g.cpt_s, g.km_s, false, ln.stream, graph, node);
        }
        if (e != hipSuccess) return e;
      }
    return hipSuccess;
  };
  if (!b->use_graph) {
    hipError_t e = emit(nullptr, nullptr);
    if (e != hipSuccess) return fail(RSG_ERR_DEVICE, "batch launch: %s", hipGetErrorString(e));
    return RSG_OK;
  }
  ln.reap(false);
  if (!ln.exec || ln.graph_layout != ln.layout) {
    // the previous block's graph may still be running (this block's control upload is queued
    // behind it): it is destroyed only after that launch has finished
    int rc = ln.retire_graph();
    if (rc) return rc;
    HIP_TRY(hipGraphCreate(&ln.graph, 0));
    hipGraphNode_t node = nullptr;
    hipError_t e = emit(ln.graph, &node);
    if (e != hipSuccess) {
      ln.drop_graph();
      return fail(RSG_ERR_DEVICE, "batch graph: %s", hipGetErrorString(e));
    }
    HIP_TRY(hipGraphInstantiate(&ln.exec, ln.graph, nullptr, nullptr, 0));
    ln.graph_layout = ln.layout;
  }
  HIP_TRY(hipGraphLaunch(ln.exec, ln.stream));
  return RSG_OK;
}

// Fills one K-step block's control data of lane l (descriptors, group work tables,
// amplitudes) for its instances with steps left, uploads it and launches the block.
template <typename T>
int batch_block(rsg_batch_s* b, int l, int32_t* unfinished) {
  int rc;
  if ((rc = batch_reserve(b, l))) return rc;
  rsg_batch_s::Lane& ln = b->lane[l];
  const int j = ln.cur;
  if (ln.busy[j]) HIP_TRY(hipEventSynchronize(ln.ev[j]));
  char* host = ln.ctl_host[j];
  auto* desc = (rsg::BatchDesc<T>*)host;
  int32_t* wv = (int32_t*)(host + ln.off_wv);
  int32_t* ws = (int32_t*)(host + ln.off_ws);
  double* amp = (double*)(host + ln.off_amp);
  size_t amp_off = 0;
  auto mine = [&](const rsg_batch_s::Slot& sl) { return active(sl) && sl.lane == l; };
  // traces that would overflow their device buffer are flushed first (needs the stream idle)
  bool flush = false;
  for (auto& sl : b->slots)
    if (mine(sl) && !sl.h->sensors.empty() && sl.h->trace_dev_steps + b->K > kTraceCapSteps) flush = true;
  if (flush) {
    HIP_TRY(hipStreamSynchronize(ln.stream));
    for (auto& sl : b->slots)
      if (sl.attached && sl.lane == l && (rc = flush_traces(sl.h))) return rc;
  }
  std::vector<int> kmax(b->slots.size(), 0);
  for (size_t i = 0; i < b->slots.size(); i++) {
    auto& sl = b->slots[i];
    rsg::BatchDesc<T>& d = desc[i];
    memset(&d, 0, sizeof(d));
    if (!mine(sl)) continue;
    rsg_instance* h = sl.h;
    if (h->tables_dirty && (rc = build_tables(h))) return rc;
    const int km = (int)std::min<int64_t>(b->K, sl.total - sl.done);
    kmax[i] = km;
    rsg::VelocityArgs<T> va;
    rsg::StressArgs<T> sa;
    fill_args<T>(h, va, sa);
    d.v = va;
    d.s = sa;
    const size_t nch = std::max<size_t>(1, h->channels.size());
    d.nch = (int32_t)nch;
    d.amp_base = (const double*)(ln.ctl_dev + ln.off_amp) + amp_off;
    d.row = (int32_t)(h->sensors.size() * 3);
    d.trace_base = d.row ? h->trace.p : nullptr;
    d.trace_row0 = h->trace_dev_steps;
    d.kmax = km;
    double t = h->time;
    for (int k = 0; k < km; k++) {
      for (size_t c = 0; c < h->channels.size(); c++) {
        double sv;
        if (!h->channels[c].sig.eval(t, &sv))
          return fail(RSG_ERR_ARG, "channel %zu: time %g precedes its first signal time", c, t);
        amp[amp_off + (size_t)k * nch + c] = 0.125 * (sv * 1.0);
      }
      t += h->dt;
    }
    amp_off += nch * (size_t)b->K;
  }
  // groups: instances of one launch configuration (coefficient form, workgroup size,
  // per-thread count and LDS staging of each kernel, so each launch runs one kernel variant;
  // a member with a plan keeps no connectivity on the device), in slot order within a
  // configuration, each group within the state budget (an instance larger than the budget
  // forms a group of its own)
  auto config = [](const rsg_instance* h) -> std::array<int32_t, 8> {
    return {h->v_lkey ? 3 : h->nkey ? 1 : 0, h->s_lkey ? 3 : h->ckey ? 1 : 0, h->block_v, h->block_s, h->cpt_v,
            h->cpt_s, (h->v_lconn || h->v_lkey) ? 1 : 0, (h->s_lconn || h->s_lkey) ? 1 : 0};
  };
  std::vector<std::array<int32_t, 8>> configs;  // distinct configurations, first seen first
  for (size_t i = 0; i < b->slots.size(); i++)
    if (kmax[i]) {
      const auto f = config(b->slots[i].h);
      if (std::find(configs.begin(), configs.end(), f) == configs.end()) configs.push_back(f);
    }
  ln.layout.clear();
  size_t ov = 0, os = 0;
  std::vector<size_t> members;
  std::vector<int32_t> natural;  // a table's items in instance order, before launch ordering
  for (const auto& f : configs) {
    const int32_t cv = f[2] * f[4], cs = f[3] * f[5];
    for (size_t i = 0; i < b->slots.size();) {
      members.clear();
      int64_t bytes = 0;
      int32_t nv = 0, ns = 0;
      for (; i < b->slots.size(); i++) {
        if (!kmax[i]) continue;
        const rsg_instance* h = b->slots[i].h;
        if (config(h) != f) continue;
        const int64_t sb = state_bytes(b, h);
        if (!members.empty() && b->group_bytes > 0 && bytes + sb > b->group_bytes) break;
        bytes += sb;
        members.push_back(i);
        nv += (h->nn + cv - 1) / cv;
        ns += (h->nc + cs - 1) / cs;
      }
      if (members.empty()) break;
      if (ov + table_words(nv) > ln.cap_wv || os + table_words(ns) > ln.cap_ws)
        return fail(RSG_ERR_STATE, "batch work tables overflow");
      const int32_t qv = (int32_t)padded_items(nv), qs = (int32_t)padded_items(ns);
      // LDS staging for a kernel when its members have plans (all or none do), sized for the
      // largest set
      bool lds_v = true, lds_s = true;
      int32_t mu_v = 0, mu_s = 0;
      for (size_t e : members) {
        const rsg_instance* h = b->slots[e].h;
        lds_v = lds_v && (h->v_lconn || h->v_lkey);
        lds_s = lds_s && (h->s_lconn || h->s_lkey);
        mu_v = std::max(mu_v, h->v_maxu);
        mu_s = std::max(mu_s, h->s_maxu);
      }
      // items (instance, block) in slot order, then placed at their launch positions
      for (int pass = 0; pass < 2; pass++) {
        natural.clear();
        for (size_t e : members) {
          const rsg_instance* h = b->slots[e].h;
          const int32_t nb = pass == 0 ? (h->nn + cv - 1) / cv : (h->nc + cs - 1) / cs;
          for (int32_t q = 0; q < nb; q++) {
            natural.push_back(rsg::batch_item_word((int32_t)e, kmax[e]));
            natural.push_back(q);
          }
        }
        const int32_t cnt = (int32_t)(natural.size() / 2), len = pass == 0 ? qv : qs;
        int32_t* t = pass == 0 ? wv + ov : ws + os;
        for (int32_t p = 0; p < len; p++) {
          const int32_t src = launch_order(p, len, b->xcd_run);
          t[2 * p] = src < cnt ? natural[2 * src] : -1;
          t[2 * p + 1] = src < cnt ? natural[2 * src + 1] : 0;
        }
      }
      rsg_batch_s::Launch g;
      g.off_v = (int64_t)ov;
      g.off_s = (int64_t)os;
      g.n_v = qv;
      g.n_s = qs;
      g.grid_v = b->dense ? qv : b->grid;
      g.grid_s = b->dense ? qs : b->grid;
      g.km_v = f[0];
      g.km_s = f[1];
      g.block_v = f[2];
      g.block_s = f[3];
      g.cpt_v = f[4];
      g.cpt_s = f[5];
      g.lds_v = lds_v;
      g.lds_s = lds_s;
      g.max_u_v = lds_v ? mu_v : 0;
      g.max_u_s = lds_s ? mu_s : 0;
      ln.layout.push_back(g);
      ov += table_words(nv);
      os += table_words(ns);
    }
  }
  if (ln.layout.empty()) {
    *unfinished = 0;
    return RSG_OK;
  }
  HIP_TRY(hipMemcpyAsync(ln.ctl_dev, host, ln.ctl_bytes, hipMemcpyHostToDevice, ln.stream));
  HIP_TRY(hipEventRecord(ln.ev[j], ln.stream));
  ln.busy[j] = true;
  ln.cur ^= 1;
  if ((rc = batch_launch_block<T>(b, ln))) return rc;
  int32_t left = 0;
  for (size_t i = 0; i < b->slots.size(); i++) {
    auto& sl = b->slots[i];
    if (kmax[i] == 0) continue;
    rsg_instance* h = sl.h;
    for (int k = 0; k < kmax[i]; k++) h->time += h->dt;
    h->steps += kmax[i];
    if (!h->sensors.empty()) h->trace_dev_steps += kmax[i];
    sl.done += kmax[i];
    if (sl.done < sl.total) {
      left++;
    } else if (!sl.fin_recorded) {  // later blocks leave this instance alone
      HIP_TRY(hipEventRecord(sl.fin, ln.stream));
      sl.fin_recorded = true;
    }
  }
  *unfinished = left;
  return RSG_OK;
}

}  // namespace

extern "C" {

int rsg_batch_create(int32_t precision, int32_t steps_per_block, int32_t grid, int32_t use_graph,
                     rsg_batch* out) {
  if (!out) return fail(RSG_ERR_ARG, "out is null");
  *out = nullptr;
  if (precision != RSG_FP64 && precision != RSG_FP32) return fail(RSG_ERR_ARG, "bad precision %d", precision);
  if (steps_per_block < 0 || steps_per_block > 4096) return fail(RSG_ERR_ARG, "steps_per_block %d", steps_per_block);
  if (grid < -1) return fail(RSG_ERR_ARG, "grid %d", grid);
  std::unique_ptr<rsg_batch_s> b(new (std::nothrow) rsg_batch_s());
  if (!b) return fail(RSG_ERR_NOMEM, "host allocation failed");
  b->precision = precision;
  b->group_bytes = (int64_t)(precision == RSG_FP32 ? 128 : 96) << 20;
  b->K = steps_per_block ? steps_per_block : 50;
  b->use_graph = use_graph != 0;
  if (grid > 0)
    b->grid = grid;
  else
    b->dense = true;
  // the lanes' streams are created together so that they land on different hardware queues
  for (auto& ln : b->lane) {
    HIP_TRY(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) HIP_TRY(hipEventCreateWithFlags(&ln.ev[i], hipEventDisableTiming));
  }
  *out = b.release();
  return RSG_OK;
}

int rsg_batch_add(rsg_batch b, rsg_handle h, int64_t nsteps, int32_t* slot) {
  if (!b || !h || !slot) return fail(RSG_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(RSG_ERR_ARG, "nsteps < 0");
  if (h->in_batch) return fail(RSG_ERR_STATE, "the handle is already in a batch");
  if (h->precision != b->precision) return fail(RSG_ERR_ARG, "precision differs from the batch's");
  for (auto& c : h->channels)
    if (!c.has_signal) return fail(RSG_ERR_ARG, "batched instances need signal tables on every channel");
  int rc;
  if (h->tables_dirty && (rc = build_tables(h))) return rc;
  // the handle's own uploads must be complete; blocks already enqueued do not involve it
  HIP_TRY(hipStreamSynchronize(h->stream));
  int32_t id = -1;
  for (size_t i = 0; i < b->slots.size(); i++)
    if (!b->slots[i].attached) {
      id = (int32_t)i;
      break;
    }
  if (id < 0) {
    if ((int64_t)b->slots.size() >= rsg::kMaxBatchSlots)
      return fail(RSG_ERR_ARG, "a batch holds at most %d instances", rsg::kMaxBatchSlots);
    id = (int32_t)b->slots.size();
    b->slots.emplace_back();
  }
  // the lane with the fewest state bytes still to step
  int64_t load[rsg_batch_s::kMaxLanes] = {};
  for (auto& o : b->slots)
    if (active(o)) load[o.lane] += state_bytes(b, o.h);
  int32_t lane = 0;
  for (int32_t l = 1; l < b->lanes; l++)
    if (load[l] < load[lane]) lane = l;
  auto& sl = b->slots[id];
  if (!sl.fin) HIP_TRY(hipEventCreateWithFlags(&sl.fin, hipEventDisableTiming));
  sl.h = h;
  sl.total = nsteps;
  sl.done = 0;
  sl.lane = lane;
  sl.attached = true;
  sl.fin_recorded = false;
  h->in_batch = true;
  h->batch_stream = b->lane[lane].stream;
  *slot = id;
  return RSG_OK;
}

int rsg_batch_remove(rsg_batch b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].attached)
    return fail(RSG_ERR_ARG, "no such slot %d", slot);
  auto& sl = b->slots[slot];
  // a finished instance waits only for the block in which it finished, so later blocks of
  // the batch can run meanwhile; an unfinished one waits for its lane
  if (sl.fin_recorded)
    HIP_TRY(hipEventSynchronize(sl.fin));
  else
    HIP_TRY(hipStreamSynchronize(b->lane[sl.lane].stream));
  sl.h->in_batch = false;
  sl.h->batch_stream = nullptr;
  sl.h = nullptr;
  sl.total = sl.done = 0;
  sl.attached = false;
  sl.fin_recorded = false;
  return RSG_OK;
}

int rsg_batch_run(rsg_batch b, int32_t nblocks, int32_t* unfinished) {
  Range range_("rsg_batch_run");
  if (!b) return fail(RSG_ERR_ARG, "batch is null");
  int32_t left = 0;
  for (auto& sl : b->slots)
    if (active(sl)) left++;
  // one block per lane in turn, so the lanes advance together
  for (int32_t q = 0; q < nblocks && left > 0; q++) {
    left = 0;
    for (int l = 0; l < rsg_batch_s::kMaxLanes; l++) {
      bool any = false;
      for (auto& sl : b->slots) any = any || (active(sl) && sl.lane == l);
      if (!any) continue;
      int32_t lane_left = 0;
      const int rc =
          b->precision == RSG_FP64 ? batch_block<double>(b, l, &lane_left) : batch_block<float>(b, l, &lane_left);
      if (rc) return rc;
      left += lane_left;
    }
  }
  if (unfinished) *unfinished = left;
  return RSG_OK;
}

int rsg_batch_set_option(rsg_batch b, int32_t option, int64_t value) {
  if (!b) return fail(RSG_ERR_ARG, "batch is null");
  switch (option) {
    case RSG_BATCH_GROUP_BYTES:
      if (value < 0) return fail(RSG_ERR_ARG, "group bytes %lld < 0", (long long)value);
      b->group_bytes = value;
      return RSG_OK;
    case RSG_BATCH_XCD_RUN:
      if (value < -1 || value > (1 << 20)) return fail(RSG_ERR_ARG, "xcd run %lld", (long long)value);
      b->xcd_run = (int32_t)value;
      return RSG_OK;
    case RSG_BATCH_LANES:
      if (value < 1 || value > rsg_batch_s::kMaxLanes)
        return fail(RSG_ERR_ARG, "lanes %lld not in [1, %d]", (long long)value, rsg_batch_s::kMaxLanes);
      b->lanes = (int32_t)value;
      return RSG_OK;
    default:
      return fail(RSG_ERR_ARG, "unknown batch option %d", option);
  }
}

int rsg_batch_synchronize(rsg_batch b) {
  if (!b) return fail(RSG_ERR_ARG, "batch is null");
  for (auto& ln : b->lane) HIP_TRY(hipStreamSynchronize(ln.stream));
  return RSG_OK;
}

int rsg_batch_slot_status(rsg_batch b, int32_t slot, int64_t* done, int64_t* total) {
  if (!b || slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].attached)
    return fail(RSG_ERR_ARG, "no such slot %d", slot);
  if (done) *done = b->slots[slot].done;
  if (total) *total = b->slots[slot].total;
  return RSG_OK;
}

int rsg_batch_destroy(rsg_batch b) {
  delete b;
  return RSG_OK;
}

}  // extern "C"
