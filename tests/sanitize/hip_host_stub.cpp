// hip_host_stub.cpp -- host-memory stand-ins for the HIP runtime calls and kernel launchers
// that rsg_host.cpp uses, so the library's host code (set-up threads, reordering, class
// keys, LDS plans, source tables, batch scheduling) can run under AddressSanitizer and
// ThreadSanitizer on a machine without a GPU.  TEST INFRASTRUCTURE ONLY: "device" memory
// is host memory, streams and events are tokens, every call completes at once, and the
// kernel launchers only check their arguments (nothing is computed).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../pannde_amd/csrc/rsg_internal.h"

namespace {
struct Token {
  int kind;
};
std::atomic<long> g_live{0};  // live allocations, for a leak check at exit

// With STUB_DIGEST=<file>, every device buffer's size and a hash of its contents when it is
// freed are written to <file> at exit, sorted: two builds of the set-up code produce the same
// file exactly when they upload the same arrays.
struct Digest {
  std::mutex mu;
  std::unordered_map<void*, size_t> size;
  std::vector<std::string> lines;
  const char* path = std::getenv("STUB_DIGEST");
  ~Digest() {
    if (!path) return;
    std::sort(lines.begin(), lines.end());
    if (FILE* f = std::fopen(path, "w")) {
      for (auto& l : lines) std::fprintf(f, "%s\n", l.c_str());
      std::fclose(f);
    }
  }
};
Digest& digest() {
  static Digest d;
  return d;
}
}  // namespace

extern "C" long stub_live_allocations() { return g_live.load(); }

hipError_t hipMalloc(void** p, size_t bytes) {
  if (!p) return hipErrorInvalidValue;
  *p = std::calloc(1, bytes ? bytes : 1);
  if (!*p) return hipErrorOutOfMemory;
  g_live++;
  Digest& d = digest();
  if (d.path) {
    std::lock_guard<std::mutex> lock(d.mu);
    d.size[*p] = bytes;
  }
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (p) {
    Digest& d = digest();
    if (d.path) {
      std::lock_guard<std::mutex> lock(d.mu);
      auto it = d.size.find(p);
      if (it != d.size.end()) {
        uint64_t h = 1469598103934665603ull;
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < it->second; i++) h = (h ^ b[i]) * 1099511628211ull;
        char line[64];
        std::snprintf(line, sizeof(line), "%zu %016llx", it->second, (unsigned long long)h);
        d.lines.push_back(line);
        d.size.erase(it);
      }
    }
    std::free(p);
    g_live--;
  }
  return hipSuccess;
}
// the stream-ordered pool: its allocations are host memory like hipMalloc's
hipError_t hipMemPoolCreate(hipMemPool_t* pool, const hipMemPoolProps*) {
  static int token;
  *pool = reinterpret_cast<hipMemPool_t>(&token);
  return hipSuccess;
}
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
hipError_t hipMallocFromPoolAsync(void** p, size_t bytes, hipMemPool_t, hipStream_t) { return hipMalloc(p, bytes); }
hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned int) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind) {
  if (n) std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(dst, src, n, k);
}
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t) {
  std::memset(dst, v, n);
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = 64 * 1024;  // the only attribute asked for: LDS bytes per workgroup
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "stub error"; }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new Token{1});
  g_live++;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<Token*>(s);
  g_live--;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(new Token{2});
  g_live++;
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<Token*>(e);
  g_live--;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.001f;
  return hipSuccess;
}

hipError_t hipGraphCreate(hipGraph_t* g, unsigned int) {
  *g = reinterpret_cast<hipGraph_t>(new Token{3});
  g_live++;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<Token*>(g);
  g_live--;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  *x = reinterpret_cast<hipGraphExec_t>(new Token{4});
  g_live++;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete reinterpret_cast<Token*>(x);
  g_live--;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }

// roctx ranges: no profiler here
extern "C" int roctxRangePushA(const char*) { return 0; }
extern "C" int roctxRangePop() { return 0; }

namespace rsg {

// The launchers check what the real ones check and touch every table entry a kernel would
// read first (so an out-of-range table shows up under the sanitizer); nothing is computed.
// links: a staged launch reads the block-local links, an unstaged one the connectivity
template <typename A>
static bool has_links(const A& a, const int32_t* conn, bool lds) {
  return lds ? (a.lconn != nullptr || a.lkey != nullptr) : conn != nullptr;
}
template <typename A>
static hipError_t check_block(const A& a, int block, const int32_t* conn) {
  if (block < 64 || block > 512 || block % 64 || a.cpt < 1) return hipErrorInvalidValue;
  if (!has_links(a, conn, a.lconn != nullptr || a.lkey != nullptr)) return hipErrorInvalidValue;
  return hipSuccess;
}
hipError_t launch_velocity_f64(const VelocityArgs<double>& a, int block, hipStream_t) {
  return check_block(a, block, a.ncell);
}
hipError_t launch_velocity_f32(const VelocityArgs<float>& a, int block, hipStream_t) {
  return check_block(a, block, a.ncell);
}
hipError_t launch_stress_f64(const StressArgs<double>& a, int block, hipStream_t) { return check_block(a, block, a.cnode); }
hipError_t launch_stress_f32(const StressArgs<float>& a, int block, hipStream_t) { return check_block(a, block, a.cnode); }

template <typename T>
static hipError_t batch(const BatchDesc<T>* d, const int32_t* table, int n, int block, bool lds, bool velocity,
                        hipGraph_t graph, hipGraphNode_t* node) {
  if (block < 64 || block > 512 || block % 64) return hipErrorInvalidValue;
  volatile long sink = 0;
  for (int i = 0; i < n; i++) {
    const int32_t w = table[2 * i];
    if (w < 0) continue;
    const BatchDesc<T>& m = d[w & (kMaxBatchSlots - 1)];
    sink = sink + m.kmax;
    // every member of a launch has what the launch's form reads
    if (velocity ? !has_links(m.v, m.v.ncell, lds) : !has_links(m.s, m.s.cnode, lds)) return hipErrorInvalidValue;
  }
  if (graph && node) *node = reinterpret_cast<hipGraphNode_t>(graph);
  return hipSuccess;
}
hipError_t launch_batch_f64(const BatchDesc<double>* d, const int32_t* t, int n, int, int, int block, bool lds, int,
                            int, int, bool velocity, hipStream_t, hipGraph_t g, hipGraphNode_t* node) {
  return batch(d, t, n, block, lds, velocity, g, node);
}
hipError_t launch_batch_f32(const BatchDesc<float>* d, const int32_t* t, int n, int, int, int block, bool lds, int,
                            int, int, bool velocity, hipStream_t, hipGraph_t g, hipGraphNode_t* node) {
  return batch(d, t, n, block, lds, velocity, g, node);
}
hipError_t energy_scratch(int device, double** out) {
  static double area[2 * kEnergyMaxBlocks + 2];
  if (device != 0 || !out) return hipErrorInvalidValue;
  *out = area;
  return hipSuccess;
}
hipError_t launch_energy_f64(const EnergyArgs<double>& a, int nb, double* out, hipStream_t) {
  if (nb <= 0 || nb > kEnergyMaxBlocks || !a.partial || !out) return hipErrorInvalidValue;
  out[0] = out[1] = 0.0;
  return hipSuccess;
}
hipError_t launch_energy_f32(const EnergyArgs<float>& a, int nb, double* out, hipStream_t) {
  if (nb <= 0 || nb > kEnergyMaxBlocks || !a.partial || !out) return hipErrorInvalidValue;
  out[0] = out[1] = 0.0;
  return hipSuccess;
}

}  // namespace rsg
