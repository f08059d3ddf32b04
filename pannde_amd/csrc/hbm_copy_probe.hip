// hbm_copy_probe.hip -- a plain device-to-device copy of known size.
//
// Used for two measurements on the target GPU: (1) the copy bandwidth HBM actually
// sustains, the practical ceiling for the gather kernels; (2) the calibration of the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters, which are read for this kernel (known
// bytes: n*8 read, n*8 written per launch) and for the RSG kernels under the same
// counter, so that counter units convert to bytes.
//
// Usage: hbm_copy_probe [MiB per buffer (1024)] [launches (20)] [blocks per CU (4; 0 = one
//        pass per thread)] [elements per thread per pass U in 1,2,4,8 (1)] [non-temporal 0/1 (0)]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include <algorithm>

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                     \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

typedef double d2 __attribute__((ext_vector_type(2)));

// Each thread moves U 16-byte elements per pass (all loads issued before the stores);
// NT selects non-temporal loads and stores.
template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const d2* __restrict__ src,
                                                   d2* __restrict__ dst, long n2) {
  const long stride = (long)gridDim.x * blockDim.x * U;
  for (long base = (long)blockIdx.x * blockDim.x * U + threadIdx.x; base < n2; base += stride) {
    d2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long i = base + (long)u * blockDim.x;
      if (i < n2) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long i = base + (long)u * blockDim.x;
      if (i < n2) {
        if (NT)
          __builtin_nontemporal_store(v[u], dst + i);
        else
          dst[i] = v[u];
      }
    }
  }
}

template <int U, bool NT>
void launch(int grid, const d2* a, d2* b, long n2) {
  hipLaunchKernelGGL((copy_kernel<U, NT>), dim3(grid), dim3(256), 0, 0, a, b, n2);
}
typedef void (*launcher)(int, const d2*, d2*, long);

int main(int argc, char** argv) {
  const long mib = argc > 1 ? atol(argv[1]) : 1024;
  const int launches = argc > 2 ? atoi(argv[2]) : 20;
  const int per_cu = argc > 3 ? atoi(argv[3]) : 4;
  const int U = argc > 4 ? atoi(argv[4]) : 1;
  const int nt = argc > 5 ? atoi(argv[5]) : 0;
  launcher L = nullptr;
  if (nt)
    L = U == 1 ? launch<1, true> : U == 2 ? launch<2, true> : U == 4 ? launch<4, true> : U == 8 ? launch<8, true> : nullptr;
  else
    L = U == 1 ? launch<1, false> : U == 2 ? launch<2, false> : U == 4 ? launch<4, false> : U == 8 ? launch<8, false> : nullptr;
  if (!L || per_cu < 0 || per_cu > 64) {
    fprintf(stderr, "bad arguments\n");
    return 2;
  }
  if (mib <= 0 || mib > 65536 || launches <= 0 || launches > 10000) {
    fprintf(stderr, "bad arguments\n");
    return 2;
  }
  const long bytes = mib << 20;
  const long n2 = bytes / 16;
  d2 *a = nullptr, *b = nullptr;
  CHECK(hipMalloc((void**)&a, bytes));
  CHECK(hipMalloc((void**)&b, bytes));
  CHECK(hipMemset(a, 0, bytes));
  CHECK(hipMemset(b, 0, bytes));
  const long one_pass = (n2 + 256L * U - 1) / (256L * U);
  const int grid = per_cu ? (int)std::min<long>(256L * per_cu, one_pass) : (int)one_pass;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  L(grid, a, b, n2);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0, 0));
  for (int k = 0; k < launches; k++) L(grid, a, b, n2);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double per = ms * 1e-3 / launches;
  printf("COPY bytes_read=%ld bytes_written=%ld launches=%d grid=%d U=%d nt=%d mean_us=%.2f GBps=%.1f\n",
         bytes, bytes, launches + 1, grid, U, nt, per * 1e6, 2.0 * bytes / per / 1e9);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
