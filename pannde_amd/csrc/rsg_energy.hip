// rsg_energy.hip -- discrete energy of an RSG state on the device, accumulated in FP64.
//
//   kinetic = 1/2 sum_n  m_n |V_n|^2,         m_n = rho_n * dxn * dyn * dzn
//   strain  = 1/2 sum_c  vol_c  S_c . C_c^-1 S_c,   vol_c = dxc * dyc * dzc
//
// rho_n is the density at the node as the velocity update uses it (HostField::atNode: the
// /8 rule, HostField.hpp:142-154) and dxn..dzn are RightHex::nodePrep's spacings
// (RightHex.hpp:258-309), which are half the extent of the node's cells, so m_n is the
// node's share of the mass: a boundary node's density already counts its missing cells as
// zero and the masses sum to the body's mass.  These weights pair the node-to-cell and
// cell-to-node differences of the scheme (RightHex.hpp:100-192) by summation by parts on a
// uniform mesh, so the sum is the energy the step conserves once sources stop.  S is the
// Voigt stress (Sxx,Syy,Szz,Syz,Sxz,Sxy) and C the 6x6 stiffness that multiplies the strain
// rates (d00, d11, d22, d12+d21, d02+d20, d01+d10) in HostRSGTDUT::updateTsAtCell
// (HostRSGTDUT.hpp:278-293), so S.C^-1.S is twice the strain-energy density.  It is found per
// cell in FP64 from an LDL^T factorisation of C; the FP32 form stores dt*C rounded to FP32 and
// the factorisation uses that value divided by dt in FP64.  A cell whose C is not positive
// definite contributes NaN.
//
// The reference computes no energy; this is the FP64 accumulation the mixed-precision form
// keeps next to its sensor traces.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rsg_internal.h"

// Work area of rsg_energy_noalloc: the partials of kEnergyMaxBlocks workgroups and the two
// totals (16 KiB), one copy per GPU for the life of the process.  Host code never uses this
// name as an address: energy_scratch() resolves each device's copy with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[2 * rsg::kEnergyMaxBlocks + 2];

namespace rsg {
namespace {

constexpr int kEnergyThreads = 256;

// The 21 stored C of cell e (upper triangle, row by row) in FP64.
template <typename T>
__device__ void cell_stiffness(const StressArgs<T>& a, double inv_scale, int e, double (&c)[21]) {
  const T* row = nullptr;
  if (a.key || a.lkey) {
    uint32_t k = a.lkey ? a.lkey[4 * (size_t)e + 3] : a.key[e];
    if (k & kSpecialKey) k = a.sp_key[k & ~kSpecialKey];
    row = a.cls + (size_t)(k >> kMatShift) * kMatRow;
#pragma unroll
    for (int j = 0; j < 21; j++) c[j] = (double)row[j] * inv_scale;
  } else {
#pragma unroll
    for (int j = 0; j < 21; j++) {
      const int64_t idx = a.ldc ? (int64_t)j * a.ldc + e : ((int64_t)(e >> 6) * 21 + j) * 64 + (e & 63);
      c[j] = (double)a.C[idx] * inv_scale;
    }
  }
}

// s . C^-1 s for a symmetric positive-definite C given by its upper triangle: with
// C = L D L^T (L unit lower triangular) and L y = s, s . C^-1 s = sum_i y_i^2 / D_i.  Fully
// unrolled, so the matrix stays in registers.  NaN when C is not positive definite.
__device__ __forceinline__ double compliance_product(const double (&c)[21], const double (&s)[6]) {
  double m[6][6];
  int t = 0;
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = i; j < 6; j++) {
      m[i][j] = c[t];
      m[j][i] = c[t];
      t++;
    }
  double L[6][6], D[6];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    double d = m[j][j];
#pragma unroll
    for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k] * D[k];
    ok = ok && d > 0.0;
    D[j] = d;
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      double v = m[i][j];
#pragma unroll
      for (int k = 0; k < j; k++) v -= L[i][k] * L[j][k] * D[k];
      L[i][j] = v / d;
    }
  }
  double y[6], q = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    double v = s[i];
#pragma unroll
    for (int k = 0; k < i; k++) v -= L[i][k] * y[k];
    y[i] = v;
    q += v * v / D[i];
  }
  return ok ? q : __builtin_nan("");
}

// One FP64 partial per workgroup: kinetic at partial[b], strain at partial[nblocks + b].
template <typename T>
__global__ __launch_bounds__(kEnergyThreads) void energy_kernel(EnergyArgs<T> a) {
  __shared__ double sk[kEnergyThreads], ss[kEnergyThreads];
  double k = 0, s = 0;
  const int stride = (int)gridDim.x * kEnergyThreads;
  for (int i = blockIdx.x * kEnergyThreads + threadIdx.x; i < a.nn; i += stride) {
    const double v0 = a.V0[i], v1 = a.V1[i], v2 = a.V2[i];
    k += 0.5 * a.node_m[i] * (v0 * v0 + v1 * v1 + v2 * v2);
  }
  const T* S[6] = {a.s.S0, a.s.S1, a.s.S2, a.s.S3, a.s.S4, a.s.S5};
  for (int e = blockIdx.x * kEnergyThreads + threadIdx.x; e < a.s.n; e += stride) {
    double c[21], sv[6];
    cell_stiffness(a.s, a.inv_scale, e, c);
#pragma unroll
    for (int j = 0; j < 6; j++) sv[j] = (double)S[j][e];
    s += 0.5 * a.cell_vol[e] * compliance_product(c, sv);
  }
  sk[threadIdx.x] = k;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int w = kEnergyThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sk[threadIdx.x] += sk[threadIdx.x + w];
      ss[threadIdx.x] += ss[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.partial[blockIdx.x] = sk[0];
    a.partial[gridDim.x + blockIdx.x] = ss[0];
  }
}

// Sums the partials of energy_kernel in a fixed order: out[0] kinetic, out[1] strain.
__global__ __launch_bounds__(kEnergyThreads) void energy_total(const double* __restrict__ partial, int nblocks,
                                                               double* __restrict__ out) {
  __shared__ double sk[kEnergyThreads], ss[kEnergyThreads];
  double k = 0, s = 0;
  for (int i = threadIdx.x; i < nblocks; i += kEnergyThreads) {
    k += partial[i];
    s += partial[nblocks + i];
  }
  sk[threadIdx.x] = k;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int w = kEnergyThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sk[threadIdx.x] += sk[threadIdx.x + w];
      ss[threadIdx.x] += ss[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sk[0];
    out[1] = ss[0];
  }
}

template <typename T>
hipError_t launch_energy(const EnergyArgs<T>& a, int nblocks, double* out, hipStream_t st) {
  if (nblocks <= 0 || nblocks > kEnergyMaxBlocks || !a.partial || !out) return hipErrorInvalidValue;
  hipLaunchKernelGGL(energy_kernel<T>, dim3(nblocks), dim3(kEnergyThreads), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(energy_total, dim3(1), dim3(kEnergyThreads), 0, st, a.partial, nblocks, out);
  return hipGetLastError();
}

}  // namespace

hipError_t energy_scratch(int device, double** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices || !out) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    // resolved once per device, with that device current
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) {
      const hipError_t r = hipSetDevice(cur);
      if (e == hipSuccess) e = r;
    }
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

hipError_t launch_energy_f64(const EnergyArgs<double>& a, int nblocks, double* out, hipStream_t s) {
  return launch_energy<double>(a, nblocks, out, s);
}
hipError_t launch_energy_f32(const EnergyArgs<float>& a, int nblocks, double* out, hipStream_t s) {
  return launch_energy<float>(a, nblocks, out, s);
}

}  // namespace rsg
