// rsg_kernels.hip -- CDNA4 (gfx950) kernels of the RSG time step.
//
// The step is two memory-bound gathers on VALU (no MFMA: ~0.5 flop per byte):
//   velocity (one thread per node): gathers the six stress components of the node's
//     8 connected cells and applies V += dt/rho * (div S + f)
//     (HostRSGTDUT::updateVsAtNode, HostRSGTDUT.hpp:338-362; RightHex::DD?_nc,
//     RightHex.hpp:151-192,225-231);
//   stress (one thread per cell): gathers the three velocity components of the cell's
//     8 vertices and applies S += dt * (C : grad V + m)
//     (HostRSGTDUT::updateTsAtCell, HostRSGTDUT.hpp:278-293; HostUTModel.hpp:270-311;
//     RightHex::DD?_cc, RightHex.hpp:100-141,241-247).
// Source terms and sensor recording are fused into the same passes (special entities,
// see rsg_internal.h).  Connectivity is 8 x int32 per entity read as two 16-byte loads;
// fields are structure-of-arrays, so every non-gather access of a wavefront is one
// contiguous 256/512-byte run.  Block ids are remapped so each of the 8 XCDs works on a
// contiguous slab of the (locality-ordered) entities and keeps its neighbours in its own
// L2.
//
// FP64 form: the reference's operation order exactly, with no contraction
// (-ffp-contract=off): value = sum_k (f_fwd - f_bwd); D = value*0.25/spacing (a true
// IEEE division); sums left to right.  The results are bit-identical to the reference
// in round-to-nearest.
// FP32 form: FP32 fields and updates; spacings enter as 0.25/spacing and C as dt*C,
// both computed in FP64 on the host and rounded once.
#include <hip/hip_runtime.h>

#include "rsg_internal.h"

namespace rsg {
namespace {

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i -> XCD i % 8).
// Give XCD x the contiguous logical range [x*q + min(x,r), ...) instead.
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  const int x = bid & 7, j = bid >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + j;
}

__device__ __forceinline__ bool is_boxed(double v) { return __builtin_isnan(v); }
__device__ __forceinline__ bool is_boxed(float v) { return __builtin_isnan(v); }
__device__ __forceinline__ int unbox(double v) {
  return (int)(uint32_t)(__double_as_longlong(v) & 0xFFFFFFFFll);
}
__device__ __forceinline__ int unbox(float v) { return (int)(__float_as_uint(v) & 0x3FFFFFu); }

// RightHex slot patterns: forward and backward vertices (or cell slots) per axis.
//   x: fwd {1,2,5,6} bwd {0,3,4,7};  y: fwd {2,3,6,7} bwd {0,1,4,5};
//   z: fwd {4,5,6,7} bwd {0,1,2,3}
template <int AX>
__device__ __forceinline__ double diffsum(const double (&f)[8]) {
  double v = 0.0;
  if (AX == 0) {
    v += (f[1] - f[0]); v += (f[2] - f[3]); v += (f[5] - f[4]); v += (f[6] - f[7]);
  } else if (AX == 1) {
    v += (f[2] - f[0]); v += (f[3] - f[1]); v += (f[6] - f[4]); v += (f[7] - f[5]);
  } else {
    v += (f[4] - f[0]); v += (f[5] - f[1]); v += (f[6] - f[2]); v += (f[7] - f[3]);
  }
  return v;
}
template <int AX>
__device__ __forceinline__ float diffsum(const float (&f)[8]) {
  float v;
  if (AX == 0) {
    v = ((f[1] - f[0]) + (f[2] - f[3])) + ((f[5] - f[4]) + (f[6] - f[7]));
  } else if (AX == 1) {
    v = ((f[2] - f[0]) + (f[3] - f[1])) + ((f[6] - f[4]) + (f[7] - f[5]));
  } else {
    v = ((f[4] - f[0]) + (f[5] - f[1])) + ((f[6] - f[2]) + (f[7] - f[3]));
  }
  return v;
}
// Derivative from a difference sum: FP64 (value*0.25)/spacing, FP32 value*(0.25/spacing).
__device__ __forceinline__ double dscale(double v, double sp) { return v * 0.25 / sp; }
__device__ __forceinline__ float dscale(float v, float inv) { return v * inv; }

template <typename T>
__device__ __forceinline__ void load8_cells(const T* __restrict__ f, const int (&c)[8],
                                            T (&out)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = (c[k] >= 0) ? f[c[k]] : T(0);
}
template <typename T>
__device__ __forceinline__ void load8_nodes(const T* __restrict__ f, const int (&c)[8],
                                            T (&out)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = f[c[k]];
}

typedef int v4i __attribute__((ext_vector_type(4)));

// 8 x int32 of one entity as two 16-byte loads; read once per step, so non-temporal.
__device__ __forceinline__ void load_conn(const int32_t* __restrict__ conn, int i, int (&c)[8]) {
  const v4i* p = reinterpret_cast<const v4i*>(conn) + 2 * (size_t)i;
  const v4i lo = __builtin_nontemporal_load(p);
  const v4i hi = __builtin_nontemporal_load(p + 1);
  c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w;
  c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
}

template <typename T>
__global__ __launch_bounds__(1024) void velocity_kernel(VelocityArgs<T> a) {
  const int n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= a.n) return;
  int c[8];
  load_conn(a.ncell, n, c);
  T idt = a.idt[n];
  const T sx = a.dnx[n], sy = a.dny[n], sz = a.dnz[n];

  T f[8];
  load8_cells(a.S0, c, f);
  const T dx_sxx = dscale(diffsum<0>(f), sx);
  load8_cells(a.S5, c, f);
  const T dx_sxy = dscale(diffsum<0>(f), sx);
  const T dy_sxy = dscale(diffsum<1>(f), sy);
  load8_cells(a.S4, c, f);
  const T dx_sxz = dscale(diffsum<0>(f), sx);
  const T dz_sxz = dscale(diffsum<2>(f), sz);
  load8_cells(a.S1, c, f);
  const T dy_syy = dscale(diffsum<1>(f), sy);
  load8_cells(a.S3, c, f);
  const T dy_syz = dscale(diffsum<1>(f), sy);
  const T dz_syz = dscale(diffsum<2>(f), sz);
  load8_cells(a.S2, c, f);
  const T dz_szz = dscale(diffsum<2>(f), sz);

  T r0 = (dx_sxx + dy_sxy) + dz_sxz;
  T r1 = (dx_sxy + dy_syy) + dz_syz;
  T r2 = (dx_sxz + dy_syz) + dz_szz;

  int sp = -1;
  if (is_boxed(idt)) {
    sp = unbox(idt);
    idt = a.sp_idt[sp];
    const int* off = a.sp_off + 4 * (size_t)sp;
    double fsum[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double acc = 0.0;
      for (int t = off[d]; t < off[d + 1]; t++) acc += a.amp[a.term[t]];
      fsum[d] = acc;
    }
    r0 = r0 + (T)fsum[0];
    r1 = r1 + (T)fsum[1];
    r2 = r2 + (T)fsum[2];
  }
  const T v0 = a.V0[n] + idt * r0;
  const T v1 = a.V1[n] + idt * r1;
  const T v2 = a.V2[n] + idt * r2;
  a.V0[n] = v0;
  a.V1[n] = v1;
  a.V2[n] = v2;
  if (sp >= 0) {
    const int slot = a.sp_sens[sp];
    if (slot >= 0) {
      a.trace[3 * slot + 0] = (double)v0;
      a.trace[3 * slot + 1] = (double)v1;
      a.trace[3 * slot + 2] = (double)v2;
    }
  }
}

template <typename T>
__device__ __forceinline__ T cload(const T* __restrict__ C, int64_t ldc, int j, int i) {
  return __builtin_nontemporal_load(C + (int64_t)j * ldc + i);
}

// Index of C_(I,J) (0-based Voigt indices) in the 21-entry upper triangle.
__host__ __device__ constexpr int tri(int i, int j) {
  return (i <= j) ? (i * 6 - i * (i - 1) / 2 + (j - i)) : (j * 6 - j * (j - 1) / 2 + (i - j));
}

template <typename T>
__global__ __launch_bounds__(1024) void stress_kernel(StressArgs<T> a) {
  const int e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  int c[8];
  load_conn(a.cnode, e, c);
  T sx = a.dcx[e];
  const T sy = a.dcy[e], sz = a.dcz[e];
  int sp = -1;
  if (is_boxed(sx)) {
    sp = unbox(sx);
    sx = a.sp_dcx[sp];
  }
  // delV[i][k] = D_i V_k (i: derivative axis, k: velocity component)
  T f[8];
  load8_nodes(a.V0, c, f);
  const T d00 = dscale(diffsum<0>(f), sx);
  const T d10 = dscale(diffsum<1>(f), sy);
  const T d20 = dscale(diffsum<2>(f), sz);
  load8_nodes(a.V1, c, f);
  const T d01 = dscale(diffsum<0>(f), sx);
  const T d11 = dscale(diffsum<1>(f), sy);
  const T d21 = dscale(diffsum<2>(f), sz);
  load8_nodes(a.V2, c, f);
  const T d02 = dscale(diffsum<0>(f), sx);
  const T d12 = dscale(diffsum<1>(f), sy);
  const T d22 = dscale(diffsum<2>(f), sz);
  const T e12 = d12 + d21, e02 = d02 + d20, e01 = d01 + d10;

  double m[6] = {0, 0, 0, 0, 0, 0};
  if (sp >= 0) {
    const int* off = a.sp_off + 7 * (size_t)sp;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      double acc = 0.0;
      for (int t = off[k]; t < off[k + 1]; t++) acc += a.amp[a.term[t]];
      m[k] = acc;
    }
  }

  T* Sout[6] = {a.S0, a.S1, a.S2, a.S3, a.S4, a.S5};
#pragma unroll
  for (int ks = 0; ks < 6; ks++) {
    const T c0 = cload(a.C, a.ldc, tri(ks, 0), e);
    const T c1 = cload(a.C, a.ldc, tri(ks, 1), e);
    const T c2 = cload(a.C, a.ldc, tri(ks, 2), e);
    const T c3 = cload(a.C, a.ldc, tri(ks, 3), e);
    const T c4 = cload(a.C, a.ldc, tri(ks, 4), e);
    const T c5 = cload(a.C, a.ldc, tri(ks, 5), e);
    T acc = c0 * d00 + c1 * d11 + c2 * d22 + c3 * e12 + c4 * e02 + c5 * e01;
    T s = Sout[ks][e];
    if constexpr (sizeof(T) == 8) {
      if (sp >= 0) acc = acc + (T)m[ks];
      s = s + a.dt * acc;  // C is plain C_IJ: S += dt*(C:gradV + m)
    } else {
      if (sp >= 0) acc = acc + (T)((double)a.dt * m[ks]);  // C holds dt*C_IJ
      s = s + acc;
    }
    Sout[ks][e] = s;
  }
}

// Energy: one FP64 partial per block for kinetic and stress-norm sums.
template <typename T>
__global__ __launch_bounds__(256) void energy_kernel(EnergyArgs<T> a) {
  __shared__ double sk[256], ss[256];
  double k = 0, s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.nn; i += gridDim.x * 256) {
    const double v0 = a.V0[i], v1 = a.V1[i], v2 = a.V2[i];
    k += 0.5 * a.node_w[i] * (v0 * v0 + v1 * v1 + v2 * v2);
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.nc; i += gridDim.x * 256) {
    double q = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const double x = a.S[j][i];
      q += x * x;
    }
    s += 0.5 * a.cell_w[i] * q;
  }
  sk[threadIdx.x] = k;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
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

template <typename Args, typename K>
hipError_t launch1d(K kernel, const Args& a, int n, int block, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (block < 64 || block > 1024 || (block % 64) != 0) return hipErrorInvalidValue;
  const int grid = (n + block - 1) / block;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_velocity_f64(const VelocityArgs<double>& a, int block, hipStream_t s) {
  return launch1d(velocity_kernel<double>, a, a.n, block, s);
}
hipError_t launch_velocity_f32(const VelocityArgs<float>& a, int block, hipStream_t s) {
  return launch1d(velocity_kernel<float>, a, a.n, block, s);
}
hipError_t launch_stress_f64(const StressArgs<double>& a, int block, hipStream_t s) {
  return launch1d(stress_kernel<double>, a, a.n, block, s);
}
hipError_t launch_stress_f32(const StressArgs<float>& a, int block, hipStream_t s) {
  return launch1d(stress_kernel<float>, a, a.n, block, s);
}
hipError_t launch_energy_f64(const EnergyArgs<double>& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel<double>, dim3(nblocks), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_energy_f32(const EnergyArgs<float>& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel<float>, dim3(nblocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace rsg
