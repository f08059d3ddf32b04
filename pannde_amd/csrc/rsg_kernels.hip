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

#include <type_traits>

#include "rsg_internal.h"

// Occupancy hints (minimum waves per SIMD, 0 = compiler default).
// Largest workgroup the gather kernels accept.  It bounds the registers the compiler may
// give a thread (512 threads -> 2 waves per SIMD -> up to 256 VGPRs); 1024 would cap them
// at 128, which makes the FP64 stress kernel spill to scratch.
#ifndef RSG_MAX_BLOCK
#define RSG_MAX_BLOCK 512
#endif
#ifndef RSG_STRESS_WPE
#define RSG_STRESS_WPE 0
#endif
#ifndef RSG_VELOCITY_WPE
#define RSG_VELOCITY_WPE 0
#endif
#if RSG_STRESS_WPE > 0
#define RSG_STRESS_ATTR __attribute__((amdgpu_waves_per_eu(RSG_STRESS_WPE)))
#else
#define RSG_STRESS_ATTR
#endif
#if RSG_VELOCITY_WPE > 0
#define RSG_VELOCITY_ATTR __attribute__((amdgpu_waves_per_eu(RSG_VELOCITY_WPE)))
#else
#define RSG_VELOCITY_ATTR
#endif

// Loads of data read once per step (links, keys, gather lists, per-entity C): non-temporal
// by default; RSG_CACHED_LINKS=1 makes them ordinary loads so they can stay cached.
#ifndef RSG_CACHED_LINKS
#define RSG_CACHED_LINKS 0
#endif
#if RSG_CACHED_LINKS
#define STREAM_LOAD(p) (*(p))
#else
#define STREAM_LOAD(p) __builtin_nontemporal_load(p)
#endif

// Velocity staging layout: 1 (default) keeps the six staged stress components as three rows
// of pairs, so one LDS read returns two components of a cell; 0 keeps six rows.
#ifndef RSG_VELOCITY_PAIRS
#define RSG_VELOCITY_PAIRS 1
#endif
// Stress staging layout: the three staged velocity components as a row of (Vx, Vy) pairs and
// a row of Vz (one read fewer per vertex) in FP32 (1), in both precisions (2), or as three
// rows (0).
#ifndef RSG_STRESS_PAIRS
#define RSG_STRESS_PAIRS 2
#endif

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

// Cells around a node: a missing cell (-1) contributes zero.  LDS-staged gathers need no
// test, because their missing links point at a staged entry that holds zero.
template <bool DIRECT, typename T>
__device__ __forceinline__ void load8_cells(const T* __restrict__ f, const int (&c)[8],
                                            T (&out)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = DIRECT ? ((c[k] >= 0) ? f[c[k]] : T(0)) : f[c[k]];
}
// The same from a staged pair row: f gets the first component of each pair, g the second.
template <typename T>
__device__ __forceinline__ void load8_pairs(const T* rows, const int (&c)[8], T (&f)[8], T (&g)[8]) {
  typedef T P __attribute__((ext_vector_type(2)));
  const P* q = reinterpret_cast<const P*>(rows);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const P v = q[c[k]];
    f[k] = v.x;
    g[k] = v.y;
  }
}
template <typename T>
__device__ __forceinline__ void load8_nodes(const T* __restrict__ f, const int (&c)[8],
                                            T (&out)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = f[c[k]];
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

// One material row (16-byte aligned) as 16-byte loads through the caches: lanes of a
// wavefront that share a material read the same addresses, which coalesce.
template <int W>
__device__ __forceinline__ void load_row(const double* __restrict__ p, double (&out)[W]) {
  static_assert(W % 2 == 0, "row width");
  const d2v* q = reinterpret_cast<const d2v*>(p);
#pragma unroll
  for (int i = 0; i < W / 2; i++) {
    const d2v v = q[i];
    out[2 * i] = v.x;
    out[2 * i + 1] = v.y;
  }
}
template <int W>
__device__ __forceinline__ void load_row(const float* __restrict__ p, float (&out)[W]) {
  static_assert(W % 4 == 0, "row width");
  const f4v* q = reinterpret_cast<const f4v*>(p);
#pragma unroll
  for (int i = 0; i < W / 4; i++) {
    const f4v v = q[i];
    out[4 * i] = v.x;
    out[4 * i + 1] = v.y;
    out[4 * i + 2] = v.z;
    out[4 * i + 3] = v.w;
  }
}

// Coefficient form of a kernel instance: KM = 0 per-entity arrays, 1 class-table keys,
// 3 keys packed with the LDS links (a batch launches each form for its own members).
template <int KM, typename A>
__device__ __forceinline__ constexpr bool packed(const A&) {
  return KM == 3;
}
template <int KM, typename A>
__device__ __forceinline__ constexpr bool keyed(const A&) {
  return KM == 1 || KM == 3;
}

// Material/spacing key of an entity, and its special index (or -1).
template <typename A>
__device__ __forceinline__ uint32_t resolve_key(const A& a, uint32_t key, int& sp) {
  if (key & kSpecialKey) {
    sp = (int)(key & ~kSpecialKey);
    return a.sp_key[sp];
  }
  sp = -1;
  return key;
}
template <typename P, typename T>
__device__ __forceinline__ void axis_spacings(P axis, uint32_t k, T& sx, T& sy, T& sz) {
  sx = axis[(k >> 14) & (kAxisClasses - 1)];
  sy = axis[kAxisClasses + ((k >> 7) & (kAxisClasses - 1))];
  sz = axis[2 * kAxisClasses + (k & (kAxisClasses - 1))];
}

// The class tables are written only by the host, so a wavefront-uniform read of them may go
// through the scalar cache: a pointer in the constant address space lets the compiler use
// scalar loads (reads only; nothing is ever stored through this path).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

// 8 x int32 of one entity as two 16-byte loads; read once per step, so non-temporal.
__device__ __forceinline__ void load_conn(const int32_t* __restrict__ conn, int i, int (&c)[8]) {
  const v4i* p = reinterpret_cast<const v4i*>(conn) + 2 * (size_t)i;
  const v4i lo = STREAM_LOAD(p);
  const v4i hi = STREAM_LOAD(p + 1);
  c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w;
  c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
}

// 8 x uint16 block-local indices of one entity as one 16-byte load (0xFFFF = none, returned
// as `none`).
__device__ __forceinline__ void load_lconn(const uint16_t* __restrict__ lconn, int i, int (&c)[8], int none) {
  const v4i w = STREAM_LOAD(reinterpret_cast<const v4i*>(lconn) + (size_t)i);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t u = (uint32_t)w[k];
    const int lo = (int)(u & 0xFFFFu), hi = (int)(u >> 16);
    c[2 * k] = lo == 0xFFFF ? none : lo;
    c[2 * k + 1] = hi == 0xFFFF ? none : hi;
  }
}

// 8 block-local indices (10 bits each, kNoLink = none, returned as `none`) and the key of one
// entity, one 16-byte load.
__device__ __forceinline__ uint32_t load_lkey(const uint32_t* __restrict__ lkey, int i, int (&c)[8], int none) {
  const v4i w = STREAM_LOAD(reinterpret_cast<const v4i*>(lkey) + (size_t)i);
  const uint32_t w0 = (uint32_t)w.x, w1 = (uint32_t)w.y, w2 = (uint32_t)w.z;
  const uint32_t l[8] = {w0 & kNoLink, (w0 >> 10) & kNoLink, (w0 >> 20) & kNoLink, w1 & kNoLink,
                         (w1 >> 10) & kNoLink, (w1 >> 20) & kNoLink, w2 & kNoLink, (w2 >> 10) & kNoLink};
#pragma unroll
  for (int k = 0; k < 8; k++) c[k] = l[k] == kNoLink ? none : (int)l[k];
  return (uint32_t)w.w;
}

// Staged rows: with PAIRS, components 2p and 2p + 1 of entry i sit side by side at pair p * ld
// + i (NCOMP / 2 rows of pairs; an odd last component keeps a row of its own after them, at
// (NCOMP - 1) * ld + i), the same bytes as NCOMP rows, so one LDS read returns two components.
template <typename T, int NCOMP, bool PAIRS>
__device__ __forceinline__ void put_staged(T* lds, int ld, int i, const T (&v)[NCOMP]) {
  if constexpr (PAIRS) {
    typedef T P __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int p = 0; p < NCOMP / 2; p++) reinterpret_cast<P*>(lds)[p * ld + i] = P{v[2 * p], v[2 * p + 1]};
    if constexpr (NCOMP % 2) lds[(NCOMP - 1) * ld + i] = v[NCOMP - 1];
  } else {
#pragma unroll
    for (int k = 0; k < NCOMP; k++) lds[k * ld + i] = v[k];
  }
}
template <typename T, int NCOMP, bool PAIRS>
__device__ __forceinline__ void zero_staged(T* lds, int ld, int zero) {
  if (zero < 0 || (int)threadIdx.x >= NCOMP) return;
  const int k = threadIdx.x;
  lds[PAIRS && k < NCOMP - NCOMP % 2 ? 2 * ((k >> 1) * ld + zero) + (k & 1) : k * ld + zero] = T(0);
}

// Stages NCOMP fields of the block's unique gather set (ulist[u_off[b]..u_off[b+1]))
// into LDS as [NCOMP][ld] (with PAIRS as [NCOMP / 2][ld] pairs), each value read from HBM
// once per block.  With U > 0 every thread owns at most U entries and all their loads are
// issued before any wait (two dependent round trips in all: the list, then the values);
// U = 0 is the general loop.  zero >= 0 also stores 0 at entry `zero` of every row (the
// target of missing links; it lies past every block's set).  With CHUNKS and `chunked` the
// list holds one word per 4 LDS slots, start << 2 | (len - 1), and slot 4c + k takes entity
// start + k when k < len (the other slots of the chunk stay unwritten: no link points at
// them).
template <typename T, int NCOMP, int U, bool CHUNKS = false, bool PAIRS = false>
__device__ __forceinline__ void stage(T* lds, int ld, const int32_t* __restrict__ u_off,
                                      const int32_t* __restrict__ ulist, int b,
                                      const T* const (&src)[NCOMP], int zero, bool chunked = false) {
  if constexpr (CHUNKS) {
    if (chunked) {
      const int c0 = u_off[b], nu = 4 * (u_off[b + 1] - c0);
      if constexpr (U > 0) {
        int g[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int i = threadIdx.x + u * blockDim.x;
          const int w = STREAM_LOAD(ulist + c0 + (i < nu ? (i >> 2) : 0));
          on[u] = i < nu && (i & 3) <= (w & 3);
          g[u] = on[u] ? (w >> 2) + (i & 3) : 0;
        }
        T v[U][NCOMP];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int k = 0; k < NCOMP; k++) v[u][k] = on[u] ? src[k][g[u]] : T(0);
#pragma unroll
        for (int u = 0; u < U; u++)
          if (on[u]) put_staged<T, NCOMP, PAIRS>(lds, ld, threadIdx.x + u * blockDim.x, v[u]);
      } else {
        for (int i = threadIdx.x; i < nu; i += blockDim.x) {
          const int w = STREAM_LOAD(ulist + c0 + (i >> 2));
          if ((i & 3) > (w & 3)) continue;
          const int g = (w >> 2) + (i & 3);
          T v[NCOMP];
#pragma unroll
          for (int k = 0; k < NCOMP; k++) v[k] = src[k][g];
          put_staged<T, NCOMP, PAIRS>(lds, ld, i, v);
        }
      }
      zero_staged<T, NCOMP, PAIRS>(lds, ld, zero);
      __syncthreads();
      return;
    }
  }
  const int u0 = u_off[b], nu = u_off[b + 1] - u0;
  if constexpr (U > 0) {
    int g[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      g[u] = STREAM_LOAD(ulist + u0 + (i < nu ? i : 0));
    }
    T v[U][NCOMP];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < NCOMP; k++) v[u][k] = src[k][g[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nu) put_staged<T, NCOMP, PAIRS>(lds, ld, i, v[u]);
    }
  } else {
    for (int i = threadIdx.x; i < nu; i += blockDim.x) {
      const int g = STREAM_LOAD(ulist + u0 + i);
      T v[NCOMP];
#pragma unroll
      for (int k = 0; k < NCOMP; k++) v[k] = src[k][g];
      put_staged<T, NCOMP, PAIRS>(lds, ld, i, v);
    }
  }
  zero_staged<T, NCOMP, PAIRS>(lds, ld, zero);
  __syncthreads();
}

// Per-node operands of the velocity update that do not depend on a gather.
template <typename T>
struct VelIn {
  int c[8];
  uint32_t key;
  T idt, sx, sy, sz, v0, v1, v2;
};

template <typename T, bool LDS, int KM>
__device__ __forceinline__ void velocity_load(const VelocityArgs<T>& a, int n, VelIn<T>& in) {
  if constexpr (LDS) {
    if (packed<KM>(a))
      in.key = load_lkey(a.lkey, n, in.c, a.max_u);  // missing cells -> the zero entry
    else
      load_lconn(a.lconn, n, in.c, a.max_u);
  } else {
    load_conn(a.ncell, n, in.c);
  }
  if (keyed<KM>(a)) {
    if (!(LDS && packed<KM>(a))) in.key = STREAM_LOAD(a.key + n);
    in.idt = in.sx = in.sy = in.sz = T(0);
  } else {
    in.key = 0;
    in.idt = a.idt[n];
    in.sx = a.dnx[n];
    in.sy = a.dny[n];
    in.sz = a.dnz[n];
  }
  in.v0 = a.V0[n];
  in.v1 = a.V1[n];
  in.v2 = a.V2[n];
}

// V += dt/rho * (div S + f) at node n, S gathered from S0..S5 (LDS or HBM).
template <typename T, bool LDS, int KM>
__device__ __forceinline__ void velocity_apply(const VelocityArgs<T>& a, int n, const VelIn<T>& in,
                                               const T* S0, const T* S1, const T* S2, const T* S3,
                                               const T* S4, const T* S5) {
  // coefficients: the node's key into the class tables, or the per-node arrays (special
  // nodes NaN-boxed)
  T idt, sx, sy, sz;
  int sp = -1;
  if (keyed<KM>(a)) {
    const uint32_t k = resolve_key(a, in.key, sp);
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
    if (__all(k == k0)) {  // one class for the whole wavefront: scalar loads
      idt = as_const(a.cls)[k0 >> kMatShift];
      axis_spacings(as_const(a.axis), k0, sx, sy, sz);
    } else {
      idt = a.cls[k >> kMatShift];
      axis_spacings(a.axis, k, sx, sy, sz);
    }
  } else {
    idt = in.idt;
    sx = in.sx;
    sy = in.sy;
    sz = in.sz;
    if (is_boxed(idt)) {
      sp = unbox(idt);
      idt = a.sp_idt[sp];
    }
  }
  T dx_sxx, dx_sxy, dy_sxy, dx_sxz, dz_sxz, dy_syy, dy_syz, dz_syz, dz_szz;
  if constexpr (LDS && RSG_VELOCITY_PAIRS) {
    // S0, S1, S2 are the staged pair rows (Sxx, Sxy), (Sxz, Syy), (Syz, Szz)
    T f[8], g[8];
    load8_pairs(S0, in.c, f, g);
    dx_sxx = dscale(diffsum<0>(f), sx);
    dx_sxy = dscale(diffsum<0>(g), sx);
    dy_sxy = dscale(diffsum<1>(g), sy);
    load8_pairs(S1, in.c, f, g);
    dx_sxz = dscale(diffsum<0>(f), sx);
    dz_sxz = dscale(diffsum<2>(f), sz);
    dy_syy = dscale(diffsum<1>(g), sy);
    load8_pairs(S2, in.c, f, g);
    dy_syz = dscale(diffsum<1>(f), sy);
    dz_syz = dscale(diffsum<2>(f), sz);
    dz_szz = dscale(diffsum<2>(g), sz);
  } else {
    T f[8];
    load8_cells<!LDS>(S0, in.c, f);
    dx_sxx = dscale(diffsum<0>(f), sx);
    load8_cells<!LDS>(S5, in.c, f);
    dx_sxy = dscale(diffsum<0>(f), sx);
    dy_sxy = dscale(diffsum<1>(f), sy);
    load8_cells<!LDS>(S4, in.c, f);
    dx_sxz = dscale(diffsum<0>(f), sx);
    dz_sxz = dscale(diffsum<2>(f), sz);
    load8_cells<!LDS>(S1, in.c, f);
    dy_syy = dscale(diffsum<1>(f), sy);
    load8_cells<!LDS>(S3, in.c, f);
    dy_syz = dscale(diffsum<1>(f), sy);
    dz_syz = dscale(diffsum<2>(f), sz);
    load8_cells<!LDS>(S2, in.c, f);
    dz_szz = dscale(diffsum<2>(f), sz);
  }

  T r0 = (dx_sxx + dy_sxy) + dz_sxz;
  T r1 = (dx_sxy + dy_syy) + dz_syz;
  T r2 = (dx_sxz + dy_syz) + dz_szz;

  if (sp >= 0) {
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
  const T v0 = in.v0 + idt * r0;
  const T v1 = in.v1 + idt * r1;
  const T v2 = in.v2 + idt * r2;
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

// One workgroup's share of the velocity update: chunk b of the instance described by a
// (CPT * blockDim.x nodes, thread t taking nodes t, t + blockDim.x, ...).  With LDS the
// chunk's unique cells are staged first; the streaming operands of the first two nodes
// are requested before the staging so their latency overlaps it.
template <typename T, bool LDS, int U, int CPT, int KM>
__device__ __forceinline__ void velocity_body(const VelocityArgs<T>& a, int b, unsigned char* smem) {
  const int bd = (int)blockDim.x;
  const int base = b * CPT * bd + (int)threadIdx.x;
  VelIn<T> in, next;
  velocity_load<T, LDS, KM>(a, base < a.n ? base : a.n - 1, in);  // tail threads load a valid node
  if constexpr (CPT > 1) velocity_load<T, LDS, KM>(a, base + bd < a.n ? base + bd : a.n - 1, next);
  const T* S0 = a.S0;
  const T* S1 = a.S1;
  const T* S2 = a.S2;
  const T* S3 = a.S3;
  const T* S4 = a.S4;
  const T* S5 = a.S5;
  if constexpr (LDS) {
    T* lds = reinterpret_cast<T*>(smem);
    // rows of max_u + 1 entries: entry max_u is the zero that missing cells read
    const int ld = a.max_u + 1;
    if constexpr (RSG_VELOCITY_PAIRS) {
      // pairs in the order the update reads them: (Sxx, Sxy), (Sxz, Syy), (Syz, Szz)
      const T* const src[6] = {a.S0, a.S5, a.S4, a.S1, a.S3, a.S2};
      stage<T, 6, U, true, true>(lds, ld, a.u_off, a.ulist, b, src, a.max_u, a.uchunk != 0);
      S0 = lds;
      S1 = lds + 2 * ld;
      S2 = lds + 4 * ld;
      S3 = S4 = S5 = nullptr;
    } else {
      const T* const src[6] = {a.S0, a.S1, a.S2, a.S3, a.S4, a.S5};
      stage<T, 6, U, true>(lds, ld, a.u_off, a.ulist, b, src, a.max_u, a.uchunk != 0);
      S0 = lds;
      S1 = lds + ld;
      S2 = lds + 2 * ld;
      S3 = lds + 3 * ld;
      S4 = lds + 4 * ld;
      S5 = lds + 5 * ld;
    }
  }
  if (base >= a.n) return;
  velocity_apply<T, LDS, KM>(a, base, in, S0, S1, S2, S3, S4, S5);
  // later nodes: each node's streaming operands are requested one node ahead
#pragma unroll 1
  for (int j = 1; j < CPT; j++) {
    const int n = base + j * bd;
    if (n >= a.n) break;
    in = next;
    if (j + 1 < CPT) velocity_load<T, LDS, KM>(a, n + bd < a.n ? n + bd : a.n - 1, next);
    velocity_apply<T, LDS, KM>(a, n, in, S0, S1, S2, S3, S4, S5);
  }
}

template <typename T>
__device__ __forceinline__ T cload(const T* __restrict__ C, int64_t ldc, int j, int i) {
  // ldc > 0: [21][ldc] arrays; ldc == 0: blocks of 64 cells, [n/64][21][64], so one
  // wavefront's 21 coefficients are a single contiguous run
  const int64_t idx = ldc ? (int64_t)j * ldc + i : ((int64_t)(i >> 6) * 21 + j) * 64 + (i & 63);
  return STREAM_LOAD(C + idx);
}

// Index of C_(I,J) (0-based Voigt indices) in the 21-entry upper triangle.
__host__ __device__ constexpr int tri(int i, int j) {
  return (i <= j) ? (i * 6 - i * (i - 1) / 2 + (j - i)) : (j * 6 - j * (j - 1) / 2 + (i - j));
}

// Per-cell operands of the stress update that do not depend on a gather: links,
// spacings, the 21 C_IJ and the 6 stresses.
template <typename T>
struct StrIn {
  int c[8];
  uint32_t key;
  T sx, sy, sz;
  T cc[21];
  T s[6];
};

template <typename T, bool LDS, int KM>
__device__ __forceinline__ void stress_load(const StressArgs<T>& a, int e, StrIn<T>& in) {
  if constexpr (LDS) {
    if (packed<KM>(a))
      in.key = load_lkey(a.lkey, e, in.c, -1);  // a cell has all 8 vertices
    else
      load_lconn(a.lconn, e, in.c, -1);
  } else {
    load_conn(a.cnode, e, in.c);
  }
  if (keyed<KM>(a)) {
    if (!(LDS && packed<KM>(a))) in.key = STREAM_LOAD(a.key + e);
    in.sx = in.sy = in.sz = T(0);
#pragma unroll
    for (int j = 0; j < 21; j++) in.cc[j] = T(0);
  } else {
    in.key = 0;
    in.sx = a.dcx[e];
    in.sy = a.dcy[e];
    in.sz = a.dcz[e];
#pragma unroll
    for (int j = 0; j < 21; j++) in.cc[j] = cload(a.C, a.ldc, j, e);
  }
  in.s[0] = a.S0[e];
  in.s[1] = a.S1[e];
  in.s[2] = a.S2[e];
  in.s[3] = a.S3[e];
  in.s[4] = a.S4[e];
  in.s[5] = a.S5[e];
}

// S += dt*(C : grad V + m) for one cell from its strain-rate terms; cc(j) is the j-th of
// the 21 stored C (a class-table row or the cell's own array).
template <typename T, typename R>
__device__ __forceinline__ void voigt_update(T dt, R cc, T d00, T d11, T d22, T e12, T e02, T e01,
                                             const double (&m)[6], int sp, T (&s)[6]) {
#pragma unroll
  for (int ks = 0; ks < 6; ks++) {
    T acc = cc[tri(ks, 0)] * d00 + cc[tri(ks, 1)] * d11 + cc[tri(ks, 2)] * d22 +
            cc[tri(ks, 3)] * e12 + cc[tri(ks, 4)] * e02 + cc[tri(ks, 5)] * e01;
    if constexpr (sizeof(T) == 8) {
      if (sp >= 0) acc = acc + (T)m[ks];
      s[ks] = s[ks] + dt * acc;  // C is plain C_IJ: S += dt*(C:gradV + m)
    } else {
      if (sp >= 0) acc = acc + (T)((double)dt * m[ks]);  // C holds dt*C_IJ
      s[ks] = s[ks] + acc;
    }
  }
}

// Strain-rate terms of one cell from its 8 vertices: the diagonal D_i V_i and the shear
// sums D_j V_k + D_k V_j, in the reference's order of operations.
// With PAIRS, V0 is the staged row of (Vx, Vy) pairs and V1 is unused.
template <typename T, bool PAIRS = false>
__device__ __forceinline__ void strain_terms(const T* V0, const T* V1, const T* V2, const int (&c)[8], T sx,
                                             T sy, T sz, T (&d)[6]) {
  // delV[i][k] = D_i V_k (i: derivative axis, k: velocity component)
  T f[8];
  T d00, d10, d20, d01, d11, d21;
  if constexpr (PAIRS) {
    T g[8];
    load8_pairs(V0, c, f, g);
    d00 = dscale(diffsum<0>(f), sx);
    d10 = dscale(diffsum<1>(f), sy);
    d20 = dscale(diffsum<2>(f), sz);
    d01 = dscale(diffsum<0>(g), sx);
    d11 = dscale(diffsum<1>(g), sy);
    d21 = dscale(diffsum<2>(g), sz);
  } else {
    load8_nodes(V0, c, f);
    d00 = dscale(diffsum<0>(f), sx);
    d10 = dscale(diffsum<1>(f), sy);
    d20 = dscale(diffsum<2>(f), sz);
    load8_nodes(V1, c, f);
    d01 = dscale(diffsum<0>(f), sx);
    d11 = dscale(diffsum<1>(f), sy);
    d21 = dscale(diffsum<2>(f), sz);
  }
  load8_nodes(V2, c, f);
  const T d02 = dscale(diffsum<0>(f), sx);
  const T d12 = dscale(diffsum<1>(f), sy);
  const T d22 = dscale(diffsum<2>(f), sz);
  d[0] = d00;
  d[1] = d11;
  d[2] = d22;
  d[3] = d12 + d21;
  d[4] = d02 + d20;
  d[5] = d01 + d10;
}

// Summed moment terms of a special cell, per Voigt component.
template <typename T>
__device__ __forceinline__ void cell_moments(const StressArgs<T>& a, int sp, double (&m)[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) m[k] = 0.0;
  if (sp < 0) return;
  const int* off = a.sp_off + 7 * (size_t)sp;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double acc = 0.0;
    for (int t = off[k]; t < off[k + 1]; t++) acc += a.amp[a.term[t]];
    m[k] = acc;
  }
}

// S += dt*(C : grad V + m) for one cell; cc[j] is the j-th of the 21 stored C (a class-table
// row or the cell's own array).
template <typename T, typename R>
__device__ __forceinline__ void voigt_update(T dt, R cc, const T (&d)[6], const double (&m)[6], int sp,
                                             T (&s)[6]) {
#pragma unroll
  for (int ks = 0; ks < 6; ks++) {
    T acc = cc[tri(ks, 0)] * d[0] + cc[tri(ks, 1)] * d[1] + cc[tri(ks, 2)] * d[2] +
            cc[tri(ks, 3)] * d[3] + cc[tri(ks, 4)] * d[4] + cc[tri(ks, 5)] * d[5];
    if constexpr (sizeof(T) == 8) {
      if (sp >= 0) acc = acc + (T)m[ks];
      s[ks] = s[ks] + dt * acc;  // C is plain C_IJ: S += dt*(C:gradV + m)
    } else {
      if (sp >= 0) acc = acc + (T)((double)dt * m[ks]);  // C holds dt*C_IJ
      s[ks] = s[ks] + acc;
    }
  }
}

// S += dt*(C : grad V + m) at cell e, V gathered from V0..V2 (LDS or HBM; with PAIRS, V0 is
// the staged row of (Vx, Vy) pairs).
template <typename T, int KM, bool PAIRS = false>
__device__ __forceinline__ void stress_apply(const StressArgs<T>& a, int e, StrIn<T>& in, const T* V0,
                                             const T* V1, const T* V2) {
  T d[6];
  double m[6];
  int sp = -1;
  if (keyed<KM>(a)) {
    // class tables: spacings and material row through the scalar cache where the
    // wavefront shares them
    const uint32_t k = resolve_key(a, in.key, sp);
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
    T sx, sy, sz;
    if (__all(k == k0))
      axis_spacings(as_const(a.axis), k0, sx, sy, sz);
    else
      axis_spacings(a.axis, k, sx, sy, sz);
    const uint32_t mat = k >> kMatShift;
    uint32_t m0 = k0 >> kMatShift;
    if constexpr (sizeof(T) == 8) {
      // FP64: the coefficients stay out of vector registers (a wave more per SIMD).  The
      // first lane's row is requested into scalar registers before the gathers, so a
      // wavefront of one material (the common case) has it when its strain rates are done;
      // otherwise one pass per distinct material, each reading its row through the scalar
      // cache.
      const auto row0 = as_const(a.cls + (size_t)m0 * kMatRow);
      T c0[21];
#pragma unroll
      for (int j = 0; j < 21; j++) c0[j] = row0[j];
      strain_terms<T, PAIRS>(V0, V1, V2, in.c, sx, sy, sz, d);
      cell_moments(a, sp, m);
      if (__all(mat == m0)) {
        voigt_update(a.dt, c0, d, m, sp, in.s);
      } else {
        for (;;) {
          if (mat == m0) {
            voigt_update(a.dt, as_const(a.cls + (size_t)m0 * kMatRow), d, m, sp, in.s);
            break;
          }
          m0 = __builtin_amdgcn_readfirstlane(mat);
        }
      }
    } else {
      // FP32: the row requested before the gathers, with scalar loads when the wavefront
      // has one material
      T cc[21];
      if (__all(mat == m0)) {
        const auto row = as_const(a.cls + (size_t)m0 * kMatRow);
#pragma unroll
        for (int j = 0; j < 21; j++) cc[j] = row[j];
      } else {
        T row[kMatRow];
        load_row(a.cls + (size_t)mat * kMatRow, row);
#pragma unroll
        for (int j = 0; j < 21; j++) cc[j] = row[j];
      }
      strain_terms<T, PAIRS>(V0, V1, V2, in.c, sx, sy, sz, d);
      cell_moments(a, sp, m);
      voigt_update(a.dt, cc, d, m, sp, in.s);
    }
  } else {
    // per-cell arrays (special cells NaN-boxed in dcx)
    T sx = in.sx;
    if (is_boxed(sx)) {
      sp = unbox(sx);
      sx = a.sp_dcx[sp];
    }
    strain_terms<T, PAIRS>(V0, V1, V2, in.c, sx, in.sy, in.sz, d);
    cell_moments(a, sp, m);
    voigt_update(a.dt, in.cc, d, m, sp, in.s);
  }
  a.S0[e] = in.s[0];
  a.S1[e] = in.s[1];
  a.S2[e] = in.s[2];
  a.S3[e] = in.s[3];
  a.S4[e] = in.s[4];
  a.S5[e] = in.s[5];
}

// One workgroup's share of the stress update: chunk b of the instance described by a.
// Every operand that does not depend on a gather is requested first, so a cell's whole
// streaming traffic (links, spacings, 21 C, 6 S) is in flight at once and its latency
// overlaps the LDS staging of the chunk's unique nodes.
template <typename T, bool LDS, int U, int CPT, int KM>
__device__ __forceinline__ void stress_body(const StressArgs<T>& a, int b, unsigned char* smem) {
  const int bd = (int)blockDim.x;
  const int base = b * CPT * bd + (int)threadIdx.x;
  StrIn<T> in, next;
  stress_load<T, LDS, KM>(a, base < a.n ? base : a.n - 1, in);  // tail threads load a valid cell
  if constexpr (CPT > 1) stress_load<T, LDS, KM>(a, base + bd < a.n ? base + bd : a.n - 1, next);
  const T* V0 = a.V0;
  const T* V1 = a.V1;
  const T* V2 = a.V2;
  constexpr bool PAIRS = LDS && (RSG_STRESS_PAIRS == 2 || (RSG_STRESS_PAIRS == 1 && sizeof(T) == 4));
  if constexpr (LDS) {
    T* lds = reinterpret_cast<T*>(smem);
    const T* const src[3] = {a.V0, a.V1, a.V2};
    stage<T, 3, U, false, PAIRS>(lds, a.max_u, a.u_off, a.ulist, b, src, -1);
    V0 = lds;
    V1 = PAIRS ? nullptr : lds + a.max_u;
    V2 = lds + 2 * a.max_u;
  }
  if (base >= a.n) return;
  stress_apply<T, KM, PAIRS>(a, base, in, V0, V1, V2);
  // later cells: each cell's streaming operands are requested one cell ahead
#pragma unroll 1
  for (int j = 1; j < CPT; j++) {
    const int e = base + j * bd;
    if (e >= a.n) break;
    in = next;
    if (j + 1 < CPT) stress_load<T, LDS, KM>(a, e + bd < a.n ? e + bd : a.n - 1, next);
    stress_apply<T, KM, PAIRS>(a, e, in, V0, V1, V2);
  }
}

// Single-instance launches: one workgroup per block, XCD-contiguous block order.
template <typename T, bool LDS, int U, int CPT, int KM>
__global__ __launch_bounds__(RSG_MAX_BLOCK) RSG_VELOCITY_ATTR void velocity_kernel(VelocityArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  velocity_body<T, LDS, U, CPT, KM>(a, a.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x, smem);
}

template <typename T, bool LDS, int U, int CPT, int KM>
__global__ __launch_bounds__(RSG_MAX_BLOCK) RSG_STRESS_ATTR void stress_kernel(StressArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  stress_body<T, LDS, U, CPT, KM>(a, a.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x, smem);
}

// Batched launches walk a table of n (instance, block) items.  The host stores the items
// in launch order (any XCD-aware reordering already applied) and pads the table with
// (-1, 0) items that do nothing, so a workgroup reads its item and then the instance's
// argument block from desc, with no table header in the chain of dependent loads.  The
// default grid has one workgroup per item and no loop (a loop around the body costs the
// FP64 velocity kernel a wave of occupancy); PERSIST grids stride over the table.
// Step k of the current K-step block runs only for instances with k < kmax, so instances
// of different lengths stop exactly at their own step count.  The table and the
// descriptors live in device memory, so a graph of these launches stays valid while
// instances join and leave.
template <typename T, bool LDS, int U, int CPT, int KM>
__device__ __forceinline__ void velocity_item(const BatchDesc<T>* __restrict__ desc, int2 it, int k,
                                              unsigned char* smem) {
  // the item carries the instance's step count for this block, so the branch needs no
  // descriptor load and the descriptor is read in one round trip after it
  if (it.x < 0 || k >= (it.x >> kItemSlotBits)) return;
  const BatchDesc<T>* d = desc + (it.x & (kMaxBatchSlots - 1));
  VelocityArgs<T> a = d->v;
  a.amp = d->amp_base + (size_t)k * d->nch;
  a.trace = d->trace_base ? d->trace_base + (size_t)(d->trace_row0 + k) * d->row : nullptr;
  velocity_body<T, LDS, U, CPT, KM>(a, it.y, smem);
}

template <typename T, bool LDS, int U, int CPT, int KM>
__device__ __forceinline__ void stress_item(const BatchDesc<T>* __restrict__ desc, int2 it, int k,
                                            unsigned char* smem) {
  if (it.x < 0 || k >= (it.x >> kItemSlotBits)) return;
  const BatchDesc<T>* d = desc + (it.x & (kMaxBatchSlots - 1));
  StressArgs<T> a = d->s;
  a.amp = d->amp_base + (size_t)k * d->nch;
  stress_body<T, LDS, U, CPT, KM>(a, it.y, smem);
}

template <typename T, bool LDS, int U, int CPT, int KM, bool PERSIST>
__global__ __launch_bounds__(RSG_MAX_BLOCK) RSG_VELOCITY_ATTR void velocity_batch_kernel(
    const BatchDesc<T>* __restrict__ desc, const int2* __restrict__ items, int n, int k) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (!PERSIST) {
    if ((int)blockIdx.x < n) velocity_item<T, LDS, U, CPT, KM>(desc, items[blockIdx.x], k, smem);
    return;
  }
  for (int p = blockIdx.x; p < n; p += gridDim.x) {
    velocity_item<T, LDS, U, CPT, KM>(desc, items[p], k, smem);
    __syncthreads();  // the next item reuses the LDS
  }
}

template <typename T, bool LDS, int U, int CPT, int KM, bool PERSIST>
__global__ __launch_bounds__(RSG_MAX_BLOCK) RSG_STRESS_ATTR void stress_batch_kernel(
    const BatchDesc<T>* __restrict__ desc, const int2* __restrict__ items, int n, int k) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (!PERSIST) {
    if ((int)blockIdx.x < n) stress_item<T, LDS, U, CPT, KM>(desc, items[blockIdx.x], k, smem);
    return;
  }
  for (int p = blockIdx.x; p < n; p += gridDim.x) {
    stress_item<T, LDS, U, CPT, KM>(desc, items[p], k, smem);
    __syncthreads();
  }
}

template <typename Args, typename K>
hipError_t launch1d(K kernel, const Args& a, int n, int block, size_t lds, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (block < 64 || block > RSG_MAX_BLOCK || (block % 64) != 0 || a.cpt < 1) return hipErrorInvalidValue;
  const int chunk = block * a.cpt;
  const int grid = (n + chunk - 1) / chunk;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, s, a);
  return hipGetLastError();
}

// Picks the kernel instance: direct gathers, or LDS staging with U = ceil(max_u/block)
// entries per thread (general loop beyond 4), for 1, 2 or 4 entities per thread.
template <typename F>
hipError_t with_variant(bool lds, int max_u, int block, int cpt, int km, F&& f) {
  auto by_u = [&](auto cptc, auto kmc) -> hipError_t {
    if (!lds) {
      if constexpr (decltype(kmc)::value == 3)
        return hipErrorInvalidValue;  // packed links exist only with LDS staging
      else
        return f(std::false_type(), std::integral_constant<int, 0>(), cptc, kmc);
    }
    switch ((max_u + block - 1) / block) {
      case 0:
      case 1: return f(std::true_type(), std::integral_constant<int, 1>(), cptc, kmc);
      case 2: return f(std::true_type(), std::integral_constant<int, 2>(), cptc, kmc);
      case 3: return f(std::true_type(), std::integral_constant<int, 3>(), cptc, kmc);
      case 4: return f(std::true_type(), std::integral_constant<int, 4>(), cptc, kmc);
      default: return f(std::true_type(), std::integral_constant<int, 0>(), cptc, kmc);
    }
  };
  auto by_cpt = [&](auto kmc) -> hipError_t {
    switch (cpt) {
      case 1: return by_u(std::integral_constant<int, 1>(), kmc);
      case 2: return by_u(std::integral_constant<int, 2>(), kmc);
      case 4: return by_u(std::integral_constant<int, 4>(), kmc);
      default: return hipErrorInvalidValue;
    }
  };
  switch (km) {
    case 0: return by_cpt(std::integral_constant<int, 0>());
    case 1: return by_cpt(std::integral_constant<int, 1>());
    case 3: return by_cpt(std::integral_constant<int, 3>());
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_velocity(const VelocityArgs<T>& a, int block, hipStream_t s) {
  const bool lds = a.lconn != nullptr || a.lkey != nullptr;
  return with_variant(lds, a.max_u, block, a.cpt, a.lkey ? 3 : a.key ? 1 : 0, [&](auto L, auto U, auto C, auto K) {
    return launch1d(velocity_kernel<T, decltype(L)::value, decltype(U)::value, decltype(C)::value, decltype(K)::value>,
                    a, a.n, block, lds ? velocity_lds_bytes(a.max_u, sizeof(T)) : 0, s);
  });
}

template <typename T>
hipError_t launch_stress(const StressArgs<T>& a, int block, hipStream_t s) {
  const bool lds = a.lconn != nullptr || a.lkey != nullptr;
  return with_variant(lds, a.max_u, block, a.cpt, a.lkey ? 3 : a.key ? 1 : 0, [&](auto L, auto U, auto C, auto K) {
    return launch1d(stress_kernel<T, decltype(L)::value, decltype(U)::value, decltype(C)::value, decltype(K)::value>,
                    a, a.n, block, lds ? (size_t)3 * a.max_u * sizeof(T) : 0, s);
  });
}

}  // namespace

// The exported launchers, and with them the kernel instances, are split over four
// translation units that the build compiles in parallel: RSG_PART = 1 FP64 single-instance,
// 2 FP32 single-instance, 3 FP64 batched, 4 FP32 batched (undefined: all of them).
#ifndef RSG_PART
#define RSG_PART 0
#endif

#if RSG_PART == 0 || RSG_PART == 1
hipError_t launch_velocity_f64(const VelocityArgs<double>& a, int block, hipStream_t s) {
  return launch_velocity<double>(a, block, s);
}
hipError_t launch_stress_f64(const StressArgs<double>& a, int block, hipStream_t s) {
  return launch_stress<double>(a, block, s);
}
#endif

#if RSG_PART == 0 || RSG_PART == 2
hipError_t launch_velocity_f32(const VelocityArgs<float>& a, int block, hipStream_t s) {
  return launch_velocity<float>(a, block, s);
}
hipError_t launch_stress_f32(const StressArgs<float>& a, int block, hipStream_t s) {
  return launch_stress<float>(a, block, s);
}
#endif

#if RSG_PART == 0 || RSG_PART >= 3
template <typename T>
hipError_t launch_batch(const BatchDesc<T>* desc, const int32_t* table, int n, int k, int grid, int block,
                        bool lds, int max_u, int cpt, int km, bool velocity, hipStream_t s, hipGraph_t graph,
                        hipGraphNode_t* node) {
  if (grid <= 0 || n <= 0) return hipSuccess;
  if (reinterpret_cast<uintptr_t>(table) % alignof(int2)) return hipErrorInvalidValue;
  const int2* items = reinterpret_cast<const int2*>(table);
  if (block < 64 || block > RSG_MAX_BLOCK || (block % 64) != 0) return hipErrorInvalidValue;
  const size_t bytes = !lds ? 0 : velocity ? velocity_lds_bytes(max_u, sizeof(T)) : (size_t)3 * max_u * sizeof(T);
  const bool persist = grid != n;
  return with_variant(lds, max_u, block, cpt, km, [&](auto L, auto U, auto C, auto K) {
    constexpr bool l = decltype(L)::value;
    constexpr int u = decltype(U)::value, c = decltype(C)::value, m = decltype(K)::value;
    const void* fn = velocity ? (persist ? reinterpret_cast<const void*>(&velocity_batch_kernel<T, l, u, c, m, true>)
                                         : reinterpret_cast<const void*>(&velocity_batch_kernel<T, l, u, c, m, false>))
                              : (persist ? reinterpret_cast<const void*>(&stress_batch_kernel<T, l, u, c, m, true>)
                                         : reinterpret_cast<const void*>(&stress_batch_kernel<T, l, u, c, m, false>));
    if (graph) {
      void* args[4] = {(void*)&desc, (void*)&items, (void*)&n, (void*)&k};
      hipKernelNodeParams p{};
      p.func = const_cast<void*>(fn);
      p.gridDim = dim3(grid);
      p.blockDim = dim3(block);
      p.sharedMemBytes = (unsigned)bytes;
      p.kernelParams = args;
      p.extra = nullptr;
      const hipGraphNode_t prev = *node;
      return hipGraphAddKernelNode(node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    }
    void* args[4] = {(void*)&desc, (void*)&items, (void*)&n, (void*)&k};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, bytes, s);
  });
}

#endif

#if RSG_PART == 0 || RSG_PART == 3
hipError_t launch_batch_f64(const BatchDesc<double>* d, const int32_t* table, int n, int k, int grid, int block,
                            bool lds, int max_u, int cpt, int km, bool velocity, hipStream_t s, hipGraph_t graph,
                            hipGraphNode_t* node) {
  return launch_batch<double>(d, table, n, k, grid, block, lds, max_u, cpt, km, velocity, s, graph, node);
}
#endif

#if RSG_PART == 0 || RSG_PART == 4
hipError_t launch_batch_f32(const BatchDesc<float>* d, const int32_t* table, int n, int k, int grid, int block,
                            bool lds, int max_u, int cpt, int km, bool velocity, hipStream_t s, hipGraph_t graph,
                            hipGraphNode_t* node) {
  return launch_batch<float>(d, table, n, k, grid, block, lds, max_u, cpt, km, velocity, s, graph, node);
}
#endif

}  // namespace rsg
