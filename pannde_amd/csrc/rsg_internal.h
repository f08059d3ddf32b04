// rsg_internal.h -- device-side layout shared by the host code (rsg_host.cpp) and the
// kernels (rsg_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>

namespace rsg {

// Kernel argument blocks.  All arrays are in the library's internal (locality) order.
//
// Coefficients come in one of two forms per kernel:
//  * class tables (key != nullptr): each entity holds a 32-bit key made of a material
//    class (bits 21..30: the cell's row of 21 C, or the node's dt/rho) and one spacing
//    class per axis (7 bits each: x at bit 14, y at bit 7, z at bit 0).  The material rows
//    and the three spacing tables are small and read through the caches; a wavefront's
//    cells usually share one material row, so those loads coalesce.  A key with bit 31 set
//    flags a special entity (source or sensor): its low bits index the special tables,
//    which hold the entity's full key;
//  * per-entity arrays (key == nullptr), for meshes with more than 1024 materials or 128
//    spacings along an axis: source and sensor entities are flagged without any extra
//    array, a node's dt/rho (idt) or a cell's x-spacing (dcx) being replaced by a quiet NaN
//    whose payload is the special index, and the special tables hold the real value.
// Either way ordinary entities cost no extra HBM bytes for the flag.
constexpr uint32_t kSpecialKey = 0x80000000u;
constexpr int kMatShift = 21;
constexpr int kMaxMaterials = 1024;
constexpr int kAxisClasses = 128;  // spacing classes per axis; tables are [3][kAxisClasses]
constexpr int kMatRow = 24;        // cell material row: the 21 C as the C array holds them, padded
constexpr uint32_t kNoLink = 0x3FF;  // packed 10-bit link: no neighbour
constexpr int kMaxPackedSet = 1022;  // largest staged set whose local indices fit 10 bits

// LDS of a velocity workgroup: 6 stress rows of max_u + 1 entries, the last one a zero that
// links to missing cells read.
inline size_t velocity_lds_bytes(int max_u, size_t elt) { return (size_t)6 * (max_u + 1) * elt; }

// First word of a batch work item: the instance slot (18 bits) and the number of steps of the
// current K-step block it takes (kmax <= 4096, bits 18..30, so the word stays non-negative);
// -1 marks an empty item.
constexpr int kItemSlotBits = 18;
constexpr int32_t kMaxBatchSlots = 1 << kItemSlotBits;
inline int32_t batch_item_word(int32_t slot, int32_t kmax) { return slot | (kmax << kItemSlotBits); }

inline uint32_t make_key(int mat, int ix, int iy, int iz) {
  return ((uint32_t)mat << kMatShift) | ((uint32_t)ix << 14) | ((uint32_t)iy << 7) | (uint32_t)iz;
}
template <typename T>
struct VelocityArgs {
  T* V0;
  T* V1;
  T* V2;
  const T* S0;  // Sxx
  const T* S1;  // Syy
  const T* S2;  // Szz
  const T* S3;  // Syz
  const T* S4;  // Sxz
  const T* S5;  // Sxy
  const T* idt;  // dt/rho_n, NaN-boxed for special nodes
  const T* dnx;  // FP64: node spacing dxn; FP32: 0.25/dxn
  const T* dny;
  const T* dnz;
  const int32_t* ncell;  // [n][8] connected cells by slot, -1 = none
  int32_t n;
  int32_t xcd;  // 1: remap workgroups so each XCD owns a contiguous range
  int32_t cpt;  // entities per thread: a workgroup covers cpt * blockDim.x nodes
  // LDS staging (lconn != nullptr): per-block unique cell lists and 16-bit local links
  const int32_t* u_off;    // [nblocks+1]
  const int32_t* ulist;    // cells staged by each block
  const uint16_t* lconn;   // [n][8] block-local cell index by slot, 0xFFFF = none
  int32_t max_u;           // largest unique set of a block (LDS rows hold max_u + 1)
  int32_t uchunk;          // 1: ulist holds chunk words, start << 2 | (len - 1): runs of up to 4
                           // consecutive cells, each taking 4 LDS slots (slot 4c + k holds
                           // start + k for k < len); max_u and the links count slots
  // special nodes (sources and/or sensors)
  const T* sp_idt;         // [nsp] real dt/rho_n
  const int32_t* sp_off;   // [nsp][4] term ranges for directions 0,1,2
  const int32_t* sp_sens;  // [nsp] sensor slot or -1
  const int32_t* term;     // channel id per term
  const double* amp;       // [n_channels] amplitudes of this step
  double* trace;           // [n_sensors][3] row of this step
  // class tables (key or lkey != nullptr: idt, dnx..dnz and sp_idt are not used)
  const uint32_t* key;     // [n] material/spacing key, or kSpecialKey | special index
  const T* cls;            // [materials] dt/rho
  const T* axis;           // [3][kAxisClasses] node spacings as dnx..dnz hold them
  const uint32_t* sp_key;  // [nsp] key of each special node
  // LDS links and key in one 16-byte record (lkey != nullptr: lconn and key are not used):
  // words 0..2 hold the 8 block-local indices, 10 bits each (3 per word, kNoLink = none),
  // word 3 the key
  const uint32_t* lkey;    // [n][4]
};

template <typename T>
struct StressArgs {
  T* S0;
  T* S1;
  T* S2;
  T* S3;
  T* S4;
  T* S5;
  const T* V0;
  const T* V1;
  const T* V2;
  const T* C;    // [21][ldc]: FP64 C_IJ; FP32 dt*C_IJ (FP64 product, rounded)
  int64_t ldc;   // leading dimension of C (>= n), or 0 for the 64-cell blocked layout
  const T* dcx;  // FP64: cell spacing dxc (NaN-boxed for special cells); FP32: 0.25/dxc
  const T* dcy;
  const T* dcz;
  const int32_t* cnode;  // [n][8] vertices
  int32_t n;
  int32_t xcd;
  int32_t cpt;             // cells per thread
  const int32_t* u_off;    // LDS staging: per-block unique node lists
  const int32_t* ulist;
  const uint16_t* lconn;   // [n][8] block-local vertex index
  int32_t max_u;
  T dt;  // used by the FP64 form only
  // special cells (stress sources)
  const T* sp_dcx;        // [nsp] real x-spacing value
  const int32_t* sp_off;  // [nsp][7] term ranges for Voigt components 0..5
  const int32_t* term;
  const double* amp;
  // class tables (key or lkey != nullptr: C, dcx..dcz and sp_dcx are not used)
  const uint32_t* key;     // [n] material/spacing key, or kSpecialKey | special index
  const T* cls;            // [materials][kMatRow]
  const T* axis;           // [3][kAxisClasses] cell spacings as dcx..dcz hold them
  const uint32_t* sp_key;  // [nsp] key of each special cell
  const uint32_t* lkey;    // [n][4] LDS links + key, as for VelocityArgs
};

// Energy arguments (rsg_energy.hip): FP64 accumulation, one partial per workgroup.
constexpr int kEnergyMaxBlocks = 1024;
template <typename T>
struct EnergyArgs {
  const T* V0;
  const T* V1;
  const T* V2;
  const double* node_m;    // node mass rho_n * dxn*dyn*dzn (FP64)
  int32_t nn;
  StressArgs<T> s;         // the stress kernel's arguments: S, n and the coefficient form
  const double* cell_vol;  // dxc*dyc*dzc (FP64)
  double inv_scale;        // 1 (FP64: C stored as C_IJ) or 1/dt (FP32: stored as dt*C_IJ)
  double* partial;         // [2][nblocks]
};

// One instance of a batch: its argument blocks plus where this K-step block's
// amplitudes and traces go.
template <typename T>
struct BatchDesc {
  VelocityArgs<T> v;
  StressArgs<T> s;
  const double* amp_base;  // [K][nch] amplitudes of the current K-step block
  double* trace_base;      // the instance's trace buffer, [rows][row] (nullptr: no sensors)
  int32_t nch;
  int32_t row;             // 3 * sensors
  int64_t trace_row0;      // first trace row of this K-step block
  int32_t kmax;            // steps of this K-step block the instance takes
  int32_t pad;
};

// Launchers (rsg_kernels.hip).  block = threads per block (64..1024, multiple of 64).
hipError_t launch_velocity_f64(const VelocityArgs<double>& a, int block, hipStream_t s);
hipError_t launch_velocity_f32(const VelocityArgs<float>& a, int block, hipStream_t s);
hipError_t launch_stress_f64(const StressArgs<double>& a, int block, hipStream_t s);
hipError_t launch_stress_f32(const StressArgs<float>& a, int block, hipStream_t s);
// Batched launch: grid workgroups walk table = n (instance, block) int32 pairs in launch
// order, 8-byte aligned, (-1, 0) for an empty item; lds selects LDS staging (max_u: the
// largest staged set of any instance), cpt the entities per thread (1, 2 or 4), km the
// coefficient form of every instance the table names (0 per-entity arrays, 1 class-table
// keys, 3 keys packed with the LDS links).  With graph != nullptr the launch is added to
// the graph as a kernel node that depends on *node (when not null) and *node becomes the
// new node; otherwise it is enqueued on s.
hipError_t launch_batch_f64(const BatchDesc<double>* d, const int32_t* table, int n, int k, int grid, int block,
                            bool lds, int max_u, int cpt, int km, bool velocity, hipStream_t s, hipGraph_t graph,
                            hipGraphNode_t* node);
hipError_t launch_batch_f32(const BatchDesc<float>* d, const int32_t* table, int n, int k, int grid, int block,
                            bool lds, int max_u, int cpt, int km, bool velocity, hipStream_t s, hipGraph_t graph,
                            hipGraphNode_t* node);
// Energy: partials into a.partial, then their fixed-order sum into out[0] (kinetic) and
// out[1] (strain); both device addresses.  nblocks <= kEnergyMaxBlocks.
hipError_t launch_energy_f64(const EnergyArgs<double>& a, int nblocks, double* out, hipStream_t s);
// Device address of the static work area of the allocation-free energy on `device` (2 *
// kEnergyMaxBlocks + 2 doubles), resolved with hipGetSymbolAddress on first use per device.
hipError_t energy_scratch(int device, double** out);
// Largest node or cell count the energy kernels' grid-stride indexing takes.
constexpr int32_t kEnergyMaxEntities = 0x7FFFFFFF - kEnergyMaxBlocks * 256;
hipError_t launch_energy_f32(const EnergyArgs<float>& a, int nblocks, double* out, hipStream_t s);

// NaN boxing of a special-table index.
inline double box_f64(int32_t idx) {
  uint64_t b = 0x7FF8000000000000ull | (uint64_t)(uint32_t)idx;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}
inline float box_f32(int32_t idx) {
  uint32_t b = 0x7FC00000u | (uint32_t)idx;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
constexpr int32_t kMaxSpecialF32 = (1 << 22) - 1;

}  // namespace rsg
