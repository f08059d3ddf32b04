/*
 * pannde_rsg.h -- C ABI of the MI355X RSG DeviceSolver library (libpannde_rsg.so).
 *
 * The library is the device half of a drop-in replacement for PanNDE's HostSolver
 * (reference: modules/HostSolver/include/HostRSGTDUT.hpp, HostUTModel.hpp).  Plain
 * pointers and sizes only; host arrays are borrowed for the duration of a call, device
 * buffers are owned by the handle.  One handle owns one HIP stream; calls on one handle
 * must not overlap.  Every entry point returns RSG_OK (0) or a negative status, and
 * rsg_last_error() returns the message of the calling thread's last failure (the
 * reference signals the same conditions with C++ exceptions: std::logic_error("DX") from
 * RightHex::nodePrep, std::out_of_range from HexMesh::setCell, std::runtime_error).
 *
 * Numbering: every array passed in or out uses the caller's node and cell numbering;
 * the library reorders internally for locality and maps back on the way out.
 * Stress is in the reference's Voigt order Sxx,Syy,Szz,Syz,Sxz,Sxy
 * (modules/PanNDE/include/StandardNames.hpp:167).
 */
#ifndef PANNDE_RSG_H
#define PANNDE_RSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSG_ABI_VERSION 1

/* status codes */
#define RSG_OK 0
#define RSG_ERR_ARG (-1)    /* bad argument (null pointer, size, index out of range) */
#define RSG_ERR_MESH (-2)   /* RightHex::nodePrep non-positive node spacing ("DX"/"DY"/"DZ") */
#define RSG_ERR_DEVICE (-3) /* HIP runtime failure */
#define RSG_ERR_STATE (-4)  /* call not valid in the handle's current state */
#define RSG_ERR_NOMEM (-5)  /* device memory exhausted / over budget */

/* field placement, PanNDE::Field::FieldType (modules/PanNDE/include/Field.hpp) */
#define RSG_FIELD_NODE 0
#define RSG_FIELD_CELL 1
#define RSG_FIELD_CONSTANT 2

/* arithmetic of a handle */
#define RSG_FP64 0 /* FP64 fields and updates in the reference's operation order */
#define RSG_FP32 1 /* FP32 fields and updates; coefficients precomputed in FP64 then
                      rounded; sensor traces and energy accumulated in FP64 */

/* source channel kinds */
#define RSG_FORCE 0  /* body force in the velocity update, HostRSGTDUT::evalXds */
#define RSG_MOMENT 1 /* stress source in the stress update, HostUTModel::evalMoments */

typedef struct rsg_instance* rsg_handle;

/* Hexahedral mesh, the data of PanNDE::Mesh (Mesh.hpp:74-248) / HostData::HexMesh:
 * node coordinates and per-cell vertex lists in CGNS order
 * [000,100,110,010,001,101,111,011].  Node-to-cell links are derived inside the
 * library with HexMesh::buildLinks' slot rule (HexMesh.hpp:300-313). */
typedef struct {
  int32_t n_nodes;
  int32_t n_cells;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* cell_nodes; /* [n_cells][8] */
} rsg_mesh_desc;

/* A material field as PanNDE::Field exposes it: data() and type().  Length is n_nodes
 * (NODE), n_cells (CELL) or 1 (CONSTANT).  Read with HostField::atNode / atCell
 * semantics (HostField.hpp:121-154). */
typedef struct {
  const double* data;
  int32_t type;
} rsg_field_desc;

/* Model constants: the domain FieldBundle of HostRSGTDUT/HostUTModel (density and
 * C11..C66 under the ElasticMaterialNames, StandardNames.hpp:75-147) plus dt. */
typedef struct {
  double dt;
  rsg_field_desc density;
  rsg_field_desc cij[21]; /* C11,C12,C13,C14,C15,C16,C22,..,C26,C33,..,C36,C44,C45,C46,
                             C55,C56,C66 */
  int32_t moment_form;    /* 0: HostRSGTDUT update; 1: HostUTModel update (stress
                             sources allowed) */
} rsg_model_desc;

/* One source channel: one non-null MultiVariate of one transducer row (force) or one
 * moment-tensor row (moment).  Channels of one kind and component are summed in the
 * order given, which must be the reference's row order.
 *   FORCE : support = cells whose centre the source covers (its checkAt at
 *           HostRSGTDUT::cellcenter); a node's term list walks its 8 connected-cell
 *           slots, then the channels, exactly as evalXds does.
 *   MOMENT: support = nodes the source covers (checkAt at the node coordinate); a
 *           cell's term list walks its 8 vertices, then the channels (evalMoments).
 * Each term adds the channel's per-step amplitude, i.e. 0.125*evalAt(point, t).
 * If n_signal > 0 the library evaluates it itself as 0.125*(s(t)*1.0) with s the
 * HostData::HostUnivariate table (signal_t, signal_v) (HostUnivariate.hpp:109-142);
 * otherwise amplitudes come with every rsg_step call. */
typedef struct {
  int32_t kind;
  int32_t component; /* FORCE: direction 0..2; MOMENT: Voigt index 0..5 */
  int32_t n_support;
  const int32_t* support;
  int32_t n_signal;
  const double* signal_t;
  const double* signal_v;
} rsg_channel_desc;

/* Replaces HostRSGTDUT::HostRSGTDUT / HostUTModel::HostUTModel set-up
 * (HostRSGTDUT.hpp:139-146,180-251; HostUTModel.hpp:113-124): spacings (RightHex
 * nodePrep/cellPrep, RightHex.hpp:253-309), density at nodes and C at cells, the
 * locality reordering, and the device upload.  The state starts at zero and time at 0
 * (the drivers' zero initial condition, TransducerModel.cpp:150-159). */
int rsg_create(const rsg_mesh_desc* mesh, const rsg_model_desc* model, int32_t precision,
               rsg_handle* out);

/* Kernel and layout choices of a handle (they never change results: every choice is
 * bit-identical).  A zero field takes the default shown. */
typedef struct {
  int32_t block_velocity;      /* workgroup size of the velocity kernel, 64..512 step 64 (256;
                                  512 for meshes of 3 x 2^20 nodes or more when
                                  per_thread_velocity is left at 0 and the 512-node staged
                                  sets fit) */
  int32_t block_stress;        /* workgroup size of the stress kernel (256) */
  int32_t per_thread_velocity; /* nodes per thread: 1, 2 or 4 (1) */
  int32_t per_thread_stress;   /* cells per thread: 1, 2 or 4 (2 for meshes of 3 x 2^20 cells
                                  or more, else 1: measured on a 3.4 M-cell plate and on
                                  32^3..128^3 blocks) */
  int32_t lds_velocity;        /* 1: stage each workgroup's gathers in LDS, -1: gather from HBM (1) */
  int32_t lds_stress;          /* the same for the stress kernel (1) */
  int32_t tile[3];             /* locality brick in cells along x, y, z (8, 8, 4) */
  int32_t c_layout;            /* 1: C in 64-cell blocks [n/64][21][64], -1: [21][n] (1) */
  int32_t xcd_remap;           /* 1: XCD-contiguous workgroup order, -1: hardware order (1) */
  int32_t coeff_classes;       /* 1: each entity carries a 4-byte key into small tables of
                                  materials (a cell's 21 C, a node's dt/rho) and of spacings
                                  per axis, instead of its own coefficients; used when there
                                  are at most 1024 materials and 128 spacings along each axis,
                                  -1: per-entity coefficient arrays (1) */
  int32_t node_classes;        /* rsg_get_tuning only: node materials in use, -1 = per-node arrays */
  int32_t cell_classes;        /* rsg_get_tuning only: cell materials in use, -1 = per-cell arrays */
  int32_t links_velocity;      /* rsg_get_tuning only: 0 32-bit global links, 1 16-bit block-local
                                  (LDS), 2 10-bit block-local packed with the key in 16 bytes */
  int32_t links_stress;        /* rsg_get_tuning only: the same for the stress kernel */
} rsg_tuning;

/* rsg_create with explicit kernel and layout choices (NULL tuning = defaults). */
int rsg_create_ex(const rsg_mesh_desc* mesh, const rsg_model_desc* model, int32_t precision,
                  const rsg_tuning* tuning, rsg_handle* out);

/* The choices a handle was created with, defaults filled in. */
int rsg_get_tuning(rsg_handle h, rsg_tuning* out);

/* Replaces the transducer list of HostRSGTDUT (setupTransducers, HostRSGTDUT.hpp:213-219)
 * and HostUTModel::applyTransducer / applyMomentExcitation (HostUTModel.hpp:138-170).
 * Replaces any previous channel set. */
int rsg_set_channels(rsg_handle h, int32_t n_channels, const rsg_channel_desc* channels);

/* Sensor nodes whose (Vx,Vy,Vz) are recorded after every step, fused into the velocity
 * kernel (the velocity table of TransducerModel.cpp:106-149).  Replaces any previous set
 * and clears recorded traces. */
int rsg_set_sensors(rsg_handle h, int32_t n_sensors, const int32_t* nodes);

/* Replaces Model::solve(nsteps) (HostRSGTDUT.hpp:162-169): nsteps of
 * {velocity update; stress update; time += dt}.  amplitudes: NULL when every channel
 * carries a signal table, else [nsteps][n_channels] values 0.125*evalAt(point, t_k),
 * t_k being the time at the start of step k.  *time_out (if not NULL) receives the
 * time after the last step, as solve() returns it. */
int rsg_step(rsg_handle h, int32_t nsteps, const double* amplitudes, double* time_out);

/* Replaces Model::getStates() (HostRSGTDUT.hpp:153): copies the live state out.
 * V: [3][n_nodes] (Vx,Vy,Vz); S: [6][n_cells] in Voigt order.  Either may be NULL. */
int rsg_get_state(rsg_handle h, double* V, double* S);

/* Writes a state (e.g. an initial condition other than zero).  Either may be NULL. */
int rsg_set_state(rsg_handle h, const double* V, const double* S);

/* Recorded sensor velocities for steps [first_step, first_step+nsteps) counted from the
 * last rsg_set_sensors: out[step][sensor][3]. */
int rsg_get_traces(rsg_handle h, int64_t first_step, int64_t nsteps, double* out);

/* Blocks until every step queued on the handle's stream has finished (rsg_step only
 * enqueues work).  The reference's solve() is synchronous; this is its completion
 * point. */
int rsg_synchronize(rsg_handle h);

/* Current simulation time (the value solve() last returned) and steps taken. */
int rsg_get_time(rsg_handle h, double* time, int64_t* steps);

/* Elastic energy of the current state, accumulated in FP64 (no counterpart in the
 * reference, which computes none; the FP64 accumulation the mixed-precision form keeps):
 *   kinetic = sum_n 0.5 * rho_n * dxn*dyn*dzn * |V_n|^2
 *   strain  = sum_c 0.5 * dxc*dyc*dzc * S_c . C_c^-1 . S_c
 * rho_n: the node density of the velocity update (HostField::atNode, HostField.hpp:142-154);
 * dxn..dzn, dxc..dzc: RightHex::nodePrep / cellPrep spacings (RightHex.hpp:253-309), so
 * rho_n*dxn*dyn*dzn is the node's share of the body's mass; C_c: the 6x6 stiffness of
 * HostRSGTDUT::updateTsAtCell (HostRSGTDUT.hpp:278-293) acting on the strain rates
 * (d00, d11, d22, d12+d21, d02+d20, d01+d10), inverted per cell in FP64 (the FP32 form
 * uses its stored dt*C divided by dt).  A cell whose C is not positive definite makes
 * strain NaN.  Either pointer may be NULL. */
int rsg_energy(rsg_handle h, double* kinetic, double* strain);

/* rsg_energy without any allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same kernels and the same result, bit for bit, with the
 * partial sums and the two totals in a static device array of 2*1024+2 doubles (16 KiB per
 * GPU, held until the process ends) instead of the handle's buffer.  Its size does not grow
 * with the mesh (a fixed 1024 partials), so the only limit is the kernels' indexing: node and
 * cell counts up to 2^31 - 2^18 - 1, larger ones are rejected with RSG_ERR_ARG.  Every call on
 * one GPU shares that GPU's array (each call overwrites all of it, nothing needs clearing):
 * calls that overlap on one GPU race.  Fails with RSG_ERR_STATE when the handle's source
 * tables would have to be rebuilt (which allocates): call rsg_step or rsg_energy once first. */
int rsg_energy_noalloc(rsg_handle h, double* kinetic, double* strain);

/* Device-kernel timing: with enabled = n > 0 the velocity and stress launches of every n-th
 * step (the first of each n) are bracketed by HIP events on the handle's stream (1: every
 * step; events between launches slow the stream, so a stride keeps the timed run close to
 * an untimed one); 0 turns timing off.  Totals and the number of timed steps accumulate
 * until read (and reset). */
int rsg_set_timing(rsg_handle h, int32_t enabled);
int rsg_get_kernel_times(rsg_handle h, double* velocity_ms, double* stress_ms,
                         int64_t* launches);

/* Bytes of device memory the handle holds (state, coefficients, the links of each kernel:
 * connectivity, or the LDS plan of a staged kernel, and the source and trace buffers). */
int rsg_device_bytes(rsg_handle h, int64_t* bytes);

/* Releases the handle and returns its device memory to the library's pool (one per device,
 * kept for later handles until the process ends) in order behind the handle's own stream,
 * without waiting for other streams. */
int rsg_destroy(rsg_handle h);

/* Copies the calling thread's last error message into buf (NUL-terminated); returns its
 * length. */
int rsg_last_error(char* buf, size_t len);

/* ---------------------------------------------------------------- batches
 * Many independent instances advanced together: one launch per kernel per step covers
 * every instance of a group (a device work table of (instance, block) items and a device
 * table of per-instance arguments), and the K consecutive steps of a block are captured
 * once in a HIP graph.  Instances are grouped so that a group's state arrays (V and S)
 * fit a byte budget (by default 128 MiB in FP32 and 96 MiB in FP64).  Each instance joins
 * one of two lanes, the one with the fewest state bytes still to step; lanes are
 * independent schedulers on streams of their own that run at once, so the two groups in
 * flight fit the MI355X's 256 MiB last-level cache and each lane's launches fill the
 * tails of the other's.  In a lane, each group takes all K steps of a block before the
 * next group starts, so the state one kernel writes is still cached when the other reads
 * it.
 * Instances join and leave between K-step blocks; each
 * takes exactly its own number of steps.  No counterpart in the reference, which runs
 * one model per process; each instance's arithmetic is that of rsg_step, so its results
 * are bit-identical alone and in any mix.  Instances are ordinary handles of the batch's
 * precision, their channels must carry signal tables, and while in a batch they are
 * stepped only by the batch; read their
 * state and traces with rsg_get_state / rsg_get_traces after rsg_batch_synchronize. */
typedef struct rsg_batch_s* rsg_batch;

/* steps_per_block: K (0: 50).  grid: 0 or -1 launches one workgroup per work item (the
 * grid sized to the group's table capacity); n > 0 launches n persistent workgroups that
 * walk the table.  use_graph: 0 launches every step directly instead of replaying a
 * captured graph. */
int rsg_batch_create(int32_t precision, int32_t steps_per_block, int32_t grid, int32_t use_graph,
                     rsg_batch* out);

/* Adds a handle that will take nsteps more steps; *slot receives its slot id (a batch holds
 * at most 2^18 instances at once).  It joins at the next K-step block; blocks already
 * enqueued keep running.  Members may differ in
 * their tuning: each launch covers the members of one kernel configuration. */
int rsg_batch_add(rsg_batch b, rsg_handle h, int64_t nsteps, int32_t* slot);

/* Detaches a slot's handle (finished or not); the handle can be stepped alone again.
 * A finished instance waits only for the block in which it finished, so later blocks of
 * the batch can run meanwhile; an unfinished one waits for its lane. */
int rsg_batch_remove(rsg_batch b, int32_t slot);

/* Runs up to nblocks K-step blocks (enqueued; see rsg_batch_synchronize).  *unfinished
 * receives the number of instances with steps left. */
int rsg_batch_run(rsg_batch b, int32_t nblocks, int32_t* unfinished);

/* Scheduling options (results do not depend on them):
 *   RSG_BATCH_GROUP_BYTES  state bytes per group (>= 0; 0: all instances of a lane in one
 *                          group; default 128 MiB in FP32, 96 MiB in FP64)
 *   RSG_BATCH_XCD_RUN      work-item order across the 8 XCDs: -1 table order, 0 one
 *                          contiguous eighth per XCD, n > 0 runs of n items per XCD in turn
 *                          (default 64)
 *   RSG_BATCH_LANES        instances added from now on join the lane with the fewest
 *                          state bytes still to step among this many (1..4, default 2);
 *                          members keep their lane */
#define RSG_BATCH_GROUP_BYTES 1
#define RSG_BATCH_XCD_RUN 2
#define RSG_BATCH_LANES 3
int rsg_batch_set_option(rsg_batch b, int32_t option, int64_t value);

int rsg_batch_synchronize(rsg_batch b);

/* Steps taken and requested of a slot's instance. */
int rsg_batch_slot_status(rsg_batch b, int32_t slot, int64_t* done, int64_t* total);

int rsg_batch_destroy(rsg_batch b);

/* HostNormalXDZ::checkAt (modules/Stubs/include/HostNormalXDZ.hpp:140-145) over a mesh:
 * at = 0 tests every cell centre (HostRSGTDUT::cellcenter), at = 1 every node.  Writes
 * the ids that pass, ascending, to ids (capacity n_cells or n_nodes) and their count to
 * *count. */
int rsg_disc_support(const rsg_mesh_desc* mesh, int32_t at, const double center[3],
                     double radius, double ds, int32_t* ids, int32_t* count);

/* HostUnivariate::evalAt (HostUnivariate.hpp:109-121) of a (t, v) table at n_eval
 * times (keys sorted, a repeated key keeps its first value).  A time below the first
 * key is undefined in the reference and returns RSG_ERR_ARG here. */
int rsg_signal_eval(int32_t n, const double* t, const double* v, int32_t n_eval,
                    const double* at, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PANNDE_RSG_H */
