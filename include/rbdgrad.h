/*
 * rbdgrad.h -- C ABI of the MI355X (gfx950) forward-dynamics-gradient library
 * for the 7-DoF KUKA iiwa 14 (librbdgrad.so).
 *
 * Per knot point:  q, qd, u  ->  qdd, dqdd/dq, dqdd/dqd   (FP32 in and out)
 *
 * The reference (plancherb1/fast-rbd-gradients) has no library or FFI: its
 * per-trajectory entry points are C++ templates and a CUDA kernel over raw
 * buffers.  Each function below names the reference interface it replaces.
 *
 * Layouts (FP32):
 *   knot input   [q(7) | qd(7) | u(7)]                                21 floats
 *   knot output  [qdd(7) | dqdd/dq (7x7) | dqdd/dqd (7x7)]            105 floats
 *                7x7 blocks column-major: entry (r, c) = d qdd_r / d x_c at 7c + r
 *                (the reference's dqdd layout, helpers_CPU/dynamicsGradient.h:839,921)
 *   reference-compatible input  [q | qd | u | qdd | Minv dense 7x7]   77 floats
 *   reference-compatible output [dqdd/dq | dqdd/dqd]                  98 floats
 *                (COMPLETELY_FUSED_VALS, helpers_GPU/dynamicsGradient.cuh:1142-1146)
 *
 * Conventions:
 *   - Status codes, never exit(); rbd_last_error() gives the detail of the last
 *     failure on the calling thread.
 *   - Device pointers (d_*) are caller-owned HIP device memory; host pointers
 *     (h_*) are caller-owned host memory.  The library allocates no device memory
 *     except in the host-pointer convenience call and the batch/serving objects,
 *     which own what they allocate.
 *   - Every launch goes on the caller's stream (a hipStream_t; NULL = default
 *     stream) and returns without synchronising, like a kernel launch.
 *   - A context is immutable after rbd_create and may be used from several
 *     threads at once.
 *   - Inputs must be finite.  A knot with a non-finite input gets unspecified
 *     output values; no other knot is affected.
 */
#ifndef RBDGRAD_H
#define RBDGRAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbd_context rbd_context;
typedef void* rbd_stream_t; /* hipStream_t */

typedef enum rbd_status {
  RBD_OK = 0,
  RBD_ERR_INVALID_ARGUMENT = 1,
  RBD_ERR_HIP = 2,
  RBD_ERR_NO_DEVICE = 3,
  RBD_ERR_OUT_OF_MEMORY = 4,
  RBD_ERR_UNSUPPORTED = 5,
  RBD_ERR_CAPACITY = 6,
  RBD_ERR_DEVICE_TIMEOUT = 7 /* a kernel of this context gave up a bounded wait (see rbd_context_status) */
} rbd_status;

/* model options: the reference's compile-time VEL_DAMPING and MPC_MODE
 * (utils/experiment_helpers.h:19-22; both off by default there) */
#define RBD_FLAG_VEL_DAMPING 1u /* joint damping 0.5*qd (helpers_CPU/dynamicsGradient.h:680,840) */
#define RBD_FLAG_NO_GRAVITY 2u  /* gravity off, MPC_MODE (helpers_CPU/dynamicsGradient.h:635-637) */

/* arithmetic precision (inputs and outputs are FP32 in every mode) */
typedef enum rbd_precision {
  RBD_PREC_FP32 = 0,         /* everything FP32 */
  RBD_PREC_FP64_PIVOT = 1,   /* FP32, but the Minv pivots D_k = S^T IA_k S and 1/D_k in FP64 */
  RBD_PREC_FP64_COLUMNS = 2, /* FP32 Minv and RNEA, FP64 derivative columns and -Minv*dtau */
  RBD_PREC_FP64 = 3,         /* everything FP64 */
  RBD_PREC_FP32_HW_TRIG = 4  /* everything FP32, joint sin/cos from the hardware's v_sin/v_cos: 1-2% less
                                time, largest relative error about 6.5e-5 instead of 4.1e-5 */
} rbd_precision;

typedef struct rbd_config {
  int device;         /* HIP device ordinal */
  unsigned flags;     /* RBD_FLAG_* */
  int precision;      /* rbd_precision */
  int lanes_per_knot; /* 0: chosen per call from the size (the context's mapping
                         table, see rbd_autotune); 1: one lane per knot
                         (throughput); 4, 7, 8: one knot per group of that many
                         lanes, 16, 9 and 8 knots per wave (full path and batches;
                         other calls use 16); 16: one knot per 16-lane group;
                         32: one knot per 16-lane group in each of two waves, Minv
                         on one and the RNEA and columns on the other (lowest
                         latency; full path and batches, other calls use 16) */
  int waves_per_group; /* workgroup size of the 4-, 7- and 8-lane group kernels in
                          waves (their waves share nothing): 0 chosen per size by
                          the mapping tables, or 1, 2, 4 */
  int reserved[3];    /* must be zero */
} rbd_config;

/* rbd_config with device 0, flags 0 (gravity on, no damping), RBD_PREC_FP32,
 * lanes_per_knot 0. */
void rbd_config_default(rbd_config* cfg);

/* Context: device, model options, precision.  The robot constants are compiled
 * into the kernels' instruction stream, so nothing is uploaded per call
 * (the reference uploads its 504-float mem_const once, time_GPU.cu:174). */
rbd_status rbd_create(const rbd_config* cfg, rbd_context** ctx);
rbd_status rbd_destroy(rbd_context* ctx);

const char* rbd_status_string(rbd_status s);
const char* rbd_last_error(void);

/* Device-side failures.  The two-wave mapping (lanes_per_knot 32, flat calls,
 * batches and servers) hands data between the two waves of a workgroup
 * through LDS; each handoff wait is bounded.  A wait that gives up marks a
 * host-visible word of the context, and the kernel still exits normally, but
 * its results are invalid.  Once the word is set, every launching call on the
 * context returns RBD_ERR_DEVICE_TIMEOUT, a server fails the batches that
 * complete after it, and rbd_context_status reports it (call it after the
 * stream has been synchronised to learn about the launches on it).
 * rbd_context_clear_status clears the word.  No reference counterpart: the
 * reference's kernels synchronise with __syncthreads and report nothing
 * (gpuErrchk aborts on API errors only, helpers_GPU/dynamicsGradient.cuh:28-35). */
rbd_status rbd_context_status(rbd_context* ctx);
rbd_status rbd_context_clear_status(rbd_context* ctx);
/* The bound of a handoff wait, in polls with a 64-cycle sleep between them
 * (default 2^24, about half a second); spins < 0 makes every wait give up at
 * once (a test hook: it forces the failure path from the host, with every
 * kernel still running to its end).  Applies to launches made after the call. */
rbd_status rbd_set_handoff_spins(rbd_context* ctx, int spins);
/* Library ABI version (incremented on incompatible changes). */
int rbd_abi_version(void);
/* Name of the robot this library's kernels are compiled for: "iiwa14" for
 * librbdgrad.so, NAME for librbdgrad_NAME.so built from a URDF
 * (make -C fast-rbd-gradients_amd MODEL=NAME URDF=robot.urdf).
 * Replaces: the reference's hard-coded iiwa tables (dynamicsGradient.h:123-229,
 * dynamicsGradient.cuh:211-310); its README.md:1 points to GRiD for other robots. */
const char* rbd_model_name(void);

/* Full path for n knot points, array-of-structures (21 in / 105 out per knot),
 * one launch on `stream`.
 * Replaces: the per-trajectory forward-dynamics gradient -- the reference
 * computes it only on the CPU, forwardDynamicsGradientSetup with
 * QDD_MINV_PASSED_IN = false + forwardDynamicsGradientThreaded
 * (helpers_CPU/dynamicsGradient.h:868-897, :924-934), driven per trajectory by
 * dynamicsGradientReusableThreaded (time_CPU.cpp:47-54). */
rbd_status rbd_fdgrad_traj_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, float* d_out,
                               rbd_stream_t stream);

/* Same, structure-of-arrays: value i of knot k at d_in[i*ld_in + k] (i < 21) and
 * output value j at d_out[j*ld_out + k] (j < 105); ld_* >= n_knots.
 * Replaces: as rbd_fdgrad_traj_f32 (a layout the reference does not have). */
rbd_status rbd_fdgrad_traj_soa_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, int64_t ld_in,
                                   float* d_out, int64_t ld_out, rbd_stream_t stream);

/* Reference-compatible mode: qdd and a dense Minv are inputs (77 per knot),
 * dqdd/dq and dqdd/dqd out (98 per knot).  d_mem_const may be NULL; if given it
 * must hold the reference's 504-float T|I constant block and is only checked
 * for size by the caller's contract (the constants are compiled in).
 * Replaces: dynamicsGradientKernel<T,N,MPC,VD><<<N,(32,7)>>>(d_dqdd, d_mem_vol,
 * d_mem_const) (helpers_GPU/dynamicsGradient.cuh:1257-1291, launched at
 * time_GPU.cu:314,325). */
rbd_status rbd_dynamics_gradient_kernel_f32(rbd_context* ctx, int64_t n_knots, float* d_dqdd,
                                            const float* d_mem_vol, const float* d_mem_const,
                                            rbd_stream_t stream);

/* Host-pointer convenience: runs the full path on h_in, writes h_out and
 * waits (the reference's "GPU Compute + I/O" timing, time_GPU.cu:311-320).
 * Calls of up to 16,384 knots copy h_in into pinned, mapped host memory that
 * the kernel reads and writes directly (kept by the context, 8.3 MB); larger
 * ones copy through a per-context device buffer.  Calls on one context are
 * serialised; they run on a stream of the context's own.
 * Replaces: the H2D copy + kernel + D2H copy sequence of time_GPU.cu:313-316. */
rbd_status rbd_fdgrad_traj_host_f32(rbd_context* ctx, int64_t n_knots, const float* h_in, float* h_out);

/* Full path plus dqdd/du = Minv: as rbd_fdgrad_traj_f32, and d_minv receives
 * the dense Minv of every knot (49 floats per knot, column-major).  The three
 * 7x7 blocks dqdd/dq, dqdd/dqd, dqdd/du are the WAFR baseline's gradient output
 * (time_GPU_WAFR.cu:18) and what iLQR/DDP callers consume.
 * Replaces: forwardDynamicsGradient + knot.Minv (dynamicsGradient.h:920-922:
 * "dqdd/dtau = Minv"). */
rbd_status rbd_fdgrad_minv_traj_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, float* d_out,
                                   float* d_minv, rbd_stream_t stream);

/* Full path in the WAFR baseline's gradient layout: d_dqdd receives 147 floats
 * per knot, [dqdd/dq | dqdd/dqd | dqdd/du = Minv], each 7x7 column-major, and
 * d_qdd the 7 of qdd per knot unless it is null.  Input as rbd_fdgrad_traj_f32
 * (q, qd, u; 21 floats per knot).
 * Replaces: the WAFR baseline's gradient output d_dqdd (time_GPU_WAFR.cu:18,
 * :134; finish_dqdd_v2, helpers_WAFR/dynamics_WAFR.cuh:2555-2579), here
 * computed by the reference method from q, qd, u instead of from a given qdd
 * and Minv. */
rbd_status rbd_fdgrad_wafr_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, float* d_qdd, float* d_dqdd,
                               rbd_stream_t stream);

/* Forward dynamics only: qdd (7 per knot) and the dense Minv (49 per knot,
 * column-major) from q, qd, u (21 per knot), one lane per knot.  The bias RNEA
 * runs with qdd = 0.
 * Replaces: forwardDynamics<T,MPC,VD,DENSE_MINV=1>(knot) (dynamicsGradient.h:685-706,
 * :936-939), which the reference runs on the CPU before every GPU timing
 * (time_GPU.cu:134). */
rbd_status rbd_forward_dynamics_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, float* d_qdd,
                                    float* d_minv, rbd_stream_t stream);

/* External forces.  d_fext holds, for each of the 7 links, the spatial force
 * the environment applies to it, in the link's own frame: 6 floats
 * [moment(3) | force(3)] at d_fext[6k .. 6k+5], 42 floats per set.  Knot j
 * uses the set at d_fext + j * fext_stride: fext_stride = 0 applies one set to
 * every knot (the reference's per-call s_fext), fext_stride >= 42 one set per
 * knot.  The forces enter the bias, qdd and the gradient (they are constant in
 * the link frames, so they move with the links).  Device memory, valid until
 * the call completes on `stream`.
 * Replaces: forwardDynamicsGradient(dqdd, knot, T, I, s_fext) and
 * forwardDynamics(knot, s_fext) (dynamicsGradient.h:916-939), whose RNEA_fp
 * subtracts s_fext from each link's force (:654-657). */
rbd_status rbd_fdgrad_traj_fext_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, const float* d_fext,
                                    int64_t fext_stride, float* d_out, rbd_stream_t stream);
rbd_status rbd_forward_dynamics_fext_f32(rbd_context* ctx, int64_t n_knots, const float* d_in, const float* d_fext,
                                         int64_t fext_stride, float* d_qdd, float* d_minv, rbd_stream_t stream);

/* The reference's staged GPU variants (time_GPU.cu:227-333), for callers
 * that keep their pipeline.  cdu is the inverse-dynamics gradient
 * [dtau/dq | dtau/dqd], 98 floats per knot, each 7x7 column-major
 * (cdu[7c + k] = dtau_k/dx_c; with VEL_DAMPING +0.5 on the dtau/dqd diagonal).
 * d_mem_const is accepted and ignored (the model constants are compiled in).
 *
 * Fused: q, qd, u, qdd (28 floats per knot; u unused) -> cdu.
 * Replaces: dynamicsGradientKernel_fused<T,N,MPC,VD><<<N,(32,7)>>>(d_cdu,
 * d_mem_vol_fused, d_mem_const) (helpers_GPU/dynamicsGradient.cuh:1223-1255). */
rbd_status rbd_dynamics_gradient_kernel_fused_f32(rbd_context* ctx, int64_t n_knots, float* d_cdu,
                                                  const float* d_mem_vol_fused, const float* d_mem_const,
                                                  rbd_stream_t stream);

/* Split: q, qd, u, qdd, v, a, f (154 floats per knot: the 28 above, then
 * link k's v, a, f 6-vectors at 28 + 6k, 70 + 6k, 112 + 6k; f after RNEA's
 * backward pass) -> cdu.  The context's gravity flag does not apply (it is
 * already in a).
 * Replaces: dynamicsGradientKernel_split<T,N,VD><<<N,(32,7)>>>(d_cdu,
 * d_mem_vol_split, d_mem_const) (.cuh:1191-1221). */
rbd_status rbd_dynamics_gradient_kernel_split_f32(rbd_context* ctx, int64_t n_knots, float* d_cdu,
                                                  const float* d_mem_vol_split, const float* d_mem_const,
                                                  rbd_stream_t stream);

/* Fill v, a, f of a split-layout buffer (154 floats per knot) from its q, qd,
 * qdd, in place, on the device.
 * Replaces: dynamicsGradientReusableThreaded_start (.cuh:1159-1166, the CPU
 * stage before the split kernel; forwardDynamicsGradientSetup_forGPU :1148-1157). */
rbd_status rbd_gradient_setup_split_f32(rbd_context* ctx, int64_t n_knots, float* d_mem_vol_split,
                                        rbd_stream_t stream);

/* dqdd = -Minv cdu per knot: Minv 49 floats per knot (column-major, the upper
 * triangle is read), cdu 98 -> dqdd 98 ([dqdd/dq | dqdd/dqd]), on the device.
 * Replaces: dynamicsGradientReusableThreaded_finish (.cuh:1181-1188, the CPU
 * stage after the split and fused kernels; forwardDynamicsGradientFinish_forGPU
 * :1168-1179). */
rbd_status rbd_gradient_finish_f32(rbd_context* ctx, int64_t n_knots, float* d_dqdd, const float* d_cdu,
                                   const float* d_minv, rbd_stream_t stream);

/* Name of the kernel the context dispatches a call of n_knots to (full path,
 * or the reference-compatible mode when reference_mode != 0), as profilers list
 * it (template arguments abbreviated).  No reference counterpart. */
const char* rbd_kernel_name(rbd_context* ctx, int reference_mode, int64_t n_knots);

/* Autotuner for the knot-to-lane mapping (lanes_per_knot = 0): times twelve
 * launch shapes -- 1, 16, 32 (two waves), 4, 7 and 8 lanes per knot, and the
 * 4-, 7- and 8-lane groups also in workgroups of 2 and 4 waves.  Lanes per
 * knot fix the knots per wave (64, 4, 4 per two waves, 16, 9, 8), the LDS
 * each wave holds (none, one Minv per knot group for 7, 8 and 16 lanes, the
 * two-wave handoff blocks) and the waves per SIMD each kernel is built for;
 * the workgroup size sets the knots per workgroup.  It times them on the
 * device at 12 call sizes from 64 to 65,536 knots (with the sizes that fill
 * the SIMDs with one wave of 9 or 16 knots), and installs the context's
 * mapping table: a call of n knots uses the shape fastest at the smallest
 * measured size >= n (the largest size's above it).  Also sets the largest
 * size at which the 16-lane mapping beats one lane per knot as the crossover
 * (threshold_out) and the largest at which two waves beat both as the two-wave
 * limit.  A context starts with a table of measured MI355X defaults.
 * Allocates and frees a temporary device buffer; synchronises `stream`.  Not
 * to be called while other threads launch on the same context.  table_us, if
 * not NULL, receives up to table_len floats as rows of 13: (n, us with 1, 16,
 * 32, 4, 7, 8 lanes per knot, then 4, 7, 8 lanes in workgroups of 2 waves,
 * then of 4 waves).  No reference counterpart (the reference fixes dim3(32,7)
 * per knot, time_GPU.cu:10-19). */
rbd_status rbd_autotune(rbd_context* ctx, rbd_stream_t stream, int64_t* threshold_out, float* table_us,
                        int table_len);
/* Current crossover (knots) and its manual override; the same for the
 * two-wave limit (calls of up to that many knots use 32 lanes per knot).
 * Setting either by hand drops the mapping table, so the two thresholds then
 * decide between the 1, 16 and 32-lane mappings. */
int64_t rbd_group_threshold(rbd_context* ctx);
rbd_status rbd_set_group_threshold(rbd_context* ctx, int64_t n_knots);
int64_t rbd_two_wave_threshold(rbd_context* ctx);
rbd_status rbd_set_two_wave_threshold(rbd_context* ctx, int64_t n_knots);

/* The batched path (rbd_fdgrad_batched_f32) has its own mapping table on a
 * batch's total knots.  rbd_autotune_batch times the twelve shapes on batches
 * of 8, 16, 32, 48, 64, 80, 96, 128, 256 and 512 trajectories of 132 knots and
 * 64 of 142 (9,088 knots) and installs the fastest at each total; it also sets the crossover (the
 * largest total at which 16-lane groups beat one lane per knot), which decides
 * alone once set by hand.  table_us receives rows of 13 as rbd_autotune's, on
 * total knots.  Same conditions as rbd_autotune.  No reference counterpart. */
rbd_status rbd_autotune_batch(rbd_context* ctx, rbd_stream_t stream, int64_t* threshold_out, float* table_us,
                              int table_len);
int64_t rbd_batch_group_threshold(rbd_context* ctx);
rbd_status rbd_set_batch_group_threshold(rbd_context* ctx, int64_t n_knots);
/* The context's mapping table, per call (batched = 0) or per batch: up to
 * `capacity` entries (largest size in knots, lanes per knot, waves per
 * workgroup), ascending; any output pointer may be NULL; returns the table's
 * length (0: the thresholds decide), -1 for a null context.  No reference
 * counterpart. */
int rbd_mapping_table(rbd_context* ctx, int batched, int64_t* n_knots, int* lanes, int* waves_per_group,
                      int capacity);

/* ------------------------------------------------------------------------
 * Batched trajectories of uneven length.
 * A batch object owns a device work queue (trajectory descriptors, chunk
 * prefix sums, a cursor) and pinned staging for the descriptors.  One call runs
 * a whole batch with one persistent-kernel launch: waves pull knot chunks from
 * the queue until it is empty.  Trajectory buffers stay caller-owned device
 * memory (AoS, 21 in / 105 out per knot).
 * Replaces: the reference's per-trajectory launch over one traj<T,N> at a time
 * (grid = N knots, time_GPU.cu:314; CPU split of a trajectory into per-thread
 * chunks, time_CPU.cpp:30-36), generalised to many trajectories per launch.
 * ------------------------------------------------------------------------ */
typedef struct rbd_batch rbd_batch;

typedef struct rbd_traj {
  const float* d_in;  /* n_knots x 21 */
  float* d_out;       /* n_knots x 105 */
  int64_t n_knots;
} rbd_traj;

rbd_status rbd_batch_create(rbd_context* ctx, int max_trajectories, rbd_batch** out);
rbd_status rbd_batch_destroy(rbd_batch* b);

/* Run the forward-dynamics gradient over n_traj trajectories on `stream`.
 * Up to 64 trajectories travel in the kernel's arguments (no upload); more
 * are copied to the device asynchronously on `stream`.  The caller's trajs
 * array may be reused on return.  The knot-to-lane mapping is chosen from the
 * total knot count by the context's batch mapping table (rbd_autotune_batch),
 * unless the context fixes lanes_per_knot (32: the two-wave mapping, 4-knot
 * chunks on workgroups of two waves; 4, 7, 8: chunks of 16, 9, 8 knots of the
 * batch's knots taken in order across trajectories).
 * use_graph != 0 replays the batch's kernels from a hipGraph.  Up to 64
 * trajectories: a graph of the one kernel with this trajectory list in its
 * arguments, captured when the list (buffers and lengths) differs from the
 * previous use_graph call's of the same mapping and replayed unchanged when it
 * is the same (new inputs written into the same buffers need no new capture).
 * More: a graph captured once per mapping, reading the descriptors the call
 * copies to the device first.  On ROCm 7.2 a graph launch costs about 15 us
 * of host time against 2.6 us for a direct launch, so direct launches are the
 * faster form where the host issues every batch.  Calls on one batch object
 * are ordered by their streams: two streams must not run the same batch
 * object at once.  Returns RBD_ERR_CAPACITY if n_traj exceeds the batch's
 * max_trajectories. */
rbd_status rbd_fdgrad_batched_f32(rbd_batch* b, const rbd_traj* trajs, int n_traj, int use_graph,
                                  rbd_stream_t stream);

/* The batched call without a batch object, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * device allocation): up to 64 trajectories (RBD_ERR_CAPACITY beyond), their
 * descriptors in the kernel's arguments, the mapping from the context's batch
 * table as for rbd_fdgrad_batched_f32, the results bitwise the same.  It calls
 * no allocating routine: the queue's last-round counters (8,320 bytes) live in
 * a file-scope device array of the library, one per device, set up by the
 * first rbd_create on that device and held until the process ends.  Every
 * call on a device shares that array: calls in one stream's order are safe,
 * calls that overlap on one device (several streams or threads at once) race.
 * No counterpart in the reference; for callers that cannot allocate. */
rbd_status rbd_fdgrad_batched_noalloc_f32(rbd_context* ctx, const rbd_traj* trajs, int n_traj, rbd_stream_t stream);

/* Persistent grid size (waves) per mapping, lanes_per_knot 1, 4, 7, 8 or 16
 * (default: the waves the device's SIMDs keep resident: per CU 6 for one lane
 * and 8 for 16 lanes with FP32 columns, 4 for either with FP64 columns, 4 for
 * 4, 7 and 8 lanes in every workgroup size; the two-wave mapping runs 2
 * workgroups of two waves per CU).  The grid of the 4-, 7- and 8-lane groups
 * is set for all three of their workgroup sizes at once. */
rbd_status rbd_batch_set_grid(rbd_batch* b, int lanes_per_knot, unsigned waves);

/* Lanes per knot used by the last rbd_fdgrad_batched_f32 call (1, 4, 7, 8, 16 or 32). */
int rbd_batch_last_lanes(const rbd_batch* b);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of gradient requests.
 * A request is one trajectory (host input n x 21, host output n x 105; both
 * caller-owned and valid until the request completes).  A dispatcher thread
 * admits requests in arrival order while the memory budget allows (a batch
 * holds one range of the budget: n x 504 bytes per request plus 32 bytes per
 * descriptor and 256 for its queue header, and the same of pinned staging),
 * packs them into batches and runs each batch on a free slot (stream + batch
 * work queue) with one upload, one launch and one download, so batches overlap
 * on the device; each slot's completion thread delivers its batches' results
 * and records latency (submit -> result in h_out).
 * All batches of one server use the same knot-to-lane mapping, so a result does
 * not depend on what it was batched with.  No reference counterpart (the
 * reference has no serving layer); the latency statistics follow its
 * printStats percentiles (utils/experiment_helpers.h:41-100).
 * ------------------------------------------------------------------------ */
typedef struct rbd_server rbd_server;

typedef struct rbd_server_config {
  int slots;                      /* batches in flight (one stream each), 1..64 */
  int lanes_per_knot;             /* 1, 4, 7, 8, 16 or 32 (two waves per knot group), used for every batch */
  int64_t memory_budget_bytes;    /* device memory for admitted requests (and as much pinned staging) */
  int64_t max_batch_trajectories; /* per batch */
  int64_t max_batch_knots;        /* per batch; also the longest admissible request */
  double batch_window_us;         /* after the oldest queued request arrived, wait this long for more (0: none) */
  int use_graph;                  /* ignored: each batch's queue is uploaded with its inputs and launched directly */
  int zero_copy;                  /* 0: requests' buffers in device memory, one upload and one download per batch;
                                     1: the kernels read inputs from and write results to the pinned staging
                                     directly (no copies: lower latency, PCIe traffic from the kernels, which caps
                                     the server at about 9e7 knots/s);
                                     2 (the default): per batch -- zero-copy up to zero_copy_max_knots knots,
                                     copies above */
  int poll_dispatch;              /* 1: the dispatcher thread polls for requests instead of sleeping (one busy
                                     core; saves the ~10 us wake-up of a sleeping thread per batch) */
  int zero_copy_max_knots;        /* zero_copy = 2: the largest zero-copy batch (0: 16384 knots) */
} rbd_server_config;

typedef struct rbd_server_stats {
  int64_t requests_completed;
  int64_t knots_completed;
  int64_t batches;
  int64_t max_inflight_batches;
  int64_t max_batch_knots;
  int64_t max_budget_bytes_used;
  int64_t queued;                 /* submitted, not yet admitted */
  double p50_latency_us;
  double p99_latency_us;
  double mean_latency_us;
  double max_latency_us;
  double knots_per_s;             /* completed knots / (last completion - first submit) */
} rbd_server_stats;

/* Per-batch trace record (the server keeps the last 65536 batches).  Times
 * are microseconds since the server was created:
 *   oldest_submit_us  submit time of the batch's first (oldest) request
 *   admit_us          the dispatcher admitted the batch (memory reserved)
 *   enqueued_us       upload, launch, download and event were enqueued
 *   done_us           the completion thread saw the batch's event
 *   delivered_us      results were in the callers' buffers */
typedef struct rbd_batch_record {
  double oldest_submit_us;
  double admit_us;
  double enqueued_us;
  double done_us;
  double delivered_us;
  int64_t requests;
  int64_t knots;
  int slot;
  int reserved;
} rbd_batch_record;

/* Copy up to max_records of the most recent batch records, oldest first;
 * *n_out receives the number copied.  No reference counterpart. */
rbd_status rbd_server_get_trace(rbd_server* s, rbd_batch_record* out, int64_t max_records, int64_t* n_out);

/* slots 4, lanes_per_knot 16, 64 MiB budget, 1024 trajectories and 2^15 knots
 * per batch, no batching window. */
void rbd_server_config_default(rbd_server_config* cfg);
rbd_status rbd_server_create(rbd_context* ctx, const rbd_server_config* cfg, rbd_server** out);
/* Finishes the batches on the device, fails the queued requests, joins the threads. */
rbd_status rbd_server_destroy(rbd_server* s);
/* Queue a request; RBD_ERR_CAPACITY if it can never fit the budget or max_batch_knots. */
rbd_status rbd_server_submit(rbd_server* s, const float* h_in, float* h_out, int64_t n_knots, uint64_t* id);
/* 1: done (result in h_out), 0: pending, -1: failed, -2: unknown id.  A finished
 * request is forgotten once polled or waited for. */
int rbd_server_poll(rbd_server* s, uint64_t id);
/* Block until the request finishes (timeout_s < 0: no limit). */
rbd_status rbd_server_wait(rbd_server* s, uint64_t id, double timeout_s);
rbd_status rbd_server_get_stats(rbd_server* s, rbd_server_stats* out);
rbd_status rbd_server_reset_stats(rbd_server* s);
/* Load generator for the served benchmark: Poisson arrivals at rate_per_s for
 * duration_s, each request one trajectory with length drawn uniformly from
 * lengths[0..n_lengths), inputs taken from pool_in (pool_knots x 21); waits for
 * all requests and returns the statistics of the run. */
rbd_status rbd_server_run_poisson(rbd_server* s, double rate_per_s, double duration_s, const int64_t* lengths,
                                  int n_lengths, const float* pool_in, int64_t pool_knots, uint64_t seed,
                                  rbd_server_stats* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* RBDGRAD_H */
