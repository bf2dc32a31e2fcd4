// Host-side internals shared by the library's translation units: the context,
// error reporting and the flat-launch helper.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/rbdgrad.h"
#include "fdgrad_knot.hpp"
#include "group_mapping.hpp"

namespace rbdi {

constexpr int kBlock = 64;   // one wave per workgroup (the 4-, 7- and 8-lane groups: 1, 2 or 4)

// With the thresholds deciding (a threshold set by hand drops the mapping
// table), calls of up to this many knots use the 16-lane-group mapping
// instead of one lane per knot: the crossover measured on MI355X (one-lane
// 15.5 us vs lane-group 10.7 us at 8192 knots, 16.0 vs 18.4 us at 16384).
constexpr int64_t kDefaultGroupThreshold = 8192;
// The two-wave mapping (Minv and the RNEA + columns of a knot on two waves)
// serves calls up to the context's two_wave_threshold knots: by default
// 16 x compute units (4,096 knots on 256 CUs, the crossover rbd_autotune
// measured on MI355X: 8.4 us against 9.0 us for one-wave lane groups at 4,096
// knots, 18.4 against 12.6 us at 8,192); rbd_autotune measures it.
constexpr int kTwoWaveLanes = 32;   // lanes_per_knot value of the two-wave mapping (16 lanes in each of 2 waves)
// lanes_per_knot values of the L-lane groups of group_mapping.hpp (full path only)
inline bool is_group_lanes(int lanes) { return lanes == 4 || lanes == 7 || lanes == 8; }
// Waves per workgroup of the L-lane group kernels: 1, 2 or 4 (a launch
// parameter; the waves of a workgroup share nothing).  Larger workgroups put
// one wave on each SIMD of a CU, which the dispatcher does not do for
// single-wave workgroups: a call or batch of about one wave per SIMD (9,014
// knots, 1,002 waves of 7-lane groups) measured 9.1 us in workgroups of 4
// against 9.9 us in workgroups of 1.
constexpr int kMaxWavesPerGroup = rbd::kMaxGroupWaves;
inline bool valid_wpg(int w) { return w == 1 || w == 2 || w == 4; }
// A knot-to-lane mapping and its workgroup shape.
struct Shape {
  int lanes = 0;   // 0: none chosen
  int wpg = 1;     // waves per workgroup (the L-lane groups; 1 for the others)
};
// Mapping tables the autotuners write: a call (or batch) of n knots uses the
// shape measured best at the smallest tabulated size >= n, and the largest
// size's shape above it.
constexpr int kMaxTuned = 16;
struct MappingTable {
  std::atomic<int> count{0};           // 0: no table, the thresholds decide
  int64_t n[kMaxTuned] = {};
  int lanes[kMaxTuned] = {};
  int wpg[kMaxTuned] = {};
  Shape lookup(int64_t knots) const {
    const int c = count.load(std::memory_order_acquire);
    if (!c) return Shape{};
    int i = 0;
    while (i < c - 1 && knots > n[i]) ++i;
    return Shape{lanes[i], wpg[i] ? wpg[i] : 1};
  }
};
// The same crossover for batches (rbd_batch_*, total knots of the batch),
// measured by rbd_autotune_batch on batches of 132-knot trajectories
// (7,957 knots on MI355X: one-lane 17.0 us vs lane-group 18.4 us at 8,448).
constexpr int64_t kDefaultBatchGroupThreshold = 8192;
// Batches of up to this many trajectories find a chunk's trajectory inside the
// wave, larger ones through the plan kernel (rbd::kSmallBatch, batch_kernels.hpp).
constexpr int kSmallBatchTrajectories = 64;
// Host-pointer calls of up to this many knots stage through pinned, mapped
// host memory that the kernel reads and writes directly (8.3 MB, allocated at
// the first such call); larger ones copy through a device buffer.
constexpr int64_t kHostZeroCopyKnots = 16384;

// Records msg as the calling thread's last error and returns s.
rbd_status fail(rbd_status s, const std::string& msg);
rbd_status hip_fail(hipError_t e, const char* what);

enum class Layout { AoS, SoA, Ref };

struct FlatCall {
  Layout layout;
  int lanes;              // 1 or rbd::kLanesPerKnotGroup
  int64_t n;
  const float* in;
  int64_t ld_in;          // SoA only
  float* out;
  int64_t ld_out;         // SoA only
  rbd::Options opt;
  hipStream_t stream;
  int simds = 0;          // SIMDs of the device (4 per CU); 0: unknown
  int wpg = 1;            // waves per workgroup of the L-lane groups (1, 2, 4)
};

hipError_t launch_flat(int precision, const FlatCall& c);

}  // namespace rbdi

struct rbd_context {
  int device = 0;
  unsigned flags = 0;
  int precision = RBD_PREC_FP32;
  int lanes_per_knot = 0;                        // 0 = by size (the mapping tables), else fixed
  int waves_per_group = 0;                       // 0 = by the tables, else fixed (L-lane groups)
  std::atomic<int64_t> group_threshold{rbdi::kDefaultGroupThreshold};
  std::atomic<int64_t> two_wave_threshold{0};    // set by rbd_create / rbd_autotune
  std::atomic<int64_t> batch_group_threshold{rbdi::kDefaultBatchGroupThreshold};
  int compute_units = 0;                         // of the device
  rbdi::MappingTable tuned;                      // per call, set by rbd_autotune
  rbdi::MappingTable batch_tuned;                // per batch (total knots), set by rbd_autotune_batch
  rbd::Options opt;
  // buffers of the host-pointer convenience call: pinned, mapped staging that
  // the kernel reads and writes directly (calls of up to kHostZeroCopyKnots
  // knots), else a device buffer between two copies; its own stream
  std::mutex host_mu;
  float* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  char* h_stage = nullptr;
  char* dh_stage = nullptr;
  size_t stage_bytes = 0;
  hipStream_t host_stream = nullptr;
  // the word a two-wave kernel's handoff wait sets when it gives up: pinned,
  // coherent host memory (opt.fault is its device address)
  volatile unsigned* fault_host = nullptr;
};

namespace rbdi {

// The shapes the autotuners compare -- lanes per knot, and waves per
// workgroup for the L-lane groups -- and the number of flat sizes.
constexpr int kAutotuneMappings = 12;
constexpr Shape kAutotuneShapes[kAutotuneMappings] = {
    {1, 1}, {rbd::kLanesPerKnotGroup, 1}, {kTwoWaveLanes, 1}, {4, 1}, {7, 1}, {8, 1},
    {4, 2}, {7, 2}, {8, 2}, {4, 4}, {7, 4}, {8, 4}};
constexpr int kAutotuneSizes = 12;

// The mapping tables a context starts with, from MI355X measurements of the
// autotuners (scaled by the device's SIMDs, 4 per CU): per call, the two-wave
// mapping up to 4 knots per SIMD (4,096 knots on 1,024 SIMDs), groups of 7
// lanes in workgroups of 4 waves up to one wave of 9 knots per SIMD (9,216),
// groups of 4 lanes up to one wave of 16 (16,384), then one lane per knot;
// per batch the same without the two-wave mapping, whose looping batch kernel
// trails the groups of 7 lanes there.
// (Measured: a call of 9,014 knots 9.1 us with 7 lanes per knot in
// workgroups of 4 waves, 9.9 in workgroups of 1, 12.9 with two waves, 15.0
// with one lane; 16,384 knots 11.5 us with 4, 15.3 with one lane; the
// 64-trajectory mixed batch of 9,014 knots 9.6-10.0 us with 7 in workgroups of
// one wave against 16.4 with one lane or 16.)
inline void set_default_tables(rbd_context* ctx) {
  const int64_t simds = 4 * int64_t(ctx->compute_units);
  const int64_t n[4] = {4 * simds, 9 * simds, 16 * simds, INT64_MAX};
  const Shape flat[4] = {{kTwoWaveLanes, 1}, {7, 4}, {4, 1}, {1, 1}};
  for (int i = 0; i < 4; ++i) {
    ctx->tuned.n[i] = n[i];
    ctx->tuned.lanes[i] = flat[i].lanes;
    ctx->tuned.wpg[i] = flat[i].wpg;
  }
  ctx->tuned.count.store(4, std::memory_order_release);
  for (int i = 0; i < 3; ++i) {
    ctx->batch_tuned.n[i] = n[i + 1];
    ctx->batch_tuned.lanes[i] = flat[i + 1].lanes;
    ctx->batch_tuned.wpg[i] = flat[i + 1].wpg;
  }
  ctx->batch_tuned.count.store(3, std::memory_order_release);
}

// Waves per workgroup for `lanes` at n knots: the context's fixed value, else
// the table's when the table chose these lanes there, else 1 (and always 1 for
// mappings other than the L-lane groups).
inline int wpg_for(const rbd_context* ctx, const MappingTable& t, int lanes, int64_t n) {
  if (!is_group_lanes(lanes)) return 1;
  if (ctx->waves_per_group) return ctx->waves_per_group;
  const Shape s = t.lookup(n);
  return s.lanes == lanes ? s.wpg : 1;
}

// Shape of a call of n knots: lanes 1, 4, 7, 8, 16 or kTwoWaveLanes -- the
// context's fixed lanes_per_knot, else the autotuned table, else the
// thresholds.  Only the full path (AoS, SoA) has two-wave and L-lane group
// kernels; other entry points pass full_path = false and get the 16-lane
// group instead.
inline Shape shape_for(const rbd_context* ctx, int64_t n, bool full_path = false) {
  int l = ctx->lanes_per_knot;
  if (!l) l = ctx->tuned.lookup(n).lanes;
  if (!l) {
    if (full_path && n <= ctx->two_wave_threshold.load(std::memory_order_relaxed)) return Shape{kTwoWaveLanes, 1};
    return Shape{n <= ctx->group_threshold.load(std::memory_order_relaxed) ? rbd::kLanesPerKnotGroup : 1, 1};
  }
  if (!full_path && l != 1) return Shape{rbd::kLanesPerKnotGroup, 1};
  return Shape{l, wpg_for(ctx, ctx->tuned, l, n)};
}
inline int lanes_for(const rbd_context* ctx, int64_t n, bool full_path = false) {
  return shape_for(ctx, n, full_path).lanes;
}

inline Shape shape_for_batch(const rbd_context* ctx, int64_t total_knots) {
  int l = ctx->lanes_per_knot;
  if (!l) l = ctx->batch_tuned.lookup(total_knots).lanes;
  if (!l) l = total_knots <= ctx->batch_group_threshold.load(std::memory_order_relaxed) ? rbd::kLanesPerKnotGroup : 1;
  return Shape{l, wpg_for(ctx, ctx->batch_tuned, l, total_knots)};
}

// One batch with a given shape (batch.hip; lanes 0: the context's choice,
// shape_for_batch); rbd_fdgrad_batched_f32 is this with lanes 0.
rbd_status batch_run(rbd_batch* b, const rbd_traj* trajs, int n_traj, int use_graph, rbd_stream_t stream,
                     Shape shape);

// For callers that place a batch's header and descriptors themselves (the
// server): batch_stage writes them (batch_stage_bytes(n_traj) bytes,
// trajectory pointers as given) to host memory and returns the chunk total;
// batch_launch_staged runs the batch from d_staged, a device-visible address
// of that block (8-byte aligned, read only: device memory the caller copied it
// to on `stream`, or the pinned host memory itself), with the queue's tail
// counters in the batch object's device memory; with trajs (the same list,
// device-visible pointers) a batch of up to 64 trajectories takes its
// descriptors from the kernel arguments instead.  Launches on one batch object
// must run one at a time (the server gives each in-flight slot its own).  With
// done_flag (a device address of host-mapped coherent memory), the kernel's
// last wave writes done_seq there once all outputs are visible to the host,
// so the caller can poll memory instead of an event.  A batch of zero knots
// launches nothing and writes nothing.
// Once per device, with that device current (rbd_create does it): the device
// address of the no-allocation batched call's scratch (batch.hip).
rbd_status noalloc_setup(int device);

size_t batch_stage_bytes(int n_traj);
int64_t batch_stage(void* host_dst, const rbd_traj* trajs, int n_traj, int lanes);
rbd_status batch_launch_staged(rbd_batch* b, const void* d_staged, const rbd_traj* trajs, int n_traj,
                               int64_t total_knots, int64_t total_chunks, Shape shape, rbd_stream_t stream,
                               unsigned long long* done_flag = nullptr, unsigned long long done_seq = 0);

inline rbd_status check_ctx(const rbd_context* ctx) {
  if (!ctx) return fail(RBD_ERR_INVALID_ARGUMENT, "null context");
  return RBD_OK;
}

// the context's device must be current for a launch on the caller's thread
inline rbd_status use_device(const rbd_context* ctx) {
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (cur != ctx->device) {
    e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  }
  return RBD_OK;
}

// RBD_ERR_DEVICE_TIMEOUT once a kernel of the context gave up a handoff wait
inline rbd_status device_status(const rbd_context* ctx) {
  if (ctx->fault_host && *ctx->fault_host != 0)
    return fail(RBD_ERR_DEVICE_TIMEOUT,
                "a two-wave kernel of this context gave up its LDS handoff wait: its results are invalid "
                "(rbd_context_clear_status to go on)");
  return RBD_OK;
}

// Before a launch: the context's device current, and no device-side failure pending.
inline rbd_status ready_to_launch(const rbd_context* ctx) {
  if (rbd_status s = device_status(ctx)) return s;
  return use_device(ctx);
}

}  // namespace rbdi
