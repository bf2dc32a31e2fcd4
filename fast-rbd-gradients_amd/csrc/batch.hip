// Batched uneven trajectories (rbd_batch_*): one persistent-kernel launch per
// batch over a device work queue of knot chunks, optionally replayed from a
// hipGraph.  Kernels in batch_kernels.hpp.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "batch_kernels.hpp"
#include "context.hpp"

using namespace rbdi;

static_assert(rbd::kSmallBatch == kSmallBatchTrajectories, "one small-batch limit");

struct rbd_batch {
  rbd_context* ctx = nullptr;
  int max_traj = 0;
  // device: [BatchHeader | TrajDesc[max_traj] | long prefix[max_traj + 1]]
  char* d_mem = nullptr;
  rbd::BatchHeader* d_hdr = nullptr;
  rbd::TrajDesc* d_desc = nullptr;
  long* d_prefix = nullptr;
  // counters of the last partial round (device memory, zero between launches)
  unsigned long long* d_tail = nullptr;
  size_t upload_bytes = 0;            // header + descriptors
  // pinned host staging ring for the descriptor upload
  static constexpr int kRing = 4;
  char* h_stage[kRing] = {};
  hipEvent_t staged[kRing] = {};
  int ring = 0;
  // persistent grid per shape, and the captured graphs per (shape, small
  // batch): index 2 * shape + small
  static constexpr int kMappings = 12;
  // waves: [0] one lane per knot, [1] 16-lane groups, [2] two-wave workgroups,
  // [3 + 3 w + g] groups of 4, 7, 8 lanes (g = 0, 1, 2) in workgroups of 1, 2,
  // 4 waves (w = 0, 1, 2)
  unsigned grid[kMappings] = {};
  hipGraphExec_t exec[2 * kMappings] = {};
  hipGraph_t graph[2 * kMappings] = {};
  // batches of up to kSmallBatch trajectories replay, per mapping, a graph
  // captured with one trajectory list in its kernel's arguments (arg_key): the
  // same list again replays it as it is; another list is captured anew, once
  // the old graph's replays (arg_done) have finished
  hipGraph_t arg_graph[kMappings] = {};
  hipGraphExec_t arg_exec[kMappings] = {};
  hipEvent_t arg_done[kMappings] = {};
  std::vector<rbd_traj> arg_key[kMappings];
  int64_t last_total_knots = 0;
  int last_lanes = 0;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

constexpr size_t kHeaderBytes = 256;   // the header's slot before the descriptors
constexpr int64_t kMaxBatchChunks = (int64_t(1) << 31) - 1;   // the kernels index chunks with int

// Where a launch finds its queue: header and descriptors (the batch's own, or
// ones a caller staged) and the prefix-sum scratch.
struct Queue {
  const rbd::BatchHeader* hdr;
  const rbd::TrajDesc* descs;
  long* prefix;
  const rbd::ArgBatch* args = nullptr;   // a small batch in the kernel arguments instead
  int64_t chunks = 0;                     // the batch's chunk total when known: no more waves than chunks
  rbd::Completion done{nullptr, 0};       // completion word, if the caller polls one
};

// the batch's own header and descriptors, uploaded per launch
Queue own_queue(rbd_batch* b) { return Queue{b->d_hdr, b->d_desc, b->d_prefix}; }

// shape index (0 one lane, 1 16-lane groups, 2 two waves, 3 + 3 w + g groups
// of 4/7/8 lanes in workgroups of 1/2/4 waves) and knots per chunk
int mapping_of(Shape sh) {
  switch (sh.lanes) {
    case 1: return 0;
    case rbd::kLanesPerKnotGroup: return 1;
    case 4: case 7: case 8: {
      const int g = sh.lanes == 4 ? 0 : sh.lanes == 7 ? 1 : 2;
      const int w = sh.wpg == 4 ? 2 : sh.wpg == 2 ? 1 : 0;
      return 3 + 3 * w + g;
    }
    default: return 2;
  }
}
long chunk_knots(int lanes) {
  return lanes == 1 ? 64 : is_group_lanes(lanes) ? 64 / lanes : rbd::kTwoWaveKnots;
}
// The chunk total of a batch.  The L-lane groups take chunks of the batch's
// flattened knots (across trajectories); the other mappings chunk each
// trajectory on its own.
template <class Len>
int64_t chunk_total(int n_traj, Len len, int lanes) {
  const int64_t chunk = chunk_knots(lanes);
  int64_t total = 0;
  if (is_group_lanes(lanes)) {
    for (int t = 0; t < n_traj; ++t) total += len(t);
    return (total + chunk - 1) / chunk;
  }
  for (int t = 0; t < n_traj; ++t) total += (len(t) + chunk - 1) / chunk;
  return total;
}

template <class P, int LANES, bool SMALL>
void launch_persistent(rbd_batch* b, const Queue& q, int wpg, hipStream_t s) {
  // a batch of fewer chunks than the persistent grid launches one wave per
  // chunk (fewer to dispatch); the captured graphs of large batches keep the
  // full grid, whose launch serves every batch (small batches' graphs are
  // captured with theirs, replay_small)
  unsigned grid = b->grid[mapping_of(Shape{LANES, wpg})];
  if (q.chunks > 0 && q.chunks < int64_t(grid)) grid = unsigned(q.chunks);
  if constexpr (LANES == 4 || LANES == 7 || LANES == 8) {
    // `grid` waves in workgroups of wpg
    const dim3 groups((grid + unsigned(wpg) - 1) / unsigned(wpg)), block(kBlock * wpg);
    static const rbd::ArgBatch kNone{};
    const rbd::ArgBatch& a = q.args ? *q.args : kNone;
    if constexpr (SMALL)
      hipLaunchKernelGGL((rbd::batch_group_kernel<P, LANES, true>), groups, block, 0, s,
                         q.args ? nullptr : q.descs, q.hdr, a, b->d_tail, q.prefix, b->ctx->opt, q.done);
    else
      hipLaunchKernelGGL((rbd::batch_group_kernel<P, LANES, false>), groups, block, 0, s, q.descs,
                         q.hdr, rbd::NoArgBatch{}, b->d_tail, q.prefix, b->ctx->opt, q.done);
  } else if constexpr (LANES == kTwoWaveLanes) {
    static const rbd::ArgBatch kNone{};
    const rbd::ArgBatch& a = q.args ? *q.args : kNone;
    if constexpr (SMALL)
      hipLaunchKernelGGL((rbd::batch_two_wave_kernel<P, true>), dim3(grid), dim3(2 * kBlock), 0, s,
                         q.args ? nullptr : q.descs, q.hdr, a, b->d_tail, q.prefix, b->ctx->opt, q.done);
    else
      hipLaunchKernelGGL((rbd::batch_two_wave_kernel<P, false>), dim3(grid), dim3(2 * kBlock), 0, s, q.descs,
                         q.hdr, rbd::NoArgBatch{}, b->d_tail, q.prefix, b->ctx->opt, q.done);
  } else if constexpr (SMALL) {
    static const rbd::ArgBatch kNone{};
    const rbd::ArgBatch& a = q.args ? *q.args : kNone;
    hipLaunchKernelGGL((rbd::batch_persistent_kernel<P, LANES, true>), dim3(grid), dim3(kBlock), 0, s,
                       q.args ? nullptr : q.descs, q.hdr, a, b->d_tail, q.prefix, b->ctx->opt, q.done);
  } else {
    hipLaunchKernelGGL((rbd::batch_persistent_kernel<P, LANES, false>), dim3(grid), dim3(kBlock), 0, s, q.descs,
                       q.hdr, rbd::NoArgBatch{}, b->d_tail, q.prefix, b->ctx->opt, q.done);
  }
}

// The persistent grids: as many waves as a CU keeps resident -- with one lane
// per knot 6 (the 26,880-byte record blocks fill the LDS) with FP32 columns,
// 4 (one per SIMD) with FP64 columns; with 16-lane groups two per SIMD with
// FP32 columns, one with FP64 columns; two-wave workgroups at 256 registers,
// one wave per SIMD; the L-lane groups one wave per SIMD, whatever their
// workgroup size.
void set_default_grid(rbd_batch* b) {
  const rbd_context* ctx = b->ctx;
  const bool fp32_columns = ctx->precision == RBD_PREC_FP32 || ctx->precision == RBD_PREC_FP64_PIVOT ||
                            ctx->precision == RBD_PREC_FP32_HW_TRIG;
  b->grid[0] = unsigned(ctx->compute_units * (fp32_columns ? 6 : 4));
  b->grid[1] = unsigned(ctx->compute_units * 4 * (fp32_columns ? 2 : 1));
  b->grid[2] = unsigned(ctx->compute_units * 2);
  for (int m = 3; m < rbd_batch::kMappings; ++m) b->grid[m] = unsigned(ctx->compute_units * 4);
}

// Forward declaration: the direct launch, which the small-batch graphs capture.
hipError_t launch_batch(rbd_batch* b, const Queue& q, Shape sh, bool small, hipStream_t s);

// A small batch replayed from its mapping's graph.  The graph holds the
// batch's one kernel with the trajectory list in its arguments and its grid
// cut to the chunk count, as a direct launch has them, so a replay of the same
// list is one hipGraphLaunch.  Another list is captured into a new graph: the
// arguments of an instantiated graph could be rewritten instead
// (hipGraphExecKernelNodeSetParams), but on ROCm 7.2 every such update keeps
// about 34 KB of device memory until the graph is destroyed and costs about
// 8 us of host time.
hipError_t replay_small(rbd_batch* b, const rbd_traj* trajs, int n_traj, const rbd::ArgBatch& a, int64_t chunks,
                        Shape sh, hipStream_t st) {
  const int mi = mapping_of(sh);
  std::vector<rbd_traj>& key = b->arg_key[mi];
  const bool same = b->arg_exec[mi] && int(key.size()) == n_traj &&
                    std::memcmp(key.data(), trajs, sizeof(rbd_traj) * size_t(n_traj)) == 0;
  hipError_t e = hipSuccess;
  if (!same) {
    if (b->arg_exec[mi]) {
      e = hipEventSynchronize(b->arg_done[mi]);        // the old graph's last replay has finished
      (void)hipGraphExecDestroy(b->arg_exec[mi]);
      (void)hipGraphDestroy(b->arg_graph[mi]);
      b->arg_exec[mi] = nullptr;
      b->arg_graph[mi] = nullptr;
      key.clear();
      if (e != hipSuccess) return e;
    }
    if (!b->arg_done[mi]) {
      e = hipEventCreateWithFlags(&b->arg_done[mi], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    hipStream_t cs = nullptr;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      Queue q{nullptr, nullptr, b->d_prefix};
      q.args = &a;
      q.chunks = chunks;
      const hipError_t le = launch_batch(b, q, sh, true, cs);
      e = hipStreamEndCapture(cs, &b->arg_graph[mi]);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->arg_exec[mi], b->arg_graph[mi], nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
      if (b->arg_graph[mi]) (void)hipGraphDestroy(b->arg_graph[mi]);
      b->arg_graph[mi] = nullptr;
      b->arg_exec[mi] = nullptr;
      return e;
    }
    key.assign(trajs, trajs + n_traj);
  }
  e = hipGraphLaunch(b->arg_exec[mi], st);
  if (e == hipSuccess) e = hipEventRecord(b->arg_done[mi], st);
  return e;
}

template <class P>
hipError_t launch_batch_p(rbd_batch* b, const Queue& q, Shape sh, bool small, hipStream_t s) {
  if (!small)
    hipLaunchKernelGGL(rbd::batch_plan_kernel, dim3(1), dim3(rbd::kPlanThreads), 0, s, q.descs, q.hdr, q.prefix);
  const int lanes = sh.lanes;
  const int wpg = is_group_lanes(lanes) && valid_wpg(sh.wpg) ? sh.wpg : 1;
  if (lanes == 1) {
    if (small) launch_persistent<P, 1, true>(b, q, 1, s); else launch_persistent<P, 1, false>(b, q, 1, s);
  } else if (lanes == kTwoWaveLanes) {
    if (small) launch_persistent<P, kTwoWaveLanes, true>(b, q, 1, s);
    else launch_persistent<P, kTwoWaveLanes, false>(b, q, 1, s);
  } else if (lanes == 4) {
    if (small) launch_persistent<P, 4, true>(b, q, wpg, s); else launch_persistent<P, 4, false>(b, q, wpg, s);
  } else if (lanes == 7) {
    if (small) launch_persistent<P, 7, true>(b, q, wpg, s); else launch_persistent<P, 7, false>(b, q, wpg, s);
  } else if (lanes == 8) {
    if (small) launch_persistent<P, 8, true>(b, q, wpg, s); else launch_persistent<P, 8, false>(b, q, wpg, s);
  } else {
    if (small) launch_persistent<P, rbd::kLanesPerKnotGroup, true>(b, q, 1, s);
    else launch_persistent<P, rbd::kLanesPerKnotGroup, false>(b, q, 1, s);
  }
  return hipGetLastError();
}

hipError_t launch_batch(rbd_batch* b, const Queue& q, Shape sh, bool small, hipStream_t s) {
  switch (b->ctx->precision) {
    case RBD_PREC_FP32: return launch_batch_p<rbd::PrecFP32>(b, q, sh, small, s);
    case RBD_PREC_FP32_HW_TRIG: return launch_batch_p<rbd::PrecFP32HwTrig>(b, q, sh, small, s);
    case RBD_PREC_FP64_PIVOT: return launch_batch_p<rbd::PrecFP64Pivot>(b, q, sh, small, s);
    case RBD_PREC_FP64_COLUMNS: return launch_batch_p<rbd::PrecFP64Col>(b, q, sh, small, s);
    default: return launch_batch_p<rbd::PrecFP64>(b, q, sh, small, s);
  }
}

// header + descriptors of a batch into host memory at dst (kHeaderBytes, then
// the descriptors); returns the chunk total
long stage(char* dst, const rbd_traj* trajs, int n_traj, int lanes) {
  rbd::BatchHeader* h = reinterpret_cast<rbd::BatchHeader*>(dst);
  std::memset(h, 0, sizeof(*h));
  rbd::TrajDesc* d = reinterpret_cast<rbd::TrajDesc*>(dst + kHeaderBytes);
  for (int t = 0; t < n_traj; ++t) d[t] = rbd::TrajDesc{trajs[t].d_in, trajs[t].d_out, long(trajs[t].n_knots), 0};
  const long total = long(chunk_total(n_traj, [trajs](int t) { return int64_t(trajs[t].n_knots); }, lanes));
  h->n_traj = n_traj;
  // the plan kernel's prefix sum counts chunks of this many knots per
  // trajectory: single knots for the L-lane groups, whose chunks cross trajectories
  h->chunk_knots = is_group_lanes(lanes) ? 1 : int(chunk_knots(lanes));
  h->total_chunks = total;
  return total;
}

int graph_index(Shape sh, bool small) { return 2 * mapping_of(sh) + (small ? 1 : 0); }

// capture the batch's kernels for one (shape, small) pair into a graph (once)
rbd_status ensure_graph(rbd_batch* b, Shape sh, bool small) {
  const int gi = graph_index(sh, small);
  if (b->exec[gi]) return RBD_OK;
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate (capture)");
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t le = launch_batch(b, own_queue(b), sh, small, cs);
    e = hipStreamEndCapture(cs, &b->graph[gi]);
    if (le != hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&b->exec[gi], b->graph[gi], nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) {
    if (b->graph[gi]) (void)hipGraphDestroy(b->graph[gi]);
    b->graph[gi] = nullptr;
    b->exec[gi] = nullptr;
    return hip_fail(e, "batch graph capture");
  }
  return RBD_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free batched call (rbd_fdgrad_batched_noalloc_f32).  Its one
// piece of device memory beyond the caller's inputs and outputs is the queue's
// last-round counters (kTailWords words, 8,320 bytes), kept in this file-scope
// array: each device the library runs on holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in, not the device
// copy, so its device address is looked up once per device
// (hipGetSymbolAddress, at the first rbd_create on that device) and only that
// address is used.  Every call on a device shares its counters; the kernels
// leave them at zero after each launch, so calls in stream order need no
// clearing, but calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned long long scratch[rbd::kTailWords];

std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

__global__ void clear_scratch_kernel(unsigned long long* p, int n) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0ull;
}

}  // namespace

// Once per device, with that device current: the device address of its copy
// of `scratch`, cleared by a kernel (the counters start at zero).
rbd_status rbdi::noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(RBD_ERR_UNSUPPORTED, "device ordinal beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return RBD_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress (no-allocation scratch)");
  unsigned long long* p = static_cast<unsigned long long*>(ptr);
  hipLaunchKernelGGL(clear_scratch_kernel, dim3((rbd::kTailWords + 255) / 256), dim3(256), 0, nullptr, p,
                     rbd::kTailWords);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "clearing the no-allocation scratch");
  g_scratch_ptr[device] = p;
  return RBD_OK;
}

extern "C" {

rbd_status rbd_batch_create(rbd_context* ctx, int max_trajectories, rbd_batch** out) {
  if (!out) return fail(RBD_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  if (rbd_status s = check_ctx(ctx)) return s;
  if (max_trajectories <= 0 || max_trajectories > (1 << 24))
    return fail(RBD_ERR_INVALID_ARGUMENT, "max_trajectories must be in [1, 2^24]");
  if (rbd_status s = use_device(ctx)) return s;
  rbd_batch* b = new (std::nothrow) rbd_batch();
  if (!b) return fail(RBD_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->ctx = ctx;
  b->max_traj = max_trajectories;
  static_assert(sizeof(rbd::BatchHeader) <= kHeaderBytes, "header slot");
  const size_t hdr = kHeaderBytes;
  const size_t desc = align_up(sizeof(rbd::TrajDesc) * size_t(max_trajectories), 256);
  const size_t pre = sizeof(long) * (size_t(max_trajectories) + 1);
  b->upload_bytes = hdr + sizeof(rbd::TrajDesc) * size_t(max_trajectories);
  hipError_t e = hipMalloc(&b->d_mem, hdr + desc + pre);
  if (e != hipSuccess) { delete b; return fail(RBD_ERR_OUT_OF_MEMORY, "hipMalloc (batch)"); }
  b->d_hdr = reinterpret_cast<rbd::BatchHeader*>(b->d_mem);
  b->d_desc = reinterpret_cast<rbd::TrajDesc*>(b->d_mem + hdr);
  b->d_prefix = reinterpret_cast<long*>(b->d_mem + hdr + desc);
  const size_t tail_bytes = sizeof(unsigned long long) * rbd::kTailWords;
  e = hipMalloc(&b->d_tail, tail_bytes);
  if (e == hipSuccess) e = hipMemset(b->d_tail, 0, tail_bytes);   // once, outside any graph; kernels keep it zero
  if (e != hipSuccess) {
    if (b->d_tail) (void)hipFree(b->d_tail);
    (void)hipFree(b->d_mem);
    delete b;
    return fail(RBD_ERR_OUT_OF_MEMORY, "batch queue counters");
  }
  for (int i = 0; i < rbd_batch::kRing && e == hipSuccess; ++i) {
    e = hipHostMalloc(reinterpret_cast<void**>(&b->h_stage[i]), b->upload_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    rbd_batch_destroy(b);
    return hip_fail(e, "batch staging");
  }
  set_default_grid(b);
  *out = b;
  return RBD_OK;
}

rbd_status rbd_batch_destroy(rbd_batch* b) {
  if (!b) return RBD_OK;
  (void)use_device(b->ctx);
  for (int mi = 0; mi < rbd_batch::kMappings; ++mi) {
    if (b->arg_done[mi]) { (void)hipEventSynchronize(b->arg_done[mi]); (void)hipEventDestroy(b->arg_done[mi]); }
    if (b->arg_exec[mi]) (void)hipGraphExecDestroy(b->arg_exec[mi]);
    if (b->arg_graph[mi]) (void)hipGraphDestroy(b->arg_graph[mi]);
  }
  for (int gi = 0; gi < 2 * rbd_batch::kMappings; ++gi) {
    if (b->exec[gi]) (void)hipGraphExecDestroy(b->exec[gi]);
    if (b->graph[gi]) (void)hipGraphDestroy(b->graph[gi]);
  }
  for (int i = 0; i < rbd_batch::kRing; ++i) {
    if (b->staged[i]) { (void)hipEventSynchronize(b->staged[i]); (void)hipEventDestroy(b->staged[i]); }
    if (b->h_stage[i]) (void)hipHostFree(b->h_stage[i]);
  }
  if (b->d_mem) (void)hipFree(b->d_mem);
  if (b->d_tail) (void)hipFree(b->d_tail);
  delete b;
  return RBD_OK;
}

rbd_status rbd_batch_set_grid(rbd_batch* b, int lanes_per_knot, unsigned waves) {
  if (!b) return fail(RBD_ERR_INVALID_ARGUMENT, "null batch");
  if (lanes_per_knot != 1 && lanes_per_knot != rbd::kLanesPerKnotGroup && !is_group_lanes(lanes_per_knot))
    return fail(RBD_ERR_INVALID_ARGUMENT, "lanes_per_knot must be 1, 4, 7, 8 or 16");
  if (waves == 0 || waves > (1u << 20)) return fail(RBD_ERR_INVALID_ARGUMENT, "waves must be in [1, 2^20]");
  // every workgroup size of the mapping
  for (int wpg : {1, 2, 4}) {
    const Shape sh{lanes_per_knot, wpg};
    const int mi = mapping_of(sh);
    if (b->grid[mi] == waves) continue;
    b->grid[mi] = waves;
    for (int small = 0; small < 2; ++small) {
      const int gi = graph_index(sh, small != 0);
      if (b->exec[gi]) { (void)hipGraphExecDestroy(b->exec[gi]); b->exec[gi] = nullptr; }
      if (b->graph[gi]) { (void)hipGraphDestroy(b->graph[gi]); b->graph[gi] = nullptr; }
    }
  }
  return RBD_OK;
}

rbd_status rbd_fdgrad_batched_f32(rbd_batch* b, const rbd_traj* trajs, int n_traj, int use_graph,
                                  rbd_stream_t stream) {
  return rbdi::batch_run(b, trajs, n_traj, use_graph, stream, Shape{});
}

}  // extern "C"

// shape_req.lanes = 0: by the context (shape_for_batch)
rbd_status rbdi::batch_run(rbd_batch* b, const rbd_traj* trajs, int n_traj, int use_graph, rbd_stream_t stream,
                           Shape shape_req) {
  if (!b) return fail(RBD_ERR_INVALID_ARGUMENT, "null batch");
  if (n_traj < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative trajectory count");
  if (n_traj > b->max_traj) return fail(RBD_ERR_CAPACITY, "more trajectories than the batch was created for");
  if (n_traj > 0 && !trajs) return fail(RBD_ERR_INVALID_ARGUMENT, "null trajectory list");
  int64_t total = 0, max_chunks = 0;
  for (int t = 0; t < n_traj; ++t) {
    if (trajs[t].n_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative trajectory length");
    if (trajs[t].n_knots > 0 && (!trajs[t].d_in || !trajs[t].d_out))
      return fail(RBD_ERR_INVALID_ARGUMENT, "null trajectory buffer");
    total += trajs[t].n_knots;
    max_chunks += (trajs[t].n_knots + 3) / 4;   // chunks of the 4-knot mapping (the larger count)
  }
  if (max_chunks > kMaxBatchChunks) return fail(RBD_ERR_CAPACITY, "batch too large (more than 2^31 chunks)");
  b->last_total_knots = total;
  if (total == 0) return RBD_OK;
  if (rbd_status s = ready_to_launch(b->ctx)) return s;
  const hipStream_t st = (hipStream_t)stream;
  const Shape sh = shape_req.lanes ? shape_req : shape_for_batch(b->ctx, total);
  const int lanes = sh.lanes;
  b->last_lanes = lanes;
  const bool small = n_traj <= rbd::kSmallBatch;
  if (small) {
    // the descriptors go in the kernel arguments: no upload
    rbd::ArgBatch a{};
    for (int t = 0; t < n_traj; ++t) a.d[t] = rbd::TrajDesc{trajs[t].d_in, trajs[t].d_out, long(trajs[t].n_knots), 0};
    const long chunks = long(chunk_total(n_traj, [trajs](int t) { return int64_t(trajs[t].n_knots); }, lanes));
    a.n_traj = n_traj;
    a.total_chunks = int(chunks);
    if (use_graph) {
      hipError_t ge = replay_small(b, trajs, n_traj, a, chunks, sh, st);
      if (ge != hipSuccess) return hip_fail(ge, "batch graph (small batch)");
      return RBD_OK;
    }
    Queue q{nullptr, nullptr, b->d_prefix};
    q.args = &a;
    q.chunks = chunks;
    hipError_t le = launch_batch(b, q, sh, true, st);
    if (le != hipSuccess) return hip_fail(le, "batch kernel launch");
    return RBD_OK;
  }
  // stage header + descriptors in pinned memory, copy to the device on `stream`
  const int r = b->ring;
  b->ring = (r + 1) % rbd_batch::kRing;
  hipError_t e = hipEventSynchronize(b->staged[r]);     // previous copy out of this slot is done
  if (e != hipSuccess) return hip_fail(e, "staging wait");
  const int64_t chunks = stage(b->h_stage[r], trajs, n_traj, lanes);
  e = hipMemcpyAsync(b->d_mem, b->h_stage[r], batch_stage_bytes(n_traj), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipEventRecord(b->staged[r], st);
  if (e != hipSuccess) return hip_fail(e, "descriptor upload");
  if (use_graph) {
    if (rbd_status s = ensure_graph(b, sh, small)) return s;
    e = hipGraphLaunch(b->exec[graph_index(sh, small)], st);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  } else {
    Queue q = own_queue(b);
    q.chunks = chunks;
    e = launch_batch(b, q, sh, small, st);
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  }
  return RBD_OK;
}

size_t rbdi::batch_stage_bytes(int n_traj) { return kHeaderBytes + sizeof(rbd::TrajDesc) * size_t(n_traj); }

int64_t rbdi::batch_stage(void* host_dst, const rbd_traj* trajs, int n_traj, int lanes) {
  return stage(static_cast<char*>(host_dst), trajs, n_traj, lanes);
}

rbd_status rbdi::batch_launch_staged(rbd_batch* b, const void* d_staged, const rbd_traj* trajs, int n_traj,
                                     int64_t total_knots, int64_t total_chunks, Shape sh, rbd_stream_t stream,
                                     unsigned long long* done_flag, unsigned long long done_seq) {
  if (n_traj > b->max_traj) return fail(RBD_ERR_CAPACITY, "more trajectories than the batch was created for");
  if (total_chunks > kMaxBatchChunks) return fail(RBD_ERR_CAPACITY, "batch too large (more than 2^31 chunks)");
  b->last_total_knots = total_knots;
  b->last_lanes = sh.lanes;
  if (total_knots == 0) return RBD_OK;
  const char* base = static_cast<const char*>(d_staged);
  Queue q{reinterpret_cast<const rbd::BatchHeader*>(base),
          reinterpret_cast<const rbd::TrajDesc*>(base + kHeaderBytes), b->d_prefix};
  q.chunks = total_chunks;
  q.done = rbd::Completion{done_flag, done_seq};
  const bool small = n_traj <= rbd::kSmallBatch;
  rbd::ArgBatch a{};
  if (small && trajs) {              // the descriptors in the kernel arguments: no reads of the staged block
    for (int t = 0; t < n_traj; ++t)
      a.d[t] = rbd::TrajDesc{trajs[t].d_in, trajs[t].d_out, long(trajs[t].n_knots), 0};
    a.n_traj = n_traj;
    a.total_chunks = int(total_chunks);
    q.args = &a;
  }
  hipError_t e = launch_batch(b, q, sh, small, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  return RBD_OK;
}

extern "C" {

int rbd_batch_last_lanes(const rbd_batch* b) { return b ? b->last_lanes : 0; }

rbd_status rbd_fdgrad_batched_noalloc_f32(rbd_context* ctx, const rbd_traj* trajs, int n_traj, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n_traj < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative trajectory count");
  if (n_traj > rbd::kSmallBatch)
    return fail(RBD_ERR_CAPACITY, "rbd_fdgrad_batched_noalloc_f32 takes at most 64 trajectories");
  if (n_traj > 0 && !trajs) return fail(RBD_ERR_INVALID_ARGUMENT, "null trajectory list");
  int64_t total = 0;
  for (int t = 0; t < n_traj; ++t) {
    if (trajs[t].n_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative trajectory length");
    if (trajs[t].n_knots > 0 && (!trajs[t].d_in || !trajs[t].d_out))
      return fail(RBD_ERR_INVALID_ARGUMENT, "null trajectory buffer");
    total += trajs[t].n_knots;
  }
  if ((total + 3) / 4 + n_traj > kMaxBatchChunks) return fail(RBD_ERR_CAPACITY, "batch too large (more than 2^31 chunks)");
  if (total == 0) return RBD_OK;
  if (rbd_status s = ready_to_launch(ctx)) return s;
  unsigned long long* tail = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) tail = g_scratch_ptr[ctx->device];
  }
  if (!tail) return fail(RBD_ERR_UNSUPPORTED, "no-allocation scratch not set up on this device");
  const Shape sh = shape_for_batch(ctx, total);
  const int lanes = sh.lanes;
  // the batch in the kernel arguments, the counters in the device's scratch,
  // the grid as a batch object's: nothing is allocated
  rbd_batch b;
  b.ctx = ctx;
  b.d_tail = tail;
  set_default_grid(&b);
  rbd::ArgBatch a{};
  for (int t = 0; t < n_traj; ++t) a.d[t] = rbd::TrajDesc{trajs[t].d_in, trajs[t].d_out, long(trajs[t].n_knots), 0};
  a.n_traj = n_traj;
  a.total_chunks = int(chunk_total(n_traj, [trajs](int t) { return int64_t(trajs[t].n_knots); }, lanes));
  Queue q{nullptr, nullptr, nullptr};
  q.args = &a;
  q.chunks = a.total_chunks;
  const hipError_t e = launch_batch(&b, q, sh, true, (hipStream_t)stream);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "batch kernel launch (no allocation)");
}

// Times every shape of the batched path (kAutotuneShapes) on batches of
// 132-knot trajectories (the mean length of lengths uniform over 8..256), 8 to
// 512 of them, and fills the context's batch mapping table with the fastest
// at each total.  One more point is 64 trajectories of 142 knots (9,088
// knots): just under one wave of 7-lane groups per SIMD on 256 CUs, in the
// kernel-argument form of batches of up to 64 trajectories.
rbd_status rbd_autotune_batch(rbd_context* ctx, rbd_stream_t stream, int64_t* threshold_out, float* table_us,
                              int table_len) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (rbd_status s = ready_to_launch(ctx)) return s;
  constexpr int64_t kLen = 132;
  static const int counts[] = {8, 16, 32, 48, 64, 64, 80, 96, 128, 256, 512};
  static const int64_t lens[] = {kLen, kLen, kLen, kLen, kLen, 142, kLen, kLen, kLen, kLen, kLen};
  constexpr int kSizes = sizeof(counts) / sizeof(counts[0]);
  static_assert(kSizes <= kMaxTuned, "mapping table size");
  const int max_traj = counts[kSizes - 1];
  const size_t knots = size_t(max_traj) * kLen;
  const hipStream_t st = (hipStream_t)stream;
  float* buf = nullptr;
  hipError_t e = hipMalloc(&buf, knots * (rbd::kInPerKnot + rbd::kOutPerKnot) * sizeof(float));
  if (e != hipSuccess) return fail(RBD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  // any finite inputs do: the work per knot does not depend on the values
  e = hipMemsetAsync(buf, 0, knots * rbd::kInPerKnot * sizeof(float), st);
  rbd_batch* b = nullptr;
  rbd_status rs = e == hipSuccess ? rbd_batch_create(ctx, max_traj, &b) : RBD_OK;
  // trajectory t of a batch of length len starts at knot t * len (len <= 142
  // for t < 64, 132 above: every batch stays inside the buffer)
  static_assert(64 * 142 <= 512 * kLen, "the longer trajectories fit the buffer");
  std::vector<rbd_traj> trajs(static_cast<size_t>(max_traj));
  auto place = [&](int64_t len) {
    for (int t = 0; t < max_traj; ++t) {
      const int64_t n = t < 64 ? len : kLen;
      const size_t k0 = t < 64 ? size_t(t) * size_t(len) : size_t(t) * size_t(kLen);
      trajs[size_t(t)] = rbd_traj{buf + k0 * rbd::kInPerKnot, buf + knots * rbd::kInPerKnot + k0 * rbd::kOutPerKnot, n};
    }
  };
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess && rs == RBD_OK) e = hipEventCreate(&t0);
  if (e == hipSuccess && rs == RBD_OK) e = hipEventCreate(&t1);
  float best_us[kAutotuneMappings][kSizes] = {};
  for (int si = 0; si < kSizes && e == hipSuccess && rs == RBD_OK; ++si) {
    place(lens[si]);
    for (int li = 0; li < kAutotuneMappings && e == hipSuccess && rs == RBD_OK; ++li) {
      const Shape sh = kAutotuneShapes[li];
      float best = 1e30f;
      for (int rep = 0; rep < 3 && e == hipSuccess && rs == RBD_OK; ++rep) {
        rs = rbdi::batch_run(b, trajs.data(), counts[si], 0, stream, sh);   // warm
        if (rs == RBD_OK) e = hipEventRecord(t0, st);
        for (int it = 0; it < 4 && e == hipSuccess && rs == RBD_OK; ++it)
          rs = rbdi::batch_run(b, trajs.data(), counts[si], 0, stream, sh);
        if (e == hipSuccess && rs == RBD_OK) e = hipEventRecord(t1, st);
        if (e == hipSuccess && rs == RBD_OK) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess && rs == RBD_OK) e = hipEventElapsedTime(&ms, t0, t1);
        if (e == hipSuccess && rs == RBD_OK && ms * 250.0f < best) best = ms * 250.0f;
      }
      best_us[li][si] = best;
    }
  }
  (void)hipStreamSynchronize(st);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (b) rbd_batch_destroy(b);
  (void)hipFree(buf);
  if (rs != RBD_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "batch autotune");
  // the table: the fastest mapping at each total; the crossover between one
  // lane and 16-lane groups keeps its meaning (the largest total at which the
  // 16-lane groups still win)
  int64_t threshold = 0;
  ctx->batch_tuned.count.store(0, std::memory_order_release);
  for (int si = 0; si < kSizes; ++si) {
    int bi = 0;
    for (int li = 1; li < kAutotuneMappings; ++li)
      if (best_us[li][si] < best_us[bi][si]) bi = li;
    ctx->batch_tuned.n[si] = counts[si] * lens[si];
    ctx->batch_tuned.lanes[si] = kAutotuneShapes[bi].lanes;
    ctx->batch_tuned.wpg[si] = kAutotuneShapes[bi].wpg;
    if (best_us[1][si] < best_us[0][si]) threshold = counts[si] * lens[si];
  }
  ctx->batch_tuned.count.store(kSizes, std::memory_order_release);
  ctx->batch_group_threshold.store(threshold);
  if (threshold_out) *threshold_out = threshold;
  constexpr int kRow = 1 + kAutotuneMappings;
  if (table_us)
    for (int i = 0; i < table_len && i < kRow * kSizes; ++i)
      table_us[i] = (i % kRow == 0) ? float(counts[i / kRow] * lens[i / kRow]) : best_us[(i % kRow) - 1][i / kRow];
  return RBD_OK;
}

int64_t rbd_batch_group_threshold(rbd_context* ctx) { return ctx ? ctx->batch_group_threshold.load() : -1; }

rbd_status rbd_set_batch_group_threshold(rbd_context* ctx, int64_t n_knots) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative threshold");
  ctx->batch_tuned.count.store(0, std::memory_order_release);   // set by hand: the threshold decides
  ctx->batch_group_threshold.store(n_knots);
  return RBD_OK;
}

}  // extern "C"
