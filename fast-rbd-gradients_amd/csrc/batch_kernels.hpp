// Batched uneven trajectories: a device work queue of knot chunks and the
// persistent kernel that drains it.
//
// A batch is a list of trajectories, each with its own input/output buffers
// and length.  A chunk is the knots one wave processes at a time: 64 in the
// one-lane mapping, 4 in the 16-lane-group mapping.  The queue holds the
// chunks in trajectory order; a chunk's trajectory comes from the prefix sum
// of chunk counts:
//   * up to 64 trajectories (SMALL): a wave scans the counts itself, one
//     trajectory per lane, and finds the chunk's trajectory with a ballot and
//     a population count -- no plan kernel, no memory traffic beyond the counts;
//   * more: the plan kernel writes the prefix sum first, and the waves
//     binary-search it.
// With G waves in the grid and T chunks, the waves pull the chunks of the full
// rounds, floor(T / G) * G of them, in a fixed order (wave w: w, w + G, ...),
// which needs no atomic: every chunk costs the same, so a fixed order is
// already balanced.  The last partial round, T mod G chunks, is pulled through
// atomic counters so that waves that start or finish early take more of it.
// Its chunks are dealt round-robin over S = min(64, G) counters, each on its
// own 128-byte line, and wave w claims from counter w mod S: one address is
// never contended by more than G / S waves (all G waves on one counter was the
// kernel's main overhead).  Each wave claims until a claim fails, so a counter
// takes exactly (its tail chunks + its waves) increments per launch, and the
// wave whose claim is the last of them sets it back to zero: the counters
// start every launch at zero with no host work, no upload and no memset, also
// in graph replays.  A batch object's launches must therefore run one at a
// time (one stream, or stream-ordered).  Every wave reaches the tail, every
// claim either returns a chunk or ends the wave's loop, and the totals are
// fixed before the kernel starts, so the grid always drains.  The header and
// descriptors are only read, so they may live in device memory or in pinned
// host memory; the counters must be device memory.
#pragma once
#include <hip/hip_runtime.h>

#include "fdgrad_knot.hpp"
#include "group_mapping.hpp"
#include "records.hpp"
#include "two_wave.hpp"

namespace rbd {

struct TrajDesc {
  const float* in;   // n x 21 (AoS)
  float* out;        // n x 105 (AoS)
  long n;            // knots
  long pad;
};

struct BatchHeader {
  int n_traj;                     // trajectories
  int chunk_knots;                // knots per chunk
  long total_chunks;              // chunk total, computed by the host (below 2^31)
};

// counters of a launch's last partial round
constexpr int kTailCounters = 64;
constexpr int kTailStride = 16;   // 64-bit words between counters: one 128-byte line each
// after them, on its own line: the count of waves that finished (completion word)
constexpr int kDoneCounter = kTailCounters * kTailStride;
constexpr int kTailWords = kDoneCounter + kTailStride;

// Completion word for a host that polls memory instead of an event: with flag
// set (host-mapped, coherent memory), the launch's last wave to finish writes
// seq to it once every wave's stores are visible to the system.
struct Completion {
  unsigned long long* flag;
  unsigned long long seq;
};

constexpr int kSmallBatch = 64;   // trajectories handled without the plan kernel

// A small batch passed by value in the kernel arguments, so that a direct
// launch needs no descriptor upload; used when descs is null.
struct ArgBatch {
  int n_traj;
  int total_chunks;
  TrajDesc d[kSmallBatch];
};
struct NoArgBatch {};

constexpr int kPlanThreads = 1024;

// One block: prefix[t] = sum of chunk counts of trajectories < t, prefix[n] = total.
__global__ __launch_bounds__(kPlanThreads) void batch_plan_kernel(const TrajDesc* __restrict__ descs,
                                                                  const BatchHeader* __restrict__ hdr,
                                                                  long* __restrict__ prefix) {
  __shared__ long partial[kPlanThreads];
  const int n = hdr->n_traj;
  const int chunk = hdr->chunk_knots;
  // each thread owns a contiguous run of trajectories
  const int per = (n + kPlanThreads - 1) / kPlanThreads;
  const int t0 = threadIdx.x * per;
  const int t1 = min(n, t0 + per);
  long sum = 0;
  for (int t = t0; t < t1; ++t) sum += (descs[t].n + chunk - 1) / chunk;
  partial[threadIdx.x] = sum;
  __syncthreads();
  // inclusive scan over the block (Hillis-Steele)
  for (int off = 1; off < kPlanThreads; off <<= 1) {
    const long v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0;
    __syncthreads();
    partial[threadIdx.x] += v;
    __syncthreads();
  }
  long run = partial[threadIdx.x] - sum;   // exclusive prefix of this thread's run
  for (int t = t0; t < t1; ++t) {
    prefix[t] = run;
    run += (descs[t].n + chunk - 1) / chunk;
  }
  if (threadIdx.x == kPlanThreads - 1) prefix[n] = partial[kPlanThreads - 1];
}

// The trajectory t of chunk c and t's first chunk.  SMALL: the counts are
// scanned in the calling wave, one trajectory per lane (every lane of the wave
// takes part); otherwise a binary search of the plan kernel's prefix sum.
template <bool SMALL>
__device__ inline void locate_chunk(const TrajDesc* descs, const long* prefix, int n_traj, int chunk, int c,
                                    int lane, int& t, int& first) {
  if constexpr (SMALL) {
    int v = lane < n_traj ? int((descs[lane].n + chunk - 1) / chunk) : 0;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    // trajectories that end at or before chunk c are the lanes with v <= c
    t = __popcll(__ballot(v <= c));
    first = __shfl(v, t > 0 ? t - 1 : 0, 64);
    if (t == 0) first = 0;
  } else {
    int lo = 0, hi = n_traj;               // prefix[lo] <= c < prefix[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= c) lo = mid; else hi = mid;
    }
    t = lo;
    first = int(prefix[lo]);
  }
}

// End of a launch that publishes a completion word (done.flag set): each
// wave's stores (zero-copy outputs in host memory included) are written back
// and acknowledged at system scope before the wave is counted (a release
// fence: no cache invalidation, which would evict the data of other launches
// running beside this one); the count is an acquire-release add, so the last
// of the launch's `waves` waves to be counted acquires every other wave's
// release, and every output is visible when it resets the count for the next
// launch and publishes seq with a release store.
__device__ inline void publish_completion(const Completion& done, unsigned long long* tail, unsigned waves,
                                          int lane) {
  if (done.flag == nullptr) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  if (lane == 0) {
    unsigned long long* fin = tail + kDoneCounter;
    if (__hip_atomic_fetch_add(fin, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
        static_cast<unsigned long long>(waves - 1)) {
      atomicExch(fin, 0ull);
      __hip_atomic_store(done.flag, done.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// The share of wave w (of the launch's G waves) of a launch's `total`
// chunks, handed to process(c) one at a time: the full rounds (all chunks when
// they fit one round) in a fixed order with no atomics, then the last partial
// round through the tail counters (tail: the kTailCounters counters
// kTailStride words apart, zero on entry and left at zero).  w is the same in
// every lane of the wave.
template <class F>
__device__ inline void drain_chunks(int w, int G, int total, unsigned long long* tail, int lane, F&& process) {
  const int fixed = total <= G ? total : total - total % G;
  const int rest = total - fixed;
  const int S = G < kTailCounters ? G : kTailCounters;
  const int s = w % S;
  const int mine = s < rest ? (rest - 1 - s) / S + 1 : 0;     // tail chunks dealt to counter s
  const int waves = (G - 1 - s) / S + 1;                      // waves that claim from counter s
  unsigned long long* ctr = tail + s * kTailStride;
  int next = w;
  for (;;) {
    int c;
    if (next < fixed) {
      c = next;
      next += G;
    } else {
      if (rest == 0) break;
      int j = 0;
      if (lane == 0) {
        j = int(atomicAdd(ctr, 1ull));
        // the launch's last claim on this counter: back to zero for the next launch
        if (j == mine + waves - 1) atomicExch(ctr, 0ull);
      }
      j = __builtin_amdgcn_readfirstlane(j);
      if (j >= mine) break;
      c = fixed + s + j * S;
    }
    process(c);
  }
}

// Persistent drain of the queue.  LANES = 1: lane i of a wave takes knot
// k0 + i of a 64-knot chunk; LANES = 16: lane group g takes knot k0 + g of a
// 4-knot chunk.  One wave per workgroup.  tail: kTailWords words (the
// kTailCounters counters kTailStride words apart, then the finished-wave
// count), zero when the launch starts (and left at zero).  SMALL batches may
// come in the arguments (descs == nullptr: arg holds them).
template <class P, int LANES, bool SMALL>
__global__ __launch_bounds__(64, kWavesPerSimd<P>) void batch_persistent_kernel(
    const TrajDesc* __restrict__ descs, const BatchHeader* __restrict__ hdr,
    const typename std::conditional<SMALL, ArgBatch, NoArgBatch>::type arg, unsigned long long* __restrict__ tail,
    const long* __restrict__ prefix, Options opt, Completion done) {
  constexpr int chunk = 64 / LANES;
  int n_traj, total;
  if constexpr (SMALL) {
    if (descs == nullptr) {
      descs = arg.d;
      n_traj = arg.n_traj;
      total = arg.total_chunks;
    } else {
      n_traj = hdr->n_traj;
      total = int(hdr->total_chunks);
    }
  } else {
    n_traj = hdr->n_traj;
    total = int(hdr->total_chunks);
  }
  const int lane = int(threadIdx.x);
  const int slot = lane / LANES;            // knot slot in the chunk
  const int role = lane % LANES;
  __shared__ typename P::Minv mi_sh[64 / LANES][kMinvPacked];   // Minv per lane group (LANES = 16)

  // lane groups: the chunk's records (4 consecutive knots of one trajectory,
  // 1,680 contiguous bytes) are assembled in LDS and written with consecutive
  // lanes on consecutive floats: whole lines instead of each lane's 7-float
  // column, which left every line partly written by each store (and, for
  // zero-copy batches, one PCIe write per partial line)
  // (the chunk's inputs, 336 contiguous bytes, are read the same way into LDS
  // first: a few whole-line reads instead of many small per-lane ones)
  // one lane per knot: the chunk's records (64 consecutive knots of one
  // trajectory, one contiguous block) are assembled in LDS and copied out
  // whole (copy_record_block, streaming stores), as the flat one-lane kernel does
  __shared__ __align__(16) float rec_sh[chunk * kOutPerKnot];
  __shared__ float in_sh[LANES == 1 ? 1 : chunk * kInPerKnot];
  auto process = [&](int c) {
    int t, first;                           // trajectory of chunk c, and its first chunk
    locate_chunk<SMALL>(descs, prefix, n_traj, chunk, c, lane, t, first);
    const TrajDesc d = descs[t];
    const long k0 = long(c - first) * chunk;
    const long k = k0 + slot;
    if constexpr (LANES == 1) {
      const int knots = int(d.n - k0 < chunk ? d.n - k0 : chunk);
      // lanes past the trajectory's end compute a copy of its last knot and store nothing
      InAoS i{d.in + kInPerKnot * (k < d.n ? k : d.n - 1)};
      OutAoS o{rec_sh + kOutPerKnot * slot};
      fdgrad_knot<P>(i, o, opt);
      wave_lds_sync();
      // streaming stores, as the flat kernel's (whole lines; made visible by
      // the kernel's end or the completion word's release like any other store)
      copy_record_block<true>(rec_sh, d.out + kOutPerKnot * k0, knots * kOutPerKnot, lane);
      wave_lds_sync();                      // read out before the next chunk overwrites it
    } else {
      const int knots = int(d.n - k0 < chunk ? d.n - k0 : chunk);
      for (int e = lane; e < knots * kInPerKnot; e += 64) in_sh[e] = d.in[kInPerKnot * k0 + e];
      wave_lds_sync();
      if (k < d.n) {
        InAoS i{in_sh + kInPerKnot * slot};
        OutAoS o{rec_sh + kOutPerKnot * slot};
        fdgrad_knot_lanes<P>(i, o, opt, role, mi_sh[slot]);
      }
      wave_lds_sync();
      for (int e = lane; e < knots * kOutPerKnot; e += 64) d.out[kOutPerKnot * k0 + e] = rec_sh[e];
      wave_lds_sync();                      // read before the next chunk overwrites it
    }
  };

  drain_chunks(int(blockIdx.x), int(gridDim.x), total, tail, lane, process);
  publish_completion(done, tail, unsigned(gridDim.x), lane);
}

// Batches on the L-lane groups (L = 4, 7, 8; group_mapping.hpp).  A chunk is
// kSlots = floor(64 / L) consecutive knots of the batch's flattened knot
// sequence (the trajectories one after another), so a chunk may span several
// trajectories and no lane idles at a trajectory's end: the 9,014 knots of 64
// trajectories of 8..256 knots are 1,002 chunks of 9 with L = 7.  Each lane
// finds its own knot's trajectory: SMALL batches from a shuffle scan of the
// (up to 64) lengths, searched per lane; larger ones by a binary search of the
// plan kernel's knot prefix sum (header chunk_knots = 1) for the chunk's first
// knot, from which each lane walks on.  Workgroups of 1, 2 or 4 waves
// (blockDim.x = 64 x waves), each wave on its own: its own chunks and LDS
// block.  Queue, descriptors in the arguments and completion word as in
// batch_persistent_kernel.
template <class P, int L, bool SMALL>
__global__ __launch_bounds__(kMaxGroupWaves * 64, 1) void batch_group_kernel(
    const TrajDesc* __restrict__ descs, const BatchHeader* __restrict__ hdr,
    const typename std::conditional<SMALL, ArgBatch, NoArgBatch>::type arg, unsigned long long* __restrict__ tail,
    const long* __restrict__ prefix, Options opt, Completion done) {
  using Gr = Group<L>;
  int n_traj, total;
  if constexpr (SMALL) {
    if (descs == nullptr) {
      descs = arg.d;
      n_traj = arg.n_traj;
      total = arg.total_chunks;
    } else {
      n_traj = hdr->n_traj;
      total = int(hdr->total_chunks);
    }
  } else {
    n_traj = hdr->n_traj;
    total = int(hdr->total_chunks);
  }
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);   // in the workgroup
  const int lane = int(threadIdx.x) % 64;
  const int slot = group_slot<L>(lane);
  const int waves_per_group = int(blockDim.x) / 64;
  // Minv per knot (L >= 7), per wave of the workgroup
  __shared__ typename P::Minv mi_sh[kMaxGroupWaves][Gr::kMinvColumns ? Gr::kSlots : 1][kMinvPacked];
  // SMALL: lane l holds the knots of trajectories 0..l (inclusive prefix sum)
  long incl = 0;
  if constexpr (SMALL) {
    incl = lane < n_traj ? descs[lane].n : 0;
    for (int off = 1; off < 64; off <<= 1) {
      const long o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
  }
  const long knots = SMALL ? __shfl(incl, 63, 64) : prefix[n_traj];
  auto process = [&](int c) {
    long kk = long(c) * Gr::kSlots + slot;      // the lane's knot in the batch
    const bool valid = kk < knots && lane < Gr::kLanesUsed;
    if (kk >= knots) kk = knots - 1;            // past the last knot: a copy of it, nothing stored
    int t;
    long first;
    if constexpr (SMALL) {
      // t = the number of trajectories that end at or before knot kk
      int lo = 0;
      for (int step = 32; step >= 1; step >>= 1) {
        const long v = __shfl(incl, lo + step - 1, 64);
        if (v <= kk) lo += step;
      }
      t = lo;
      first = __shfl(incl, t > 0 ? t - 1 : 0, 64);
      if (t == 0) first = 0;
    } else {
      const long kk0 = long(c) * Gr::kSlots;    // the chunk's first knot: wave-uniform search
      int lo = 0, hi = n_traj;                  // prefix[lo] <= kk0 < prefix[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= kk0) lo = mid; else hi = mid;
      }
      t = lo;
      while (t + 1 < n_traj && prefix[t + 1] <= kk) ++t;
      first = prefix[t];
    }
    const TrajDesc d = descs[t];
    const long k = kk - first;
    InAoS i{d.in + kInPerKnot * k};
    OutMasked<OutAoS> o{OutAoS{d.out + kOutPerKnot * k}, valid};
    fdgrad_knot_group<P, L>(i, o, opt, lane, mi_sh[wave][Gr::kMinvColumns ? slot : 0]);
    wave_lds_sync();                            // every lane is done with mi_sh before the next chunk
  };
  const int G = int(gridDim.x) * waves_per_group;
  drain_chunks(int(blockIdx.x) * waves_per_group + wave, G, total, tail, lane, process);
  publish_completion(done, tail, unsigned(G), lane);
}

// Batches on the two-wave mapping: each workgroup of two waves takes
// kTwoWaveKnots-knot chunks w, w + G, ... (G workgroups) in a fixed order --
// both waves need the same chunk, and every chunk costs the same -- and runs
// each as two_wave_group.  Chunk lookup, descriptors in the arguments and the
// completion word as in batch_persistent_kernel.
template <class P, bool SMALL>
__global__ __launch_bounds__(128, 1) void batch_two_wave_kernel(
    const TrajDesc* __restrict__ descs, const BatchHeader* __restrict__ hdr,
    const typename std::conditional<SMALL, ArgBatch, NoArgBatch>::type arg, unsigned long long* __restrict__ tail,
    const long* __restrict__ prefix, Options opt, Completion done) {
  int n_traj, total;
  if constexpr (SMALL) {
    if (descs == nullptr) {
      descs = arg.d;
      n_traj = arg.n_traj;
      total = arg.total_chunks;
    } else {
      n_traj = hdr->n_traj;
      total = int(hdr->total_chunks);
    }
  } else {
    n_traj = hdr->n_traj;
    total = int(hdr->total_chunks);
  }
  __shared__ TwoWaveShared<P> sh[kTwoWaveKnots];
  __shared__ int flags[2];                 // [0]: bias ready (R -> M), [1]: Minv and corrections ready (M -> R)
  // the chunk's records, assembled from both waves and written with
  // consecutive threads on consecutive floats (as in batch_persistent_kernel)
  __shared__ float rec_sh[kTwoWaveKnots * kOutPerKnot];
  __shared__ float in_sh[kTwoWaveKnots * kInPerKnot];   // the chunk's inputs, read the same way
  const int lane = int(threadIdx.x % 64);
  const int g = lane / kLanesPerKnotGroup;
  for (int c = int(blockIdx.x); c < total; c += int(gridDim.x)) {
    __syncthreads();                       // the previous chunk is done with sh, flags and rec_sh
    if (threadIdx.x == 0) { flags[0] = 0; flags[1] = 0; }
    __syncthreads();
    int t, first;
    locate_chunk<SMALL>(descs, prefix, n_traj, kTwoWaveKnots, c, lane, t, first);
    const TrajDesc d = descs[t];           // d.n >= 1: chunk c lies in t
    const long k0 = long(c - first) * kTwoWaveKnots;
    const long k = k0 + g;
    const bool valid = k < d.n;
    const int knots = int(d.n - k0 < kTwoWaveKnots ? d.n - k0 : kTwoWaveKnots);
    for (int e = int(threadIdx.x); e < knots * kInPerKnot; e += int(blockDim.x))
      in_sh[e] = d.in[kInPerKnot * k0 + e];
    __syncthreads();
    // lanes past the trajectory's end compute a copy of its last knot and store nothing
    InAoS i{in_sh + kInPerKnot * (valid ? g : knots - 1)};
    OutMasked<OutAoS> o{OutAoS{rec_sh + kOutPerKnot * g}, valid};
    two_wave_group<P>(i, o, opt, sh[g], flags, int(threadIdx.x) / 64);
    __syncthreads();                       // both waves' parts of the records are in rec_sh
    for (int e = int(threadIdx.x); e < knots * kOutPerKnot; e += int(blockDim.x))
      d.out[kOutPerKnot * k0 + e] = rec_sh[e];
  }
  publish_completion(done, tail, 2u * gridDim.x, lane);
}

}  // namespace rbd
