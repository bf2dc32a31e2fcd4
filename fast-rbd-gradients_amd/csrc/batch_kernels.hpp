// Batched uneven trajectories: a device work queue of knot chunks and the
// persistent kernel that drains it.
//
// A batch is a list of trajectories, each with its own input/output buffers
// and length.  A chunk is the knots one wave processes per iteration: 64 in the
// one-lane mapping, 4 in the 16-lane-group mapping.  Wave w of the persistent
// grid takes chunk w first; after that each wave takes chunk
// gridDim.x + atomicAdd(cursor, 1) - cursor_base, so a batch no larger than the grid
// costs one atomic per busy wave and idle waves none (all waves contending for
// one address was the kernel's main overhead).  Every processed chunk is
// followed by exactly one atomic add, so a launch advances the cursor by
// exactly its chunk total: a caller that keeps the cursor in device memory
// passes the running sum as cursor_base and never resets it, one that uploads
// a zeroed cursor with the header passes 0.  A chunk's trajectory comes from
// the prefix sum of chunk counts:
//   * up to 64 trajectories (SMALL): every wave scans the counts itself, one
//     trajectory per lane, and finds a chunk's trajectory with a ballot and a
//     population count -- no plan kernel, no memory traffic per chunk;
//   * more: the plan kernel writes the prefix sum and the total first, and the
//     waves binary-search it.
// Every wave leaves the loop once its chunk index passes the total, which is
// fixed before the persistent kernel starts, and the cursor only grows, so the
// grid always drains.  The header and descriptors are only read, so they may
// live in device memory or in pinned host memory.
#pragma once
#include <hip/hip_runtime.h>

#include "fdgrad_knot.hpp"

namespace rbd {

struct TrajDesc {
  const float* in;   // n x 21 (AoS)
  float* out;        // n x 105 (AoS)
  long n;            // knots
  long pad;
};

struct BatchHeader {
  int n_traj;                     // written by the host copy with the descriptors
  int chunk_knots;                // knots per chunk
  long total_chunks;              // chunk total, computed by the host
  unsigned long long cursor;      // dynamic chunks handed out, when the caller uploads it zeroed
};

constexpr int kSmallBatch = 64;   // trajectories handled without the plan kernel

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

// Persistent drain of the queue.  LANES = 1: lane i of a wave takes knot
// k0 + i of a 64-knot chunk; LANES = 16: lane group g takes knot k0 + g of a
// 4-knot chunk.  One wave per workgroup.
template <class P, int LANES, bool SMALL>
__global__ __launch_bounds__(64, kWavesPerSimd<P>) void batch_persistent_kernel(
    const TrajDesc* __restrict__ descs, const BatchHeader* __restrict__ hdr, unsigned long long* __restrict__ cursor,
    unsigned long long cursor_base, const long* __restrict__ prefix, Options opt) {
  const int n_traj = hdr->n_traj;
  const int chunk = 64 / LANES;
  const int lane = threadIdx.x;
  const int slot = lane / LANES;            // knot slot in the chunk
  const int role = lane % LANES;
  long total, incl = 0;                     // SMALL: inclusive prefix of chunk counts up to trajectory `lane`
  if constexpr (SMALL) {
    long v = lane < n_traj ? (descs[lane].n + chunk - 1) / chunk : 0;
    for (int off = 1; off < 64; off <<= 1) {
      const long o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    incl = v;
    total = __shfl(v, 63, 64);
  } else {
    total = hdr->total_chunks;
  }
  for (long c = blockIdx.x; c < total;) {
    int t;
    long first;                             // first chunk of trajectory t
    if constexpr (SMALL) {
      // trajectories that end at or before chunk c are the lanes with incl <= c
      t = __popcll(__ballot(incl <= c));
      first = __shfl(incl, t > 0 ? t - 1 : 0, 64);
      if (t == 0) first = 0;
    } else {
      int lo = 0, hi = n_traj;               // prefix[lo] <= c < prefix[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= c) lo = mid; else hi = mid;
      }
      t = lo;
      first = prefix[lo];
    }
    const TrajDesc d = descs[t];
    const long k = (c - first) * chunk + slot;
    if (k < d.n) {
      InAoS i{d.in + kInPerKnot * k};
      OutAoS o{d.out + kOutPerKnot * k};
      if constexpr (LANES == 1) fdgrad_knot<P>(i, o, opt);
      else fdgrad_knot_lanes<P>(i, o, opt, role);
    }
    unsigned long long next = 0;
    if (lane == 0) next = atomicAdd(cursor, 1ull) - cursor_base;
    next = __shfl(next, 0, 64);
    // a cursor out of step with cursor_base (a caller error) ends the loop
    // instead of producing an index outside [0, total)
    c = long(gridDim.x) + long(next < (unsigned long long)total ? next : (unsigned long long)total);
  }
}

}  // namespace rbd
