// Batched uneven trajectories: a device work queue of knot chunks and the
// persistent kernel that drains it.
//
// A batch is a list of trajectories, each with its own input/output buffers
// and length.  The plan kernel turns the lengths into an exclusive prefix sum
// of chunk counts (a chunk is the knots one wave processes per iteration: 64 in
// the one-lane mapping, 4 in the 16-lane-group mapping) and resets the queue
// cursor.  The persistent kernel's waves then take chunk indices from the
// cursor with one atomic add per wave per chunk, find the trajectory by binary
// search in the prefix sum, and run the per-knot solver.  Every wave leaves the
// loop once the cursor passes the chunk total, which is fixed before the
// persistent kernel starts, so the grid always drains.
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
  long total_chunks;              // written by the plan kernel
  unsigned long long cursor;      // queue head, reset by the plan kernel
  int n_traj;                     // written by the host copy with the descriptors
  int chunk_knots;                // knots per chunk
};

constexpr int kPlanThreads = 1024;

// One block: prefix[t] = sum of chunk counts of trajectories < t, prefix[n] = total.
__global__ __launch_bounds__(kPlanThreads) void batch_plan_kernel(const TrajDesc* __restrict__ descs,
                                                                  BatchHeader* __restrict__ hdr,
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
  if (threadIdx.x == kPlanThreads - 1) {
    prefix[n] = partial[kPlanThreads - 1];
    hdr->total_chunks = partial[kPlanThreads - 1];
    hdr->cursor = 0;
  }
}

// Persistent drain of the queue.  LANES = 1: lane i of a wave takes knot
// k0 + i of a 64-knot chunk; LANES = 16: lane group g takes knot k0 + g of a
// 4-knot chunk.  One wave per workgroup.
template <class P, int LANES>
__global__ __launch_bounds__(64) void batch_persistent_kernel(const TrajDesc* __restrict__ descs,
                                                              BatchHeader* __restrict__ hdr,
                                                              const long* __restrict__ prefix, Options opt) {
  const long total = hdr->total_chunks;
  const int n_traj = hdr->n_traj;
  const int chunk = 64 / LANES;
  const int lane = threadIdx.x;
  const int slot = lane / LANES;            // knot slot in the chunk
  const int role = lane % LANES;
  for (;;) {
    unsigned long long c = 0;
    if (lane == 0) c = atomicAdd(&hdr->cursor, 1ull);
    c = __shfl(c, 0, 64);
    if ((long)c >= total) break;
    int lo = 0, hi = n_traj;                 // prefix[lo] <= c < prefix[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= (long)c) lo = mid; else hi = mid;
    }
    const TrajDesc d = descs[lo];
    const long k = ((long)c - prefix[lo]) * chunk + slot;
    if (k < d.n) {
      InAoS i{d.in + kInPerKnot * k};
      OutAoS o{d.out + kOutPerKnot * k};
      if constexpr (LANES == 1) fdgrad_knot<P>(i, o, opt);
      else fdgrad_knot_lanes<P>(i, o, opt, role);
    }
  }
}

}  // namespace rbd
