// Two-wave latency mapping (device code): a workgroup of two waves takes
// kTwoWaveKnots knots, 16 lanes per knot in each wave; wave 0 runs Minv,
// wave 1 the RNEA and the columns (two_wave_* in fdgrad_knot.hpp).  The waves
// hand data over in LDS and signal with one LDS flag per direction: each wave
// sets its flag unconditionally before it waits for the other's, so every
// wait ends.  A wait is also bounded (Options::handoff_spins polls): one that
// gives up writes 1 to the host-visible word Options::fault, and the host
// turns that into an error status (RBD_ERR_DEVICE_TIMEOUT) instead of
// returning the invalid results as good ones.  Used by the flat kernel (one
// knot group per workgroup, rbdgrad.hip) and the batch kernel (knot groups
// pulled in a loop, batch_kernels.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "fdgrad_knot.hpp"

namespace rbd {

constexpr int kTwoWaveKnots = 4;

// The fences cover LDS only: a release that also waited for the wave's
// outstanding global stores would hold the flag back by microseconds.  The
// scheduling barriers keep the flag store and the polling loop where they are
// in the instruction stream.
__device__ inline void lds_signal(int* flag) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  if (threadIdx.x % 64 == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __builtin_amdgcn_sched_barrier(0);
}
__device__ inline void lds_wait(int* flag, const Options& opt) {
  __builtin_amdgcn_sched_barrier(0);
  const bool seen = bounded_wait(
      [flag] { return __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0; },
      [] { __builtin_amdgcn_s_sleep(1); }, opt.handoff_spins);
  // the flag is read from LDS by every lane, so `seen` is the same in all of them
  if (!seen && opt.fault != nullptr && threadIdx.x % 64 == 0)
    __hip_atomic_store(opt.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Materialise a value before a wait: the compiler would otherwise sink work
// whose results are only used after the wait below it, and the two waves
// would run one after the other instead of side by side.
template <typename S>
__device__ inline void pin(S x) { asm volatile("" ::"v"(x)); }

// One knot per 16-lane group of each wave of a pair of waves: i, o are the
// knot of this lane's group (o masked for lanes past the last knot, which
// compute a copy of it); sh is the group's LDS block and flags the pair's two
// flags, both zero on entry; wave is the calling wave's place in the pair (0:
// Minv, 1: the RNEA and the columns).
template <class P, class In, class O>
__device__ RBD_INLINE void two_wave_group(const In& i, O& o, const Options& opt, TwoWaveShared<P>& sh, int* flags,
                                          int wave) {
  const int role = (threadIdx.x % 64) % kLanesPerKnotGroup;
  if (wave == 0) {
    CS<TwoWaveTrig<P>> tw[NQ];
    typename P::Minv Mi[NQ][NQ];
    two_wave_m1a<P>(i, role, tw, sh);
    wave_lds_sync();
    two_wave_m1b<P>(sh, Mi);
    for (int r = 0; r < NQ; ++r) {
      pin(tw[r].c);
      pin(tw[r].s);
      for (int c = r; c < NQ; ++c) pin(Mi[r][c]);
    }
    lds_wait(&flags[0], opt);
    two_wave_m2<P>(role, tw, Mi, sh, o);
    lds_signal(&flags[1]);
  } else {
    TwoWaveR<P> st;
    typename P::Col tau_col[NQ];
    two_wave_r1a<P>(i, opt, role, sh, st);
    lds_signal(&flags[0]);
    two_wave_r1b<P>(opt, role, st, tau_col);
    for (int r = 0; r < NQ; ++r) pin(tau_col[r]);
    lds_wait(&flags[1], opt);
    two_wave_r2<P>(role, sh, tau_col, o);
  }
}

}  // namespace rbd
