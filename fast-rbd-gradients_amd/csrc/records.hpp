// Whole-record output through LDS for the one-lane-per-knot kernels (device
// code).  A wave's 64 consecutive knots have one contiguous block of records
// (64 x 105 floats = 26,880 bytes).  Each lane writes its record into an LDS
// copy of that block (at the record stride of 105 floats, which is odd, so the
// lanes' writes fall in different banks), and the wave then copies the block
// out with consecutive lanes on consecutive 16-byte pieces: every store covers
// 1 KB of contiguous output instead of 64 records' cache lines.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdgrad_knot.hpp"

namespace rbd {

constexpr int kRecordBlockFloats = 64 * kOutPerKnot;   // one wave's records

// A store that the caches treat as streaming (non-temporal) when NT: the
// outputs are written once and not read again by the kernel.
template <bool NT, typename T>
__device__ inline void put(T* dst, T x) {
  if constexpr (NT) __builtin_nontemporal_store(x, dst);
  else *dst = x;
}

// Copy `words` floats from the LDS block to dst (the first record of the
// wave's knots): 16-byte pieces when dst is 16-byte aligned, single floats
// otherwise.  Called by the whole wave after a wave_lds_sync.  NT: streaming
// stores (whole lines here, unlike a lane's own record stores, which leave
// lines partly written); the flat one-lane kernel at 2^20 knots measured 189
// against 201 us with them.
template <bool NT = false>
__device__ inline void copy_record_block(const float* rec, float* dst, int words, int lane) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  if ((reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
    v4f* d4 = reinterpret_cast<v4f*>(dst);
    const v4f* s4 = reinterpret_cast<const v4f*>(rec);
    for (int e = lane; e < words / 4; e += 64) put<NT>(d4 + e, s4[e]);
    for (int e = (words / 4) * 4 + lane; e < words; e += 64) put<NT>(dst + e, rec[e]);
  } else {
    for (int e = lane; e < words; e += 64) put<NT>(dst + e, rec[e]);
  }
}

}  // namespace rbd
