// librbdgrad.so: HIP kernels for gfx950 and the C ABI declared in include/rbdgrad.h.
//
// Mappings of knot points onto lanes, all running the per-knot solver
// (fdgrad_knot.hpp) out of registers; the memory traffic is the knot's 84
// input bytes and its 420 output bytes, plus LDS exchanges inside a knot's
// lane group:
//   * one lane per knot (throughput): a wave computes 64 knots in lock step
//     with uniform control flow (the derivative columns are unrolled at
//     compile time, the same column in every lane);
//   * one knot per group of 4, 7 or 8 lanes (group_mapping.hpp): every lane
//     runs the shared steps and a share of the derivative columns;
//   * one knot per 16-lane group: the shared steps, then one derivative column
//     per lane, so a knot's critical path is the shared steps plus one column
//     instead of all thirteen;
//   * two waves per knot group (lowest latency, two_wave.hpp): Minv on one
//     wave, the RNEA and the columns on the other.
// The context picks the mapping per call from the size (its mapping table,
// measured by rbd_autotune, or the built-in defaults).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "context.hpp"
#include "group_mapping.hpp"
#include "records.hpp"
#include "two_wave.hpp"

namespace rbdi {

namespace {
thread_local std::string g_last_error;
}

rbd_status fail(rbd_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}
rbd_status hip_fail(hipError_t e, const char* what) {
  return fail(RBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace rbdi

using rbdi::kBlock;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
namespace rbd {

// One lane per knot, array-of-structures, records staged in LDS.  A wave takes
// 64 consecutive knots, whose 105-float records are one contiguous block of
// 26,880 bytes: each lane writes its record into LDS at the record's own
// stride (105 floats, odd: no bank conflicts), and the wave copies the block
// out with consecutive lanes on consecutive 16-byte pieces, so each store
// covers 1 KB of contiguous output (8 cache lines) instead of 64 records'
// lines.  Per-lane record stores had the cache and address units stalling the
// waves (at 2^20 knots: 5.2x the L1 accesses and 4.9x the L2 write requests
// of the structure-of-arrays layout).  The block limits a CU to 6 waves.  The
// copy uses streaming (non-temporal) stores.
//
// A wave's vector-memory counter retires loads and stores in the order they
// were issued, so waiting for a block's input loads also waits for every store
// issued before them.  Each loop iteration therefore issues its block's loads
// first and only then copies out the previous block's records (2^20 knots:
// 179.5-181.4 against 183.3-184.9 us).  The copy's LDS reads precede the next
// records' LDS writes in program order, on the same array, and a wave's LDS
// operations execute in order, so no fence separates them; a fence there
// changed how the compiler contracted the knot's products into FMAs, and with
// it the results' last bits.
template <class P>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fdgrad_aos_kernel(long n, const float* __restrict__ in,
                                                            float* __restrict__ out, Options opt) {
  __shared__ __align__(16) float rec[kRecordBlockFloats];
  const int lane = int(threadIdx.x);
  const long stride = long(gridDim.x) * kBlock;
  long prev = 0;          // first knot of the block whose records rec holds
  int prev_words = 0;     // and their words (0: none yet)
  for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
    const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
    // lanes past the last knot compute a copy of it; only nk records are written out
    const float* src = in + kInPerKnot * (k0 + (lane < nk ? lane : nk - 1));
    InRegs i;
    for (int j = 0; j < kInPerKnot; ++j) i.x[j] = src[j];
    __builtin_amdgcn_sched_barrier(0);
    copy_record_block<true>(rec, out + prev * kOutPerKnot, prev_words, lane);
    __builtin_amdgcn_sched_barrier(0);
    OutAoS o{rec + kOutPerKnot * lane};
    fdgrad_knot<P>(i, o, opt);
    wave_lds_sync();
    prev = k0;
    prev_words = nk * kOutPerKnot;
  }
  copy_record_block<true>(rec, out + prev * kOutPerKnot, prev_words, lane);
}

// Structure-of-arrays output of the one-lane kernel with streaming
// (non-temporal) stores: each store writes one output row for the wave's 64
// consecutive knots, 256 contiguous bytes, written once.  At 2^20 knots the
// kernel measured 162 against 186 us with plain stores, whose write-allocated
// lines kept the waves waiting.  The last column computed (d/dqd column 6,
// the record's last 7 rows) is kept in `last` instead of stored: the kernel
// stores it after the next knot's input loads have issued.
struct OutSoAStream {
  float* p; long ld;
  float* last;
  __device__ void st(long row, float x) const { put<true>(p + row * ld, x); }
  __device__ void qdd(int i, float x) const { st(i, x); }
  __device__ void dq(int r, int c, float x) const { st(NQ + NQ * c + r, x); }
  __device__ void dqd(int r, int c, float x) const {
    if (c == NQ - 1) last[r] = x;
    else st(NQ + NQ * NQ + NQ * c + r, x);
  }
};

// One lane per knot, structure of arrays.  As in the AoS kernel, each loop
// iteration issues its knot's input loads before any store: here the previous
// knot's last column, so the wait for the loads covers only stores issued a
// column's work earlier, which have retired.  At 2^20 knots: 156.0-157.0
// against 162.0-162.2 us with the last column stored at the end of its own
// iteration, bitwise the same results.
template <class P>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fdgrad_soa_kernel(long n, const float* __restrict__ in, long ld_in,
                                                            float* __restrict__ out, long ld_out, Options opt) {
  constexpr int kLastRow = kOutPerKnot - NQ;
  const long stride = long(gridDim.x) * blockDim.x;
  float last[NQ];
  long prev = -1;
  for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
    InRegs i;
    for (int j = 0; j < kInPerKnot; ++j) i.x[j] = in[j * ld_in + k];
    __builtin_amdgcn_sched_barrier(0);
    if (prev >= 0)
      for (int r = 0; r < NQ; ++r) put<true>(out + (kLastRow + r) * ld_out + prev, last[r]);
    __builtin_amdgcn_sched_barrier(0);
    OutSoAStream o{out + k, ld_out, last};
    fdgrad_knot<P>(i, o, opt);
    prev = k;
  }
  if (prev >= 0)
    for (int r = 0; r < NQ; ++r) put<true>(out + (kLastRow + r) * ld_out + prev, last[r]);
}

// Reference-compatible mode (qdd and Minv given), one lane per knot: a wave's
// 64 consecutive knots write one contiguous block of 64 x 98 floats, staged in
// LDS and copied out whole as in the AoS kernel above: 215.6 us at 2^20
// knots, against 381 us with per-lane stores of the 98-float records.  As
// there, each iteration issues its block's input loads (the 49 of a record's
// 77 floats the gradient reads: q, qd, qdd and Minv's upper triangle) before
// it copies out the previous block's records: 200 against 216-218 us, with
// the same FP instruction mix and bitwise the same results.
constexpr int kDqddBlockFloats = 64 * kDqddPerKnot;
struct InVolRegs {
  static constexpr int kUpper = NQ * (NQ + 1) / 2;
  float x[3 * NQ + kUpper];
  const float* fx = nullptr;
  // row r of the upper triangle starts at r * NQ - r * (r - 1) / 2
  static constexpr int upper(int r, int c) { return 3 * NQ + r * NQ - r * (r - 1) / 2 + (c - r); }
  __device__ void load(const float* p) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      x[k] = p[k];
      x[NQ + k] = p[NQ + k];
      x[2 * NQ + k] = p[3 * NQ + k];
    }
#pragma unroll
    for (int r = 0; r < NQ; ++r)
#pragma unroll
      for (int c = r; c < NQ; ++c) x[upper(r, c)] = p[4 * NQ + r + NQ * c];
  }
  __device__ float q(int i) const { return x[i]; }
  __device__ float qd(int i) const { return x[NQ + i]; }
  __device__ float qdd(int i) const { return x[2 * NQ + i]; }
  __device__ float minv(int r, int c) const { return x[upper(r, c)]; }   // r <= c: only the upper triangle is read
};

template <class P>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fdgrad_refmode_kernel(long n, const float* __restrict__ vol,
                                                                float* __restrict__ dqdd, Options opt) {
  __shared__ __align__(16) float rec[kDqddBlockFloats];
  const int lane = int(threadIdx.x);
  const long stride = long(gridDim.x) * kBlock;
  long prev = 0;          // first knot of the block whose records rec holds
  int prev_words = 0;     // and their words (0: none yet)
  for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
    const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
    // lanes past the last knot compute a copy of it; only nk records are written out
    InVolRegs i;
    i.load(vol + kVolPerKnot * (k0 + (lane < nk ? lane : nk - 1)));
    __builtin_amdgcn_sched_barrier(0);
    copy_record_block<true>(rec, dqdd + prev * kDqddPerKnot, prev_words, lane);
    __builtin_amdgcn_sched_barrier(0);
    OutDqddAoS o{rec + kDqddPerKnot * lane};
    fdgrad_knot_given_qdd_minv<P>(i, o, opt);
    wave_lds_sync();
    prev = k0;
    prev_words = nk * kDqddPerKnot;
  }
  copy_record_block<true>(rec, dqdd + prev * kDqddPerKnot, prev_words, lane);
}

// One knot per group of kLanesPerKnotGroup lanes (4 knots per wave): the
// low-latency mapping for small batches.  Every lane of a group runs the
// shared steps, then one derivative column; the group's lanes finish together.
// W: waves per SIMD the kernel is compiled for.  W = 1 lets the compiler use
// up to 512 registers (no spills: the latency variant, for calls whose waves
// do not outnumber the SIMDs); W = kWavesPerSimd<P> is the throughput variant.
template <class P, int W>
__global__ __launch_bounds__(kBlock, W) void fdgrad_aos_lanes_kernel(long n, const float* __restrict__ in,
                                                                  float* __restrict__ out, Options opt) {
  const int role = threadIdx.x % kLanesPerKnotGroup;
  __shared__ typename P::Minv mi_sh[kBlock / kLanesPerKnotGroup][kMinvPacked];   // Minv per lane group
  const long groups = long(gridDim.x) * (blockDim.x / kLanesPerKnotGroup);
  for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / kLanesPerKnotGroup; k < n; k += groups) {
    InAoS i{in + kInPerKnot * k};
    OutAoS o{out + kOutPerKnot * k};
    fdgrad_knot_lanes<P>(i, o, opt, role, mi_sh[threadIdx.x / kLanesPerKnotGroup]);
  }
}

template <class P, int W>
__global__ __launch_bounds__(kBlock, W) void fdgrad_soa_lanes_kernel(long n, const float* __restrict__ in, long ld_in,
                                                                  float* __restrict__ out, long ld_out, Options opt) {
  const int role = threadIdx.x % kLanesPerKnotGroup;
  __shared__ typename P::Minv mi_sh[kBlock / kLanesPerKnotGroup][kMinvPacked];   // Minv per lane group
  const long groups = long(gridDim.x) * (blockDim.x / kLanesPerKnotGroup);
  for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / kLanesPerKnotGroup; k < n; k += groups) {
    InSoA i{in + k, ld_in};
    OutSoA o{out + k, ld_out};
    fdgrad_knot_lanes<P>(i, o, opt, role, mi_sh[threadIdx.x / kLanesPerKnotGroup]);
  }
}

template <class P, int W>
__global__ __launch_bounds__(kBlock, W) void fdgrad_refmode_lanes_kernel(long n, const float* __restrict__ vol,
                                                                      float* __restrict__ dqdd, Options opt) {
  const int role = threadIdx.x % kLanesPerKnotGroup;
  const long groups = long(gridDim.x) * (blockDim.x / kLanesPerKnotGroup);
  for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / kLanesPerKnotGroup; k < n; k += groups) {
    InAoS i{vol + kVolPerKnot * k};
    OutDqddAoS o{dqdd + kDqddPerKnot * k};
    fdgrad_knot_given_qdd_minv_lanes<P>(i, o, opt, role);
  }
}

// Two-wave latency mapping: one group of kTwoWaveKnots knots per workgroup
// (two_wave_group, two_wave.hpp).
template <class P, bool SOA>
__global__ __launch_bounds__(2 * kBlock, 1) void fdgrad_two_wave_kernel(long n, const float* __restrict__ in,
                                                                       long ld_in, float* __restrict__ out,
                                                                       long ld_out, Options opt) {
  __shared__ TwoWaveShared<P> sh[kTwoWaveKnots];
  __shared__ int flags[2];                 // [0]: bias ready (R -> M), [1]: Minv and corrections ready (M -> R)
  if (threadIdx.x == 0) { flags[0] = 0; flags[1] = 0; }
  __syncthreads();
  const int g = (threadIdx.x % kBlock) / kLanesPerKnotGroup;
  const long k = long(blockIdx.x) * kTwoWaveKnots + g;
  const bool valid = k < n;
  const long kk = valid ? k : n - 1;
  using In = typename std::conditional<SOA, InSoA, InAoS>::type;
  using O = typename std::conditional<SOA, OutSoA, OutAoS>::type;
  In i;
  O o0;
  if constexpr (SOA) { i = InSoA{in + kk, ld_in}; o0 = OutSoA{out + kk, ld_out}; }
  else { i = InAoS{in + kInPerKnot * kk}; o0 = OutAoS{out + kOutPerKnot * kk}; }
  OutMasked<O> o{o0, valid};
  two_wave_group<P>(i, o, opt, sh[g], flags, int(threadIdx.x) / 64);
}

// L lanes per knot (L = 4, 7, 8; group_mapping.hpp): floor(64 / L) knots per
// wave, the waves striding over the call's knots.  Workgroups of 1, 2 or 4
// waves (blockDim.x = 64 x waves), each wave on its own: its own knots and its
// own LDS block.  Lanes of knots past the end compute a copy of the last knot
// and store nothing.
template <class P, int L, bool SOA>
__global__ __launch_bounds__(kMaxGroupWaves * kBlock, 1) void fdgrad_group_kernel(long n, const float* __restrict__ in,
                                                                                 long ld_in, float* __restrict__ out,
                                                                                 long ld_out, Options opt) {
  using G = Group<L>;
  // Minv per knot (L >= 7), per wave of the workgroup
  __shared__ typename P::Minv mi_sh[kMaxGroupWaves][G::kMinvColumns ? G::kSlots : 1][kMinvPacked];
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kBlock);   // the same in every lane
  const int lane = int(threadIdx.x) % kBlock;
  const int slot = group_slot<L>(lane);
  const long waves_per_group = long(blockDim.x) / kBlock;
  using In = typename std::conditional<SOA, InSoA, InAoS>::type;
  using O = typename std::conditional<SOA, OutSoA, OutAoS>::type;
  for (long w = long(blockIdx.x) * waves_per_group + wave; w * G::kSlots < n; w += long(gridDim.x) * waves_per_group) {
    const long k = w * G::kSlots + slot;
    const long kk = k < n ? k : n - 1;
    In i;
    O o0;
    if constexpr (SOA) { i = InSoA{in + kk, ld_in}; o0 = OutSoA{out + kk, ld_out}; }
    else { i = InAoS{in + kInPerKnot * kk}; o0 = OutAoS{out + kOutPerKnot * kk}; }
    OutMasked<O> o{o0, k < n && lane < G::kLanesUsed};
    fdgrad_knot_group<P, L>(i, o, opt, lane, mi_sh[wave][G::kMinvColumns ? slot : 0]);
    wave_lds_sync();                        // every lane is done with mi_sh before the next knots
  }
}

// Full path plus dqdd/du = Minv (49 per knot, dense), both mappings.  WAFR:
// out receives the WAFR baseline's 147-float layout (OutWafr) and minv is the
// optional qdd buffer.
template <class P, int LANES, bool WAFR = false>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fdgrad_minv_kernel(long n, const float* __restrict__ in,
                                                             float* __restrict__ out, float* __restrict__ minv,
                                                             Options opt) {
  using O = typename std::conditional<WAFR, OutWafr, OutAoSMinv>::type;
  if constexpr (LANES == 1) {
    // the wave's 64 consecutive knots: both outputs staged in LDS, each copied
    // out as one contiguous block (as fdgrad_aos_kernel does), by one round of
    // resident waves: at 2^20 knots 252 against 385 us (+ Minv) and 252
    // against 444 us (WAFR layout) with per-lane stores and a wave per block.
    // The two blocks (39,424 bytes) leave a CU 4 waves.
    constexpr int kRec = WAFR ? kWafrPerKnot : kOutPerKnot;
    constexpr int kAux = WAFR ? NQ : NQ * NQ;     // qdd (WAFR, optional) or Minv per knot
    __shared__ __align__(16) float rec[64 * kRec];
    __shared__ __align__(16) float aux[64 * kAux];
    // Each iteration issues its block's input loads before it copies out the
    // previous block's records, as fdgrad_aos_kernel does: 236 against 252 us
    // for either output, with bitwise the same results.
    const int lane = int(threadIdx.x);
    const long stride = long(gridDim.x) * kBlock;
    long prev = 0;          // first knot of the block whose records rec, aux hold
    int prev_nk = 0;        // and their knots (0: none yet)
    for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
      const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
      // lanes past the last knot compute a copy of it; only nk knots are written out
      const float* src = in + kInPerKnot * (k0 + (lane < nk ? lane : nk - 1));
      InRegs i;
      for (int j = 0; j < kInPerKnot; ++j) i.x[j] = src[j];
      __builtin_amdgcn_sched_barrier(0);
      copy_record_block<true>(rec, out + prev * kRec, prev_nk * kRec, lane);
      if (minv) copy_record_block<true>(aux, minv + prev * kAux, prev_nk * kAux, lane);
      __builtin_amdgcn_sched_barrier(0);
      O o;
      if constexpr (WAFR) o = OutWafr{minv ? aux + kAux * lane : nullptr, rec + kRec * lane};
      else o = OutAoSMinv{rec + kRec * lane, aux + kAux * lane};
      fdgrad_knot<P>(i, o, opt);
      wave_lds_sync();
      prev = k0;
      prev_nk = nk;
    }
    copy_record_block<true>(rec, out + prev * kRec, prev_nk * kRec, lane);
    if (minv) copy_record_block<true>(aux, minv + prev * kAux, prev_nk * kAux, lane);
  } else {
    const int role = threadIdx.x % LANES;
    __shared__ typename P::Minv mi_sh[kBlock / kLanesPerKnotGroup][kMinvPacked];   // Minv per lane group
    const long groups = long(gridDim.x) * (blockDim.x / LANES);
    for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / LANES; k < n; k += groups) {
      InAoS i{in + kInPerKnot * k};
      O o;
      if constexpr (WAFR) o = OutWafr{minv ? minv + NQ * k : nullptr, out + kWafrPerKnot * k};
      else o = OutAoSMinv{out + kOutPerKnot * k, minv + NQ * NQ * k};
      fdgrad_knot_lanes<P>(i, o, opt, role, mi_sh[threadIdx.x / kLanesPerKnotGroup]);
    }
  }
}

// Full path with external link forces: knot k reads its 42 forces at
// fext + k * fext_stride (stride 0: one set for every knot).  Both mappings.
template <class P, int LANES>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fdgrad_fext_kernel(long n, const float* __restrict__ in,
                                                             const float* __restrict__ fext, long fext_stride,
                                                             float* __restrict__ out, Options opt) {
  if constexpr (LANES == 1) {
    // records staged in LDS and copied out whole, as fdgrad_aos_kernel does
    // by one round of resident waves (2^20 knots: 181 against 273 us with
    // per-lane stores and a wave per block)
    // (q, qd, u issued before the previous block's copy, as fdgrad_aos_kernel
    // does; the forces are read where the RNEA needs them): 186 against 191 us
    // with one set of forces per knot, bitwise the same results
    __shared__ __align__(16) float rec[kRecordBlockFloats];
    const int lane = int(threadIdx.x);
    const long stride = long(gridDim.x) * kBlock;
    long prev = 0;          // first knot of the block whose records rec holds
    int prev_words = 0;     // and their words (0: none yet)
    for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
      const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
      const long k = k0 + (lane < nk ? lane : nk - 1);
      const float* src = in + kInPerKnot * k;
      InRegs i;
      for (int j = 0; j < kInPerKnot; ++j) i.x[j] = src[j];
      i.fx = fext + fext_stride * k;
      __builtin_amdgcn_sched_barrier(0);
      copy_record_block<true>(rec, out + prev * kOutPerKnot, prev_words, lane);
      __builtin_amdgcn_sched_barrier(0);
      OutAoS o{rec + kOutPerKnot * lane};
      fdgrad_knot<P>(i, o, opt);
      wave_lds_sync();
      prev = k0;
      prev_words = nk * kOutPerKnot;
    }
    copy_record_block<true>(rec, out + prev * kOutPerKnot, prev_words, lane);
  } else {
    const int role = threadIdx.x % LANES;
    __shared__ typename P::Minv mi_sh[kBlock / kLanesPerKnotGroup][kMinvPacked];   // Minv per lane group
    const long groups = long(gridDim.x) * (blockDim.x / LANES);
    for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / LANES; k < n; k += groups) {
      InAoS i{in + kInPerKnot * k, fext + fext_stride * k};
      OutAoS o{out + kOutPerKnot * k};
      fdgrad_knot_lanes<P>(i, o, opt, role, mi_sh[threadIdx.x / kLanesPerKnotGroup]);
    }
  }
}

// Forward dynamics only (qdd and dense Minv), one lane per knot; FEXT: with
// external link forces as above.
template <class P, bool FEXT>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void fd_kernel(long n, const float* __restrict__ in,
                                                    const float* __restrict__ fext, long fext_stride,
                                                    float* __restrict__ qdd, float* __restrict__ minv, Options opt) {
  // qdd and Minv of the wave's 64 consecutive knots staged in LDS and copied
  // out as two contiguous blocks (2^20 knots: 63 against 119 us with per-lane
  // stores)
  __shared__ __align__(16) float qrec[64 * NQ];
  __shared__ __align__(16) float mrec[64 * NQ * NQ];
  const int lane = int(threadIdx.x);
  const long stride = long(gridDim.x) * kBlock;
  for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
    const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
    const long k = k0 + (lane < nk ? lane : nk - 1);
    InAoS i{in + kInPerKnot * k};
    if constexpr (FEXT) i.fx = fext + fext_stride * k;
    OutFD o{qrec + NQ * lane, mrec + NQ * NQ * lane};
    fd_knot<P>(i, o, opt);
    wave_lds_sync();
    copy_record_block<true>(qrec, qdd + NQ * k0, nk * NQ, lane);
    copy_record_block<true>(mrec, minv + NQ * NQ * k0, nk * NQ * NQ, lane);
    wave_lds_sync();                        // read out before the next knots overwrite them
  }
}

// Inverse-dynamics gradient (cdu, 98 per knot) from q, qd, qdd (SPLIT = false:
// 28 per knot, the fused layout) or from q, qd and precomputed v, a, f
// (SPLIT = true: 154 per knot, the split layout).  Both mappings.
template <class P, int LANES, bool SPLIT>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void id_gradient_kernel(long n, const float* __restrict__ vol,
                                                             float* __restrict__ cdu, Options opt) {
  if constexpr (LANES == 1 && SPLIT && std::is_same<P, PrecFP32>::value) {
    // FP32, split layout: the wave's 64 consecutive input records (154 floats
    // each, one contiguous block) are read into LDS as whole lines, 16 bytes
    // per lane, and each lane then reads its record from there.  Read by each
    // lane at the 616-byte record stride, they were fetched 5x over.
    typedef float v4f __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) float inrec[64 * kSplitPerKnot];
    const int lane = int(threadIdx.x);
    const long stride = long(gridDim.x) * kBlock;
    for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
      const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
      const float* src = vol + k0 * kSplitPerKnot;
      const int words = nk * kSplitPerKnot;
      if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
        for (int e = lane; e < words / 4; e += 64)
          reinterpret_cast<v4f*>(inrec)[e] = reinterpret_cast<const v4f*>(src)[e];
        for (int e = (words / 4) * 4 + lane; e < words; e += 64) inrec[e] = src[e];
      } else {
        for (int e = lane; e < words; e += 64) inrec[e] = src[e];
      }
      wave_lds_sync();
      if (lane < nk) {
        InSplit i{inrec + kSplitPerKnot * lane};
        OutDqddAoS o{cdu + kDqddPerKnot * (k0 + lane)};
        id_gradient_split_knot<P>(i, o, opt);
      }
      wave_lds_sync();                      // every lane has read its record before the next block overwrites inrec
    }
    return;
  }
  const int role = threadIdx.x % LANES;
  const long groups = long(gridDim.x) * (blockDim.x / LANES);
  for (long k = (long(blockIdx.x) * blockDim.x + threadIdx.x) / LANES; k < n; k += groups) {
    OutDqddAoS o{cdu + kDqddPerKnot * k};
    if constexpr (SPLIT) {
      InSplit i{vol + kSplitPerKnot * k};
      if constexpr (LANES == 1) id_gradient_split_knot<P>(i, o, opt);
      else id_gradient_split_knot_lanes<P>(i, o, opt, role);
    } else {
      InAoS i{vol + kFusedPerKnot * k};
      if constexpr (LANES == 1) id_gradient_knot<P>(i, o, opt);
      else id_gradient_knot_lanes<P>(i, o, opt, role);
    }
  }
}

// v, a, f of every knot of a split-layout buffer, written in place after its
// q, qd, u, qdd.  One lane per knot.  In FP32 the wave's 64 consecutive
// records (154 floats each, one contiguous block) are assembled whole in LDS,
// their first 28 floats copied from the input, and written back as one block
// of whole lines: storing only floats 28..153 of each record from each lane
// left every line partly written, which made the memory side fetch and merge
// it (3.3x the algorithmic bytes at 2^20 knots).
template <class P>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<P>) void vaf_kernel(long n, float* vol, Options opt) {
  if constexpr (std::is_same<P, PrecFP32>::value) {
    __shared__ __align__(16) float rec[64 * kSplitPerKnot];
    const int lane = int(threadIdx.x);
    const long stride = long(gridDim.x) * kBlock;
    for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
      const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
      // lanes past the last knot compute a copy of it; only nk records are written out
      const float* src = vol + kSplitPerKnot * (k0 + (lane < nk ? lane : nk - 1));
      float* r = rec + kSplitPerKnot * lane;
      for (int j = 0; j < 4 * NQ; ++j) r[j] = src[j];
      InAoS i{src};
      OutVaf o{r};
      vaf_knot<P>(i, o, opt);
      wave_lds_sync();
      copy_record_block<true>(rec, vol + k0 * kSplitPerKnot, nk * kSplitPerKnot, lane);
      wave_lds_sync();                      // read out before the next knots overwrite rec
    }
  } else {
    const long stride = long(gridDim.x) * blockDim.x;
    for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
      InAoS i{vol + kSplitPerKnot * k};
      OutVaf o{vol + kSplitPerKnot * k};
      vaf_knot<P>(i, o, opt);
    }
  }
}

// dqdd = -Minv cdu per knot.  One lane per knot; the wave's 64 consecutive
// knots' results are staged in LDS and copied out as one contiguous block of
// 64 x 98 floats, as fdgrad_refmode_kernel does.
__global__ __launch_bounds__(kBlock) void finish_kernel(long n, const float* __restrict__ minv,
                                                        const float* __restrict__ cdu, float* __restrict__ dqdd) {
  __shared__ __align__(16) float rec[kDqddBlockFloats];
  const int lane = int(threadIdx.x);
  const long stride = long(gridDim.x) * kBlock;
  for (long k0 = long(blockIdx.x) * kBlock; k0 < n; k0 += stride) {
    const int nk = int(n - k0 < kBlock ? n - k0 : kBlock);
    // lanes past the last knot compute a copy of it; only nk records are written out
    const long k = k0 + (lane < nk ? lane : nk - 1);
    InFinish i{minv + NQ * NQ * k, cdu + kDqddPerKnot * k};
    OutFinish o{rec + kDqddPerKnot * lane};
    finish_knot(i, o);
    wave_lds_sync();
    copy_record_block<true>(rec, dqdd + k0 * kDqddPerKnot, nk * kDqddPerKnot, lane);
    wave_lds_sync();                        // read out before the next knots overwrite rec
  }
}

}  // namespace rbd

namespace rbdi {

// grid for n knots at `lanes` lanes per knot (and, for the L-lane groups, wpg
// waves per workgroup), capped so huge calls grid-stride
unsigned grid_for(int64_t n, int lanes, int wpg = 1) {
  const int64_t knots_per_block = lanes == 1 ? kBlock : (kBlock / lanes) * (is_group_lanes(lanes) ? wpg : 1);
  const int64_t blocks = (n + knots_per_block - 1) / knots_per_block;
  const int64_t cap = int64_t(1) << 20;
  return unsigned(blocks < cap ? blocks : cap);
}

template <class P, int W>
void launch_lanes(const FlatCall& c, dim3 grid, dim3 block) {
  const long n = long(c.n);
  switch (c.layout) {
    case Layout::AoS:
      hipLaunchKernelGGL((rbd::fdgrad_aos_lanes_kernel<P, W>), grid, block, 0, c.stream, n, c.in, c.out, c.opt);
      break;
    case Layout::SoA:
      hipLaunchKernelGGL((rbd::fdgrad_soa_lanes_kernel<P, W>), grid, block, 0, c.stream, n, c.in, long(c.ld_in),
                         c.out, long(c.ld_out), c.opt);
      break;
    case Layout::Ref:
      hipLaunchKernelGGL((rbd::fdgrad_refmode_lanes_kernel<P, W>), grid, block, 0, c.stream, n, c.in, c.out, c.opt);
      break;
  }
}

// Workgroups (one wave each) of a one-lane kernel that a CU keeps resident,
// asked of the runtime once per kernel: 8 for the FP32-column SoA kernel (two
// waves per SIMD), 6 for the array-of-structures one (its LDS record block),
// 4 for the kernels that stage two output blocks.
template <auto Kernel>
hipError_t resident_per_cu(int* per_cu) {
  static std::atomic<int> cached{0};
  int v = cached.load(std::memory_order_relaxed);
  if (v <= 0) {
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, Kernel, kBlock, 0);
    if (e != hipSuccess) return e;
    cached.store(v, std::memory_order_relaxed);
  }
  *per_cu = v;
  return hipSuccess;
}

// A one-lane launch of more 64-knot blocks than the device keeps resident
// runs one round of resident waves, each looping over blocks w, w + G, ...
// instead of dispatching a new wave per block: at 2^20 knots 157 against 160
// us (structure of arrays) and 176 against 183 us (array of structures, whose
// last round of dispatched waves left SIMDs idle).  cus: the device's compute
// units (0: unknown, the grid stays as it is).
template <auto Kernel>
hipError_t one_lane_grid(int cus, dim3* grid) {
  if (cus <= 0) return hipSuccess;
  int per_cu = 0;
  if (hipError_t e = resident_per_cu<Kernel>(&per_cu)) return e;
  const unsigned resident = unsigned(per_cu) * unsigned(cus);
  if (resident > 0 && grid->x > resident) grid->x = resident;
  return hipSuccess;
}

template <class P>
hipError_t launch_flat_p(const FlatCall& c) {
  dim3 grid(grid_for(c.n, c.lanes)), block(kBlock);
  const long n = long(c.n);
  if (c.lanes == 1) {
    hipError_t e = hipSuccess;
    switch (c.layout) {
      case Layout::AoS: e = one_lane_grid<rbd::fdgrad_aos_kernel<P>>(c.simds / 4, &grid); break;   // 4 SIMDs per CU
      case Layout::SoA: e = one_lane_grid<rbd::fdgrad_soa_kernel<P>>(c.simds / 4, &grid); break;
      case Layout::Ref: e = one_lane_grid<rbd::fdgrad_refmode_kernel<P>>(c.simds / 4, &grid); break;
    }
    if (e != hipSuccess) return e;
    switch (c.layout) {
      case Layout::AoS:
        hipLaunchKernelGGL(rbd::fdgrad_aos_kernel<P>, grid, block, 0, c.stream, n, c.in, c.out, c.opt);
        break;
      case Layout::SoA:
        hipLaunchKernelGGL(rbd::fdgrad_soa_kernel<P>, grid, block, 0, c.stream, n, c.in, long(c.ld_in), c.out,
                           long(c.ld_out), c.opt);
        break;
      case Layout::Ref:
        hipLaunchKernelGGL(rbd::fdgrad_refmode_kernel<P>, grid, block, 0, c.stream, n, c.in, c.out, c.opt);
        break;
    }
  } else if (is_group_lanes(c.lanes) && c.layout != Layout::Ref) {
    // L lanes per knot: floor(64 / L) knots per wave, c.wpg waves per workgroup
    const int wpg = valid_wpg(c.wpg) ? c.wpg : 1;
    const dim3 ggrid(grid_for(c.n, c.lanes, wpg)), gblock(kBlock * wpg);
    auto go = [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      if (c.layout == Layout::AoS)
        hipLaunchKernelGGL((rbd::fdgrad_group_kernel<P, L, false>), ggrid, gblock, 0, c.stream, n, c.in, 0L, c.out,
                           0L, c.opt);
      else
        hipLaunchKernelGGL((rbd::fdgrad_group_kernel<P, L, true>), ggrid, gblock, 0, c.stream, n, c.in,
                           long(c.ld_in), c.out, long(c.ld_out), c.opt);
    };
    if (c.lanes == 4) go(rbd::IC<4>{});
    else if (c.lanes == 7) go(rbd::IC<7>{});
    else go(rbd::IC<8>{});
  } else if (c.lanes == kTwoWaveLanes && c.layout != Layout::Ref) {
    {
      // latency: Minv and the RNEA + columns on two waves per knot
      const int64_t two_wave_blocks = (c.n + rbd::kTwoWaveKnots - 1) / rbd::kTwoWaveKnots;
      const dim3 g2{unsigned(two_wave_blocks), 1, 1}, b2{2 * kBlock, 1, 1};
      if (c.layout == Layout::AoS)
        hipLaunchKernelGGL((rbd::fdgrad_two_wave_kernel<P, false>), g2, b2, 0, c.stream, n, c.in, 0L, c.out, 0L,
                           c.opt);
      else
        hipLaunchKernelGGL((rbd::fdgrad_two_wave_kernel<P, true>), g2, b2, 0, c.stream, n, c.in, long(c.ld_in),
                           c.out, long(c.ld_out), c.opt);
    }
  } else {
    const int64_t waves = (c.n * rbd::kLanesPerKnotGroup + 63) / 64;
    if (c.simds > 0 && waves <= c.simds) launch_lanes<P, 1>(c, grid, block);
    else launch_lanes<P, rbd::kWavesPerSimd<P>>(c, grid, block);
  }
  return hipGetLastError();
}

hipError_t launch_flat(int precision, const FlatCall& c) {
  switch (precision) {
    case RBD_PREC_FP32: return launch_flat_p<rbd::PrecFP32>(c);
    case RBD_PREC_FP32_HW_TRIG: return launch_flat_p<rbd::PrecFP32HwTrig>(c);
    case RBD_PREC_FP64_PIVOT: return launch_flat_p<rbd::PrecFP64Pivot>(c);
    case RBD_PREC_FP64_COLUMNS: return launch_flat_p<rbd::PrecFP64Col>(c);
    default: return launch_flat_p<rbd::PrecFP64>(c);
  }
}

}  // namespace rbdi

using namespace rbdi;


// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int rbd_abi_version(void) { return 3; }

const char* rbd_model_name(void) { return rbd::kModelName; }

void rbd_config_default(rbd_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->flags = 0;
  cfg->precision = RBD_PREC_FP32;
}

const char* rbd_status_string(rbd_status s) {
  switch (s) {
    case RBD_OK: return "ok";
    case RBD_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RBD_ERR_HIP: return "HIP runtime error";
    case RBD_ERR_NO_DEVICE: return "no HIP device";
    case RBD_ERR_OUT_OF_MEMORY: return "out of device memory";
    case RBD_ERR_UNSUPPORTED: return "unsupported";
    case RBD_ERR_CAPACITY: return "capacity exceeded";
    case RBD_ERR_DEVICE_TIMEOUT: return "device wait timed out";
  }
  return "unknown status";
}

const char* rbd_last_error(void) { return rbdi::g_last_error.c_str(); }

rbd_status rbd_create(const rbd_config* cfg, rbd_context** out) {
  if (!out) return fail(RBD_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  rbd_config c;
  if (cfg) c = *cfg; else rbd_config_default(&c);
  if (c.precision < RBD_PREC_FP32 || c.precision > RBD_PREC_FP32_HW_TRIG)
    return fail(RBD_ERR_INVALID_ARGUMENT, "unknown precision mode");
  if (c.flags & ~(RBD_FLAG_VEL_DAMPING | RBD_FLAG_NO_GRAVITY))
    return fail(RBD_ERR_INVALID_ARGUMENT, "unknown flag bits");
  for (int r : c.reserved)
    if (r != 0) return fail(RBD_ERR_INVALID_ARGUMENT, "reserved config fields must be zero");
  if (c.lanes_per_knot != 0 && c.lanes_per_knot != 1 && c.lanes_per_knot != rbd::kLanesPerKnotGroup &&
      c.lanes_per_knot != kTwoWaveLanes && !is_group_lanes(c.lanes_per_knot))
    return fail(RBD_ERR_INVALID_ARGUMENT, "lanes_per_knot must be 0, 1, 4, 7, 8, 16 or 32");
  if (c.waves_per_group != 0 && !valid_wpg(c.waves_per_group))
    return fail(RBD_ERR_INVALID_ARGUMENT, "waves_per_group must be 0, 1, 2 or 4");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return fail(RBD_ERR_NO_DEVICE, "no HIP device visible");
  if (c.device < 0 || c.device >= ndev) return fail(RBD_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  rbd_context* ctx = new (std::nothrow) rbd_context();
  if (!ctx) return fail(RBD_ERR_OUT_OF_MEMORY, "host allocation failed");
  if (hipDeviceGetAttribute(&ctx->compute_units, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess ||
      ctx->compute_units <= 0)
    ctx->compute_units = 256;
  ctx->device = c.device;
  ctx->flags = c.flags;
  ctx->precision = c.precision;
  ctx->lanes_per_knot = c.lanes_per_knot;
  ctx->waves_per_group = c.waves_per_group;
  ctx->two_wave_threshold.store(int64_t(16) * ctx->compute_units);
  set_default_tables(ctx);
  ctx->opt.damping = (c.flags & RBD_FLAG_VEL_DAMPING) ? 0.5f : 0.0f;
  ctx->opt.gravity = (c.flags & RBD_FLAG_NO_GRAVITY) ? 0.0f : float(rbd::kGravity);
  // the handoff-failure word: pinned, coherent host memory that the kernels
  // write through its device address
  int cur = -1;
  (void)hipGetDevice(&cur);
  e = hipSetDevice(c.device);
  void* word = nullptr;
  if (e == hipSuccess)
    e = hipHostMalloc(&word, 64, hipHostMallocMapped | hipHostMallocCoherent | hipHostMallocPortable);
  void* dword = nullptr;
  if (e == hipSuccess) e = hipHostGetDevicePointer(&dword, word, 0);
  // the no-allocation batched call's scratch on this device (once per device)
  const rbd_status ns = e == hipSuccess ? noalloc_setup(c.device) : RBD_OK;
  if (cur >= 0 && cur != c.device) (void)hipSetDevice(cur);
  if (e != hipSuccess || ns != RBD_OK) {
    if (word) (void)hipHostFree(word);
    delete ctx;
    return ns != RBD_OK ? ns : hip_fail(e, "rbd_create: handoff-failure word");
  }
  *static_cast<volatile unsigned*>(word) = 0u;
  ctx->fault_host = static_cast<volatile unsigned*>(word);
  ctx->opt.fault = static_cast<unsigned*>(dword);
  *out = ctx;
  return RBD_OK;
}

rbd_status rbd_destroy(rbd_context* ctx) {
  if (!ctx) return RBD_OK;
  if (ctx->d_scratch || ctx->host_stream) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(ctx->device);
    if (ctx->host_stream) {
      (void)hipStreamSynchronize(ctx->host_stream);
      (void)hipStreamDestroy(ctx->host_stream);
    }
    if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
    if (cur >= 0) (void)hipSetDevice(cur);
  }
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  if (ctx->fault_host) (void)hipHostFree(const_cast<unsigned*>(ctx->fault_host));
  delete ctx;
  return RBD_OK;
}

rbd_status rbd_context_status(rbd_context* ctx) {
  if (rbd_status s = check_ctx(ctx)) return s;
  return device_status(ctx);
}

rbd_status rbd_context_clear_status(rbd_context* ctx) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (ctx->fault_host) *ctx->fault_host = 0u;
  return RBD_OK;
}

rbd_status rbd_set_handoff_spins(rbd_context* ctx, int spins) {
  if (rbd_status s = check_ctx(ctx)) return s;
  ctx->opt.handoff_spins = spins;
  return RBD_OK;
}

rbd_status rbd_fdgrad_traj_f32(rbd_context* ctx, int64_t n, const float* d_in, float* d_out, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const Shape sh = shape_for(ctx, n, true);
  FlatCall c{Layout::AoS, sh.lanes, n, d_in, 0, d_out, 0, ctx->opt, (hipStream_t)stream, 4 * ctx->compute_units};
  c.wpg = sh.wpg;
  hipError_t e = launch_flat(ctx->precision, c);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad kernel launch");
}

rbd_status rbd_fdgrad_traj_soa_f32(rbd_context* ctx, int64_t n, const float* d_in, int64_t ld_in, float* d_out,
                                   int64_t ld_out, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (ld_in < n || ld_out < n) return fail(RBD_ERR_INVALID_ARGUMENT, "leading dimension smaller than knot count");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const Shape sh = shape_for(ctx, n, true);
  FlatCall c{Layout::SoA, sh.lanes, n, d_in, ld_in, d_out, ld_out, ctx->opt, (hipStream_t)stream,
             4 * ctx->compute_units};
  c.wpg = sh.wpg;
  hipError_t e = launch_flat(ctx->precision, c);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad kernel launch");
}

rbd_status rbd_dynamics_gradient_kernel_f32(rbd_context* ctx, int64_t n, float* d_dqdd, const float* d_mem_vol,
                                            const float* d_mem_const, rbd_stream_t stream) {
  (void)d_mem_const;  // the constants are compiled into the kernel
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_dqdd || !d_mem_vol) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const FlatCall c{Layout::Ref, lanes_for(ctx, n), n, d_mem_vol, 0, d_dqdd, 0, ctx->opt,
                   (hipStream_t)stream, 4 * ctx->compute_units};
  hipError_t e = launch_flat(ctx->precision, c);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad kernel launch");
}

rbd_status rbd_fdgrad_traj_host_f32(rbd_context* ctx, int64_t n, const float* h_in, float* h_out) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!h_in || !h_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  std::lock_guard<std::mutex> lock(ctx->host_mu);
  const size_t in_bytes = size_t(n) * rbd::kInPerKnot * sizeof(float);
  const size_t out_bytes = size_t(n) * rbd::kOutPerKnot * sizeof(float);
  hipError_t e = hipSuccess;
  if (!ctx->host_stream) {
    e = hipStreamCreateWithFlags(&ctx->host_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { ctx->host_stream = nullptr; return hip_fail(e, "hipStreamCreate (host call)"); }
  }
  const Shape sh = shape_for(ctx, n, true);
  if (n <= kHostZeroCopyKnots) {
    // the kernel reads the inputs from and writes the outputs to pinned,
    // mapped host memory: one launch and one wait, no copy engine
    if (ctx->stage_bytes < in_bytes + out_bytes) {
      if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
      ctx->h_stage = ctx->dh_stage = nullptr;
      ctx->stage_bytes = 0;
      const size_t bytes = size_t(kHostZeroCopyKnots) * (rbd::kInPerKnot + rbd::kOutPerKnot) * sizeof(float);
      void* h = nullptr;
      e = hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent);
      if (e != hipSuccess) return fail(RBD_ERR_OUT_OF_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
      void* d = nullptr;
      e = hipHostGetDevicePointer(&d, h, 0);
      if (e != hipSuccess) { (void)hipHostFree(h); return hip_fail(e, "hipHostGetDevicePointer (host call)"); }
      ctx->h_stage = static_cast<char*>(h);
      ctx->dh_stage = static_cast<char*>(d);
      ctx->stage_bytes = bytes;
    }
    std::memcpy(ctx->h_stage, h_in, in_bytes);
    const float* d_in = reinterpret_cast<const float*>(ctx->dh_stage);
    float* d_out = reinterpret_cast<float*>(ctx->dh_stage + in_bytes);
    FlatCall c{Layout::AoS, sh.lanes, n, d_in, 0, d_out, 0, ctx->opt, ctx->host_stream, 4 * ctx->compute_units};
    c.wpg = sh.wpg;
    e = launch_flat(ctx->precision, c);
    if (e != hipSuccess) return hip_fail(e, "fdgrad kernel launch");
    e = hipStreamSynchronize(ctx->host_stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize (host call)");
    std::memcpy(h_out, ctx->h_stage + in_bytes, out_bytes);
    return RBD_OK;
  }
  if (ctx->scratch_bytes < in_bytes + out_bytes) {
    if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    e = hipMalloc(&ctx->d_scratch, in_bytes + out_bytes);
    if (e != hipSuccess) return fail(RBD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    ctx->scratch_bytes = in_bytes + out_bytes;
  }
  float* d_in = ctx->d_scratch;
  float* d_out = ctx->d_scratch + size_t(n) * rbd::kInPerKnot;
  e = hipMemcpyAsync(d_in, h_in, in_bytes, hipMemcpyHostToDevice, ctx->host_stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
  FlatCall c{Layout::AoS, sh.lanes, n, d_in, 0, d_out, 0, ctx->opt, ctx->host_stream, 4 * ctx->compute_units};
  c.wpg = sh.wpg;
  e = launch_flat(ctx->precision, c);
  if (e != hipSuccess) return hip_fail(e, "fdgrad kernel launch");
  e = hipMemcpyAsync(h_out, d_out, out_bytes, hipMemcpyDeviceToHost, ctx->host_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->host_stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy D2H");
  return RBD_OK;
}

const char* rbd_kernel_name(rbd_context* ctx, int reference_mode, int64_t n_knots) {
  if (!ctx) return "";
  static const char* prec[5] = {"PrecFP32", "PrecFP64Pivot", "PrecFP64Col", "PrecFP64", "PrecFP32HwTrig"};
  thread_local std::string name;
  const int lanes = lanes_for(ctx, n_knots, !reference_mode);
  const char* base = reference_mode ? (lanes == 1 ? "fdgrad_refmode_kernel" : "fdgrad_refmode_lanes_kernel")
                     : lanes == 1 ? "fdgrad_aos_kernel"
                     : lanes == kTwoWaveLanes ? "fdgrad_two_wave_kernel"
                     : is_group_lanes(lanes) ? "fdgrad_group_kernel" : "fdgrad_aos_lanes_kernel";
  name = std::string(base) + "<" + prec[ctx->precision] +
         (is_group_lanes(lanes) && !reference_mode ? ", " + std::to_string(lanes) : std::string()) + ">";
  return name.c_str();
}

extern "C++" {
namespace {

// the Minv-output kernels of both layouts (WAFR: out is the 147-float layout,
// minv the optional qdd buffer)
template <bool WAFR>
rbd_status launch_minv(rbd_context* ctx, int64_t n, const float* d_in, float* d_out, float* d_minv,
                       rbd_stream_t stream) {
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const int lanes = lanes_for(ctx, n);
  const dim3 grid(grid_for(n, lanes)), block(kBlock);
  const hipStream_t st = (hipStream_t)stream;
  hipError_t ge = hipSuccess;
  auto go = [&](auto prec_tag) {
    using P = decltype(prec_tag);
    if (lanes == 1) {
      dim3 g = grid;
      ge = one_lane_grid<rbd::fdgrad_minv_kernel<P, 1, WAFR>>(ctx->compute_units, &g);
      if (ge == hipSuccess)
        hipLaunchKernelGGL((rbd::fdgrad_minv_kernel<P, 1, WAFR>), g, block, 0, st, long(n), d_in, d_out, d_minv,
                           ctx->opt);
    } else
      hipLaunchKernelGGL((rbd::fdgrad_minv_kernel<P, rbd::kLanesPerKnotGroup, WAFR>), grid, block, 0, st, long(n),
                         d_in, d_out, d_minv, ctx->opt);
  };
  switch (ctx->precision) {
    case RBD_PREC_FP32: go(rbd::PrecFP32{}); break;
    case RBD_PREC_FP32_HW_TRIG: go(rbd::PrecFP32HwTrig{}); break;
    case RBD_PREC_FP64_PIVOT: go(rbd::PrecFP64Pivot{}); break;
    case RBD_PREC_FP64_COLUMNS: go(rbd::PrecFP64Col{}); break;
    default: go(rbd::PrecFP64{}); break;
  }
  if (ge != hipSuccess) return hip_fail(ge, "occupancy query");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad_minv_kernel launch");
}

}  // namespace
}  // extern "C++"

rbd_status rbd_fdgrad_minv_traj_f32(rbd_context* ctx, int64_t n, const float* d_in, float* d_out, float* d_minv,
                                   rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out || !d_minv) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  return launch_minv<false>(ctx, n, d_in, d_out, d_minv, stream);
}

rbd_status rbd_fdgrad_wafr_f32(rbd_context* ctx, int64_t n, const float* d_in, float* d_qdd, float* d_dqdd,
                               rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_dqdd) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  return launch_minv<true>(ctx, n, d_in, d_dqdd, d_qdd, stream);
}

namespace {

rbd_status check_fext(const float* d_fext, int64_t fext_stride) {
  if (!d_fext) return fail(RBD_ERR_INVALID_ARGUMENT, "null external-force buffer");
  if (fext_stride != 0 && fext_stride < rbd::kFextPerSet)
    return fail(RBD_ERR_INVALID_ARGUMENT, "fext_stride must be 0 or at least 42");
  return RBD_OK;
}

rbd_status forward_dynamics(rbd_context* ctx, int64_t n, const float* d_in, const float* d_fext,
                            int64_t fext_stride, float* d_qdd, float* d_minv, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_qdd || !d_minv) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const dim3 grid(grid_for(n, 1)), block(kBlock);
  const hipStream_t st = (hipStream_t)stream;
  // one wave per 64-knot block: a grid of resident waves looping over the
  // blocks took 77.6 against 63 us at 2^20 knots here (11 resident per CU,
  // so the blocks do not divide evenly over them)
  auto go = [&](auto prec_tag) {
    using P = decltype(prec_tag);
    if (d_fext)
      hipLaunchKernelGGL((rbd::fd_kernel<P, true>), grid, block, 0, st, long(n), d_in, d_fext, long(fext_stride),
                         d_qdd, d_minv, ctx->opt);
    else
      hipLaunchKernelGGL((rbd::fd_kernel<P, false>), grid, block, 0, st, long(n), d_in, nullptr, 0L, d_qdd, d_minv,
                         ctx->opt);
  };
  // forward dynamics has no derivative columns: FP64_COLUMNS computes it as FP32 does
  switch (ctx->precision) {
    case RBD_PREC_FP32:
    case RBD_PREC_FP64_COLUMNS: go(rbd::PrecFP32{}); break;
    case RBD_PREC_FP32_HW_TRIG: go(rbd::PrecFP32HwTrig{}); break;
    case RBD_PREC_FP64_PIVOT: go(rbd::PrecFP64Pivot{}); break;
    default: go(rbd::PrecFP64{}); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fd_kernel launch");
}

}  // namespace

rbd_status rbd_forward_dynamics_f32(rbd_context* ctx, int64_t n, const float* d_in, float* d_qdd, float* d_minv,
                                    rbd_stream_t stream) {
  return forward_dynamics(ctx, n, d_in, nullptr, 0, d_qdd, d_minv, stream);
}

rbd_status rbd_forward_dynamics_fext_f32(rbd_context* ctx, int64_t n, const float* d_in, const float* d_fext,
                                         int64_t fext_stride, float* d_qdd, float* d_minv, rbd_stream_t stream) {
  if (rbd_status s = check_fext(d_fext, fext_stride)) return s;
  return forward_dynamics(ctx, n, d_in, d_fext, fext_stride, d_qdd, d_minv, stream);
}

rbd_status rbd_fdgrad_traj_fext_f32(rbd_context* ctx, int64_t n, const float* d_in, const float* d_fext,
                                    int64_t fext_stride, float* d_out, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = check_fext(d_fext, fext_stride)) return s;
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const int lanes = lanes_for(ctx, n);
  const dim3 grid(grid_for(n, lanes)), block(kBlock);
  const hipStream_t st = (hipStream_t)stream;
  hipError_t ge = hipSuccess;
  auto go = [&](auto prec_tag) {
    using P = decltype(prec_tag);
    if (lanes == 1) {
      dim3 g = grid;
      ge = one_lane_grid<rbd::fdgrad_fext_kernel<P, 1>>(ctx->compute_units, &g);
      if (ge == hipSuccess)
        hipLaunchKernelGGL((rbd::fdgrad_fext_kernel<P, 1>), g, block, 0, st, long(n), d_in, d_fext,
                           long(fext_stride), d_out, ctx->opt);
    } else
      hipLaunchKernelGGL((rbd::fdgrad_fext_kernel<P, rbd::kLanesPerKnotGroup>), grid, block, 0, st, long(n), d_in,
                         d_fext, long(fext_stride), d_out, ctx->opt);
  };
  switch (ctx->precision) {
    case RBD_PREC_FP32: go(rbd::PrecFP32{}); break;
    case RBD_PREC_FP32_HW_TRIG: go(rbd::PrecFP32HwTrig{}); break;
    case RBD_PREC_FP64_PIVOT: go(rbd::PrecFP64Pivot{}); break;
    case RBD_PREC_FP64_COLUMNS: go(rbd::PrecFP64Col{}); break;
    default: go(rbd::PrecFP64{}); break;
  }
  if (ge != hipSuccess) return hip_fail(ge, "occupancy query");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad_fext_kernel launch");
}

extern "C++" {
namespace {

template <bool SPLIT>
rbd_status id_gradient(rbd_context* ctx, int64_t n, float* d_cdu, const float* d_vol, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_cdu || !d_vol) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const int lanes = lanes_for(ctx, n);
  const dim3 grid(grid_for(n, lanes)), block(kBlock);
  const hipStream_t st = (hipStream_t)stream;
  auto go = [&](auto prec_tag) {
    using P = decltype(prec_tag);
    if (lanes == 1)
      hipLaunchKernelGGL((rbd::id_gradient_kernel<P, 1, SPLIT>), grid, block, 0, st, long(n), d_vol, d_cdu, ctx->opt);
    else
      hipLaunchKernelGGL((rbd::id_gradient_kernel<P, rbd::kLanesPerKnotGroup, SPLIT>), grid, block, 0, st, long(n),
                         d_vol, d_cdu, ctx->opt);
  };
  switch (ctx->precision) {
    case RBD_PREC_FP32: go(rbd::PrecFP32{}); break;
    case RBD_PREC_FP32_HW_TRIG: go(rbd::PrecFP32HwTrig{}); break;
    case RBD_PREC_FP64_PIVOT: go(rbd::PrecFP64Pivot{}); break;   // no pivots here: same as FP32
    case RBD_PREC_FP64_COLUMNS: go(rbd::PrecFP64Col{}); break;
    default: go(rbd::PrecFP64{}); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "id_gradient_kernel launch");
}

}  // namespace
}  // extern "C++"

rbd_status rbd_dynamics_gradient_kernel_fused_f32(rbd_context* ctx, int64_t n, float* d_cdu,
                                                  const float* d_mem_vol_fused, const float* d_mem_const,
                                                  rbd_stream_t stream) {
  (void)d_mem_const;  // the constants are compiled into the kernel
  return id_gradient<false>(ctx, n, d_cdu, d_mem_vol_fused, stream);
}

rbd_status rbd_dynamics_gradient_kernel_split_f32(rbd_context* ctx, int64_t n, float* d_cdu,
                                                  const float* d_mem_vol_split, const float* d_mem_const,
                                                  rbd_stream_t stream) {
  (void)d_mem_const;
  return id_gradient<true>(ctx, n, d_cdu, d_mem_vol_split, stream);
}

rbd_status rbd_gradient_setup_split_f32(rbd_context* ctx, int64_t n, float* d_mem_vol_split, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_mem_vol_split) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const dim3 grid(grid_for(n, 1)), block(kBlock);
  const hipStream_t st = (hipStream_t)stream;
  if (ctx->precision == RBD_PREC_FP64)
    hipLaunchKernelGGL(rbd::vaf_kernel<rbd::PrecFP64>, grid, block, 0, st, long(n), d_mem_vol_split, ctx->opt);
  else
    hipLaunchKernelGGL(rbd::vaf_kernel<rbd::PrecFP32>, grid, block, 0, st, long(n), d_mem_vol_split, ctx->opt);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "vaf_kernel launch");
}

rbd_status rbd_gradient_finish_f32(rbd_context* ctx, int64_t n, float* d_dqdd, const float* d_cdu,
                                   const float* d_minv, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_dqdd || !d_cdu || !d_minv) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = ready_to_launch(ctx)) return s;
  const dim3 grid(grid_for(n, 1)), block(kBlock);
  hipLaunchKernelGGL(rbd::finish_kernel, grid, block, 0, (hipStream_t)stream, long(n), d_minv, d_cdu, d_dqdd);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RBD_OK : hip_fail(e, "finish_kernel launch");
}

int64_t rbd_group_threshold(rbd_context* ctx) { return ctx ? ctx->group_threshold.load() : -1; }

int64_t rbd_two_wave_threshold(rbd_context* ctx) { return ctx ? ctx->two_wave_threshold.load() : -1; }

// Setting a threshold by hand drops the mapping table: the thresholds decide.
rbd_status rbd_set_two_wave_threshold(rbd_context* ctx, int64_t n_knots) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative threshold");
  ctx->tuned.count.store(0, std::memory_order_release);
  ctx->two_wave_threshold.store(n_knots);
  return RBD_OK;
}

rbd_status rbd_set_group_threshold(rbd_context* ctx, int64_t n_knots) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative threshold");
  ctx->tuned.count.store(0, std::memory_order_release);
  ctx->group_threshold.store(n_knots);
  return RBD_OK;
}

int rbd_mapping_table(rbd_context* ctx, int batched, int64_t* n_knots, int* lanes, int* waves_per_group,
                      int capacity) {
  if (!ctx) return -1;
  const MappingTable& t = batched ? ctx->batch_tuned : ctx->tuned;
  const int c = t.count.load(std::memory_order_acquire);
  for (int i = 0; i < c && i < capacity; ++i) {
    if (n_knots) n_knots[i] = t.n[i];
    if (lanes) lanes[i] = t.lanes[i];
    if (waves_per_group) waves_per_group[i] = t.wpg[i] ? t.wpg[i] : 1;
  }
  return c;
}

rbd_status rbd_autotune(rbd_context* ctx, rbd_stream_t stream, int64_t* threshold_out, float* table_us,
                        int table_len) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (rbd_status s = ready_to_launch(ctx)) return s;
  // sizes, including the knot counts that fill the device's SIMDs with one
  // wave of 7-lane groups (9 knots each) or of 4-lane groups (16)
  const int64_t simds = 4 * int64_t(ctx->compute_units);
  int64_t sizes[kAutotuneSizes] = {64,        256,       1024,       2048,  4096,  8192,
                                   9 * simds, 12 * simds, 16 * simds, 24 * simds, 32 * simds, 65536};
  std::sort(sizes, sizes + kAutotuneSizes);
  int64_t nmax = 0;
  for (int64_t n : sizes) nmax = n > nmax ? n : nmax;
  float* buf = nullptr;
  hipError_t e = hipMalloc(&buf, size_t(nmax) * (rbd::kInPerKnot + rbd::kOutPerKnot) * sizeof(float));
  if (e != hipSuccess) return fail(RBD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  float* in = buf;
  float* out = buf + size_t(nmax) * rbd::kInPerKnot;
  // any finite inputs do: the work per knot does not depend on the values
  e = hipMemsetAsync(in, 0, size_t(nmax) * rbd::kInPerKnot * sizeof(float), (hipStream_t)stream);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  float best_us[kAutotuneMappings][kAutotuneSizes] = {};
  for (int si = 0; si < kAutotuneSizes && e == hipSuccess; ++si) {
    for (int li = 0; li < kAutotuneMappings && e == hipSuccess; ++li) {
      FlatCall c{Layout::AoS, kAutotuneShapes[li].lanes, sizes[si], in, 0, out, 0, ctx->opt,
                 (hipStream_t)stream, 4 * ctx->compute_units};
      c.wpg = kAutotuneShapes[li].wpg;
      float best = 1e30f;
      for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) {
        e = launch_flat(ctx->precision, c);                       // warm
        if (e == hipSuccess) e = hipEventRecord(t0, (hipStream_t)stream);
        for (int it = 0; it < 4 && e == hipSuccess; ++it) e = launch_flat(ctx->precision, c);
        if (e == hipSuccess) e = hipEventRecord(t1, (hipStream_t)stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e == hipSuccess && ms * 250.0f < best) best = ms * 250.0f;   // us per launch
      }
      best_us[li][si] = best;
    }
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipFree(buf);
  if (e != hipSuccess) return hip_fail(e, "autotune");
  // the table: at each size the fastest mapping; the thresholds keep their
  // meaning for the one-lane, 16-lane and two-wave mappings
  int64_t threshold = 0, two_wave = 0;
  ctx->tuned.count.store(0, std::memory_order_release);
  for (int si = 0; si < kAutotuneSizes; ++si) {
    int bi = 0;
    for (int li = 1; li < kAutotuneMappings; ++li)
      if (best_us[li][si] < best_us[bi][si]) bi = li;
    ctx->tuned.n[si] = sizes[si];
    ctx->tuned.lanes[si] = kAutotuneShapes[bi].lanes;
    ctx->tuned.wpg[si] = kAutotuneShapes[bi].wpg;
    if (best_us[1][si] < best_us[0][si]) threshold = sizes[si];
    if (best_us[2][si] < best_us[1][si] && best_us[2][si] < best_us[0][si]) two_wave = sizes[si];
  }
  ctx->tuned.count.store(kAutotuneSizes, std::memory_order_release);
  ctx->group_threshold.store(threshold);
  ctx->two_wave_threshold.store(two_wave);
  if (threshold_out) *threshold_out = threshold;
  constexpr int kRow = 1 + kAutotuneMappings;
  if (table_us)
    for (int i = 0; i < table_len && i < kRow * kAutotuneSizes; ++i)
      table_us[i] = (i % kRow == 0) ? float(sizes[i / kRow]) : best_us[(i % kRow) - 1][i / kRow];
  return RBD_OK;
}

}  // extern "C"
