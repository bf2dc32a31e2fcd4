// C++ host interface over the C ABI (include/rbdgrad.h) that keeps the
// reference's per-trajectory entry points and data structures, so code written
// against plancherb1/fast-rbd-gradients switches by including this header and
// linking librbdgrad.so:
//   knot<T>, traj<T,N>                 helpers_CPU/dynamicsGradient.h:18-30
//   COMPLETELY_FUSED_VALS, CONST_VALS  helpers_GPU/dynamicsGradient.cuh:1142-1146
//   dynamicsGradientKernel             helpers_GPU/dynamicsGradient.cuh:1257-1291
//   dynamicsGradientKernel_fused/_split helpers_GPU/dynamicsGradient.cuh:1191-1255
//   dynamicsGradientReusableThreaded_start/_finish (as device stages)
//                                      helpers_GPU/dynamicsGradient.cuh:1148-1188
//   forwardDynamics                    helpers_CPU/dynamicsGradient.h:685-706, :936-939
//   gpuErrchk-style abort on error     helpers_GPU/dynamicsGradient.cuh:28-35
// Template options keep the reference's names and defaults (MPC_MODE = false,
// VEL_DAMPING = true); each option set maps to one library context, created on
// first use on the current device.  Only T = float is provided, the only type
// the reference instantiates (time_GPU.cu:344-351).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "../include/rbdgrad.h"

namespace rbdgrad {

constexpr int NUM_POS = 7;
constexpr int SPLIT_VALS = 4 * NUM_POS + 18 * NUM_POS;                  // 154
constexpr int FUSED_VALS = 4 * NUM_POS;                                 // 28
constexpr int COMPLETELY_FUSED_VALS = 4 * NUM_POS + NUM_POS * NUM_POS;  // 77
constexpr int CONST_VALS = 72 * NUM_POS;                                // 504
constexpr int DQDD_VALS = 2 * NUM_POS * NUM_POS;                        // 98 per knot (also cdu)

template <typename T>
struct knot {
  T q[NUM_POS];
  T qd[NUM_POS];
  T u[NUM_POS];
  T qdd[NUM_POS];
  T Minv[NUM_POS * NUM_POS];
};

template <typename T, int KNOT_POINTS>
struct traj {
  knot<T> knots[KNOT_POINTS];
};

// Print the failure and exit, as the reference's gpuErrchk does.
inline void check(rbd_status s, const char* file, int line) {
  if (s != RBD_OK) {
    std::fprintf(stderr, "rbdgrad error: %s (%s) %s %d\n", rbd_status_string(s), rbd_last_error(), file, line);
    std::exit(int(s));
  }
}
#define rbdErrchk(x) rbdgrad::check((x), __FILE__, __LINE__)

// One context per option set, on the device current at first use.
template <bool MPC_MODE, bool VEL_DAMPING>
inline rbd_context* context() {
  static std::once_flag once;
  static rbd_context* ctx = nullptr;
  std::call_once(once, [] {
    rbd_config cfg;
    rbd_config_default(&cfg);
    int dev = 0;
    (void)hipGetDevice(&dev);
    cfg.device = dev;
    cfg.flags = (MPC_MODE ? RBD_FLAG_NO_GRAVITY : 0u) | (VEL_DAMPING ? RBD_FLAG_VEL_DAMPING : 0u);
    rbdErrchk(rbd_create(&cfg, &ctx));
  });
  return ctx;
}

// dynamicsGradientKernel<T,KNOT_POINTS,MPC_MODE,VEL_DAMPING><<<KNOT_POINTS,dim3(32,7)>>>(
//     d_dqdd, d_mem_vol_completely_fused, d_mem_const): qdd and dense Minv given
// (77 per knot), dqdd/dq | dqdd/dqd out (98 per knot); enqueued on `stream`.
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status dynamicsGradientKernel(T* d_dqdd, T* d_mem_vol_completely_fused, T* d_mem_const,
                                         hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_dynamics_gradient_kernel_f32(context<MPC_MODE, VEL_DAMPING>(), KNOT_POINTS, d_dqdd,
                                          d_mem_vol_completely_fused, d_mem_const, stream);
}

// dynamicsGradientKernel_fused<T,KNOT_POINTS,MPC_MODE,VEL_DAMPING>(d_cdu,
// d_mem_vol_fused, d_mem_const): q, qd, u, qdd (28 per knot) -> cdu (98 per knot).
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status dynamicsGradientKernel_fused(T* d_cdu, T* d_mem_vol_fused, T* d_mem_const, hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_dynamics_gradient_kernel_fused_f32(context<MPC_MODE, VEL_DAMPING>(), KNOT_POINTS, d_cdu,
                                                d_mem_vol_fused, d_mem_const, stream);
}

// dynamicsGradientKernel_split<T,KNOT_POINTS,VEL_DAMPING>(d_cdu, d_mem_vol_split,
// d_mem_const): q, qd, u, qdd, v, a, f (154 per knot) -> cdu.
template <typename T, int KNOT_POINTS, bool VEL_DAMPING = true>
inline rbd_status dynamicsGradientKernel_split(T* d_cdu, T* d_mem_vol_split, T* d_mem_const, hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_dynamics_gradient_kernel_split_f32(context<false, VEL_DAMPING>(), KNOT_POINTS, d_cdu, d_mem_vol_split,
                                                d_mem_const, stream);
}

// The split pipeline's CPU stages, run on the device: v, a, f into a split
// buffer (the reference's _start), and dqdd = -Minv cdu (its _finish; Minv 49
// per knot, column-major, upper triangle read).
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status dynamicsGradientStart(T* d_mem_vol_split, hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_gradient_setup_split_f32(context<MPC_MODE, VEL_DAMPING>(), KNOT_POINTS, d_mem_vol_split, stream);
}
template <typename T, int KNOT_POINTS>
inline rbd_status dynamicsGradientFinish(T* d_dqdd, const T* d_cdu, const T* d_Minv, hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_gradient_finish_f32(context<false, true>(), KNOT_POINTS, d_dqdd, d_cdu, d_Minv, stream);
}

// The full path for one trajectory on the device: 21 in (q, qd, u) and 105 out
// (qdd, dqdd/dq, dqdd/dqd) per knot -- what the reference composes on the CPU in
// forwardDynamicsGradientSetup<QDD_MINV_PASSED_IN=false> + forwardDynamicsGradientThreaded
// (dynamicsGradient.h:868-897, :924-934).
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status forwardDynamicsGradient(T* d_out, const T* d_in, hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_fdgrad_traj_f32(context<MPC_MODE, VEL_DAMPING>(), KNOT_POINTS, d_in, d_out, stream);
}

// The same with external link forces (the reference's s_fext argument,
// dynamicsGradient.h:916-934): d_fext holds 42 floats per set (link k's
// [moment | force] in its own frame at 6k), one set for every knot
// (fext_stride = 0) or one per knot (fext_stride = FEXT_VALS).
constexpr int FEXT_VALS = 6 * NUM_POS;
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status forwardDynamicsGradient(T* d_out, const T* d_in, const T* d_fext, int64_t fext_stride = 0,
                                          hipStream_t stream = 0) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  return rbd_fdgrad_traj_fext_f32(context<MPC_MODE, VEL_DAMPING>(), KNOT_POINTS, d_in, d_fext, fext_stride, d_out,
                                  stream);
}

// forwardDynamics for every knot of a host trajectory, computed on the GPU:
// fills knot.qdd and knot.Minv (dense, DENSE_MINV = 1) from q, qd, u.  Blocking.
// h_fext (host, FEXT_VALS floats, may be null): external link forces applied
// at every knot (forwardDynamics(knot, s_fext), dynamicsGradient.h:936-939).
template <typename T, int KNOT_POINTS, bool MPC_MODE = false, bool VEL_DAMPING = true>
inline rbd_status forwardDynamics(traj<T, KNOT_POINTS>* h_traj, const T* h_fext = nullptr) {
  static_assert(std::is_same<T, float>::value, "librbdgrad provides the FP32 interface");
  constexpr size_t n = KNOT_POINTS;
  float* d = nullptr;
  if (hipMalloc(&d, (n * (21 + 7 + 49) + FEXT_VALS) * sizeof(float)) != hipSuccess) return RBD_ERR_OUT_OF_MEMORY;
  float* d_fext = d + n * (21 + 7 + 49);
  float* h = nullptr;
  if (hipHostMalloc(reinterpret_cast<void**>(&h), n * (21 + 7 + 49) * sizeof(float), hipHostMallocDefault) !=
      hipSuccess) {
    (void)hipFree(d);
    return RBD_ERR_OUT_OF_MEMORY;
  }
  for (size_t k = 0; k < n; ++k)
    for (int i = 0; i < NUM_POS; ++i) {
      h[21 * k + i] = h_traj->knots[k].q[i];
      h[21 * k + NUM_POS + i] = h_traj->knots[k].qd[i];
      h[21 * k + 2 * NUM_POS + i] = h_traj->knots[k].u[i];
    }
  rbd_status st = RBD_OK;
  if (hipMemcpy(d, h, n * 21 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = RBD_ERR_HIP;
  if (st == RBD_OK && h_fext &&
      hipMemcpy(d_fext, h_fext, FEXT_VALS * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    st = RBD_ERR_HIP;
  if (st == RBD_OK)
    st = h_fext ? rbd_forward_dynamics_fext_f32(context<MPC_MODE, VEL_DAMPING>(), int64_t(n), d, d_fext, 0,
                                                d + 21 * n, d + 28 * n, 0)
                : rbd_forward_dynamics_f32(context<MPC_MODE, VEL_DAMPING>(), int64_t(n), d, d + 21 * n, d + 28 * n, 0);
  if (st == RBD_OK &&
      hipMemcpy(h + 21 * n, d + 21 * n, n * 56 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    st = RBD_ERR_HIP;
  if (st == RBD_OK)
    for (size_t k = 0; k < n; ++k) {
      for (int i = 0; i < NUM_POS; ++i) h_traj->knots[k].qdd[i] = h[21 * n + 7 * k + i];
      for (int i = 0; i < NUM_POS * NUM_POS; ++i) h_traj->knots[k].Minv[i] = h[28 * n + 49 * k + i];
    }
  (void)hipHostFree(h);
  (void)hipFree(d);
  return st;
}

// Pack q, qd, u, qdd, Minv of every knot into the reference's COMPLETELY_FUSED
// layout (time_GPU.cu:141-158).
template <typename T, int KNOT_POINTS>
inline void packCompletelyFused(const traj<T, KNOT_POINTS>& t, T* h_mem_vol) {
  for (int k = 0; k < KNOT_POINTS; ++k) {
    T* v = h_mem_vol + COMPLETELY_FUSED_VALS * k;
    for (int i = 0; i < NUM_POS; ++i) {
      v[i] = t.knots[k].q[i];
      v[NUM_POS + i] = t.knots[k].qd[i];
      v[2 * NUM_POS + i] = t.knots[k].u[i];
      v[3 * NUM_POS + i] = t.knots[k].qdd[i];
    }
    for (int i = 0; i < NUM_POS * NUM_POS; ++i) v[4 * NUM_POS + i] = t.knots[k].Minv[i];
  }
}

}  // namespace rbdgrad
