// TEST INFRASTRUCTURE ONLY -- C entry points of the CPU oracle / CPU baseline.
// Built into oracle/liboracle.so by oracle/Makefile.  Loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
//
// Per-knot layouts:
//   in   [q(7) | qd(7) | u(7)]                               21 values
//   out  [qdd(7) | dqdd/dq (7x7 col-major) | dqdd/dqd (7x7)]  105 values
//   vol  [q | qd | u | qdd | Minv dense 7x7]                  77 values
//        (the reference's COMPLETELY_FUSED layout, time_GPU.cu:141-158)
//   dqdd [dqdd/dq | dqdd/dqd]                                98 values
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "rbd_oracle.hpp"
#include "reusable_pool.hpp"

namespace rbd_oracle {
thread_local OpCount g_ops;
thread_local bool g_skip_structural_zeros = false;
thread_local int g_mutation = 0;
}

using namespace rbd_oracle;

namespace {
template <typename T>
void fdgrad_range(long k0, long k1, const T* in, T* out, int flags) {
  for (long k = k0; k < k1; ++k) {
    const T* x = in + 21 * k;
    fdgrad_knot<T>(x, x + 7, x + 14, out + 105 * k, flags);
  }
}
}  // namespace

extern "C" {

int oracle_abi_version(void) { return 1; }

// Plant mistake m (see g_mutation) in this thread's oracle; 0 restores it.
void oracle_set_mutation(int m) { g_mutation = m; }

// Replace the robot model for every later call (not while a call runs).
//   links:  7 x {m, cx, cy, cz, ixx, ixy, ixz, iyy, iyz, izz}: mass, COM and the
//           inertia about the COM, in the link frame;
//   joints: 7 x {R00 R01 R02 R10 R11 R12 R20 R21 R22, rx, ry, rz}: the joint frame
//           is R Rz(q) at offset r in the parent link's frame (R row-major).
// links == nullptr restores the iiwa 14.  Entries of R within 1e-12 of an
// integer are snapped to it, as for the iiwa's rpy angles.
void oracle_set_model(const double* links, const double* joints) {
  if (!links || !joints) { model() = iiwa_model(); return; }
  Model m;
  for (int k = 0; k < NQ; ++k) {
    const double* l = links + 10 * k;
    m.links[k] = LinkInertial{l[0], l[1], l[2], l[3], l[4], l[5], l[6], l[7], l[8], l[9]};
    const double* j = joints + 12 * k;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) m.frames[k].F[a][b] = snap_unit(j[3 * b + a]);
    for (int i = 0; i < 3; ++i) m.frames[k].r[i] = j[9 + i];
  }
  model() = m;
}

// 7 spatial inertias, 6x6 column-major each.
void oracle_inertias(double* I) {
  for (int k = 0; k < NQ; ++k) spatial_inertia<double>(k, I + 36 * k);
}

// 7 joint motion transforms T_k(q_k), 6x6 column-major each.
void oracle_transforms(const double* q, double* X) {
  for (int k = 0; k < NQ; ++k) joint_transform<double>(k, std::sin(q[k]), std::cos(q[k]), X + 36 * k);
}

// tau = RNEA(q, qd, qdd).
void oracle_rnea(const double* q, const double* qd, const double* qdd, double* tau, int flags) {
  rnea_knot<double>(q, qd, qdd, tau, flags);
}

// qdd = FD(q, qd, u) and the dense Minv (may be null).
void oracle_fd(const double* q, const double* qd, const double* u, double* qdd, double* Minv, int flags) {
  fd_knot<double>(q, qd, u, qdd, Minv, flags);
}

// The same with external forces fext (42 values: link k's spatial force in its
// own frame at fext[6k .. 6k+5]; may be null).
void oracle_rnea_fext(const double* q, const double* qd, const double* qdd, const double* fext, double* tau,
                      int flags) {
  rnea_knot<double>(q, qd, qdd, tau, flags, fext);
}
void oracle_fd_fext(const double* q, const double* qd, const double* u, const double* fext, double* qdd,
                    double* Minv, int flags) {
  fd_knot<double>(q, qd, u, qdd, Minv, flags, fext);
}

// Full path with external forces: knot k uses fext + k * fext_stride (stride 0:
// one set of 42 for every knot, as the reference's per-call s_fext).
void oracle_fdgrad_fext_f64(long n, const double* in, const double* fext, long fext_stride, double* out,
                            int flags) {
  for (long k = 0; k < n; ++k) {
    const double* x = in + 21 * k;
    fdgrad_knot<double>(x, x + 7, x + 14, out + 105 * k, flags, nullptr, fext + k * fext_stride);
  }
}

// The articulated-inertia pivots D_k = S^T IA_k S of the Minv recursion.
void oracle_pivots(const double* q, double* D) {
  Knot<double> K;
  setup_model(K, q);
  compute_minv(K, D);
}

void oracle_fdgrad_f64(long n, const double* in, double* out, int flags) {
  fdgrad_range<double>(0, n, in, out, flags);
}

void oracle_fdgrad_f32(long n, const float* in, float* out, int flags) {
  fdgrad_range<float>(0, n, in, out, flags);
}

// Reference-compatible mode: vol (77/knot) -> dqdd (98/knot).
void oracle_fdgrad_refmode_f64(long n, const double* vol, double* dqdd, int flags) {
  for (long k = 0; k < n; ++k) fdgrad_knot_given_qdd_minv<double>(vol + 77 * k, dqdd + 98 * k, flags);
}
void oracle_fdgrad_refmode_f32(long n, const float* vol, float* dqdd, int flags) {
  for (long k = 0; k < n; ++k) fdgrad_knot_given_qdd_minv<float>(vol + 77 * k, dqdd + 98 * k, flags);
}

// Complex-step derivative of forward dynamics: for one knot, out = [qdd | dqdd/dq | dqdd/dqd]
// with each derivative column Im(FD(x + i h e_j)) / h -- no subtractive cancellation,
// so h = 1e-20 gives derivatives to machine precision.  Pins the analytic gradient.
void oracle_fd_complex_step_fext(const double* in, const double* fext, double h, double* out, int flags) {
  using C = std::complex<double>;
  C q[NQ], qd[NQ], u[NQ], qdd[NQ];
  for (int j = 0; j < 2 * NQ + 1; ++j) {
    for (int i = 0; i < NQ; ++i) { q[i] = in[i]; qd[i] = in[7 + i]; u[i] = in[14 + i]; }
    if (j < NQ) q[j] += C(0, h);
    else if (j < 2 * NQ) qd[j - NQ] += C(0, h);
    fd_knot<C>(q, qd, u, qdd, nullptr, flags, fext);
    if (j == 2 * NQ) for (int i = 0; i < NQ; ++i) out[i] = qdd[i].real();
    else for (int i = 0; i < NQ; ++i) out[NQ + NQ * j + i] = qdd[i].imag() / h;
  }
}
void oracle_fd_complex_step(const double* in, double h, double* out, int flags) {
  oracle_fd_complex_step_fext(in, nullptr, h, out, flags);
}

// The reference's staged variants (time_GPU.cu:227-333).
// Fused layout: vol28 = [q | qd | u | qdd]; split layout: vol154 = [q | qd | u | qdd |
// v (6 per link) | a | f], f after RNEA's backward pass.  cdu = [dtau/dq | dtau/dqd],
// each 7x7 column-major.
void oracle_id_gradient_f64(long n, const double* vol, double* cdu, int flags) {
  for (long k = 0; k < n; ++k) {
    const double* x = vol + 28 * k;
    Knot<double> K;
    setup_model(K, x);
    rnea_forward(K, x + 7, x + 21, flags);
    rnea_backward(K, (double*)nullptr);
    inverse_dynamics_gradient(K, x + 7, cdu + 98 * k, cdu + 98 * k + 49, flags);
  }
}
void oracle_vaf_f64(long n, double* vol, int flags) {
  for (long k = 0; k < n; ++k) {
    double* x = vol + 154 * k;
    Knot<double> K;
    setup_model(K, x);
    rnea_forward(K, x + 7, x + 21, flags);
    rnea_backward(K, (double*)nullptr);
    for (int j = 0; j < NQ; ++j)
      for (int i = 0; i < 6; ++i) {
        x[28 + 6 * j + i] = K.v[j][i];
        x[70 + 6 * j + i] = K.a[j][i];
        x[112 + 6 * j + i] = K.f[j][i];
      }
  }
}
void oracle_id_gradient_split_f64(long n, const double* vol, double* cdu, int flags) {
  for (long k = 0; k < n; ++k) {
    const double* x = vol + 154 * k;
    Knot<double> K;
    setup_model(K, x);
    for (int j = 0; j < NQ; ++j)
      for (int i = 0; i < 6; ++i) {
        K.v[j][i] = x[28 + 6 * j + i];
        K.a[j][i] = x[70 + 6 * j + i];
        K.f[j][i] = x[112 + 6 * j + i];
      }
    inverse_dynamics_gradient(K, x + 7, cdu + 98 * k, cdu + 98 * k + 49, flags);
  }
}
// dqdd = -Minv cdu (upper triangle of the 49-value column-major Minv).
void oracle_finish_f64(long n, const double* minv, const double* cdu, double* dqdd) {
  for (long k = 0; k < n; ++k)
    for (int c = 0; c < 2 * NQ; ++c) symv_upper<true>(dqdd + 98 * k + NQ * c, minv + 49 * k, cdu + 98 * k + NQ * c);
}
// cdu of one knot (vol28) by complex-step differentiation of RNEA at fixed qdd.
void oracle_rnea_complex_step(const double* vol, double h, double* cdu, int flags) {
  using C = std::complex<double>;
  C q[NQ], qd[NQ], qdd[NQ], tau[NQ];
  for (int j = 0; j < 2 * NQ; ++j) {
    for (int i = 0; i < NQ; ++i) { q[i] = vol[i]; qd[i] = vol[7 + i]; qdd[i] = vol[21 + i]; }
    if (j < NQ) q[j] += C(0, h);
    else qd[j - NQ] += C(0, h);
    rnea_knot<C>(q, qd, qdd, tau, flags);
    for (int i = 0; i < NQ; ++i) cdu[NQ * j + i] = tau[i].imag() / h;
  }
}

// Operation count of one knot point of the FP32 method.
// mode 0: full path (q, qd, u -> qdd, dqdd/dq, dqdd/dqd);
// mode 1: reference-compatible mode (qdd and Minv given).
// skip_zeros: do not count operations with a structural-zero operand.
// counts = [add/sub, mul, div, sin+cos]
void oracle_op_count(int mode, int skip_zeros, long long* counts) {
  g_ops = OpCount();
  g_skip_structural_zeros = skip_zeros != 0;
  Counted in[77], out[105];
  for (int i = 0; i < 77; ++i) in[i] = Counted(0.1 + 0.01 * i);
  if (mode == 0) fdgrad_knot<Counted>(in, in + 7, in + 14, out, 0);
  else fdgrad_knot_given_qdd_minv<Counted>(in, out, 0);
  counts[0] = g_ops.add; counts[1] = g_ops.mul; counts[2] = g_ops.div; counts[3] = g_ops.trig;
  g_skip_structural_zeros = false;
}

// CPU baseline: the reference's dynamicsGradientReusableThreaded timing loop
// (time_CPU.cpp:47-54, :131-140) on this pool, FP32.  Each of `iters` calls
// splits the n knots into contiguous chunks, kStart = n/P*tid with the
// remainder on the last thread (time_CPU.cpp:32), hands one chunk to each
// worker and waits.  mode 0: full path (in 21/knot -> out 105/knot); mode 1:
// reference-compatible (vol 77/knot -> dqdd 98/knot).  Stops early once
// time_cap_s has elapsed.  times_us receives per-call wall times; returns the
// number of calls made.
long oracle_cpu_bench(int threads, long n, const float* in, float* out, long iters, int mode,
                      int flags, double time_cap_s, double* times_us) {
  ReusablePool pool(threads);
  const auto t_start = std::chrono::steady_clock::now();
  long done = 0;
  for (long it = 0; it < iters; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int tid = 0; tid < threads; ++tid) {
      const long k0 = n / threads * tid;
      const long k1 = (tid == threads - 1) ? n : n / threads * (tid + 1);
      pool.add_task(tid, [=] {
        if (mode == 0) fdgrad_range<float>(k0, k1, in, out, flags);
        else
          for (long k = k0; k < k1; ++k) fdgrad_knot_given_qdd_minv<float>(in + 77 * k, out + 98 * k, flags);
      });
    }
    pool.sync();
    const auto t1 = std::chrono::steady_clock::now();
    times_us[done++] = std::chrono::duration<double, std::micro>(t1 - t0).count();
    if (std::chrono::duration<double>(t1 - t_start).count() > time_cap_s) break;
  }
  return done;
}

// Multithreaded FP64 oracle over many knots (used by tests at large sizes).
void oracle_fdgrad_f64_mt(int threads, long n, const double* in, double* out, int flags) {
  ReusablePool pool(threads);
  for (int tid = 0; tid < threads; ++tid) {
    const long k0 = n / threads * tid;
    const long k1 = (tid == threads - 1) ? n : n / threads * (tid + 1);
    pool.add_task(tid, [=] { fdgrad_range<double>(k0, k1, in, out, flags); });
  }
  pool.sync();
}

}  // extern "C"
