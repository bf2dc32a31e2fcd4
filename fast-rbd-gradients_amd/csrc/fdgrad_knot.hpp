// Per-knot-point forward-dynamics gradient of the KUKA iiwa 14, register resident.
//
// One call computes, for one knot point, q, qd, u -> qdd, dqdd/dq, dqdd/dqd with
// the reference's method (plancherb1/fast-rbd-gradients):
//   1. sin/cos of q and the seven joint transforms  (dynamicsGradient.h:507-587)
//   2. Minv by the articulated-body recursion         (minv.h:423-478)
//   3. RNEA with qdd = 0 -> bias c                    (dynamicsGradient.h:619-683)
//   4. qdd = Minv (u - c)                             (dynamicsGradient.h:704)
//   5. RNEA with qdd: v, a, f                         (dynamicsGradient.h:893-895)
//   6. dtau/dq, dtau/dqd: forward sweep over links per column, backward sweep
//                                                     (dynamicsGradient.h:786-857)
//   7. dqdd/d(q,qd) = -Minv dtau/d(q,qd)              (dynamicsGradient.h:921)
// Every link index is a compile-time constant (static_for), so each joint's
// transform, inertia and offset are folded into straight-line code: structural
// zeros of E, r and I cost nothing when the file is compiled with
// -fno-signed-zeros -ffinite-math-only (these let x*0 and x+0 fold away).
// In one lane the thirteen columns of step 6 are unrolled at compile time, so
// each column's sweep starts at its own link with its seed folded into the
// code (a lane group instead gives each lane a run-time column of its own,
// dtau_column_uniform); each output column -Minv dtau[:, c] is stored as soon
// as it is known, so dtau never has to be held whole.
//
// The file is plain C++17 when RBD_HD is empty, so the same arithmetic is also
// compiled for the host by the CPU test harness.
#pragma once
// The robot: the KUKA iiwa 14 unless the build names a header generated from a
// URDF by urdf_model.py (-DRBD_MODEL_HEADER='"model_NAME.hpp"').
#ifdef RBD_MODEL_HEADER
#include RBD_MODEL_HEADER
#else
#include "iiwa_model.hpp"
#endif
#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RBD_HD __host__ __device__
#else
#include <cmath>
#define RBD_HD
#endif
#define RBD_INLINE __attribute__((always_inline)) inline

namespace rbd {

template <int N> struct IC { static constexpr int value = N; };

// f(IC<B>), f(IC<B+1>), ..., f(IC<E-1>)
template <int B, int E, class F>
RBD_HD RBD_INLINE void static_for(F&& f) {
  if constexpr (B < E) { f(IC<B>{}); static_for<B + 1, E>(f); }
}
// f(IC<B-1>), f(IC<B-2>), ..., f(IC<E>)
template <int B, int E, class F>
RBD_HD RBD_INLINE void static_for_down(F&& f) {
  if constexpr (B > E) { f(IC<B - 1>{}); static_for_down<B - 1, E>(f); }
}

template <typename S> struct CS { S c, s; };

RBD_HD RBD_INLINE void sincos_s(float x, float* s, float* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  ::sincosf(x, s, c);
#else
  *s = std::sin(x); *c = std::cos(x);
#endif
}
RBD_HD RBD_INLINE void sincos_s(double x, double* s, double* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  ::sincos(x, s, c);
#else
  *s = std::sin(x); *c = std::cos(x);
#endif
}

// ---------------------------------------------------------------------------
// Joint transform of link K: T = [E 0; -E r^ E]
// ---------------------------------------------------------------------------
template <int K, int I, int J, typename S>
RBD_HD RBD_INLINE S Eij(const CS<S>& t) {
  constexpr S a = S(ecoef_c(K, I, J)), b = S(ecoef_s(K, I, J)), d = S(ecoef_1(K, I, J));
  return a * t.c + b * t.s + d;
}
// y = E x
template <int K, typename S>
RBD_HD RBD_INLINE void rot(const CS<S>& t, const S* x, S* y) {
  static_for<0, 3>([&](auto Ic) {
    constexpr int i = decltype(Ic)::value;
    y[i] = Eij<K, i, 0>(t) * x[0] + Eij<K, i, 1>(t) * x[1] + Eij<K, i, 2>(t) * x[2];
  });
}
// y = E^T x
template <int K, typename S>
RBD_HD RBD_INLINE void rot_t(const CS<S>& t, const S* x, S* y) {
  static_for<0, 3>([&](auto Jc) {
    constexpr int j = decltype(Jc)::value;
    y[j] = Eij<K, 0, j>(t) * x[0] + Eij<K, 1, j>(t) * x[1] + Eij<K, 2, j>(t) * x[2];
  });
}
template <int K, typename S> constexpr S roff(int i) { return S(kJointOffset[K][i]); }

// y = T x for a motion vector: [E w; E (u - r x w)]
template <int K, typename S>
RBD_HD RBD_INLINE void xmotion(const CS<S>& t, const S* x, S* y) {
  constexpr S rx = roff<K, S>(0), ry = roff<K, S>(1), rz = roff<K, S>(2);
  // u - r x w, each product subtracted on its own
  S d[3] = {x[3], x[4], x[5]};
  d[0] -= ry * x[2]; d[0] += rz * x[1];
  d[1] -= rz * x[0]; d[1] += rx * x[2];
  d[2] -= rx * x[1]; d[2] += ry * x[0];
  S y0[3], y1[3];
  rot<K>(t, x, y0);
  rot<K>(t, d, y1);
  y[0] = y0[0]; y[1] = y0[1]; y[2] = y0[2];
  y[3] = y1[0]; y[4] = y1[1]; y[5] = y1[2];
}
// y = T^T x (+ z) for a force vector: [E^T n + r x (E^T f); E^T f] (+ z).
// The angular rows add each product to the row on its own (one FMA each),
// starting from z when there is one.
template <int K, bool ADD, typename S>
RBD_HD RBD_INLINE void xforce_t_into(const CS<S>& t, const S* x, const S* z, S* y) {
  constexpr S rx = roff<K, S>(0), ry = roff<K, S>(1), rz = roff<K, S>(2);
  S fl[3];
  rot_t<K>(t, x + 3, fl);
  const S r[3] = {rx, ry, rz};
  static_for<0, 3>([&](auto Jc) {
    constexpr int j = decltype(Jc)::value;
    constexpr int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    S acc = ADD ? z[j] : S(0);
    acc += Eij<K, 0, j>(t) * x[0];
    acc += Eij<K, 1, j>(t) * x[1];
    acc += Eij<K, 2, j>(t) * x[2];
    acc += r[j1] * fl[j2];                    // (r x fl)_j
    acc -= r[j2] * fl[j1];
    y[j] = acc;
  });
  for (int j = 0; j < 3; ++j) y[3 + j] = ADD ? z[3 + j] + fl[j] : fl[j];
}
template <int K, typename S>
RBD_HD RBD_INLINE void xforce_t(const CS<S>& t, const S* x, S* y) {
  xforce_t_into<K, false>(t, x, static_cast<const S*>(nullptr), y);
}
// y = T^T x + z
template <int K, typename S>
RBD_HD RBD_INLINE void xforce_t_add(const CS<S>& t, const S* x, const S* z, S* y) {
  xforce_t_into<K, true>(t, x, z, y);
}

// entry (R, C) of link K's spatial inertia, as a compile-time constant
template <typename S, int K, int R, int C>
constexpr S kInertia = S(inertia_entry(K, R, C));

// y (+)= I_K x, I_K the constant spatial inertia of link K
template <int K, bool ACC = false, typename S>
RBD_HD RBD_INLINE void imv(const S* x, S* y) {
  static_for<0, 6>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    S acc = kInertia<S, K, r, 0> * x[0];
    static_for<1, 6>([&](auto Cc) {
      constexpr int c = decltype(Cc)::value;
      acc += kInertia<S, K, r, c> * x[c];
    });
    y[r] = ACC ? y[r] + acc : acc;
  });
}
// A = I_K
template <int K, typename S>
RBD_HD RBD_INLINE void iload(S A[6][6]) {
  static_for<0, 6>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    static_for<0, 6>([&](auto Cc) {
      constexpr int c = decltype(Cc)::value;
      A[r][c] = kInertia<S, K, r, c>;
    });
  });
}

// y (+)= x x* z  =  [w x za + u x zl ; w x zl]   (spatial force cross product).
// Accumulating, each product is added to y on its own (one FMA each) rather
// than summed first and added once: the same terms, one instruction fewer
// per component.
template <bool ACC = false, typename S>
RBD_HD RBD_INLINE void crf(const S* x, const S* z, S* y) {
  if (ACC) {
    S y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4], y5 = y[5];
    y0 += x[1] * z[2]; y0 -= x[2] * z[1]; y0 += x[4] * z[5]; y0 -= x[5] * z[4];
    y1 += x[2] * z[0]; y1 -= x[0] * z[2]; y1 += x[5] * z[3]; y1 -= x[3] * z[5];
    y2 += x[0] * z[1]; y2 -= x[1] * z[0]; y2 += x[3] * z[4]; y2 -= x[4] * z[3];
    y3 += x[1] * z[5]; y3 -= x[2] * z[4];
    y4 += x[2] * z[3]; y4 -= x[0] * z[5];
    y5 += x[0] * z[4]; y5 -= x[1] * z[3];
    y[0] = y0; y[1] = y1; y[2] = y2; y[3] = y3; y[4] = y4; y[5] = y5;
  } else {
    y[0] = x[1] * z[2] - x[2] * z[1] + x[4] * z[5] - x[5] * z[4];
    y[1] = x[2] * z[0] - x[0] * z[2] + x[5] * z[3] - x[3] * z[5];
    y[2] = x[0] * z[1] - x[1] * z[0] + x[3] * z[4] - x[4] * z[3];
    y[3] = x[1] * z[5] - x[2] * z[4];
    y[4] = x[2] * z[3] - x[0] * z[5];
    y[5] = x[0] * z[4] - x[1] * z[3];
  }
}

// ---------------------------------------------------------------------------
// Articulated-inertia transport: P = T^T Ia T for a symmetric Ia whose row and
// column 2 are zero (revolute joint about z).  T = [E 0; 0 E][1 0; -r^ 1]:
// rotate the 3x3 blocks (A' = E^T A E, C' = E^T C E, M' = E^T M E), then
// translate: TL = A' + r^ C'^T - C' r^ - r^ M' r^, TR = C' + r^ M', BR = M'.
// (Same product as the reference's ctransform_Ia_revolute, minv.h:221-386.)
// ---------------------------------------------------------------------------
template <int K, typename S>
RBD_HD RBD_INLINE void transport(const CS<S>& t, const S Ia[6][6], S P[6][6]) {
  S A[3][3], C[3][3], M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { A[i][j] = Ia[i][j]; C[i][j] = Ia[i][j + 3]; M[i][j] = Ia[i + 3][j + 3]; }
  // X' = E^T X E : first Y = X E (rows of X times E), then E^T Y
  auto sandwich = [&](const S X[3][3], S out[3][3]) {
    S Y[3][3];
    for (int i = 0; i < 3; ++i) {
      S row_e[3];
      // (X E)[i][j] = sum_l X[i][l] E[l][j] = (E^T x_i)_j  with x_i the i-th row
      const S xi[3] = {X[i][0], X[i][1], X[i][2]};
      rot_t<K>(t, xi, row_e);
      Y[i][0] = row_e[0]; Y[i][1] = row_e[1]; Y[i][2] = row_e[2];
    }
    for (int j = 0; j < 3; ++j) {
      const S yj[3] = {Y[0][j], Y[1][j], Y[2][j]};
      S col[3];
      rot_t<K>(t, yj, col);            // (E^T Y)[:, j] = E^T y_j
      out[0][j] = col[0]; out[1][j] = col[1]; out[2][j] = col[2];
    }
  };
  S Ar[3][3], Cr[3][3], Mr[3][3];
  sandwich(A, Ar);
  sandwich(C, Cr);
  sandwich(M, Mr);
  constexpr S rx = roff<K, S>(0), ry = roff<K, S>(1), rz = roff<K, S>(2);
  const S R[3][3] = {{S(0), -rz, ry}, {rz, S(0), -rx}, {-ry, rx, S(0)}};
  // TR = Cr + R Mr
  S TR[3][3], RM[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      RM[i][j] = R[i][0] * Mr[0][j] + R[i][1] * Mr[1][j] + R[i][2] * Mr[2][j];
      TR[i][j] = Cr[i][j] + RM[i][j];
    }
  // TL = Ar + R Cr^T - Cr R - R Mr R = Ar + R Cr^T - (Cr + R Mr) R = Ar + R Cr^T - TR R
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j <= i; ++j) {
      const S rct = R[i][0] * Cr[j][0] + R[i][1] * Cr[j][1] + R[i][2] * Cr[j][2];
      const S trr = TR[i][0] * R[0][j] + TR[i][1] * R[1][j] + TR[i][2] * R[2][j];
      const S e = Ar[i][j] + rct - trr;
      P[i][j] = e; P[j][i] = e;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      P[i][j + 3] = TR[i][j];
      P[j + 3][i] = TR[i][j];
      P[i + 3][j + 3] = Mr[i][j];
    }
}


// ---------------------------------------------------------------------------
// Arithmetic precision per phase.  Inputs and outputs are FP32 in every mode.
//   Minv  : the articulated-body recursion for Minv
//   Pivot : the pivots D_k = S^T IA_k S and 1/D_k inside it
//   Rnea  : bias RNEA, qdd and the RNEA with qdd
//   Col   : the derivative columns dtau/d(q,qd) and -Minv dtau
// ---------------------------------------------------------------------------
//   HW_TRIG: FP32 sin/cos from the hardware's v_sin_f32/v_cos_f32 (knot_sincos)
template <typename SMinv, typename SPivot, typename SRnea, typename SCol, bool HW_TRIG = false>
struct Precision {
  using Minv = SMinv; using Pivot = SPivot; using Rnea = SRnea; using Col = SCol;
  static constexpr bool kHwTrig = HW_TRIG;
};
using PrecFP32      = Precision<float, float, float, float>;     // all FP32
using PrecFP32HwTrig = Precision<float, float, float, float, true>;  // all FP32, hardware sin/cos
using PrecFP64Pivot = Precision<float, double, float, float>;    // FP64 pivots only
using PrecFP64Col   = Precision<float, float, float, double>;    // FP64 derivative columns
using PrecFP64      = Precision<double, double, double, double>; // all FP64

// Waves each SIMD keeps resident for the kernels of precision P.  With FP32
// RNEA and columns a knot's state fits in 256 registers per lane, so two waves
// share a SIMD and each hides the other's latency (1.44x the throughput of one
// wave at 2^20 knots when measured); with FP64 columns it needs about 500 and
// runs one.
template <class P>
constexpr int kWavesPerSimd = (sizeof(typename P::Col) == 4 && sizeof(typename P::Rnea) == 4) ? 2 : 1;

// sin/cos of a joint angle under precision P: the math library's sincos, or,
// with P::kHwTrig and an FP32 angle, the hardware's v_sin_f32/v_cos_f32 (whose
// argument is in revolutions).  The angle is first reduced to [-pi, pi] as
// x - k 2pi with 2pi split into two FP32 parts (exact to about 2^-40 |x|), so
// small and large angles keep their precision.  Measured on MI355X: 2% less
// per-call latency at 128 knots and 1% less time at 2^20 knots, and the
// largest relative output error over 2^20 random and stretched-arm poses
// 6.5e-5 instead of 4.1e-5.  The host build uses the math library's.
template <class P, typename S>
RBD_HD RBD_INLINE void knot_sincos(S x, S* s, S* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (P::kHwTrig && std::is_same<S, float>::value) {
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    constexpr float kHi = float(kTwoPi);
    constexpr float kLo = float(kTwoPi - double(kHi));
    constexpr float kInv = float(1.0 / kTwoPi);
    const float k = __builtin_rintf(x * kInv);
    const float a = __builtin_fmaf(-k, kLo, __builtin_fmaf(-k, kHi, x));
    *s = __builtin_amdgcn_sinf(a * kInv);
    *c = __builtin_amdgcn_cosf(a * kInv);
    return;
  }
#endif
  sincos_s(x, s, c);
}

// ---------------------------------------------------------------------------
// Minv = M^-1 (upper triangle, Mi[r][c] for r <= c) by the articulated-body
// recursion (computeMinv, minv.h:423-478; equation numbers as in its comments).
// The pivots D_k and 1/D_k are formed in SP (FP64 in the pivot variant).
// ---------------------------------------------------------------------------
template <typename S, typename SP>
RBD_HD RBD_INLINE void compute_minv(const CS<S>* t, S Mi[NQ][NQ]) {
  S IA[6][6];
  SP IA22;                    // IA[2][2] (the pivot) carried in SP
  iload<NQ - 1>(IA);
  IA22 = SP(IA[2][2]);
  S F[NQ][6];                 // F[j] = F_k[:, j], columns j >= k meaningful
  S U[NQ][6], Dinv[NQ];
  for (int j = 0; j < NQ; ++j)
    for (int r = 0; r < 6; ++r) F[j][r] = S(0);
  static_for_down<NQ, 0>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    for (int r = 0; r < 6; ++r) U[k][r] = IA[r][2];        // U = IA S
    const SP dinv = SP(1) / IA22;                            // 1 / (S^T IA S)
    Dinv[k] = S(dinv);
    Mi[k][k] = Dinv[k];                                      // (21)
    static_for<k + 1, NQ>([&](auto Jc) {                     // (12)
      constexpr int j = decltype(Jc)::value;
      Mi[k][j] = -Dinv[k] * F[j][2];
    });
    static_for<k, NQ>([&](auto Jc) {                         // (14), (19)
      constexpr int j = decltype(Jc)::value;
      for (int r = 0; r < 6; ++r) F[j][r] += U[k][r] * Mi[k][j];
    });
    if constexpr (k > 0) {
      static_for<k, NQ>([&](auto Jc) {                       // (15): F_{k-1} = T^T F_k
        constexpr int j = decltype(Jc)::value;
        S y[6];
        xforce_t<k>(t[k], F[j], y);
        for (int r = 0; r < 6; ++r) F[j][r] = y[r];
      });
      S Ia[6][6];                                            // IA - U U^T / D
      S UD[6];
      for (int r = 0; r < 6; ++r) UD[r] = U[k][r] * Dinv[k];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c)
          Ia[r][c] = (r == 2 || c == 2) ? S(0) : IA[r][c] - U[k][r] * UD[c];
      S P[6][6];
      transport<k>(t[k], Ia, P);
      iload<k - 1>(IA);
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) IA[r][c] += P[r][c];
      if constexpr (sizeof(SP) > sizeof(S)) {
        // the next pivot S^T IA_{k-1} S formed in SP from the same (rounded) terms;
        // only entry (2,2) of this second transport is live, the rest is dead code
        SP IaW[6][6], PW[6][6];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) IaW[r][c] = SP(Ia[r][c]);
        const CS<SP> tw{SP(t[k].c), SP(t[k].s)};
        transport<k>(tw, IaW, PW);
        IA22 = SP(inertia_entry(k - 1, 2, 2)) + PW[2][2];
      } else {
        IA22 = SP(IA[2][2]);
      }
    }
  });
  // forward pass (Minv_fp, minv.h:439-469): F_k[:, j] = S Minv[k, j] + T F_{k-1}[:, j]
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if constexpr (k == 0) {
      for (int j = 0; j < NQ; ++j) {
        for (int r = 0; r < 6; ++r) F[j][r] = S(0);
        F[j][2] = Mi[0][j];
      }
    } else {
      static_for<k, NQ>([&](auto Jc) {
        constexpr int j = decltype(Jc)::value;
        S w[6];
        xmotion<k>(t[k], F[j], w);                           // T F_{k-1}[:, j]
        const S uw = U[k][0] * w[0] + U[k][1] * w[1] + U[k][2] * w[2] +
                     U[k][3] * w[3] + U[k][4] * w[4] + U[k][5] * w[5];
        Mi[k][j] -= Dinv[k] * uw;                            // (30)
        w[2] += Mi[k][j];                                    // (32) + (34)
        for (int r = 0; r < 6; ++r) F[j][r] = w[r];
      });
    }
  });
}

// Column j of Minv, for a lane that computes one column: col[k] = Minv[k][j]
// for k <= j (entries k > j are left unspecified).  The articulated inertias
// (U_k, D_k, the IA transports) are computed in full, the F and Minv updates
// for column j only -- the recursion of compute_minv restricted to one column.
// Through the backward pass column j's F stays exactly zero for every link
// k > j (its Minv entry there is -D_k^-1 * 0), so one code path serves every j.
template <typename S, typename SP>
RBD_HD RBD_INLINE void compute_minv_column(const CS<S>* t, int j, S col[NQ]) {
  S IA[6][6];
  SP IA22;
  iload<NQ - 1>(IA);
  IA22 = SP(IA[2][2]);
  S F[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
  S U[NQ][6], Dinv[NQ];
  static_for_down<NQ, 0>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    for (int r = 0; r < 6; ++r) U[k][r] = IA[r][2];
    const SP dinv = SP(1) / IA22;
    Dinv[k] = S(dinv);
    col[k] = (j == k) ? Dinv[k] : -Dinv[k] * F[2];          // (21), (12)
    for (int r = 0; r < 6; ++r) F[r] += U[k][r] * col[k];   // (14), (19)
    if constexpr (k > 0) {
      S y[6];
      xforce_t<k>(t[k], F, y);                               // (15)
      for (int r = 0; r < 6; ++r) F[r] = y[r];
      S Ia[6][6];
      S UD[6];
      for (int r = 0; r < 6; ++r) UD[r] = U[k][r] * Dinv[k];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c)
          Ia[r][c] = (r == 2 || c == 2) ? S(0) : IA[r][c] - U[k][r] * UD[c];
      S P[6][6];
      transport<k>(t[k], Ia, P);
      iload<k - 1>(IA);
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) IA[r][c] += P[r][c];
      if constexpr (sizeof(SP) > sizeof(S)) {
        SP IaW[6][6], PW[6][6];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) IaW[r][c] = SP(Ia[r][c]);
        const CS<SP> tw{SP(t[k].c), SP(t[k].s)};
        transport<k>(tw, IaW, PW);
        IA22 = SP(inertia_entry(k - 1, 2, 2)) + PW[2][2];
      } else {
        IA22 = SP(IA[2][2]);
      }
    }
  });
  // forward pass: F_k = S Minv[k, j] + T F_{k-1}
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if constexpr (k == 0) {
      for (int r = 0; r < 6; ++r) F[r] = S(0);
      F[2] = col[0];
    } else {
      S w[6];
      xmotion<k>(t[k], F, w);
      const S uw = U[k][0] * w[0] + U[k][1] * w[1] + U[k][2] * w[2] +
                   U[k][3] * w[3] + U[k][4] * w[4] + U[k][5] * w[5];
      col[k] -= Dinv[k] * uw;                                 // (30)
      w[2] += col[k];                                         // (32) + (34)
      for (int r = 0; r < 6; ++r) F[r] = w[r];
    }
  });
}

// ---------------------------------------------------------------------------
// RNEA forward pass (RNEA_fp, dynamicsGradient.h:619-659): v, a, f per link.
// The gravity trick sets the base acceleration to +g along z.
// ---------------------------------------------------------------------------
template <typename S>
RBD_HD RBD_INLINE void rnea_fwd(const CS<S>* t, const S* qd, const S* qdd, S gravity,
                                S v[NQ][6], S a[NQ][6], S f[NQ][6]) {
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if constexpr (k == 0) {
      for (int i = 0; i < 6; ++i) { v[0][i] = S(0); a[0][i] = S(0); }
      v[0][2] = qd[0];
      a[0][2] = qdd[0];
      const S g[3] = {S(0), S(0), gravity};
      S ga[3];
      rot<0>(t[0], g, ga);                   // T_0 [0,0,0,0,0,g] = [0; E g]
      a[0][3] = ga[0]; a[0][4] = ga[1]; a[0][5] = ga[2];
    } else {
      xmotion<k>(t[k], v[k - 1], v[k]);
      v[k][2] += qd[k];
      xmotion<k>(t[k], a[k - 1], a[k]);
      a[k][0] += v[k][1] * qd[k];            // + (v x S) qd
      a[k][1] -= v[k][0] * qd[k];
      a[k][3] += v[k][4] * qd[k];
      a[k][4] -= v[k][3] * qd[k];
      a[k][2] += qdd[k];
    }
    S Iv[6];
    imv<k>(v[k], Iv);
    imv<k>(a[k], f[k]);                      // f = I a + v x* I v
    crf<true>(v[k], Iv, f[k]);
  });
}

// RNEA backward pass (RNEA_bp, dynamicsGradient.h:661-670): f_{k-1} += T^T f_k
template <typename S>
RBD_HD RBD_INLINE void rnea_bwd(const CS<S>* t, S f[NQ][6]) {
  static_for_down<NQ, 1>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    S y[6];
    xforce_t_add<k>(t[k], f[k], f[k - 1], y);
    for (int i = 0; i < 6; ++i) f[k - 1][i] = y[i];
  });
}

// External forces (RNEA_fp's s_fext, dynamicsGradient.h:654-657): f_k -= fext_k
// with fext_k link k's spatial force in its own frame at fx[6k .. 6k+5].  The
// input accessor's fx is null (a compile-time constant in the plain kernels,
// so this folds away) unless the call passes forces.
template <class In, typename S>
RBD_HD RBD_INLINE void subtract_fext(const In& in, S f[NQ][6]) {
  if (in.fx)
    for (int k = 0; k < NQ; ++k)
      for (int i = 0; i < 6; ++i) f[k][i] -= S(in.fx[6 * k + i]);
}

// ---------------------------------------------------------------------------
// One derivative column.  Forward sweep (compute_vafdu, dynamicsGradient.h:786-829)
// from link c: at k == c the seed is dv = S, da = v_c x S for d/dqd, and
// dv = v_c x S, da = a_c x S for d/dq (the reference's (T v_{c-1}) x S and
// (v x S) qd + (T a_{c-1}) x S reduce to these because S x S = 0);
// for k > c: dv = T dv, da = T da + (dv x S) qd_k.  Then
// df_k = I da + dv x* (I v_k) + v_k x* (I dv).
// Backward sweep (compute_cdu, :831-857): g_{k-1} = df_{k-1} + T^T g_k, with the
// extra T^T (-(f_k x S)) at k == c for d/dq; dtau_k = g_k[2].
// ---------------------------------------------------------------------------
template <typename S>
struct PhaseC {
  CS<S> t[NQ];
  S v[NQ][6];      // link velocities
  S Iv[NQ][6];     // I_k v_k
  S axs[NQ][4];    // a_k x S  -> (a1, -a0, a4, -a3)   (entries 2 and 5 are zero)
  S nfxs[NQ][4];   // -(f_k x S) -> (-f1, f0, -f4, f3)
  S qd[NQ];
  S Mi[NQ][NQ];    // upper triangle of Minv
};

// dtau[:, c] for the d/dqd (DQ == false) or d/dq (DQ == true) column c.
template <bool DQ, typename S>
RBD_HD RBD_INLINE void dtau_column(const PhaseC<S>& P, int c, S tau[NQ]) {
  S dv[6], da[6], df[NQ][6];
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if (k < c) {
      for (int i = 0; i < 6; ++i) df[k][i] = S(0);
      return;
    }
    if (k == c) {
      if constexpr (DQ) {
        dv[0] = P.v[k][1]; dv[1] = -P.v[k][0]; dv[2] = S(0);
        dv[3] = P.v[k][4]; dv[4] = -P.v[k][3]; dv[5] = S(0);
        da[0] = P.axs[k][0]; da[1] = P.axs[k][1]; da[2] = S(0);
        da[3] = P.axs[k][2]; da[4] = P.axs[k][3]; da[5] = S(0);
      } else {
        for (int i = 0; i < 6; ++i) dv[i] = S(0);
        dv[2] = S(1);
        da[0] = P.v[k][1]; da[1] = -P.v[k][0]; da[2] = S(0);
        da[3] = P.v[k][4]; da[4] = -P.v[k][3]; da[5] = S(0);
      }
    } else {
      if constexpr (k > 0) {
        S y[6];
        xmotion<k>(P.t[k], dv, y);
        for (int i = 0; i < 6; ++i) dv[i] = y[i];
        xmotion<k>(P.t[k], da, y);
        const S qk = P.qd[k];
        da[0] = y[0] + dv[1] * qk; da[1] = y[1] - dv[0] * qk; da[2] = y[2];
        da[3] = y[3] + dv[4] * qk; da[4] = y[4] - dv[3] * qk; da[5] = y[5];
      }
    }
    S Idv[6];
    imv<k>(dv, Idv);
    imv<k>(da, df[k]);
    crf<true>(dv, P.Iv[k], df[k]);
    crf<true>(P.v[k], Idv, df[k]);
  });
  S g[6];
  for (int i = 0; i < 6; ++i) g[i] = df[NQ - 1][i];
  static_for_down<NQ, 1>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    tau[k] = g[2];
    if constexpr (DQ) {
      if (k == c) {
        g[0] += P.nfxs[k][0]; g[1] += P.nfxs[k][1];
        g[3] += P.nfxs[k][2]; g[4] += P.nfxs[k][3];
      }
    }
    S y[6];
    if (k - 1 >= c) xforce_t_add<k>(P.t[k], g, df[k - 1], y);
    else            xforce_t<k>(P.t[k], g, y);
    for (int i = 0; i < 6; ++i) g[i] = y[i];
  });
  tau[0] = g[2];
}

// tau[c] += x for a run-time c without indexing a register array
template <typename S>
RBD_HD RBD_INLINE void add_on_diagonal(S tau[NQ], int c, S x) {
  static_for<0, NQ>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    tau[r] += (r == c) ? x : S(0);
  });
}

// out[r] = (Minv x)[r] (NEG: negated), Minv upper-stored
template <bool NEG, typename S, typename SX>
RBD_HD RBD_INLINE void minv_times(const S Mi[NQ][NQ], const SX x[NQ], SX out[NQ]) {
  static_for<0, NQ>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    SX acc = SX(0);
    static_for<0, NQ>([&](auto Cc) {
      constexpr int c = decltype(Cc)::value;
      acc += SX(r <= c ? Mi[r][c] : Mi[c][r]) * x[c];
    });
    out[r] = NEG ? -acc : acc;
  });
}

// Uniform-sweep variant of dtau_column for lane-parallel use: the column c and
// its kind (dq) may differ from lane to lane while every lane executes the same
// instructions.  dv, da start at zero and stay zero until k == c, where the seed
// replaces them by select; the zero df_k for k < c then adds exactly nothing in
// the backward sweep, so the result equals dtau_column's.
template <typename S>
RBD_HD RBD_INLINE void dtau_column_uniform(const PhaseC<S>& P, int c, bool dq, S tau[NQ]) {
  S dv[6], da[6], df[NQ][6];
  for (int i = 0; i < 6; ++i) { dv[i] = S(0); da[i] = S(0); }
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    S pv[6], pa[6];
    if constexpr (k > 0) {
      xmotion<k>(P.t[k], dv, pv);
      xmotion<k>(P.t[k], da, pa);
      const S qk = P.qd[k];
      pa[0] += pv[1] * qk; pa[1] -= pv[0] * qk; pa[3] += pv[4] * qk; pa[4] -= pv[3] * qk;
    } else {
      for (int i = 0; i < 6; ++i) { pv[i] = S(0); pa[i] = S(0); }
    }
    // seeds: d/dqd -> dv = S, da = v_k x S;  d/dq -> dv = v_k x S, da = a_k x S
    const S vx[6] = {P.v[k][1], -P.v[k][0], S(0), P.v[k][4], -P.v[k][3], S(0)};
    const S ax[6] = {P.axs[k][0], P.axs[k][1], S(0), P.axs[k][2], P.axs[k][3], S(0)};
    const bool seed = (k == c);
    for (int i = 0; i < 6; ++i) {
      const S sv = dq ? vx[i] : (i == 2 ? S(1) : S(0));
      const S sa = dq ? ax[i] : vx[i];
      dv[i] = seed ? sv : pv[i];
      da[i] = seed ? sa : pa[i];
    }
    S Idv[6];
    imv<k>(dv, Idv);
    imv<k>(da, df[k]);
    crf<true>(dv, P.Iv[k], df[k]);
    crf<true>(P.v[k], Idv, df[k]);
  });
  S g[6];
  for (int i = 0; i < 6; ++i) g[i] = df[NQ - 1][i];
  static_for_down<NQ, 1>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    tau[k] = g[2];
    const bool extra = dq && (k == c);
    g[0] += extra ? P.nfxs[k][0] : S(0); g[1] += extra ? P.nfxs[k][1] : S(0);
    g[3] += extra ? P.nfxs[k][2] : S(0); g[4] += extra ? P.nfxs[k][3] : S(0);
    S y[6];
    xforce_t_add<k>(P.t[k], g, df[k - 1], y);
    for (int i = 0; i < 6; ++i) g[i] = y[i];
  });
  tau[0] = g[2];
}

// Phase-C inputs from v, a (with qdd), f (after the backward pass) and Minv.
template <typename SC, typename SR, typename SM>
RBD_HD RBD_INLINE void make_phase_c(PhaseC<SC>& P, const CS<SC>* t, const SR* qd, const SR v[NQ][6],
                                    const SR a[NQ][6], const SR f[NQ][6], const SM Mi[NQ][NQ]) {
  for (int k = 0; k < NQ; ++k) { P.t[k] = t[k]; P.qd[k] = SC(qd[k]); }
  for (int r = 0; r < NQ; ++r)
    for (int c = r; c < NQ; ++c) P.Mi[r][c] = SC(Mi[r][c]);
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    for (int i = 0; i < 6; ++i) P.v[k][i] = SC(v[k][i]);
    imv<k>(P.v[k], P.Iv[k]);
    P.axs[k][0] = SC(a[k][1]); P.axs[k][1] = -SC(a[k][0]); P.axs[k][2] = SC(a[k][4]); P.axs[k][3] = -SC(a[k][3]);
    P.nfxs[k][0] = -SC(f[k][1]); P.nfxs[k][1] = SC(f[k][0]); P.nfxs[k][2] = -SC(f[k][4]); P.nfxs[k][3] = SC(f[k][3]);
  });
}

// The d/dq columns in one lane (one lane per knot) and their stores.  MINV =
// false stores the inverse-dynamics gradient dtau/dq itself (the reference's
// cdu) instead of -Minv times it.
template <bool MINV = true, typename SC, class Out>
RBD_HD RBD_INLINE void gradient_dq_columns(const PhaseC<SC>& P, Out& out) {
  // d/dq column 0 is identically zero (rotation about the gravity axis)
  for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  static_for<1, NQ>([&](auto Cc) {
    constexpr int c = decltype(Cc)::value;
    SC tau[NQ], col[NQ];
    dtau_column<true>(P, c, tau);
    if constexpr (MINV) minv_times<true>(P.Mi, tau, col);
    else for (int r = 0; r < NQ; ++r) col[r] = tau[r];
    for (int r = 0; r < NQ; ++r) out.dq(r, c, float(col[r]));
  });
}

// The d/dqd columns in one lane and their stores (damping on the diagonal).
template <bool MINV = true, typename SC, class Out>
RBD_HD RBD_INLINE void gradient_dqd_columns(const PhaseC<SC>& P, float damping, Out& out) {
  static_for<0, NQ>([&](auto Cc) {
    constexpr int c = decltype(Cc)::value;
    SC tau[NQ], col[NQ];
    dtau_column<false>(P, c, tau);
    add_on_diagonal(tau, c, SC(damping));
    if constexpr (MINV) minv_times<true>(P.Mi, tau, col);
    else for (int r = 0; r < NQ; ++r) col[r] = tau[r];
    for (int r = 0; r < NQ; ++r) out.dqd(r, c, float(col[r]));
  });
}

// All derivative columns in one lane and the stores of dqdd/dq, dqdd/dqd: the
// d/dq columns, then the d/dqd columns -- the order of the outputs in memory,
// so that each knot's AoS record fills front to back and few of its cache
// lines are partly written at a time.
template <bool MINV = true, typename SC, class Out>
RBD_HD RBD_INLINE void gradient_columns(const PhaseC<SC>& P, float damping, Out& out) {
  gradient_dq_columns<MINV>(P, out);
  gradient_dqd_columns<MINV>(P, damping, out);
}

// Lane roles when one knot is spread over a group of kLanesPerKnotGroup lanes:
// role j < 7: d/dqd column j; 7 <= j < 13: d/dq column j - 6; j == 13: qdd and
// the zero d/dq column 0; j >= 14: no stores.
constexpr int kLanesPerKnotGroup = 16;
constexpr int kRoleQdd = 13;

// One derivative column for lane role `role` of a knot group and its store
// (MINV as in gradient_columns).
template <bool MINV = true, typename SC, class Out>
RBD_HD RBD_INLINE void gradient_column_lane(const PhaseC<SC>& P, float damping, int role, Out& out) {
  const bool dq = role >= NQ;
  const int c = dq ? role - (NQ - 1) : role;
  SC tau[NQ], col[NQ];
  dtau_column_uniform(P, c, dq, tau);
  add_on_diagonal(tau, dq ? -1 : c, SC(damping));
  if constexpr (MINV) minv_times<true>(P.Mi, tau, col);
  else for (int r = 0; r < NQ; ++r) col[r] = tau[r];
  if (role < NQ) {
    for (int r = 0; r < NQ; ++r) out.dqd(r, c, float(col[r]));
  } else if (role < 2 * NQ - 1) {
    for (int r = 0; r < NQ; ++r) out.dq(r, c, float(col[r]));
  }
}

// Largest of two scalar types
template <typename A, typename B>
using Wider = typename std::conditional<(sizeof(A) >= sizeof(B)), A, B>::type;

// ---------------------------------------------------------------------------
// Accessors.  In: q(i), qd(i), u(i) [, qdd(i), minv(r, c) in the reference-
// compatible mode].  Out: qdd(i, x), dq(r, c, x), dqd(r, c, x).  All FP32.
// ---------------------------------------------------------------------------

// Run-time model options (the reference's compile-time VEL_DAMPING and MPC_MODE),
// and the bound on the two-wave mapping's LDS handoff waits.
struct Options {
  float damping = 0.0f;               // joint viscous damping: 0.5 with VEL_DAMPING (dynamicsGradient.h:680,840)
  float gravity = float(kGravity);    // 0 in MPC_MODE (dynamicsGradient.h:635-637)
  // A handoff wait polls its flag at most handoff_spins + 1 times, with a
  // short sleep between polls (2^24 sleeps of 64 cycles: about half a second);
  // a negative value makes every wait give up at once.  A wait that gives up
  // writes 1 to *fault, a word the host can read (null: nowhere), and the
  // kernel carries on and exits normally; its results are then invalid.
  unsigned* fault = nullptr;
  int handoff_spins = 1 << 24;
};

// The bounded wait of a handoff, for any poll and pause: true when poll()
// returned true within spins + 1 polls; false at once for spins < 0.
template <class Poll, class Pause>
RBD_HD RBD_INLINE bool bounded_wait(Poll&& poll, Pause&& pause, int spins) {
  if (spins < 0) return false;
  for (int spin = 0;; ++spin) {
    if (poll()) return true;
    if (spin >= spins) return false;
    pause();
  }
}

// sin/cos type of the full path: computed once in the widest arithmetic type
template <class P>
using KnotTrig = Wider<typename P::Minv, Wider<typename P::Rnea, typename P::Col>>;

template <class P, class In>
RBD_HD RBD_INLINE void knot_trig(const In& in, CS<KnotTrig<P>> tw[NQ]) {
  using ST = KnotTrig<P>;
  for (int k = 0; k < NQ; ++k) {
    ST s, c;
    knot_sincos<P>(ST(in.q(k)), &s, &c);
    tw[k] = CS<ST>{c, s};
  }
}

// The same for a knot's lane group: lane role j < 7 computes the sin/cos of
// q_j alone and the group's lanes read all seven from each other by
// shuffles, instead of every lane computing all seven (a sincos is about 100
// instructions).  The values are those of knot_trig; the host build, which
// runs the lane roles one after another, calls knot_trig.
template <class P, class In>
RBD_HD RBD_INLINE void knot_trig_group(const In& in, int role, CS<KnotTrig<P>> tw[NQ]) {
#if defined(__HIP_DEVICE_COMPILE__)
  using ST = KnotTrig<P>;
  ST s, c;
  knot_sincos<P>(ST(in.q(role < NQ ? role : 0)), &s, &c);
  for (int k = 0; k < NQ; ++k) tw[k] = CS<ST>{__shfl(c, k, kLanesPerKnotGroup), __shfl(s, k, kLanesPerKnotGroup)};
#else
  (void)role;
  knot_trig<P>(in, tw);
#endif
}

// Steps 2-5 of the full path, given sin/cos and Minv: bias RNEA, qdd = Minv
// (u - c), RNEA with qdd; fills the phase-C inputs and qdd.
template <class P, class In>
RBD_HD RBD_INLINE void phase_a_after_minv(const In& in, const Options& opt, const CS<KnotTrig<P>> tw[NQ],
                                          const typename P::Minv Mi[NQ][NQ], PhaseC<typename P::Col>& PC,
                                          typename P::Rnea qdd[NQ]) {
  using SR = typename P::Rnea; using SC = typename P::Col;
  CS<SR> tr[NQ]; CS<SC> tc[NQ];
  SR qd[NQ], u[NQ];
  for (int k = 0; k < NQ; ++k) {
    tr[k] = CS<SR>{SR(tw[k].c), SR(tw[k].s)}; tc[k] = CS<SC>{SC(tw[k].c), SC(tw[k].s)};
    qd[k] = SR(in.qd(k));
    u[k] = SR(in.u(k));
  }
  // bias: RNEA with qdd = 0
  SR v[NQ][6], a[NQ][6], f[NQ][6];
  const SR zero[NQ] = {SR(0), SR(0), SR(0), SR(0), SR(0), SR(0), SR(0)};
  rnea_fwd(tr, qd, zero, SR(opt.gravity), v, a, f);
  subtract_fext(in, f);
  rnea_bwd(tr, f);
  SR bias[NQ];
  for (int k = 0; k < NQ; ++k) bias[k] = u[k] - f[k][2] - SR(opt.damping) * qd[k];
  minv_times<false>(Mi, bias, qdd);
  // RNEA with qdd, added incrementally: a += da, f += bwd(I da), da_k = T da_{k-1} + S qdd_k
  {
    SR dA[6] = {SR(0), SR(0), qdd[0], SR(0), SR(0), SR(0)};
    SR dF[NQ][6];
    static_for<0, NQ>([&](auto Kc) {
      constexpr int k = decltype(Kc)::value;
      if constexpr (k > 0) {
        SR y[6];
        xmotion<k>(tr[k], dA, y);
        for (int i = 0; i < 6; ++i) dA[i] = y[i];
        dA[2] += qdd[k];
      }
      for (int i = 0; i < 6; ++i) a[k][i] += dA[i];
      imv<k>(dA, dF[k]);
    });
    rnea_bwd(tr, dF);
    for (int k = 0; k < NQ; ++k)
      for (int i = 0; i < 6; ++i) f[k][i] += dF[k][i];
  }
  make_phase_c(PC, tc, qd, v, a, f, Mi);
}

// Steps 1-5 of the full path: sin/cos, Minv, bias RNEA, qdd = Minv (u - c), RNEA
// with qdd; fills the phase-C inputs and qdd.
template <class P, class In>
RBD_HD RBD_INLINE void phase_a(const In& in, const Options& opt, PhaseC<typename P::Col>& PC,
                               typename P::Rnea qdd[NQ]) {
  using SM = typename P::Minv; using SP = typename P::Pivot;
  CS<KnotTrig<P>> tw[NQ];
  knot_trig<P>(in, tw);
  CS<SM> tm[NQ];
  for (int k = 0; k < NQ; ++k) tm[k] = CS<SM>{SM(tw[k].c), SM(tw[k].s)};
  SM Mi[NQ][NQ];
  compute_minv<SM, SP>(tm, Mi);
  phase_a_after_minv<P>(in, opt, tw, Mi, PC, qdd);
}

// Steps 1 and 5 of the reference-compatible modes: qdd given, and with
// WITH_MINV the dense Minv too (otherwise PC.Mi is left zero and unused).
template <class P, bool WITH_MINV = true, class In>
RBD_HD RBD_INLINE void phase_a_given_qdd_minv(const In& in, const Options& opt, PhaseC<typename P::Col>& PC) {
  using SR = typename P::Rnea; using SC = typename P::Col;
  using ST = Wider<SR, SC>;
  CS<SR> tr[NQ]; CS<SC> tc[NQ];
  SR qd[NQ], qdd[NQ];
  for (int k = 0; k < NQ; ++k) {
    ST s, c;
    knot_sincos<P>(ST(in.q(k)), &s, &c);
    tr[k] = CS<SR>{SR(c), SR(s)}; tc[k] = CS<SC>{SC(c), SC(s)};
    qd[k] = SR(in.qd(k));
    qdd[k] = SR(in.qdd(k));
  }
  SC Mi[NQ][NQ];
  for (int r = 0; r < NQ; ++r)
    for (int c = r; c < NQ; ++c) {
      if constexpr (WITH_MINV) Mi[r][c] = SC(in.minv(r, c));
      else Mi[r][c] = SC(0);
    }
  SR v[NQ][6], a[NQ][6], f[NQ][6];
  rnea_fwd(tr, qd, qdd, SR(opt.gravity), v, a, f);
  subtract_fext(in, f);
  rnea_bwd(tr, f);
  make_phase_c(PC, tc, qd, v, a, f, Mi);
}

// Output accessors that also want dqdd/du = Minv (dense) declare kMinv = true.
template <class Out, class = void> struct WantsMinv { static constexpr bool value = false; };
template <class Out> struct WantsMinv<Out, typename std::enable_if<Out::kMinv>::type> {
  static constexpr bool value = true;
};

// column c of the symmetric Minv (upper-stored) to out.minv
template <class Out, typename S>
RBD_HD RBD_INLINE void store_minv_column(Out& out, const S Mi[NQ][NQ], int c) {
  static_for<0, NQ>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    S m = S(0);
    static_for<0, NQ>([&](auto Cc) {
      constexpr int cc = decltype(Cc)::value;
      m = (cc == c) ? (r <= cc ? Mi[r][cc] : Mi[cc][r]) : m;
    });
    out.minv(r, c, float(m));
  });
}

// Full path for one knot in one lane.  P: arithmetic precision per phase.
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void fdgrad_knot(const In& in, Out& out, const Options& opt) {
  PhaseC<typename P::Col> PC;
  typename P::Rnea qdd[NQ];
  phase_a<P>(in, opt, PC, qdd);
  for (int r = 0; r < NQ; ++r) out.qdd(r, float(qdd[r]));
  if constexpr (WantsMinv<Out>::value)
    for (int r = 0; r < NQ; ++r)
      for (int c = 0; c < NQ; ++c) out.minv(r, c, float(r <= c ? PC.Mi[r][c] : PC.Mi[c][r]));
  gradient_columns(PC, opt.damping, out);
}

// Forward dynamics only: qdd = FD(q, qd, u) and the dense Minv, one lane per
// knot (forwardDynamics with DENSE_MINV, dynamicsGradient.h:685-706, :936-939,
// with the bias RNEA at qdd = 0).
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void fd_knot(const In& in, Out& out, const Options& opt) {
  using SM = typename P::Minv; using SP = typename P::Pivot; using SR = typename P::Rnea;
  using ST = Wider<SM, SR>;
  CS<SM> tm[NQ]; CS<SR> tr[NQ];
  SR qd[NQ], u[NQ];
  for (int k = 0; k < NQ; ++k) {
    ST s, c;
    knot_sincos<P>(ST(in.q(k)), &s, &c);
    tm[k] = CS<SM>{SM(c), SM(s)}; tr[k] = CS<SR>{SR(c), SR(s)};
    qd[k] = SR(in.qd(k));
    u[k] = SR(in.u(k));
  }
  SM Mi[NQ][NQ];
  compute_minv<SM, SP>(tm, Mi);
  SR v[NQ][6], a[NQ][6], f[NQ][6];
  const SR zero[NQ] = {SR(0), SR(0), SR(0), SR(0), SR(0), SR(0), SR(0)};
  rnea_fwd(tr, qd, zero, SR(opt.gravity), v, a, f);
  subtract_fext(in, f);
  rnea_bwd(tr, f);
  SR bias[NQ], qdd[NQ];
  for (int k = 0; k < NQ; ++k) bias[k] = u[k] - f[k][2] - SR(opt.damping) * qd[k];
  minv_times<false>(Mi, bias, qdd);
  for (int r = 0; r < NQ; ++r) out.qdd(r, float(qdd[r]));
  for (int r = 0; r < NQ; ++r)
    for (int c = 0; c < NQ; ++c) out.minv(r, c, float(r <= c ? Mi[r][c] : Mi[c][r]));
}

// Minv of a knot shared by its lane group: the packed upper triangle
// (upper_index), in LDS in the kernels.
constexpr int kMinvPacked = NQ * (NQ + 1) / 2;

// index of (r, c), r <= c, in the packed upper triangle
RBD_HD constexpr int upper_index(int r, int c) { return r * NQ - r * (r - 1) / 2 + (c - r); }

// Lanes of one wave see each other's LDS writes made before this point.  The
// fences order the LDS accesses of the wave (they cost no instruction); in the
// host build, which runs the lane roles of a stage one after another, there
// is nothing to do.
RBD_HD RBD_INLINE void wave_lds_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#endif
}

// Lane-group stage 1: sin/cos (kept in tw for stage 2) and lane role's column
// of Minv (roles 0..6), its upper entries stored into the group's mi.
template <class P, class In>
RBD_HD RBD_INLINE void lanes_minv_stage(const In& in, int role, CS<KnotTrig<P>> tw[NQ], typename P::Minv* mi) {
  using SM = typename P::Minv; using SP = typename P::Pivot;
  knot_trig_group<P>(in, role, tw);
  CS<SM> tm[NQ];
  for (int k = 0; k < NQ; ++k) tm[k] = CS<SM>{SM(tw[k].c), SM(tw[k].s)};
  const int j = role < NQ ? role : NQ - 1;
  SM col[NQ];
  compute_minv_column<SM, SP>(tm, j, col);
  if (role < NQ)
    for (int k = 0; k < NQ; ++k)
      if (k <= j) mi[upper_index(k, j)] = col[k];
}

// Minv (upper triangle) from a group's packed mi
template <typename S>
RBD_HD RBD_INLINE void load_packed_minv(const S* mi, S Mi[NQ][NQ]) {
  static_for<0, NQ>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    static_for<r, NQ>([&](auto Cc) {
      constexpr int c = decltype(Cc)::value;
      Mi[r][c] = mi[upper_index(r, c)];
    });
  });
}

// Lane-group stage 2 (after every lane's stage 1): steps 2-5 with the shared
// Minv (identical values in every lane), then lane `role` computes one
// derivative column (steps 6-7) and stores it.
template <class P, class In, class Out>
RBD_HD RBD_INLINE void lanes_finish_stage(const In& in, Out& out, const Options& opt, int role,
                                          const CS<KnotTrig<P>> tw[NQ], const typename P::Minv* mi) {
  typename P::Minv Mi[NQ][NQ];
  load_packed_minv(mi, Mi);
  PhaseC<typename P::Col> PC;
  typename P::Rnea qdd[NQ];
  phase_a_after_minv<P>(in, opt, tw, Mi, PC, qdd);
  if (role == kRoleQdd) {
    for (int r = 0; r < NQ; ++r) out.qdd(r, float(qdd[r]));
    for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  }
  if constexpr (WantsMinv<Out>::value)
    if (role < NQ) store_minv_column(out, PC.Mi, role);
  gradient_column_lane(PC, opt.damping, role, out);
}

// Full path for one knot spread over kLanesPerKnotGroup lanes of one wave:
// lanes 0..6 each compute one column of Minv and share it through the group's
// LDS slot mi (kMinvPacked values), then every lane runs steps 2-5 and one
// derivative column.  Device code only: the host build runs the two stages
// for every lane role in turn.
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void fdgrad_knot_lanes(const In& in, Out& out, const Options& opt, int role,
                                         typename P::Minv* mi) {
  CS<KnotTrig<P>> tw[NQ];
  lanes_minv_stage<P>(in, role, tw, mi);
  wave_lds_sync();
  lanes_finish_stage<P>(in, out, opt, role, tw, mi);
}

// ---------------------------------------------------------------------------
// Two-wave latency mapping.  For one knot, one wave (M) runs the Minv
// recursion, lane j of the knot's group computing column j, while a second
// wave (R) runs the RNEA at qdd = 0 and all thirteen derivative columns with
// the bias accelerations and forces a_b, f_b; the two overlap instead of
// adding up.  The columns
// need a, f at the true qdd only in the d/dq seeds (a_c x S at link c) and the
// backward extra term (f_c at link c), in which a d/dq column is linear.  So
// with da = a - a_b and df = f - f_b, both linear in qdd:
//   dtau/dq_c = column(a_b, f_b) + column'(da_c x S, df_c)
// where column' is the sweep with v = 0 (no velocity terms, ~400
// instructions).  After M has Minv it forms qdd = Minv (u - c - damping qd),
// the chain da_k = T_k da_{k-1} + S qdd_k, dF_k = I_k da_k, df = RNEA
// backward pass of dF, and each d/dq lane's correction column'; R then adds
// the corrections and forms -Minv dtau.  Stage functions are per (knot, lane
// role); wave M runs M1a (its Minv column into LDS), M1b (the whole Minv back
// from LDS), waits for R1a's bias, runs M2 and signals; wave R runs R1a,
// signals, runs R1b, waits for M2, runs R2.  The critical path is the longer
// of M1 + M2 and R1, plus R2.
// ---------------------------------------------------------------------------

// Exchange between the waves for one knot (LDS in the kernel).
template <class P>
struct TwoWaveShared {
  typename P::Rnea bias[NQ];                        // u - c - damping qd (R1 -> M2)
  typename P::Minv mi[kMinvPacked];                 // upper Minv, packed (M1a -> M1b, R2)
  typename P::Col corr[kLanesPerKnotGroup][NQ];     // d/dq corrections by lane role (M2 -> R2)
};

// column'(da_c x S, df_c) for lane role's d/dq column c, uniform over lanes
// like dtau_column_uniform: dax[k] = (da_k x S) as (a1, -a0, a4, -a3),
// ndfx[k] = -(df_k x S) as (-f1, f0, -f4, f3).
template <typename S>
RBD_HD RBD_INLINE void dq_correction_uniform(const CS<S>* t, const S dax[NQ][4], const S ndfx[NQ][4], int c,
                                             S tau[NQ]) {
  S da[6], df[NQ][6];
  for (int i = 0; i < 6; ++i) da[i] = S(0);
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    S pa[6];
    if constexpr (k > 0) {
      xmotion<k>(t[k], da, pa);
    } else {
      for (int i = 0; i < 6; ++i) pa[i] = S(0);
    }
    const S ax[6] = {dax[k][0], dax[k][1], S(0), dax[k][2], dax[k][3], S(0)};
    const bool seed = (k == c);
    for (int i = 0; i < 6; ++i) da[i] = seed ? ax[i] : pa[i];
    imv<k>(da, df[k]);
  });
  S g[6];
  for (int i = 0; i < 6; ++i) g[i] = df[NQ - 1][i];
  static_for_down<NQ, 1>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    tau[k] = g[2];
    const bool extra = (k == c);
    g[0] += extra ? ndfx[k][0] : S(0); g[1] += extra ? ndfx[k][1] : S(0);
    g[3] += extra ? ndfx[k][2] : S(0); g[4] += extra ? ndfx[k][3] : S(0);
    S y[6];
    xforce_t_add<k>(t[k], g, df[k - 1], y);
    for (int i = 0; i < 6; ++i) g[i] = y[i];
  });
  tau[0] = g[2];
}

// Wave R's state between its stages.
template <class P>
struct TwoWaveR {
  using SR = typename P::Rnea; using SC = typename P::Col;
  CS<SC> tc[NQ];
  SR qd[NQ], v[NQ][6], a[NQ][6], f[NQ][6];
};

// Stage R1a (wave R): RNEA at qdd = 0 and the bias into sh.
template <class P, class In>
RBD_HD RBD_INLINE void two_wave_r1a(const In& in, const Options& opt, int role, TwoWaveShared<P>& sh,
                                    TwoWaveR<P>& st) {
  using SR = typename P::Rnea; using SC = typename P::Col;
  CS<KnotTrig<P>> tw[NQ];
  knot_trig_group<P>(in, role, tw);
  CS<SR> tr[NQ];
  SR u[NQ];
  for (int k = 0; k < NQ; ++k) {
    tr[k] = CS<SR>{SR(tw[k].c), SR(tw[k].s)}; st.tc[k] = CS<SC>{SC(tw[k].c), SC(tw[k].s)};
    st.qd[k] = SR(in.qd(k));
    u[k] = SR(in.u(k));
  }
  const SR zero[NQ] = {SR(0), SR(0), SR(0), SR(0), SR(0), SR(0), SR(0)};
  rnea_fwd(tr, st.qd, zero, SR(opt.gravity), st.v, st.a, st.f);
  subtract_fext(in, st.f);
  rnea_bwd(tr, st.f);
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if (role == k) sh.bias[k] = u[k] - st.f[k][2] - SR(opt.damping) * st.qd[k];
  });
}

// Stage R1b (wave R): lane role's column with a_b, f_b (damping on the
// d/dqd diagonal) into tau_col.
template <class P>
RBD_HD RBD_INLINE void two_wave_r1b(const Options& opt, int role, const TwoWaveR<P>& st,
                                    typename P::Col tau_col[NQ]) {
  using SR = typename P::Rnea; using SC = typename P::Col;
  SR Mz[NQ][NQ];
  for (int r = 0; r < NQ; ++r)
    for (int c = 0; c < NQ; ++c) Mz[r][c] = SR(0);
  PhaseC<SC> PC;
  make_phase_c(PC, st.tc, st.qd, st.v, st.a, st.f, Mz);
  const bool dq = role >= NQ;
  const int c = dq ? role - (NQ - 1) : role;
  dtau_column_uniform(PC, c, dq, tau_col);
  add_on_diagonal(tau_col, dq ? -1 : c, SC(opt.damping));
}

// sin/cos type of wave M (shared by its two stages)
template <class P>
using TwoWaveTrig = KnotTrig<P>;

// Stage M1a (wave M): sin/cos (kept for M2) and lane role's column of Minv
// (roles 0..6), its upper entries stored into sh.mi.
template <class P, class In>
RBD_HD RBD_INLINE void two_wave_m1a(const In& in, int role, CS<TwoWaveTrig<P>> tw[NQ], TwoWaveShared<P>& sh) {
  lanes_minv_stage<P>(in, role, tw, sh.mi);
}

// Stage M1b (wave M, after every lane's M1a): the whole upper Minv from sh.mi.
template <class P>
RBD_HD RBD_INLINE void two_wave_m1b(const TwoWaveShared<P>& sh, typename P::Minv Mi[NQ][NQ]) {
  load_packed_minv(sh.mi, Mi);
}

// Stage M2 (wave M, after R1): qdd (stored by lane role kRoleQdd), the
// qdd-dependent chain and lane role's d/dq correction into sh.
template <class P, class Out>
RBD_HD RBD_INLINE void two_wave_m2(int role, const CS<TwoWaveTrig<P>> tw[NQ], const typename P::Minv Mi[NQ][NQ],
                                   TwoWaveShared<P>& sh, Out& out) {
  using SR = typename P::Rnea; using SC = typename P::Col;
  CS<SR> tr[NQ]; CS<SC> tc[NQ];
  for (int k = 0; k < NQ; ++k) {
    tr[k] = CS<SR>{SR(tw[k].c), SR(tw[k].s)};
    tc[k] = CS<SC>{SC(tw[k].c), SC(tw[k].s)};
  }
  SR bias[NQ], qdd[NQ];
  for (int k = 0; k < NQ; ++k) bias[k] = sh.bias[k];
  minv_times<false>(Mi, bias, qdd);
  if (role == kRoleQdd)
    for (int r = 0; r < NQ; ++r) out.qdd(r, float(qdd[r]));
  // da_k = T_k da_{k-1} + S qdd_k, dF_k = I_k da_k, then the backward pass
  SR dA[NQ][6], dF[NQ][6];
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    if constexpr (k == 0) {
      for (int i = 0; i < 6; ++i) dA[0][i] = SR(0);
    } else {
      xmotion<k>(tr[k], dA[k - 1], dA[k]);
    }
    dA[k][2] += qdd[k];
    imv<k>(dA[k], dF[k]);
  });
  rnea_bwd(tr, dF);
  SC dax[NQ][4], ndfx[NQ][4];
  for (int k = 0; k < NQ; ++k) {
    dax[k][0] = SC(dA[k][1]); dax[k][1] = -SC(dA[k][0]); dax[k][2] = SC(dA[k][4]); dax[k][3] = -SC(dA[k][3]);
    ndfx[k][0] = -SC(dF[k][1]); ndfx[k][1] = SC(dF[k][0]); ndfx[k][2] = -SC(dF[k][4]); ndfx[k][3] = SC(dF[k][3]);
  }
  SC corr[NQ];
  dq_correction_uniform(tc, dax, ndfx, role - (NQ - 1), corr);
  for (int r = 0; r < NQ; ++r) sh.corr[role][r] = corr[r];
}

// Stage R2 (wave R, after M2): corrections added, -Minv dtau, stores.
template <class P, class Out>
RBD_HD RBD_INLINE void two_wave_r2(int role, const TwoWaveShared<P>& sh, typename P::Col tau_col[NQ], Out& out) {
  using SC = typename P::Col;
  const bool dq = role >= NQ;
  const int c = dq ? role - (NQ - 1) : role;
  if (dq)
    for (int r = 0; r < NQ; ++r) tau_col[r] += sh.corr[role][r];
  SC Mi[NQ][NQ];
  static_for<0, NQ>([&](auto Rc) {
    constexpr int r = decltype(Rc)::value;
    static_for<r, NQ>([&](auto Cc) {
      constexpr int cc = decltype(Cc)::value;
      Mi[r][cc] = SC(sh.mi[upper_index(r, cc)]);
    });
  });
  SC col[NQ];
  minv_times<true>(Mi, tau_col, col);
  if (role < NQ) {
    for (int r = 0; r < NQ; ++r) out.dqd(r, c, float(col[r]));
  } else if (role < 2 * NQ - 1) {
    for (int r = 0; r < NQ; ++r) out.dq(r, c, float(col[r]));
  } else if (role == kRoleQdd) {
    for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  }
}

// Reference-compatible mode (QDD_MINV_PASSED_IN, the reference's timed
// configuration; dynamicsGradientKernel, helpers_GPU/dynamicsGradient.cuh:1257-1291):
// qdd and a dense Minv are inputs; out is dqdd/dq, dqdd/dqd.  The upper triangle
// of the given Minv is read (the same matrix for the symmetric Minv that
// forwardDynamics produces).
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void fdgrad_knot_given_qdd_minv(const In& in, Out& out, const Options& opt) {
  PhaseC<typename P::Col> PC;
  phase_a_given_qdd_minv<P>(in, opt, PC);
  gradient_columns(PC, opt.damping, out);
}

// The same spread over a lane group.
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void fdgrad_knot_given_qdd_minv_lanes(const In& in, Out& out, const Options& opt, int role) {
  PhaseC<typename P::Col> PC;
  phase_a_given_qdd_minv<P>(in, opt, PC);
  if (role == kRoleQdd)
    for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  gradient_column_lane(PC, opt.damping, role, out);
}

// Inverse-dynamics gradient cdu = [dtau/dq | dtau/dqd] (98 per knot, the
// reference's "fused" GPU stage, dynamicsGradientKernel_fused,
// helpers_GPU/dynamicsGradient.cuh:1223-1255): q, qd, qdd given (u unused),
// RNEA at qdd, then the derivative sweeps; no Minv.
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void id_gradient_knot(const In& in, Out& out, const Options& opt) {
  PhaseC<typename P::Col> PC;
  phase_a_given_qdd_minv<P, false>(in, opt, PC);
  gradient_columns<false>(PC, opt.damping, out);
}
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void id_gradient_knot_lanes(const In& in, Out& out, const Options& opt, int role) {
  PhaseC<typename P::Col> PC;
  phase_a_given_qdd_minv<P, false>(in, opt, PC);
  if (role == kRoleQdd)
    for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  gradient_column_lane<false>(PC, opt.damping, role, out);
}

// The same from precomputed v, a, f (the reference's "split" GPU stage,
// dynamicsGradientKernel_split, .cuh:1191-1221): in.v/a/f(k, i) are RNEA's
// link velocities, accelerations and forces, f after the backward pass.
template <class P, class In>
RBD_HD RBD_INLINE void phase_split(const In& in, PhaseC<typename P::Col>& PC) {
  using SC = typename P::Col;
  CS<SC> tc[NQ];
  SC qd[NQ], v[NQ][6], a[NQ][6], f[NQ][6], Mi[NQ][NQ];
  for (int k = 0; k < NQ; ++k) {
    SC s, c;
    knot_sincos<P>(SC(in.q(k)), &s, &c);
    tc[k] = CS<SC>{c, s};
    qd[k] = SC(in.qd(k));
    for (int i = 0; i < 6; ++i) { v[k][i] = SC(in.v(k, i)); a[k][i] = SC(in.a(k, i)); f[k][i] = SC(in.f(k, i)); }
    for (int c2 = 0; c2 < NQ; ++c2) Mi[k][c2] = SC(0);
  }
  make_phase_c(PC, tc, qd, v, a, f, Mi);
}
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void id_gradient_split_knot(const In& in, Out& out, const Options& opt) {
  PhaseC<typename P::Col> PC;
  phase_split<P>(in, PC);
  gradient_columns<false>(PC, opt.damping, out);
}
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void id_gradient_split_knot_lanes(const In& in, Out& out, const Options& opt, int role) {
  PhaseC<typename P::Col> PC;
  phase_split<P>(in, PC);
  if (role == kRoleQdd)
    for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
  gradient_column_lane<false>(PC, opt.damping, role, out);
}

// RNEA's v, a, f at (q, qd, qdd), f after the backward pass: what the
// reference's CPU "start" stage writes into the split layout
// (forwardDynamicsGradientSetup_forGPU, .cuh:1148-1157, with QDD_MINV_PASSED_IN).
template <class P = PrecFP32, class In, class Out>
RBD_HD RBD_INLINE void vaf_knot(const In& in, Out& out, const Options& opt) {
  using SR = typename P::Rnea;
  CS<SR> tr[NQ];
  SR qd[NQ], qdd[NQ];
  for (int k = 0; k < NQ; ++k) {
    SR s, c;
    knot_sincos<P>(SR(in.q(k)), &s, &c);
    tr[k] = CS<SR>{c, s};
    qd[k] = SR(in.qd(k));
    qdd[k] = SR(in.qdd(k));
  }
  SR v[NQ][6], a[NQ][6], f[NQ][6];
  rnea_fwd(tr, qd, qdd, SR(opt.gravity), v, a, f);
  subtract_fext(in, f);
  rnea_bwd(tr, f);
  for (int k = 0; k < NQ; ++k)
    for (int i = 0; i < 6; ++i) out.vaf(k, i, float(v[k][i]), float(a[k][i]), float(f[k][i]));
}

// The reference's CPU "finish" stage (forwardDynamicsGradientFinish_forGPU,
// .cuh:1168-1179; matMultSym, helpers_CPU/math.h:16-25): dqdd = -Minv cdu
// with the upper triangle of Minv (49, column-major) and cdu 98 per knot.
template <typename S = float, class In, class Out>
RBD_HD RBD_INLINE void finish_knot(const In& in, Out& out) {
  S Mi[NQ][NQ];
  for (int r = 0; r < NQ; ++r)
    for (int c = r; c < NQ; ++c) Mi[r][c] = S(in.minv(r, c));
  for (int c = 0; c < 2 * NQ; ++c) {
    S x[NQ], y[NQ];
    for (int r = 0; r < NQ; ++r) x[r] = S(in.cdu(r, c));
    minv_times<true>(Mi, x, y);
    for (int r = 0; r < NQ; ++r) out.dqdd(r, c, float(y[r]));
  }
}

// ---------------------------------------------------------------------------
// Memory layouts
// ---------------------------------------------------------------------------
constexpr int kInPerKnot = 21;     // q, qd, u
constexpr int kOutPerKnot = 105;   // qdd, dqdd/dq (col-major), dqdd/dqd
constexpr int kVolPerKnot = 77;    // q, qd, u, qdd, Minv dense (COMPLETELY_FUSED_VALS, .cuh:1145)
constexpr int kDqddPerKnot = 98;   // dqdd/dq, dqdd/dqd
constexpr int kFextPerSet = 42;    // external forces: 7 links x (moment, force) in the link frame
constexpr int kFusedPerKnot = 28;  // q, qd, u, qdd (FUSED_VALS, .cuh:1144)
constexpr int kSplitPerKnot = 154; // q, qd, u, qdd, v (42), a (42), f (42) (SPLIT_VALS, .cuh:1143)

// array-of-structures: one knot's values contiguous (the reference's knot<T> order)
struct InAoS {
  const float* p;
  const float* fx = nullptr;   // external link forces (42 floats) or null
  RBD_HD float q(int i) const { return p[i]; }
  RBD_HD float qd(int i) const { return p[NQ + i]; }
  RBD_HD float u(int i) const { return p[2 * NQ + i]; }
  RBD_HD float qdd(int i) const { return p[3 * NQ + i]; }
  RBD_HD float minv(int r, int c) const { return p[4 * NQ + r + NQ * c]; }
};
struct OutAoS {
  float* p;
  RBD_HD void qdd(int i, float x) const { p[i] = x; }
  RBD_HD void dq(int r, int c, float x) const { p[NQ + NQ * c + r] = x; }
  RBD_HD void dqd(int r, int c, float x) const { p[NQ + NQ * NQ + NQ * c + r] = x; }
};
// full-path output plus dqdd/du = Minv (dense, column-major, 49 per knot)
struct OutAoSMinv {
  static constexpr bool kMinv = true;
  float* p; float* m;
  RBD_HD void qdd(int i, float x) const { p[i] = x; }
  RBD_HD void dq(int r, int c, float x) const { p[NQ + NQ * c + r] = x; }
  RBD_HD void dqd(int r, int c, float x) const { p[NQ + NQ * NQ + NQ * c + r] = x; }
  RBD_HD void minv(int r, int c, float x) const { m[NQ * c + r] = x; }
};
// The WAFR baseline's gradient layout (time_GPU_WAFR.cu:18; finish_dqdd_v2,
// helpers_WAFR/dynamics_WAFR.cuh:2555-2579): per knot kWafrPerKnot floats,
// [dqdd/dq | dqdd/dqd | dqdd/du = Minv], each 7x7 column-major; qdd (7 per
// knot) to its own buffer when q is not null.
constexpr int kWafrPerKnot = 3 * NQ * NQ;
struct OutWafr {
  static constexpr bool kMinv = true;
  float* q; float* p;
  RBD_HD void qdd(int i, float x) const { if (q) q[i] = x; }
  RBD_HD void dq(int r, int c, float x) const { p[NQ * c + r] = x; }
  RBD_HD void dqd(int r, int c, float x) const { p[NQ * NQ + NQ * c + r] = x; }
  RBD_HD void minv(int r, int c, float x) const { p[2 * NQ * NQ + NQ * c + r] = x; }
};
// forward dynamics only: qdd (7 per knot) and Minv (49 per knot, dense column-major)
// an output accessor whose stores are dropped when `on` is false (lanes that
// only take part in a workgroup's barriers)
template <class Out>
struct OutMasked {
  Out o; bool on;
  RBD_HD void qdd(int i, float x) const { if (on) o.qdd(i, x); }
  RBD_HD void dq(int r, int c, float x) const { if (on) o.dq(r, c, x); }
  RBD_HD void dqd(int r, int c, float x) const { if (on) o.dqd(r, c, x); }
};
struct OutFD {
  float* qdd_p; float* minv_p;
  RBD_HD void qdd(int i, float x) const { qdd_p[i] = x; }
  RBD_HD void minv(int r, int c, float x) const { minv_p[NQ * c + r] = x; }
};
struct OutDqddAoS {   // 98 per knot, no qdd
  float* p;
  RBD_HD void qdd(int, float) const {}
  RBD_HD void dq(int r, int c, float x) const { p[NQ * c + r] = x; }
  RBD_HD void dqd(int r, int c, float x) const { p[NQ * NQ + NQ * c + r] = x; }
};
// structure-of-arrays: value i of knot k at p[i * ld + k]
// split layout: [q | qd | u | qdd | v | a | f], link k's 6-vectors at 6k
struct InSplit {
  const float* p;
  const float* fx = nullptr;
  RBD_HD float q(int i) const { return p[i]; }
  RBD_HD float qd(int i) const { return p[NQ + i]; }
  RBD_HD float v(int k, int i) const { return p[4 * NQ + 6 * k + i]; }
  RBD_HD float a(int k, int i) const { return p[10 * NQ + 6 * k + i]; }
  RBD_HD float f(int k, int i) const { return p[16 * NQ + 6 * k + i]; }
};
struct OutVaf {   // v, a, f into the split layout
  float* p;
  RBD_HD void vaf(int k, int i, float v, float a, float f) const {
    p[4 * NQ + 6 * k + i] = v; p[10 * NQ + 6 * k + i] = a; p[16 * NQ + 6 * k + i] = f;
  }
};
// finish stage: Minv (49, column-major, upper triangle read) and cdu (98) in, dqdd (98) out
struct InFinish {
  const float* m; const float* c;
  RBD_HD float minv(int r, int col) const { return m[r + NQ * col]; }
  RBD_HD float cdu(int r, int col) const { return c[r + NQ * col]; }
};
struct OutFinish {
  float* p;
  RBD_HD void dqdd(int r, int col, float x) const { p[r + NQ * col] = x; }
};
struct InSoA {
  const float* p; long ld;
  const float* fx = nullptr;   // external link forces (42 floats) or null
  RBD_HD float q(int i) const { return p[i * ld]; }
  RBD_HD float qd(int i) const { return p[(NQ + i) * ld]; }
  RBD_HD float u(int i) const { return p[(2 * NQ + i) * ld]; }
};
// A knot's 21 inputs already in registers: the one-lane kernels load them
// before anything else of a loop iteration, ahead of the stores they defer.
struct InRegs {
  float x[kInPerKnot];
  const float* fx = nullptr;
  RBD_HD float q(int i) const { return x[i]; }
  RBD_HD float qd(int i) const { return x[NQ + i]; }
  RBD_HD float u(int i) const { return x[2 * NQ + i]; }
};
struct OutSoA {
  float* p; long ld;
  RBD_HD void qdd(int i, float x) const { p[i * ld] = x; }
  RBD_HD void dq(int r, int c, float x) const { p[(NQ + NQ * c + r) * ld] = x; }
  RBD_HD void dqd(int r, int c, float x) const { p[(NQ + NQ * NQ + NQ * c + r) * ld] = x; }
};

}  // namespace rbd
