// TEST INFRASTRUCTURE ONLY -- not part of the product.
//
// CPU restatement of the reference's 7-DoF KUKA iiwa forward-dynamics gradient
// (plancherb1/fast-rbd-gradients, helpers_CPU/*), written from the algorithm's
// steps without Eigen.  It is templated on the scalar type so the same code is
//   * the FP64 oracle the HIP kernels are checked against,
//   * the FP32 CPU baseline timed beside the GPU (bench.py cpu_baseline leg),
//   * a complex-step instantiation (std::complex<double>) whose imaginary parts
//     give machine-precision derivatives of forward dynamics, used to pin the
//     analytic gradient,
//   * an operation-counting instantiation that fixes the algorithmic FLOP
//     count per knot point.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Conventions (reference helpers_CPU/dynamicsGradient.h:77-118, :217-219,
// helpers_CPU/minv.h:3):
//   spatial vectors are [angular; linear]; 6x6 matrices are column-major,
//   A(r,c) = A[6c+r]; 7x7 matrices M(r,c) = M[r+7c]; T_k = kX_{k-1}(q_k) is the
//   motion transform parent->child and T_k^T the force transform child->parent;
//   every joint is revolute about its local z axis, S = e_2.
#pragma once
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdint>

namespace rbd_oracle {

constexpr int NQ = 7;
constexpr double kGravity = 9.81;          // dynamicsGradient.h:7-9

// Planted mistakes, for tests that show the oracle's checks catch them
// (0 = none).  1: drop dv x* (I v) from df; 2: flip the sign of the extra
// T^T(-(f x S)) backward term; 3: use T instead of T^T in the derivative
// backward sweep; 4: drop the Coriolis term (v x S) qd from the RNEA
// acceleration; 5: skip the (15) transport of F in the Minv backward pass.
extern thread_local int g_mutation;

enum Flags : int {
  kVelDamping = 1,   // bias -= 0.5*qd, dtau/dqd += 0.5*I   (dynamicsGradient.h:680,840)
  kNoGravity  = 2,   // MPC_MODE: base acceleration 0        (dynamicsGradient.h:635-637)
};

// ---------------------------------------------------------------------------
// Model data: utils/iiwa_14.urdf.  Link k (0-based) is URDF iiwa_link_{k+1},
// joint k is URDF iiwa_joint_{k+1}.  Inertia entries are about the COM, as in
// the URDF; the parallel-axis shift happens in spatial_inertia().
// (Same numbers as the reference's tables, dynamicsGradient.h:127-202, :223-229.)
// These are the default; oracle_set_model() replaces them with another 7-joint
// serial chain (a robot read from a URDF).
// ---------------------------------------------------------------------------
struct LinkInertial { double mass, cx, cy, cz, Ixx, Ixy, Ixz, Iyy, Iyz, Izz; };
struct JointOrigin  { double x, y, z, roll, pitch, yaw; };

constexpr double kPi = 3.141592653589793;
constexpr double kHalfPi = 1.570796326794897;

static const LinkInertial kIiwaLinks[NQ] = {
  {5.76, 0.0,    -0.03,  0.12,   0.121,  0.0,      0.0,      0.116,    -0.021,   0.0175 },
  {6.35, 0.0003,  0.059, 0.042,  0.0638, 0.0001,   0.00008,  0.0416,    0.0157,  0.0331 },
  {3.5,  0.0,     0.03,  0.13,   0.0873, 0.0,      0.0,      0.083,     0.014,   0.0108 },
  {3.5,  0.0,     0.067, 0.034,  0.0368, 0.0,      0.0,      0.02045,   0.008,   0.02171},
  {3.5,  0.0001,  0.021, 0.076,  0.0318, 0.000007, 0.000027, 0.028916,  0.005586,0.006  },
  {1.8,  0.0,     0.0006,0.0004, 0.0049, 0.0,      0.0,      0.0047,    0.0,     0.0036 },
  {1.2,  0.0,     0.0,   0.02,   0.0055, 0.0,      0.0,      0.0055,    0.0,     0.005  },
};

static const JointOrigin kIiwaJoints[NQ] = {
  {0.0, 0.0,    0.1575, 0.0,      0.0, 0.0 },
  {0.0, 0.0,    0.2025, kHalfPi,  0.0, kPi },
  {0.0, 0.2045, 0.0,    kHalfPi,  0.0, kPi },
  {0.0, 0.0,    0.2155, kHalfPi,  0.0, 0.0 },
  {0.0, 0.1845, 0.0,   -kHalfPi,  kPi, 0.0 },
  {0.0, 0.0,    0.2155, kHalfPi,  0.0, 0.0 },
  {0.0, 0.081,  0.0,   -kHalfPi,  kPi, 0.0 },
};

// Fixed part of each joint frame: F_k = R^T, where the joint frame is R Rz(q)
// at offset r in the parent frame.  For a URDF joint R = R_rpy = Rz(yaw)
// Ry(pitch) Rx(roll).  Entries within 1e-12 of an integer are snapped to it:
// for the iiwa, where every angle is a multiple of pi/2, this reproduces the
// exact 0/+-1 pattern of the reference's hand-written table
// (dynamicsGradient.h:507-587).
struct FixedFrame { double F[3][3]; double r[3]; };

inline double snap_unit(double x) {
  const double r = std::round(x);
  return std::fabs(x - r) < 1e-12 ? r : x;
}

inline FixedFrame make_fixed_frame(const JointOrigin& j) {
  const double cr = std::cos(j.roll), sr = std::sin(j.roll);
  const double cp = std::cos(j.pitch), sp = std::sin(j.pitch);
  const double cy = std::cos(j.yaw), sy = std::sin(j.yaw);
  const double R[3][3] = {
    {cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr},
    {sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr},
    {-sp,     cp * sr,                cp * cr},
  };
  FixedFrame out;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out.F[a][b] = snap_unit(R[b][a]);
  out.r[0] = j.x; out.r[1] = j.y; out.r[2] = j.z;
  return out;
}

// The robot every function below computes on.
struct Model {
  LinkInertial links[NQ];
  FixedFrame frames[NQ];
};
inline Model iiwa_model() {
  Model m;
  for (int k = 0; k < NQ; ++k) {
    m.links[k] = kIiwaLinks[k];
    m.frames[k] = make_fixed_frame(kIiwaJoints[k]);
  }
  return m;
}
// Process-wide; replaced only between computations (oracle_set_model).
inline Model& model() {
  static Model m = iiwa_model();
  return m;
}
inline const FixedFrame& fixed_frame(int k) { return model().frames[k]; }

// ---------------------------------------------------------------------------
// Scalar helpers (work for double, float, std::complex<double> and the
// operation-counting type below).
// ---------------------------------------------------------------------------
template <typename T> inline T tsin(const T& x) { using std::sin; return sin(x); }
template <typename T> inline T tcos(const T& x) { using std::cos; return cos(x); }

// Spatial inertia of link k about its joint frame origin:
//   [ Ibar + m(c^T c 1 - c c^T)   m c x ;  m c x^T   m 1 ]
// (the parallel-axis shift of buildI6x6, dynamicsGradient.h:77-118).
template <typename T>
void spatial_inertia(int k, T* I) {
  const LinkInertial& L = model().links[k];
  const double m = L.mass, cx = L.cx, cy = L.cy, cz = L.cz;
  double A[36] = {0};
  A[0]  = L.Ixx + m * (cy * cy + cz * cz);
  A[7]  = L.Iyy + m * (cx * cx + cz * cz);
  A[14] = L.Izz + m * (cx * cx + cy * cy);
  A[1] = A[6]  = L.Ixy - m * cx * cy;
  A[2] = A[12] = L.Ixz - m * cx * cz;
  A[8] = A[13] = L.Iyz - m * cy * cz;
  // upper-right block m*[c]x  (rows 0..2, cols 3..5); lower-left = its transpose
  const double mc[3] = {m * cx, m * cy, m * cz};
  const double cxm[3][3] = {{0, -mc[2], mc[1]}, {mc[2], 0, -mc[0]}, {-mc[1], mc[0], 0}};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[6 * (c + 3) + r] = cxm[r][c];
      A[6 * c + (r + 3)] = cxm[c][r];
    }
  A[6 * 3 + 3] = A[6 * 4 + 4] = A[6 * 5 + 5] = m;
  for (int i = 0; i < 36; ++i) I[i] = T(A[i]);
}

// Motion transform T_k(q_k) = [E 0; -E r^ E], E = Rz(q)^T F (see fixed_frame).
// Entries that are structurally zero are never computed, so the operation
// count matches the reference's entry-wise updateTransforms (dynamicsGradient.h:507-587).
template <typename T>
void joint_transform(int k, const T& s, const T& c, T* X) {
  const FixedFrame& ff = fixed_frame(k);
  for (int i = 0; i < 36; ++i) X[i] = T(0);
  // Rz(q)^T = [[c, s, 0], [-s, c, 0], [0, 0, 1]]
  T E[3][3];
  bool nz[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T e = T(0); bool any = false;
      for (int l = 0; l < 3; ++l) {
        const double f = ff.F[l][j];
        if (f == 0.0) continue;
        // Rz^T entry (i,l)
        int code = 0;  // 0: zero, 1: c, 2: s, 3: -s, 4: one
        if (i < 2 && l < 2) code = (i == l) ? 1 : (i == 0 ? 2 : 3);
        else if (i == 2 && l == 2) code = 4;
        if (code == 0) continue;
        T term = (code == 1) ? c : (code == 2) ? s : (code == 3) ? -s : T(1);
        if (f == -1.0) term = -term;
        else if (f != 1.0) term = term * T(f);
        e = any ? e + term : term;   // one term per entry for the iiwa's 0/+-1 frames
        any = true;
      }
      E[i][j] = e; nz[i][j] = any;
    }
  // r^ (skew of the joint origin); B = -E r^
  const double rx[3][3] = {{0, -ff.r[2], ff.r[1]}, {ff.r[2], 0, -ff.r[0]}, {-ff.r[1], ff.r[0], 0}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      X[6 * j + i] = E[i][j];
      X[6 * (j + 3) + (i + 3)] = E[i][j];
      T b = T(0); bool any = false;
      for (int l = 0; l < 3; ++l) {
        if (!nz[i][l] || rx[l][j] == 0.0) continue;
        T term = E[i][l] * T(-rx[l][j]);
        b = any ? b + term : term;
        any = true;
      }
      X[6 * j + (i + 3)] = b;
    }
}

// out (+)= T x using the transform's sparsity (top-right block zero):
// rows 0-2 read 3 entries, rows 3-5 read 6 (matVMult_Tmat, math.h:65-88).
template <typename T, bool ACC>
void tmat_mv(T* out, const T* X, const T* x) {
  for (int r = 0; r < 6; ++r) {
    const int n = r < 3 ? 3 : 6;
    T acc = X[r] * x[0];
    for (int i = 1; i < n; ++i) acc = acc + X[r + 6 * i] * x[i];
    out[r] = ACC ? out[r] + acc : acc;
  }
}
// out (+)= T^T x: rows 0-2 read 6 entries, rows 3-5 read entries 3..5.
template <typename T, bool ACC>
void tmat_tmv(T* out, const T* X, const T* x) {
  for (int r = 0; r < 6; ++r) {
    const int i0 = r < 3 ? 0 : 3;
    T acc = X[6 * r + i0] * x[i0];
    for (int i = i0 + 1; i < 6; ++i) acc = acc + X[6 * r + i] * x[i];
    out[r] = ACC ? out[r] + acc : acc;
  }
}
// out (+)= A x for a dense 6x6 column-major A (matVMult, math.h:4-13)
template <typename T, bool ACC>
void mat6_mv(T* out, const T* A, const T* x) {
  for (int r = 0; r < 6; ++r) {
    T acc = A[r] * x[0];
    for (int i = 1; i < 6; ++i) acc = acc + A[r + 6 * i] * x[i];
    out[r] = ACC ? out[r] + acc : acc;
  }
}
template <typename T>
inline void cross3(T* out, const T* a, const T* b) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}
// out = x x* y  (spatial force cross product):  [w x ya + u x yl ; w x yl]
// (computeFxMatTimesVector, dynamicsGradient.h:766-781; vxIv :602-616)
template <typename T>
void crf_mv(T* out, const T* x, const T* y) {
  T t0[3], t1[3], t2[3];
  cross3(t0, x, y);
  cross3(t1, x + 3, y + 3);
  cross3(t2, x, y + 3);
  for (int i = 0; i < 3; ++i) { out[i] = t0[i] + t1[i]; out[i + 3] = t2[i]; }
}
// x x S scaled by alpha, S = e_2:  [x1, -x0, 0, x4, -x3, 0]*alpha
// (MxMatCol3, dynamicsGradient.h:712-746)
template <typename T>
inline void mx3(T* out, const T* x, const T& alpha) {
  out[0] = x[1] * alpha; out[1] = -(x[0] * alpha); out[2] = T(0);
  out[3] = x[4] * alpha; out[4] = -(x[3] * alpha); out[5] = T(0);
}
// out (+)= (T x) x S, computing only rows 0, 1, 3, 4 of T x
// (MxMatCol3_Tmat, dynamicsGradient.h:748-764)
template <typename T, bool ACC>
void tmat_mv_mx3(T* out, const T* X, const T* x) {
  T y[6];
  for (int r : {0, 1, 3, 4}) {
    const int n = r < 3 ? 3 : 6;
    T acc = X[r] * x[0];
    for (int i = 1; i < n; ++i) acc = acc + X[r + 6 * i] * x[i];
    y[r] = acc;
  }
  if (ACC) {
    out[0] = out[0] + y[1]; out[1] = out[1] - y[0];
    out[3] = out[3] + y[4]; out[4] = out[4] - y[3];
  } else {
    out[0] = y[1]; out[1] = -y[0]; out[2] = T(0);
    out[3] = y[4]; out[4] = -y[3]; out[5] = T(0);
  }
}
template <typename T>
inline void mx3_acc(T* out, const T* x) {
  out[0] = out[0] + x[1]; out[1] = out[1] - x[0];
  out[3] = out[3] + x[4]; out[4] = out[4] - x[3];
}

// ---------------------------------------------------------------------------
// Per-knot working set
// ---------------------------------------------------------------------------
template <typename T>
struct Knot {
  T X[NQ][36];       // T_k(q_k)
  T I[NQ][36];       // spatial inertias
  T v[NQ][6], a[NQ][6], f[NQ][6];
  T Minv[NQ * NQ];   // upper triangle stored, Minv(r,c) = Minv[r + 7c], r <= c
};

template <typename T>
void setup_model(Knot<T>& K, const T* q) {
  for (int k = 0; k < NQ; ++k) {
    const T s = tsin(q[k]), c = tcos(q[k]);
    joint_transform<T>(k, s, c, K.X[k]);
    spatial_inertia<T>(k, K.I[k]);
  }
}

// RNEA forward pass (RNEA_fp, dynamicsGradient.h:619-659).  qdd may be null
// (treated as zero -- the reference reads it unconditionally, :634,:647).
// fext (may be null): external spatial force on each link in its own frame,
// fext[6k + i], subtracted from f_k (dynamicsGradient.h:654-657).
template <typename T>
void rnea_forward(Knot<T>& K, const T* qd, const T* qdd, int flags, const double* fext = nullptr) {
  for (int k = 0; k < NQ; ++k) {
    T* v = K.v[k]; T* a = K.a[k]; const T* X = K.X[k];
    if (k == 0) {
      for (int i = 0; i < 6; ++i) { v[i] = T(0); a[i] = T(0); }
      v[2] = qd[0];
      if (qdd) a[2] = qdd[0];
      if (!(flags & kNoGravity))
        for (int i = 3; i < 6; ++i) a[i] = X[6 * 5 + i] * T(kGravity);
    } else {
      tmat_mv<T, false>(v, X, K.v[k - 1]);
      v[2] = v[2] + qd[k];
      mx3(a, v, g_mutation == 4 ? T(0) : qd[k]);
      tmat_mv<T, true>(a, X, K.a[k - 1]);
      if (qdd) a[2] = a[2] + qdd[k];
    }
    T Iv[6];
    mat6_mv<T, false>(Iv, K.I[k], v);
    crf_mv(K.f[k], v, Iv);
    mat6_mv<T, true>(K.f[k], K.I[k], a);
    if (fext)
      for (int i = 0; i < 6; ++i) K.f[k][i] = K.f[k][i] - T(fext[6 * k + i]);
  }
}

// RNEA backward pass (RNEA_bp, dynamicsGradient.h:661-670): f_{k-1} += T_k^T f_k,
// tau_k = f_k[2].
template <typename T>
void rnea_backward(Knot<T>& K, T* tau) {
  for (int k = NQ - 1; k >= 0; --k) {
    if (tau) tau[k] = K.f[k][2];
    if (k > 0) tmat_tmv<T, true>(K.f[k - 1], K.X[k], K.f[k]);
  }
}

// Minv by the articulated-body recursion (computeMinv, minv.h:423-478):
// backward pass (eqs. 12, 14, 15, 19, 21 of the minv.h comments) building the
// articulated inertias IA_k, then forward pass (eqs. 30, 32, 34).
// The articulated-inertia transport IA_{k-1} += T^T (IA_k - U U^T / D) T is done
// here as rotate-then-translate on 3x3 blocks (the reference uses the RobCoGen
// sparse expansion, minv.h:146-386, which is the same product).
template <typename T>
void articulated_transport(T* IAp, const T* X, const T* Ia) {
  // X = [E 0; B E] with B = -E r^.  X^T Ia X with Ia = [A C; C^T M]:
  //   top-left  = E^T A E + B^T C^T E + E^T C B + B^T M B
  //   top-right = E^T C E + B^T M E          (bottom-left = its transpose)
  //   bot-right = E^T M E
  T E[3][3], B[3][3], A[3][3], C[3][3], M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      E[i][j] = X[6 * j + i];
      B[i][j] = X[6 * j + i + 3];
      A[i][j] = Ia[6 * j + i];
      C[i][j] = Ia[6 * (j + 3) + i];
      M[i][j] = Ia[6 * (j + 3) + i + 3];
    }
  auto mm = [](T (*out)[3], T (*x)[3], T (*y)[3], bool tx) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T acc = (tx ? x[0][i] : x[i][0]) * y[0][j];
        for (int l = 1; l < 3; ++l) acc = acc + (tx ? x[l][i] : x[i][l]) * y[l][j];
        out[i][j] = acc;
      }
  };
  T AE[3][3], CE[3][3], MB[3][3], ME[3][3], CtE[3][3], CB[3][3];
  T Ct[3][3];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Ct[i][j] = C[j][i];
  mm(AE, A, E, false);    // A E
  mm(CE, C, E, false);    // C E
  mm(ME, M, E, false);    // M E
  mm(MB, M, B, false);    // M B
  mm(CtE, Ct, E, false);  // C^T E
  mm(CB, C, B, false);    // C B
  // top-left: E^T (A E + C B) + B^T (C^T E + M B)
  T P[3][3], Q[3][3], TL[3][3], TR[3][3], BR[3][3], t1[3][3], t2[3][3];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
    P[i][j] = AE[i][j] + CB[i][j];
    Q[i][j] = CtE[i][j] + MB[i][j];
  }
  mm(t1, E, P, true); mm(t2, B, Q, true);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) TL[i][j] = t1[i][j] + t2[i][j];
  // top-right: E^T C E + B^T M E
  mm(t1, E, CE, true); mm(t2, B, ME, true);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) TR[i][j] = t1[i][j] + t2[i][j];
  mm(BR, E, ME, true);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      IAp[6 * j + i] = IAp[6 * j + i] + TL[i][j];
      IAp[6 * (j + 3) + i] = IAp[6 * (j + 3) + i] + TR[i][j];
      IAp[6 * j + i + 3] = IAp[6 * j + i + 3] + TR[j][i];
      IAp[6 * (j + 3) + i + 3] = IAp[6 * (j + 3) + i + 3] + BR[i][j];
    }
}

template <typename T>
void compute_minv(Knot<T>& K, T* Dout = nullptr) {
  T IA[NQ][36];                      // the recursion overwrites IA (dynamicsGradient.h:887-889)
  for (int k = 0; k < NQ; ++k) for (int i = 0; i < 36; ++i) IA[k][i] = K.I[k][i];
  T F[NQ][6 * NQ];                   // F_k[:, j] at F[k][6j + r], zero-initialised (minv.h:473)
  for (int k = 0; k < NQ; ++k) for (int i = 0; i < 6 * NQ; ++i) F[k][i] = T(0);
  T U[NQ][6], Dinv[NQ];
  T* Mi = K.Minv;
  for (int i = 0; i < NQ * NQ; ++i) Mi[i] = T(0);
  // backward pass (Minv_bp, minv.h:423-436)
  for (int k = NQ - 1; k >= 0; --k) {
    for (int i = 0; i < 6; ++i) U[k][i] = IA[k][12 + i];   // U = IA S
    const T D = U[k][2];                                    // D = S^T IA S
    if (Dout) Dout[k] = D;
    Dinv[k] = T(1) / D;
    Mi[k + NQ * k] = Dinv[k];                                // (21)
    for (int j = k + 1; j < NQ; ++j)                          // (12)
      Mi[k + NQ * j] = -(Dinv[k] * F[k][6 * j + 2]);
    for (int j = k; j < NQ; ++j)                              // (14)/(19)
      for (int r = 0; r < 6; ++r) F[k][6 * j + r] = F[k][6 * j + r] + U[k][r] * Mi[k + NQ * j];
    if (k > 0) {
      if (g_mutation != 5)
        for (int j = k; j < NQ; ++j)                          // (15); columns j < k of F_k are zero
          tmat_tmv<T, true>(&F[k - 1][6 * j], K.X[k], &F[k][6 * j]);
      // Ia = IA - U U^T / D (compute_Ia_revolute, minv.h:146-197): symmetric,
      // row and column 2 are exactly zero for a revolute joint about z.
      T UD[6], Ia[36];
      for (int i = 0; i < 6; ++i) UD[i] = U[k][i] * Dinv[k];
      for (int c = 0; c < 6; ++c)
        for (int r = 0; r <= c; ++r) {
          const T e = (r == 2 || c == 2) ? T(0) : IA[k][6 * c + r] - U[k][r] * UD[c];
          Ia[6 * c + r] = e; Ia[6 * r + c] = e;
        }
      articulated_transport(IA[k - 1], K.X[k], Ia);
    }
  }
  // forward pass (Minv_fp, minv.h:439-469)
  for (int k = 0; k < NQ; ++k) {
    if (k > 0) {
      T tU[6];                                                // (30): T^T U
      tmat_tmv<T, false>(tU, K.X[k], U[k]);
      for (int j = k; j < NQ; ++j) {
        T acc = tU[0] * F[k - 1][6 * j];
        for (int r = 1; r < 6; ++r) acc = acc + tU[r] * F[k - 1][6 * j + r];
        Mi[k + NQ * j] = Mi[k + NQ * j] + (-Dinv[k]) * acc;
      }
    }
    for (int j = k; j < NQ; ++j) {                            // (32) F = S Minv[k, k..]
      for (int r = 0; r < 6; ++r) F[k][6 * j + r] = T(0);
      F[k][6 * j + 2] = Mi[k + NQ * j];
    }
    if (k > 0)                                                // (34) F += T F_{k-1}
      for (int j = k; j < NQ; ++j) tmat_mv<T, true>(&F[k][6 * j], K.X[k], &F[k - 1][6 * j]);
  }
}

// y (+)= alpha * Minv_sym * x, Minv upper-stored (matVMultSym / matMultSym, math.h:16-37)
template <bool NEG, typename T>
void symv_upper(T* y, const T* Mi, const T* x) {
  for (int r = 0; r < NQ; ++r) {
    T acc = Mi[NQ * r] * x[0];            // entry (r, 0) is stored at (0, r)
    for (int c = 1; c < NQ; ++c) {
      const T m = r <= c ? Mi[r + NQ * c] : Mi[c + NQ * r];
      acc = acc + m * x[c];
    }
    y[r] = NEG ? -acc : acc;
  }
}

// d(tau)/d(q), d(tau)/d(qd) (inverseDynamicsGradient = compute_vafdu + compute_cdu,
// dynamicsGradient.h:786-866).  Column c of the outputs is cdq[7c + k] = dtau_k/dq_c.
template <typename T>
void inverse_dynamics_gradient(const Knot<T>& K, const T* qd, T* cdq, T* cdqd, int flags) {
  // packed lower-triangular column storage: entry (k, c<=k) at 3k(k+1) + 6c
  T vdq[3 * NQ * (NQ + 1)], vdqd[3 * NQ * (NQ + 1)], adq[3 * NQ * (NQ + 1)], adqd[3 * NQ * (NQ + 1)];
  T fdq[6 * NQ * NQ], fdqd[6 * NQ * NQ];
  for (int k = 0; k < NQ; ++k) {
    const T* vk = K.v[k]; const T* X = K.X[k]; const T* Ik = K.I[k];
    T Ivk[6]; mat6_mv<T, false>(Ivk, Ik, vk);
    T FxvI[36];
    for (int c = 0; c < 6; ++c) crf_mv(&FxvI[6 * c], vk, &Ik[6 * c]);
    for (int c = 0; c <= k; ++c) {
      const int ikc = 3 * k * (k + 1) + 6 * c, ilc = ikc - 6 * k, ikc6 = 6 * NQ * k + 6 * c;
      if (c > 0) {
        // d/dq: column 0 is identically zero (rotation about the gravity axis)
        if (c < k) tmat_mv<T, false>(&vdq[ikc], X, &vdq[ilc]);
        else tmat_mv_mx3<T, false>(&vdq[ikc], X, K.v[k - 1]);   // (T v_{k-1}) x S
        mx3(&adq[ikc], &vdq[ikc], qd[k]);
        if (c < k) tmat_mv<T, true>(&adq[ikc], X, &adq[ilc]);
        else tmat_mv_mx3<T, true>(&adq[ikc], X, K.a[k - 1]);    // + (T a_{k-1}) x S
        crf_mv(&fdq[ikc6], &vdq[ikc], Ivk);
        if (g_mutation == 1) for (int i = 0; i < 6; ++i) fdq[ikc6 + i] = T(0);
        mat6_mv<T, true>(&fdq[ikc6], Ik, &adq[ikc]);
        mat6_mv<T, true>(&fdq[ikc6], FxvI, &vdq[ikc]);
      } else {
        for (int i = 0; i < 6; ++i) fdq[ikc6 + i] = T(0);
      }
      // d/dqd
      if (c < k - 1) tmat_mv<T, false>(&vdqd[ikc], X, &vdqd[ilc]);
      else if (c == k - 1) for (int i = 0; i < 6; ++i) vdqd[ikc + i] = X[12 + i];   // T S
      else { for (int i = 0; i < 6; ++i) vdqd[ikc + i] = T(0); vdqd[ikc + 2] = T(1); }  // S
      mx3(&adqd[ikc], &vdqd[ikc], qd[k]);
      if (c < k) tmat_mv<T, true>(&adqd[ikc], X, &adqd[ilc]);
      else mx3_acc(&adqd[ikc], vk);                  // + v_k x S
      crf_mv(&fdqd[ikc6], &vdqd[ikc], Ivk);
      if (g_mutation == 1) for (int i = 0; i < 6; ++i) fdqd[ikc6 + i] = T(0);
      mat6_mv<T, true>(&fdqd[ikc6], Ik, &adqd[ikc]);
      mat6_mv<T, true>(&fdqd[ikc6], FxvI, &vdqd[ikc]);
    }
  }
  // backward sweep (compute_cdu, dynamicsGradient.h:831-857)
  for (int k = NQ - 1; k >= 0; --k) {
    const T* X = K.X[k]; const T* fk = K.f[k];
    for (int c = 0; c < NQ; ++c) {
      const int ikc = 6 * NQ * k + 6 * c, ilc = ikc - 6 * NQ;
      cdq[NQ * c + k] = fdq[ikc + 2];
      cdqd[NQ * c + k] = fdqd[ikc + 2];
      if ((flags & kVelDamping) && c == k) cdqd[NQ * c + k] = cdqd[NQ * c + k] + T(0.5);
      if (k > 0) {
        if (g_mutation == 3) {
          if (c >= k) for (int i = 0; i < 6; ++i) { fdq[ilc + i] = T(0); fdqd[ilc + i] = T(0); }
          tmat_mv<T, true>(&fdq[ilc], X, &fdq[ikc]);
          tmat_mv<T, true>(&fdqd[ilc], X, &fdqd[ikc]);
        } else if (c < k) {
          tmat_tmv<T, true>(&fdq[ilc], X, &fdq[ikc]);
          tmat_tmv<T, true>(&fdqd[ilc], X, &fdqd[ikc]);
        } else {
          tmat_tmv<T, false>(&fdq[ilc], X, &fdq[ikc]);
          tmat_tmv<T, false>(&fdqd[ilc], X, &fdqd[ikc]);
          if (c == k) {                              // + T^T (-(f_k x S))
            for (int r = 0; r < 6; ++r) {
              const T* Xc = &X[6 * r];
              const T extra = Xc[1] * fk[0] - Xc[0] * fk[1] + Xc[4] * fk[3] - Xc[3] * fk[4];
              fdq[ilc + r] = g_mutation == 2 ? fdq[ilc + r] - extra : fdq[ilc + r] + extra;
            }
          }
        }
      }
    }
  }
}

// Full path for one knot point: q, qd, u -> qdd, dqdd/dq, dqdd/dqd.
// Order follows forwardDynamicsGradientSetup with QDD_MINV_PASSED_IN = false
// (dynamicsGradient.h:868-897) and forwardDynamicsGradientThreaded (:924-934),
// except that the bias RNEA uses qdd = 0 (the reference reads a stale knot.qdd).
// out: [qdd(7) | dqdd/dq (49, col-major) | dqdd/dqd (49)].
template <typename T>
void fdgrad_knot(const T* q, const T* qd, const T* u, T* out, int flags, Knot<T>* keep = nullptr,
                 const double* fext = nullptr) {
  Knot<T> local;
  Knot<T>& K = keep ? *keep : local;
  setup_model(K, q);
  rnea_forward(K, qd, (const T*)nullptr, flags, fext);
  T cbias[NQ];
  rnea_backward(K, cbias);
  T bias[NQ];
  for (int k = 0; k < NQ; ++k) {
    bias[k] = u[k] - cbias[k];
    if (flags & kVelDamping) bias[k] = bias[k] - T(0.5) * qd[k];
  }
  compute_minv(K);
  T* qdd = out;
  symv_upper<false>(qdd, K.Minv, bias);
  rnea_forward(K, qd, qdd, flags, fext);
  rnea_backward(K, (T*)nullptr);
  T cdu[2 * NQ * NQ];
  inverse_dynamics_gradient(K, qd, cdu, cdu + NQ * NQ, flags);
  for (int c = 0; c < 2 * NQ; ++c) symv_upper<true>(out + NQ + NQ * c, K.Minv, cdu + NQ * c);
}

// Reference-compatible mode (QDD_MINV_PASSED_IN = true, the reference's timed
// configuration; time_GPU.cu:141-158 COMPLETELY_FUSED layout): per knot
// vol = [q | qd | u | qdd | Minv dense 7x7], out = [dqdd/dq | dqdd/dqd].
template <typename T>
void fdgrad_knot_given_qdd_minv(const T* vol, T* out, int flags, const double* fext = nullptr) {
  Knot<T> K;
  const T* q = vol; const T* qd = vol + NQ; const T* qdd = vol + 3 * NQ; const T* Md = vol + 4 * NQ;
  setup_model(K, q);
  rnea_forward(K, qd, qdd, flags, fext);
  rnea_backward(K, (T*)nullptr);
  T cdu[2 * NQ * NQ];
  inverse_dynamics_gradient(K, qd, cdu, cdu + NQ * NQ, flags);
  for (int i = 0; i < NQ * NQ; ++i) K.Minv[i] = Md[i];   // the upper triangle is what symv reads
  for (int c = 0; c < 2 * NQ; ++c) symv_upper<true>(out + NQ * c, K.Minv, cdu + NQ * c);
}

// Forward dynamics only: qdd and dense Minv (forwardDynamics, dynamicsGradient.h:685-706).
template <typename T>
void fd_knot(const T* q, const T* qd, const T* u, T* qdd, T* Minv_dense, int flags, const double* fext = nullptr) {
  Knot<T> K;
  setup_model(K, q);
  rnea_forward(K, qd, (const T*)nullptr, flags, fext);
  T cbias[NQ];
  rnea_backward(K, cbias);
  T bias[NQ];
  for (int k = 0; k < NQ; ++k) {
    bias[k] = u[k] - cbias[k];
    if (flags & kVelDamping) bias[k] = bias[k] - T(0.5) * qd[k];
  }
  compute_minv(K);
  symv_upper<false>(qdd, K.Minv, bias);
  if (Minv_dense)
    for (int r = 0; r < NQ; ++r)
      for (int c = 0; c < NQ; ++c) Minv_dense[r + NQ * c] = r <= c ? K.Minv[r + NQ * c] : K.Minv[c + NQ * r];
}

// Inverse dynamics tau = RNEA(q, qd, qdd) (+0.5 qd with damping).
template <typename T>
void rnea_knot(const T* q, const T* qd, const T* qdd, T* tau, int flags, const double* fext = nullptr) {
  Knot<T> K;
  setup_model(K, q);
  rnea_forward(K, qd, qdd, flags, fext);
  rnea_backward(K, tau);
  if (flags & kVelDamping)
    for (int k = 0; k < NQ; ++k) tau[k] = tau[k] + T(0.5) * qd[k];
}

// ---------------------------------------------------------------------------
// Operation-counting scalar: every binary +, -, *, / on it increments a
// counter (unary minus is a sign flip and is not counted).  Instantiating the
// float path with it gives the algorithmic FLOP count per knot point.
// ---------------------------------------------------------------------------
struct OpCount { long long add = 0, mul = 0, div = 0, trig = 0; };
extern thread_local OpCount g_ops;
// When set, an operand that is an exact structural zero (a literal 0 of the
// method or of the model data) makes the operation free: the count is then
// the number of operations the method needs on this model's sparsity.
extern thread_local bool g_skip_structural_zeros;

struct Counted {
  double v = 0;
  bool z = true;        // structurally zero
  Counted() = default;
  Counted(double x) : v(x), z(x == 0.0) {}
  static Counted make(double x, bool zero) { Counted c; c.v = x; c.z = zero; return c; }
};
inline Counted operator-(Counted a) { return Counted::make(-a.v, a.z); }
inline Counted operator+(Counted a, Counted b) {
  if (g_skip_structural_zeros) { if (a.z) return b; if (b.z) return a; }
  ++g_ops.add; return Counted::make(a.v + b.v, false);
}
inline Counted operator-(Counted a, Counted b) {
  if (g_skip_structural_zeros) { if (b.z) return a; if (a.z) return -b; }
  ++g_ops.add; return Counted::make(a.v - b.v, false);
}
inline Counted operator*(Counted a, Counted b) {
  if (g_skip_structural_zeros && (a.z || b.z)) return Counted::make(0.0, true);
  ++g_ops.mul; return Counted::make(a.v * b.v, false);
}
inline Counted operator/(Counted a, Counted b) {
  if (g_skip_structural_zeros && a.z) return Counted::make(0.0, true);
  ++g_ops.div; return Counted::make(a.v / b.v, false);
}
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline Counted sin(Counted a) { ++g_ops.trig; return Counted::make(std::sin(a.v), false); }
inline Counted cos(Counted a) { ++g_ops.trig; return Counted::make(std::cos(a.v), false); }

}  // namespace rbd_oracle
