// KUKA iiwa 14 model constants, as compile-time tables for the device code.
//
// Source: utils/iiwa_14.urdf of the reference (link k = URDF iiwa_link_{k+1},
// joint k = URDF iiwa_joint_{k+1}); the reference hard-codes the same numbers in
// helpers_CPU/dynamicsGradient.h:123-229 and helpers_GPU/dynamicsGradient.cuh:211-310.
// Every joint is revolute about its local z axis.  The joint frame is
// R_rpy * Rz(q) at offset r from the parent; with E = (R_rpy Rz(q))^T the motion
// transform parent->child is T = [E 0; -E r^ E].  kFixedRot[k] = R_rpy^T, whose
// entries are exact integers because every rpy angle of this robot is a
// multiple of pi/2.
#pragma once

namespace rbd {

constexpr const char* kModelName = "iiwa14";
constexpr int NQ = 7;
constexpr double kGravity = 9.81;

// mass, COM (x, y, z) in the link frame, inertia about the COM (URDF <inertia>)
struct LinkInertial { double m, cx, cy, cz, ixx, ixy, ixz, iyy, iyz, izz; };

constexpr LinkInertial kLinks[NQ] = {
  {5.76, 0.0,    -0.03,   0.12,   0.121,  0.0,      0.0,      0.116,    -0.021,    0.0175 },
  {6.35, 0.0003,  0.059,  0.042,  0.0638, 0.0001,   0.00008,  0.0416,    0.0157,   0.0331 },
  {3.5,  0.0,     0.03,   0.13,   0.0873, 0.0,      0.0,      0.083,     0.014,    0.0108 },
  {3.5,  0.0,     0.067,  0.034,  0.0368, 0.0,      0.0,      0.02045,   0.008,    0.02171},
  {3.5,  0.0001,  0.021,  0.076,  0.0318, 0.000007, 0.000027, 0.028916,  0.005586, 0.006  },
  {1.8,  0.0,     0.0006, 0.0004, 0.0049, 0.0,      0.0,      0.0047,    0.0,      0.0036 },
  {1.2,  0.0,     0.0,    0.02,   0.0055, 0.0,      0.0,      0.0055,    0.0,      0.005  },
};

// R_rpy^T per joint (URDF rpy: (0,0,0), (pi/2,0,pi), (pi/2,0,pi), (pi/2,0,0),
// (-pi/2,pi,0), (pi/2,0,0), (-pi/2,pi,0)).
constexpr int kFixedRot[NQ][3][3] = {
  {{ 1, 0, 0}, {0, 1, 0}, {0,  0, 1}},
  {{-1, 0, 0}, {0, 0, 1}, {0,  1, 0}},
  {{-1, 0, 0}, {0, 0, 1}, {0,  1, 0}},
  {{ 1, 0, 0}, {0, 0, 1}, {0, -1, 0}},
  {{-1, 0, 0}, {0, 0, 1}, {0,  1, 0}},
  {{ 1, 0, 0}, {0, 0, 1}, {0, -1, 0}},
  {{-1, 0, 0}, {0, 0, 1}, {0,  1, 0}},
};

// joint origin xyz in the parent frame
constexpr double kJointOffset[NQ][3] = {
  {0.0, 0.0,    0.1575},
  {0.0, 0.0,    0.2025},
  {0.0, 0.2045, 0.0   },
  {0.0, 0.0,    0.2155},
  {0.0, 0.1845, 0.0   },
  {0.0, 0.0,    0.2155},
  {0.0, 0.081,  0.0   },
};

// E = Rz(q)^T R_rpy^T = [[c, s, 0], [-s, c, 0], [0, 0, 1]] * F, so
// E(i,j) = ec(i,j)*cos q + es(i,j)*sin q + e1(i,j), each coefficient in {-1,0,1}.
constexpr int ecoef_c(int k, int i, int j) {
  return i == 0 ? kFixedRot[k][0][j] : (i == 1 ? kFixedRot[k][1][j] : 0);
}
constexpr int ecoef_s(int k, int i, int j) {
  return i == 0 ? kFixedRot[k][1][j] : (i == 1 ? -kFixedRot[k][0][j] : 0);
}
constexpr int ecoef_1(int k, int i, int j) { return i == 2 ? kFixedRot[k][2][j] : 0; }

// Spatial inertia about the joint frame origin, entry (r, c):
//   [ Ibar + m(|c|^2 1 - c c^T)   m c^ ;  (m c^)^T   m 1 ]
// (the parallel-axis shift of the reference's buildI6x6, dynamicsGradient.h:77-118)
constexpr double inertia_entry(int k, int r, int c) {
  const LinkInertial L = kLinks[k];
  const double com[3] = {L.cx, L.cy, L.cz};
  if (r < 3 && c < 3) {
    const double ibar[3][3] = {{L.ixx, L.ixy, L.ixz}, {L.ixy, L.iyy, L.iyz}, {L.ixz, L.iyz, L.izz}};
    const double cc = L.cx * L.cx + L.cy * L.cy + L.cz * L.cz;
    return ibar[r][c] + L.m * ((r == c ? cc : 0.0) - com[r] * com[c]);
  }
  if (r >= 3 && c >= 3) return r == c ? L.m : 0.0;
  // skew(m c): [[0, -mcz, mcy], [mcz, 0, -mcx], [-mcy, mcx, 0]]
  // upper-right (r, c) = skew[r][c-3]; lower-left (r, c) = skew^T[r-3][c] = skew[c][r-3]
  const int i = r < 3 ? r : c, j = r < 3 ? c - 3 : r - 3;
  const double mc[3] = {L.m * L.cx, L.m * L.cy, L.m * L.cz};
  const double skew[3][3] = {{0.0, -mc[2], mc[1]}, {mc[2], 0.0, -mc[0]}, {-mc[1], mc[0], 0.0}};
  return skew[i][j];
}

}  // namespace rbd
