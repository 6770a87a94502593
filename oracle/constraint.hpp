// oracle/constraint.hpp — TEST INFRASTRUCTURE (the oracle).  Not product
// code: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker or the timed CPU baseline.
//
// CPU restatement of lsd-slam's keyframe constraint test
// (SlamSystem::testConstraint and tryTrackSim3, GlobalMapping), the consumer
// of keyframe-to-keyframe Sim3 alignment whose accepted pairs become the
// pose-graph edges behind PangolinOutputIOWrapper::publishKeyframeGraph
// (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:148-169).  lsd-slam is
// an external dependency absent from the tree; its published algorithm is
// restated: three stages (levels 4..3, 2, 1), in each B's keyframe tracked on
// A then A's keyframe on B from their previous results, a failed direction
// ending the stage at 1e20, else the reciprocal consistency
// log(AtoB BtoA)' (Sigma_AB + Adj_AB Sigma_BA Adj_AB')^-1 log(AtoB BtoA)
// against 3000, 4000, 6000 x strictness.  Sophus' Sim3 log and adjoint are
// restated from their definitions (T exp(x) T^-1 = exp(Adj_T x)).
#pragma once
#include <cmath>
#include <cstring>

#include "lie.hpp"

namespace orc {

// log of a Sim3 as (upsilon, omega, sigma): omega from the quaternion,
// sigma = log s, upsilon = V^-1 t by Cramer's rule.
inline void sim3_log(const Sim3& T, double* xi) {
  Quat q = T.q;
  if (q.w < 0.0) q = {-q.x, -q.y, -q.z, -q.w};
  const double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  const double th = 2.0 * std::atan2(n, q.w);
  const double k = n > 1e-12 ? th / n : 2.0 / q.w;
  double* w = xi + 3;
  w[0] = k * q.x;
  w[1] = k * q.y;
  w[2] = k * q.z;
  xi[6] = std::log(T.s);
  double A, B, C;
  sim3_abc(xi[6], w, A, B, C);
  double V[3][3];
  for (int j = 0; j < 3; ++j) {
    double e[3] = {0, 0, 0}, col[3];
    e[j] = 1.0;
    apply_V(A, B, C, w, e, col);
    for (int i = 0; i < 3; ++i) V[i][j] = col[i];
  }
  const double det = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) -
                     V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                     V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
  for (int c = 0; c < 3; ++c) {
    double M[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[i][j] = j == c ? T.t[i] : V[i][j];
    const double dc = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                      M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                      M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    xi[c] = dc / det;
  }
}

// Adjoint of a Sim3 on (upsilon, omega, sigma).
inline void sim3_adj(const Sim3& T, double* M) {
  double R[9];
  T.R(R);
  std::memset(M, 0, 49 * sizeof(double));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      M[7 * i + j] = T.s * R[3 * i + j];
      M[7 * (3 + i) + 3 + j] = R[3 * i + j];
      // (t x R e_j)_i
      const double c[3] = {R[j], R[3 + j], R[6 + j]};
      const double cr[3] = {T.t[1] * c[2] - T.t[2] * c[1], T.t[2] * c[0] - T.t[0] * c[2], T.t[0] * c[1] - T.t[1] * c[0]};
      M[7 * i + 3 + j] = cr[i];
    }
  for (int i = 0; i < 3; ++i) M[7 * i + 6] = -T.t[i];
  M[48] = 1.0;
}

// Inverse of a symmetric 7x7 matrix, one LDL^T solve per column.
inline void inv7_sym(const double* A, double* out) {
  for (int j = 0; j < 7; ++j) {
    double e[7] = {0, 0, 0, 0, 0, 0, 0}, x[7];
    e[j] = 1.0;
    ldlt_solve<7>(A, e, x);
    for (int i = 0; i < 7; ++i) out[7 * i + j] = x[i];
  }
}

inline double reciprocal_consistency(const Sim3& AtoB, const double* AtoBInfo, const Sim3& BtoA,
                                     const double* BtoAInfo) {
  double Sa[49], Sb[49], J[49], M[49], H[49];
  inv7_sym(AtoBInfo, Sa);
  inv7_sym(BtoAInfo, Sb);
  sim3_adj(AtoB, J);
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 7; ++j) {
      double v = Sa[7 * i + j];
      for (int a = 0; a < 7; ++a)
        for (int b = 0; b < 7; ++b) v += J[7 * i + a] * Sb[7 * a + b] * J[7 * j + b];
      M[7 * i + j] = v;
    }
  inv7_sym(M, H);
  double d[7];
  sim3_log(AtoB * BtoA, d);
  double e = 0.0;
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 7; ++j) e += d[i] * H[7 * i + j] * d[j];
  return e;
}

}  // namespace orc
