// oracle/lie.hpp — TEST INFRASTRUCTURE (the oracle).  Not product code: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// SE3 / Sim3 group algebra and the LDL^T solve in double, restated from the
// textbook closed forms (Rodrigues; Sim3 exp with the V matrix, Strasdat 2012 /
// Eade's Lie-group notes).  The reference uses Sophus and Eigen for these
// (lsd-slam util/SophusUtil.h; wrapper evidence: PangolinOutputIOWrapper.cpp:10,
// Keyframe.h:13), neither of which is present here.
#pragma once
#include <cfloat>
#include <cmath>

namespace orc {

struct Quat {
  double x, y, z, w;
};

inline Quat qmul(const Quat& a, const Quat& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
inline Quat qnormalize(Quat q) {
  double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return {q.x / n, q.y / n, q.z / n, q.w / n};
}
inline Quat qconj(const Quat& q) { return {-q.x, -q.y, -q.z, q.w}; }

// R (row-major) of a unit quaternion.
inline void qmat(const Quat& q, double* R) {
  const double x = q.x, y = q.y, z = q.z, w = q.w;
  R[0] = 1.0 - 2.0 * (y * y + z * z);
  R[1] = 2.0 * (x * y - z * w);
  R[2] = 2.0 * (x * z + y * w);
  R[3] = 2.0 * (x * y + z * w);
  R[4] = 1.0 - 2.0 * (x * x + z * z);
  R[5] = 2.0 * (y * z - x * w);
  R[6] = 2.0 * (x * z - y * w);
  R[7] = 2.0 * (y * z + x * w);
  R[8] = 1.0 - 2.0 * (x * x + y * y);
}

// Similarity X -> s R X + t (an SE3 when s == 1).
struct Sim3 {
  Quat q{0, 0, 0, 1};
  double t[3]{0, 0, 0};
  double s = 1.0;

  void R(double* M) const { qmat(q, M); }
  Sim3 operator*(const Sim3& b) const {
    Sim3 r;
    r.q = qnormalize(qmul(q, b.q));
    double M[9];
    qmat(q, M);
    for (int i = 0; i < 3; ++i)
      r.t[i] = s * (M[3 * i] * b.t[0] + M[3 * i + 1] * b.t[1] + M[3 * i + 2] * b.t[2]) + t[i];
    r.s = s * b.s;
    return r;
  }
  Sim3 inverse() const {
    Sim3 r;
    r.q = qconj(q);
    double M[9];
    qmat(q, M);
    r.s = 1.0 / s;
    for (int i = 0; i < 3; ++i)
      r.t[i] = -r.s * (M[i] * t[0] + M[3 + i] * t[1] + M[6 + i] * t[2]);
    return r;
  }
  static Sim3 from7(const double* p) {
    Sim3 r;
    r.q = qnormalize({p[0], p[1], p[2], p[3]});
    r.t[0] = p[4]; r.t[1] = p[5]; r.t[2] = p[6];
    r.s = 1.0;
    return r;
  }
  static Sim3 from8(const double* p) {
    Sim3 r = from7(p);
    r.s = p[7];
    return r;
  }
  void to7(double* p) const {
    p[0] = q.x; p[1] = q.y; p[2] = q.z; p[3] = q.w;
    p[4] = t[0]; p[5] = t[1]; p[6] = t[2];
  }
  void to8(double* p) const { to7(p); p[7] = s; }
};

inline Quat so3_exp(const double* w) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = std::sqrt(th2);
  double k, c;
  if (th < 1e-10) {
    k = 0.5 - th2 / 48.0;
    c = 1.0 - th2 / 8.0;
  } else {
    k = std::sin(0.5 * th) / th;
    c = std::cos(0.5 * th);
  }
  return qnormalize({k * w[0], k * w[1], k * w[2], c});
}

// t = (a I + b W + c W^2) u with W = hat(w).
inline void apply_V(double a, double b, double c, const double* w, const double* u, double* out) {
  double wxu[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
  double wxwxu[3] = {w[1] * wxu[2] - w[2] * wxu[1], w[2] * wxu[0] - w[0] * wxu[2],
                     w[0] * wxu[1] - w[1] * wxu[0]};
  for (int i = 0; i < 3; ++i) out[i] = a * u[i] + b * wxu[i] + c * wxwxu[i];
}

// SE3 exp of (v1,v2,v3,w1,w2,w3).
inline Sim3 se3_exp(const double* xi) {
  Sim3 r;
  const double* w = xi + 3;
  r.q = so3_exp(w);
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = std::sqrt(th2);
  double b, c;
  if (th < 1e-5) {
    b = 0.5 - th2 / 24.0;
    c = 1.0 / 6.0 - th2 / 120.0;
  } else {
    b = (1.0 - std::cos(th)) / th2;
    c = (th - std::sin(th)) / (th2 * th);
  }
  apply_V(1.0, b, c, w, xi, r.t);
  r.s = 1.0;
  return r;
}

// Coefficients of V = A I + B W + C W^2 = int_0^1 e^{s sigma} exp(s W) ds.
inline void sim3_abc(double sg, const double* w, double& A, double& B, double& C) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = std::sqrt(th2);
  double es = std::exp(sg);
  if (std::fabs(sg) < 1e-6) {
    A = 1.0 + 0.5 * sg + sg * sg / 6.0;
    if (th < 1e-5) {
      B = 0.5 + sg / 3.0 - th2 / 24.0;
      C = 1.0 / 6.0 + sg / 8.0 - th2 / 120.0;
    } else {
      // first order in sigma
      double s = std::sin(th), c = std::cos(th);
      B = (1.0 - c) / th2 + sg * (s - th * c) / (th2 * th);
      C = (th - s) / (th2 * th) + sg * (0.5 * th2 - th * s + 1.0 - c) / (th2 * th2);
    }
  } else {
    A = (es - 1.0) / sg;
    if (th < 1e-5) {
      B = (es * (sg - 1.0) + 1.0) / (sg * sg);
      C = (es * (sg * sg - 2.0 * sg + 2.0) - 2.0) / (2.0 * sg * sg * sg);
    } else {
      double s = std::sin(th), c = std::cos(th);
      double den = sg * sg + th2;
      B = (es * (sg * s - th * c) + th) / (th * den);
      C = (A - (es * (sg * c + th * s) - sg) / den) / th2;
    }
  }
}

// Sim3 exp of (v1,v2,v3,w1,w2,w3,sigma): s = e^sigma, R = exp(w), t = V u.
inline Sim3 sim3_exp(const double* xi) {
  Sim3 r;
  const double* w = xi + 3;
  r.q = so3_exp(w);
  double A, B, C;
  sim3_abc(xi[6], w, A, B, C);
  apply_V(A, B, C, w, xi, r.t);
  r.s = std::exp(xi[6]);
  return r;
}

// Solve A x = b for symmetric A (N x N, row-major) by LDL^T without
// pivoting; zero pivots give zero solution components (pseudo-inverse on
// the null directions, as Eigen's LDLT::solve does).
template <int N>
inline void ldlt_solve(const double* A, const double* b, double* x) {
  double L[N][N];
  double D[N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) L[i][j] = (i == j) ? 1.0 : 0.0;
  for (int j = 0; j < N; ++j) {
    double d = A[j * N + j];
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k] * D[k];
    D[j] = d;
    for (int i = j + 1; i < N; ++i) {
      double v = A[i * N + j];
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k] * D[k];
      L[i][j] = (std::fabs(d) > DBL_MIN) ? v / d : 0.0;
    }
  }
  double y[N];
  for (int i = 0; i < N; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
    y[i] = v;
  }
  for (int i = 0; i < N; ++i) y[i] = (std::fabs(D[i]) > DBL_MIN) ? y[i] / D[i] : 0.0;
  for (int i = N - 1; i >= 0; --i) {
    double v = y[i];
    for (int k = i + 1; k < N; ++k) v -= L[k][i] * x[k];
    x[i] = v;
  }
}

}  // namespace orc
