// SE3 / Sim3 algebra and the 6x6 / 7x7 LDL^T solve in FP64, usable on the host
// (plain LM driver) and on the device (batched LM control blocks).
//
// Replaces the Sophus SE3 / Sim3 exp, product and inverse and Eigen's
// Matrix6x6::ldlt().solve that the reference trackers use (lsd-slam
// util/SophusUtil.h; wrapper evidence PangolinOutputIOWrapper.cpp:10,
// Keyframe.h:13).  Poses are (unit quaternion, translation, scale) acting as
// X -> s R X + t; updates are left-multiplicative, T <- exp(d) T.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#define LSDG_HD __host__ __device__ inline

namespace lsdg {

struct Pose {
  double q[4];  // x, y, z, w
  double t[3];
  double s;
};

LSDG_HD Pose pose_identity() {
  Pose p;
  p.q[0] = p.q[1] = p.q[2] = 0.0;
  p.q[3] = 1.0;
  p.t[0] = p.t[1] = p.t[2] = 0.0;
  p.s = 1.0;
  return p;
}

LSDG_HD void quat_normalize(double* q) {
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}

// Rotation matrix (row-major) of a unit quaternion.
LSDG_HD void quat_R(const double* q, double* R) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
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

LSDG_HD Pose pose_mul(const Pose& a, const Pose& b) {
  Pose r;
  r.q[0] = a.q[3] * b.q[0] + a.q[0] * b.q[3] + a.q[1] * b.q[2] - a.q[2] * b.q[1];
  r.q[1] = a.q[3] * b.q[1] - a.q[0] * b.q[2] + a.q[1] * b.q[3] + a.q[2] * b.q[0];
  r.q[2] = a.q[3] * b.q[2] + a.q[0] * b.q[1] - a.q[1] * b.q[0] + a.q[2] * b.q[3];
  r.q[3] = a.q[3] * b.q[3] - a.q[0] * b.q[0] - a.q[1] * b.q[1] - a.q[2] * b.q[2];
  quat_normalize(r.q);
  double M[9];
  quat_R(a.q, M);
  for (int i = 0; i < 3; ++i)
    r.t[i] = a.s * (M[3 * i] * b.t[0] + M[3 * i + 1] * b.t[1] + M[3 * i + 2] * b.t[2]) + a.t[i];
  r.s = a.s * b.s;
  return r;
}

LSDG_HD Pose pose_inv(const Pose& a) {
  Pose r;
  r.q[0] = -a.q[0]; r.q[1] = -a.q[1]; r.q[2] = -a.q[2]; r.q[3] = a.q[3];
  double M[9];
  quat_R(a.q, M);
  r.s = 1.0 / a.s;
  for (int i = 0; i < 3; ++i) r.t[i] = -r.s * (M[i] * a.t[0] + M[3 + i] * a.t[1] + M[6 + i] * a.t[2]);
  return r;
}

LSDG_HD Pose pose_from7(const double* p) {
  Pose r;
  for (int i = 0; i < 4; ++i) r.q[i] = p[i];
  quat_normalize(r.q);
  for (int i = 0; i < 3; ++i) r.t[i] = p[4 + i];
  r.s = 1.0;
  return r;
}
LSDG_HD Pose pose_from8(const double* p) {
  Pose r = pose_from7(p);
  r.s = p[7];
  return r;
}
LSDG_HD void pose_to7(const Pose& r, double* p) {
  for (int i = 0; i < 4; ++i) p[i] = r.q[i];
  for (int i = 0; i < 3; ++i) p[4 + i] = r.t[i];
}
LSDG_HD void pose_to8(const Pose& r, double* p) {
  pose_to7(r, p);
  p[7] = r.s;
}

LSDG_HD void so3_exp_q(const double* w, double* q) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = sqrt(th2);
  double k, c;
  if (th < 1e-10) {
    k = 0.5 - th2 / 48.0;
    c = 1.0 - th2 / 8.0;
  } else {
    k = sin(0.5 * th) / th;
    c = cos(0.5 * th);
  }
  q[0] = k * w[0]; q[1] = k * w[1]; q[2] = k * w[2]; q[3] = c;
  quat_normalize(q);
}

// out = (a I + b W + c W^2) u, W = hat(w)
LSDG_HD void apply_abc(double a, double b, double c, const double* w, const double* u, double* out) {
  double wu[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
  double wwu[3] = {w[1] * wu[2] - w[2] * wu[1], w[2] * wu[0] - w[0] * wu[2], w[0] * wu[1] - w[1] * wu[0]};
  for (int i = 0; i < 3; ++i) out[i] = a * u[i] + b * wu[i] + c * wwu[i];
}

// exp of an se3 twist (v, w): R = exp(w), t = V v,
// V = I + (1 - cos th)/th^2 W + (th - sin th)/th^3 W^2.
LSDG_HD Pose se3_exp(const double* xi) {
  Pose r;
  const double* w = xi + 3;
  so3_exp_q(w, r.q);
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = sqrt(th2);
  double b, c;
  if (th < 1e-5) {
    b = 0.5 - th2 / 24.0;
    c = 1.0 / 6.0 - th2 / 120.0;
  } else {
    b = (1.0 - cos(th)) / th2;
    c = (th - sin(th)) / (th2 * th);
  }
  apply_abc(1.0, b, c, w, xi, r.t);
  r.s = 1.0;
  return r;
}

// Coefficients of V = A I + B W + C W^2 = int_0^1 e^{u sigma} exp(u W) du
// for the rotation vector w (W = hat(w)) and log-scale sigma.
LSDG_HD void sim3_v_coeffs(double sg, const double* w, double* Ao, double* Bo, double* Co) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = sqrt(th2);
  double es = exp(sg);
  double A, B, C;
  if (fabs(sg) < 1e-6) {
    A = 1.0 + 0.5 * sg + sg * sg / 6.0;
    if (th < 1e-5) {
      B = 0.5 + sg / 3.0 - th2 / 24.0;
      C = 1.0 / 6.0 + sg / 8.0 - th2 / 120.0;
    } else {
      double sn = sin(th), cs = cos(th);
      B = (1.0 - cs) / th2 + sg * (sn - th * cs) / (th2 * th);
      C = (th - sn) / (th2 * th) + sg * (0.5 * th2 - th * sn + 1.0 - cs) / (th2 * th2);
    }
  } else {
    A = (es - 1.0) / sg;
    if (th < 1e-5) {
      B = (es * (sg - 1.0) + 1.0) / (sg * sg);
      C = (es * (sg * sg - 2.0 * sg + 2.0) - 2.0) / (2.0 * sg * sg * sg);
    } else {
      double sn = sin(th), cs = cos(th);
      double den = sg * sg + th2;
      B = (es * (sg * sn - th * cs) + th) / (th * den);
      C = (A - (es * (sg * cs + th * sn) - sg) / den) / th2;
    }
  }
  *Ao = A;
  *Bo = B;
  *Co = C;
}

// exp of a sim3 twist (v, w, sigma): s = e^sigma, R = exp(w), t = V v.
LSDG_HD Pose sim3_exp(const double* xi) {
  Pose r;
  const double* w = xi + 3;
  so3_exp_q(w, r.q);
  double A, B, C;
  sim3_v_coeffs(xi[6], w, &A, &B, &C);
  apply_abc(A, B, C, w, xi, r.t);
  r.s = exp(xi[6]);
  return r;
}

// FP32 rotation (scaled by s for Sim3) and translation of a pose, as the
// per-point kernels consume them.
LSDG_HD void pose_rt_f(const Pose& T, bool scaled, float* R, float* t) {
  double M[9];
  quat_R(T.q, M);
  for (int i = 0; i < 9; ++i) R[i] = (float)(scaled ? M[i] * T.s : M[i]);
  for (int i = 0; i < 3; ++i) t[i] = (float)T.t[i];
}

// In-plane roll of R for Sim3Tracker's ESM gradient rotation: rollMat = Q^T R,
// Q the shortest rotation taking the optical axis f = (0,0,-1) to R f;
// r4 = (roll00, roll01, roll10, roll11).
LSDG_HD void pose_roll_f(const Pose& T, float* r4) {
  double R[9];
  quat_R(T.q, R);
  const double f[3] = {0.0, 0.0, -1.0};
  const double g[3] = {-R[2], -R[5], -R[8]};
  double cth = f[0] * g[0] + f[1] * g[1] + f[2] * g[2];
  double ax[3] = {f[1] * g[2] - f[2] * g[1], f[2] * g[0] - f[0] * g[2], f[0] * g[1] - f[1] * g[0]};
  double q[4];
  if (cth < -1.0 + 1e-12) {
    q[0] = 1.0; q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
  } else {
    double s = sqrt((1.0 + cth) * 2.0);
    q[0] = ax[0] / s; q[1] = ax[1] / s; q[2] = ax[2] / s; q[3] = 0.5 * s;
  }
  quat_normalize(q);
  double Q[9];
  quat_R(q, Q);
  double M[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[3 * i + j] = Q[i] * R[j] + Q[3 + i] * R[3 + j] + Q[6 + i] * R[6 + j];
  r4[0] = (float)M[0]; r4[1] = (float)M[1]; r4[2] = (float)M[3]; r4[3] = (float)M[4];
}

// LDL^T solve of a symmetric N x N system (row-major); zero pivots give zero
// solution components.
// The 31-bit draw DepthMap::initializeRandomly takes from rand() upstream,
// here deterministic: the high bits of SplitMix64's output number i + 1 from
// state `seed` (Steele, Lea, Flood 2014).
LSDG_HD int random_init_draw(unsigned long long seed, int i) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (int)(z >> 33);
}

template <int N>
LSDG_HD void ldlt_solve(const double* A, const double* b, double* x) {
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
      L[i][j] = (fabs(d) > DBL_MIN) ? v / d : 0.0;
    }
  }
  double y[N];
  for (int i = 0; i < N; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
    y[i] = v;
  }
  for (int i = 0; i < N; ++i) y[i] = (fabs(D[i]) > DBL_MIN) ? y[i] / D[i] : 0.0;
  for (int i = N - 1; i >= 0; --i) {
    double v = y[i];
    for (int k = i + 1; k < N; ++k) v -= L[k][i] * x[k];
    x[i] = v;
  }
}

}  // namespace lsdg
