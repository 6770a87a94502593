// C-ABI of the keyframe constraint search: lsd-slam's SlamSystem::testConstraint
// and tryTrackSim3 (GlobalMapping), the caller of keyframe-to-keyframe Sim3
// alignment whose accepted pairs become the pose-graph edges behind
// PangolinOutputIOWrapper::publishKeyframeGraph
// (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:148-169).  The
// alignments of every pair still in the running go through one batch per
// stage; the consistency test is host-side 7x7 algebra in double.
#include <cmath>
#include <cstring>
#include <vector>

#include "capi_common.h"

using namespace lsdg;

namespace {

constexpr double kSim3Eps = 1e-10;  // Sophus' epsilon<double>: the scale bounds of a usable alignment
constexpr float kFailed = 1e20f;    // tryTrackSim3's error for a failed alignment

struct Stage {
  int start, final;
  float threshold;
};
// testConstraint: levels SIM3TRACKING_MAX_LEVEL-1..3, then 2, then 1
constexpr Stage kStages[3] = {{4, 3, 3000.0f}, {2, 2, 4000.0f}, {1, 1, 6000.0f}};

// x = A^-1 b for a 3x3 system, Gaussian elimination with partial pivoting.
bool solve3(const double* Ain, const double* b, double* x) {
  double A[3][4];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A[i][j] = Ain[3 * i + j];
    A[i][3] = b[i];
  }
  for (int c = 0; c < 3; ++c) {
    int p = c;
    for (int r = c + 1; r < 3; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (A[p][c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < 4; ++j) std::swap(A[p][j], A[c][j]);
    for (int r = c + 1; r < 3; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int j = c; j < 4; ++j) A[r][j] -= f * A[c][j];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double v = A[i][3];
    for (int j = i + 1; j < 3; ++j) v -= A[i][j] * x[j];
    x[i] = v / A[i][i];
  }
  return true;
}

// log of a Sim3: (upsilon, omega, sigma) with t = V(sigma, omega) upsilon.
void sim3_log(const Pose& T, double* xi) {
  double q[4] = {T.q[0], T.q[1], T.q[2], T.q[3]};
  if (q[3] < 0.0)
    for (double& v : q) v = -v;
  const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  // rotation angle 2 atan2(|v|, w); near zero its series 2 / w (1 - n^2 / (3 w^2))
  const double k = n < 1e-10 ? 2.0 / q[3] * (1.0 - n * n / (3.0 * q[3] * q[3])) : 2.0 * std::atan2(n, q[3]) / n;
  double* w = xi + 3;
  for (int i = 0; i < 3; ++i) w[i] = k * q[i];
  xi[6] = std::log(T.s);
  double A, B, C;
  sim3_v_coeffs(xi[6], w, &A, &B, &C);
  // V = A I + B W + C W^2
  const double W[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  double V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double w2 = 0.0;
      for (int m = 0; m < 3; ++m) w2 += W[3 * i + m] * W[3 * m + j];
      V[3 * i + j] = (i == j ? A : 0.0) + B * W[3 * i + j] + C * w2;
    }
  if (!solve3(V, T.t, xi))
    for (int i = 0; i < 3; ++i) xi[i] = T.t[i];
}

// Adjoint of a Sim3 on (upsilon, omega, sigma):
// [[s R, hat(t) R, -t], [0, R, 0], [0, 0, 1]].
void sim3_adj(const Pose& T, double* M) {
  double R[9];
  quat_R(T.q, R);
  std::memset(M, 0, 49 * sizeof(double));
  const double* t = T.t;
  const double H[9] = {0.0, -t[2], t[1], t[2], 0.0, -t[0], -t[1], t[0], 0.0};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double hr = 0.0;
      for (int m = 0; m < 3; ++m) hr += H[3 * i + m] * R[3 * m + j];
      M[7 * i + j] = T.s * R[3 * i + j];
      M[7 * i + 3 + j] = hr;
      M[7 * (3 + i) + 3 + j] = R[3 * i + j];
    }
  for (int i = 0; i < 3; ++i) M[7 * i + 6] = -t[i];
  M[48] = 1.0;
}

// Inverse of a 7x7 matrix, Gauss-Jordan with partial pivoting; false if singular.
bool inv7(const double* Ain, double* out) {
  double A[7][14];
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 14; ++j) A[i][j] = j < 7 ? Ain[7 * i + j] : (j - 7 == i ? 1.0 : 0.0);
  for (int c = 0; c < 7; ++c) {
    int p = c;
    for (int r = c + 1; r < 7; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (!(std::fabs(A[p][c]) > 0.0) || !std::isfinite(A[p][c])) return false;
    if (p != c)
      for (int j = 0; j < 14; ++j) std::swap(A[p][j], A[c][j]);
    const double d = A[c][c];
    for (int j = 0; j < 14; ++j) A[c][j] /= d;
    for (int r = 0; r < 7; ++r) {
      if (r == c) continue;
      const double f = A[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 14; ++j) A[r][j] -= f * A[c][j];
    }
  }
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 7; ++j) out[7 * i + j] = A[i][7 + j];
  return true;
}

// tryTrackSim3's consistency of the two directions A -> B and B -> A.
float reciprocal_consistency(const Pose& AtoB, const double* AtoBInfo, const Pose& BtoA, const double* BtoAInfo) {
  double Sa[49], Sb[49], J[49], M[49], H[49];
  if (!inv7(AtoBInfo, Sa) || !inv7(BtoAInfo, Sb)) return kFailed;
  sim3_adj(AtoB, J);
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 7; ++j) {
      double v = 0.0;
      for (int a = 0; a < 7; ++a) {
        double ja = 0.0;
        for (int b = 0; b < 7; ++b) ja += Sb[7 * a + b] * J[7 * j + b];
        v += J[7 * i + a] * ja;
      }
      M[7 * i + j] = Sa[7 * i + j] + v;
    }
  if (!inv7(M, H)) return kFailed;
  double d[7];
  sim3_log(pose_mul(AtoB, BtoA), d);
  double e = 0.0;
  for (int i = 0; i < 7; ++i) {
    double hd = 0.0;
    for (int j = 0; j < 7; ++j) hd += H[7 * i + j] * d[j];
    e += hd * d[i];
  }
  return (float)e;
}

// The checks tryTrackSim3 applies to one direction's alignment.
bool usable(const lsdg_job_result& r) {
  return r.status != LSDG_DIVERGED && !r.stats.diverged && r.frame_to_ref[7] <= 1.0 / kSim3Eps &&
         r.frame_to_ref[7] >= kSim3Eps && r.stats.hessian[0] != 0.0 && r.stats.hessian[48] != 0.0;
}

void edge_of(const lsdg_job_result& r, int first, int second, lsdg_constraint_edge& e) {
  e.first_id = first;
  e.second_id = second;
  std::memcpy(e.second_to_first, r.frame_to_ref, sizeof(e.second_to_first));
  std::memcpy(e.information, r.stats.hessian, sizeof(e.information));
  e.mean_residual = r.stats.last_residual;
  e.mean_residual_d = r.stats.last_depth_residual;
  e.mean_residual_p = r.stats.last_photometric_residual;
  e.usage = r.stats.point_usage;
}

}  // namespace

extern "C" lsdg_status lsdg_constraint_search(lsdg_ctx c, const lsdg_constraint_candidate* cand, int32_t n,
                                              float strictness, lsdg_constraint_result* out) {
  if (!c || n < 0 || (n > 0 && (!cand || !out))) return LSDG_EINVAL;
  for (int i = 0; i < n; ++i)
    if (!cand[i].frame_ref || !cand[i].candidate_ref || cand[i].frame_ref->ctx != c ||
        cand[i].candidate_ref->ctx != c)
      return fail(c, LSDG_EINVAL, "constraint candidate without tracking references of this context");
  std::vector<Pose> ftoc(n), ctof(n);
  std::vector<char> alive(n, 1);
  for (int i = 0; i < n; ++i) {
    std::memset(&out[i], 0, sizeof(out[i]));
    ctof[i] = pose_from8(cand[i].candidate_to_frame_init);
    ftoc[i] = pose_inv(ctof[i]);
  }
  std::vector<lsdg_job> jobs;
  std::vector<lsdg_job_result> res;
  std::vector<int> who;
  for (int s = 0; s < 3; ++s) {
    jobs.clear();
    who.clear();
    for (int i = 0; i < n; ++i) {
      if (!alive[i]) continue;
      who.push_back(i);
      lsdg_job j{};
      j.kind = LSDG_JOB_SIM3;
      j.start_level = kStages[s].start;
      j.final_level = kStages[s].final;
      j.rebuild_pyramid = 0;
      // B's keyframe on A (candidate to frame), then A's keyframe on B
      j.ref = cand[i].frame_ref;
      j.frame = cand[i].candidate_ref->kf;
      pose_to8(ctof[i], j.frame_to_ref_init);
      jobs.push_back(j);
      j.ref = cand[i].candidate_ref;
      j.frame = cand[i].frame_ref->kf;
      pose_to8(ftoc[i], j.frame_to_ref_init);
      jobs.push_back(j);
    }
    if (who.empty()) break;
    res.assign(jobs.size(), lsdg_job_result{});
    lsdg_status st = lsdg_batch_track(c, jobs.data(), (int32_t)jobs.size(), res.data());
    if (st != LSDG_OK) return st;
    for (size_t k = 0; k < who.size(); ++k) {
      const int i = who[k];
      const lsdg_job_result& rb = res[2 * k];      // BtoA: candidate on frame
      const lsdg_job_result& ra = res[2 * k + 1];  // AtoB: frame on candidate
      float err;
      // tryTrackSim3 keeps the first direction's result and returns before
      // tracking the second when the first is not usable
      ctof[i] = pose_from8(rb.frame_to_ref);
      if (!usable(rb)) {
        err = kFailed;
      } else {
        ftoc[i] = pose_from8(ra.frame_to_ref);
        err = usable(ra) ? reciprocal_consistency(ftoc[i], ra.stats.hessian, ctof[i], rb.stats.hessian) : kFailed;
      }
      out[i].error[s] = err;
      if (!(err <= kStages[s].threshold * strictness)) {
        alive[i] = 0;
        out[i].rejected_at = s + 1;
      } else if (s == 2) {
        out[i].accepted = 1;
        edge_of(rb, cand[i].frame_ref->kf->id, cand[i].candidate_ref->kf->id, out[i].candidate_on_frame);
        edge_of(ra, cand[i].candidate_ref->kf->id, cand[i].frame_ref->kf->id, out[i].frame_on_candidate);
      }
    }
  }
  for (int i = 0; i < n; ++i) {
    pose_to8(ctof[i], out[i].candidate_to_frame);
    pose_to8(ftoc[i], out[i].frame_to_candidate);
  }
  return LSDG_OK;
}

// ------------------------------------------------------------- candidates
extern "C" lsdg_status lsdg_keyframe_candidates(const lsdg_camera* cam, const lsdg_keyframe_pose* kf,
                                                const lsdg_keyframe_pose* others, int32_t n, float closeness_th,
                                                float kf_dist_weight, int32_t* out_index, double* dist2,
                                                double* cos_angle, int32_t cap, int32_t* count) {
  if (!cam || !kf || (!others && n > 0) || n < 0 || !count || cap < 0 || (cap > 0 && !out_index) ||
      cam->fx <= 0.0f || cam->fy <= 0.0f || !(kf_dist_weight > 0.0f))
    return LSDG_EINVAL;
  // the thresholds findCandidates passes to findEuclideanOverlapFrames
  const double distTh = (double)closeness_th * 15.0 / ((double)kf_dist_weight * kf_dist_weight);
  const double angleTh = 1.0 - 0.25 * (double)closeness_th;
  const double fovX = 2.0 * std::atan((double)cam->w / cam->fx / 2.0);
  const double fovY = 2.0 * std::atan((double)cam->h / cam->fy / 2.0);
  const double cosTh = std::cos(angleTh * 0.5 * (fovX + fovY));
  const Pose P = pose_from8(kf->cam_to_world);
  double R[9];
  quat_R(P.q, R);
  const double zx = R[2], zy = R[5], zz = R[8];  // optical axis in the world
  const double facSelf = (double)kf->mean_idepth / P.s;
  int found = 0;
  for (int i = 0; i < n; ++i) {
    const Pose Q = pose_from8(others[i].cam_to_world);
    double fac = (double)others[i].mean_idepth / Q.s;
    if (facSelf < fac) fac = facSelf;
    const double dx = (P.t[0] - Q.t[0]) * fac, dy = (P.t[1] - Q.t[1]) * fac, dz = (P.t[2] - Q.t[2]) * fac;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 <= distTh)) continue;
    double S[9];
    quat_R(Q.q, S);
    const double c = S[2] * zx + S[5] * zy + S[8] * zz;
    if (c < cosTh) continue;
    if (found < cap) {
      out_index[found] = i;
      if (dist2) dist2[found] = d2;
      if (cos_angle) cos_angle[found] = c;
    }
    ++found;
  }
  *count = found;
  return LSDG_OK;
}
