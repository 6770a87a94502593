// oracle/oracle_capi.cpp — TEST INFRASTRUCTURE (the oracle).  C entry points of
// the CPU restatement for the tests (ctypes), smoke() and bench.py's
// cpu_baseline leg.  Handles carry the precision they were created with
// (0 = FP32 restated reference path, 1 = FP64 oracle).
#include <cstring>
#include <memory>

#include "constraint.hpp"
#include "ingest.hpp"
#include "keyframe_export.hpp"
#include "lsd_oracle.hpp"

namespace orc {

// Upstream defaults: settings.h / settings.cpp and DenseDepthTrackerSettings.
void params_default(lsdg_params* p) {
  std::memset(p, 0, sizeof(*p));
  const int its[PYR] = {5, 20, 50, 100, 100};
  for (int l = 0; l < PYR; ++l) {
    p->lambda_initial[l] = 0.0f;
    p->lambda_success_fac[l] = 0.5f;
    p->lambda_fail_fac[l] = 2.0f;
    p->step_size_min[l] = 1e-8f;
    p->convergence_eps[l] = 0.999f;
    p->max_its_per_lvl[l] = its[l];
  }
  p->var_weight = 1.0f;
  p->huber_d = 3.0f;
  p->camera_pixel_noise2 = 4.0f * 4.0f;
  p->se3_min_level = 1;
  p->se3_max_level = 5;
  p->min_goodperall_pixel = 0.04f;
  p->min_goodperall_pixel_absmin = 0.01f;
  p->min_goodpergoodbad_pixel = 0.5f;
  p->max_diff_constant = 40.0f * 40.0f;
  p->max_diff_grad_mult = 0.5f * 0.5f;
  p->use_affine_lightning_estimation = 0;
  p->use_esm_tracking = 1;
  p->min_abs_grad_create = 5.0f;
  p->min_abs_grad_decrease = 5.0f;
  p->min_epl_length_squared = 1.0f * 1.0f;
  p->min_epl_grad_squared = 2.0f * 2.0f;
  p->min_epl_angle_squared = 0.3f * 0.3f;
  p->gradient_sample_dist = 1.0f;
  p->sample_point_to_border = 7.0f;
  p->max_epl_length_crop = 30.0f;
  p->min_epl_length_crop = 3.0f;
  p->max_error_stereo = 1300.0f;
  p->min_distance_error_stereo = 1.5f;
  p->stereo_epl_var_fac = 2.0f;
  p->min_depth = 0.05f;
  p->max_var = 0.5f * 0.5f;
  p->var_random_init_initial = 0.5f * 0.5f * 0.5f;
  p->var_gt_init_initial = 0.01f * 0.01f;
  p->succ_var_inc_fac = 1.01f;
  p->fail_var_inc_fac = 1.1f;
  p->diff_fac_observe = 1.0f;
  p->diff_fac_prop_merge = 1.0f;
  p->diff_fac_smoothing = 1.0f;
  p->depth_smoothing_factor = 1.0f;
  p->division_eps = 1e-10f;
  p->validity_counter_inc = 5;
  p->validity_counter_dec = 5;
  p->validity_counter_initial_observe = 5;
  p->validity_counter_max = 5;
  p->validity_counter_max_variable = 250;
  p->val_sum_min_for_create = 30;
  p->val_sum_min_for_keep = 24;
  p->val_sum_min_for_unblacklist = 100;
  p->min_blacklist = -1;
  p->allow_negative_idepths = 1;
  p->use_subpixel_stereo = 1;
  p->stereo_subpixel_oob_guard = 1;
  p->stereo_max_steps = 1024;
}

}  // namespace orc

using namespace orc;

struct OFrame {
  int prec;
  bool dbl;
  Frame<float> f32;
  Frame<double> f64;
};
struct ORef {
  int prec;
  bool dbl;
  OFrame* kf;
  TrackingReference<float> r32;
  TrackingReference<double> r64;
};
struct ODepth {
  int prec;
  bool dbl;
  DepthMap<float> d32;
  DepthMap<double> d64;
};

template <class S>
static Frame<S>& fr(OFrame* f);
template <>
Frame<float>& fr<float>(OFrame* f) { return f->f32; }
template <>
Frame<double>& fr<double>(OFrame* f) { return f->f64; }
template <class S>
static TrackingReference<S>& rf(ORef* r);
template <>
TrackingReference<float>& rf<float>(ORef* r) { return r->r32; }
template <>
TrackingReference<double>& rf<double>(ORef* r) { return r->r64; }
template <class S>
static DepthMap<S>& dm(ODepth* d);
template <>
DepthMap<float>& dm<float>(ODepth* d) { return d->d32; }
template <>
DepthMap<double>& dm<double>(ODepth* d) { return d->d64; }

#define DISPATCH(prec, call) ((prec) ? call<double> : call<float>)

template <class S, class Acc>
static void fill_stats(lsdg_track_stats* st, const SE3Tracker<S, Acc>& t) {
  std::memset(st, 0, sizeof(*st));
  st->last_residual = (float)t.lastResidual;
  st->last_mean_res = (float)t.lastMeanRes;
  st->point_usage = (float)t.pointUsage;
  st->affine_a = (float)t.aff_a;
  st->affine_b = (float)t.aff_b;
  st->last_good_count = (int)t.lastGoodCount;
  st->last_bad_count = (int)t.lastBadCount;
  st->diverged = t.diverged;
  st->tracking_was_good = t.trackingWasGood;
  for (int l = 0; l < PYR; ++l) {
    st->num_residual_calls[l] = t.numResidualCalls[l];
    st->num_warp_update_calls[l] = t.numWarpCalls[l];
  }
}

template <class S>
static void fill_eval(lsdg_eval* o, const EvalSums<S>& e) {
  std::memset(o, 0, sizeof(*o));
  for (int i = 0; i < 49; ++i) o->A[i] = (double)e.A[i];
  for (int i = 0; i < 7; ++i) o->b[i] = (double)e.b[i];
  o->error = (double)e.err;
  o->sum_res_p = (double)e.sumResP;
  o->sum_res_d = (double)e.sumResD;
  o->num_terms_p = e.numTermsP;
  o->num_terms_d = e.numTermsD;
  o->n_constraints = e.n;
  o->warped_count = e.warped;
  o->good_count = e.good;
  o->bad_count = e.bad;
  o->sum_res_unweighted = (double)e.sumResUnw;
  o->sum_signed_res = (double)e.sumSigned;
  o->usage = (double)e.usage;
  o->sxx = (double)e.sxx; o->syy = (double)e.syy; o->sx = (double)e.sx; o->sy = (double)e.sy; o->sw = (double)e.sw;
}

extern "C" void orc_params_default(lsdg_params* p) { params_default(p); }

extern "C" void* orc_frame_new(int prec, int id, int w, int h, float fx, float fy, float cx, float cy, const float* img,
                    const lsdg_params* P) {
  // propagateDepth indexes the level-1 tracking mask with (x >> 1, y >> 1):
  // odd sizes would read past it (the C-ABI refuses them as well)
  if (w < 2 || h < 2 || (w & 1) || (h & 1)) return nullptr;
  OFrame* f = new OFrame();
  f->prec = prec;
  f->dbl = prec == 1;
  if (f->dbl) {
    f->f64.init(id, w, h, fx, fy, cx, cy);
    f->f64.setImage(img, *P);
  } else {
    f->f32.init(id, w, h, fx, fy, cx, cy);
    f->f32.setImage(img, *P);
  }
  return f;
}
extern "C" void orc_frame_free(void* f) { delete (OFrame*)f; }

template <class S>
static int frame_read_t(OFrame* F, int what, int l, double* out64, float* out32) {
  Frame<S>& f = fr<S>(F);
  const std::vector<S>* v = nullptr;
  switch (what) {
    case LSDG_WHAT_IMAGE: v = &f.image[l]; break;
    case LSDG_WHAT_GRADIENTS: v = &f.grad[l]; break;
    case LSDG_WHAT_MAXGRAD: v = &f.maxg[l]; break;
    case LSDG_WHAT_IDEPTH: v = &f.idepth[l]; break;
    case LSDG_WHAT_IDEPTH_VAR: v = &f.ivar[l]; break;
    case LSDG_WHAT_REF_PIXEL_WAS_GOOD: {
      const uint8_t* m = f.refPixelWasGoodNoCreate();
      size_t n = (size_t)f.w[1] * f.h[1];
      for (size_t i = 0; i < n; ++i) {
        double x = m ? m[i] : 1;
        if (out64) out64[i] = x;
        if (out32) out32[i] = (float)x;
      }
      return 0;
    }
    default: return LSDG_EINVAL;
  }
  for (size_t i = 0; i < v->size(); ++i) {
    if (out64) out64[i] = (double)(*v)[i];
    if (out32) out32[i] = (float)(*v)[i];
  }
  return 0;
}
extern "C" int orc_frame_read_f32(void* f, int what, int level, float* out) {
  OFrame* F = (OFrame*)f;
  return F->dbl ? frame_read_t<double>(F, what, level, nullptr, out) : frame_read_t<float>(F, what, level, nullptr, out);
}
extern "C" int orc_frame_read_f64(void* f, int what, int level, double* out) {
  OFrame* F = (OFrame*)f;
  return F->dbl ? frame_read_t<double>(F, what, level, out, nullptr) : frame_read_t<float>(F, what, level, out, nullptr);
}
extern "C" void orc_frame_num_mappable(void* f, float* out) {
  OFrame* F = (OFrame*)f;
  *out = F->dbl ? F->f64.numMappablePixels : F->f32.numMappablePixels;
}

template <class S>
static void set_depth_t(OFrame* F, const float* id, const float* var) {
  Frame<S>& f = fr<S>(F);
  size_t n = (size_t)f.w[0] * f.h[0];
  std::vector<S> a(n), b(n);
  for (size_t i = 0; i < n; ++i) {
    bool ok = std::isfinite(id[i]) && var[i] > 0 && id[i] > 0;
    a[i] = ok ? (S)id[i] : (S)-1;
    b[i] = ok ? (S)var[i] : (S)-1;
  }
  f.setDepth0(a.data(), b.data());
}
extern "C" void orc_frame_set_depth(void* f, const float* id, const float* var) {
  OFrame* F = (OFrame*)f;
  if (F->dbl) set_depth_t<double>(F, id, var);
  else set_depth_t<float>(F, id, var);
}

template <class S>
static void set_tracking_t(OFrame* F, OFrame* parent, const double* tp, float itr, const uint8_t* wg) {
  Frame<S>& f = fr<S>(F);
  f.parent = parent ? &fr<S>(parent) : nullptr;
  f.parentId = parent ? fr<S>(parent).id : -1;
  f.thisToParent = Sim3::from8(tp);
  f.initialTrackedResidual = (S)itr;
  if (wg) {
    uint8_t* m = f.refPixelWasGood();
    std::memcpy(m, wg, (size_t)f.w[1] * f.h[1]);
  }
}
extern "C" void orc_frame_set_tracking(void* f, void* parent, const double* tp, float itr, const uint8_t* wg) {
  OFrame* F = (OFrame*)f;
  if (F->dbl) set_tracking_t<double>(F, (OFrame*)parent, tp, itr, wg);
  else set_tracking_t<float>(F, (OFrame*)parent, tp, itr, wg);
}
extern "C" void orc_frame_get_tracking(void* f, int* pid, double* tp, float* itr) {
  OFrame* F = (OFrame*)f;
  if (F->dbl) {
    *pid = F->f64.parentId; F->f64.thisToParent.to8(tp); *itr = (float)F->f64.initialTrackedResidual;
  } else {
    *pid = F->f32.parentId; F->f32.thisToParent.to8(tp); *itr = (float)F->f32.initialTrackedResidual;
  }
}
extern "C" void orc_frame_set_cam_to_world(void* f, const double* c2w) {
  OFrame* F = (OFrame*)f;
  if (F->dbl) F->f64.camToWorld = Sim3::from8(c2w);
  else F->f32.camToWorld = Sim3::from8(c2w);
}
extern "C" void orc_frame_counters(void* f, int* tracked, int* mapped, int set) {
  OFrame* F = (OFrame*)f;
  int& t = F->dbl ? F->f64.numFramesTrackedOnThis : F->f32.numFramesTrackedOnThis;
  int& m = F->dbl ? F->f64.numMappedOnThis : F->f32.numMappedOnThis;
  if (set) { t = *tracked; m = *mapped; }
  else { *tracked = t; *mapped = m; }
}

extern "C" void* orc_ref_new(void* kf) {
  OFrame* F = (OFrame*)kf;
  ORef* r = new ORef();
  r->prec = F->prec;
  r->dbl = F->dbl;
  r->kf = F;
  r->r32.kf = &F->f32;
  r->r64.kf = &F->f64;
  return r;
}
extern "C" void orc_ref_free(void* r) { delete (ORef*)r; }
extern "C" int orc_ref_num_points(void* r, int level) {
  ORef* R = (ORef*)r;
  if (R->dbl) { R->r64.makePointCloud(level); return R->r64.num[level]; }
  R->r32.makePointCloud(level);
  return R->r32.num[level];
}
template <class S>
static void read_points_t(TrackingReference<S>& r, int l, float* out) {
  r.makePointCloud(l);
  for (int i = 0; i < r.num[l]; ++i) {
    float* o = out + 8 * i;
    o[0] = (float)r.pos[l][3 * i]; o[1] = (float)r.pos[l][3 * i + 1]; o[2] = (float)r.pos[l][3 * i + 2];
    o[3] = (float)r.colVar[l][2 * i]; o[4] = (float)r.colVar[l][2 * i + 1];
    o[5] = (float)r.gradd[l][2 * i]; o[6] = (float)r.gradd[l][2 * i + 1];
    o[7] = (float)r.idx[l][i];
  }
}
extern "C" void orc_ref_read_points(void* r, int level, float* out) {
  ORef* R = (ORef*)r;
  if (R->dbl) read_points_t<double>(R->r64, level, out);
  else read_points_t<float>(R->r32, level, out);
}

template <class S, class Acc>
static int se3_track_t(ORef* R, OFrame* F, const lsdg_params* P, const double* init, double* out,
                       lsdg_track_stats* st) {
  SE3Tracker<S, Acc> t;
  t.P = *P;
  Sim3 res = t.trackFrame(rf<S>(R), fr<S>(F), Sim3::from7(init));
  res.to7(out);
  if (st) fill_stats(st, t);
  return t.diverged ? LSDG_DIVERGED : LSDG_OK;
}
extern "C" int orc_se3_track(void* r, void* f, const lsdg_params* P, const double* init, double* out, lsdg_track_stats* st) {
  ORef* R = (ORef*)r;
  if (R->prec == 1) return se3_track_t<double, double>(R, (OFrame*)f, P, init, out, st);
  if (R->prec == 2) return se3_track_t<float, double>(R, (OFrame*)f, P, init, out, st);
  return se3_track_t<float, float>(R, (OFrame*)f, P, init, out, st);
}

template <class S, class Acc>
static int se3_eval_t(ORef* R, OFrame* F, const lsdg_params* P, int lvl, const double* pose, float a, float b,
                      lsdg_eval* out, uint8_t* isGood) {
  SE3Tracker<S, Acc> t;
  t.P = *P;
  Sim3 T = Sim3::from7(pose);
  EvalSums<Acc> e = t.evaluate(rf<S>(R), fr<S>(F), T, lvl, isGood, (S)a, (S)b);
  fill_eval(out, e);
  return 0;
}
extern "C" int orc_se3_evaluate(void* r, void* f, const lsdg_params* P, int lvl, const double* pose, float a, float b,
                     lsdg_eval* out, uint8_t* isGood) {
  ORef* R = (ORef*)r;
  if (R->prec == 1) return se3_eval_t<double, double>(R, (OFrame*)f, P, lvl, pose, a, b, out, isGood);
  if (R->prec == 2) return se3_eval_t<float, double>(R, (OFrame*)f, P, lvl, pose, a, b, out, isGood);
  return se3_eval_t<float, float>(R, (OFrame*)f, P, lvl, pose, a, b, out, isGood);
}

template <class S, class Acc>
static int sim3_track_t(ORef* R, OFrame* F, const lsdg_params* P, const double* init, int sl, int fl, double* out,
                        lsdg_track_stats* st) {
  Sim3Tracker<S, Acc> t;
  t.P = *P;
  Sim3 res = t.trackFrameSim3(rf<S>(R), fr<S>(F), Sim3::from8(init), sl, fl);
  res.to8(out);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->last_residual = (float)t.lastResidual;
    st->last_depth_residual = (float)t.lastDepthResidual;
    st->last_photometric_residual = (float)t.lastPhotometricResidual;
    st->point_usage = (float)t.pointUsage;
    st->diverged = t.diverged;
    st->tracking_was_good = !t.diverged;
    for (int l = 0; l < PYR; ++l) {
      st->num_residual_calls[l] = t.numResidualCalls[l];
      st->num_warp_update_calls[l] = t.numWarpCalls[l];
    }
    for (int i = 0; i < 49; ++i) st->hessian[i] = t.lastHessian[i];
  }
  return t.diverged ? LSDG_DIVERGED : LSDG_OK;
}
extern "C" int orc_sim3_track(void* r, void* f, const lsdg_params* P, const double* init, int sl, int fl, double* out,
                   lsdg_track_stats* st) {
  ORef* R = (ORef*)r;
  if (R->prec == 1) return sim3_track_t<double, double>(R, (OFrame*)f, P, init, sl, fl, out, st);
  if (R->prec == 2) return sim3_track_t<float, double>(R, (OFrame*)f, P, init, sl, fl, out, st);
  return sim3_track_t<float, float>(R, (OFrame*)f, P, init, sl, fl, out, st);
}
template <class S, class Acc>
static int sim3_eval_t(ORef* R, OFrame* F, const lsdg_params* P, int lvl, const double* pose, lsdg_eval* out) {
  Sim3Tracker<S, Acc> t;
  t.P = *P;
  EvalSums<Acc> e = t.evaluate(rf<S>(R), fr<S>(F), Sim3::from8(pose), lvl);
  fill_eval(out, e);
  return 0;
}
extern "C" int orc_sim3_evaluate(void* r, void* f, const lsdg_params* P, int lvl, const double* pose, lsdg_eval* out) {
  ORef* R = (ORef*)r;
  if (R->prec == 1) return sim3_eval_t<double, double>(R, (OFrame*)f, P, lvl, pose, out);
  if (R->prec == 2) return sim3_eval_t<float, double>(R, (OFrame*)f, P, lvl, pose, out);
  return sim3_eval_t<float, float>(R, (OFrame*)f, P, lvl, pose, out);
}

// Energy (frozen weights) and its gradient; grad_d has one entry per point.
extern "C" int orc_energy_grad(void* r, void* f, const lsdg_params* P, int lvl, const double* pose, double* energy,
                    double* gxi, double* grad_d) {
  ORef* R = (ORef*)r;
  Sim3 T = Sim3::from7(pose);
  if (R->dbl) {
    Adjoint<double> a;
    a.P = *P;
    std::vector<double> g((size_t)orc_ref_num_points(r, lvl));
    a.eval(R->r64, ((OFrame*)f)->f64, T, lvl, energy, gxi, g.data());
    if (grad_d) std::copy(g.begin(), g.end(), grad_d);
  } else if (R->prec == 2) {
    Adjoint<float, double> a;
    a.P = *P;
    std::vector<float> g((size_t)orc_ref_num_points(r, lvl));
    a.eval(R->r32, ((OFrame*)f)->f32, T, lvl, energy, gxi, g.data());
    if (grad_d)
      for (size_t i = 0; i < g.size(); ++i) grad_d[i] = g[i];
  } else {
    Adjoint<float> a;
    a.P = *P;
    std::vector<float> g((size_t)orc_ref_num_points(r, lvl));
    a.eval(R->r32, ((OFrame*)f)->f32, T, lvl, energy, gxi, g.data());
    if (grad_d)
      for (size_t i = 0; i < g.size(); ++i) grad_d[i] = g[i];
  }
  return 0;
}

// Sim3 adjoint (frozen-weight reweighted least-squares energy); grad_d has
// one entry per point.
extern "C" int orc_energy_grad_sim3(void* r, void* f, const lsdg_params* P, int lvl, const double* pose8,
                                    double* energy, double* gxi, double* grad_d) {
  ORef* R = (ORef*)r;
  Sim3 T = Sim3::from8(pose8);
  const size_t n = (size_t)orc_ref_num_points(r, lvl);
  if (R->dbl) {
    AdjointSim3<double> a;
    a.P = *P;
    std::vector<double> g(n);
    a.eval(R->r64, ((OFrame*)f)->f64, T, lvl, energy, gxi, g.data());
    if (grad_d) std::copy(g.begin(), g.end(), grad_d);
    return 0;
  }
  std::vector<float> g(n);
  if (R->prec == 2) {
    AdjointSim3<float, double> a;
    a.P = *P;
    a.eval(R->r32, ((OFrame*)f)->f32, T, lvl, energy, gxi, g.data());
  } else {
    AdjointSim3<float> a;
    a.P = *P;
    a.eval(R->r32, ((OFrame*)f)->f32, T, lvl, energy, gxi, g.data());
  }
  if (grad_d)
    for (size_t i = 0; i < n; ++i) grad_d[i] = g[i];
  return 0;
}

// Sim3 energy E = 1/2 sum_i wh_i (w_p,i rp_i^2 + w_d,i rd_i^2) at exp(xi7) * T
// and point inverse depths d_i + delta_i, with the per-point weights
// wp = wh w_p and wd = wh w_d, the nearest-pixel frame depth sample and the
// in-image set frozen at T (FP64 only; for finite differences).  frozen holds
// 3 doubles per point (wp, wd, frame idepth sample) and mask one byte; with
// compute_frozen they are computed at T (xi and delta must then be zero).
extern "C" int orc_energy_frozen_sim3(void* r, void* f, const lsdg_params* P, int lvl, const double* pose8,
                                      const double* xi7, const double* delta_d, double* frozen,
                                      uint8_t* mask_frozen, int compute_frozen, double* energy) {
  ORef* R = (ORef*)r;
  if (!R->dbl) return LSDG_EINVAL;
  TrackingReference<double>& ref = R->r64;
  Frame<double>& frame = ((OFrame*)f)->f64;
  ref.makePointCloud(lvl);
  Sim3 T0 = Sim3::from8(pose8);
  Sim3 T = sim3_exp(xi7) * T0;
  double Rm[9], t[3], R0[9], t0[3];
  pose_RT(T, Rm, t, true);
  pose_RT(T0, R0, t0, true);
  const int W = frame.w[lvl], H = frame.h[lvl];
  const double fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
  const double hd = (double)P->huber_d;
  const double* grad = frame.grad[lvl].data();
  double E = 0;
  const int N = ref.num[lvl];
  for (int i = 0; i < N; ++i) {
    const double* p0 = &ref.pos[lvl][3 * i];
    const double* cv = &ref.colVar[lvl][2 * i];
    double* fz = frozen + 3 * (size_t)i;
    if (compute_frozen) {
      double Wx = R0[0] * p0[0] + R0[1] * p0[1] + R0[2] * p0[2] + t0[0];
      double Wy = R0[3] * p0[0] + R0[4] * p0[1] + R0[5] * p0[2] + t0[1];
      double Wz = R0[6] * p0[0] + R0[7] * p0[1] + R0[8] * p0[2] + t0[2];
      double u = (Wx / Wz) * fx + cx, v = (Wy / Wz) * fy + cy;
      bool in = (u > 1 && v > 1 && u < W - 2 && v < H - 2);
      mask_frozen[i] = in;
      fz[0] = fz[1] = fz[2] = 0;
      if (in) {
        double ri[3];
        interp4(grad, u, v, W, 3, ri);
        double gx = fx * ri[0], gy = fy * ri[1], d = 1.0 / p0[2];
        double rp = cv[0] - ri[2];
        int ir = (int)(u + 0.5) + W * (int)(v + 0.5);
        double varF = frame.ivar[lvl][ir];
        double rd = varF > 0 ? 1.0 / Wz - frame.idepth[lvl][ir] : -1.0;
        double vard = varF > 0 ? varF : -1.0;
        double s = (double)P->var_weight * cv[1], sv = (double)P->var_weight * vard;
        double g0 = (t0[0] * Wz - t0[2] * Wx) / (Wz * Wz * d);
        double g1 = (t0[1] * Wz - t0[2] * Wy) / (Wz * Wz * d);
        double g2 = (Wz - t0[2]) / (Wz * Wz * d);
        double drpdd = gx * g0 + gy * g1;
        double w_p = 1.0 / ((double)P->camera_pixel_noise2 + s * drpdd * drpdd);
        double w_d = 1.0 / (sv + g2 * g2 * s);
        double wabs = std::fabs(rp * std::sqrt(w_p)) + (sv > 0 ? std::fabs(rd * std::sqrt(w_d)) : 0.0);
        double wh = wabs < hd ? 1.0 : hd / wabs;
        fz[0] = wh * w_p;
        fz[1] = sv > 0 ? wh * w_d : 0.0;
        fz[2] = frame.idepth[lvl][ir];
      }
    }
    if (!mask_frozen[i]) continue;
    double d = 1.0 / p0[2] + (delta_d ? delta_d[i] : 0.0);
    double sc = (1.0 / d) / p0[2];  // moves the point along its ray to the new inverse depth
    double p[3] = {p0[0] * sc, p0[1] * sc, p0[2] * sc};
    double Wx = Rm[0] * p[0] + Rm[1] * p[1] + Rm[2] * p[2] + t[0];
    double Wy = Rm[3] * p[0] + Rm[4] * p[1] + Rm[5] * p[2] + t[1];
    double Wz = Rm[6] * p[0] + Rm[7] * p[1] + Rm[8] * p[2] + t[2];
    double u = (Wx / Wz) * fx + cx, v = (Wy / Wz) * fy + cy;
    if (!(u > 1 && v > 1 && u < W - 2 && v < H - 2)) return LSDG_ESTATE;  // left the image: step too large
    double ri[3];
    interp4(grad, u, v, W, 3, ri);
    double rp = cv[0] - ri[2];
    double rd = 1.0 / Wz - fz[2];
    E += 0.5 * (fz[0] * rp * rp + fz[1] * rd * rd);
  }
  *energy = E;
  return 0;
}

// Energy with frozen per-point weights w_p and frozen in-image set, at a
// perturbed pose exp(xi) * T and perturbed point inverse depths d_i + delta_i
// (FP64 only; for finite differences).  w_frozen/mask_frozen come from the
// linearisation point (pass NULL to compute them here and return them).
extern "C" int orc_energy_frozen(void* r, void* f, const lsdg_params* P, int lvl, const double* pose, const double* xi,
                      const double* delta_d, double* w_frozen, uint8_t* mask_frozen, int compute_frozen,
                      double* energy) {
  ORef* R = (ORef*)r;
  if (!R->dbl) return LSDG_EINVAL;
  TrackingReference<double>& ref = R->r64;
  Frame<double>& frame = ((OFrame*)f)->f64;
  ref.makePointCloud(lvl);
  Sim3 T0 = Sim3::from7(pose);
  Sim3 T = se3_exp(xi) * T0;
  double Rm[9], t[3];
  pose_RT(T, Rm, t, false);
  double R0[9], t0[3];
  pose_RT(T0, R0, t0, false);
  int W = frame.w[lvl], H = frame.h[lvl];
  double fx = frame.fx[lvl], fy = frame.fy[lvl], cx = frame.cx[lvl], cy = frame.cy[lvl];
  double k = (double)P->huber_d / 2.0;
  double E = 0;
  int N = ref.num[lvl];
  for (int i = 0; i < N; ++i) {
    const double* p0 = &ref.pos[lvl][3 * i];
    const double* cv = &ref.colVar[lvl][2 * i];
    if (compute_frozen) {
      double Wx = R0[0] * p0[0] + R0[1] * p0[1] + R0[2] * p0[2] + t0[0];
      double Wy = R0[3] * p0[0] + R0[4] * p0[1] + R0[5] * p0[2] + t0[1];
      double Wz = R0[6] * p0[0] + R0[7] * p0[1] + R0[8] * p0[2] + t0[2];
      double u = (Wx / Wz) * fx + cx, v = (Wy / Wz) * fy + cy;
      bool in = (u > 1 && v > 1 && u < W - 2 && v < H - 2);
      mask_frozen[i] = in;
      w_frozen[i] = 0;
      if (in) {
        double ri[3];
        interp4(frame.grad[lvl].data(), u, v, W, 3, ri);
        double gx = fx * ri[0], gy = fy * ri[1], d = 1.0 / p0[2];
        double g0 = (t0[0] * Wz - t0[2] * Wx) / (Wz * Wz * d);
        double g1 = (t0[1] * Wz - t0[2] * Wy) / (Wz * Wz * d);
        double drpdd = gx * g0 + gy * g1;
        w_frozen[i] = 1.0 / ((double)P->camera_pixel_noise2 + (double)P->var_weight * cv[1] * drpdd * drpdd);
      }
    }
    if (!mask_frozen[i]) continue;
    double d = 1.0 / p0[2] + (delta_d ? delta_d[i] : 0.0);
    double s = (1.0 / d) / p0[2];  // rescale of the ray point to the new inverse depth
    double p[3] = {p0[0] * s, p0[1] * s, p0[2] * s};
    double Wx = Rm[0] * p[0] + Rm[1] * p[1] + Rm[2] * p[2] + t[0];
    double Wy = Rm[3] * p[0] + Rm[4] * p[1] + Rm[5] * p[2] + t[1];
    double Wz = Rm[6] * p[0] + Rm[7] * p[1] + Rm[8] * p[2] + t[2];
    double u = (Wx / Wz) * fx + cx, v = (Wy / Wz) * fy + cy;
    if (!(u > 1 && v > 1 && u < W - 2 && v < H - 2)) return LSDG_ESTATE;  // left the image: step too large
    double ri[3];
    interp4(frame.grad[lvl].data(), u, v, W, 3, ri);
    double rr = cv[0] - ri[2];
    double e = std::fabs(rr * std::sqrt(w_frozen[i]));
    E += e < k ? 0.5 * e * e : k * e - 0.5 * k * k;
  }
  *energy = E;
  return 0;
}

extern "C" void* orc_depth_new(int prec, int w, int h, float fx, float fy, float cx, float cy, const lsdg_params* P,
                    int threads) {
  if (w < 2 || h < 2 || (w & 1) || (h & 1)) return nullptr;
  ODepth* d = new ODepth();
  d->prec = prec;
  d->dbl = prec == 1;
  if (d->dbl) {
    d->d64.P = *P; d->d64.threads = threads; d->d64.init(w, h, fx, fy, cx, cy);
  } else {
    d->d32.P = *P; d->d32.threads = threads; d->d32.accDouble = (prec == 2); d->d32.init(w, h, fx, fy, cx, cy);
  }
  return d;
}
extern "C" void orc_depth_free(void* d) { delete (ODepth*)d; }
extern "C" void orc_depth_set_threads(void* d, int threads) {
  ODepth* D = (ODepth*)d;
  D->d32.threads = threads;
  D->d64.threads = threads;
}
extern "C" void orc_depth_set_prior(void* d, void* kf, const float* id, const float* var, int validity) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) D->d64.setPrior(&((OFrame*)kf)->f64, id, var, validity);
  else D->d32.setPrior(&((OFrame*)kf)->f32, id, var, validity);
}
extern "C" void orc_depth_init_gt(void* d, void* kf) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) D->d64.initializeFromGT(&((OFrame*)kf)->f64);
  else D->d32.initializeFromGT(&((OFrame*)kf)->f32);
}
extern "C" void orc_depth_init_random(void* d, void* kf, uint64_t seed) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) D->d64.initializeRandomly(&((OFrame*)kf)->f64, seed);
  else D->d32.initializeRandomly(&((OFrame*)kf)->f32, seed);
}
extern "C" uint64_t orc_splitmix64(uint64_t seed, int i) { return splitmix64_output(seed, i); }
extern "C" void orc_depth_update(void* d, void** refs, int n) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) {
    std::vector<Frame<double>*> v;
    for (int i = 0; i < n; ++i) v.push_back(&((OFrame*)refs[i])->f64);
    D->d64.updateKeyframe(v);
  } else {
    std::vector<Frame<float>*> v;
    for (int i = 0; i < n; ++i) v.push_back(&((OFrame*)refs[i])->f32);
    D->d32.updateKeyframe(v);
  }
}
extern "C" void orc_depth_create_keyframe(void* d, void* nkf) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) D->d64.createKeyFrame(&((OFrame*)nkf)->f64);
  else D->d32.createKeyFrame(&((OFrame*)nkf)->f32);
}
extern "C" void orc_depth_finalize(void* d) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) D->d64.finalizeKeyFrame();
  else D->d32.finalizeKeyFrame();
}
// Sub-steps, for per-kernel parity.
extern "C" void orc_depth_step(void* d, int step, int arg) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) {
    if (step == 0) D->d64.observeDepth();
    if (step == 1) D->d64.fillHoles();
    if (step == 2) D->d64.regularize(arg != 0, D->d64.P.val_sum_min_for_keep);
  } else {
    if (step == 0) D->d32.observeDepth();
    if (step == 1) D->d32.fillHoles();
    if (step == 2) D->d32.regularize(arg != 0, D->d32.P.val_sum_min_for_keep);
  }
}
template <class S>
static void depth_read_t(DepthMap<S>& D, lsdg_hypotheses* o) {
  for (size_t i = 0; i < D.cur.size(); ++i) {
    const Hyp<S>& h = D.cur[i];
    if (o->is_valid) o->is_valid[i] = h.isValid;
    if (o->blacklisted) o->blacklisted[i] = h.blacklisted;
    if (o->next_stereo_frame_min_id) o->next_stereo_frame_min_id[i] = h.nextStereoFrameMinID;
    if (o->validity_counter) o->validity_counter[i] = h.validity_counter;
    if (o->idepth) o->idepth[i] = (float)h.idepth;
    if (o->idepth_var) o->idepth_var[i] = (float)h.idepth_var;
    if (o->idepth_smoothed) o->idepth_smoothed[i] = (float)h.idepth_smoothed;
    if (o->idepth_var_smoothed) o->idepth_var_smoothed[i] = (float)h.idepth_var_smoothed;
    if (o->observe_code) o->observe_code[i] = D.obsCode[i];
  }
}
extern "C" void orc_depth_read(void* d, lsdg_hypotheses* o) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) depth_read_t(D->d64, o);
  else depth_read_t(D->d32, o);
}
// Full-precision read of the FP64 oracle's idepth / var (double arrays).
extern "C" void orc_depth_read_f64(void* d, double* idepth, double* var, double* ids, double* vars) {
  ODepth* D = (ODepth*)d;
  for (size_t i = 0; i < D->d64.cur.size(); ++i) {
    const Hyp<double>& h = D->d64.cur[i];
    if (idepth) idepth[i] = h.idepth;
    if (var) var[i] = h.idepth_var;
    if (ids) ids[i] = h.idepth_smoothed;
    if (vars) vars[i] = h.idepth_var_smoothed;
  }
}
template <class S>
static void depth_write_t(DepthMap<S>& D, const lsdg_hypotheses* o) {
  for (size_t i = 0; i < D.cur.size(); ++i) {
    Hyp<S>& h = D.cur[i];
    if (o->is_valid) h.isValid = o->is_valid[i] != 0;
    if (o->blacklisted) h.blacklisted = o->blacklisted[i];
    if (o->next_stereo_frame_min_id) h.nextStereoFrameMinID = o->next_stereo_frame_min_id[i];
    if (o->validity_counter) h.validity_counter = o->validity_counter[i];
    if (o->idepth) h.idepth = (S)o->idepth[i];
    if (o->idepth_var) h.idepth_var = (S)o->idepth_var[i];
    if (o->idepth_smoothed) h.idepth_smoothed = (S)o->idepth_smoothed[i];
    if (o->idepth_var_smoothed) h.idepth_var_smoothed = (S)o->idepth_var_smoothed[i];
  }
}
extern "C" void orc_depth_write(void* d, const lsdg_hypotheses* o) {
  ODepth* D = (ODepth*)d;
  if (D->dbl) depth_write_t(D->d64, o);
  else depth_write_t(D->d32, o);
}

// Exposed closed forms for the oracle's own pinning tests.
extern "C" void orc_se3_exp(const double* xi, double* out7) { se3_exp(xi).to7(out7); }
extern "C" void orc_sim3_exp(const double* xi, double* out8) { sim3_exp(xi).to8(out8); }
extern "C" void orc_sim3_mul(const double* a8, const double* b8, double* out8) { (Sim3::from8(a8) * Sim3::from8(b8)).to8(out8); }
extern "C" void orc_sim3_inv(const double* a8, double* out8) { Sim3::from8(a8).inverse().to8(out8); }
extern "C" void orc_ldlt6(const double* A, const double* b, double* x) { ldlt_solve<6>(A, b, x); }
extern "C" void orc_ldlt7(const double* A, const double* b, double* x) { ldlt_solve<7>(A, b, x); }
extern "C" float orc_interp_f32(const float* img, float x, float y, int w) { return interp(img, x, y, w); }

// PangolinOutputIOWrapper::publishKeyframe's records of a frame's level.
template <class S>
static int keyframe_points_t(OFrame* F, int l, lsdg_point_dense* out) {
  Frame<S>& f = fr<S>(F);
  if (!f.hasDepth) return LSDG_ESTATE;
  const size_t n = (size_t)f.w[l] * f.h[l];
  std::vector<float> id(n), var(n), img(n);
  for (size_t i = 0; i < n; ++i) {
    id[i] = (float)f.idepth[l][i];
    var[i] = (float)f.ivar[l][i];
    img[i] = (float)f.image[l][i];
  }
  publish_points(id.data(), var.data(), img.data(), (int)n, out);
  return 0;
}
extern "C" int orc_keyframe_points(void* f, int level, lsdg_point_dense* out) {
  OFrame* F = (OFrame*)f;
  if (level < 0 || level >= PYR) return LSDG_EINVAL;
  return F->dbl ? keyframe_points_t<double>(F, level, out) : keyframe_points_t<float>(F, level, out);
}
extern "C" int orc_compute_vbo(const lsdg_point_dense* in, int w, int h, float fx, float fy, float cx, float cy,
                               const lsdg_vbo_params* p, lsdg_vertex* out) {
  return compute_vbo(in, w, h, fx, fy, cx, cy, *p, out);
}
extern "C" void orc_frame_camera(void* f, int level, float* k4) {
  OFrame* F = (OFrame*)f;
  if (F->dbl) {
    k4[0] = (float)F->f64.fx[level]; k4[1] = (float)F->f64.fy[level];
    k4[2] = (float)F->f64.cx[level]; k4[3] = (float)F->f64.cy[level];
  } else {
    k4[0] = F->f32.fx[level]; k4[1] = F->f32.fy[level]; k4[2] = F->f32.cx[level]; k4[3] = F->f32.cy[level];
  }
}

// Ingest (Undistorter::undistort), the FOV model's table, the pose.txt line.
extern "C" void orc_undistort(const uint8_t* src, int inW, const float* map, int outW, int outH, uint8_t* out) {
  undistort(src, inW, map, outW, outH, out);
}
extern "C" void orc_fov_table(const float* ik, float omega, int inW, int inH, const float* ok, int outW, int outH,
                              float* map) {
  fov_table(ik, omega, inW, inH, ok, outW, outH, map);
}
extern "C" int orc_pose_txt_line(int id, const double* c2wT, const double* t2pT, char* buf, int cap) {
  const std::string s = pose_txt_line(id, c2wT, t2pT);
  if ((int)s.size() >= cap) return -1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

// testConstraint for one pair (tracking references of the new keyframe A and
// a candidate B, keyframe ids idA / idB), sequential as upstream.
static bool usable(const double* pose8, const lsdg_track_stats& st) {
  return !st.diverged && pose8[7] <= 1.0 / 1e-10 && pose8[7] >= 1e-10 && st.hessian[0] != 0.0 && st.hessian[48] != 0.0;
}
static void edge_of(const double* pose8, const lsdg_track_stats& st, int first, int second, lsdg_constraint_edge& e) {
  e.first_id = first;
  e.second_id = second;
  std::memcpy(e.second_to_first, pose8, sizeof(e.second_to_first));
  std::memcpy(e.information, st.hessian, sizeof(e.information));
  e.mean_residual = st.last_residual;
  e.mean_residual_d = st.last_depth_residual;
  e.mean_residual_p = st.last_photometric_residual;
  e.usage = st.point_usage;
}
extern "C" void orc_constraint_test(void* refA, void* refB, int idA, int idB, const double* init8,
                                    const lsdg_params* P, float strictness, lsdg_constraint_result* out) {
  ORef* A = (ORef*)refA;
  ORef* B = (ORef*)refB;
  std::memset(out, 0, sizeof(*out));
  static const int lv[3][2] = {{4, 3}, {2, 2}, {1, 1}};
  static const float th[3] = {3000.0f, 4000.0f, 6000.0f};
  double ctof[8], ftoc[8];
  std::memcpy(ctof, init8, sizeof(ctof));
  Sim3::from8(init8).inverse().to8(ftoc);
  lsdg_track_stats sb, sa;
  for (int s = 0; s < 3; ++s) {
    double o[8];
    orc_sim3_track(A, B->kf, P, ctof, lv[s][0], lv[s][1], o, &sb);
    std::memcpy(ctof, o, sizeof(o));
    float err;
    if (!usable(ctof, sb)) {
      err = 1e20f;
    } else {
      orc_sim3_track(B, A->kf, P, ftoc, lv[s][0], lv[s][1], o, &sa);
      std::memcpy(ftoc, o, sizeof(o));
      err = usable(ftoc, sa) ? (float)reciprocal_consistency(Sim3::from8(ftoc), sa.hessian, Sim3::from8(ctof),
                                                              sb.hessian)
                             : 1e20f;
    }
    out->error[s] = err;
    if (err > th[s] * strictness) {
      out->rejected_at = s + 1;
      break;
    }
    if (s == 2) {
      out->accepted = 1;
      edge_of(ctof, sb, idA, idB, out->candidate_on_frame);
      edge_of(ftoc, sa, idB, idA, out->frame_on_candidate);
    }
  }
  std::memcpy(out->candidate_to_frame, ctof, sizeof(ctof));
  std::memcpy(out->frame_to_candidate, ftoc, sizeof(ftoc));
}
// The Sim3 log, adjoint and reciprocal consistency, for their pinning tests.
extern "C" void orc_sim3_log(const double* p8, double* xi) { sim3_log(Sim3::from8(p8), xi); }
extern "C" void orc_sim3_adj(const double* p8, double* M) { sim3_adj(Sim3::from8(p8), M); }
extern "C" double orc_reciprocal_consistency(const double* atob8, const double* atobInfo, const double* btoa8,
                                             const double* btoaInfo) {
  return reciprocal_consistency(Sim3::from8(atob8), atobInfo, Sim3::from8(btoa8), btoaInfo);
}
