// Profiling driver: the batched-tracking workload of bench.py (64 SE3
// alignments of 640x480 frames, pyramids rebuilt per run, LM levels 4..1, at
// most 20 iterations per level, seeds 1001 + 17 i) through the C-ABI alone,
// with no Python in the process, for counter passes of rocprofv3.  Optionally
// first runs a plain device copy of a known byte count, the reference kernel
// for scaling the HBM byte counters.
//
//   lsdg_prof [runs] [copy_mib] [precision: 0 FP32 (default), 1 FP16, 2 BF16]
//             [sums: 0 FP64 per thread (default), 1 Kahan-compensated FP32]
//             [tune: 1 = run the plan tuner first and use its shape, as bench.py does]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/lsdg.h"
#include "../../include/lsdg_tune.h"

extern "C" {
size_t synth_scene_size(void);
void synth_scene_init(void* S, uint64_t seed, double contrast, double freq_scale);
void synth_render(const void* S, const double* cam_to_world, int w, int h, double fx, double fy, double cx, double cy,
                  double noise_sigma, uint64_t noise_seed, float* intensity_out, uint8_t* img_out, float* depth_out);
void synth_sample_pose(uint64_t seed, double trans_m, double rot_rad, int uniform_magnitude, double* out7);
void synth_uniform(uint64_t seed, size_t n, double lo, double hi, float* out);
}

namespace {

constexpr double kContrast = 0.53;  // synth.CONTRAST
constexpr double kPi = 3.14159265358979323846;

__global__ void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

#define CHECK(x)                                                                 \
  do {                                                                           \
    int s_ = (int)(x);                                                           \
    if (s_ < 0) {                                                                \
      std::fprintf(stderr, "%s failed (%d) at line %d\n", #x, s_, __LINE__);    \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

struct Instance {
  std::vector<uint8_t> kf, fr;
  std::vector<float> idepth, var;
};

// synth.tracking_instance(seed) at 640x480: 2 cm / 1 deg, noise sigma 2,
// keyframe depth on pixels with |grad I| > 5, variance ~ U[1e-3, 1e-2].
Instance make_instance(uint64_t seed, const lsdg_camera& cam) {
  const int w = cam.w, h = cam.h, n = w * h;
  std::vector<char> scene(synth_scene_size());
  synth_scene_init(scene.data(), seed, kContrast, 1.0);
  Instance in;
  in.kf.resize(n);
  in.fr.resize(n);
  std::vector<float> dep(n), dep2(n);
  const double I7[7] = {0, 0, 0, 1, 0, 0, 0};
  synth_render(scene.data(), I7, w, h, cam.fx, cam.fy, cam.cx, cam.cy, 0.0, 0, nullptr, in.kf.data(), dep.data());
  double pose[7];
  synth_sample_pose(seed * 7919 + 1, 0.02, 1.0 * kPi / 180.0, 0, pose);
  synth_render(scene.data(), pose, w, h, cam.fx, cam.fy, cam.cx, cam.cy, 2.0, seed * 7919 + 2, nullptr, in.fr.data(),
               dep2.data());
  std::vector<float> u(n);
  synth_uniform(seed * 7919 + 3, n, 1e-3, 1e-2, u.data());
  in.idepth.assign(n, -1.0f);
  in.var.assign(n, -1.0f);
  for (int y = 1; y < h - 1; ++y)
    for (int x = 1; x < w - 1; ++x) {
      const int i = x + y * w;
      float gx = 0.5f * ((float)in.kf[i + 1] - (float)in.kf[i - 1]);
      float gy = 0.5f * ((float)in.kf[i + w] - (float)in.kf[i - w]);
      if (std::sqrt(gx * gx + gy * gy) > 5.0f) {
        in.idepth[i] = 1.0f / dep[i];
        in.var[i] = u[i];
      }
    }
  return in;
}

}  // namespace

int main(int argc, char** argv) {
  const int runs = argc > 1 ? std::atoi(argv[1]) : 5;
  const int copy_mib = argc > 2 ? std::atoi(argv[2]) : 0;
  const int prec = argc > 3 ? std::atoi(argv[3]) : LSDG_FP32;
  const int sums = argc > 4 ? std::atoi(argv[4]) : LSDG_SUMS_FP64;
  const int tune = argc > 5 ? std::atoi(argv[5]) : 0;
  if (prec < LSDG_FP32 || prec > LSDG_BF16) {
    std::fprintf(stderr, "precision must be 0, 1 or 2\n");
    return 2;
  }
  static const char* kPrecName[] = {"FP32", "FP16", "BF16"};
  if (copy_mib > 0) {
    const size_t n4 = (size_t)copy_mib * (1u << 20) / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n4 * 16) != hipSuccess || hipMalloc(&b, n4 * 16) != hipSuccess) return 1;
    if (hipMemset(a, 0, n4 * 16) != hipSuccess) return 1;
    for (int k = 0; k < 3; ++k) copy_f4<<<4096, 256>>>(a, b, n4);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copy: %d MiB read + %d MiB written per launch, 3 launches\n", copy_mib, copy_mib);
    (void)hipFree(a);
    (void)hipFree(b);
  }
  lsdg_params P;
  lsdg_params_default(&P);
  for (int l = 0; l < LSDG_PYRAMID_LEVELS; ++l) P.max_its_per_lvl[l] = 20;
  lsdg_ctx ctx = nullptr;
  CHECK(lsdg_ctx_create(0, &P, &ctx));
  CHECK(lsdg_ctx_set_sums(ctx, sums));
  const lsdg_camera cam{640, 480, 525.0f, 525.0f, 319.5f, 239.5f};
  const int B = 64;
  std::vector<lsdg_frame> kfs(B), frs(B);
  std::vector<lsdg_ref> refs(B);
  std::vector<lsdg_job> jobs(B);
  for (int i = 0; i < B; ++i) {
    Instance in = make_instance(1001 + 17 * i, cam);
    CHECK(lsdg_frame_create_u8(ctx, 2 * i, &cam, in.kf.data(), cam.w, prec, &kfs[i]));
    CHECK(lsdg_frame_set_depth(kfs[i], in.idepth.data(), in.var.data()));
    CHECK(lsdg_frame_create_u8(ctx, 2 * i + 1, &cam, in.fr.data(), cam.w, prec, &frs[i]));
    CHECK(lsdg_ref_create(ctx, kfs[i], &refs[i]));
    lsdg_job& J = jobs[i];
    J.ref = refs[i];
    J.frame = frs[i];
    const double init[8] = {0, 0, 0, 1, 0, 0, 0, 1};
    for (int k = 0; k < 8; ++k) J.frame_to_ref_init[k] = init[k];
    J.kind = LSDG_JOB_SE3;
    J.start_level = 4;
    J.final_level = 1;
    J.rebuild_pyramid = 1;
  }
  if (tune) {  // as the bench does: the tuner's shape, recorded in the table, is the plan's
    lsdg_batch_config best;
    double t_default = 0, t_best = 0;
    CHECK(lsdg_batch_tune(ctx, jobs.data(), B, 5, &best, &t_default, &t_best));
    std::printf("tuned shape: block %d, K %d, solo %d, xcd groups %d (%.4f -> %.4f ms)\n", best.coop_block,
                best.coop_k, best.solo_max, best.xcd_groups, t_default, t_best);
  }
  lsdg_batch b = nullptr;
  CHECK(lsdg_batch_create(ctx, jobs.data(), B, &b));
  std::vector<lsdg_job_result> res(B);
  CHECK(lsdg_batch_run(b, res.data()));
  CHECK(lsdg_batch_run(b, res.data()));
  CHECK(lsdg_ctx_synchronize(ctx));
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < runs; ++r) CHECK(lsdg_batch_run(b, nullptr));
  CHECK(lsdg_ctx_synchronize(ctx));
  auto t1 = std::chrono::steady_clock::now();
  const double s = std::chrono::duration<double>(t1 - t0).count();
  int ok = 0;
  for (int i = 0; i < B; ++i) ok += res[i].status == LSDG_OK;
  std::printf("batched SE3 640x480 x %d, %s: %d runs, %.3f ms per run, %.1f frames/s, %d/%d jobs converged\n", B,
              kPrecName[prec], runs, 1e3 * s / runs, B * runs / s, ok, B);
  lsdg_batch_destroy(b);
  for (int i = 0; i < B; ++i) {
    lsdg_ref_destroy(refs[i]);
    lsdg_frame_destroy(frs[i]);
    lsdg_frame_destroy(kfs[i]);
  }
  lsdg_ctx_destroy(ctx);
  return 0;
}
