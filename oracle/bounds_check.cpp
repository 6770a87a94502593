// oracle/bounds_check.cpp — TEST INFRASTRUCTURE.  Runs the oracle's depth-map
// pipeline under AddressSanitizer (the Makefile's bounds_check target) at image
// sizes that are even but not multiples of the 16x16 tiles, with inverse depth
// on every pixel up to the last row and column, so any access past a per-pixel
// array or the level-1 tracking mask is reported:
//   initializeFromGTDepth -> updateKeyframe (one tracked reference frame) ->
//   createKeyFrame (propagation through the tracking mask) -> updateKeyframe
//   -> createKeyFrame (propagation through the photometric gate: the new
//   keyframe was not tracked on the current one).
// Also checks that odd sizes are refused.  Prints "ok WxH" per size; any
// sanitizer report exits non-zero.
//
//   bounds_check [w h]...   (default: 202x138 330x250 66x40)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../include/lsdg.h"

extern "C" {
void orc_params_default(lsdg_params* p);
void* orc_frame_new(int prec, int id, int w, int h, float fx, float fy, float cx, float cy, const float* img,
                    const lsdg_params* P);
void orc_frame_free(void* f);
void orc_frame_set_depth(void* f, const float* id, const float* var);
void orc_frame_set_tracking(void* f, void* parent, const double* tp, float itr, const uint8_t* wg);
void* orc_depth_new(int prec, int w, int h, float fx, float fy, float cx, float cy, const lsdg_params* P,
                    int threads);
void orc_depth_free(void* d);
void orc_depth_init_gt(void* d, void* kf);
void orc_depth_update(void* d, void** refs, int n);
void orc_depth_create_keyframe(void* d, void* nkf);
}

namespace {

// A smooth texture seen by a camera translated by (tx, ty): plenty of
// gradient everywhere, so every gate of the pipeline has pixels to act on.
std::vector<float> texture(int w, int h, double tx, double ty) {
  std::vector<float> img((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const double u = x + 40.0 * tx, v = y + 40.0 * ty;
      img[(size_t)x + (size_t)y * w] =
          (float)(128.0 + 60.0 * std::sin(0.31 * u + 0.7 * std::cos(0.11 * v)) + 40.0 * std::cos(0.23 * v - 0.05 * u));
    }
  return img;
}

int run(int w, int h) {
  lsdg_params P;
  orc_params_default(&P);
  P.max_epl_length_crop = 100.0f;
  const float f = 0.8f * (float)w, cx = 0.5f * (float)(w - 1), cy = 0.5f * (float)(h - 1);
  std::vector<float> img0 = texture(w, h, 0.0, 0.0), img1 = texture(w, h, 0.05, 0.01), img2 = texture(w, h, 0.1, 0.02),
                     img3 = texture(w, h, 0.12, 0.03);
  // inverse depth on every pixel, border rows and columns included
  std::vector<float> id((size_t)w * h), var((size_t)w * h, 1e-3f);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) id[(size_t)x + (size_t)y * w] = 0.5f + 0.001f * (float)x + 0.0005f * (float)y;
  for (int prec : {0, 1}) {
    void* kf = orc_frame_new(prec, 0, w, h, f, f, cx, cy, img0.data(), &P);
    void* fr1 = orc_frame_new(prec, 1, w, h, f, f, cx, cy, img1.data(), &P);
    void* nkf = orc_frame_new(prec, 2, w, h, f, f, cx, cy, img2.data(), &P);
    void* nkf2 = orc_frame_new(prec, 3, w, h, f, f, cx, cy, img3.data(), &P);
    void* dm = orc_depth_new(prec, w, h, f, f, cx, cy, &P, 1);
    if (!kf || !fr1 || !nkf || !nkf2 || !dm) {
      std::fprintf(stderr, "oracle refused %dx%d\n", w, h);
      return 1;
    }
    orc_frame_set_depth(kf, id.data(), var.data());
    orc_depth_init_gt(dm, kf);
    // tracked on the keyframe, mask all good: thisToParent = small translation
    const double tp1[8] = {0, 0, 0, 1, -0.001, 0.0, 0.0, 1.0};
    std::vector<uint8_t> good((size_t)(w / 2) * (h / 2), 1);
    orc_frame_set_tracking(fr1, kf, tp1, 0.5f, good.data());
    void* refs[1] = {fr1};
    orc_depth_update(dm, refs, 1);
    // new keyframe tracked on the keyframe: propagation through the mask
    // 3 cm right and down: sources on the last row and column land about 3 px
    // inside the image, past the destination gate, and read the mask
    const double tp2[8] = {0, 0, 0, 1, 0.03, 0.03, 0.0, 1.0};
    orc_frame_set_tracking(nkf, kf, tp2, 0.5f, good.data());
    orc_depth_create_keyframe(dm, nkf);
    refs[0] = fr1;
    orc_frame_set_tracking(fr1, nkf, tp1, 0.5f, good.data());
    orc_depth_update(dm, refs, 1);
    // not tracked on the current keyframe: the photometric gate instead
    const double tp3[8] = {0, 0, 0, 1, -0.001, 0.0, 0.0, 1.0};
    orc_frame_set_tracking(nkf2, fr1, tp3, 0.5f, nullptr);
    orc_depth_create_keyframe(dm, nkf2);
    orc_depth_free(dm);
    for (void* p : {kf, fr1, nkf, nkf2}) orc_frame_free(p);
  }
  // odd sizes are refused
  std::vector<float> odd((size_t)(w + 1) * (h + 1), 100.0f);
  if (orc_frame_new(0, 9, w + 1, h, f, f, cx, cy, odd.data(), &P) ||
      orc_frame_new(0, 9, w, h + 1, f, f, cx, cy, odd.data(), &P) ||
      orc_depth_new(0, w + 1, h + 1, f, f, cx, cy, &P, 1)) {
    std::fprintf(stderr, "odd size accepted next to %dx%d\n", w, h);
    return 1;
  }
  std::printf("ok %dx%d\n", w, h);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::pair<int, int>> sizes;
  for (int i = 1; i + 1 < argc; i += 2) sizes.push_back({std::atoi(argv[i]), std::atoi(argv[i + 1])});
  if (sizes.empty()) sizes = {{202, 138}, {330, 250}, {66, 40}};
  for (auto& s : sizes)
    if (run(s.first, s.second)) return 1;
  return 0;
}
