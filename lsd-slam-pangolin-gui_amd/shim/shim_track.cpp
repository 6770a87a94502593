// Tracks one frame against one keyframe through the lsd_slam shim classes:
//
//   shim_track <dir> <w> <h> <fx> <fy> <cx> <cy>
//
// reads <dir>/kf.u8, <dir>/fr.u8 (w*h bytes), <dir>/idepth.f32 and
// <dir>/var.f32 (w*h floats), runs SE3Tracker::trackFrame from the identity,
// a DepthMap update of the keyframe with the tracked frame, and prints the
// pose (7 numbers), the tracker's result fields and the number of valid
// hypotheses.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <fstream>
#include <memory>
#include <vector>

#include "lsd_slam_shim.h"

template <class T>
static bool load(const std::string& path, std::vector<T>& v) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
  return (size_t)f.gcount() == v.size() * sizeof(T);
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: shim_track dir w h fx fy cx cy\n");
    return 2;
  }
  const std::string dir = argv[1];
  const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
  const float fx = std::atof(argv[4]), fy = std::atof(argv[5]), cx = std::atof(argv[6]), cy = std::atof(argv[7]);
  const size_t n = (size_t)w * h;
  std::vector<unsigned char> kf(n), fr(n);
  std::vector<float> idepth(n), var(n);
  if (!load(dir + "/kf.u8", kf) || !load(dir + "/fr.u8", fr) || !load(dir + "/idepth.f32", idepth) ||
      !load(dir + "/var.f32", var)) {
    std::fprintf(stderr, "missing inputs in %s\n", dir.c_str());
    return 2;
  }
  lsd_slam::Context ctx(0);
  if (!ctx.ok()) {
    std::fprintf(stderr, "no device: %s\n", ctx.lastError().c_str());
    return 1;
  }
  lsd_slam::Frame keyframe(ctx, 0, w, h, fx, fy, cx, cy, 0.0, kf.data());
  auto frame = std::make_shared<lsd_slam::Frame>(ctx, 1, w, h, fx, fy, cx, cy, 0.033, fr.data());
  if (!keyframe.valid() || !frame->valid() || !keyframe.setDepth(idepth.data(), var.data())) {
    std::fprintf(stderr, "frame setup failed: %s\n", ctx.lastError().c_str());
    return 1;
  }
  lsd_slam::TrackingReference ref(ctx);
  ref.importFrame(&keyframe);
  lsd_slam::SE3Tracker tracker(ctx);
  lsd_slam::SE3 pose = tracker.trackFrame(&ref, frame, lsd_slam::SE3());
  std::printf("pose");
  for (double v : pose.v) std::printf(" %.17g", v);
  std::printf("\nresidual %.9g usage %.9g good %d bad %d diverged %d good_tracking %d\n", tracker.lastResidual,
              tracker.pointUsage, tracker.lastGoodCount, tracker.lastBadCount, (int)tracker.diverged,
              (int)tracker.trackingWasGood);
  lsd_slam::DepthMap map(ctx, w, h, fx, fy, cx, cy);
  map.initializeFromGTDepth(&keyframe);
  std::deque<std::shared_ptr<lsd_slam::Frame>> refs{frame};
  map.updateKeyframe(refs);
  std::vector<uint8_t> valid(n);
  lsdg_hypotheses hy{};
  hy.is_valid = valid.data();
  if (lsdg_depth_read(map.handle(), &hy) != LSDG_OK) return 1;
  size_t nv = 0;
  for (uint8_t v : valid) nv += v;
  std::printf("depth_valid %zu valid_map %d\n", nv, (int)map.isValid());
  return 0;
}
