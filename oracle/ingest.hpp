// oracle/ingest.hpp — TEST INFRASTRUCTURE (the oracle).  Not product code:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and only as the checker or the timed CPU baseline.
//
// CPU restatement of the step before the hot path and of the trajectory
// writer after it:
//   undistort      libvideoio's Undistorter::undistort, which InputThread
//                  applies to every camera image (lib/App/InputThread.cpp:59-62);
//                  libvideoio is an external dependency (tools/LSD.cpp:33,
//                  fips.yml) absent from the tree, so its published algorithm
//                  (inherited from lsd_slam's Undistorter) is restated: per
//                  output pixel a source position from a table; no source
//                  (x < 0) gives 0, otherwise the 8-bit input is bilinearly
//                  interpolated and truncated back to 8 bits.
//   fov_table      the table of the FOV (ATAN) lens model of LSD-SLAM's
//                  calibration files (Devernay and Faugeras 2001): distorted
//                  radius = atan(2 r tan(omega / 2)) / omega.
//   pose_txt_line  TextOutputIOWrapper::publishTrackedFrame
//                  (lib/Pangolin_IOWrapper/TextOutputIOWrapper.cpp:100-120),
//                  written through an ostream as the reference writes its
//                  ofstream.
#pragma once
#include <cmath>
#include <cstdint>
#include <sstream>
#include <string>

namespace orc {

// Undistorter::undistort for an out_w x out_h table `map` (x, y per pixel).
inline void undistort(const uint8_t* src, int inW, const float* map, int outW, int outH, uint8_t* out) {
  for (int idx = 0; idx < outW * outH; ++idx) {
    float xx = map[2 * idx], yy = map[2 * idx + 1];
    if (xx < 0) {
      out[idx] = 0;
      continue;
    }
    const int xxi = (int)xx, yyi = (int)yy;
    xx -= (float)xxi;
    yy -= (float)yyi;
    const float xxyy = xx * yy;
    const uint8_t* s = src + xxi + yyi * inW;
    const float v = xxyy * (float)s[1 + inW] + (yy - xxyy) * (float)s[inW] + (xx - xxyy) * (float)s[1] +
                    (1.0f - xx - yy + xxyy) * (float)s[0];
    out[idx] = (uint8_t)v;
  }
}

// FOV model: output pinhole camera ok = (fx, fy, cx, cy), input camera ik;
// sources outside 0 < x < in_w - 1, 0 < y < in_h - 1 become (-1, -1).
inline void fov_table(const float* ik, double omega, int inW, int inH, const float* ok, int outW, int outH,
                      float* map) {
  const double d2t = 2.0 * std::tan(omega / 2.0);
  for (int y = 0; y < outH; ++y)
    for (int x = 0; x < outW; ++x) {
      double ix = (x - (double)ok[2]) / ok[0];
      double iy = (y - (double)ok[3]) / ok[1];
      const double r = std::sqrt(ix * ix + iy * iy);
      const double fac = (r == 0.0 || omega == 0.0) ? 1.0 : std::atan(r * d2t) / (omega * r);
      const float sx = (float)(ik[0] * fac * ix + ik[2]);
      const float sy = (float)(ik[1] * fac * iy + ik[3]);
      const bool in = sx > 0.0f && sy > 0.0f && sx < (float)(inW - 1) && sy < (float)(inH - 1);
      map[2 * (x + y * outW)] = in ? sx : -1.0f;
      map[2 * (x + y * outW) + 1] = in ? sy : -1.0f;
    }
}

// TextOutputIOWrapper.cpp:104-119: id, camToWorld translation,
// thisToParent_raw translation.
inline std::string pose_txt_line(int id, const double* camToWorldT, const double* thisToParentT) {
  std::ostringstream os;
  os << id;
  os << "," << camToWorldT[0] << "," << camToWorldT[1] << "," << camToWorldT[2];
  os << "," << thisToParentT[0] << "," << thisToParentT[1] << "," << thisToParentT[2];
  os << std::endl;
  return os.str();
}

}  // namespace orc
