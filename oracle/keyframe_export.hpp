// oracle/keyframe_export.hpp — TEST INFRASTRUCTURE (the oracle).  Not product
// code: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker or the timed CPU baseline.
//
// CPU restatement of the reference's keyframe point export, which lives in
// the reference tree itself:
//   publish_points   PangolinOutputIOWrapper::publishKeyframe
//                    (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:56-89)
//   compute_vbo      Keyframe::computeVbo (lib/Pangolin_IOWrapper/Keyframe.h:76-142)
// The reference files need GL, Sophus and lsd-slam headers that this machine
// lacks, so they are restated here, not compiled.
#pragma once
#include <cmath>
#include <cstdint>

#include "../include/lsdg.h"

namespace orc {

// PangolinOutputIOWrapper.cpp:85-88 assigns the float image value to an
// unsigned char (truncation); out-of-range values, undefined there, clamp.
inline uint8_t color_byte(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)(int)v;
}

// PangolinOutputIOWrapper.cpp:81-89.
inline void publish_points(const float* idepth, const float* idepthVar, const float* color, int n,
                           lsdg_point_dense* pc) {
  for (int idx = 0; idx < n; idx++) {
    pc[idx].idepth = idepth[idx];
    pc[idx].idepth_var = idepthVar[idx];
    const uint8_t c = color_byte(color[idx]);
    pc[idx].color[0] = c;
    pc[idx].color[1] = c;
    pc[idx].color[2] = c;
    pc[idx].color[3] = c;
  }
}

// Keyframe.h:76-142 (my_sparsifyFactor = 1): returns the number of vertices.
inline int compute_vbo(const lsdg_point_dense* originalInput, int width, int height, float fx, float fy, float cx,
                       float cy, const lsdg_vbo_params& p, lsdg_vertex* tmpBuffer) {
  const float my_scaledTH = p.scaled_th;
  const float my_absTH = p.abs_th;
  const float my_scale = p.scale;
  const int my_minNearSupport = p.min_near_support;
  const float fxi = 1 / fx;
  const float fyi = 1 / fy;
  const float cxi = -cx / fx;
  const float cyi = -cy / fy;
  int points = 0;
  for (int y = 1; y < height - 1; y++) {
    for (int x = 1; x < width - 1; x++) {
      const lsdg_point_dense& c = originalInput[x + y * width];
      if (c.idepth <= 0) continue;
      float depth = 1 / c.idepth;
      float depth4 = depth * depth;
      depth4 *= depth4;
      if (c.idepth_var * depth4 > my_scaledTH) continue;
      if (c.idepth_var * depth4 * my_scale * my_scale > my_absTH) continue;
      if (my_minNearSupport > 1) {
        int nearSupport = 0;
        for (int dx = -1; dx < 2; dx++) {
          for (int dy = -1; dy < 2; dy++) {
            int idx = x + dx + (y + dy) * width;
            if (originalInput[idx].idepth > 0) {
              float diff = originalInput[idx].idepth - 1.0f / depth;
              if (diff * diff < 2 * c.idepth_var) nearSupport++;
            }
          }
        }
        if (nearSupport < my_minNearSupport) continue;
      }
      lsdg_vertex& v = tmpBuffer[points];
      v.point[0] = (x * fxi + cxi) * depth;
      v.point[1] = (y * fyi + cyi) * depth;
      v.point[2] = depth;
      v.color[3] = 100;
      v.color[2] = c.color[0];
      v.color[1] = c.color[1];
      v.color[0] = c.color[2];
      points++;
    }
  }
  return points;
}

}  // namespace orc
