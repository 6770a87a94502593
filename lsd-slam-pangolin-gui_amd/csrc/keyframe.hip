// Keyframe export for display: the pointcloud records the reference's
// Pangolin output wrapper builds from a keyframe and the filter its viewer
// runs before uploading a vertex buffer.
//
//   k_kf_dense     PangolinOutputIOWrapper::publishKeyframe
//                  (lib/Pangolin_IOWrapper/PangolinOutputIOWrapper.cpp:81-89):
//                  idepth, idepth variance and the image as four color bytes.
//   k_vbo_count,   Keyframe::computeVbo (lib/Pangolin_IOWrapper/Keyframe.h:
//   k_vbo_scan,    91-142) as a stream compaction that keeps the viewer's
//   k_vbo_write    raster order: per-block counts of the kept pixels, one
//                  scan, then each block writes its kept pixels at its offset
//                  in lane order.  Every float expression keeps the
//                  reference's operation order (no contraction), so the
//                  vertices match a sequential loop bit for bit.
#include "lsdg_internal.h"

namespace lsdg {

// The reference's float-to-byte conversion truncates; values it leaves
// undefined (negative, >= 256, NaN) are clamped.
__device__ inline uint8_t color_byte(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)(int)v;
}

__global__ __launch_bounds__(256) void k_kf_dense(const float* __restrict__ id, const float* __restrict__ var,
                                                  const float* __restrict__ img, int n,
                                                  lsdg_point_dense* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  lsdg_point_dense p;
  p.idepth = id[i];
  p.idepth_var = var[i];
  const uint8_t c = color_byte(img[i]);
  p.color[0] = p.color[1] = p.color[2] = p.color[3] = c;
  out[i] = p;
}

__device__ inline bool vbo_keep(const lsdg_point_dense* __restrict__ in, int x, int y, const VboArgs& a,
                                float& depth) {
  const int w = a.w;
  if (x < 1 || x >= w - 1 || y < 1 || y >= a.h - 1) return false;
  const lsdg_point_dense c = in[x + y * w];
  if (c.idepth <= 0) return false;
  depth = 1 / c.idepth;
  float depth4 = depth * depth;
  depth4 *= depth4;
  if (c.idepth_var * depth4 > a.scaledTH) return false;
  if (c.idepth_var * depth4 * a.scale * a.scale > a.absTH) return false;
  if (a.minNear > 1) {
    int near = 0;
    for (int dx = -1; dx < 2; dx++)
      for (int dy = -1; dy < 2; dy++) {
        const float nid = in[x + dx + (y + dy) * w].idepth;
        if (nid > 0) {
          const float diff = nid - 1.0f / depth;
          if (diff * diff < 2 * c.idepth_var) near++;
        }
      }
    if (near < a.minNear) return false;
  }
  return true;
}

// Kept pixels of block b (pixels 256 b .. 256 b + 255 in raster order).
__global__ __launch_bounds__(256) void k_vbo_count(const lsdg_point_dense* __restrict__ in, VboArgs a,
                                                   int* __restrict__ blk) {
  __shared__ int wc[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float depth;
  const bool keep = i < a.w * a.h && vbo_keep(in, i % a.w, i / a.w, a, depth);
  const unsigned long long m = __ballot(keep);
  if (__lane_id() == 0) wc[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) blk[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Exclusive scan of blk[0, nb) in place, blk[nb] = total; one block.
__global__ __launch_bounds__(1024) void k_vbo_scan(int* blk, int nb) {
  __shared__ int sh[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? blk[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) blk[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[nb] = carry;
}

__global__ __launch_bounds__(256) void k_vbo_write(const lsdg_point_dense* __restrict__ in, VboArgs a,
                                                   const int* __restrict__ blk, lsdg_vertex* __restrict__ out) {
  __shared__ int wc[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int x = i % a.w, y = i / a.w;
  float depth = 0.0f;
  const bool keep = i < a.w * a.h && vbo_keep(in, x, y, a, depth);
  const unsigned long long m = __ballot(keep);
  const int wave = threadIdx.x >> 6;
  if (__lane_id() == 0) wc[wave] = __popcll(m);
  __syncthreads();
  if (!keep) return;
  int r = blk[blockIdx.x] +
          (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
  for (int v = 0; v < wave; ++v) r += wc[v];
  const lsdg_point_dense c = in[i];
  lsdg_vertex o;
  o.point[0] = ((float)x * a.fxi + a.cxi) * depth;
  o.point[1] = ((float)y * a.fyi + a.cyi) * depth;
  o.point[2] = depth;
  o.color[3] = 100;
  o.color[2] = c.color[0];
  o.color[1] = c.color[1];
  o.color[0] = c.color[2];
  out[r] = o;
}

// ------------------------------------------------------------ launchers
hipError_t launch_kf_dense(const float* id, const float* var, const float* img, int n, lsdg_point_dense* out,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_kf_dense<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(id, var, img, n, out);
  return hipGetLastError();
}

size_t vbo_block_count(int w, int h) { return ((size_t)w * h + 255) / 256; }

hipError_t launch_vbo(const lsdg_point_dense* in, const VboArgs& a, int* blk, lsdg_vertex* out, hipStream_t s) {
  const int nb = (int)vbo_block_count(a.w, a.h);
  if (nb == 0) return hipSuccess;
  k_vbo_count<<<dim3(nb), dim3(256), 0, s>>>(in, a, blk);
  k_vbo_scan<<<dim3(1), dim3(1024), 0, s>>>(blk, nb);
  k_vbo_write<<<dim3(nb), dim3(256), 0, s>>>(in, a, blk, out);
  return hipGetLastError();
}

}  // namespace lsdg
