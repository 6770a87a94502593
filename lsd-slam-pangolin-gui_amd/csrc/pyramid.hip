// Frame pyramids and TrackingReference point clouds on gfx950.
//
//  k_pyr_images  Frame::buildImage for all levels in one launch: a block owns a
//                32x32 level-0 tile and reduces it through levels 1-4 in LDS
//                (2x2 box, ((TL + TR) + BL) + BR, times 0.25).
//  k_pyr_grads   Frame::buildGradients + buildMaxGradients for every level in
//                one launch: one thread per pixel writes (gx, gy, I, 0), the
//                optional FP16/BF16 packed copy and the 3x3 max of |grad I|
//                (recomputed from the image; identical to the reference's
//                two-pass vertical-then-horizontal max on its linear index).
//  k_pyr_depth   Frame::buildIDepthAndIDepthVar for levels 1-4 in one launch.
//  k_ingest      undistortion of an 8-bit camera image, its conversion to
//                float and buildImage's levels 1-4, in one launch.
//  k_pc_*        TrackingReference::makePointCloud as a stream compaction
//                (count, scan, write) in row-major pixel order.
#include "lsdg_internal.h"

namespace lsdg {

struct PyrImgArgs {
  float* img[PYR];
  int w[PYR], h[PYR];
  const uint8_t* u8;  // the 8-bit level 0 when it holds the image (read instead of the float copy), or null
};

// Levels 2-4 from a block's 16x16 level-1 tile L1 (level 1 already stored).
__device__ inline void pyr_levels_2to4(const PyrImgArgs& p, const float (*L1)[17], int bx, int by) {
  __shared__ float L2[8][9];
  __shared__ float L3[4][5];
  const int t = threadIdx.x;
  if (t < 64) {
    int lx = t & 7, ly = t >> 3;
    int x = bx * 8 + lx, y = by * 8 + ly;
    float v = 0.f;
    if (x < p.w[2] && y < p.h[2]) {
      v = (L1[2 * ly][2 * lx] + L1[2 * ly][2 * lx + 1] + L1[2 * ly + 1][2 * lx] + L1[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[2][x + y * p.w[2]] = v;
    }
    L2[ly][lx] = v;
  }
  __syncthreads();
  if (t < 16) {
    int lx = t & 3, ly = t >> 2;
    int x = bx * 4 + lx, y = by * 4 + ly;
    float v = 0.f;
    if (x < p.w[3] && y < p.h[3]) {
      v = (L2[2 * ly][2 * lx] + L2[2 * ly][2 * lx + 1] + L2[2 * ly + 1][2 * lx] + L2[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[3][x + y * p.w[3]] = v;
    }
    L3[ly][lx] = v;
  }
  __syncthreads();
  if (t < 4) {
    int lx = t & 1, ly = t >> 1;
    int x = bx * 2 + lx, y = by * 2 + ly;
    if (x < p.w[4] && y < p.h[4]) {
      float v = (L3[2 * ly][2 * lx] + L3[2 * ly][2 * lx + 1] + L3[2 * ly + 1][2 * lx] + L3[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[4][x + y * p.w[4]] = v;
    }
  }
}

// Levels 1-4 of a 32x32 level-0 tile, read from the 8-bit image u8 when it
// is set (writing level 0 from it too when writeL0) or from the float level 0.
__device__ inline void pyr_images_tile(const PyrImgArgs& p, const uint8_t* __restrict__ u8, bool writeL0, int bx,
                                       int by) {
  __shared__ float L1[16][17];
  const int t = threadIdx.x;
  const int w0 = p.w[0], h0 = p.h[0];
  if (u8 && writeL0) {
    for (int k = t; k < 1024; k += 256) {
      int x = bx * 32 + (k & 31), y = by * 32 + (k >> 5);
      if (x < w0 && y < h0) p.img[0][x + y * w0] = (float)u8[x + y * w0];
    }
  }
  // level 1
  {
    int lx = t & 15, ly = t >> 4;
    int x = bx * 16 + lx, y = by * 16 + ly;
    float v = 0.f;
    if (x < p.w[1] && y < p.h[1]) {
      int s = 2 * x + 2 * y * w0;
      float a, b, c, d;
      if (u8) {
        a = (float)u8[s]; b = (float)u8[s + 1]; c = (float)u8[s + w0]; d = (float)u8[s + w0 + 1];
      } else {
        const float* I = p.img[0];
        a = I[s]; b = I[s + 1]; c = I[s + w0]; d = I[s + w0 + 1];
      }
      v = (a + b + c + d) * 0.25f;
      p.img[1][x + y * p.w[1]] = v;
    }
    L1[ly][lx] = v;
  }
  __syncthreads();
  pyr_levels_2to4(p, L1, bx, by);
}

// Levels 1-4 of a 64x64 level-0 tile (the batched rebuild: a quarter of the
// blocks of the 32x32 tiling, which leaves most threads idle at the coarse
// levels), read from the 8-bit level 0 when u8 is set, else from the float
// copy.  Same operations per pixel as pyr_images_tile.
__device__ inline void pyr_images_tile64(const PyrImgArgs& p, const uint8_t* __restrict__ u8, int bx, int by) {
  __shared__ float L1[32][33];
  __shared__ float L2[16][17];
  __shared__ float L3[8][9];
  const int t = threadIdx.x;
  const int w0 = p.w[0];
  // level 1: 32x32, four pixels per thread, every load issued first.  An
  // 8-bit level 0 of even width is read as byte pairs: s is even, so both
  // rows' pairs are 2-byte aligned (the buffer itself is 256-byte aligned).
  const bool pairs = u8 && (w0 & 1) == 0;
  float q[4][4];
  bool ok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t + 256 * i, lx = k & 31, ly = k >> 5;
    const int x = bx * 32 + lx, y = by * 32 + ly;
    ok[i] = x < p.w[1] && y < p.h[1];
    const int s = ok[i] ? 2 * x + 2 * y * w0 : 0;
    if (pairs) {
      const unsigned a = *(const unsigned short*)(u8 + s), b = *(const unsigned short*)(u8 + s + w0);
      q[i][0] = (float)(a & 0xffu); q[i][1] = (float)(a >> 8); q[i][2] = (float)(b & 0xffu); q[i][3] = (float)(b >> 8);
    } else if (u8) {
      q[i][0] = (float)u8[s]; q[i][1] = (float)u8[s + 1]; q[i][2] = (float)u8[s + w0]; q[i][3] = (float)u8[s + w0 + 1];
    } else {
      const float* I = p.img[0];
      q[i][0] = I[s]; q[i][1] = I[s + 1]; q[i][2] = I[s + w0]; q[i][3] = I[s + w0 + 1];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t + 256 * i, lx = k & 31, ly = k >> 5;
    float v = 0.f;
    if (ok[i]) {
      v = (q[i][0] + q[i][1] + q[i][2] + q[i][3]) * 0.25f;
      p.img[1][bx * 32 + lx + (by * 32 + ly) * p.w[1]] = v;
    }
    L1[ly][lx] = v;
  }
  __syncthreads();
  {
    const int lx = t & 15, ly = t >> 4;
    const int x = bx * 16 + lx, y = by * 16 + ly;
    float v = 0.f;
    if (x < p.w[2] && y < p.h[2]) {
      v = (L1[2 * ly][2 * lx] + L1[2 * ly][2 * lx + 1] + L1[2 * ly + 1][2 * lx] + L1[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[2][x + y * p.w[2]] = v;
    }
    L2[ly][lx] = v;
  }
  __syncthreads();
  if (t < 64) {
    const int lx = t & 7, ly = t >> 3;
    const int x = bx * 8 + lx, y = by * 8 + ly;
    float v = 0.f;
    if (x < p.w[3] && y < p.h[3]) {
      v = (L2[2 * ly][2 * lx] + L2[2 * ly][2 * lx + 1] + L2[2 * ly + 1][2 * lx] + L2[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[3][x + y * p.w[3]] = v;
    }
    L3[ly][lx] = v;
  }
  __syncthreads();
  if (t < 16) {
    const int lx = t & 3, ly = t >> 2;
    const int x = bx * 4 + lx, y = by * 4 + ly;
    if (x < p.w[4] && y < p.h[4]) {
      const float v =
          (L3[2 * ly][2 * lx] + L3[2 * ly][2 * lx + 1] + L3[2 * ly + 1][2 * lx] + L3[2 * ly + 1][2 * lx + 1]) * 0.25f;
      p.img[4][x + y * p.w[4]] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_pyr_images(PyrImgArgs p, int writeL0) {
  pyr_images_tile(p, p.u8, writeL0 != 0, blockIdx.x, blockIdx.y);
}

// Undistorter::undistort for one output pixel: the source position m (from
// the undistortion table) is bilinearly interpolated in the 8-bit input and
// truncated back to 8 bits; m.x < 0 has no source and gives 0.  Table entries
// are checked on the host (0 <= m < size - 1), so both taps are in bounds.
__device__ inline uint8_t undistort_pixel(const uint8_t* __restrict__ src, int inW, float2 m) {
  if (m.x < 0.0f) return 0;
  const int xi = (int)m.x, yi = (int)m.y;
  const float xx = m.x - (float)xi, yy = m.y - (float)yi;
  const float xxyy = xx * yy;
  const uint8_t* b = src + xi + yi * inW;
  const float v = xxyy * (float)b[1 + inW] + (yy - xxyy) * (float)b[inW] + (xx - xxyy) * (float)b[1] +
                  (1.0f - xx - yy + xxyy) * (float)b[0];
  return (uint8_t)v;
}

// InputThread's ingest of one camera image (lib/App/InputThread.cpp:59-71):
// undistortion, the 8-bit to float conversion of Frame's constructor and
// Frame::buildImage for levels 1-4, in one launch.  A block owns a 16x16
// level-0 tile (one pixel per thread: the table entry, then the source taps,
// two waits on memory), writes level 0 (and, when `undist` is set, the 8-bit
// undistorted image), and reduces the tile level by level in shared memory:
// a 16x16 tile holds whole 2^4 blocks of level 0, so levels 1-4 follow from
// it alone.
__global__ __launch_bounds__(256) void k_ingest(PyrImgArgs p, const uint8_t* __restrict__ src, int inW,
                                                const float2* __restrict__ map, uint8_t* __restrict__ undist) {
  static_assert(PYR == 5, "a 16x16 tile holds whole 2^(PYR-1) blocks");
  __shared__ float T[16][17];
  const int t = threadIdx.x, lx = t & 15, ly = t >> 4;
  const int x = blockIdx.x * 16 + lx, y = blockIdx.y * 16 + ly;
  const bool inside = x < p.w[0] && y < p.h[0];
  const int i = inside ? x + y * p.w[0] : 0;
  const uint8_t u = undistort_pixel(src, inW, inside ? map[i] : make_float2(-1.0f, -1.0f));
  float v = 0.0f;
  if (inside) {
    v = (float)u;
    p.img[0][i] = v;
    if (undist) undist[i] = u;
  }
  T[ly][lx] = v;
  __syncthreads();
#pragma unroll
  for (int l = 1, n = 8; l < PYR; ++l, n >>= 1) {
    const int tx = t % n, ty = t / n;
    float c = 0.0f;
    if (t < n * n) {
      const int cx = blockIdx.x * n + tx, cy = blockIdx.y * n + ty;
      if (cx < p.w[l] && cy < p.h[l]) {
        c = (T[2 * ty][2 * tx] + T[2 * ty][2 * tx + 1] + T[2 * ty + 1][2 * tx] + T[2 * ty + 1][2 * tx + 1]) * 0.25f;
        p.img[l][cx + cy * p.w[l]] = c;
      }
    }
    __syncthreads();  // the level's reads end before its values replace them
    if (t < n * n) T[ty][tx] = c;
    __syncthreads();
  }
}

struct PyrGradArgs {
  const float* img[PYR];
  float4* grad[PYR];
  uint2* gradh[PYR];
  float* maxg[PYR];
  int w[PYR], h[PYR];
  int blk0[PYR + 1];
  int prec;
  int ppt;  // pixels per thread: a block covers ppt x 256 consecutive pixels of one level
};

__device__ inline float absgrad_at(const float* __restrict__ I, int j, int w, int h) {
  if (j >= w && j < w * (h - 1)) {
    float gx = 0.5f * (I[j + 1] - I[j - 1]);
    float gy = 0.5f * (I[j + w] - I[j - w]);
    return sqrtf(gx * gx + gy * gy);
  }
  return 0.f;
}
__device__ inline float vmax_at(const float* __restrict__ I, int j, int w, int h) {
  if (j >= w + 1 && j < w * (h - 1) - 1) {
    float g1 = absgrad_at(I, j - w, w, h), g2 = absgrad_at(I, j, w, h);
    if (g1 < g2) g1 = g2;
    float g3 = absgrad_at(I, j + w, w, h);
    return (g1 < g3) ? g3 : g1;
  }
  return 0.f;
}

__device__ inline unsigned short f32_to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline uint2 pack_grad(float4 g, int prec) {
  unsigned short a, b, c;
  if (prec == LSDG_BF16) {
    a = f32_to_bf16(g.x); b = f32_to_bf16(g.y); c = f32_to_bf16(g.z);
  } else {
    a = __half_as_ushort(__float2half_rn(g.x));
    b = __half_as_ushort(__float2half_rn(g.y));
    c = __half_as_ushort(__float2half_rn(g.z));
  }
  return make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)c);
}

// Pixels per thread of the batched gradient stage (a single frame's keeps
// one per thread, for more blocks on its smaller grid).
constexpr int GRAD_PPT_BATCH = 4;

__device__ inline void pyr_grads_block(const PyrGradArgs& p, int b) {
  int l = 0;
  while (l < PYR - 1 && b >= p.blk0[l + 1]) l++;
  const int w = p.w[l], h = p.h[l];
  const float* I = p.img[l];
  // pixel i of thread t at t + 256 i
  const int base = (b - p.blk0[l]) * 256 * p.ppt + threadIdx.x;
  for (int i = 0; i < p.ppt; ++i) {
    const int idx = base + 256 * i;
    if (idx >= w * h) break;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (idx >= w && idx < w * (h - 1)) {
      g.x = 0.5f * (I[idx + 1] - I[idx - 1]);
      g.y = 0.5f * (I[idx + w] - I[idx - w]);
      g.z = I[idx];
    }
    if (p.grad[l]) p.grad[l][idx] = g;
    if (p.gradh[l]) p.gradh[l][idx] = pack_grad(g, p.prec);
    if (!p.maxg[l]) continue;
    float m;
    if (idx >= w + 1 && idx < w * (h - 1) - 1) {
      float g1 = vmax_at(I, idx - 1, w, h), g2 = vmax_at(I, idx, w, h);
      if (g1 < g2) g1 = g2;
      float g3 = vmax_at(I, idx + 1, w, h);
      m = (g1 < g3) ? g3 : g1;
    } else {
      m = absgrad_at(I, idx, w, h);
    }
    p.maxg[l][idx] = m;
  }
}

__global__ __launch_bounds__(256) void k_pyr_grads(PyrGradArgs p) { pyr_grads_block(p, blockIdx.x); }

// Batched pyramids: every frame of a batch in one launch per stage.  Block b
// belongs to the frame f with off[f] <= b < off[f + 1].
__device__ inline int frame_of_block(const int* __restrict__ off, int nf, int b) {
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_pyr_images_batch(const PyrImgArgs* __restrict__ args,
                                                          const int* __restrict__ off, const int* __restrict__ tx,
                                                          int nf) {
  int f = frame_of_block(off, nf, blockIdx.x);
  int local = blockIdx.x - off[f];
  pyr_images_tile64(args[f], args[f].u8, local % tx[f], local / tx[f]);
}

__global__ __launch_bounds__(256) void k_pyr_grads_batch(const PyrGradArgs* __restrict__ args,
                                                         const int* __restrict__ off, int nf) {
  int f = frame_of_block(off, nf, blockIdx.x);
  pyr_grads_block(args[f], blockIdx.x - off[f]);
}

__global__ __launch_bounds__(256) void k_pyr_depth(PyrDepthArgs p) {
  __shared__ float I1[16][17], V1[16][17], I2[8][9], V2[8][9], I3[4][5], V3[4][5];
  const int t = threadIdx.x, bx = blockIdx.x, by = blockIdx.y;
  const int w0 = p.w[0];
  {
    int lx = t & 15, ly = t >> 4, x = bx * 16 + lx, y = by * 16 + ly;
    float oi = -1.f, ov = -1.f;
    if (x < p.w[1] && y < p.h[1]) {
      int s = 2 * x + 2 * y * w0;
      float ia[4] = {p.id[0][s], p.id[0][s + 1], p.id[0][s + w0], p.id[0][s + w0 + 1]};
      float va[4] = {p.var[0][s], p.var[0][s + 1], p.var[0][s + w0], p.var[0][s + w0 + 1]};
      fuse4(ia, va, oi, ov);
      p.id[1][x + y * p.w[1]] = oi;
      p.var[1][x + y * p.w[1]] = ov;
    }
    I1[ly][lx] = oi;
    V1[ly][lx] = ov;
  }
  __syncthreads();
  if (t < 64) {
    int lx = t & 7, ly = t >> 3, x = bx * 8 + lx, y = by * 8 + ly;
    float oi = -1.f, ov = -1.f;
    if (x < p.w[2] && y < p.h[2]) {
      float ia[4] = {I1[2 * ly][2 * lx], I1[2 * ly][2 * lx + 1], I1[2 * ly + 1][2 * lx], I1[2 * ly + 1][2 * lx + 1]};
      float va[4] = {V1[2 * ly][2 * lx], V1[2 * ly][2 * lx + 1], V1[2 * ly + 1][2 * lx], V1[2 * ly + 1][2 * lx + 1]};
      fuse4(ia, va, oi, ov);
      p.id[2][x + y * p.w[2]] = oi;
      p.var[2][x + y * p.w[2]] = ov;
    }
    I2[ly][lx] = oi;
    V2[ly][lx] = ov;
  }
  __syncthreads();
  if (t < 16) {
    int lx = t & 3, ly = t >> 2, x = bx * 4 + lx, y = by * 4 + ly;
    float oi = -1.f, ov = -1.f;
    if (x < p.w[3] && y < p.h[3]) {
      float ia[4] = {I2[2 * ly][2 * lx], I2[2 * ly][2 * lx + 1], I2[2 * ly + 1][2 * lx], I2[2 * ly + 1][2 * lx + 1]};
      float va[4] = {V2[2 * ly][2 * lx], V2[2 * ly][2 * lx + 1], V2[2 * ly + 1][2 * lx], V2[2 * ly + 1][2 * lx + 1]};
      fuse4(ia, va, oi, ov);
      p.id[3][x + y * p.w[3]] = oi;
      p.var[3][x + y * p.w[3]] = ov;
    }
    I3[ly][lx] = oi;
    V3[ly][lx] = ov;
  }
  __syncthreads();
  if (t < 4) {
    int lx = t & 1, ly = t >> 1, x = bx * 2 + lx, y = by * 2 + ly;
    if (x < p.w[4] && y < p.h[4]) {
      float ia[4] = {I3[2 * ly][2 * lx], I3[2 * ly][2 * lx + 1], I3[2 * ly + 1][2 * lx], I3[2 * ly + 1][2 * lx + 1]};
      float va[4] = {V3[2 * ly][2 * lx], V3[2 * ly][2 * lx + 1], V3[2 * ly + 1][2 * lx], V3[2 * ly + 1][2 * lx + 1]};
      float oi, ov;
      fuse4(ia, va, oi, ov);
      p.id[4][x + y * p.w[4]] = oi;
      p.var[4][x + y * p.w[4]] = ov;
    }
  }
}

// Pixels without a usable depth become (-1, -1).
__global__ void k_depth_normalize(float* __restrict__ id, float* __restrict__ var, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a = id[i], v = var[i];
  bool ok = isfinite(a) && v > 0.f && a > 0.f;
  if (!ok) {
    id[i] = -1.f;
    var[i] = -1.f;
  }
}

// ---------------------------------------------------------- point cloud
__device__ inline bool pc_valid(const float* id, const float* var, int idx, int w, int h) {
  int y = idx / w, x = idx - y * w;
  if (x < 1 || x >= w - 1 || y < 1 || y >= h - 1) return false;
  return !(var[idx] <= 0.f || id[idx] == 0.f);
}

__global__ __launch_bounds__(256) void k_pc_count(const float* __restrict__ id, const float* __restrict__ var,
                                                  int w, int h, int* __restrict__ counts) {
  int idx = blockIdx.x * 256 + threadIdx.x;
  bool v = idx < w * h && pc_valid(id, var, idx, w, h);
  int c = __syncthreads_count(v);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// Exclusive scan of n block counts in one 1024-thread block; total at counts[n].
__global__ __launch_bounds__(1024) void k_pc_scan(int* counts, int n) {
  __shared__ int sh[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < n ? counts[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int a = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < n) counts[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[n] = carry;
}

struct PcWriteArgs {
  const float* id;
  const float* var;
  const float* img;
  const float4* grad;
  int w, h;
  float fxi, fyi, cxi, cyi;
  float4* posI;
  float* pvar;
  float2* pgrad;
  int* pidx;
};

__global__ __launch_bounds__(256) void k_pc_write(PcWriteArgs a, const int* __restrict__ offsets) {
  __shared__ int wave_count[4];
  int idx = blockIdx.x * 256 + threadIdx.x;
  bool v = idx < a.w * a.h && pc_valid(a.id, a.var, idx, a.w, a.h);
  unsigned long long m = __ballot(v);
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int rank = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_count[wv] = __popcll(m);
  __syncthreads();
  int base = offsets[blockIdx.x];
  for (int k = 0; k < wv; ++k) base += wave_count[k];
  if (!v) return;
  int o = base + rank;
  int y = idx / a.w, x = idx - y * a.w;
  float inv = 1.0f / a.id[idx];
  float px = inv * (a.fxi * (float)x + a.cxi);
  float py = inv * (a.fyi * (float)y + a.cyi);
  float pz = inv * 1.0f;
  a.posI[o] = make_float4(px, py, pz, a.img[idx]);
  a.pvar[o] = a.var[idx];
  float4 g = a.grad[idx];
  a.pgrad[o] = make_float2(g.x, g.y);
  a.pidx[o] = idx;
}

// ------------------------------------------------------------ launchers
hipError_t launch_pyr_images(lsdg_frame f, bool fromU8, hipStream_t s) {
  PyrImgArgs p;
  for (int l = 0; l < PYR; ++l) {
    p.img[l] = f->d_image[l];
    p.w[l] = f->w[l];
    p.h[l] = f->h[l];
  }
  // level 0 from the 8-bit staging (converted and written), or levels 1-4
  // only, read from the 8-bit copy when it still holds level 0
  p.u8 = (fromU8 || f->u8Level0) ? f->d_input_u8 : nullptr;
  dim3 grid((f->w[0] + 31) / 32, (f->h[0] + 31) / 32);
  hipLaunchKernelGGL(k_pyr_images, grid, dim3(256), 0, s, p, fromU8 ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ingest(lsdg_frame f, const uint8_t* d_src, int inW, const float2* d_map, uint8_t* d_undist,
                         hipStream_t s) {
  PyrImgArgs p;
  for (int l = 0; l < PYR; ++l) {
    p.img[l] = f->d_image[l];
    p.w[l] = f->w[l];
    p.h[l] = f->h[l];
  }
  dim3 grid((f->w[0] + 15) / 16, (f->h[0] + 15) / 16);
  k_ingest<<<grid, dim3(256), 0, s>>>(p, d_src, inW, d_map, d_undist);
  return hipGetLastError();
}

// Fills the gradient-stage arguments of frame f for the data in `need`, ppt
// pixels per thread; returns the number of blocks (levels with nothing to
// build get none).
static int pyr_grad_args(lsdg_frame f, unsigned need, int ppt, PyrGradArgs& p) {
  p.ppt = ppt;
  int blk = 0;
  for (int l = 0; l < PYR; ++l) {
    p.img[l] = f->d_image[l];
    p.grad[l] = (need & fd_bit(FD_GRAD, l)) ? f->d_grad[l] : nullptr;
    p.gradh[l] = (need & fd_bit(FD_GRADH, l)) ? f->d_gradh[l] : nullptr;
    p.maxg[l] = (need & fd_bit(FD_MAXG, l)) ? f->d_maxg[l] : nullptr;
    p.w[l] = f->w[l];
    p.h[l] = f->h[l];
    p.blk0[l] = blk;
    if (p.grad[l] || p.gradh[l] || p.maxg[l]) blk += (f->w[l] * f->h[l] + 256 * ppt - 1) / (256 * ppt);
  }
  p.blk0[PYR] = blk;
  p.prec = f->precision;
  return blk;
}

hipError_t launch_pyr_grads(lsdg_frame f, unsigned need, hipStream_t s) {
  PyrGradArgs p;
  int blk = pyr_grad_args(f, need, 1, p);
  if (blk == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pyr_grads, dim3(blk), dim3(256), 0, s, p);
  return hipGetLastError();
}

// Host tables for the batched launches (filled once per batch plan).
size_t pyr_batch_bytes(int nf) {
  return (sizeof(PyrImgArgs) + sizeof(PyrGradArgs)) * (size_t)nf + sizeof(int) * (size_t)(3 * nf + 2) + 64;
}

// Writes the argument tables for frames fs[0..nf) into host memory `h` (laid
// out as the device copy at `d`), returns the block totals of both stages.
void pyr_batch_tables(lsdg_frame const* fs, const unsigned* need, int nf, void* h, void* d, PyrBatch* out) {
  char* hb = (char*)h;
  PyrImgArgs* ia = (PyrImgArgs*)hb;
  PyrGradArgs* ga = (PyrGradArgs*)(hb + sizeof(PyrImgArgs) * nf);
  int* ioff = (int*)(hb + (sizeof(PyrImgArgs) + sizeof(PyrGradArgs)) * nf);
  int* itx = ioff + nf + 1;
  int* goff = itx + nf;
  int ib = 0, gb = 0;
  for (int k = 0; k < nf; ++k) {
    lsdg_frame f = fs[k];
    PyrImgArgs p;
    PyrGradArgs q;
    for (int l = 0; l < PYR; ++l) {
      p.img[l] = f->d_image[l];
      p.w[l] = f->w[l];
      p.h[l] = f->h[l];
    }
    p.u8 = f->u8Level0 ? f->d_input_u8 : nullptr;
    int blk = pyr_grad_args(f, need[k], GRAD_PPT_BATCH, q);
    ia[k] = p;
    ga[k] = q;
    int tx = (f->w[0] + 63) / 64, ty = (f->h[0] + 63) / 64;  // 64x64 tiles (pyr_images_tile64)
    ioff[k] = ib;
    itx[k] = tx;
    ib += tx * ty;
    goff[k] = gb;
    gb += blk;
  }
  ioff[nf] = ib;
  goff[nf] = gb;
  char* db = (char*)d;
  out->img_args = db;
  out->grad_args = db + sizeof(PyrImgArgs) * nf;
  out->img_off = (const int*)(db + (sizeof(PyrImgArgs) + sizeof(PyrGradArgs)) * nf);
  out->img_tx = out->img_off + nf + 1;
  out->grad_off = out->img_tx + nf;
  out->img_blocks = ib;
  out->grad_blocks = gb;
  out->nf = nf;
}

hipError_t launch_pyr_batch(const PyrBatch& b, hipStream_t s) {
  if (b.nf == 0) return hipSuccess;
  k_pyr_images_batch<<<dim3(b.img_blocks), dim3(256), 0, s>>>((const PyrImgArgs*)b.img_args, b.img_off, b.img_tx,
                                                               b.nf);
  if (b.grad_blocks > 0)
    k_pyr_grads_batch<<<dim3(b.grad_blocks), dim3(256), 0, s>>>((const PyrGradArgs*)b.grad_args, b.grad_off, b.nf);
  return hipGetLastError();
}

hipError_t launch_pyr_depth(lsdg_frame f, hipStream_t s) {
  PyrDepthArgs p;
  for (int l = 0; l < PYR; ++l) {
    p.id[l] = f->d_idepth[l];
    p.var[l] = f->d_ivar[l];
    p.w[l] = f->w[l];
    p.h[l] = f->h[l];
  }
  dim3 grid((f->w[0] + 31) / 32, (f->h[0] + 31) / 32);
  hipLaunchKernelGGL(k_pyr_depth, grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_depth_normalize(float* id, float* var, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_depth_normalize, dim3((n + 255) / 256), dim3(256), 0, s, id, var, n);
  return hipGetLastError();
}

hipError_t launch_point_cloud(lsdg_ref r, int l, hipStream_t s) {
  lsdg_frame f = r->kf;
  int w = f->w[l], h = f->h[l];
  int nb = (w * h + 255) / 256;
  hipError_t e;
  if (r->blockcounts_cap < nb + 1) {
    if (r->d_blockcounts) (void)hipFree(r->d_blockcounts);
    r->d_blockcounts = nullptr;
    e = hipMalloc(&r->d_blockcounts, sizeof(int) * (nb + 1));
    if (e != hipSuccess) return e;
    r->blockcounts_cap = nb + 1;
  }
  hipLaunchKernelGGL(k_pc_count, dim3(nb), dim3(256), 0, s, f->d_idepth[l], f->d_ivar[l], w, h, r->d_blockcounts);
  hipLaunchKernelGGL(k_pc_scan, dim3(1), dim3(1024), 0, s, r->d_blockcounts, nb);
  int total = 0;
  e = hipMemcpyAsync(&total, r->d_blockcounts + nb, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  if (total > r->cap[l]) {
    if (r->d_pool[l]) (void)hipFree(r->d_pool[l]);
    r->d_pool[l] = nullptr;
    size_t bytes = (size_t)total * (sizeof(float4) + sizeof(float) + sizeof(float2) + sizeof(int)) + 256;
    e = hipMalloc(&r->d_pool[l], bytes);
    if (e != hipSuccess) return e;
    r->cap[l] = total;
  }
  char* base = (char*)r->d_pool[l];
  PointsLevel& pl = r->lv[l];
  pl.posI = (const float4*)base;
  pl.grad = (const float2*)(base + (size_t)total * sizeof(float4));
  pl.var = (const float*)(base + (size_t)total * (sizeof(float4) + sizeof(float2)));
  pl.idx = (const int*)(base + (size_t)total * (sizeof(float4) + sizeof(float2) + sizeof(float)));
  pl.n = total;
  r->num[l] = total;
  if (total == 0) return hipSuccess;
  PcWriteArgs a;
  a.id = f->d_idepth[l];
  a.var = f->d_ivar[l];
  a.img = f->d_image[l];
  a.grad = f->d_grad[l];
  a.w = w;
  a.h = h;
  a.fxi = f->fxi[l];
  a.fyi = f->fyi[l];
  a.cxi = f->cxi[l];
  a.cyi = f->cyi[l];
  a.posI = (float4*)pl.posI;
  a.pvar = (float*)pl.var;
  a.pgrad = (float2*)pl.grad;
  a.pidx = (int*)pl.idx;
  hipLaunchKernelGGL(k_pc_write, dim3(nb), dim3(256), 0, s, a, r->d_blockcounts);
  return hipGetLastError();
}

}  // namespace lsdg
