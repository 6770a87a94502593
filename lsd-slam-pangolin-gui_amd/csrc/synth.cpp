// Deterministic synthetic scenes for the LSD-SLAM front-end workloads.
//
// A scene is three textured planes (an unbounded back wall, a floor and a
// bounded slanted disc) carrying seeded improved-Perlin (Perlin 2002) fractal
// textures in plane coordinates, so every camera that sees a surface point sees
// the same intensity.  Images are rendered by analytic ray/plane intersection
// at pixel centres; sensor noise is Gaussian (Box-Muller over a counter-based
// hash), then rounded and clamped to uint8 like the CV_8UC1 frames the
// reference feeds SlamSystem::nextImage (lib/App/InputThread.cpp:59-71).
//
// Everything is a pure function of (seed, pose, camera), so any host
// regenerates bit-identical inputs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// splitmix64 stream
struct Stream {
  uint64_t s;
  explicit Stream(uint64_t seed) : s(seed) {}
  uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};
// counter-based: value for (key, counter) without state
inline uint64_t hash2(uint64_t key, uint64_t ctr) {
  return mix64(mix64(key + 0x9E3779B97F4A7C15ull) ^ (ctr * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
}
inline double u01(uint64_t x) { return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

struct Perlin {
  int p[512];
  void init(uint64_t seed) {
    Stream st(seed);
    int perm[256];
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {
      int j = (int)(st.next() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int hsh, double x, double y) {
    switch (hsh & 7) {
      case 0: return x + y;
      case 1: return -x + y;
      case 2: return x - y;
      case 3: return -x - y;
      case 4: return x;
      case 5: return -x;
      case 6: return y;
      default: return -y;
    }
  }
  double noise(double x, double y) const {
    double fx = std::floor(x), fy = std::floor(y);
    int X = ((int)(long long)fx) & 255, Y = ((int)(long long)fy) & 255;
    x -= fx;
    y -= fy;
    double u = fade(x), v = fade(y);
    int A = p[X] + Y, B = p[X + 1] + Y;
    return lerp(v, lerp(u, grad(p[A], x, y), grad(p[B], x - 1, y)),
                lerp(u, grad(p[A + 1], x, y - 1), grad(p[B + 1], x - 1, y - 1)));
  }
};

struct Plane {
  double n[3];   // unit normal
  double c;      // n . X = c
  double e1[3], e2[3];
  double centre[3];
  double radius; // <= 0: unbounded
  double freq, amp, base;
  double off[8][2];
  Perlin perlin;
};

}  // namespace

extern "C" {

struct synth_scene {
  Plane pl[3];
  int octaves;
};

static void normalize3(double* v) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n; v[1] /= n; v[2] /= n;
}
static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static void basis(Plane& P) {
  double a[3] = {1, 0, 0};
  if (std::fabs(P.n[0]) > 0.8) { a[0] = 0; a[1] = 1; }
  cross3(P.n, a, P.e1); normalize3(P.e1);
  cross3(P.n, P.e1, P.e2); normalize3(P.e2);
}

// contrast scales the texture amplitude; freq_scale the spatial frequency.
size_t synth_scene_size(void) { return sizeof(synth_scene); }

void synth_scene_init(synth_scene* S, uint64_t seed, double contrast, double freq_scale) {
  Stream st(mix64(seed ^ 0x5CE9E5EEDull));
  S->octaves = 4;
  // back wall: z ~ 3.6..4.4 m with a small tilt, unbounded
  {
    Plane& P = S->pl[0];
    P.n[0] = (st.uniform() - 0.5) * 0.5; P.n[1] = (st.uniform() - 0.5) * 0.4; P.n[2] = -1.0;
    normalize3(P.n);
    double z0 = 3.6 + 0.8 * st.uniform();
    P.c = P.n[2] * z0;  // passes through (0,0,z0)
    P.radius = -1;
    P.centre[0] = 0; P.centre[1] = 0; P.centre[2] = z0;
  }
  // floor: y = y0 (camera y points down), slightly tilted, unbounded
  {
    Plane& P = S->pl[1];
    P.n[0] = (st.uniform() - 0.5) * 0.2; P.n[1] = -1.0; P.n[2] = (st.uniform() - 0.5) * 0.2;
    normalize3(P.n);
    double y0 = 0.45 + 0.3 * st.uniform();
    P.c = P.n[1] * y0;
    P.radius = -1;
    P.centre[0] = 0; P.centre[1] = y0; P.centre[2] = 2.0;
  }
  // slanted disc at 1.0..1.8 m in the upper half
  {
    Plane& P = S->pl[2];
    double ang = (0.35 + 0.35 * st.uniform()) * (st.uniform() < 0.5 ? -1 : 1);
    P.n[0] = std::sin(ang); P.n[1] = 0.25 * (st.uniform() - 0.5); P.n[2] = -std::cos(ang);
    normalize3(P.n);
    double z0 = 1.0 + 0.8 * st.uniform();
    double x0 = (st.uniform() - 0.5) * 0.8 * z0;
    double y0 = -(0.05 + 0.2 * st.uniform()) * z0;
    P.centre[0] = x0; P.centre[1] = y0; P.centre[2] = z0;
    P.c = P.n[0] * x0 + P.n[1] * y0 + P.n[2] * z0;
    P.radius = 0.35 + 0.25 * st.uniform();
  }
  for (int i = 0; i < 3; ++i) {
    Plane& P = S->pl[i];
    basis(P);
    P.perlin.init(st.next());
    P.freq = (4.0 + 3.0 * st.uniform()) * freq_scale;  // cycles per metre, base octave
    P.amp = contrast * (90.0 + 30.0 * st.uniform());
    P.base = 110.0 + 40.0 * st.uniform();
    for (int k = 0; k < 8; ++k) {
      P.off[k][0] = 1000.0 * st.uniform();
      P.off[k][1] = 1000.0 * st.uniform();
    }
  }
}

static double shade(const synth_scene* S, const Plane& P, const double* X) {
  double s = X[0] * P.e1[0] + X[1] * P.e1[1] + X[2] * P.e1[2];
  double t = X[0] * P.e2[0] + X[1] * P.e2[1] + X[2] * P.e2[2];
  double v = 0.0, a = 1.0, f = P.freq;
  for (int k = 0; k < S->octaves; ++k) {
    v += a * P.perlin.noise(s * f + P.off[k][0], t * f + P.off[k][1]);
    a *= 0.5;
    f *= 2.0;
  }
  return P.base + P.amp * v;
}

static void quat_to_R(const double* q, double* R) {
  double x = q[0], y = q[1], z = q[2], w = q[3];
  double n = std::sqrt(x * x + y * y + z * z + w * w);
  x /= n; y /= n; z /= n; w /= n;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w); R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w); R[7] = 2 * (y * z + x * w); R[8] = 1 - 2 * (x * x + y * y);
}

// Renders the scene from a camera with pose cam_to_world (qx,qy,qz,qw,tx,ty,tz).
// intensity_out (optional): noise-free float intensity; img_out (optional):
// uint8 after N(0, noise_sigma^2) noise, rounding and clamping; depth_out
// (optional): camera-frame z of the visible surface.
void synth_render(const synth_scene* S, const double* cam_to_world, int w, int h, double fx,
                  double fy, double cx, double cy, double noise_sigma, uint64_t noise_seed,
                  float* intensity_out, uint8_t* img_out, float* depth_out) {
  double R[9];
  quat_to_R(cam_to_world, R);
  const double* o = cam_to_world + 4;
#pragma omp parallel for schedule(static)
  for (int v = 0; v < h; ++v) {
    for (int u = 0; u < w; ++u) {
      double dc[3] = {(u - cx) / fx, (v - cy) / fy, 1.0};
      double d[3] = {R[0] * dc[0] + R[1] * dc[1] + R[2] * dc[2],
                     R[3] * dc[0] + R[4] * dc[1] + R[5] * dc[2],
                     R[6] * dc[0] + R[7] * dc[1] + R[8] * dc[2]};
      double best = 1e30;
      int bi = -1;
      for (int i = 0; i < 3; ++i) {
        const Plane& P = S->pl[i];
        double den = P.n[0] * d[0] + P.n[1] * d[1] + P.n[2] * d[2];
        if (std::fabs(den) < 1e-12) continue;
        double s = (P.c - (P.n[0] * o[0] + P.n[1] * o[1] + P.n[2] * o[2])) / den;
        if (!(s > 1e-6) || s >= best) continue;
        if (P.radius > 0) {
          double X[3] = {o[0] + s * d[0] - P.centre[0], o[1] + s * d[1] - P.centre[1],
                         o[2] + s * d[2] - P.centre[2]};
          if (X[0] * X[0] + X[1] * X[1] + X[2] * X[2] > P.radius * P.radius) continue;
        }
        best = s;
        bi = i;
      }
      double val = 128.0, z = 1e30;
      if (bi >= 0) {
        double X[3] = {o[0] + best * d[0], o[1] + best * d[1], o[2] + best * d[2]};
        val = shade(S, S->pl[bi], X);
        z = best;  // dc has unit z, so the ray parameter is the camera-frame depth
      }
      size_t idx = (size_t)u + (size_t)v * (size_t)w;
      if (intensity_out) intensity_out[idx] = (float)val;
      if (depth_out) depth_out[idx] = (float)z;
      if (img_out) {
        double nv = 0.0;
        if (noise_sigma > 0) {
          uint64_t a = hash2(noise_seed, 2 * idx), b = hash2(noise_seed, 2 * idx + 1);
          nv = noise_sigma * std::sqrt(-2.0 * std::log(u01(a))) * std::cos(6.283185307179586 * u01(b));
        }
        double q = std::floor(val + nv + 0.5);
        if (q < 0) q = 0;
        if (q > 255) q = 255;
        img_out[idx] = (uint8_t)q;
      }
    }
  }
}

// Uniform unit vector and uniform [0,1) values from a seeded stream.
// Pose with translation |t| = trans_m and rotation angle rot_rad (if
// uniform_magnitude, both scaled by independent U[0,1)).
void synth_sample_pose(uint64_t seed, double trans_m, double rot_rad, int uniform_magnitude,
                       double* out7) {
  Stream st(mix64(seed ^ 0x905Eull));
  auto unit = [&](double* v) {
    double z = 2.0 * st.uniform() - 1.0, phi = 6.283185307179586 * st.uniform();
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    v[0] = r * std::cos(phi); v[1] = r * std::sin(phi); v[2] = z;
  };
  double td[3], ax[3];
  unit(td);
  unit(ax);
  double tm = trans_m, rm = rot_rad;
  if (uniform_magnitude) {
    tm *= st.uniform();
    rm *= st.uniform();
  }
  double s = std::sin(0.5 * rm), c = std::cos(0.5 * rm);
  out7[0] = ax[0] * s; out7[1] = ax[1] * s; out7[2] = ax[2] * s; out7[3] = c;
  out7[4] = td[0] * tm; out7[5] = td[1] * tm; out7[6] = td[2] * tm;
}

// Counter-based uniform / Gaussian fields (per-pixel seeds for variances and
// depth-prior noise).
void synth_uniform(uint64_t seed, size_t n, double lo, double hi, float* out) {
  for (size_t i = 0; i < n; ++i) out[i] = (float)(lo + (hi - lo) * u01(hash2(seed, i)));
}
void synth_gaussian(uint64_t seed, size_t n, double sigma, float* out) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t a = hash2(seed, 2 * i), b = hash2(seed, 2 * i + 1);
    out[i] = (float)(sigma * std::sqrt(-2.0 * std::log(u01(a))) * std::cos(6.283185307179586 * u01(b)));
  }
}

}  // extern "C"
