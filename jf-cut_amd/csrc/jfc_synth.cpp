// jfc_synth.cpp -- deterministic synthetic workloads for the benchmark
// configurations (2D images, video volumes, uint16 volumes) with scribble
// seeds.  Host code; no reference counterpart (the reference loads images and
// brushed seeds from the GUI, QGCWidget.cpp:1865-1892 and 2243-2287).
//
// Every random quantity comes from splitmix64 streams keyed by the seed, and
// per-pixel noise from a counter-based hash of (seed, pixel, channel), so the
// output does not depend on the number of threads.  Noise is the sum of eight
// 16-bit uniforms (Irwin-Hall), scaled to unit variance.
#include "jfcut.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ULL;
    return mix64(s);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
};

// Approximately standard normal deviate for counter `idx` of stream `key`.
inline double gauss(uint64_t key, uint64_t idx) {
  const uint64_t a = mix64(key ^ mix64(2 * idx + 1));
  const uint64_t b = mix64(key ^ mix64(2 * idx + 2));
  int64_t s = 0;
  for (int i = 0; i < 4; ++i) s += (int64_t)((a >> (16 * i)) & 0xffff);
  for (int i = 0; i < 4; ++i) s += (int64_t)((b >> (16 * i)) & 0xffff);
  // mean 8 * 32767.5, variance 8 * (65536^2 - 1) / 12
  return ((double)s - 262140.0) / 65536.0 / std::sqrt(8.0 / 12.0);
}

template <class F>
void parallel_rows(int64_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (n < 64) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int64_t r = t; r < n; r += nt) f(r);
    });
  for (auto &x : th) x.join();
}

struct Blob {
  double cx, cy, sigma, amp[3];
};

struct Ellipse2 {
  double cx, cy, a, b, ang;  // ang in radians
  bool inside(double x, double y) const {
    const double dx = x - cx, dy = y - cy;
    const double c = std::cos(ang), s = std::sin(ang);
    const double u = dx * c + dy * s, v = -dx * s + dy * c;
    return (u * u) / (a * a) + (v * v) / (b * b) <= 1.0;
  }
};

const double kFg[3] = {200, 80, 60};
const double kBg[3] = {40, 90, 160};

std::vector<Blob> make_blobs(Rng &r, int W, int H, int n) {
  std::vector<Blob> b(n);
  for (auto &x : b) {
    x.cx = r.uni(0, W);
    x.cy = r.uni(0, H);
    x.sigma = r.uni(20, 60);
    for (int c = 0; c < 3; ++c) x.amp[c] = r.uni(-40, 40);
  }
  return b;
}

// Render one RGB frame (rows in parallel).
void render_frame(uint8_t *rgb, int W, int H, const Ellipse2 &E,
                  const std::vector<Blob> &blobs, double noise, uint64_t key) {
  parallel_rows(H, [&](int64_t y) {
    std::vector<double> acc((size_t)W * 3, 0.0);
    for (const Blob &b : blobs) {
      const double R = 4 * b.sigma;
      if (std::fabs(y - b.cy) > R) continue;
      const int x0 = std::max(0, (int)std::ceil(b.cx - R));
      const int x1 = std::min(W - 1, (int)std::floor(b.cx + R));
      const double dy2 = (y - b.cy) * (y - b.cy);
      for (int x = x0; x <= x1; ++x) {
        const double g = std::exp(-((x - b.cx) * (x - b.cx) + dy2) / (2 * b.sigma * b.sigma));
        for (int c = 0; c < 3; ++c) acc[(size_t)x * 3 + c] += b.amp[c] * g;
      }
    }
    for (int x = 0; x < W; ++x) {
      const double *base = E.inside(x, (double)y) ? kFg : kBg;
      for (int c = 0; c < 3; ++c) {
        const uint64_t idx = ((uint64_t)y * W + x) * 3 + c;
        double v = base[c] + acc[(size_t)x * 3 + c] + noise * gauss(key, idx);
        v = std::floor(v + 0.5);
        v = std::min(255.0, std::max(0.0, v));
        rgb[idx] = (uint8_t)v;
      }
    }
  });
}

// Mark pixels within `radius` of segment (ax,ay)-(bx,by).
void stroke2(uint8_t *seeds, int W, int H, double ax, double ay, double bx, double by,
             double radius, uint8_t val) {
  const int x0 = std::max(0, (int)std::floor(std::min(ax, bx) - radius - 1));
  const int x1 = std::min(W - 1, (int)std::ceil(std::max(ax, bx) + radius + 1));
  const int y0 = std::max(0, (int)std::floor(std::min(ay, by) - radius - 1));
  const int y1 = std::min(H - 1, (int)std::ceil(std::max(ay, by) + radius + 1));
  const double vx = bx - ax, vy = by - ay, L2 = vx * vx + vy * vy;
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      double t = L2 > 0 ? ((x - ax) * vx + (y - ay) * vy) / L2 : 0.0;
      t = std::min(1.0, std::max(0.0, t));
      const double dx = x - (ax + t * vx), dy = y - (ay + t * vy);
      if (dx * dx + dy * dy <= radius * radius) seeds[(size_t)y * W + x] = val;
    }
}

// Two 4-vertex polylines inside the ellipse (scaled by 0.6) and one in each
// background band above and below it; 5 px wide.
void scribbles2(uint8_t *seeds, int W, int H, const Ellipse2 &E, Rng &r) {
  const double rad = 2.5;
  for (int p = 0; p < 2; ++p) {
    double px = 0, py = 0;
    for (int k = 0; k < 4; ++k) {
      const double th = r.uni(0, 2 * M_PI), rho = std::sqrt(r.uni());
      const double u = 0.6 * E.a * rho * std::cos(th), v = 0.6 * E.b * rho * std::sin(th);
      const double c = std::cos(E.ang), s = std::sin(E.ang);
      const double x = E.cx + u * c - v * s, y = E.cy + u * s + v * c;
      if (k) stroke2(seeds, W, H, px, py, x, y, rad, 1);
      px = x;
      py = y;
    }
  }
  const double m = 3 + 0.01 * std::min(W, H);
  const double top0 = m, top1 = E.cy - E.b - m;
  const double bot0 = E.cy + E.b + m, bot1 = H - 1 - m;
  const double bands[2][2] = {{top0, top1}, {bot0, bot1}};
  for (int p = 0; p < 2; ++p) {
    const double y0 = bands[p][0], y1 = std::max(bands[p][0], bands[p][1]);
    double px = 0, py = 0;
    for (int k = 0; k < 4; ++k) {
      const double x = r.uni(m, W - 1 - m), y = r.uni(y0, y1);
      if (k) stroke2(seeds, W, H, px, py, x, y, rad, 2);
      px = x;
      py = y;
    }
  }
}

Ellipse2 base_ellipse(int W, int H) {
  const double s = std::min(W, H) / 512.0;
  return Ellipse2{W / 2.0, H / 2.0, 150 * s, 100 * s, 0.0};
}

}  // namespace

extern "C" int jfc_synth_image(uint8_t *rgb, uint8_t *seeds, int W, int H, int nblobs,
                               double noise, uint64_t seed) {
  if (!rgb || !seeds || W < 16 || H < 16 || nblobs < 0) return JFC_ERR_INVALID;
  Rng r(seed);
  const Ellipse2 E = base_ellipse(W, H);
  const std::vector<Blob> blobs = make_blobs(r, W, H, nblobs);
  render_frame(rgb, W, H, E, blobs, noise, mix64(seed ^ 0x5eedULL));
  std::memset(seeds, 0, (size_t)W * H);
  scribbles2(seeds, W, H, E, r);
  return JFC_OK;
}

extern "C" int jfc_synth_video(uint8_t *rgb, uint8_t *seeds, int W, int H, int T,
                               int nblobs, double noise, uint64_t seed) {
  if (!rgb || !seeds || W < 16 || H < 16 || T < 2 || nblobs < 0) return JFC_ERR_INVALID;
  Rng r(seed);
  const std::vector<Blob> blobs = make_blobs(r, W, H, nblobs);
  const double phi = r.uni(0, 2 * M_PI);
  const size_t frame = (size_t)W * H;
  std::memset(seeds, 0, frame * T);
  for (int t = 0; t < T; ++t) {
    Ellipse2 E = base_ellipse(W, H);
    E.cx += 4.0 * t * std::cos(phi);
    E.cy += 4.0 * t * std::sin(phi);
    E.ang = t * M_PI / 180.0;
    render_frame(rgb + frame * 3 * t, W, H, E, blobs, noise,
                 mix64(seed ^ mix64(0x7000ULL + (uint64_t)t)));
    if (t == 0 || t == T - 1) {
      Rng rs(mix64(seed + 31ULL * (uint64_t)(t + 1)));
      scribbles2(seeds + frame * t, W, H, E, rs);
    }
  }
  return JFC_OK;
}

namespace {

struct Ellipsoid {
  double c[3], r[3];
  int val;
  bool inside(double x, double y, double z, double grow) const {
    const double a = (x - c[0]) / (r[0] + grow), b = (y - c[1]) / (r[1] + grow),
                 d = (z - c[2]) / (r[2] + grow);
    return a * a + b * b + d * d <= 1.0;
  }
};

void stroke3(uint8_t *seeds, int nx, int ny, int nz, const double *A, const double *B,
             double radius, uint8_t val) {
  int lo[3], hi[3];
  const int n[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::max(0, (int)std::floor(std::min(A[k], B[k]) - radius - 1));
    hi[k] = std::min(n[k] - 1, (int)std::ceil(std::max(A[k], B[k]) + radius + 1));
  }
  const double v[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double L2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const double p[3] = {(double)x, (double)y, (double)z};
        double t = L2 > 0 ? ((p[0] - A[0]) * v[0] + (p[1] - A[1]) * v[1] + (p[2] - A[2]) * v[2]) / L2 : 0;
        t = std::min(1.0, std::max(0.0, t));
        double d2 = 0;
        for (int k = 0; k < 3; ++k) {
          const double q = p[k] - (A[k] + t * v[k]);
          d2 += q * q;
        }
        if (d2 <= radius * radius) seeds[x + (size_t)nx * (y + (size_t)ny * z)] = val;
      }
}

}  // namespace

extern "C" int jfc_synth_volume(uint16_t *vol, uint8_t *seeds, int nx, int ny, int nz,
                                int nell, uint64_t seed) {
  if (!vol || !seeds || nx < 16 || ny < 16 || nz < 16 || nell < 1) return JFC_ERR_INVALID;
  Rng r(seed);
  std::vector<Ellipsoid> E(nell);
  int largest = 0;
  double best = -1;
  for (int i = 0; i < nell; ++i) {
    const int n[3] = {nx, ny, nz};
    for (int k = 0; k < 3; ++k) {
      E[i].r[k] = r.uni(20, 120) * std::min({nx, ny, nz}) / 512.0;
      E[i].r[k] = std::max(E[i].r[k], 6.0);
      E[i].c[k] = r.uni(0.15 * n[k], 0.85 * n[k]);
    }
    E[i].val = (int)r.uni(1200, 3000);
    const double v = E[i].r[0] * E[i].r[1] * E[i].r[2];
    if (v > best) { best = v; largest = i; }
  }
  const size_t N = (size_t)nx * ny * nz;
  std::vector<float> a(N), b(N);
  parallel_rows((int64_t)ny * nz, [&](int64_t row) {
    const int y = (int)(row % ny), z = (int)(row / ny);
    for (int x = 0; x < nx; ++x) {
      int val = 400;
      for (const Ellipsoid &e : E)
        if (e.inside(x, y, z, 0)) val = e.val;
      a[x + (size_t)nx * row] = (float)val;
    }
  });
  // separable Gaussian smoothing, sigma 2, radius 8, clamp-to-edge
  const int R = 8;
  float w[2 * R + 1];
  double ws = 0;
  for (int i = -R; i <= R; ++i) ws += std::exp(-(double)i * i / 8.0);
  for (int i = -R; i <= R; ++i) w[i + R] = (float)(std::exp(-(double)i * i / 8.0) / ws);
  const int n3[3] = {nx, ny, nz};
  const size_t strides[3] = {1, (size_t)nx, (size_t)nx * ny};
  for (int axis = 0; axis < 3; ++axis) {
    const std::vector<float> &src = axis == 1 ? b : a;
    std::vector<float> &dst = axis == 1 ? a : b;
    parallel_rows((int64_t)ny * nz, [&](int64_t row) {
      const int y = (int)(row % ny), z = (int)(row / ny);
      for (int x = 0; x < nx; ++x) {
        const int p[3] = {x, y, z};
        float s = 0.f;
        for (int i = -R; i <= R; ++i) {
          int q = p[axis] + i;
          q = std::min(n3[axis] - 1, std::max(0, q));
          const size_t idx = x + (size_t)nx * row + (size_t)(q - p[axis]) * strides[axis];
          s += w[i + R] * src[idx];
        }
        dst[x + (size_t)nx * row] = s;
      }
    });
  }
  // after three passes the result is in b
  const uint64_t key = mix64(seed ^ 0x7001ULL);
  parallel_rows((int64_t)ny * nz, [&](int64_t row) {
    for (int x = 0; x < nx; ++x) {
      const size_t i = x + (size_t)nx * row;
      double v = b[i] + 80.0 * gauss(key, i);
      v = std::floor(v + 0.5);
      v = std::min(65535.0, std::max(0.0, v));
      vol[i] = (uint16_t)v;
    }
  });
  std::memset(seeds, 0, N);
  const Ellipsoid &L = E[largest];
  for (int s = 0; s < 3; ++s) {
    double P[2][3];
    for (int e = 0; e < 2; ++e) {
      // uniform in the unit ball, scaled by half the radii
      double u[3], n2;
      do {
        for (int k = 0; k < 3; ++k) u[k] = r.uni(-1, 1);
        n2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      } while (n2 > 1.0);
      for (int k = 0; k < 3; ++k) {
        P[e][k] = L.c[k] + 0.5 * L.r[k] * u[k];
        P[e][k] = std::min((double)n3[k] - 1, std::max(0.0, P[e][k]));
      }
    }
    stroke3(seeds, nx, ny, nz, P[0], P[1], 3.5, 1);
  }
  const double scale = std::min({nx, ny, nz}) / 512.0;
  const double grow = std::max(3.0, 12.0 * scale), reach = std::max(6.0, 60.0 * scale);
  for (int s = 0, tries = 0; s < 3 && tries < 100000; ++tries) {
    double P[2][3];
    for (int k = 0; k < 3; ++k) P[0][k] = r.uni(4, n3[k] - 5);
    for (int k = 0; k < 3; ++k)
      P[1][k] = std::min(n3[k] - 5.0, std::max(4.0, P[0][k] + r.uni(-reach, reach)));
    bool clear = true;
    for (int q = 0; q <= 32 && clear; ++q) {
      const double t = q / 32.0;
      const double x = P[0][0] + t * (P[1][0] - P[0][0]), y = P[0][1] + t * (P[1][1] - P[0][1]),
                   z = P[0][2] + t * (P[1][2] - P[0][2]);
      for (const Ellipsoid &e : E)
        if (e.inside(x, y, z, grow)) { clear = false; break; }
    }
    if (!clear) continue;
    stroke3(seeds, nx, ny, nz, P[0], P[1], 3.5, 2);
    ++s;
  }
  return JFC_OK;
}
