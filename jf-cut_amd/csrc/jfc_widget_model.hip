// jfc_widget_model.hip -- the widget's own data terms on the device: the
// "Model" button of QGCWidget (slotModelClicked,
// src/QVisualizer/Graph Cut/QGCWidget.cpp:2560-2606) with lazy snapping
// (k-means unary, inverse-distance pairwise) or BK (histogram unary,
// Gaussian pairwise).
//
// Reference counterparts (src/QVisualizer/cl/GraphCut/):
//   k_wm_seed_list   graphcut_count_cut / graphcut_scatter_cut   graphcut_count_block.cl:67-104
//   k_wm_levels      graphcut_cut_histogram, _distribution       graphcut_histogram.cl:19-60
//   k_wm_assign      graphcut_kmeans                             graphcut_kmeans.cl:21-159
//   k_wm_unary<lazy> graphcut_compute_excessflow_lazy / _bk      graphcut_compute_excessflow.cl:19-139
//   k_wm_pairwise    graphcut_compute_capacity_lazy / _bk        graphcut_compute_capacity.cl:21-254
//
// Float semantics.  Every float operation is the reference's, in its order,
// IEEE-rounded: no contraction into fused multiply-adds (the pragma below),
// correctly rounded division and square root, and exp / log evaluated in
// double and rounded once to float.  A CPU restatement with the same
// operation order therefore reproduces every integer output.
//
// Readings fixed here where the reference is undefined or racy:
//   - the k-means initial centroid of a seeded level is the pixel with the
//     highest node index at that level (graphcut_cut_distribution lets any
//     writer win);
//   - k-means centroid sums are exact (per-byte counts, summed in double in
//     byte order) before the division, instead of float partial sums in
//     tile order; a cluster with no member becomes 0 as in QGCW:2085-2096;
//   - the BK unary clamps its histogram level to 0..255 (the reference reads
//     past the table when the weighted dot product exceeds 1), a NaN unary
//     (0 * inf, inf - inf) becomes 0, and unaries saturate at +-2^30, the
//     library's capacity bound.
#pragma clang fp contract(off)

#include "jfcut.h"
#include "jfc_device.h"

#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace jfc {
int model_fail(int code, const std::string &msg);
int graph_instance(jfc_graph *g, int inst, InstDesc *I, int **cs, int **ct, int **c0,
                   long long *stride);
void graph_mark_inputs(jfc_graph *g, int inst);
void graph_inputs_failed(jfc_graph *g, int inst);
hipStream_t graph_stream(jfc_graph *g);
int graph_stage(jfc_graph *g, size_t bytes, void **p);
int graph_scratch(jfc_graph *g, size_t bytes, void **p);
int graph_work_plane(jfc_graph *g, int inst, int **plane, long long *n);
}  // namespace jfc

using namespace jfc;

namespace {

constexpr int kMaxFlow = 32767;  // QGCSetting::cl_max_flow (QGCSetting.cpp:36)
constexpr int kLevels = 256;     // QGCSetting::sizeCutHistogram (QGCSetting.cpp:20)
constexpr int kMaxK = 32;        // spinBoxK maximum (QGCPanel.ui)
constexpr int kSat = 1 << 30;    // library capacity bound

struct W4 {
  float v[4];
};

// read_imagef of a CL_UNORM_INT8 texel: byte / 255
__device__ __host__ __forceinline__ float unorm8(unsigned b) { return (float)b / 255.0f; }

__device__ __forceinline__ void texel(const uchar4 *img, long long v, unsigned t[4]) {
  const uchar4 p = img[v];
  t[0] = p.x;
  t[1] = p.y;
  t[2] = p.z;
  t[3] = p.w;
}

// OpenCL dot(a, b), left to right
__device__ __forceinline__ float dot4(const float a[4], const W4 &w) {
  return ((a[0] * w.v[0] + a[1] * w.v[1]) + a[2] * w.v[2]) + a[3] * w.v[3];
}

// OpenCL distance(c, p) = length(c - p)
__device__ __forceinline__ float dist4(const float *c, const float p[4]) {
  const float d0 = c[0] - p[0], d1 = c[1] - p[1], d2 = c[2] - p[2], d3 = c[3] - p[3];
  return sqrtf(((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3);
}

// histogram level of a seeded pixel: clamp(dot, 0, 1) * (size - 1) + 0.5,
// truncated (graphcut_histogram.cl:33-35)
__device__ __forceinline__ int level_clamped(float data) {
  data = fminf(fmaxf(data, 0.0f), 1.0f);
  return (int)(data * (float)(kLevels - 1) + 0.5f);
}

// float unary -> int: round half away from zero, NaN -> 0, saturate
__device__ __forceinline__ int unary_int(float v) {
  if (v != v) return 0;
  const float r = roundf(v);
  if (r >= (float)kSat) return kSat;
  if (r <= -(float)kSat) return -kSat;
  return (int)r;
}

__global__ void k_wm_seed_list(const unsigned char *seeds, long long n, int *list, int *count) {
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n;
       base += (long long)gridDim.x * blockDim.x) {
    const long long i = base + threadIdx.x;
    wave_append(i < n && seeds[i] != 0, (int)i, list, count, nullptr, nullptr);
  }
}

// Per seeded pixel of one side: the level histogram (graphcut_cut_histogram)
// and the highest node index + 1 per level (graphcut_cut_distribution).
__global__ void k_wm_levels(const uchar4 *img, const int *list, const int *count, W4 w,
                            unsigned *hist, unsigned *rep) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = list[i];
    unsigned t[4];
    texel(img, v, t);
    float f[4];
    for (int c = 0; c < 4; ++c) f[c] = unorm8(t[c]);
    const int lv = level_clamped(dot4(f, w));
    atomicAdd(&hist[lv], 1u);
    atomicMax(&rep[lv], (unsigned)v + 1u);
  }
}

// One Lloyd assignment over the seeded pixels of a side: the nearest
// centroid (first on ties, graphcut_kmeans.cl:37-47) and, per cluster and
// channel, a count per byte value, from which the host forms exact sums.
__global__ void k_wm_assign(const uchar4 *img, const int *list, const int *count, W4 w,
                            const float *cent, int k, unsigned *acc) {
  __shared__ float s_c[kMaxK * 4];
  for (int i = threadIdx.x; i < k * 4; i += blockDim.x) s_c[i] = cent[i];
  __syncthreads();
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = list[i];
    unsigned t[4];
    texel(img, v, t);
    float p[4];
    for (int c = 0; c < 4; ++c) p[c] = unorm8(t[c]) * w.v[c];
    float best = 3.402823466e38f;
    int idx = 0;
    for (int j = 0; j < k; ++j) {
      const float d = dist4(s_c + 4 * j, p);
      if (d < best) {
        best = d;
        idx = j;
      }
    }
    for (int c = 0; c < 4; ++c) atomicAdd(&acc[(idx * 4 + c) * kLevels + t[c]], 1u);
  }
}

// Unary terms, one row of an instance per loop step.  LAZY: distance to the
// nearest centroid of each side (graphcut_compute_excessflow.cl:47-77,
// 109-139); BK: negative log-likelihood from the seeded level histograms
// (graphcut_compute_excessflow.cl:19-45, 79-107).  Seeds: +-cl_max_flow,
// foreground first.
template <bool LAZY>
__global__ __launch_bounds__(256) void k_wm_unary(
    const uchar4 *img, const unsigned char *fg, const unsigned char *bg, long long n, W4 w,
    const float *cent, int k, const unsigned *hist, unsigned sum_f, unsigned sum_b, int shift,
    float lambda, int nx, int px, int *cs, int *ct) {
  __shared__ float s_c[2 * kMaxK * 4];
  __shared__ unsigned s_h[2 * kLevels];
  if (LAZY)
    for (int i = threadIdx.x; i < 2 * k * 4; i += blockDim.x) s_c[i] = cent[i];
  else
    for (int i = threadIdx.x; i < 2 * kLevels; i += blockDim.x) s_h[i] = hist[i];
  __syncthreads();
  const float gain = (float)(kMaxFlow >> shift) * lambda;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n;
       v += (long long)gridDim.x * blockDim.x) {
    int ef;
    if (fg[v]) {
      ef = kMaxFlow;
    } else if (bg[v]) {
      ef = -kMaxFlow;
    } else {
      unsigned t[4];
      texel(img, v, t);
      float f[4];
      for (int c = 0; c < 4; ++c) f[c] = unorm8(t[c]);
      float val;
      if (LAZY) {
        float p[4];
        for (int c = 0; c < 4; ++c) p[c] = f[c] * w.v[c];
        float dx = 3.402823466e38f, dy = 3.402823466e38f;
        for (int j = 0; j < k; ++j) {
          dx = fminf(dx, dist4(s_c + 4 * j, p));
          dy = fminf(dy, dist4(s_c + 4 * (k + j), p));
        }
        const float r = (dx + dy == 0.0f) ? 0.0f : (dy - dx) / (dx + dy);
        val = gain * r;
      } else {
        int lv = (int)(dot4(f, w) * (float)(kLevels - 1) + 0.5f);
        lv = lv < 0 ? 0 : (lv > kLevels - 1 ? kLevels - 1 : lv);
        const float qf = (float)s_h[lv] / (float)sum_f;
        const float qb = (float)s_h[kLevels + lv] / (float)sum_b;
        const float df = -(float)log((double)qf), db = -(float)log((double)qb);
        val = gain * (db - df);
      }
      ef = unary_int(val);
    }
    // the planes store row r at r * px
    const long long sv = (v / nx) * px + v % nx;
    cs[sv] = ef > 0 ? ef : 0;
    ct[sv] = ef < 0 ? -ef : 0;
  }
}

// convert_int4(read_imagef(.) * weight * SHRT_MAX): truncation toward zero
__device__ __forceinline__ void scaled_texel(const uchar4 *img, long long v, const W4 &w,
                                             int d[4]) {
  unsigned t[4];
  texel(img, v, t);
  for (int c = 0; c < 4; ++c) d[c] = (int)((unorm8(t[c]) * w.v[c]) * (float)SHRT_MAX);
}

// One arc of graphcut_compute_capacity_lazy / _bk: the squared distance in
// unsigned arithmetic, then top / (1 + dist) or top * exp(-t^2 / 2), + 0.5,
// truncated; top = SHRT_MAX in 2D, CHAR_MAX in 3D.
template <bool LAZY, bool D3>
__device__ __forceinline__ int wm_arc(const int a[4], const int b[4], float scale) {
  unsigned s = 0;
  for (int c = 0; c < 4; ++c) {
    const unsigned d = (unsigned)(a[c] > b[c] ? a[c] - b[c] : b[c] - a[c]);
    s += d * d;
  }
  const float top = D3 ? 127.0f : (float)SHRT_MAX;
  const float x = sqrtf((float)s) * scale;
  if (LAZY) return (int)(top / (1.0f + x) + 0.5f);
  const float e = (float)exp((double)((-0.5f * x) * x));
  return (int)(top * e + 0.5f);
}

// Planes of one instance; plane d of node v is the arc toward v + dir(d)
// (0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z); border arcs are 0.  The reference's
// c1 / c2 pairs are the same values (c1 = (+x, +y[, +z]), c2 = (-y, -x) in 2D
// and (-z, -y, -x) in 3D).
template <bool LAZY, bool D3>
__global__ __launch_bounds__(256) void k_wm_pairwise(InstDesc I, const uchar4 *img, W4 w,
                                                     float scale, int *c0, long long stride) {
  const long long rows = (long long)I.ny * I.nz;
  for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
    const int y = (int)(row % I.ny), z = (int)(row / I.ny);
    for (int x = threadIdx.x; x < I.nx; x += blockDim.x) {
      const long long v = x + (long long)I.nx * row;
      int a[4];
      scaled_texel(img, v, w, a);
      for (int d = 0; d < (D3 ? 6 : 4); ++d) {
        int dx, dy, dz;
        dir_of<D3>(d, dx, dy, dz);
        const int X = x + dx, Y = y + dy, Z = z + dz;
        int cap = 0;
        if (X >= 0 && X < I.nx && Y >= 0 && Y < I.ny && Z >= 0 && Z < I.nz) {
          int b[4];
          scaled_texel(img, X + (long long)I.nx * (Y + (long long)I.ny * Z), w, b);
          cap = wm_arc<LAZY, D3>(a, b, scale);
        }
        c0[d * stride + I.node_off + x + (long long)I.px * row] = cap;
      }
    }
  }
}

int grid1(long long n) {
  long long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// glm::normalize(vec4): v * (1 / sqrt(dot(v, v))), dot as (xx + yy) + (zz + ww)
W4 normalize_weight(const float in[4]) {
  const float s = (in[0] * in[0] + in[1] * in[1]) + (in[2] * in[2] + in[3] * in[3]);
  const float inv = 1.0f / std::sqrt(s);
  W4 w;
  for (int c = 0; c < 4; ++c) w.v[c] = in[c] * inv;
  return w;
}

}  // namespace

extern "C" void jfc_default_widget_model_params(jfc_widget_model_params *p) {
  if (!p) return;
  p->lazy = 1;
  p->weight[0] = 0.2126f;
  p->weight[1] = 0.7152f;
  p->weight[2] = 0.0722f;
  p->weight[3] = 0.0f;
  p->lazy_neighbor = 1.0f;
  p->lazy_terminal = 1.0f;
  p->bk_neighbor = 0.002f;
  p->bk_terminal = 0.0f;
  p->k = 8;
  p->kmeans_iterations = 32;
}

#define WM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      rc = model_fail(JFC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
      goto done;                                                                      \
    }                                                                                 \
  } while (0)

extern "C" int jfc_model_widget(jfc_graph *g, int inst, const uint8_t *rgba, const uint8_t *fg,
                                const uint8_t *bg, int on_device,
                                const jfc_widget_model_params *prm,
                                jfc_widget_model_stats *stats) {
  InstDesc I;
  int *cs = nullptr, *ct = nullptr, *c0 = nullptr;
  long long stride = 0;
  int rc = graph_instance(g, inst, &I, &cs, &ct, &c0, &stride);
  if (rc) return rc;
  if (!rgba || !fg || !bg || !prm) return model_fail(JFC_ERR_INVALID, "null argument");
  const bool d3 = I.conn == 6;
  if (!(I.conn == 4 || I.conn == 6))
    return model_fail(JFC_ERR_INVALID,
                      "the widget's data terms are 4-connected in 2D and 6-connected in 3D");
  if (prm->k < 2 || prm->k > kMaxK || prm->kmeans_iterations < 1 || prm->kmeans_iterations > 256)
    return model_fail(JFC_ERR_INVALID, "k must be 2..32 and kmeans_iterations 1..256");
  for (int c = 0; c < 4; ++c)
    if (!(prm->weight[c] >= 0.0f) || !std::isfinite(prm->weight[c]))
      return model_fail(JFC_ERR_INVALID, "weights must be finite and >= 0");
  if (!(prm->weight[0] + prm->weight[1] + prm->weight[2] + prm->weight[3] > 0.0f))
    return model_fail(JFC_ERR_INVALID, "at least one weight must be positive");
  if (prm->lazy ? !(prm->lazy_neighbor >= 0.0f && prm->lazy_terminal >= 0.0f &&
                    std::isfinite(prm->lazy_neighbor) && std::isfinite(prm->lazy_terminal))
                : !(prm->bk_neighbor > 0.0f && prm->bk_terminal >= 0.0f &&
                    std::isfinite(prm->bk_neighbor) && std::isfinite(prm->bk_terminal)))
    return model_fail(JFC_ERR_INVALID, "invalid lambda / sigma");
  const long long n = (long long)I.nx * I.ny * I.nz;
  if (n >= (1LL << 31) - 1) return model_fail(JFC_ERR_INVALID, "instance too large for the model path");

  // Everything runs on the graph's stream.  Host inputs cross in one copy
  // each through the graph's staging buffer; the small scratch (seed count,
  // histograms, level representatives, k-means sums, centroids) is the
  // graph's data-term block; the seed list borrows the instance's slice of
  // the excess plane, which every solve rewrites from the inputs.
  const hipStream_t st = graph_stream(g);
  const W4 w = normalize_weight(prm->weight);
  const int k = prm->k;
  const uchar4 *img = reinterpret_cast<const uchar4 *>(rgba);
  const unsigned char *dfg = fg, *dbg = bg;
  int *d_list = nullptr, *d_cnt = nullptr;
  unsigned *d_hist = nullptr, *d_rep = nullptr, *d_acc = nullptr;
  float *d_cent = nullptr;
  long long list_cap = 0;
  bool written = false;
  jfc_widget_model_stats st_local;
  jfc_widget_model_stats &ST = stats ? *stats : st_local;
  std::vector<float> cent_used(2 * k * 4, 0.0f);
  std::vector<unsigned> hist(2 * kLevels, 0);
  memset(&ST, 0, sizeof(ST));

  if (!on_device) {
    void *stage = nullptr;
    if ((rc = graph_stage(g, (size_t)n * 6 + 256, &stage))) return rc;
    unsigned char *sp = (unsigned char *)stage;
    WM_HIP(hipMemcpyAsync(sp, rgba, n * sizeof(uchar4), hipMemcpyHostToDevice, st));
    WM_HIP(hipMemcpyAsync(sp + n * 4, fg, n, hipMemcpyHostToDevice, st));
    WM_HIP(hipMemcpyAsync(sp + n * 5, bg, n, hipMemcpyHostToDevice, st));
    img = reinterpret_cast<const uchar4 *>(sp);
    dfg = sp + n * 4;
    dbg = sp + n * 5;
  }
  {
    const size_t o_cnt = 0, o_hist = 256, o_rep = o_hist + 2 * kLevels * 4,
                 o_acc = o_rep + kLevels * 4, o_cent = o_acc + (size_t)kMaxK * 4 * kLevels * 4,
                 total = o_cent + 2 * kMaxK * 4 * 4;
    void *sc = nullptr;
    if ((rc = graph_scratch(g, total, &sc))) return rc;
    unsigned char *b = (unsigned char *)sc;
    d_cnt = (int *)(b + o_cnt);
    d_hist = (unsigned *)(b + o_hist);
    d_rep = (unsigned *)(b + o_rep);
    d_acc = (unsigned *)(b + o_acc);
    d_cent = (float *)(b + o_cent);
    if ((rc = graph_work_plane(g, inst, &d_list, &list_cap))) return rc;
  }
  WM_HIP(hipMemsetAsync(d_hist, 0, 2 * kLevels * sizeof(unsigned), st));

  for (int side = 0; side < 2; ++side) {
    int cnt = 0;
    WM_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int), st));
    k_wm_seed_list<<<grid1(n), 256, 0, st>>>(side ? dbg : dfg, n, d_list, d_cnt);
    WM_HIP(hipGetLastError());
    WM_HIP(hipMemcpyAsync(&cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    WM_HIP(hipStreamSynchronize(st));
    ST.seeds[side] = cnt;
    if (cnt == 0) {
      // computeBFS / kmeans "no labels." (QGCW:2015-2016)
      rc = model_fail(JFC_ERR_NO_SEEDS, side ? "no background seeds" : "no foreground seeds");
      goto done;
    }
    WM_HIP(hipMemsetAsync(d_rep, 0, kLevels * sizeof(unsigned), st));
    k_wm_levels<<<grid1(cnt), 256, 0, st>>>(img, d_list, d_cnt, w, d_hist + side * kLevels, d_rep);
    WM_HIP(hipGetLastError());
    if (!prm->lazy) continue;

    // initial centroids: the representatives of the seeded levels, picked at
    // (count - 1) * i / (k - 1) (QGCW:2011-2039), scaled by weight / 255
    std::vector<unsigned> rep(kLevels);
    WM_HIP(hipMemcpyAsync(rep.data(), d_rep, kLevels * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    WM_HIP(hipStreamSynchronize(st));
    std::vector<unsigned> dist;
    for (int l = 0; l < kLevels; ++l)
      if (rep[l]) dist.push_back(rep[l]);
    const unsigned dsize = (unsigned)dist.size() - 1u;
    float scale[4];
    for (int c = 0; c < 4; ++c) scale[c] = w.v[c] / 255.0f;
    std::vector<float> cent(k * 4), prev(k * 4);
    std::vector<uchar4> picked(k);
    for (int i = 0; i < k; ++i) {
      const long long idx = (long long)dist[dsize * (unsigned)i / (unsigned)(k - 1)] - 1;
      WM_HIP(hipMemcpyAsync(&picked[i], img + idx, sizeof(uchar4), hipMemcpyDeviceToHost, st));
    }
    WM_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < k; ++i) {
      const uchar4 t = picked[i];
      const unsigned tb[4] = {t.x, t.y, t.z, t.w};
      for (int c = 0; c < 4; ++c) cent[i * 4 + c] = (float)tb[c] * scale[c];
    }
    // weighted feature value of each byte per channel: unorm8(b) * w
    float val[4][kLevels];
    for (int c = 0; c < 4; ++c)
      for (int b = 0; b < kLevels; ++b) val[c][b] = unorm8(b) * w.v[c];
    // Lloyd iterations (QGCW:2073-2102): the centroids of the last
    // assignment are the ones the unary uses
    std::vector<unsigned> acc((size_t)k * 4 * kLevels);
    int it = 0;
    for (; it < prm->kmeans_iterations;) {
      prev = cent;
      WM_HIP(hipMemcpyAsync(d_cent, prev.data(), k * 4 * sizeof(float), hipMemcpyHostToDevice, st));
      WM_HIP(hipMemsetAsync(d_acc, 0, acc.size() * sizeof(unsigned), st));
      k_wm_assign<<<std::min(grid1(cnt), 1024), 256, 0, st>>>(img, d_list, d_cnt, w, d_cent, k, d_acc);
      WM_HIP(hipGetLastError());
      WM_HIP(hipMemcpyAsync(acc.data(), d_acc, acc.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
      WM_HIP(hipStreamSynchronize(st));
      ++it;
      float err = 0.0f;
      for (int i = 0; i < k; ++i) {
        long long members = 0;
        for (int b = 0; b < kLevels; ++b) members += acc[(i * 4) * kLevels + b];
        for (int c = 0; c < 4; ++c) {
          if (!members) {
            cent[i * 4 + c] = 0.0f;
            continue;
          }
          double sum = 0.0;
          for (int b = 0; b < kLevels; ++b) {
            const unsigned m = acc[(i * 4 + c) * kLevels + b];
            if (m) sum += (double)m * (double)val[c][b];
          }
          cent[i * 4 + c] = (float)(sum / (double)members);
        }
        // glm::length(centroid - previous)
        float d[4];
        for (int c = 0; c < 4; ++c) d[c] = cent[i * 4 + c] - prev[i * 4 + c];
        const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + (d[2] * d[2] + d[3] * d[3]));
        err = std::max(err, len);
      }
      if (err < 1.0f / 1024) break;
    }
    ST.iterations[side] = it;
    for (int i = 0; i < k * 4; ++i) {
      cent_used[side * k * 4 + i] = prev[i];
      ST.centroids[side][i / 4][i % 4] = prev[i];
    }
  }
  {
    WM_HIP(hipMemcpyAsync(hist.data(), d_hist, hist.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    WM_HIP(hipStreamSynchronize(st));
    for (int side = 0; side < 2; ++side)
      for (int l = 0; l < kLevels; ++l) ST.hist_sum[side] += hist[side * kLevels + l];
    const int shift = d3 ? 3 : 2;
    written = true;
    if (prm->lazy) {
      WM_HIP(hipMemcpyAsync(d_cent, cent_used.data(), cent_used.size() * sizeof(float),
                            hipMemcpyHostToDevice, st));
      k_wm_unary<true><<<grid1(n), 256, 0, st>>>(img, dfg, dbg, n, w, d_cent, k, nullptr, 0, 0,
                                                 shift, prm->lazy_terminal, I.nx, I.px,
                                                 cs + I.node_off, ct + I.node_off);
    } else {
      k_wm_unary<false><<<grid1(n), 256, 0, st>>>(img, dfg, dbg, n, w, nullptr, 0, d_hist,
                                                  ST.hist_sum[0], ST.hist_sum[1], shift,
                                                  prm->bk_terminal, I.nx, I.px, cs + I.node_off,
                                                  ct + I.node_off);
    }
    WM_HIP(hipGetLastError());
    const int grid = (int)std::min<long long>((long long)I.ny * I.nz, 65536);
    if (prm->lazy) {
      // lambda / SHRT_MAX (graphcut_compute_capacity.cl:82)
      const float scale = prm->lazy_neighbor / (float)SHRT_MAX;
      if (d3) k_wm_pairwise<true, true><<<grid, 256, 0, st>>>(I, img, w, scale, c0, stride);
      else k_wm_pairwise<true, false><<<grid, 256, 0, st>>>(I, img, w, scale, c0, stride);
    } else {
      // 1 / (SHRT_MAX * sigma) (graphcut_compute_capacity.cl:29)
      const float scale = 1.0f / ((float)SHRT_MAX * prm->bk_neighbor);
      if (d3) k_wm_pairwise<false, true><<<grid, 256, 0, st>>>(I, img, w, scale, c0, stride);
      else k_wm_pairwise<false, false><<<grid, 256, 0, st>>>(I, img, w, scale, c0, stride);
    }
    WM_HIP(hipGetLastError());
    WM_HIP(hipStreamSynchronize(st));
    graph_mark_inputs(g, inst);
  }
done:
  if (rc && written) graph_inputs_failed(g, inst);
  (void)list_cap;
  return rc;
}
