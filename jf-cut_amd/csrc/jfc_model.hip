// jfc_model.hip -- data term on the device: integer k-means per seed side,
// unary and pairwise capacities (QGCWidget::slotModelClicked,
// src/QVisualizer/Graph Cut/QGCWidget.cpp:2560-2606).
//
// Reference counterparts (src/QVisualizer/cl/GraphCut/):
//   k_seed_list   graphcut_count_cut / graphcut_scatter_cut  graphcut_count_block.cl:67-104
//   k_key_hist    graphcut_cut_distribution                  graphcut_histogram.cl:41-60
//   k_km_assign   graphcut_kmeans                            graphcut_kmeans.cl:21-89
//   k_unary       graphcut_compute_excessflow_lazy           graphcut_compute_excessflow.cl:47-77
//   k_pairwise    graphcut_compute_capacity_bk               graphcut_compute_capacity.cl:21-69
//
// The reference runs k-means in float32 with host-side float division; here
// every quantity is an integer so CPU and GPU agree bit for bit: centroids
// are fixed point (value * 2^8) rounded from exact int64 sums, distances are
// exact int64, unary rounding is integer, and the pairwise exp() is replaced
// by a table of integer thresholds on the squared difference, computed once
// on the host.
#include "jfcut.h"
#include "jfc_device.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace jfc {
int model_fail(int code, const std::string &msg);
}

using namespace jfc;

namespace {

constexpr int KM_FRAC = 8;  // fixed-point bits of centroids
constexpr int KMAX_CLUSTERS = 16;
constexpr int TMAX = 256;   // max thresholds per table (lambda < 255.5)

__device__ __forceinline__ int feat(const void *img, int bpc, long long i) {
  return bpc == 1 ? (int)((const unsigned char *)img)[i]
                  : (int)((const unsigned short *)img)[i];
}

__global__ void k_seed_list(const unsigned char *seeds, long long n, int side,
                            int *list, int *count) {
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n;
       base += (long long)gridDim.x * blockDim.x) {
    const long long i = base + threadIdx.x;
    const bool p = i < n && seeds[i] == side;
    wave_append(p, (int)i, list, count, nullptr, nullptr);
  }
}

__global__ void k_key_hist(const void *img, int bpc, int C, const int *list,
                           const int *count, long long *kcnt, long long *ksum) {
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const long long v = list[i];
    int key = 0;
    int x[3];
    for (int c = 0; c < C; ++c) {
      x[c] = feat(img, bpc, v * C + c);
      key += x[c];
    }
    if (bpc == 2) key = x[0] >> 4;
    atomicAdd((unsigned long long *)&kcnt[key], 1ull);
    for (int c = 0; c < C; ++c)
      atomicAdd((unsigned long long *)&ksum[(long long)key * C + c], (unsigned long long)x[c]);
  }
}

__device__ __forceinline__ long long km_d2(const int *x, const long long *c, int C) {
  long long d = 0;
  for (int i = 0; i < C; ++i) {
    const long long t = (long long)x[i] * (1 << KM_FRAC) - c[i];
    d += t * t;
  }
  return d;
}

// Assignment to the nearest centroid (lowest index on ties) and per-cluster
// sums, reduced in LDS before one global atomic per cluster entry.
__global__ __launch_bounds__(256) void k_km_assign(const void *img, int bpc, int C,
                                                   const int *list, const int *count,
                                                   const long long *cent, int K,
                                                   long long *acc) {
  __shared__ long long s_c[KMAX_CLUSTERS * 3];
  __shared__ long long s_acc[KMAX_CLUSTERS * 4];
  for (int i = threadIdx.x; i < K * C; i += blockDim.x) s_c[i] = cent[i];
  for (int i = threadIdx.x; i < K * (C + 1); i += blockDim.x) s_acc[i] = 0;
  __syncthreads();
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const long long v = list[i];
    int x[3];
    for (int c = 0; c < C; ++c) x[c] = feat(img, bpc, v * C + c);
    int best = 0;
    long long bd = km_d2(x, s_c, C);
    for (int k = 1; k < K; ++k) {
      const long long d = km_d2(x, s_c + k * C, C);
      if (d < bd) { bd = d; best = k; }
    }
    for (int c = 0; c < C; ++c)
      atomicAdd((unsigned long long *)&s_acc[best * (C + 1) + c], (unsigned long long)x[c]);
    atomicAdd((unsigned long long *)&s_acc[best * (C + 1) + C], 1ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * (C + 1); i += blockDim.x)
    if (s_acc[i]) atomicAdd((unsigned long long *)&acc[i], (unsigned long long)s_acc[i]);
}

// cs = round(scale * d2_bg / den), ct = round(scale * d2_fg / den); seeds get
// (2^30, 0) / (0, 2^30).
// v: the image's (unpadded) node index; the planes store row r at r * px.
__global__ void k_unary(const void *img, int bpc, int C, long long n, int nx, int px,
                        const unsigned char *seeds, const long long *cfg, int Kf,
                        const long long *cbg, int Kb, long long den, long long scale,
                        int *cs, int *ct) {
  __shared__ long long s_f[KMAX_CLUSTERS * 3], s_b[KMAX_CLUSTERS * 3];
  for (int i = threadIdx.x; i < Kf * C; i += blockDim.x) s_f[i] = cfg[i];
  for (int i = threadIdx.x; i < Kb * C; i += blockDim.x) s_b[i] = cbg[i];
  __syncthreads();
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n;
       v += (long long)gridDim.x * blockDim.x) {
    const long long sv = (v / nx) * px + v % nx;
    const int sd = seeds ? seeds[v] : 0;
    if (sd == 1) { cs[sv] = 1 << 30; ct[sv] = 0; continue; }
    if (sd == 2) { cs[sv] = 0; ct[sv] = 1 << 30; continue; }
    int x[3];
    for (int c = 0; c < C; ++c) x[c] = feat(img, bpc, v * C + c);
    long long df = km_d2(x, s_f, C), db = km_d2(x, s_b, C);
    for (int k = 1; k < Kf; ++k) df = min(df, km_d2(x, s_f + k * C, C));
    for (int k = 1; k < Kb; ++k) db = min(db, km_d2(x, s_b + k * C, C));
    cs[sv] = (int)((scale * db + den / 2) / den);
    ct[sv] = (int)((scale * df + den / 2) / den);
  }
}

template <bool D3>
__device__ __forceinline__ bool nb_at(const InstDesc &I, int x, int y, int z, int d,
                                      long long &u) {
  int dx, dy, dz;
  dir_of<D3>(d, dx, dy, dz);
  const int X = x + dx, Y = y + dy, Z = z + dz;
  if (X < 0 || X >= I.nx || Y < 0 || Y >= I.ny || Z < 0 || Z >= I.nz) return false;
  u = X + (long long)I.nx * (Y + (long long)I.ny * Z);
  return true;
}

__device__ __forceinline__ long long sqdiff(const void *img, int bpc, int C, long long a,
                                            long long b) {
  long long D = 0;
  for (int c = 0; c < C; ++c) {
    const long long t = feat(img, bpc, a * C + c) - feat(img, bpc, b * C + c);
    D += t * t;
  }
  return D;
}

// Sum of squared neighbour differences per edge group (0: x/y and
// diagonals, 1: z when it forms its own group), one arc per undirected edge.
template <bool D3>
__global__ __launch_bounds__(256) void k_pair_sums(InstDesc I, const void *img, int bpc,
                                                   int C, int zsep, long long *sums) {
  __shared__ long long buf[4][4];
  long long s0 = 0, p0 = 0, s1 = 0, p1 = 0;
  const long long rows = (long long)I.ny * I.nz;
  for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
    const int y = (int)(row % I.ny), z = (int)(row / I.ny);
    for (int x = threadIdx.x; x < I.nx; x += blockDim.x) {
      const long long v = x + (long long)I.nx * row;
      for (int d = 0; d < I.K; d += 2) {
        long long u;
        if (!nb_at<D3>(I, x, y, z, d, u)) continue;
        const long long D = sqdiff(img, bpc, C, v, u);
        if (D3 && d == 4 && zsep) { s1 += D; p1 += 1; }
        else { s0 += D; p0 += 1; }
      }
    }
  }
  long long vals[4] = {s0, p0, s1, p1};
  const int w = threadIdx.x >> 6;
  for (int j = 0; j < 4; ++j) {
    long long x = vals[j];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane_id() == 0) buf[w][j] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    long long t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += buf[i][threadIdx.x];
    if (t) atomicAdd((unsigned long long *)&sums[threadIdx.x], (unsigned long long)t);
  }
}

// w(D) = #{k : D <= T[k]} with tables 0: x/y, 1: 2D diagonals, 2: z.  Every
// table is non-increasing (thresholds() below), so the count is the length
// of the prefix with D <= T[k], found by binary search.
template <bool D3>
__global__ __launch_bounds__(256) void k_pairwise(InstDesc I, const void *img, int bpc,
                                                  int C, const long long *T,
                                                  const int *nT, int *c0,
                                                  long long stride) {
  __shared__ long long s_T[3 * TMAX];
  __shared__ int s_n[3];
  for (int i = threadIdx.x; i < 3 * TMAX; i += blockDim.x) s_T[i] = T[i];
  if (threadIdx.x < 3) s_n[threadIdx.x] = nT[threadIdx.x];
  __syncthreads();
  const long long rows = (long long)I.ny * I.nz;
  for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
    const int y = (int)(row % I.ny), z = (int)(row / I.ny);
    for (int x = threadIdx.x; x < I.nx; x += blockDim.x) {
      const long long v = x + (long long)I.nx * row;
      for (int d = 0; d < I.K; ++d) {
        long long u;
        int w = 0;
        if (nb_at<D3>(I, x, y, z, d, u)) {
          const long long D = sqdiff(img, bpc, C, v, u);
          const int tb = D3 ? (d >= 4 ? 2 : 0) : (d >= 4 ? 1 : 0);
          const long long *t = s_T + tb * TMAX;
          int hi = s_n[tb];
          while (w < hi) {
            const int mid = (w + hi) >> 1;
            if (D <= t[mid]) w = mid + 1;
            else hi = mid;
          }
        }
        c0[d * stride + I.node_off + x + (long long)I.px * row] = w;
      }
    }
  }
}

int thresholds(double lambda, long long S, long long P, long long *T) {
  int kmax = (int)std::floor(lambda + 0.5);
  if (kmax > TMAX) kmax = TMAX;
  for (int k = 1; k <= kmax; ++k) {
    if (S == 0 || P == 0) { T[k - 1] = INT64_MAX; continue; }
    const double v = (2.0 * (double)S / (double)P) * std::log(lambda / (k - 0.5));
    T[k - 1] = (long long)std::floor(v);
  }
  return kmax;
}

int grid1(long long n) {
  long long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

// Implemented in jfc_api.hip: accessors of the opaque graph for this file.
namespace jfc {
int graph_instance(jfc_graph *g, int inst, InstDesc *I, int **cs, int **ct, int **c0,
                   long long *stride);
void graph_mark_inputs(jfc_graph *g, int inst);
}

#define MODEL_HIP(call)                                                         \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      rc = model_fail(JFC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
      goto done;                                                                \
    }                                                                           \
  } while (0)

// Parameter checks that need no graph.  The unary of a node is
// round(scale * d2 / den) with d2 <= C * (maxval * 2^KM_FRAC)^2 (a feature and
// a centroid both in [0, maxval]) and den = range^2 * 2^(2 KM_FRAC): the
// product scale * d2 must fit int64 and the unary must not exceed the
// terminal-capacity limit, or the solve's int32 arithmetic would not be exact.
static int check_model_params(int channels, int bpc, const jfc_model_params *prm) {
  if (!prm) return model_fail(JFC_ERR_INVALID, "null argument");
  if (!((channels == 1 || channels == 3) && (bpc == 1 || (bpc == 2 && channels == 1))))
    return model_fail(JFC_ERR_INVALID, "image must be 1 or 3 channels of uint8, or 1 channel of uint16");
  if (prm->k_fg < 1 || prm->k_fg > KMAX_CLUSTERS || prm->k_bg < 1 || prm->k_bg > KMAX_CLUSTERS)
    return model_fail(JFC_ERR_INVALID, "cluster count must be 1..16");
  if (prm->range <= 0 || prm->range > (1 << 20) || prm->unary_scale < 0 ||
      !(prm->lambda_xy >= 0) || !(prm->lambda_z >= 0) || prm->lambda_xy > 255 ||
      prm->lambda_z > 255)
    return model_fail(JFC_ERR_INVALID, "invalid model parameters");
  const long double maxval = bpc == 1 ? 255.0L : 65535.0L;
  const long double d2max = (long double)channels * (maxval * (1 << KM_FRAC)) * (maxval * (1 << KM_FRAC));
  const long double den = (long double)prm->range * prm->range * (long double)(1LL << (2 * KM_FRAC));
  const long double prod = (long double)prm->unary_scale * d2max + den;
  if (prod >= 9.0e18L || (long double)prm->unary_scale * d2max / den + 0.5L > (long double)JFC_MAX_TERMINAL_CAP)
    return model_fail(JFC_ERR_OVERFLOW,
                      "unary_scale too large for the image range: a unary could exceed " +
                          std::to_string(JFC_MAX_TERMINAL_CAP));
  return JFC_OK;
}

extern "C" int jfc_model_image(jfc_graph *g, int inst, const void *image, int channels,
                               int bpc, const uint8_t *seeds, int on_device,
                               const jfc_model_params *prm, jfc_model_stats *stats) {
  int rc = check_model_params(channels, bpc, prm);
  if (rc) return rc;
  InstDesc I;
  int *cs = nullptr, *ct = nullptr, *c0 = nullptr;
  long long stride = 0;
  rc = graph_instance(g, inst, &I, &cs, &ct, &c0, &stride);
  if (rc) return rc;
  if (!image || !seeds) return model_fail(JFC_ERR_INVALID, "null argument");
  const long long n = (long long)I.nx * I.ny * I.nz;
  if (n >= (1LL << 31)) return model_fail(JFC_ERR_INVALID, "instance too large for the model path");
  const int C = channels;
  const size_t img_bytes = (size_t)n * C * bpc;
  const bool d3 = I.conn == 6;
  const int nkeys = bpc == 2 ? 4096 : 255 * C + 1;
  void *d_img = nullptr;
  unsigned char *d_seeds = nullptr;
  int *d_list = nullptr, *d_cnt = nullptr, *d_nT = nullptr;
  long long *d_kcnt = nullptr, *d_ksum = nullptr, *d_cent = nullptr, *d_acc = nullptr,
            *d_cf = nullptr, *d_cb = nullptr, *d_sums = nullptr, *d_T = nullptr;
  jfc_model_stats st_local;
  jfc_model_stats &ST = stats ? *stats : st_local;
  memset(&ST, 0, sizeof(ST));
  std::vector<long long> cent_side[2];
  long long sums[4] = {0, 0, 0, 0};
  std::vector<long long> Tab(3 * TMAX, 0);
  int nT[3] = {0, 0, 0};
  const void *img = image;
  const unsigned char *sd = seeds;
  if (!on_device) {
    MODEL_HIP(hipMalloc(&d_img, img_bytes));
    MODEL_HIP(hipMalloc((void **)&d_seeds, n));
    MODEL_HIP(hipMemcpy(d_img, image, img_bytes, hipMemcpyHostToDevice));
    MODEL_HIP(hipMemcpy(d_seeds, seeds, n, hipMemcpyHostToDevice));
    img = d_img;
    sd = d_seeds;
  }
  MODEL_HIP(hipMalloc((void **)&d_list, n * sizeof(int)));
  MODEL_HIP(hipMalloc((void **)&d_cnt, sizeof(int)));
  MODEL_HIP(hipMalloc((void **)&d_kcnt, nkeys * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_ksum, (size_t)nkeys * C * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_cent, KMAX_CLUSTERS * 3 * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_acc, KMAX_CLUSTERS * 4 * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_cf, KMAX_CLUSTERS * 3 * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_cb, KMAX_CLUSTERS * 3 * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_sums, 4 * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_T, 3 * TMAX * sizeof(long long)));
  MODEL_HIP(hipMalloc((void **)&d_nT, 3 * sizeof(int)));

  for (int side = 1; side <= 2; ++side) {
    const int K = side == 1 ? prm->k_fg : prm->k_bg;
    int cnt = 0;
    MODEL_HIP(hipMemset(d_cnt, 0, sizeof(int)));
    k_seed_list<<<grid1(n), 256>>>(sd, n, side, d_list, d_cnt);
    MODEL_HIP(hipMemcpy(&cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost));
    if (side == 1) ST.seeds_fg = cnt; else ST.seeds_bg = cnt;
    if (cnt == 0) {
      rc = model_fail(JFC_ERR_NO_SEEDS, side == 1 ? "no foreground seeds" : "no background seeds");
      goto done;
    }
    // initial centroids: distinct seeded keys, evenly spaced picks
    // (graphcut_cut_distribution + QGCWidget.cpp:2001-2039)
    MODEL_HIP(hipMemset(d_kcnt, 0, nkeys * sizeof(long long)));
    MODEL_HIP(hipMemset(d_ksum, 0, (size_t)nkeys * C * sizeof(long long)));
    k_key_hist<<<grid1(cnt), 256>>>(img, bpc, C, d_list, d_cnt, d_kcnt, d_ksum);
    std::vector<long long> kc(nkeys), ks((size_t)nkeys * C);
    MODEL_HIP(hipMemcpy(kc.data(), d_kcnt, nkeys * sizeof(long long), hipMemcpyDeviceToHost));
    MODEL_HIP(hipMemcpy(ks.data(), d_ksum, ks.size() * sizeof(long long), hipMemcpyDeviceToHost));
    std::vector<int> levels;
    for (int k = 0; k < nkeys; ++k)
      if (kc[k]) levels.push_back(k);
    const int m = (int)levels.size();
    std::vector<long long> cent((size_t)K * C, 0);
    for (int j = 0; j < K; ++j) {
      const int pick = K > 1 ? levels[(long long)j * (m - 1) / (K - 1)] : levels[0];
      const long long nn = kc[pick];
      for (int c = 0; c < C; ++c)
        cent[j * C + c] = (ks[(size_t)pick * C + c] * (1 << KM_FRAC) + nn / 2) / nn;
    }
    // Lloyd iterations (graphcut_kmeans.cl:21-89 + host update QGCWidget.cpp:2073-2102)
    int it = 0;
    while (it < prm->kmeans_iters) {
      ++it;
      MODEL_HIP(hipMemcpy(d_cent, cent.data(), cent.size() * sizeof(long long), hipMemcpyHostToDevice));
      MODEL_HIP(hipMemset(d_acc, 0, KMAX_CLUSTERS * 4 * sizeof(long long)));
      k_km_assign<<<std::min(grid1(cnt), 1024), 256>>>(img, bpc, C, d_list, d_cnt, d_cent, K, d_acc);
      std::vector<long long> acc((size_t)K * (C + 1));
      MODEL_HIP(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(long long), hipMemcpyDeviceToHost));
      bool changed = false;
      for (int k = 0; k < K; ++k) {
        const long long nk = acc[k * (C + 1) + C];
        if (!nk) continue;
        for (int c = 0; c < C; ++c) {
          const long long nc = (acc[k * (C + 1) + c] * (1 << KM_FRAC) + nk / 2) / nk;
          if (nc != cent[k * C + c]) { cent[k * C + c] = nc; changed = true; }
        }
      }
      if (!changed) break;
    }
    if (side == 1) ST.iters_fg = it; else ST.iters_bg = it;
    for (int i = 0; i < K * C; ++i)
      (side == 1 ? ST.centroids_fg : ST.centroids_bg)[i] = cent[i];
    cent_side[side - 1] = cent;
  }
  {
    // unary terms
    MODEL_HIP(hipMemcpy(d_cf, cent_side[0].data(), cent_side[0].size() * sizeof(long long), hipMemcpyHostToDevice));
    MODEL_HIP(hipMemcpy(d_cb, cent_side[1].data(), cent_side[1].size() * sizeof(long long), hipMemcpyHostToDevice));
    const long long den = prm->range * prm->range * (1LL << (2 * KM_FRAC));
    k_unary<<<grid1(n), 256>>>(img, bpc, C, n, I.nx, I.px, sd, d_cf, prm->k_fg, d_cb, prm->k_bg, den,
                               prm->unary_scale, cs + I.node_off, ct + I.node_off);
    // pairwise terms
    MODEL_HIP(hipMemset(d_sums, 0, 4 * sizeof(long long)));
    const int rgrid = (int)std::min<long long>((long long)I.ny * I.nz, 65536);
    if (d3) k_pair_sums<true><<<rgrid, 256>>>(I, img, bpc, C, prm->z_separate, d_sums);
    else k_pair_sums<false><<<rgrid, 256>>>(I, img, bpc, C, 0, d_sums);
    MODEL_HIP(hipMemcpy(sums, d_sums, sizeof(sums), hipMemcpyDeviceToHost));
    ST.pair_sum[0] = sums[0]; ST.pair_count[0] = sums[1];
    ST.pair_sum[1] = sums[2]; ST.pair_count[1] = sums[3];
    nT[0] = thresholds(prm->lambda_xy, sums[0], sums[1], Tab.data());
    nT[1] = thresholds(prm->lambda_xy / std::sqrt(2.0), sums[0], sums[1], Tab.data() + TMAX);
    if (d3 && prm->z_separate) nT[2] = thresholds(prm->lambda_z, sums[2], sums[3], Tab.data() + 2 * TMAX);
    else nT[2] = thresholds(prm->lambda_z, sums[0], sums[1], Tab.data() + 2 * TMAX);
    MODEL_HIP(hipMemcpy(d_T, Tab.data(), Tab.size() * sizeof(long long), hipMemcpyHostToDevice));
    MODEL_HIP(hipMemcpy(d_nT, nT, sizeof(nT), hipMemcpyHostToDevice));
    if (d3) k_pairwise<true><<<rgrid, 256>>>(I, img, bpc, C, d_T, d_nT, c0, stride);
    else k_pairwise<false><<<rgrid, 256>>>(I, img, bpc, C, d_T, d_nT, c0, stride);
    MODEL_HIP(hipDeviceSynchronize());
    graph_mark_inputs(g, inst);
  }
done:
  if (d_img) (void)hipFree(d_img);
  if (d_seeds) (void)hipFree(d_seeds);
  for (void *p : {(void *)d_list, (void *)d_cnt, (void *)d_kcnt, (void *)d_ksum,
                  (void *)d_cent, (void *)d_acc, (void *)d_cf, (void *)d_cb,
                  (void *)d_sums, (void *)d_T, (void *)d_nT})
    if (p) (void)hipFree(p);
  return rc;
}
