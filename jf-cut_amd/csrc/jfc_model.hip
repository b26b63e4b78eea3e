// jfc_model.hip -- data term on the device: integer k-means per seed side,
// unary and pairwise capacities (QGCWidget::slotModelClicked,
// src/QVisualizer/Graph Cut/QGCWidget.cpp:2560-2606).
//
// Reference counterparts (src/QVisualizer/cl/GraphCut/):
//   k_seed_scan   graphcut_count_cut / graphcut_scatter_cut  graphcut_count_block.cl:67-104
//                 + graphcut_cut_distribution                graphcut_histogram.cl:41-60
//   k_km_init     the host pick of initial centres           QGCWidget.cpp:2001-2039
//   k_km_assign   graphcut_kmeans                            graphcut_kmeans.cl:21-89
//   k_km_update   the host centre update                     QGCWidget.cpp:2073-2102
//   k_unary       graphcut_compute_excessflow_lazy           graphcut_compute_excessflow.cl:47-77
//   k_pairwise    graphcut_compute_capacity_bk               graphcut_compute_capacity.cl:21-69
//
// The reference runs k-means in float32 with host-side float division; here
// every quantity is an integer so CPU and GPU agree bit for bit: centroids
// are fixed point (value * 2^8) rounded from exact int64 sums, distances are
// exact integers, unary rounding is integer, and the pairwise exp() becomes a
// table of integer thresholds on the squared difference, computed once per
// call on the host.
//
// The whole model runs on the caller's stream with three host waits (the
// seed counts, the pair sums for the thresholds, the end); the k-means
// iterations run on the device (each update raises a done flag that turns
// the later launches into no-ops).  The scratch is one small block plus a
// seed list: the graph's own (allocated on its first model call, the list
// borrowing the excess plane) for jfc_model_image, or a static device array
// for jfc_model_image_static, which allocates nothing.
#include "jfcut.h"
#include "jfc_device.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

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

constexpr int KM_FRAC = 8;  // fixed-point bits of centroids
constexpr int KMAX_CLUSTERS = 16;
constexpr int TMAX = 256;        // max thresholds per table (lambda < 255.5)
constexpr int NKEYS_MAX = 4096;  // 255 * 3 + 1 keys (RGB) or 4096 (uint16 >> 4)

// Scratch of one model call: a small block (histograms, centroids, sums,
// thresholds, control words) and a seed list.
struct Scratch {
  long long *kcnt;  // [2][NKEYS_MAX] seeded pixels per key
  long long *ksum;  // [2][NKEYS_MAX] feature sums per key and channel (key * C + c)
  long long *cent;  // [2][KMAX_CLUSTERS * 3] centroids, fixed point
  long long *acc;   // [2][KMAX_CLUSTERS * 4] cluster sums and counts of an iteration
  long long *sums;  // [4] S0, P0, S1, P1: squared differences and edges per group
  unsigned *T;      // [3][TMAX] pairwise thresholds
  int *ctl;         // [0,1] seed counts, [2,3] done flags, [4,5] iterations
  int *list;        // seeded node indices: foreground from the front, background
                    // from the back (list[cap - 1 - i])
  long long cap;    // list entries (both sides together)
};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kKcntBytes = align256(2 * NKEYS_MAX * 8);
constexpr size_t kKsumBytes = align256(2 * NKEYS_MAX * 8);
constexpr size_t kCentBytes = align256(2 * KMAX_CLUSTERS * 3 * 8);
constexpr size_t kAccBytes = align256(2 * KMAX_CLUSTERS * 4 * 8);
constexpr size_t kSumsBytes = align256(4 * 8);
constexpr size_t kTBytes = align256(3 * TMAX * 4);
constexpr size_t kCtlBytes = align256(8 * 4);
constexpr size_t kSmallBytes =
    kKcntBytes + kKsumBytes + kCentBytes + kAccBytes + kSumsBytes + kTBytes + kCtlBytes;
// The part that must start at zero in every call: histograms, accumulators,
// sums and control words (cent and T are written before they are read).
constexpr size_t kClearInts = (kKcntBytes + kKsumBytes) / 4;

// The small part of the scratch from `base`; the seed list is given apart.
Scratch carve(void *base, int *list, long long cap) {
  unsigned char *p = (unsigned char *)base;
  Scratch s;
  s.kcnt = (long long *)p; p += kKcntBytes;
  s.ksum = (long long *)p; p += kKsumBytes;
  s.cent = (long long *)p; p += kCentBytes;
  s.acc = (long long *)p; p += kAccBytes;
  s.sums = (long long *)p; p += kSumsBytes;
  s.T = (unsigned *)p; p += kTBytes;
  s.ctl = (int *)p;
  s.list = list;
  s.cap = cap;
  return s;
}

// Seeded pixels per key: the plain sum of the channels (8-bit) or the top 12
// bits (16-bit), the levels graphcut_cut_distribution bins.
template <int BPC>
__device__ __forceinline__ int pixel(const void *img, long long i) {
  return BPC == 1 ? (int)((const unsigned char *)img)[i] : (int)((const unsigned short *)img)[i];
}

// Zero the histograms and, in one launch, the accumulators, sums and control
// words (the scratch keeps the previous call's contents).
__global__ void k_model_clear(Scratch s, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = tid; i < (int)kClearInts; i += stride) ((int *)s.kcnt)[i] = 0;
  for (int i = tid; i < 2 * KMAX_CLUSTERS * 4; i += stride) s.acc[i] = 0;
  if (tid < 4) s.sums[tid] = 0;
  if (tid < 4) s.ctl[tid] = 0;             // seed counts, done flags
  if (tid == 4 || tid == 5) s.ctl[tid] = iters;  // iterations unless done earlier
}

// One pass over the seed map: seeded node lists per side (an entry past the
// list's end is dropped; the count keeps counting so that the host sees the
// overflow) and the key histograms of the seeded pixels.
template <int BPC>
__global__ __launch_bounds__(256) void k_seed_scan(const void *img, int C,
                                                   const unsigned char *seeds, long long n,
                                                   Scratch s) {
  const int lane = lane_id();
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n;
       base += (long long)gridDim.x * blockDim.x) {
    const long long i = base + threadIdx.x;
    const int sd = i < n ? seeds[i] : 0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const bool p = sd == side + 1;
      const unsigned long long m = __ballot(p);
      if (!m) continue;
      const int leader = __ffsll((long long)m) - 1;
      int b = 0;
      if (lane == leader) b = atomicAdd(&s.ctl[side], __popcll(m));
      b = __shfl(b, leader);
      if (p) {
        const long long at = b + __popcll(m & ((1ull << lane) - 1ull));
        if (at < s.cap) s.list[side ? s.cap - 1 - at : at] = (int)i;
        int x[3] = {0, 0, 0}, key = 0;
        for (int c = 0; c < C; ++c) {
          x[c] = pixel<BPC>(img, i * C + c);
          key += x[c];
        }
        if (BPC == 2) key = x[0] >> 4;
        atomicAdd((unsigned long long *)&s.kcnt[side * NKEYS_MAX + key], 1ull);
        for (int c = 0; c < C; ++c)
          atomicAdd((unsigned long long *)&s.ksum[side * NKEYS_MAX + key * C + c],
                    (unsigned long long)x[c]);
      }
    }
  }
}

// Initial centroids per side (one workgroup of 1024 lanes, blockIdx.x =
// side): the distinct seeded keys in ascending order, the j-th centre taken
// from key number j * (m - 1) / (K - 1), rounded to fixed point
// (QGCWidget.cpp:2001-2039).
__global__ __launch_bounds__(1024) void k_km_init(Scratch s, int C, int nkeys, int Kf, int Kb) {
  __shared__ int s_cnt[1024];
  __shared__ long long s_pick[KMAX_CLUSTERS];
  const int side = blockIdx.x, K = side ? Kb : Kf, tid = threadIdx.x;
  const long long *kc = s.kcnt + side * NKEYS_MAX;
  const long long *ks = s.ksum + side * NKEYS_MAX;
  long long *cent = s.cent + side * KMAX_CLUSTERS * 3;
  int mine = 0;
  for (int q = 0; q < 4; ++q) {
    const int k = tid * 4 + q;
    if (k < nkeys && kc[k]) ++mine;
  }
  s_cnt[tid] = mine;
  __syncthreads();
  // inclusive scan over the 1024 counts
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? s_cnt[tid - o] : 0;
    __syncthreads();
    s_cnt[tid] += v;
    __syncthreads();
  }
  const int m = s_cnt[1023];
  if (tid < K) s_pick[tid] = K > 1 ? (long long)tid * (m - 1) / (K - 1) : 0;
  if (tid < K * C) cent[tid] = 0;
  __syncthreads();
  if (m == 0) return;  // no seeds: reported by the host
  int r = s_cnt[tid] - mine;  // rank of this lane's first non-empty key
  for (int q = 0; q < 4; ++q) {
    const int k = tid * 4 + q;
    if (k >= nkeys || !kc[k]) continue;
    for (int j = 0; j < K; ++j)
      if (s_pick[j] == r) {
        const long long nn = kc[k];
        for (int c = 0; c < C; ++c)
          cent[j * C + c] = (ks[k * C + c] * (1 << KM_FRAC) + nn / 2) / nn;
      }
    ++r;
  }
}

// Squared distance between a feature and a fixed-point centroid.  8-bit
// channels: |x * 256 - c| < 2^16, so each square fits 32 bits unsigned;
// 16-bit: 64-bit products.
template <int BPC>
__device__ __forceinline__ long long km_d2(const int *x, const long long *c, int C) {
  if (BPC == 1) {
    unsigned long long d = 0;
    for (int i = 0; i < C; ++i) {
      const int t = x[i] * (1 << KM_FRAC) - (int)c[i];
      d += (unsigned)(t * t);
    }
    return (long long)d;
  } else {
    const long long t = (long long)x[0] * (1 << KM_FRAC) - c[0];
    return t * t;
  }
}

// Lloyd assignment (graphcut_kmeans.cl:21-89): nearest centroid, lowest
// index on ties; cluster sums and counts reduced in LDS, then one global
// atomic per entry and workgroup.  blockIdx.y = side; a side whose
// iteration has converged (done flag) returns at once.
template <int BPC>
__global__ __launch_bounds__(256) void k_km_assign(const void *img, int C, Scratch s, int Kf,
                                                   int Kb) {
  const int side = blockIdx.y, K = side ? Kb : Kf;
  if (s.ctl[2 + side]) return;
  __shared__ long long s_c[KMAX_CLUSTERS * 3];
  __shared__ long long s_acc[KMAX_CLUSTERS * 4];
  for (int i = threadIdx.x; i < K * C; i += blockDim.x) s_c[i] = s.cent[side * KMAX_CLUSTERS * 3 + i];
  for (int i = threadIdx.x; i < K * (C + 1); i += blockDim.x) s_acc[i] = 0;
  __syncthreads();
  const long long n = min((long long)s.ctl[side], s.cap);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = s.list[side ? s.cap - 1 - i : i];
    int x[3] = {0, 0, 0};
    for (int c = 0; c < C; ++c) x[c] = pixel<BPC>(img, v * C + c);
    int best = 0;
    long long bd = km_d2<BPC>(x, s_c, C);
    for (int k = 1; k < K; ++k) {
      const long long d = km_d2<BPC>(x, s_c + k * C, C);
      if (d < bd) { bd = d; best = k; }
    }
    for (int c = 0; c < C; ++c)
      atomicAdd((unsigned long long *)&s_acc[best * (C + 1) + c], (unsigned long long)x[c]);
    atomicAdd((unsigned long long *)&s_acc[best * (C + 1) + C], 1ull);
  }
  __syncthreads();
  long long *acc = s.acc + side * KMAX_CLUSTERS * 4;
  for (int i = threadIdx.x; i < K * (C + 1); i += blockDim.x)
    if (s_acc[i]) atomicAdd((unsigned long long *)&acc[i], (unsigned long long)s_acc[i]);
}

// Centre update of iteration `it` (QGCWidget.cpp:2088-2101): the rounded mean
// of each non-empty cluster; a side where no centre moved is done after
// `it` iterations.  The accumulators are cleared for the next iteration.
// One workgroup of 64 lanes per side (blockIdx.x).
__global__ __launch_bounds__(64) void k_km_update(Scratch s, int C, int Kf, int Kb, int it) {
  const int side = blockIdx.x, K = side ? Kb : Kf;
  if (s.ctl[2 + side]) return;
  long long *acc = s.acc + side * KMAX_CLUSTERS * 4;
  long long *cent = s.cent + side * KMAX_CLUSTERS * 3;
  const int j = threadIdx.x;  // (cluster, channel) = (j / C, j % C)
  int changed = 0;
  long long nc = 0;
  if (j < K * C) {
    const int k = j / C, c = j % C;
    const long long nk = acc[k * (C + 1) + C];
    nc = cent[j];
    if (nk) {
      const long long v = (acc[k * (C + 1) + c] * (1 << KM_FRAC) + nk / 2) / nk;
      changed = v != nc;
      nc = v;
    }
  }
  const int any = __ballot(changed) != 0;
  __syncthreads();
  if (j < K * C) cent[j] = nc;
  for (int i = j; i < K * (C + 1); i += 64) acc[i] = 0;
  if (j == 0 && !any) {
    s.ctl[2 + side] = 1;
    s.ctl[4 + side] = it;
  }
}

// round(q) of a non-negative integer quotient x / den, exactly: the double
// quotient is within one of the floor (x / den < 2^31 here), then one integer
// correction step.
__device__ __forceinline__ long long div_floor(long long x, long long den) {
  long long q = (long long)((double)x / (double)den);
  if (q * den > x) --q;
  else if ((q + 1) * den <= x) ++q;
  return q;
}

// The undirected edges a node owns: the forward directions (even d) whose
// neighbour lies inside the grid; each edge is owned by exactly one node.
// 2D: +x, +y and, 8-connected, (+x,+y), (-x,+y); 3D: +x, +y, +z.
template <int CONN>
struct Fwd {
  static constexpr int N = CONN == 8 ? 4 : (CONN == 6 ? 3 : 2);
  static __device__ __forceinline__ int dir(int j) { return 2 * j; }
  static __device__ __forceinline__ bool inside(int j, int x, int y, int z, int nx, int ny,
                                                int nz) {
    switch (2 * j) {
      case 0: return x + 1 < nx;
      case 2: return y + 1 < ny;
      case 4: return CONN == 6 ? z + 1 < nz : (x + 1 < nx && y + 1 < ny);
      default: return x >= 1 && y + 1 < ny;  // 6: (-x, +y)
    }
  }
  // neighbour offset in the unpadded numbering (row r = y + ny z)
  static __device__ __forceinline__ long long off(int j, int nx, long long plane) {
    switch (2 * j) {
      case 0: return 1;
      case 2: return nx;
      case 4: return CONN == 6 ? plane : nx + 1;
      default: return nx - 1;
    }
  }
  // the same in the stored numbering (rows of pitch px)
  static __device__ __forceinline__ long long soff(int j, int px, long long splane) {
    switch (2 * j) {
      case 0: return 1;
      case 2: return px;
      case 4: return CONN == 6 ? splane : px + 1;
      default: return px - 1;
    }
  }
  // threshold table of the edge: 0 = x/y, 1 = 2D diagonal, 2 = z
  static __device__ __forceinline__ int table(int j) {
    return 2 * j >= 4 ? (CONN == 6 ? 2 : 1) : 0;
  }
};

// ---- row staging -----------------------------------------------------------------
//
// The per-node data-term kernels work on segments of SEG pixels of one image
// row (lane l takes pixels x0 + l + 256 j, so every global store of a wave is
// 64 consecutive words), with the image bytes of the rows a segment reads
// staged in LDS through aligned 4-byte loads: a pixel is 1 to 3 bytes, and
// per-channel byte loads from HBM would spend one memory instruction per
// byte per lane.
constexpr int SEG = 1024;                          // pixels per work item
constexpr int ROWW = (3 * (SEG + 2) + 3) / 4 + 2;  // words of one staged row

struct RowStage {
  int off;  // byte of the first staged pixel inside the staged words
};

// Stage pixels [a, b) of image row `row` (nx pixels per row, bpp bytes per
// pixel, `total` bytes in the image) into `words`.  Words that end past the
// image are assembled from single bytes, so no load leaves the buffer.
__device__ __forceinline__ RowStage stage_row(const unsigned char *img, long long total,
                                              long long row, int nx, int bpp, int a, int b,
                                              unsigned *words) {
  const long long first = (row * nx + a) * bpp, last = (row * nx + b) * bpp;
  const long long w0 = first >> 2, w1 = (last + 3) >> 2;
  for (long long w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
    unsigned v;
    if ((w << 2) + 4 <= total) {
      v = ((const unsigned *)img)[w];
    } else {
      v = 0;
      for (int k = 0; k < 4; ++k)
        if ((w << 2) + k < total) v |= (unsigned)img[(w << 2) + k] << (8 * k);
    }
    words[w - w0] = v;
  }
  RowStage s;
  s.off = (int)(first - (w0 << 2));
  return s;
}

// Channel c of staged pixel p (p counted from the stage's first pixel).
template <int BPC>
__device__ __forceinline__ int staged(const unsigned *words, const RowStage &s, int p, int bpp,
                                      int c) {
  const unsigned char *b = (const unsigned char *)words + s.off + p * bpp;
  return BPC == 1 ? (int)b[c] : (int)((const unsigned short *)b)[c];
}

template <int BPC>
__device__ __forceinline__ unsigned long long staged_sqdiff(const unsigned *wa,
                                                            const RowStage &sa, int pa,
                                                            const unsigned *wb,
                                                            const RowStage &sb, int pb,
                                                            int bpp, int C) {
  unsigned long long D = 0;
  for (int c = 0; c < C; ++c) {
    const long long t = staged<BPC>(wa, sa, pa, bpp, c) - staged<BPC>(wb, sb, pb, bpp, c);
    D += (unsigned long long)(t * t);
  }
  return D;
}

// cs = round(scale * d2_bg / den), ct = round(scale * d2_fg / den); seeds get
// (2^30, 0) / (0, 2^30).  Rows of the image (unpadded, nx nodes) go to rows
// of the planes (pitch px).
template <int BPC>
__global__ __launch_bounds__(256) void k_unary(const void *img, int C, const unsigned char *seeds,
                                               int nx, long long rows, int px, Scratch s,
                                               int Kf, int Kb, long long den, long long scale,
                                               int *cs, int *ct) {
  __shared__ long long s_f[KMAX_CLUSTERS * 3], s_b[KMAX_CLUSTERS * 3];
  __shared__ unsigned s_row[ROWW];
  for (int i = threadIdx.x; i < Kf * C; i += blockDim.x) s_f[i] = s.cent[i];
  for (int i = threadIdx.x; i < Kb * C; i += blockDim.x) s_b[i] = s.cent[KMAX_CLUSTERS * 3 + i];
  const int bpp = C * BPC;
  const long long total = rows * nx * bpp;
  const long long half = den / 2;
  const int segs = (nx + SEG - 1) / SEG;
  for (long long item = blockIdx.x; item < rows * segs; item += gridDim.x) {
    const long long r = item / segs;
    const int x0 = (int)(item % segs) * SEG, x1 = min(nx, x0 + SEG);
    __syncthreads();  // the previous item's reads of s_row are done
    const RowStage rs = stage_row((const unsigned char *)img, total, r, nx, bpp, x0, x1, s_row);
    __syncthreads();
    for (int x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
      const long long v = r * nx + x;
      const int sd = seeds[v];
      int a, b;
      if (sd == 1) { a = 1 << 30; b = 0; }
      else if (sd == 2) { a = 0; b = 1 << 30; }
      else {
        int f[3] = {0, 0, 0};
        for (int c = 0; c < C; ++c) f[c] = staged<BPC>(s_row, rs, x - x0, bpp, c);
        long long df = km_d2<BPC>(f, s_f, C), db = km_d2<BPC>(f, s_b, C);
        for (int k = 1; k < Kf; ++k) df = min(df, km_d2<BPC>(f, s_f + k * C, C));
        for (int k = 1; k < Kb; ++k) db = min(db, km_d2<BPC>(f, s_b + k * C, C));
        a = (int)div_floor(scale * db + half, den);
        b = (int)div_floor(scale * df + half, den);
      }
      cs[r * px + x] = a;
      ct[r * px + x] = b;
    }
  }
}

// The rows a segment of row r = y + ny z reads for its forward edges: its own
// row (x0 .. x1, plus the pixel right of it), the next row of the slice
// (x0 - 1 .. x1 + 1) and, in 3D, the same row of the next slice.
struct Stages {
  RowStage own, next, up;
  int a_own, a_next, a_up;  // first staged pixel of each
};

template <int CONN>
__device__ __forceinline__ Stages stage_rows(const unsigned char *img, long long total, int nx,
                                             int ny, int nz, int bpp, long long r, int x0,
                                             int x1, unsigned *w_own, unsigned *w_next,
                                             unsigned *w_up) {
  const int y = (int)(r % ny), z = (int)(r / ny);
  Stages S;
  S.a_own = x0;
  S.own = stage_row(img, total, r, nx, bpp, x0, min(nx, x1 + 1), w_own);
  S.a_next = max(0, x0 - 1);
  if (y + 1 < ny) S.next = stage_row(img, total, r + 1, nx, bpp, S.a_next, min(nx, x1 + 1), w_next);
  S.a_up = x0;
  if (CONN == 6 && z + 1 < nz) S.up = stage_row(img, total, r + ny, nx, bpp, x0, x1, w_up);
  return S;
}

// Squared difference of forward edge j of node (x, y, z), all from LDS.
template <int CONN, int BPC>
__device__ __forceinline__ unsigned long long edge_d(const Stages &S, const unsigned *w_own,
                                                     const unsigned *w_next, const unsigned *w_up,
                                                     int x, int j, int bpp, int C) {
  const int p = x - S.a_own;
  switch (2 * j) {
    case 0: return staged_sqdiff<BPC>(w_own, S.own, p, w_own, S.own, p + 1, bpp, C);
    case 2: return staged_sqdiff<BPC>(w_own, S.own, p, w_next, S.next, x - S.a_next, bpp, C);
    case 4:
      if (CONN == 6) return staged_sqdiff<BPC>(w_own, S.own, p, w_up, S.up, x - S.a_up, bpp, C);
      return staged_sqdiff<BPC>(w_own, S.own, p, w_next, S.next, x + 1 - S.a_next, bpp, C);
    default: return staged_sqdiff<BPC>(w_own, S.own, p, w_next, S.next, x - 1 - S.a_next, bpp, C);
  }
}

// Sum of squared neighbour differences per edge group (0: x/y and the 2D
// diagonals, and z unless it forms its own group; 1: z when separate), one
// term per undirected edge.
template <int CONN, int BPC>
__global__ __launch_bounds__(256) void k_pair_sums(const void *img, int C, int nx, int ny,
                                                   int nz, int zsep, long long *sums) {
  __shared__ long long buf[4][4];
  __shared__ unsigned w_own[ROWW], w_next[ROWW], w_up[CONN == 6 ? ROWW : 1];
  unsigned long long s0 = 0, s1 = 0;
  long long p0 = 0, p1 = 0;
  const int bpp = C * BPC;
  const long long rows = (long long)ny * nz, total = rows * nx * bpp;
  const int segs = (nx + SEG - 1) / SEG;
  for (long long item = blockIdx.x; item < rows * segs; item += gridDim.x) {
    const long long r = item / segs;
    const int x0 = (int)(item % segs) * SEG, x1 = min(nx, x0 + SEG);
    const int y = (int)(r % ny), z = (int)(r / ny);
    __syncthreads();
    const Stages S = stage_rows<CONN>((const unsigned char *)img, total, nx, ny, nz, bpp, r, x0, x1,
                                      w_own, w_next, w_up);
    __syncthreads();
    for (int x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
#pragma unroll
      for (int j = 0; j < Fwd<CONN>::N; ++j) {
        if (!Fwd<CONN>::inside(j, x, y, z, nx, ny, nz)) continue;
        const unsigned long long D = edge_d<CONN, BPC>(S, w_own, w_next, w_up, x, j, bpp, C);
        if (CONN == 6 && j == 2 && zsep) { s1 += D; ++p1; }
        else { s0 += D; ++p0; }
      }
    }
  }
  long long vals[4] = {(long long)s0, p0, (long long)s1, p1};
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

// Pairwise weights w(D) = #{k : D <= T[k]} for each undirected edge, written
// to both of its arcs (the weight is symmetric): plane d at the owner and
// plane d^1 at the neighbour; arcs that leave the grid get 0 from the node
// they start at.  Every table is non-increasing, so w is the length of the
// prefix with D <= T[k]: a guess from the continuous form lambda exp(-D /
// (2 s2)) + 1/2 is corrected against the exact integer table.
struct PairParams {
  int n[3];          // thresholds per table
  float lam[3];      // lambda per table (the guess)
  float inv2s2[3];   // 1 / (2 s2) per table (the guess)
};

template <int CONN, int BPC>
__global__ __launch_bounds__(256) void k_pairwise(const void *img, int C, int nx, int ny,
                                                  int nz, int px, const unsigned *T,
                                                  PairParams pp, int *c0, long long stride) {
  __shared__ unsigned s_T[3 * TMAX];
  __shared__ unsigned w_own[ROWW], w_next[ROWW], w_up[CONN == 6 ? ROWW : 1];
  for (int i = threadIdx.x; i < 3 * TMAX; i += blockDim.x) s_T[i] = T[i];
  const int bpp = C * BPC;
  const long long rows = (long long)ny * nz, total = rows * nx * bpp;
  const long long splane = (long long)px * ny;
  const int segs = (nx + SEG - 1) / SEG;
  for (long long item = blockIdx.x; item < rows * segs; item += gridDim.x) {
    const long long r = item / segs;
    const int x0 = (int)(item % segs) * SEG, x1 = min(nx, x0 + SEG);
    const int y = (int)(r % ny), z = (int)(r / ny);
    __syncthreads();
    const Stages S = stage_rows<CONN>((const unsigned char *)img, total, nx, ny, nz, bpp, r, x0, x1,
                                      w_own, w_next, w_up);
    __syncthreads();
    for (int x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
      const long long sv = r * px + x;
#pragma unroll
      for (int j = 0; j < Fwd<CONN>::N; ++j) {
        const int d = Fwd<CONN>::dir(j);
        int w = 0;
        if (Fwd<CONN>::inside(j, x, y, z, nx, ny, nz)) {
          const unsigned long long D = edge_d<CONN, BPC>(S, w_own, w_next, w_up, x, j, bpp, C);
          const int tb = Fwd<CONN>::table(j);
          const int n = pp.n[tb];
          const unsigned *t = s_T + tb * TMAX;
          w = (int)(pp.lam[tb] * __expf(-(float)D * pp.inv2s2[tb]) + 0.5f);
          w = max(0, min(w, n));
          while (w < n && D <= t[w]) ++w;
          while (w > 0 && D > t[w - 1]) --w;
          c0[(d ^ 1) * stride + sv + Fwd<CONN>::soff(j, px, splane)] = w;
        }
        c0[d * stride + sv] = w;
      }
      // reverse arcs of this node that leave the grid
      if (x == 0) c0[1 * stride + sv] = 0;
      if (y == 0) c0[3 * stride + sv] = 0;
      if (CONN == 6 && z == 0) c0[5 * stride + sv] = 0;
      if (CONN == 8) {
        if (x == 0 || y == 0) c0[5 * stride + sv] = 0;
        if (x + 1 == nx || y == 0) c0[7 * stride + sv] = 0;
      }
    }
  }
}

int thresholds(double lambda, long long S, long long P, unsigned *T) {
  int kmax = (int)std::floor(lambda + 0.5);
  if (kmax > TMAX) kmax = TMAX;
  for (int k = 1; k <= kmax; ++k) {
    if (S == 0 || P == 0) { T[k - 1] = 0xffffffffu; continue; }
    const double v = (2.0 * (double)S / (double)P) * std::log(lambda / (k - 0.5));
    // differences are below 2^32, so a threshold at or above 2^32 - 1 admits
    // every difference, like the int64 value it stands for
    const double f = std::floor(v);
    T[k - 1] = f >= 4294967295.0 ? 0xffffffffu : (unsigned)f;
  }
  return kmax;
}

int grid1(long long n, int cap = 8192) {
  long long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// Parameter checks that need no graph.  The unary of a node is
// round(scale * d2 / den) with d2 <= C * (maxval * 2^KM_FRAC)^2 (a feature and
// a centroid both in [0, maxval]) and den = range^2 * 2^(2 KM_FRAC): the
// product scale * d2 must fit int64 and the unary must not exceed the
// terminal-capacity limit, or the solve's int32 arithmetic would not be exact.
int check_model_params(int channels, int bpc, const jfc_model_params *prm) {
  if (!prm) return model_fail(JFC_ERR_INVALID, "null argument");
  if (!((channels == 1 || channels == 3) && (bpc == 1 || (bpc == 2 && channels == 1))))
    return model_fail(JFC_ERR_INVALID, "image must be 1 or 3 channels of uint8, or 1 channel of uint16");
  if (prm->k_fg < 1 || prm->k_fg > KMAX_CLUSTERS || prm->k_bg < 1 || prm->k_bg > KMAX_CLUSTERS)
    return model_fail(JFC_ERR_INVALID, "cluster count must be 1..16");
  if (prm->range <= 0 || prm->range > (1 << 20) || prm->unary_scale < 0 ||
      !(prm->lambda_xy >= 0) || !(prm->lambda_z >= 0) || prm->lambda_xy > 255 ||
      prm->lambda_z > 255 || prm->kmeans_iters < 0)
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

#define MODEL_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return model_fail(JFC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// The model of one instance from device-resident image and seeds, with the
// given scratch, on stream st.  Writes the instance's cs, ct and c0 planes.
template <int BPC>
int run_model(const InstDesc &I, const void *img, int C, const unsigned char *seeds,
              const jfc_model_params *prm, jfc_model_stats &ST, Scratch s, int *cs, int *ct,
              int *c0, long long stride, hipStream_t st, bool *planes_written) {
  *planes_written = false;
  const long long n = (long long)I.nx * I.ny * I.nz;
  const long long rows = (long long)I.ny * I.nz;
  const int nkeys = BPC == 2 ? 4096 : 255 * C + 1;
  const int iters = prm->kmeans_iters;
  k_model_clear<<<64, 256, 0, st>>>(s, iters);
  k_seed_scan<BPC><<<grid1(n, 16384), 256, 0, st>>>(img, C, seeds, n, s);
  int ctl[6];
  MODEL_HIP(hipMemcpyAsync(ctl, s.ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipStreamSynchronize(st));
  ST.seeds_fg = ctl[0];
  ST.seeds_bg = ctl[1];
  if (ctl[0] == 0 || ctl[1] == 0)
    return model_fail(JFC_ERR_NO_SEEDS, ctl[0] == 0 ? "no foreground seeds" : "no background seeds");
  if ((long long)ctl[0] + ctl[1] > s.cap)
    return model_fail(JFC_ERR_INVALID, "more seeded nodes (" +
                                           std::to_string((long long)ctl[0] + ctl[1]) +
                                           ") than the seed list holds (" +
                                           std::to_string(s.cap) + ")");
  // k-means on the device: iteration `it` assigns and updates; a side whose
  // centres stop moving is done and its later launches return at once
  k_km_init<<<2, 1024, 0, st>>>(s, C, nkeys, prm->k_fg, prm->k_bg);
  const long long most = std::max(ctl[0], ctl[1]);
  const dim3 ag(std::min(grid1(most), 1024), 2);
  for (int it = 1; it <= iters; ++it) {
    k_km_assign<BPC><<<ag, 256, 0, st>>>(img, C, s, prm->k_fg, prm->k_bg);
    k_km_update<<<2, 64, 0, st>>>(s, C, prm->k_fg, prm->k_bg, it);
  }
  // unary terms from the final centres, straight into the stored rows
  const long long den = prm->range * prm->range * (1LL << (2 * KM_FRAC));
  // one workgroup per row segment of SEG pixels, up to 8192 in flight
  const int rgrid = (int)std::min<long long>(rows * ((I.nx + SEG - 1) / SEG), 8192);
  k_unary<BPC><<<rgrid, 256, 0, st>>>(img, C, seeds, I.nx, rows, I.px, s, prm->k_fg,
                                      prm->k_bg, den, prm->unary_scale, cs + I.node_off,
                                      ct + I.node_off);
  *planes_written = true;
  // pairwise: sums of squared differences, then the thresholds on the host
  const bool d3 = I.conn == 6;
  const int zsep = d3 && prm->z_separate ? 1 : 0;
  if (I.conn == 6) k_pair_sums<6, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, zsep, s.sums);
  else if (I.conn == 8) k_pair_sums<8, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, 0, s.sums);
  else k_pair_sums<4, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, 0, s.sums);
  long long sums[4];
  long long cent[2][KMAX_CLUSTERS * 3];
  MODEL_HIP(hipMemcpyAsync(sums, s.sums, sizeof(sums), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipMemcpyAsync(cent, s.cent, sizeof(cent), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipMemcpyAsync(ctl, s.ctl, 6 * sizeof(int), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipStreamSynchronize(st));
  ST.iters_fg = ctl[4];
  ST.iters_bg = ctl[5];
  for (int i = 0; i < prm->k_fg * C; ++i) ST.centroids_fg[i] = cent[0][i];
  for (int i = 0; i < prm->k_bg * C; ++i) ST.centroids_bg[i] = cent[1][i];
  ST.pair_sum[0] = sums[0]; ST.pair_count[0] = sums[1];
  ST.pair_sum[1] = sums[2]; ST.pair_count[1] = sums[3];
  static_assert(sizeof(unsigned) == 4, "thresholds are 32-bit");
  unsigned Tab[3 * TMAX];
  std::memset(Tab, 0, sizeof(Tab));
  PairParams pp;
  const double lam[3] = {prm->lambda_xy, prm->lambda_xy / std::sqrt(2.0), prm->lambda_z};
  const int zg = zsep ? 2 : 0;  // sums of the z table: own group or the x/y one
  const long long S[3] = {sums[0], sums[0], sums[zg]}, P[3] = {sums[1], sums[1], sums[zg + 1]};
  for (int t = 0; t < 3; ++t) {
    pp.n[t] = thresholds(lam[t], S[t], P[t], Tab + t * TMAX);
    pp.lam[t] = (float)lam[t];
    pp.inv2s2[t] = (S[t] == 0 || P[t] == 0) ? 0.0f : (float)((double)P[t] / (2.0 * (double)S[t]));
  }
  MODEL_HIP(hipMemcpyAsync(s.T, Tab, sizeof(Tab), hipMemcpyHostToDevice, st));
  if (I.conn == 6)
    k_pairwise<6, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, I.px, s.T, pp,
                                              c0 + I.node_off, stride);
  else if (I.conn == 8)
    k_pairwise<8, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, I.px, s.T, pp,
                                              c0 + I.node_off, stride);
  else
    k_pairwise<4, BPC><<<rgrid, 256, 0, st>>>(img, C, I.nx, I.ny, I.nz, I.px, s.T, pp,
                                              c0 + I.node_off, stride);
  MODEL_HIP(hipGetLastError());
  MODEL_HIP(hipStreamSynchronize(st));
  return JFC_OK;
}

int model_common(jfc_graph *g, int inst, const void *img, int C, int bpc,
                 const unsigned char *seeds, const jfc_model_params *prm,
                 jfc_model_stats *stats, void *scratch, int *list, long long cap) {
  InstDesc I;
  int *cs = nullptr, *ct = nullptr, *c0 = nullptr;
  long long stride = 0;
  int rc = graph_instance(g, inst, &I, &cs, &ct, &c0, &stride);
  if (rc) return rc;
  jfc_model_stats st_local;
  jfc_model_stats &ST = stats ? *stats : st_local;
  std::memset(&ST, 0, sizeof(ST));
  const Scratch s = carve(scratch, list, cap);
  bool written = false;
  rc = bpc == 2 ? run_model<2>(I, img, C, seeds, prm, ST, s, cs, ct, c0, stride, graph_stream(g), &written)
                : run_model<1>(I, img, C, seeds, prm, ST, s, cs, ct, c0, stride, graph_stream(g), &written);
  if (rc) {
    if (written) graph_inputs_failed(g, inst);
    return rc;
  }
  graph_mark_inputs(g, inst);
  return JFC_OK;
}

}  // namespace

extern "C" int jfc_model_image(jfc_graph *g, int inst, const void *image, int channels,
                               int bpc, const uint8_t *seeds, int on_device,
                               const jfc_model_params *prm, jfc_model_stats *stats) {
  int rc = check_model_params(channels, bpc, prm);
  if (rc) return rc;
  InstDesc I;
  int *cs, *ct, *c0;
  long long stride;
  if ((rc = graph_instance(g, inst, &I, &cs, &ct, &c0, &stride))) return rc;
  if (!image || !seeds) return model_fail(JFC_ERR_INVALID, "null argument");
  const long long n = (long long)I.nx * I.ny * I.nz;
  if (n >= (1LL << 31)) return model_fail(JFC_ERR_INVALID, "instance too large for the model path");
  const void *img = image;
  const unsigned char *sd = seeds;
  if (!on_device) {
    // host inputs cross in one copy each through the graph's staging buffer
    const size_t ib = align256((size_t)n * channels * bpc);
    void *stage = nullptr;
    if ((rc = graph_stage(g, ib + (size_t)n, &stage))) return rc;
    hipStream_t st = graph_stream(g);
    MODEL_HIP(hipMemcpyAsync(stage, image, (size_t)n * channels * bpc, hipMemcpyHostToDevice, st));
    MODEL_HIP(hipMemcpyAsync((unsigned char *)stage + ib, seeds, (size_t)n, hipMemcpyHostToDevice, st));
    img = stage;
    sd = (const unsigned char *)stage + ib;
  }
  // the graph's small scratch block (allocated on its first model call);
  // the seed list borrows the instance's slice of the excess plane, which
  // every solve rewrites from the inputs before reading it
  void *scratch = nullptr;
  if ((rc = graph_scratch(g, kSmallBytes, &scratch))) return rc;
  int *list = nullptr;
  long long cap = 0;
  if ((rc = graph_work_plane(g, inst, &list, &cap))) return rc;
  return model_common(g, inst, img, channels, bpc, sd, prm, stats, scratch, list, cap);
}
// ---- the entry point that allocates nothing --------------------------------------
//
// jfc_model_image_static keeps its scratch in this file-scope device array:
// each GPU holds one copy for the life of the process, at an address that
// hipGetSymbolAddress returns for the device that is current.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address); it passes the per-device address instead.
constexpr long long kStaticSeeds = JFC_MODEL_STATIC_MAX_SEEDS;  // both sides together
constexpr size_t kStaticBytes = kSmallBytes + (size_t)kStaticSeeds * 4;
static_assert(kStaticBytes == (size_t)JFC_MODEL_STATIC_BYTES, "jfcut.h states the array's size");
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// Address of the current device's copy of `scratch`, looked up once per
// device with that device current.
int static_scratch(void **p) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return model_fail(JFC_ERR_DEVICE, "no current HIP device for the static scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return model_fail(JFC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *p = g_scratch_ptr[dev];
  return JFC_OK;
}
}  // namespace

extern "C" int jfc_model_image_static(jfc_graph *g, int inst, const void *image, int channels,
                                      int bpc, const uint8_t *seeds,
                                      const jfc_model_params *prm, jfc_model_stats *stats) {
  int rc = check_model_params(channels, bpc, prm);
  if (rc) return rc;
  InstDesc I;
  int *cs, *ct, *c0;
  long long stride;
  if ((rc = graph_instance(g, inst, &I, &cs, &ct, &c0, &stride))) return rc;
  if (!image || !seeds) return model_fail(JFC_ERR_INVALID, "null argument");
  const long long n = (long long)I.nx * I.ny * I.nz;
  if (n >= (1LL << 31)) return model_fail(JFC_ERR_INVALID, "instance too large for the model path");
  void *base = nullptr;
  if ((rc = static_scratch(&base))) return rc;
  return model_common(g, inst, image, channels, bpc, seeds, prm, stats, base,
                      (int *)((unsigned char *)base + kSmallBytes), kStaticSeeds);
}
