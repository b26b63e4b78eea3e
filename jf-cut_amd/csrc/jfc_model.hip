// jfc_model.hip -- data term on the device: integer k-means per seed side,
// unary and pairwise capacities (QGCWidget::slotModelClicked,
// src/QVisualizer/Graph Cut/QGCWidget.cpp:2560-2606).
//
// Reference counterparts (src/QVisualizer/cl/GraphCut/):
//   k_seed_scan   graphcut_count_cut / graphcut_scatter_cut  graphcut_count_block.cl:67-104
//                 + graphcut_cut_distribution                graphcut_histogram.cl:41-60
//   k_km_init     the host pick of initial centres           QGCWidget.cpp:2001-2039
//   k_km_iter     graphcut_kmeans + the host centre update   graphcut_kmeans.cl:21-89,
//   k_km_final                                               QGCWidget.cpp:2073-2102
//   k_unary       graphcut_compute_excessflow_lazy           graphcut_compute_excessflow.cl:47-77
//   k_pair_sums   the squared differences behind the pairwise sigma
//   k_pairwise    graphcut_compute_capacity_bk               graphcut_compute_capacity.cl:21-69
//
// The reference runs k-means in float32 with host-side float division; here
// every quantity is an integer so CPU and GPU agree bit for bit: centroids
// are fixed point (value * 2^8) rounded from exact int64 sums, distances are
// exact integers, unary rounding is integer, and the pairwise exp() becomes a
// table of integer thresholds on the squared difference, computed once per
// call on the host.
//
// The model runs on the caller's stream with three host waits (the seed
// counts, the pair sums for the thresholds, the end); the pair sums and the
// pairwise weights, which need neither seeds nor centres, run on a second
// stream of the graph beside the seed scan, the k-means and the unary terms.
// The k-means iterations run on the device (an update that moves no centre
// raises a done flag that turns the later launches into no-ops).  The
// scratch is one small block plus a seed list: the graph's own (allocated on its first model call, the list
// borrowing the excess plane) for jfc_model_image, or a static device array
// for jfc_model_image_static, which allocates nothing (and so uses the second
// stream only once a jfc_model_image call on the graph has created it).
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
int graph_aux(jfc_graph *g, bool create, size_t pin_bytes, GraphAux *a);
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
  long long *cent;  // [2][KMAX_CLUSTERS * 3] final centroids, fixed point
  long long *cbuf;  // [2][2][KMAX_CLUSTERS * 3] centroids of the running iterations
  double *q;        // [2][KMAX_CLUSTERS][4] per centre: |c|^2, -2 c_0, -2 c_1, -2 c_2
  long long *part;  // [2][2][KM_GRID][KMAX_CLUSTERS * 4] cluster sums and counts of
                    // an iteration (two, alternating), per side and workgroup
  long long *sums;  // [NSUM][2] squared differences of edge groups 0 and 1, in
                    // NSUM slots that spread the workgroups' atomics
  unsigned *T;      // [3][TMAX] pairwise thresholds
  int *ctl;         // [0,1] seed counts, [2,3] done flags, [4,5] iterations
  int *list;        // seeded pixels' features, packed (8-bit channel c in bits
                    // 8c .. 8c + 7, or the 16-bit value): foreground from the
                    // front, background from the back (list[cap - 1 - i])
  long long cap;    // list entries (both sides together)
};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kKcntBytes = align256(2 * NKEYS_MAX * 8);
constexpr size_t kKsumBytes = align256(2 * NKEYS_MAX * 8);
constexpr size_t kCentBytes = align256(2 * KMAX_CLUSTERS * 3 * 8);
constexpr size_t kCbufBytes = align256(4 * KMAX_CLUSTERS * 3 * 8);
constexpr size_t kQBytes = align256(2 * KMAX_CLUSTERS * 4 * 8);
constexpr int KM_GRID = 256;  // assigning workgroups per side, at most
constexpr size_t kPartBytes = align256(4 * KM_GRID * KMAX_CLUSTERS * 4 * 8);
constexpr int NSUM = 256;
constexpr size_t kSumsBytes = align256(NSUM * 2 * 8);
// pinned words of the second stream: the sum slots, then the thresholds
constexpr size_t kAuxPinBytes = NSUM * 2 * 8 + 3 * TMAX * 4;
constexpr size_t kTBytes = align256(3 * TMAX * 4);
constexpr size_t kCtlBytes = align256(8 * 4);
constexpr size_t kSmallBytes =
    kKcntBytes + kKsumBytes + kCentBytes + kCbufBytes + kQBytes + kPartBytes + kSumsBytes + kTBytes +
    kCtlBytes;
// The part that must start at zero in every call: histograms, sums and
// control words (cent, cbuf, q, part and T are written before they are read).
constexpr size_t kClearInts = (kKcntBytes + kKsumBytes) / 4;

// The small part of the scratch from `base`; the seed list is given apart.
Scratch carve(void *base, int *list, long long cap) {
  unsigned char *p = (unsigned char *)base;
  Scratch s;
  s.kcnt = (long long *)p; p += kKcntBytes;
  s.ksum = (long long *)p; p += kKsumBytes;
  s.cent = (long long *)p; p += kCentBytes;
  s.cbuf = (long long *)p; p += kCbufBytes;
  s.q = (double *)p; p += kQBytes;
  s.part = (long long *)p; p += kPartBytes;
  s.sums = (long long *)p; p += kSumsBytes;
  s.T = (unsigned *)p; p += kTBytes;
  s.ctl = (int *)p;
  s.list = list;
  s.cap = cap;
  return s;
}

// Zero the histograms and, in one launch, the sums and control words (the
// scratch keeps the previous call's contents).
__global__ void k_model_clear(Scratch s, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = tid; i < (int)kClearInts; i += stride) ((int *)s.kcnt)[i] = 0;
  for (int i = tid; i < NSUM * 2; i += stride) s.sums[i] = 0;
  if (tid < 4) s.ctl[tid] = 0;                   // seed counts, done flags
  if (tid == 4 || tid == 5) s.ctl[tid] = iters;  // iterations unless done earlier
}

// One pass over the seed map, 16 nodes per lane (one 16-byte load where the
// map is 16-byte aligned): the seeded pixels' packed features in a list per
// side (an entry past the list's end is dropped; the count keeps counting so
// that the host sees the overflow) and the key histograms of the seeded
// pixels (key: the plain sum of the channels for 8-bit images, the top 12
// bits for 16-bit ones, the levels graphcut_cut_distribution bins).  A lane
// with seeds loads its 16 pixels' bytes at once (16 * C *
// BPC <= 48 bytes).  List space is claimed once per wave and side, at
// offsets from a wave prefix sum of the lanes' seed counts.  Each workgroup
// histograms its pixels in LDS (32-bit: a workgroup sees fewer than 2^22
// pixels) and adds the non-zero bins to the global histograms; 16-bit images
// keep per key the sum of the 4 bits below the key (value = 16 key + low
// bits).
constexpr int kScanGrid = 1536;  // 6 workgroups per CU (LDS-bound)

template <int BPC, int C>
__global__ __launch_bounds__(256) void k_seed_scan(const void *img, const unsigned char *seeds,
                                                   long long n, Scratch s) {
  constexpr int NK = BPC == 2 ? 4096 : 255 * C + 1;
  constexpr int NS = BPC == 2 ? NK : C * NK;
  constexpr int BPP = C * BPC, PW = (16 * BPP) / 4;  // words of a lane's 16 pixels
  __shared__ unsigned h_cnt[2][NK];
  __shared__ unsigned h_sum[2][NS];
  __shared__ int s_any;
  for (int i = threadIdx.x; i < 2 * NK; i += blockDim.x) (&h_cnt[0][0])[i] = 0;
  for (int i = threadIdx.x; i < 2 * NS; i += blockDim.x) (&h_sum[0][0])[i] = 0;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  const int lane = lane_id();
  const long long groups = (n + 15) / 16;
  const bool s_al = ((size_t)seeds & 15) == 0, i_al = ((size_t)img & 15) == 0;
  const unsigned char *im = (const unsigned char *)img;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = (long long)blockIdx.x * blockDim.x; base < groups; base += 4 * stride) {
    // four groups' seed words in flight together
    uint4 sv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long gi = base + u * stride + threadIdx.x;
      const long long i0 = gi * 16;
      sv[u] = make_uint4(0u, 0u, 0u, 0u);
      if (gi < groups) {
        if (s_al && i0 + 16 <= n) {
          sv[u] = *(const uint4 *)(seeds + i0);
        } else {
          unsigned t[4] = {0u, 0u, 0u, 0u};
          for (int j = 0; j < 16; ++j)
            if (i0 + j < n) t[j >> 2] |= (unsigned)seeds[i0 + j] << (8 * (j & 3));
          sv[u] = make_uint4(t[0], t[1], t[2], t[3]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (base + u * stride >= groups) break;
      const long long gi = base + u * stride + threadIdx.x;
      const long long i0 = gi * 16;
      unsigned m[2] = {0u, 0u};  // bit j: node i0 + j is seeded on that side
      {
        const unsigned w[4] = {sv[u].x, sv[u].y, sv[u].z, sv[u].w};
        if (w[0] | w[1] | w[2] | w[3]) {
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const unsigned b = (w[j >> 2] >> (8 * (j & 3))) & 255u;
            m[0] |= (unsigned)(b == 1u) << j;
            m[1] |= (unsigned)(b == 2u) << j;
          }
        }
      }
      const unsigned any = __ballot((m[0] | m[1]) != 0u) != 0;
      if (!any) continue;
      if (lane == 0) s_any = 1;
      // the lane's 16 pixels, all loads in flight together
      unsigned pw[PW];
#pragma unroll
      for (int k = 0; k < PW; ++k) pw[k] = 0;
      if (m[0] | m[1]) {
        const long long b0 = i0 * BPP;
        if (i_al && i0 + 16 <= n) {
#pragma unroll
          for (int k = 0; k < PW / 4; ++k) {
            const uint4 v = *(const uint4 *)(im + b0 + 16 * k);
            pw[4 * k] = v.x; pw[4 * k + 1] = v.y; pw[4 * k + 2] = v.z; pw[4 * k + 3] = v.w;
          }
        } else {
          const long long nbytes = min(16LL, n - i0) * BPP;
#pragma unroll
          for (int k = 0; k < 4 * PW; ++k)
            if (k < nbytes) pw[k >> 2] |= (unsigned)im[b0 + k] << (8 * (k & 3));
        }
      }
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        if (!__ballot(m[side] != 0u)) continue;
        const int cnt = __popc(m[side]);
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
          const int t = __shfl_up(incl, o);
          if (lane >= o) incl += t;
        }
        int b = 0;
        if (lane == 63) b = atomicAdd(&s.ctl[side], incl);
        b = __shfl(b, 63);
        long long at = (long long)b + incl - cnt;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          if (!((m[side] >> j) & 1u)) continue;
          int x[C], key = 0;
          unsigned packed = 0;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            if (BPC == 1) {
              const int k = j * C + c;
              x[c] = (int)((pw[k >> 2] >> (8 * (k & 3))) & 255u);
            } else {
              x[c] = (int)((pw[j >> 1] >> (16 * (j & 1))) & 65535u);
            }
            key += x[c];
            packed |= (unsigned)x[c] << (8 * c);
          }
          if (at < s.cap) s.list[side ? s.cap - 1 - at : at] = (int)packed;
          ++at;
          if (BPC == 2) {
            key = x[0] >> 4;
            atomicAdd(&h_sum[side][key], (unsigned)(x[0] & 15));
          } else {
#pragma unroll
            for (int c = 0; c < C; ++c) atomicAdd(&h_sum[side][key * C + c], (unsigned)x[c]);
          }
          atomicAdd(&h_cnt[side][key], 1u);
        }
      }
    }
  }
  __syncthreads();
  if (!s_any) return;
  for (int i = threadIdx.x; i < 2 * NK; i += blockDim.x) {
    const int side = i / NK, key = i % NK;
    const unsigned cn = h_cnt[side][key];
    if (!cn) continue;
    atomicAdd((unsigned long long *)&s.kcnt[side * NKEYS_MAX + key], (unsigned long long)cn);
    if (BPC == 2) {
      const unsigned long long v = 16ull * key * cn + h_sum[side][key];
      atomicAdd((unsigned long long *)&s.ksum[side * NKEYS_MAX + key], v);
    } else {
      for (int c = 0; c < C; ++c)
        atomicAdd((unsigned long long *)&s.ksum[side * NKEYS_MAX + key * C + c],
                  (unsigned long long)h_sum[side][key * C + c]);
    }
  }
}

// The centre table of one side from its centroids (in LDS): per centre
// |c|^2 and -2 c, as float64 (exact: |c| < 2^24, so |c|^2 < 2^50).  The
// distance searches read it with uniform loads.
__device__ __forceinline__ void put_table(double *q, const long long *c, int K, int C) {
  const int k = threadIdx.x;
  if (k >= K) return;
  double n2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double v = i < C ? (double)c[k * C + i] : 0.0;
    n2 = fma(v, v, n2);
    q[k * 4 + 1 + i] = -2.0 * v;
  }
  q[k * 4] = n2;
}

// Initial centroids per side (one workgroup of 256 lanes, blockIdx.x =
// side): the distinct seeded keys in ascending order, the j-th centre taken
// from key number j * (m - 1) / (K - 1), rounded to fixed point
// (QGCWidget.cpp:2001-2039).
__global__ __launch_bounds__(256) void k_km_init(Scratch s, int C, int nkeys, int Kf, int Kb) {
  constexpr int KPL = NKEYS_MAX / 256;  // keys per lane
  __shared__ int s_cnt[256];
  __shared__ long long s_pick[KMAX_CLUSTERS];
  __shared__ long long s_cent[KMAX_CLUSTERS * 3];
  const int side = blockIdx.x, K = side ? Kb : Kf, tid = threadIdx.x;
  const long long *kc = s.kcnt + side * NKEYS_MAX;
  const long long *ks = s.ksum + side * NKEYS_MAX;
  long long *cent = s.cent + side * KMAX_CLUSTERS * 3;
  int mine = 0;
  for (int q = 0; q < KPL; ++q) {
    const int k = tid * KPL + q;
    if (k < nkeys && kc[k]) ++mine;
  }
  s_cnt[tid] = mine;
  __syncthreads();
  // inclusive scan over the 256 counts
  for (int o = 1; o < 256; o <<= 1) {
    const int v = tid >= o ? s_cnt[tid - o] : 0;
    __syncthreads();
    s_cnt[tid] += v;
    __syncthreads();
  }
  const int m = s_cnt[255];
  if (tid < K) s_pick[tid] = K > 1 ? (long long)tid * (m - 1) / (K - 1) : 0;
  if (tid < K * C) s_cent[tid] = 0;
  __syncthreads();
  if (m == 0) return;  // no seeds: reported by the host
  int r = s_cnt[tid] - mine;  // rank of this lane's first non-empty key
  for (int q = 0; q < KPL; ++q) {
    const int k = tid * KPL + q;
    if (k >= nkeys || !kc[k]) continue;
    for (int j = 0; j < K; ++j)
      if (s_pick[j] == r) {
        const long long nn = kc[k];
        for (int c = 0; c < C; ++c)
          s_cent[j * C + c] = (ks[k * C + c] * (1 << KM_FRAC) + nn / 2) / nn;
      }
    ++r;
  }
  __syncthreads();
  if (tid < K * C) {
    cent[tid] = s_cent[tid];
    s.cbuf[side * KMAX_CLUSTERS * 3 + tid] = s_cent[tid];  // c_0 for the first iteration
  }
  put_table(s.q + side * KMAX_CLUSTERS * 4, s_cent, K, C);
}

// The centre table read through the constant address space: its reads are
// uniform, so they become scalar loads into SGPRs, which the float64 FMAs
// take as operands (the table is written by an earlier launch only).
typedef const __attribute__((address_space(4))) double ctab_t;

// Nearest centre of a feature x (C channels), exactly, in float64: with
// X = 2^KM_FRAC x, |X - c|^2 = |X|^2 + (|c|^2 - 2 X.c), and every product
// and partial sum is an integer below 2^53 (|X|, |c| < 2^24 and C = 1 for
// 16-bit images, |X|, |c| < 2^16 for 8-bit ones), so the float64 arithmetic
// is exact and the lowest-index minimum is the integer one.  q is the side's
// centre table (put_table).
template <int C, typename Q>
__device__ __forceinline__ double near_part(const double *X, Q *q, int K, int *best) {
  double be = 0.0;
  int bk = 0;
  for (int k = 0; k < K; ++k) {
    double e = q[k * 4];
#pragma unroll
    for (int c = 0; c < C; ++c) e = fma(X[c], q[k * 4 + 1 + c], e);
    if (k == 0 || e < be) {
      be = e;
      bk = k;
    }
  }
  *best = bk;
  return be;
}

template <int C>
__device__ __forceinline__ void scaled_feature(const int *x, double *X, double *n2) {
  double t = 0.0;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    X[c] = (double)(x[c] << KM_FRAC);
    t = fma(X[c], X[c], t);
  }
  *n2 = t;
}

// The side's nearest-centre squared distance of P pixels at once: the centre
// loop is outer, so each centre's table entry is loaded (into SGPRs) once
// for all P, and only the minimum is kept (float64 min of exact integers).
template <int C, int P>
__device__ __forceinline__ void near_d2(const double (*X)[C], const double *n2, ctab_t *q, int K,
                                        double *D) {
  double m[P];
#pragma unroll
  for (int p = 0; p < P; ++p) m[p] = __builtin_huge_val();
  for (int k = 0; k < K; ++k) {
    double qk[C + 1];
#pragma unroll
    for (int c = 0; c <= C; ++c) qk[c] = q[k * 4 + c];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      double e = qk[0];
#pragma unroll
      for (int c = 0; c < C; ++c) e = fma(X[p][c], qk[1 + c], e);
      m[p] = fmin(m[p], e);
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) D[p] = n2[p] + m[p];
}

// The centre update (QGCWidget.cpp:2088-2101) from one iteration's partial
// sums (nb workgroups' rows of `part`), by a whole workgroup: the rounded
// mean of each non-empty cluster, else the old centre, into s_new.  Returns
// whether any centre moved.
__device__ bool km_update(const long long *part, int nb, const long long *old, int K, int C,
                          long long *s_new, unsigned long long *s_sum) {
  const int j = threadIdx.x, E = K * (C + 1);
  if (j < E) s_sum[j] = 0;
  __syncthreads();
  for (int i = j; i < nb * E; i += blockDim.x) {
    const unsigned long long v = (unsigned long long)part[(i / E) * (KMAX_CLUSTERS * 4) + i % E];
    if (v) atomicAdd(&s_sum[i % E], v);
  }
  __syncthreads();
  int changed = 0;
  if (j < K * C) {
    const int k = j / C, c = j % C;
    const long long nk = (long long)s_sum[k * (C + 1) + C];
    long long nc = old[j];
    if (nk) {
      const long long v = ((long long)s_sum[k * (C + 1) + c] * (1 << KM_FRAC) + nk / 2) / nk;
      changed = v != nc;
      nc = v;
    }
    s_new[j] = nc;
  }
  return __syncthreads_or(changed) != 0;
}

// The final centres of a side: stored with their table, and the done flag
// with the iteration count when they converged.
__device__ void km_finish(Scratch &s, int side, const long long *c, int K, int C, bool done,
                          int iters) {
  const int j = threadIdx.x;
  if (j < K * C) s.cent[side * KMAX_CLUSTERS * 3 + j] = c[j];
  put_table(s.q + side * KMAX_CLUSTERS * 4, c, K, C);
  if (j == 0 && done) {
    s.ctl[2 + side] = 1;
    s.ctl[4 + side] = iters;
  }
}

// Lloyd iteration `it` of a side (blockIdx.y), one launch per iteration.
// Each workgroup first forms the iteration's centres c_{it-1}: the initial
// ones for it = 1, else the update of iteration it - 1 from its partial
// sums, computed alike by every workgroup (workgroup 0 stores them for the
// next launch, in the other of two buffers).  When no centre moved, the side
// converged after it - 1 iterations: workgroup 0 stores the final centres
// and the done flag, and later launches return at once.  Then the
// assignment (graphcut_kmeans.cl:21-89): nearest centroid, lowest index on
// ties; each lane sums its seeds per cluster in registers (KB >= K
// clusters), loading 4 list entries ahead; the sums meet in a wave
// reduction, and the workgroup's are stored (no atomics) in its row of the
// iteration's partial sums.
template <int BPC, int C, int KB>
__global__ __launch_bounds__(256) void k_km_iter(Scratch s, int Kf, int Kb, int nb, int it) {
  const int side = blockIdx.y, K = side ? Kb : Kf;
  if (s.ctl[2 + side]) return;
  __shared__ long long s_c[KMAX_CLUSTERS * 3];
  __shared__ double s_q[KMAX_CLUSTERS * 4];
  __shared__ unsigned long long s_sum[KMAX_CLUSTERS * 4];
  constexpr int kPartSide = KM_GRID * KMAX_CLUSTERS * 4;
  const long long *old = s.cbuf + (((it - 1) & 1) * 2 + side) * KMAX_CLUSTERS * 3;
  if (it == 1) {
    if ((int)threadIdx.x < K * C) s_c[threadIdx.x] = old[threadIdx.x];
    __syncthreads();
  } else {
    const long long *prev = s.part + (long long)(((it - 1) & 1) * 2 + side) * kPartSide;
    if (!km_update(prev, nb, old, K, C, s_c, s_sum)) {
      if (blockIdx.x == 0) km_finish(s, side, s_c, K, C, true, it - 1);
      return;
    }
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < K * C)
    s.cbuf[((it & 1) * 2 + side) * KMAX_CLUSTERS * 3 + threadIdx.x] = s_c[threadIdx.x];
  put_table(s_q, s_c, K, C);
  __syncthreads();
  unsigned acc[KB][C + 1];
#pragma unroll
  for (int k = 0; k < KB; ++k)
#pragma unroll
    for (int c = 0; c <= C; ++c) acc[k][c] = 0;
  const long long n = min((long long)s.ctl[side], s.cap);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
    unsigned pk[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0 + u * stride;
      pk[u] = i < n ? (unsigned)s.list[side ? s.cap - 1 - i : i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + u * stride >= n) break;
      int x[C];
#pragma unroll
      for (int c = 0; c < C; ++c) x[c] = BPC == 2 ? (int)pk[u] : (int)((pk[u] >> (8 * c)) & 255u);
      double X[C], n2;
      scaled_feature<C>(x, X, &n2);
      int best;
      near_part<C>(X, (const double *)s_q, K, &best);
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const bool hit = k == best;
#pragma unroll
        for (int c = 0; c < C; ++c) acc[k][c] += hit ? (unsigned)x[c] : 0u;
        acc[k][C] += hit ? 1u : 0u;
      }
    }
  }
  // a lane holds at most n / (grid * 256) < 2^16 seeds, so its sums fit 32
  // bits; the wave's and the workgroup's are summed in 64 bits
  __shared__ unsigned long long s_red[4][KB * (C + 1)];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < KB; ++k) {
#pragma unroll
    for (int c = 0; c <= C; ++c) {
      if (k < K) {
        unsigned long long t = acc[k][c];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (lane_id() == 0) s_red[w][k * (C + 1) + c] = t;
      }
    }
  }
  __syncthreads();
  const int E = K * (C + 1);
  if ((int)threadIdx.x < E) {
    const unsigned long long t = s_red[0][threadIdx.x] + s_red[1][threadIdx.x] +
                                 s_red[2][threadIdx.x] + s_red[3][threadIdx.x];
    s.part[(long long)((it & 1) * 2 + side) * kPartSide + (long long)blockIdx.x * (KMAX_CLUSTERS * 4) +
           threadIdx.x] = (long long)t;
  }
}

// The update of the last iteration, one workgroup of 256 lanes per side
// (blockIdx.x): the final centres, and the done flag if none moved.
__global__ __launch_bounds__(256) void k_km_final(Scratch s, int C, int Kf, int Kb, int nb,
                                                  int iters) {
  const int side = blockIdx.x, K = side ? Kb : Kf;
  if (s.ctl[2 + side]) return;
  __shared__ long long s_c[KMAX_CLUSTERS * 3];
  __shared__ unsigned long long s_sum[KMAX_CLUSTERS * 4];
  constexpr int kPartSide = KM_GRID * KMAX_CLUSTERS * 4;
  const long long *old = s.cbuf + ((iters & 1) * 2 + side) * KMAX_CLUSTERS * 3;
  const long long *last = s.part + (long long)((iters & 1) * 2 + side) * kPartSide;
  const bool moved = km_update(last, nb, old, K, C, s_c, s_sum);
  km_finish(s, side, s_c, K, C, !moved, iters);
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

// A row segment's staged words held in registers between the global loads
// and the LDS stores: lane l holds words l + 256 j.  A workgroup loads the
// next segment's rows while it computes on the current one.
constexpr int SW = (ROWW + 255) / 256;
struct RowRegs {
  unsigned v[SW];
  int nw;  // words of the segment
};

__device__ __forceinline__ RowStage fetch_row(const unsigned char *img, long long total,
                                              long long row, int nx, int bpp, int a, int b,
                                              RowRegs &R) {
  const long long first = (row * nx + a) * bpp, last = (row * nx + b) * bpp;
  const long long w0 = first >> 2, w1 = (last + 3) >> 2;
  R.nw = (int)(w1 - w0);
#pragma unroll
  for (int j = 0; j < SW; ++j) {
    const long long w = w0 + threadIdx.x + 256 * j;
    unsigned v = 0;
    if (w < w1) {
      if ((w << 2) + 4 <= total) {
        v = ((const unsigned *)img)[w];
      } else {
        for (int k = 0; k < 4; ++k)
          if ((w << 2) + k < total) v |= (unsigned)img[(w << 2) + k] << (8 * k);
      }
    }
    R.v[j] = v;
  }
  RowStage s;
  s.off = (int)(first - (w0 << 2));
  return s;
}

__device__ __forceinline__ void put_row(const RowRegs &R, unsigned *words) {
#pragma unroll
  for (int j = 0; j < SW; ++j) {
    const int i = threadIdx.x + 256 * j;
    if (i < R.nw) words[i] = R.v[j];
  }
}

// Pixel p of a stage (p counted from the stage's first pixel), C channels.
// A 3-byte pixel is cut from the two aligned words it touches: byte loads
// the compiler merges into 16-bit LDS loads at odd addresses run on a slow
// unaligned path (measured 10x slower on the whole kernel).  The staged
// rows keep a spare word past their end for the second word.
template <int BPC, int C>
__device__ __forceinline__ void staged_px(const unsigned *words, const RowStage &s, int p,
                                          int *x) {
  const unsigned B = (unsigned)(s.off + p * (C * BPC));
  if (BPC == 1 && C == 3) {
    const unsigned lo = words[B >> 2], hi = words[(B >> 2) + 1];
    const unsigned v = __builtin_amdgcn_alignbyte(hi, lo, B & 3u);
    x[0] = (int)(v & 255u);
    x[C > 1 ? 1 : 0] = (int)((v >> 8) & 255u);
    x[C > 2 ? 2 : 0] = (int)((v >> 16) & 255u);
  } else {
    const unsigned char *b = (const unsigned char *)words + B;
#pragma unroll
    for (int c = 0; c < C; ++c) x[c] = BPC == 1 ? (int)b[c] : (int)((const unsigned short *)b)[c];
  }
}

// Squared difference of two pixels, exactly in 32 bits: |a - b| < 2^16 per
// channel, so each square is the low word of a 24-bit product, and the sum
// stays below 2^32 (three channels only for 8-bit images).
template <int C>
__device__ __forceinline__ unsigned sqdiff(const int *a, const int *b) {
  unsigned D = 0;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int t = a[c] - b[c];
    D += (unsigned)__mul24(t, t);
  }
  return D;
}

// Unary rounding: round(scale * D / den) = floor((scale * D + den / 2) / den)
// for an integer D (den = range^2 2^16 is even).  y = D r + 1/2 with r =
// scale / den in float64, in one rounding, lies within 2^-21 of the exact
// quotient (which is at most 2^30 + 1/2, the terminal-capacity bound the
// parameters are checked against); unless y falls within 2^-16 of an
// integer, floor(y) is the exact floor, else the integer division decides.
struct Unary {
  long long den, half, scale;
  double r;
};

__device__ __forceinline__ int unary_round(double D, const Unary &u) {
  const double y = fma(D, u.r, 0.5);
  const double f = floor(y);
  const double fr = y - f;
  if (fr >= 0x1p-16 && fr <= 1.0 - 0x1p-16) return (int)f;
  return (int)div_floor(u.scale * (long long)D + u.half, u.den);
}

// The rows a segment of row r = y + ny z reads for its forward edges: its own
// row (x0 .. x1, plus the pixel right of it), the next row of the slice
// (x0 - 1 .. x1 + 1) and, in 3D, the same row of the next slice.
struct Stages {
  RowStage own, next, up;
  int a_own, a_next, a_up;  // first staged pixel of each
};

// Row segments per workgroup of the stencil kernels: consecutive segments,
// the next one's rows loaded while the current one is processed.  Short-lived
// workgroups let the other stream's small kernels in between.
constexpr int kItemsPerBlock = 4;

// A segment of the row-segment list: row r = y + ny z, pixels [x0, x1).
struct Seg {
  long long r;
  int x0, x1, y, z;
};

__device__ __forceinline__ Seg seg_of(long long item, int segs, int nx, int ny) {
  Seg g;
  g.r = item / segs;
  g.x0 = (int)(item % segs) * SEG;
  g.x1 = min(nx, g.x0 + SEG);
  g.y = (int)(g.r % ny);
  g.z = (int)(g.r / ny);
  return g;
}

// The segment after g in the list (no division).
__device__ __forceinline__ Seg seg_next(Seg g, int nx, int ny) {
  g.x0 += SEG;
  if (g.x0 >= nx) {
    g.x0 = 0;
    ++g.r;
    if (++g.y == ny) {
      g.y = 0;
      ++g.z;
    }
  }
  g.x1 = min(nx, g.x0 + SEG);
  return g;
}

// The rows of stage_rows, into registers.
template <int CONN>
struct RowsInFlight {
  Stages S;
  RowRegs own, next, up;
  __device__ __forceinline__ void fetch(const unsigned char *img, long long total, int nx, int ny,
                                        int nz, int bpp, const Seg &g) {
    S.a_own = g.x0;
    S.own = fetch_row(img, total, g.r, nx, bpp, g.x0, min(nx, g.x1 + 1), own);
    S.a_next = max(0, g.x0 - 1);
    next.nw = 0;
    if (g.y + 1 < ny) S.next = fetch_row(img, total, g.r + 1, nx, bpp, S.a_next, min(nx, g.x1 + 1), next);
    S.a_up = g.x0;
    up.nw = 0;
    if (CONN == 6 && g.z + 1 < nz) S.up = fetch_row(img, total, g.r + ny, nx, bpp, g.x0, g.x1, up);
  }
  __device__ __forceinline__ void put(unsigned *w_own, unsigned *w_next, unsigned *w_up) const {
    put_row(own, w_own);
    put_row(next, w_next);
    if (CONN == 6) put_row(up, w_up);
  }
};

// The far pixel of forward edge j of node (x, y, z), from LDS.
template <int CONN, int BPC, int C>
__device__ __forceinline__ void edge_px(const Stages &S, const unsigned *w_own,
                                        const unsigned *w_next, const unsigned *w_up, int x,
                                        int j, int *nb) {
  switch (2 * j) {
    case 0: staged_px<BPC, C>(w_own, S.own, x + 1 - S.a_own, nb); break;
    case 2: staged_px<BPC, C>(w_next, S.next, x - S.a_next, nb); break;
    case 4:
      if (CONN == 6) staged_px<BPC, C>(w_up, S.up, x - S.a_up, nb);
      else staged_px<BPC, C>(w_next, S.next, x + 1 - S.a_next, nb);
      break;
    default: staged_px<BPC, C>(w_next, S.next, x - 1 - S.a_next, nb); break;
  }
}

// Unary terms, 4 pixels per lane over the image in its byte order (pixel
// p = x + nx r): one load of the 4 pixels' bytes (4 * C * BPC, a multiple of
// 4) and one of their seeds, the centre search of the 4 together (each
// centre's table entry loaded once), and the results written to rows of
// pitch px, one 16-byte store per plane where the 4 lie in one row at a
// multiple of 4.  cs = round(scale * d2_bg / den), ct = round(scale * d2_fg /
// den), with d2 the squared distance to the side's nearest centre; seeds get
// (2^30, 0) / (0, 2^30).  n < 2^31.
template <int BPC, int C>
__global__ __launch_bounds__(256) void k_unary(const void *img, const unsigned char *seeds, int n,
                                               int nx, int px, Scratch s, int Kf, int Kb, Unary u,
                                               int *cs, int *ct) {
  constexpr int W = C * BPC;  // words of 4 pixels
  ctab_t *qf = (ctab_t *)s.q, *qb = (ctab_t *)(s.q + KMAX_CLUSTERS * 4);
  const bool al = ((size_t)img & 3) == 0 && ((size_t)seeds & 3) == 0;
  const bool vec_out = (px & 3) == 0 && (((size_t)cs | (size_t)ct) & 15) == 0;
  const unsigned char *im = (const unsigned char *)img;
  for (unsigned q = blockIdx.x * 256u + threadIdx.x; q < (unsigned)(n + 3) / 4u;
       q += gridDim.x * 256u) {
    const int p0 = (int)(q * 4u);
    const int np = min(4, n - p0);
    unsigned w[W], sw = 0;
    if (al && np == 4) {
#pragma unroll
      for (int k = 0; k < W; ++k) w[k] = ((const unsigned *)im)[(long long)p0 * W / 4 + k];
      sw = *(const unsigned *)(seeds + p0);
    } else {
#pragma unroll
      for (int k = 0; k < W; ++k) w[k] = 0;
      for (int k = 0; k < 4 * W; ++k)
        if (k < np * W) w[k >> 2] |= (unsigned)im[(long long)p0 * W + k] << (8 * (k & 3));
      for (int k = 0; k < np; ++k) sw |= (unsigned)seeds[p0 + k] << (8 * k);
    }
    int f[4][C];
    double X[4][C], n2[4], Df[4], Db[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int k = j * C + c;  // channel value index
        f[j][c] = BPC == 1 ? (int)((w[k >> 2] >> (8 * (k & 3))) & 255u)
                           : (int)((w[k >> 1] >> (16 * (k & 1))) & 65535u);
      }
      scaled_feature<C>(f[j], X[j], &n2[j]);
    }
    near_d2<C, 4>(X, n2, qf, Kf, Df);
    near_d2<C, 4>(X, n2, qb, Kb, Db);
    int a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned sd = (sw >> (8 * j)) & 255u;
      a[j] = 1 << 30;
      b[j] = 0;
      if (sd == 2u) {
        a[j] = 0;
        b[j] = 1 << 30;
      } else if (sd != 1u) {
        a[j] = unary_round(Db[j], u);
        b[j] = unary_round(Df[j], u);
      }
    }
    const unsigned r = (unsigned)p0 / (unsigned)nx;
    const int x = p0 - (int)(r * (unsigned)nx);
    const long long o = (long long)r * px + x;
    if (vec_out && np == 4 && (x & 3) == 0 && x + 4 <= nx) {
      *(int4 *)(cs + o) = make_int4(a[0], a[1], a[2], a[3]);
      *(int4 *)(ct + o) = make_int4(b[0], b[1], b[2], b[3]);
    } else {
      long long oo = o;
      int xx = x;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < np) {
          cs[oo] = a[j];
          ct[oo] = b[j];
        }
        if (++xx == nx) {
          xx = 0;
          oo += px - nx + 1;
        } else {
          ++oo;
        }
      }
    }
  }
}

// Pair sums: per edge group the sum of squared neighbour differences (0: x/y
// and the 2D diagonals, and z unless it forms its own group; 1: z when
// separate), one term per undirected edge; the edge counts have a closed form
// (edge_counts).  Row segments are staged in LDS, the next one loaded while
// the current one is summed.
template <int CONN, int BPC, int C>
__global__ __launch_bounds__(256) void k_pair_sums(const void *img, int nx, int ny, int nz,
                                                   int zsep, long long *sums) {
  __shared__ unsigned long long buf[4][2];
  __shared__ unsigned w_own[ROWW], w_next[ROWW], w_up[CONN == 6 ? ROWW : 1];
  unsigned long long s0 = 0, s1 = 0;
  const int bpp = C * BPC;
  const long long rows = (long long)ny * nz, total = rows * nx * bpp;
  const int segs = (nx + SEG - 1) / SEG;
  const long long nitems = rows * segs;
  const unsigned char *im = (const unsigned char *)img;
  RowsInFlight<CONN> F;
  const long long first = (long long)blockIdx.x * kItemsPerBlock;
  const long long last = min(nitems, first + kItemsPerBlock);
  Seg g = seg_of(first, segs, nx, ny);
  if (first < last) F.fetch(im, total, nx, ny, nz, bpp, g);
  for (long long item = first; item < last; ++item, g = seg_next(g, nx, ny)) {
    const Stages S = F.S;
    __syncthreads();  // the previous item's reads of the stages are done
    F.put(w_own, w_next, w_up);
    __syncthreads();
    if (item + 1 < last) F.fetch(im, total, nx, ny, nz, bpp, seg_next(g, nx, ny));
    // per item at most 4 * 4 * 3 * 255^2 < 2^32 for 8-bit images
    unsigned i0 = 0, i1 = 0;
    for (int x = g.x0 + threadIdx.x; x < g.x1; x += blockDim.x) {
      int f[C];
      staged_px<BPC, C>(w_own, S.own, x - S.a_own, f);
#pragma unroll
      for (int j = 0; j < Fwd<CONN>::N; ++j) {
        if (!Fwd<CONN>::inside(j, x, g.y, g.z, nx, ny, nz)) continue;
        int nb[C];
        edge_px<CONN, BPC, C>(S, w_own, w_next, w_up, x, j, nb);
        const unsigned D = sqdiff<C>(f, nb);
        const bool g1 = CONN == 6 && j == 2 && zsep;
        if (BPC == 1) {
          if (g1) i1 += D;
          else i0 += D;
        } else {
          if (g1) s1 += D;
          else s0 += D;
        }
      }
    }
    s0 += i0;
    s1 += i1;
  }
  unsigned long long vals[2] = {s0, s1};
  const int w = threadIdx.x >> 6;
  for (int j = 0; j < 2; ++j) {
    unsigned long long x = vals[j];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane_id() == 0) buf[w][j] = x;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += buf[i][threadIdx.x];
    // workgroup b adds to slot b % NSUM: device atomics on one address
    // serialise, and there is a workgroup per kItemsPerBlock row segments
    if (t) atomicAdd((unsigned long long *)&sums[2 * (blockIdx.x % NSUM) + threadIdx.x], t);
  }
}

// Undirected edges per group, in closed form: group 0 holds the x and y
// edges, the 2D diagonals and, unless separate, the z edges; group 1 the
// separate z edges.  Every instance with nz > 1 and 4 or 8 neighbours is a
// stack of independent slices.
void edge_counts(const InstDesc &I, int zsep, long long P[2]) {
  const long long nx = I.nx, ny = I.ny, nz = I.nz;
  long long e0 = (nx - 1) * ny * nz + nx * (ny - 1) * nz;
  long long e1 = 0;
  if (I.conn == 8) e0 += 2 * (nx - 1) * (ny - 1) * nz;
  if (I.conn == 6) (zsep ? e1 : e0) += nx * ny * (nz - 1);
  P[0] = e0;
  P[1] = e1;
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

// w(D) of one table: the guess and, in the same step, the two entries that
// confirm it (T[w] < D <= T[w - 1]); only where they do not (a guess off by
// one near a threshold) the exact search runs.
__device__ __forceinline__ int weight(unsigned D, const unsigned *t, int n, float lam, float inv) {
  if (n == 0) return 0;
  int w = (int)(lam * __expf(-(float)D * inv) + 0.5f);
  w = max(0, min(w, n));
  const unsigned th = t[min(w, n - 1)], tl = t[max(w - 1, 0)];
  const bool up = w < n && D <= th;  // the weight exceeds w
  const bool dn = w > 0 && D > tl;   // the weight is below w
  if (up || dn) {
    while (w < n && D <= t[w]) ++w;
    while (w > 0 && D > t[w - 1]) --w;
  }
  return w;
}

template <int CONN, int BPC, int C>
__global__ __launch_bounds__(256) void k_pairwise(const void *img, int nx, int ny, int nz, int px,
                                                  const unsigned *T, PairParams pp, int *c0,
                                                  long long stride) {
  __shared__ unsigned s_T[3 * TMAX];
  __shared__ unsigned w_own[ROWW], w_next[ROWW], w_up[CONN == 6 ? ROWW : 1];
  for (int i = threadIdx.x; i < 3 * TMAX; i += blockDim.x) s_T[i] = T[i];
  const int bpp = C * BPC;
  const long long rows = (long long)ny * nz, total = rows * nx * bpp;
  const long long splane = (long long)px * ny;
  const int segs = (nx + SEG - 1) / SEG;
  const long long nitems = rows * segs;
  const unsigned char *im = (const unsigned char *)img;
  RowsInFlight<CONN> F;
  const long long first = (long long)blockIdx.x * kItemsPerBlock;
  const long long last = min(nitems, first + kItemsPerBlock);
  Seg g = seg_of(first, segs, nx, ny);
  if (first < last) F.fetch(im, total, nx, ny, nz, bpp, g);
  for (long long item = first; item < last; ++item, g = seg_next(g, nx, ny)) {
    const long long r = g.r;
    const int x0 = g.x0, x1 = g.x1, y = g.y, z = g.z;
    const Stages S = F.S;
    __syncthreads();
    F.put(w_own, w_next, w_up);
    __syncthreads();
    if (item + 1 < last) F.fetch(im, total, nx, ny, nz, bpp, seg_next(g, nx, ny));
    for (int x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
      const long long sv = r * px + x;
      int f[C];
      staged_px<BPC, C>(w_own, S.own, x - S.a_own, f);
#pragma unroll
      for (int j = 0; j < Fwd<CONN>::N; ++j) {
        const int d = Fwd<CONN>::dir(j);
        int w = 0;
        if (Fwd<CONN>::inside(j, x, y, z, nx, ny, nz)) {
          int nb[C];
          edge_px<CONN, BPC, C>(S, w_own, w_next, w_up, x, j, nb);
          const unsigned D = sqdiff<C>(f, nb);
          const int tb = Fwd<CONN>::table(j);
          w = weight(D, s_T + tb * TMAX, pp.n[tb], pp.lam[tb], pp.inv2s2[tb]);
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

// The streams of one model call: the caller's (st) and, when the graph has
// one, a second stream for the pair sums and the pairwise weights, which
// need neither the seeds nor the k-means.  join() orders st after the second
// stream's work; every return after the fork goes through it.
struct Streams {
  hipStream_t st, aux;
  hipEvent_t fork, join_ev, ready;
  void *pin;  // pinned host words of the second stream's copies
  bool forked = false;
  hipStream_t side() const { return aux ? aux : st; }
  hipError_t fork_now() {
    if (!aux) return hipSuccess;
    hipError_t e = hipEventRecord(fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(aux, fork, 0);
    forked = e == hipSuccess;
    return e;
  }
  hipError_t join() {
    if (!forked) return hipSuccess;
    forked = false;
    hipError_t e = hipEventRecord(join_ev, aux);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, join_ev, 0);
    return e;
  }
  ~Streams() { (void)join(); }
};

template <int CONN, int BPC, int C>
void launch_pair_sums(int grid, const InstDesc &I, const void *img, int zsep, const Scratch &s,
                      hipStream_t st) {
  k_pair_sums<CONN, BPC, C><<<grid, 256, 0, st>>>(img, I.nx, I.ny, I.nz, zsep, s.sums);
}

template <int CONN, int BPC, int C>
void launch_pairwise(int grid, const InstDesc &I, const void *img, const Scratch &s,
                     const PairParams &pp, int *c0, long long stride, hipStream_t st) {
  k_pairwise<CONN, BPC, C><<<grid, 256, 0, st>>>(img, I.nx, I.ny, I.nz, I.px, s.T, pp, c0, stride);
}

// The model of one instance from device-resident image and seeds, with the
// given scratch.  Writes the instance's cs, ct and c0 planes.  The pair sums
// and the pairwise weights run on the second stream (when there is one)
// beside the seed scan, the k-means and the unary terms.
template <int BPC, int C>
int run_model(const InstDesc &I, const void *img, const unsigned char *seeds,
              const jfc_model_params *prm, jfc_model_stats &ST, Scratch s, int *cs, int *ct,
              int *c0, long long stride, Streams &S, bool *planes_written) {
  *planes_written = false;
  const hipStream_t st = S.st;
  const long long n = (long long)I.nx * I.ny * I.nz;
  const long long rows = (long long)I.ny * I.nz;
  const int nkeys = BPC == 2 ? 4096 : 255 * C + 1;
  const int iters = prm->kmeans_iters;
  const bool d3 = I.conn == 6;
  const int zsep = d3 && prm->z_separate ? 1 : 0;
  // kItemsPerBlock row segments of SEG pixels per workgroup
  const int rgrid = (int)((rows * ((I.nx + SEG - 1) / SEG) + kItemsPerBlock - 1) / kItemsPerBlock);
  k_model_clear<<<64, 256, 0, st>>>(s, iters);
  k_seed_scan<BPC, C><<<grid1((n + 15) / 16, kScanGrid), 256, 0, st>>>(img, seeds, n, s);
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
  // the pair sums on the second stream from here on: launched earlier, they
  // would hold the copy of the seed counts back behind their workgroups
  MODEL_HIP(S.fork_now());
  const hipStream_t sb = S.side();
  if (I.conn == 6) launch_pair_sums<6, BPC, C>(rgrid, I, img, zsep, s, sb);
  else if (I.conn == 8) launch_pair_sums<8, BPC, C>(rgrid, I, img, 0, s, sb);
  else launch_pair_sums<4, BPC, C>(rgrid, I, img, 0, s, sb);
  // with a second stream, its sums come back to pinned words while the
  // host issues the k-means and the unary terms
  long long host_slots[NSUM * 2];
  long long *slots = S.aux ? (long long *)S.pin : host_slots;
  if (S.aux) {
    MODEL_HIP(hipMemcpyAsync(slots, s.sums, NSUM * 2 * sizeof(long long), hipMemcpyDeviceToHost, sb));
    MODEL_HIP(hipEventRecord(S.ready, sb));
  }
  // k-means on the device: iteration `it` assigns and updates; a side whose
  // centres stop moving is done and its later launches return at once
  k_km_init<<<2, 256, 0, st>>>(s, C, nkeys, prm->k_fg, prm->k_bg);
  // up to KM_GRID workgroups per side, about 4 seeds per lane
  const long long most = std::max(ctl[0], ctl[1]);
  const int nb = (int)std::min<long long>(KM_GRID, std::max<long long>(1, (most + 1023) / 1024));
  const dim3 ag(nb, 2);
  const int kmax = std::max(prm->k_fg, prm->k_bg);
  for (int it = 1; it <= iters; ++it) {
    if (kmax <= 4) k_km_iter<BPC, C, 4><<<ag, 256, 0, st>>>(s, prm->k_fg, prm->k_bg, nb, it);
    else if (kmax <= 8) k_km_iter<BPC, C, 8><<<ag, 256, 0, st>>>(s, prm->k_fg, prm->k_bg, nb, it);
    else k_km_iter<BPC, C, 16><<<ag, 256, 0, st>>>(s, prm->k_fg, prm->k_bg, nb, it);
  }
  if (iters > 0) k_km_final<<<2, 256, 0, st>>>(s, C, prm->k_fg, prm->k_bg, nb, iters);
  // unary terms from the final centres, straight into the stored rows
  Unary u;
  u.den = prm->range * prm->range * (1LL << (2 * KM_FRAC));
  u.half = u.den / 2;
  u.scale = prm->unary_scale;
  u.r = (double)u.scale / (double)u.den;
  const int ugrid = (int)std::max<long long>(1, (n + 1023) / 1024);  // 4 pixels per lane
  k_unary<BPC, C><<<ugrid, 256, 0, st>>>(img, seeds, (int)n, I.nx, I.px, s, prm->k_fg, prm->k_bg, u,
                                          cs + I.node_off, ct + I.node_off);
  *planes_written = true;
  // the thresholds on the host, from the pair sums
  long long sums[4] = {0, 0, 0, 0};
  if (S.aux) {
    MODEL_HIP(hipEventSynchronize(S.ready));
  } else {
    MODEL_HIP(hipMemcpyAsync(slots, s.sums, NSUM * 2 * sizeof(long long), hipMemcpyDeviceToHost, sb));
    MODEL_HIP(hipStreamSynchronize(sb));
  }
  for (int i = 0; i < NSUM; ++i) {
    sums[0] += slots[2 * i];
    sums[2] += slots[2 * i + 1];
  }
  long long P2[2];
  edge_counts(I, zsep, P2);
  sums[1] = P2[0];
  sums[3] = P2[1];
  ST.pair_sum[0] = sums[0]; ST.pair_count[0] = sums[1];
  ST.pair_sum[1] = sums[2]; ST.pair_count[1] = sums[3];
  static_assert(sizeof(unsigned) == 4, "thresholds are 32-bit");
  unsigned host_tab[3 * TMAX];
  // the thresholds cross from pinned words after the slots (read by now)
  unsigned *Tab = S.aux ? (unsigned *)((char *)S.pin + NSUM * 2 * sizeof(long long)) : host_tab;
  std::memset(Tab, 0, 3 * TMAX * sizeof(unsigned));
  PairParams pp;
  const double lam[3] = {prm->lambda_xy, prm->lambda_xy / std::sqrt(2.0), prm->lambda_z};
  const int zg = zsep ? 2 : 0;  // sums of the z table: own group or the x/y one
  const long long Sg[3] = {sums[0], sums[0], sums[zg]}, Pg[3] = {sums[1], sums[1], sums[zg + 1]};
  for (int t = 0; t < 3; ++t) {
    pp.n[t] = thresholds(lam[t], Sg[t], Pg[t], Tab + t * TMAX);
    pp.lam[t] = (float)lam[t];
    pp.inv2s2[t] = (Sg[t] == 0 || Pg[t] == 0) ? 0.0f : (float)((double)Pg[t] / (2.0 * (double)Sg[t]));
  }
  MODEL_HIP(hipMemcpyAsync(s.T, Tab, 3 * TMAX * sizeof(unsigned), hipMemcpyHostToDevice, sb));
  int *pc0 = c0 + I.node_off;
  if (I.conn == 6) launch_pairwise<6, BPC, C>(rgrid, I, img, s, pp, pc0, stride, sb);
  else if (I.conn == 8) launch_pairwise<8, BPC, C>(rgrid, I, img, s, pp, pc0, stride, sb);
  else launch_pairwise<4, BPC, C>(rgrid, I, img, s, pp, pc0, stride, sb);
  MODEL_HIP(hipGetLastError());
  MODEL_HIP(S.join());
  long long cent[2][KMAX_CLUSTERS * 3];
  MODEL_HIP(hipMemcpyAsync(cent, s.cent, sizeof(cent), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipMemcpyAsync(ctl, s.ctl, 6 * sizeof(int), hipMemcpyDeviceToHost, st));
  MODEL_HIP(hipStreamSynchronize(st));
  ST.iters_fg = ctl[4];
  ST.iters_bg = ctl[5];
  for (int i = 0; i < prm->k_fg * C; ++i) ST.centroids_fg[i] = cent[0][i];
  for (int i = 0; i < prm->k_bg * C; ++i) ST.centroids_bg[i] = cent[1][i];
  return JFC_OK;
}

int model_common(jfc_graph *g, int inst, const void *img, int C, int bpc,
                 const unsigned char *seeds, const jfc_model_params *prm,
                 jfc_model_stats *stats, void *scratch, int *list, long long cap,
                 bool create_aux) {
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
  Streams S;
  S.st = graph_stream(g);
  GraphAux A;
  if ((rc = graph_aux(g, create_aux, kAuxPinBytes, &A))) return rc;
  S.aux = A.st;
  S.fork = A.fork;
  S.join_ev = A.join;
  S.ready = A.ready;
  S.pin = A.pin;
  if (bpc == 2) rc = run_model<2, 1>(I, img, seeds, prm, ST, s, cs, ct, c0, stride, S, &written);
  else if (C == 3) rc = run_model<1, 3>(I, img, seeds, prm, ST, s, cs, ct, c0, stride, S, &written);
  else rc = run_model<1, 1>(I, img, seeds, prm, ST, s, cs, ct, c0, stride, S, &written);
  if (hipError_t e = S.join()) {
    if (!rc) rc = model_fail(JFC_ERR_DEVICE, std::string("stream join: ") + hipGetErrorString(e));
  }
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
  return model_common(g, inst, img, channels, bpc, sd, prm, stats, scratch, list, cap, true);
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
                      (int *)((unsigned char *)base + kSmallBytes), kStaticSeeds, false);
}
