/*
 * jfc_oracle.c -- CPU oracle for the JF-Cut grid max-flow path.
 *
 * TEST INFRASTRUCTURE ONLY (see jfc_oracle.h): the checker, never the thing
 * measured or shipped.  Written from the mathematics of push-relabel
 * (Goldberg & Tarjan 1988) and from the semantics of the reference kernels;
 * citations are relative to the JF-Cut source tree.
 */
#include "jfc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 0x3fffffff

/* Direction tables.  2D: E W S N SE NW SW NE; 3D: +x -x +y -y +z -z.
 * Opposite of d is d^1. */
static const int k_d2[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                               {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
static const int k_d3[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                               {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

int orc_ndirs(int conn) {
  return (conn == 4 || conn == 8 || conn == 6) ? conn : 0;
}

void orc_dir(int conn, int d, int *dx, int *dy, int *dz) {
  if (conn == 6) {
    *dx = k_d3[d][0]; *dy = k_d3[d][1]; *dz = k_d3[d][2];
  } else {
    *dx = k_d2[d][0]; *dy = k_d2[d][1]; *dz = 0;
  }
}

typedef struct {
  int nx, ny, nz, K;
  int64_t N;
  int64_t off[8];
  uint8_t *mask; /* bit d set iff the neighbour in direction d exists */
} grid_t;

static int grid_init(grid_t *g, int nx, int ny, int nz, int conn) {
  g->nx = nx; g->ny = ny; g->nz = nz;
  g->K = orc_ndirs(conn);
  g->N = (int64_t)nx * ny * nz;
  if (g->K == 0 || g->N <= 0) return -1;
  for (int d = 0; d < g->K; ++d) {
    int dx, dy, dz;
    orc_dir(conn, d, &dx, &dy, &dz);
    g->off[d] = dx + (int64_t)dy * nx + (int64_t)dz * nx * ny;
  }
  g->mask = (uint8_t *)malloc((size_t)g->N);
  if (!g->mask) return -1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        uint8_t m = 0;
        for (int d = 0; d < g->K; ++d) {
          int dx, dy, dz;
          orc_dir(conn, d, &dx, &dy, &dz);
          int X = x + dx, Y = y + dy, Z = z + dz;
          if (X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz)
            m |= (uint8_t)(1u << d);
        }
        g->mask[x + (int64_t)nx * (y + (int64_t)ny * z)] = m;
      }
  return 0;
}

/* Exact distance labels: BFS from deficit nodes (h = 0) backwards along
 * residual arcs; nodes that cannot reach a deficit get ORC_INF.  Follows the
 * init + relaxation of graphcut_global_relabel.cl:19-48 and :50-124, done
 * here as one exact breadth-first search. */
static void global_relabel(const grid_t *g, const int32_t *e, const int32_t *r,
                           int32_t *h, int64_t *queue) {
  const int64_t N = g->N;
  int64_t head = 0, tail = 0;
  for (int64_t v = 0; v < N; ++v) {
    if (e[v] < 0) { h[v] = 0; queue[tail++] = v; }
    else h[v] = ORC_INF;
  }
  while (head < tail) {
    int64_t v = queue[head++];
    uint8_t m = g->mask[v];
    for (int d = 0; d < g->K; ++d) {
      if (!(m & (1u << d))) continue;
      int64_t u = v + g->off[d];
      /* arc u -> v runs in direction d^1 from u */
      if (h[u] == ORC_INF && r[(int64_t)(d ^ 1) * N + u] > 0) {
        h[u] = h[v] + 1;
        queue[tail++] = u;
      }
    }
  }
}

int orc_canonical_cut(int nx, int ny, int nz, int conn, const int32_t *e,
                      const int32_t *res, uint8_t *labels) {
  grid_t g;
  if (grid_init(&g, nx, ny, nz, conn)) return -1;
  const int64_t N = g.N;
  int64_t *queue = (int64_t *)malloc((size_t)N * sizeof(int64_t));
  if (!queue) { free(g.mask); return -1; }
  int64_t head = 0, tail = 0;
  int ok = 1;
  /* graphcut_init_cut: object = CHAR_MAX where excess > 0 */
  for (int64_t v = 0; v < N; ++v) {
    labels[v] = e[v] > 0;
    if (labels[v]) queue[tail++] = v;
  }
  /* graphcut_compute_cut: a node joins when a member has residual toward it;
   * reaching a deficit node means an augmenting path still exists. */
  while (head < tail) {
    int64_t u = queue[head++];
    uint8_t m = g.mask[u];
    for (int d = 0; d < g.K; ++d) {
      if (!(m & (1u << d))) continue;
      int64_t v = u + g.off[d];
      if (!labels[v] && res[(int64_t)d * N + u] > 0) {
        labels[v] = 1;
        if (e[v] < 0) ok = 0;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
  free(g.mask);
  return ok;
}

/* graphcut_init_tag + graphcut_compute_tag (graphcut_compute_tag.cl:19-49,
 * 51-140 2D, :142-242 3D): foreground seeds are tagged; an object node with a
 * tagged axis neighbour (4 in 2D whatever the graph's connectivity, 6 in 3D)
 * becomes tagged, to the fixpoint.  Breadth-first over a queue here. */
int orc_tag(int nx, int ny, int nz, const uint8_t *fg, const uint8_t *object,
            uint8_t *tag) {
  grid_t g;
  if (grid_init(&g, nx, ny, nz, nz > 1 ? 6 : 4)) return -1;
  const int64_t N = g.N;
  int64_t *queue = (int64_t *)malloc((size_t)N * sizeof(int64_t));
  if (!queue) { free(g.mask); return -1; }
  int64_t head = 0, tail = 0;
  for (int64_t v = 0; v < N; ++v) {
    tag[v] = fg[v] != 0;
    if (tag[v]) queue[tail++] = v;
  }
  while (head < tail) {
    const int64_t u = queue[head++];
    for (int d = 0; d < g.K; ++d) {
      if (!(g.mask[u] & (1u << d))) continue;
      const int64_t v = u + g.off[d];
      if (!tag[v] && object[v]) {
        tag[v] = 1;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
  free(g.mask);
  return 0;
}

int64_t orc_energy(int nx, int ny, int nz, int conn, const int32_t *cs,
                   const int32_t *ct, const int32_t *caps,
                   const uint8_t *labels) {
  grid_t g;
  if (grid_init(&g, nx, ny, nz, conn)) return -1;
  const int64_t N = g.N;
  int64_t E = 0;
  for (int64_t v = 0; v < N; ++v) {
    if (labels[v]) {
      E += ct[v];
      uint8_t m = g.mask[v];
      for (int d = 0; d < g.K; ++d)
        if ((m & (1u << d)) && !labels[v + g.off[d]]) E += caps[(int64_t)d * N + v];
    } else {
      E += cs[v];
    }
  }
  free(g.mask);
  return E;
}

int64_t orc_maxflow(int nx, int ny, int nz, int conn, const int32_t *cs,
                    const int32_t *ct, const int32_t *caps, int32_t *e_out,
                    int32_t *res_out, uint8_t *labels_out,
                    int *certificate_ok) {
  grid_t g;
  if (grid_init(&g, nx, ny, nz, conn)) return -1;
  const int64_t N = g.N;
  const int K = g.K;
  int32_t *e = (int32_t *)malloc((size_t)N * 4);
  int32_t *h = (int32_t *)malloc((size_t)N * 4);
  int32_t *r = (int32_t *)malloc((size_t)N * 4 * K);
  int64_t *queue = (int64_t *)malloc((size_t)N * 8);
  int64_t *fifo = (int64_t *)malloc((size_t)N * 8);
  uint8_t *inq = (uint8_t *)calloc((size_t)N, 1);
  if (!e || !h || !r || !queue || !fifo || !inq) {
    free(e); free(h); free(r); free(queue); free(fifo); free(inq); free(g.mask);
    return -1;
  }
  /* Folding of the terminal arcs (QGCWidget.cpp:3383-3399, BK303 add_tweights):
   * e = cs - ct, and min(cs, ct) flows straight through the node. */
  int64_t flow_const = 0, deficit0 = 0;
  for (int64_t v = 0; v < N; ++v) {
    e[v] = cs[v] - ct[v];
    flow_const += cs[v] < ct[v] ? cs[v] : ct[v];
    if (e[v] < 0) deficit0 -= e[v];
  }
  for (int d = 0; d < K; ++d)
    for (int64_t v = 0; v < N; ++v)
      r[(int64_t)d * N + v] = (g.mask[v] & (1u << d)) ? caps[(int64_t)d * N + v] : 0;

  global_relabel(&g, e, r, h, queue);
  /* FIFO of active nodes: e > 0 and finite height (graphcut_push_relabel.cl:80) */
  int64_t qh = 0, qt = 0, qn = 0;
#define PUSHQ(v) do { fifo[qt] = (v); qt = (qt + 1 == N) ? 0 : qt + 1; ++qn; inq[v] = 1; } while (0)
  for (int64_t v = 0; v < N; ++v)
    if (e[v] > 0 && h[v] < ORC_INF) PUSHQ(v);
  int64_t relabels = 0;
  const int64_t gr_period = N / 2 + 64;
  while (qn > 0) {
    int64_t u = fifo[qh];
    qh = (qh + 1 == N) ? 0 : qh + 1;
    --qn;
    inq[u] = 0;
    if (h[u] >= ORC_INF) continue;
    uint8_t m = g.mask[u];
    while (e[u] > 0) {
      /* push along admissible arcs: residual > 0 and h(u) == h(v) + 1
       * (graphcut_push_relabel.cl:104-180) */
      for (int d = 0; d < K && e[u] > 0; ++d) {
        if (!(m & (1u << d))) continue;
        int64_t v = u + g.off[d];
        int32_t *ruv = &r[(int64_t)d * N + u];
        if (*ruv > 0 && h[u] == h[v] + 1) {
          int32_t t = e[u] < *ruv ? e[u] : *ruv;
          *ruv -= t;
          r[(int64_t)(d ^ 1) * N + v] += t;
          e[u] -= t;
          e[v] += t;
          if (e[v] > 0 && !inq[v] && h[v] < ORC_INF) PUSHQ(v);
        }
      }
      if (e[u] <= 0) break;
      /* relabel: min over residual arcs of h_nb + 1 (graphcut_push_relabel.cl:82-87) */
      int32_t hm = ORC_INF;
      for (int d = 0; d < K; ++d) {
        if (!(m & (1u << d))) continue;
        int64_t v = u + g.off[d];
        if (r[(int64_t)d * N + u] > 0 && h[v] < ORC_INF && h[v] + 1 < hm) hm = h[v] + 1;
      }
      h[u] = hm;
      ++relabels;
      if (hm >= ORC_INF) break;
    }
    if (relabels > gr_period) {
      relabels = 0;
      global_relabel(&g, e, r, h, queue);
      memset(inq, 0, (size_t)N);
      qh = qt = qn = 0;
      for (int64_t v = 0; v < N; ++v)
        if (e[v] > 0 && h[v] < ORC_INF) PUSHQ(v);
    }
  }
#undef PUSHQ
  int64_t deficit = 0;
  for (int64_t v = 0; v < N; ++v)
    if (e[v] < 0) deficit -= e[v];
  /* flow absorbed by the sink side plus the folded constant */
  int64_t flow = flow_const + deficit0 - deficit;

  uint8_t *lab = labels_out ? labels_out : (uint8_t *)malloc((size_t)N);
  int ok = lab ? orc_canonical_cut(nx, ny, nz, conn, e, r, lab) : -1;
  if (certificate_ok) *certificate_ok = ok;
  if (!labels_out) free(lab);
  if (e_out) memcpy(e_out, e, (size_t)N * 4);
  if (res_out) memcpy(res_out, r, (size_t)N * 4 * K);
  free(e); free(h); free(r); free(queue); free(fifo); free(inq); free(g.mask);
  return flow;
}

/* ---- integer data term -------------------------------------------------- */

static int km_key(const int32_t *x, int C, int wide) {
  if (wide) return x[0] >> 4;
  int s = 0;
  for (int c = 0; c < C; ++c) s += x[c];
  return s;
}

/* Initial centroids, in the spirit of graphcut_cut_distribution
 * (graphcut_histogram.cl:41-60) and the host pick at QGCWidget.cpp:2001-2039:
 * distinct seeded levels, evenly spaced picks. */
void orc_kmeans_init(const int32_t *feat, int n, int C, int wide, int K,
                     int64_t *centroids) {
  const int nkeys = wide ? 4096 : 255 * C + 1;
  int64_t *cnt = (int64_t *)calloc((size_t)nkeys, sizeof(int64_t));
  int *levels = (int *)malloc((size_t)nkeys * sizeof(int));
  for (int i = 0; i < n; ++i) cnt[km_key(feat + (int64_t)i * C, C, wide)]++;
  int m = 0;
  for (int k = 0; k < nkeys; ++k)
    if (cnt[k]) levels[m++] = k;
  for (int j = 0; j < K; ++j) {
    for (int c = 0; c < C; ++c) centroids[j * C + c] = 0;
    if (m == 0) continue;
    int pick = (K > 1) ? levels[(int64_t)j * (m - 1) / (K - 1)] : levels[0];
    int64_t s[4] = {0, 0, 0, 0}, nn = 0;
    for (int i = 0; i < n; ++i) {
      const int32_t *x = feat + (int64_t)i * C;
      if (km_key(x, C, wide) == pick) {
        for (int c = 0; c < C; ++c) s[c] += x[c];
        ++nn;
      }
    }
    for (int c = 0; c < C; ++c)
      centroids[j * C + c] = (s[c] * (1 << ORC_KM_FRAC) + nn / 2) / nn;
  }
  free(cnt);
  free(levels);
}

static int64_t km_d2(const int32_t *x, const int64_t *c, int C) {
  int64_t d = 0;
  for (int i = 0; i < C; ++i) {
    int64_t t = (int64_t)x[i] * (1 << ORC_KM_FRAC) - c[i];
    d += t * t;
  }
  return d;
}

int orc_kmeans_iterate(const int32_t *feat, int n, int C, int K, int max_iters,
                       int64_t *centroids) {
  int64_t *sum = (int64_t *)malloc((size_t)K * C * sizeof(int64_t));
  int64_t *cnt = (int64_t *)malloc((size_t)K * sizeof(int64_t));
  int it = 0;
  while (it < max_iters) {
    ++it;
    memset(sum, 0, (size_t)K * C * sizeof(int64_t));
    memset(cnt, 0, (size_t)K * sizeof(int64_t));
    /* assignment: nearest centroid, lowest index on ties (graphcut_kmeans.cl:21-60) */
    for (int i = 0; i < n; ++i) {
      const int32_t *x = feat + (int64_t)i * C;
      int best = 0;
      int64_t bd = km_d2(x, centroids, C);
      for (int k = 1; k < K; ++k) {
        int64_t dk = km_d2(x, centroids + k * C, C);
        if (dk < bd) { bd = dk; best = k; }
      }
      for (int c = 0; c < C; ++c) sum[best * C + c] += x[c];
      cnt[best]++;
    }
    /* update: rounded mean; empty clusters keep their centroid (QGCWidget.cpp:2088-2101) */
    int changed = 0;
    for (int k = 0; k < K; ++k) {
      if (!cnt[k]) continue;
      for (int c = 0; c < C; ++c) {
        int64_t nc = (sum[k * C + c] * (1 << ORC_KM_FRAC) + cnt[k] / 2) / cnt[k];
        if (nc != centroids[k * C + c]) { centroids[k * C + c] = nc; changed = 1; }
      }
    }
    if (!changed) break;
  }
  free(sum);
  free(cnt);
  return it;
}

void orc_unary(const int32_t *feat, int64_t N, int C, const int64_t *cfg,
               int Kf, const int64_t *cbg, int Kb, int64_t R,
               const uint8_t *seeds, int32_t *cs, int32_t *ct) {
  const int64_t den = R * R * ((int64_t)1 << (2 * ORC_KM_FRAC));
  for (int64_t v = 0; v < N; ++v) {
    if (seeds && seeds[v] == 1) { cs[v] = 1 << 30; ct[v] = 0; continue; }
    if (seeds && seeds[v] == 2) { cs[v] = 0; ct[v] = 1 << 30; continue; }
    const int32_t *x = feat + v * C;
    int64_t df = km_d2(x, cfg, C), db = km_d2(x, cbg, C);
    for (int k = 1; k < Kf; ++k) { int64_t t = km_d2(x, cfg + k * C, C); if (t < df) df = t; }
    for (int k = 1; k < Kb; ++k) { int64_t t = km_d2(x, cbg + k * C, C); if (t < db) db = t; }
    /* source capacity = cost of the background label = distance to the bg model */
    cs[v] = (int32_t)((100 * db + den / 2) / den);
    ct[v] = (int32_t)((100 * df + den / 2) / den);
  }
}

int orc_pair_thresholds(double lambda, int64_t S, int64_t P, int64_t *T) {
  int kmax = (int)floor(lambda + 0.5);
  for (int k = 1; k <= kmax; ++k) {
    if (S == 0 || P == 0) { T[k - 1] = INT64_MAX; continue; }
    double v = (2.0 * (double)S / (double)P) * log(lambda / (k - 0.5));
    T[k - 1] = (int64_t)floor(v);
  }
  return kmax;
}

static int pair_w(int64_t D, const int64_t *T, int n) {
  int w = 0;
  while (w < n && D <= T[w]) ++w;
  return w;
}

void orc_pairwise(int nx, int ny, int nz, int conn, const int32_t *feat, int C,
                  double lambda_xy, double lambda_z, int z_separate,
                  int32_t *caps) {
  grid_t g;
  if (grid_init(&g, nx, ny, nz, conn)) return;
  const int64_t N = g.N;
  const int K = g.K;
  /* mean squared neighbour difference over each group of undirected edges
   * (even directions = one arc per edge) */
  int64_t S[2] = {0, 0}, P[2] = {0, 0};
  for (int64_t v = 0; v < N; ++v)
    for (int d = 0; d < K; d += 2) {
      if (!(g.mask[v] & (1u << d))) continue;
      int grp = (conn == 6 && d == 4 && z_separate) ? 1 : 0;
      const int32_t *a = feat + v * C, *b = feat + (v + g.off[d]) * C;
      int64_t D = 0;
      for (int c = 0; c < C; ++c) D += (int64_t)(a[c] - b[c]) * (a[c] - b[c]);
      S[grp] += D;
      P[grp] += 1;
    }
  int64_t Txy[256], Td[256], Tz[256];
  int nxy = orc_pair_thresholds(lambda_xy, S[0], P[0], Txy);
  int nd = orc_pair_thresholds(lambda_xy / sqrt(2.0), S[0], P[0], Td);
  int nzz = z_separate ? orc_pair_thresholds(lambda_z, S[1], P[1], Tz)
                       : orc_pair_thresholds(lambda_z, S[0], P[0], Tz);
  for (int d = 0; d < K; ++d)
    for (int64_t v = 0; v < N; ++v) {
      int32_t w = 0;
      if (g.mask[v] & (1u << d)) {
        const int32_t *a = feat + v * C, *b = feat + (v + g.off[d]) * C;
        int64_t D = 0;
        for (int c = 0; c < C; ++c) D += (int64_t)(a[c] - b[c]) * (a[c] - b[c]);
        if (conn == 6 && d >= 4) w = pair_w(D, Tz, nzz);
        else if (conn == 8 && d >= 4) w = pair_w(D, Td, nd);
        else w = pair_w(D, Txy, nxy);
      }
      caps[(int64_t)d * N + v] = w;
    }
  free(g.mask);
}
