// jfc_fastcut.hip -- the Fast-Cut comparator of JF-Cut's widget
// (relabelPushFast, src/QVisualizer/Graph Cut/QGCWidget.cpp:2156-2205, with
// the kernels of src/QVisualizer/cl/GraphCut/graphcut_push_relabel_fast.cl),
// selected by jfc_options.method = 1 in place of the tile push-relabel sweep.
//
// One Fast-Cut sweep over the active tiles:
//   k_fc_relabel  x2  graphcut_fast_relabel (:54-77, :228-256): every node of
//                     one colour of the node checkerboard with e >= 0 and a
//                     finite label takes min over residual arcs of h + 1;
//                     the other colour is read only, so each launch is a
//                     Jacobi step without races (4- and 6-connected grids are
//                     bipartite).
//   k_fc_push     x2D graphcut_fast_push_x/_y/_z (:81-170, :260-387): one
//                     thread owns a tile line along the axis, pushes
//                     left-to-right through the line (flow travels the whole
//                     line in one pass), into the first node of the next
//                     tile, then right-to-left and into the last node of the
//                     previous tile.  Tiles of one parity along the axis run
//                     together, so every boundary node has one writer.
//   k_fc_flags        the tiles' active flags (graphcut_count_node).
// Heights are frozen during the pushes and every push is along an admissible
// arc (h(u) = h(v) + 1), so the labelling stays valid and the result is the
// same exact max-flow and canonical cut as the tile solver's.
#include "jfc_device.h"
#include "jfc_internal.h"

namespace jfc {

namespace {

// count of the list a compaction just filled
__device__ __forceinline__ int list_count(const Queue &q) { return q.ctr[QSLOT * q.slot]; }

template <class S>
__global__ __launch_bounds__(256) void k_fc_relabel(Tiles T, Planes P, Queue q, int parity,
                                                    int clear_other) {
  // the next compaction appends to the other slot
  if (clear_other && blockIdx.x == 0 && threadIdx.x < 9)
    q.ctr[QSLOT * (q.slot ^ 1) + QWORD * threadIdx.x] = 0;
  const int count = list_count(q);
  const Lane<S> L(threadIdx.x);
  for (int i = blockIdx.x; i < count; i += gridDim.x) {
    const int t = q.list[i] & kEntryTile;
    const TileRef r = tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    if (gx >= I.nx || gy >= I.ny || gz >= I.nz || ((gx + gy + gz) & 1) != parity) continue;
    const long long v = node_index(I, gx, gy, gz);
    const int h = P.h[v] & HMASK;
    if (P.e[v] < 0 || h >= HINF) continue;
    const unsigned m = P.rmask[v];
    const long long nn = (long long)I.nx * I.ny * I.nz;
    const int hcap = nn < HINF ? (int)nn : HINF;
    int nh = HINF;
    for (int d = 0; d < I.K; ++d) {
      if (!((m >> d) & 1u)) continue;
      int dx, dy, dz;
      dir_of<S::D3>(d, dx, dy, dz);
      nh = min(nh, (P.h[node_index(I, gx + dx, gy + dy, gz + dz)] & HMASK) + 1);
    }
    if (nh >= hcap) nh = HINF;
    if (nh != h) P.h[v] = nh;
  }
}

// Push along one line of one tile.  Planes: F = +axis, B = -axis.
template <class S, int AXIS>
__global__ __launch_bounds__(256) void k_fc_push(Tiles T, Planes P, Queue q, int parity) {
  constexpr int LEN = AXIS == 0 ? S::TW : (AXIS == 1 ? S::TH : S::TD);
  constexpr int NLINES = S::NT / LEN;
  constexpr int PER_BLOCK = 256 / NLINES;  // tiles per workgroup
  constexpr int F = 2 * AXIS, B = 2 * AXIS + 1;
  const int count = list_count(q);
  const int slot = threadIdx.x / NLINES, line = threadIdx.x % NLINES;
  const long long st = P.stride;
  for (int i = blockIdx.x * PER_BLOCK + slot; i < count; i += gridDim.x * PER_BLOCK) {
    const int t = q.list[i] & kEntryTile;
    const TileRef r = tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int tc = AXIS == 0 ? r.ltx : (AXIS == 1 ? r.lty : r.ltz);
    if ((tc & 1) != parity) continue;
    // fixed coordinates of the line and its first node
    int x0 = r.ox, y0 = r.oy, z0 = r.oz;
    if (AXIS == 0) { y0 += line % S::TH; z0 += line / S::TH; }
    if (AXIS == 1) { x0 += line % S::TW; z0 += line / S::TW; }
    if (AXIS == 2) { x0 += line % S::TW; y0 += line / S::TW; }
    if (x0 >= I.nx || y0 >= I.ny || z0 >= I.nz) continue;
    const int extent = AXIS == 0 ? I.nx - x0 : (AXIS == 1 ? I.ny - y0 : I.nz - z0);
    const int n = min(LEN, extent);
    const long long v0 = node_index(I, x0, y0, z0);
    const long long step = AXIS == 0 ? 1 : (AXIS == 1 ? (long long)I.px : (long long)I.px * I.ny);
    int e[LEN], e0[LEN], h[LEN], cf[LEN], cb[LEN], cf0[LEN], cb0[LEN];
    for (int k = 0; k < n; ++k) {
      const long long v = v0 + k * step;
      e[k] = e0[k] = P.e[v];
      h[k] = P.h[v] & HMASK;
      cf[k] = cf0[k] = P.cap[F * st + v];
      cb[k] = cb0[k] = P.cap[B * st + v];
    }
    // forward (graphcut_push_relabel_fast.cl:115-119)
    for (int k = 0; k + 1 < n; ++k) {
      if (h[k] - h[k + 1] != 1) continue;
      const int f = min(cf[k], e[k]);
      if (f > 0) {
        cf[k] -= f;
        e[k] -= f;
        cb[k + 1] += f;
        e[k + 1] += f;
      }
    }
    // into the first node of the next tile (push_fast_tt, :120)
    const bool has_next = extent > LEN;
    if (has_next) {
      const int f = min(cf[n - 1], e[n - 1]);
      if (f > 0) {
        const long long u = v0 + n * step;
        if (h[n - 1] - (P.h[u] & HMASK) == 1) {
          cf[n - 1] -= f;
          e[n - 1] -= f;
          const int c = P.cap[B * st + u];
          P.cap[B * st + u] = c + f;
          P.e[u] += f;
          if (c <= 0 && c + f > 0) P.rmask[u] |= 1u << B;
        }
      }
    }
    // backward (:122-126)
    for (int k = n - 1; k > 0; --k) {
      if (h[k] - h[k - 1] != 1) continue;
      const int f = min(cb[k], e[k]);
      if (f > 0) {
        cb[k] -= f;
        e[k] -= f;
        cf[k - 1] += f;
        e[k - 1] += f;
      }
    }
    // into the last node of the previous tile (:127)
    const bool has_prev = (AXIS == 0 ? x0 : (AXIS == 1 ? y0 : z0)) > 0;
    if (has_prev) {
      const int f = min(cb[0], e[0]);
      if (f > 0) {
        const long long u = v0 - step;
        if (h[0] - (P.h[u] & HMASK) == 1) {
          cb[0] -= f;
          e[0] -= f;
          const int c = P.cap[F * st + u];
          P.cap[F * st + u] = c + f;
          P.e[u] += f;
          if (c <= 0 && c + f > 0) P.rmask[u] |= 1u << F;
        }
      }
    }
    // write back what changed (:129-132), with the two mask bits
    for (int k = 0; k < n; ++k) {
      const long long v = v0 + k * step;
      if (e[k] != e0[k]) P.e[v] = e[k];
      if (cf[k] != cf0[k]) P.cap[F * st + v] = cf[k];
      if (cb[k] != cb0[k]) P.cap[B * st + v] = cb[k];
      if ((cf[k] > 0) != (cf0[k] > 0) || (cb[k] > 0) != (cb0[k] > 0)) {
        const unsigned m = P.rmask[v] & ~((1u << F) | (1u << B));
        P.rmask[v] = m | ((unsigned)(cf[k] > 0) << F) | ((unsigned)(cb[k] > 0) << B);
      }
    }
  }
}

// Active flag of every tile (a node with excess and a finite label) and its
// deficit flag (a node with e < 0, the relabel's ordering seeds).
template <class S>
__global__ __launch_bounds__(256) void k_fc_flags(Tiles T, Planes P, unsigned char *flag,
                                                  unsigned char *defi) {
  __shared__ int s_any, s_def;
  const Lane<S> L(threadIdx.x);
  for (int t = blockIdx.x; t < T.ntiles; t += gridDim.x) {
    if (threadIdx.x == 0) s_any = s_def = 0;
    __syncthreads();
    const TileRef r = tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    if (gx < I.nx && gy < I.ny && gz < I.nz) {
      const long long v = node_index(I, gx, gy, gz);
      const int e = P.e[v];
      if (e > 0 && (P.h[v] & HMASK) < HINF) s_any = 1;
      if (e < 0) s_def = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      flag[t] = (unsigned char)s_any;
      defi[t] = (unsigned char)s_def;
    }
    __syncthreads();
  }
}

// Every residual plane written from the inputs and the fresh bits cleared:
// the line pushes read and write cap directly.
__global__ void k_fc_materialize(Planes P, const unsigned *mask0, int K) {
  const long long n = P.stride;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n;
       v += (long long)gridDim.x * blockDim.x) {
    for (int d = 0; d < K; ++d) P.cap[d * n + v] = P.c0[d * n + v];
    P.rmask[v] = mask0[v] & 0xffu;
  }
}

}  // namespace

void launch_fc_materialize(const Planes &P, const unsigned *mask0, int K, hipStream_t st) {
  const long long g = std::min<long long>((P.stride + 255) / 256, 16384);
  k_fc_materialize<<<(int)std::max<long long>(g, 1), 256, 0, st>>>(P, mask0, K);
}

void launch_fc_sweep(int shape, int nwg, const Tiles &T, const Planes &P, const Queue &q,
                     unsigned char *flag, unsigned char *defi, hipStream_t st) {
  const int d3 = shape_is_3d(shape);
  JFC_DISPATCH(shape, ({
    for (int par = 0; par < 2; ++par)
      k_fc_relabel<S><<<nwg, 256, 0, st>>>(T, P, q, par, par == 0);
    for (int par = 0; par < 2; ++par)
      k_fc_push<S, 0><<<nwg, 256, 0, st>>>(T, P, q, par);
    for (int par = 0; par < 2; ++par)
      k_fc_push<S, 1><<<nwg, 256, 0, st>>>(T, P, q, par);
    if (d3)
      for (int par = 0; par < 2; ++par)
        k_fc_push<S, 2><<<nwg, 256, 0, st>>>(T, P, q, par);
    k_fc_flags<S><<<std::min(T.ntiles, 8192), 256, 0, st>>>(T, P, flag, defi);
  }));
}

}  // namespace jfc
