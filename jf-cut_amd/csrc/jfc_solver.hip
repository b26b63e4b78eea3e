// jfc_solver.hip -- gfx950 kernels of the grid max-flow solve.
//
// Reference counterparts (paths relative to src/QVisualizer/):
//   k_push     graphcut_relabel_push   cl/GraphCut/graphcut_push_relabel.cl:25-203 (2D), :215-439 (3D)
//   k_relax    graphcut_global_relabel cl/GraphCut/graphcut_global_relabel.cl:50-124, :128-249
//   k_gr_init  graphcut_init_global_relabel  graphcut_global_relabel.cl:19-48
//   k_reach    graphcut_compute_cut    cl/GraphCut/graphcut_compute_cut.cl:51-140, :182-282
//   k_cut_init graphcut_init_cut       graphcut_compute_cut.cl:19-49
//   k_compact  graphcut_count_node + clogs scan + graphcut_scatter_node
//                               cl/GraphCut/graphcut_count_block.cl:19-65
//   k_jf_*     graphcut_bfs_jump_flooding / _histogram / _mapping
//              cl/GraphCut/graphcut_bfs_block.cl:43-91 (+ reduce_histogram, INF reduce.cl:56-141)
//   k_cut_init also reduce_flow  bin/x64/cl/Infrastructures/reduce.cl:17-42 (int64 here)
//
// Layout: every graph stores int32 planes (e, h, K residual planes) in
// structure-of-arrays order over all instances.  Kernels work on tiles of
// 256 nodes (one node per lane) staged in LDS with a one-node halo.  Tiles of
// one colour class never share a node; a halo node shared by two of them is
// reached through different arcs, so within a push-relabel launch every
// residual plane entry has one writer and only excess and mask words need
// atomics.
#include "jfc_device.h"

#include <algorithm>
#include <atomic>
#include <climits>

namespace jfc {

bool shape_is_3d(int s) { return s == SH_8x8x4 || s == SH_32x4x2; }

void shape_dims(int s, int *tw, int *th, int *td) {
  int a = 0, b = 0, c = 0;
  JFC_DISPATCH(s, { a = S::TW; b = S::TH; c = S::TD; });
  *tw = a; *th = b; *td = c;
}

// ---- small utilities --------------------------------------------------------

__global__ void k_fill_i32(int *p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u8(unsigned char *p, long long n, unsigned char v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

static int grid_for(long long n, int per_block = 256, int cap = 8192) {
  long long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

void launch_fill_i32(int *p, long long n, int v, hipStream_t st) {
  if (n > 0) k_fill_i32<<<grid_for(n), 256, 0, st>>>(p, n, v);
}
void launch_fill_u8(unsigned char *p, long long n, unsigned char v, hipStream_t st) {
  if (n > 0) k_fill_u8<<<grid_for(n), 256, 0, st>>>(p, n, v);
}

// Several fills in one launch: blockIdx.y picks the segment.
__global__ void k_fill_multi(FillSet f) {
  if ((int)blockIdx.y >= f.n) return;
  const FillSeg s = f.seg[blockIdx.y];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < s.n;
       i += (long long)gridDim.x * blockDim.x) {
    if (s.bytes == 1) static_cast<unsigned char *>(s.p)[i] = (unsigned char)s.v;
    else static_cast<int *>(s.p)[i] = s.v;
  }
}

void launch_fill_multi(const FillSet &f, hipStream_t st) {
  if (f.n <= 0) return;
  long long most = 1;
  for (int k = 0; k < f.n; ++k) most = std::max(most, f.seg[k].n);
  k_fill_multi<<<dim3(grid_for(most, 256, 1024), f.n), 256, 0, st>>>(f);
}

__global__ void k_validate(const int *p, long long n, int lo, int hi, int *bad) {
  int b = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = p[i];
    b |= (v < lo) | (v > hi);
  }
  if (__ballot(b) && lane_id() == 0) atomicOr(bad, 1);
}

void launch_validate(const int *p, long long n, int lo, int hi, int *bad,
                     hipStream_t st) {
  if (n > 0) k_validate<<<grid_for(n), 256, 0, st>>>(p, n, lo, hi, bad);
}

// Hard seeds from the caller's unpadded seed bytes (row r of nx bytes) onto
// the stored planes (row r at r * px): fg wins where both are set, the rule
// of graphcut_compute_excessflow.cl:71-74.
__global__ void k_seeds(int *cs, int *ct, const unsigned char *fg, const unsigned char *bg,
                        int nx, int px, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = (i / nx) * px + i % nx;
    if (fg && fg[i]) { cs[v] = 1 << 30; ct[v] = 0; }
    else if (bg && bg[i]) { cs[v] = 0; ct[v] = 1 << 30; }
  }
}

void launch_seeds(int *cs, int *ct, const unsigned char *fg, const unsigned char *bg, int nx,
                  int px, long long rows, hipStream_t st) {
  const long long n = rows * nx;
  if (n > 0 && (fg || bg)) k_seeds<<<grid_for(n), 256, 0, st>>>(cs, ct, fg, bg, nx, px, n);
}

__global__ void k_excess_split(const int *ef, int *cs, int *ct, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = ef[i];
    cs[i] = v > 0 ? v : 0;
    ct[i] = v < 0 ? -v : 0;
  }
}

void launch_excess_split(const int *ef, int *cs, int *ct, long long n,
                         hipStream_t st) {
  if (n > 0) k_excess_split<<<grid_for(n), 256, 0, st>>>(ef, cs, ct, n);
}

// Zero the arcs of plane d that leave the grid of one instance.
template <class S>
__global__ void k_zero_border(InstDesc I, int *plane, int d) {
  int dx, dy, dz;
  dir_of<S::D3>(d, dx, dy, dz);
  const long long n = (long long)I.nx * I.ny * I.nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % I.nx);
    const long long r = i / I.nx;
    const int y = (int)(r % I.ny);
    const int z = (int)(r / I.ny);
    const int X = x + dx, Y = y + dy, Z = z + dz;
    if (X < 0 || X >= I.nx || Y < 0 || Y >= I.ny || Z < 0 || Z >= I.nz)
      plane[node_index(I, x, y, z)] = 0;
  }
}

void launch_zero_border(int shape, const InstDesc &I, int *plane, int d,
                        hipStream_t st) {
  const long long n = (long long)I.nx * I.ny * I.nz;
  JFC_DISPATCH(shape, (k_zero_border<S><<<grid_for(n), 256, 0, st>>>(I, plane, d)));
}

// ---- block reductions --------------------------------------------------------

__device__ __forceinline__ long long block_sum_ll(long long v, long long *buf) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if (lane_id() == 0) buf[w] = v;
  __syncthreads();
  long long r = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += buf[i];
  return r;
}

// Tile-parallel kernels: workgroup b walks a contiguous chunk of tiles so that
// per-instance sums are flushed once per instance change.
struct Chunk {
  int t0, t1;
  __device__ __forceinline__ Chunk(int ntiles) {
    const int per = (ntiles + gridDim.x - 1) / gridDim.x;
    t0 = blockIdx.x * per;
    t1 = min(ntiles, t0 + per);
  }
};

static int tile_grid(int ntiles) { return ntiles < 4096 ? (ntiles > 0 ? ntiles : 1) : 4096; }

// Streaming tile kernels keep TU tiles in flight per lane (their loads are
// issued before any use) so that enough bytes are outstanding to cover HBM
// latency with one 256-lane group per tile.
constexpr int TU = 4;

// Bitwise OR of TU per-tile bits over the 256-lane workgroup (4 waves):
// one ballot per bit, one LDS word per wave.  The buffer half alternates with
// `parity`, so the words read in one group are never overwritten before
// every lane has passed the next group's barrier.  (__syncthreads_or only
// tells whether any lane is non-zero.)
__device__ __forceinline__ int block_or_bits(int bits, int *s_w, int parity) {
  int m = 0;
#pragma unroll
  for (int u = 0; u < TU; ++u)
    if (__ballot((bits >> u) & 1)) m |= 1 << u;
  int *buf = s_w + 4 * parity;
  if (lane_id() == 0) buf[threadIdx.x >> 6] = m;
  __syncthreads();
  return buf[0] | buf[1] | buf[2] | buf[3];
}

// Any lane of the 256-lane workgroup with pred set: one ballot per wave, one
// LDS word per wave, one barrier.  The buffer half alternates with `parity`
// (as in block_or_bits), so a second barrier is not needed before reuse.
// The barrier also orders the workgroup's LDS accesses like __syncthreads.
__device__ __forceinline__ int block_any(int pred, int *s_w, int &parity) {
  const int w = __ballot(pred) != 0;
  int *buf = s_w + 4 * parity;
  parity ^= 1;
  if (lane_id() == 0) buf[threadIdx.x >> 6] = w;
  __syncthreads();
  return buf[0] | buf[1] | buf[2] | buf[3];
}

// block_any for two bits at once: bit b of the result is set when some lane
// has bit b of `bits` set (one ballot per bit, one barrier).
__device__ __forceinline__ int block_any2(int bits, int *s_w, int &parity) {
  const int w = (__ballot(bits & 1) != 0 ? 1 : 0) | (__ballot(bits & 2) != 0 ? 2 : 0);
  int *buf = s_w + 4 * parity;
  parity ^= 1;
  if (lane_id() == 0) buf[threadIdx.x >> 6] = w;
  __syncthreads();
  return buf[0] | buf[1] | buf[2] | buf[3];
}

// Node of lane L in tile t; inst = -1 past the end of the chunk.
struct TileNode {
  int inst;
  bool valid;
  long long v;
};

template <class S>
__device__ __forceinline__ TileNode tile_node(const Tiles &T, int t, int t_end,
                                              const Lane<S> &L) {
  TileNode n{-1, false, 0};
  if (t >= t_end) return n;
  const TileRef r = tile_ref<S>(T, t);
  n.inst = T.tile_inst[t];
  const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
  n.valid = gx < r.I.nx && gy < r.I.ny && gz < r.I.nz;
  n.v = node_index(r.I, gx, gy, gz);
  return n;
}

// ---- row-streaming passes ----------------------------------------------------
//
// Whole-graph passes (fold, relabel / cut initialisation, deficit, masks) walk
// grid rows, not tiles: row r of the graph (the rows of all instances
// concatenated) is nx consecutive nodes, so a workgroup streams it with
// 16-byte loads and stores, several in flight per lane.  A lane handles a
// group of 4 consecutive nodes (rows start at multiples of 32 nodes).  A
// workgroup step takes one row, or several short rows of a one-instance
// graph (RowStep), so that every lane has groups to stream.  Per-tile flags
// are derived from the column and the row's tile row; per-instance sums are
// flushed when a workgroup's step changes instance and at its end.

// Groups in flight per lane.
constexpr int RS_U = 4;

template <class S>
struct RowRef {
  int inst, nx;
  long long v0;  // first node of the row
  int tile0;     // tile of column 0
};

template <class S>
__device__ __forceinline__ RowRef<S> row_ref(const Tiles &T, int r) {
  int lo = 0, hi = T.ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.row_off[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  const InstDesc &I = T.inst[lo];
  const int rr = r - T.row_off[lo];
  const int y = rr % I.ny, z = rr / I.ny;
  RowRef<S> R;
  R.inst = lo;
  R.nx = I.nx;
  R.v0 = I.node_off + (long long)rr * I.px;
  R.tile0 = I.tile_off + I.tx * (y / S::TH + I.ty * (z / S::TD));
  return R;
}

// Row r of a one-instance graph (no search over the instances).
template <class S>
__device__ __forceinline__ RowRef<S> row_ref_one(const Tiles &T, int r) {
  const InstDesc &I = T.inst[0];
  const int y = r % I.ny, z = r / I.ny;
  RowRef<S> R;
  R.inst = 0;
  R.nx = I.nx;
  R.v0 = I.node_off + (long long)r * I.px;
  R.tile0 = I.tile_off + I.tx * (y / S::TH + I.ty * (z / S::TD));
  return R;
}

// Group k of a row: first column x and node count n (0 past the end).
struct Group {
  int x, n;
};

// A workgroup step of a whole-graph pass: T.row_step rows from row r0 (one
// instance), seen as slots of 4-node groups, ng per row.  Slot k is group
// k % ng of row r0 + k / ng.  Rows start at multiples of 32 nodes, so a row's
// groups are 4-aligned.
template <class S>
struct RowStep {
  RowRef<S> R0;  // first row of the step
  int r0, ng, slots;
  __device__ __forceinline__ RowStep(const Tiles &T, int step) {
    r0 = step * T.row_step;
    R0 = row_ref<S>(T, r0);
    ng = (R0.nx + 3) / 4;
    slots = min(T.row_step, T.nrows - r0) * ng;
  }
  // The row and group of slot k (k < slots).
  __device__ __forceinline__ void at(const Tiles &T, int k, RowRef<S> &R, Group &g) const {
    const int dr = T.row_step > 1 ? k / ng : 0;
    const int gk = T.row_step > 1 ? k - dr * ng : k;
    R = dr ? row_ref_one<S>(T, r0 + dr) : R0;
    g.x = 4 * gk;
    g.n = min(4, R.nx - g.x);
  }
};

__device__ __forceinline__ int row_steps(const Tiles &T) {
  return (T.nrows + T.row_step - 1) / T.row_step;
}

// Load / store up to 4 consecutive ints (16-byte access when the group is
// full and aligned).
__device__ __forceinline__ int4 load4(const int *p, long long v, int n) {
  int4 r = make_int4(0, 0, 0, 0);
  if (n == 4 && (((size_t)(p + v)) & 15) == 0) return *(const int4 *)(p + v);
  if (n > 0) r.x = p[v];
  if (n > 1) r.y = p[v + 1];
  if (n > 2) r.z = p[v + 2];
  if (n > 3) r.w = p[v + 3];
  return r;
}

__device__ __forceinline__ void store4(int *p, long long v, int n, int4 a) {
  if (n == 4 && (((size_t)(p + v)) & 15) == 0) { *(int4 *)(p + v) = a; return; }
  if (n > 0) p[v] = a.x;
  if (n > 1) p[v + 1] = a.y;
  if (n > 2) p[v + 2] = a.z;
  if (n > 3) p[v + 3] = a.w;
}

__device__ __forceinline__ void store4_u8(unsigned char *p, long long v, int n, unsigned bits) {
  if (n == 4 && (((size_t)(p + v)) & 3) == 0) {
    *(unsigned *)(p + v) = (bits & 1u) | ((bits >> 1) & 1u) << 8 | ((bits >> 2) & 1u) << 16 |
                           ((bits >> 3) & 1u) << 24;
    return;
  }
  for (int i = 0; i < n; ++i) p[v + i] = (unsigned char)((bits >> i) & 1u);
}

__device__ __forceinline__ int comp(const int4 &a, int i) {
  return i == 0 ? a.x : i == 1 ? a.y : i == 2 ? a.z : a.w;
}

// Set flag[tile] = 1 for the tiles of the nodes of a group with a set bit.
template <class S>
__device__ __forceinline__ void flag_group(unsigned char *flag, int tile0, Group g,
                                           unsigned bits) {
  if (!bits) return;
  const int t0 = tile0 + g.x / S::TW, t1 = tile0 + (g.x + g.n - 1) / S::TW;
  if (t0 == t1) {
    flag[t0] = 1;
    return;
  }
  for (int i = 0; i < g.n; ++i)
    if ((bits >> i) & 1u) flag[tile0 + (g.x + i) / S::TW] = 1;
}

// Per-instance int64 sums of a workgroup: accumulated per lane, flushed with
// one atomic per workgroup and instance.
template <int NS>
struct InstSums {
  long long a[NS];
  int cur = -1;
  __device__ __forceinline__ void reset() {
#pragma unroll
    for (int k = 0; k < NS; ++k) a[k] = 0;
  }
  __device__ __forceinline__ void flush(long long *const *dst, const int *stride,
                                        long long *buf) {
    if (cur < 0) return;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const long long A = block_sum_ll(a[k], buf);
      if (threadIdx.x == 0 && A != 0)
        atomicAdd((unsigned long long *)&dst[k][(long long)cur * stride[k]], (unsigned long long)A);
    }
  }
  // Called with the (uniform) instance of the next row.
  __device__ __forceinline__ void enter(int inst, long long *const *dst, const int *stride,
                                        long long *buf) {
    if (inst == cur) return;
    flush(dst, stride, buf);
    cur = inst;
    reset();
  }
};

// Grid of a persistent (queue-consuming) launch: every workgroup must be
// resident from the start -- one dispatched late finds the queue drained but
// still runs its static first run after the others are done, stretching the
// launch -- so the grid is capped at what the device holds at once for this
// kernel (occupancy per CU, from its register and LDS use, times CUs).
// Compute units of the current device, queried once per device (every
// persistent launch sizes its grid with it).
static int device_cus() {
  static std::atomic<int> cache[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev >= 0 && dev < 16) {
    const int c = cache[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 1;
  if (dev >= 0 && dev < 16) cache[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

template <auto Kernel, int Block = 256>
static int resident_grid_b(int want) {
  static const int per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, Kernel, Block, 0) != hipSuccess || n < 1)
      n = 1;
    return n;
  }();
  return std::max(1, std::min(want, per_cu * device_cus()));
}

template <auto Kernel>
static int resident_grid(int want) {
  return resident_grid_b<Kernel, 256>(want);
}

// Grid of a pass over the rows (nsteps row steps): one workgroup per step at
// most, and no more than the device holds at once -- a second, partial round
// of workgroups would run with most of the device idle.
template <auto Kernel>
static int row_grid(int nsteps) {
  return resident_grid<Kernel>(std::max(1, nsteps));
}

// Fold terminal arcs: e = cs - ct; const += min(cs, ct);
// deficit0 += max(ct - cs, 0); supply0 += max(cs - ct, 0).
// With `active` set it flags the tiles holding excess (`active` zeroed by the
// launcher); with `zero_h` it starts the labelling at zero heights over all
// nodes (the reference clears h, QGCWidget.cpp:2400; the push kernel reads
// the heights of tiles it has not visited in the solve as 0 instead).  It
// always flags the tiles holding a deficit node in `defi`, the seeds of a
// global relabel's tile ordering (the seed flags of
// graphcut_init_global_relabel, graphcut_global_relabel.cl:19-48, kept
// current by the push visits instead of a pass over e).
//
// The same pass starts the residual-arc masks of the solve (MASKS): 0 leaves
// them alone, 1 copies mask0 to rmask, 2 builds mask0 from the input planes
// (residual bit d = c0[d] > 0, every plane fresh) and writes it to both, 3
// builds it into rmask only (mask0 is then built by the next solve of the
// same inputs, if there is one).  One pass instead of a fold, a mask build
// and a copy: the mask words are written once and the terminal planes read
// once.
template <class S, int MASKS>
__global__ __launch_bounds__(256) void k_fold(Tiles T, Planes P, long long *red,
                                              unsigned char *active, unsigned char *defi,
                                              int zero_h, unsigned *mask0) {
  // planes of eight directions in flight per group: one group per lane
  constexpr int U = MASKS >= 2 ? 1 : RS_U;
  __shared__ long long buf[4];
  long long *dst[3] = {red + RED_CONST, red + RED_DEF0, red + RED_SUP0};
  const int stride[3] = {RED_SLOTS, RED_SLOTS, RED_SLOTS};
  InstSums<3> sums;
  sums.reset();
  const int nsteps = row_steps(T);
  for (int step = blockIdx.x; step < nsteps; step += gridDim.x) {
    const RowStep<S> RS(T, step);
    sums.enter(RS.R0.inst, dst, stride, buf);
    const int K = MASKS >= 2 ? T.inst[RS.R0.inst].K : 0;
    for (int k0 = threadIdx.x; k0 < RS.slots; k0 += U * 256) {
      Group g[U];
      RowRef<S> R[U];
      int4 a[U], b[U], m[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * 256;
        if (k < RS.slots) RS.at(T, k, R[u], g[u]);
        else { R[u] = RS.R0; g[u] = Group{0, 0}; }
        a[u] = load4(P.cs, R[u].v0 + g[u].x, g[u].n);
        b[u] = load4(P.ct, R[u].v0 + g[u].x, g[u].n);
        if (MASKS == 1) m[u] = load4((const int *)mask0, R[u].v0 + g[u].x, g[u].n);
        if (MASKS >= 2) {
          int4 c[S::KMAX];
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d)
            c[d] = d < K ? load4(P.c0 + d * P.stride, R[u].v0 + g[u].x, g[u].n) : make_int4(0, 0, 0, 0);
          m[u] = make_int4(0, 0, 0, 0);
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d) {
            if (d >= K) break;
            const int fresh = 1 << (8 + d);
            m[u].x |= (c[d].x > 0 ? 1 << d : 0) | fresh; m[u].y |= (c[d].y > 0 ? 1 << d : 0) | fresh;
            m[u].z |= (c[d].z > 0 ? 1 << d : 0) | fresh; m[u].w |= (c[d].w > 0 ? 1 << d : 0) | fresh;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (g[u].n == 0) continue;
        const long long v = R[u].v0 + g[u].x;
        int4 e;
        e.x = a[u].x - b[u].x; e.y = a[u].y - b[u].y;
        e.z = a[u].z - b[u].z; e.w = a[u].w - b[u].w;
        store4(P.e, v, g[u].n, e);
        if (zero_h) store4(P.h, v, g[u].n, make_int4(0, 0, 0, 0));
        if (MASKS) store4((int *)P.rmask, v, g[u].n, m[u]);
        if (MASKS == 2) store4((int *)mask0, v, g[u].n, m[u]);
        unsigned bits = 0, dbits = 0;
        for (int i = 0; i < g[u].n; ++i) {
          const int s = comp(a[u], i), t = comp(b[u], i);
          sums.a[0] += min(s, t);
          sums.a[1] += max(t - s, 0);
          sums.a[2] += max(s - t, 0);
          bits |= (unsigned)(s > t) << i;
          dbits |= (unsigned)(s < t) << i;
        }
        if (active) flag_group<S>(active, R[u].tile0, g[u], bits);
        flag_group<S>(defi, R[u].tile0, g[u], dbits);
      }
    }
  }
  sums.flush(dst, stride, buf);
}

void launch_fold(int shape, const Tiles &T, const Planes &P, long long *red,
                 unsigned char *active, unsigned char *defi, bool zero_h, unsigned *mask0,
                 int masks, hipStream_t st, bool cleared) {
  if (!cleared) {
    if (active) launch_fill_u8(active, T.ntiles, 0, st);
    launch_fill_u8(defi, T.ntiles, 0, st);
  }
  const int z = zero_h ? 1 : 0;
  const int n = (T.nrows + T.row_step - 1) / T.row_step;
  if (masks == 3) {
    JFC_DISPATCH(shape, (k_fold<S, 3><<<row_grid<&k_fold<S, 3>>(n), 256, 0, st>>>(
                            T, P, red, active, defi, z, mask0)));
  } else if (masks == 2) {
    JFC_DISPATCH(shape, (k_fold<S, 2><<<row_grid<&k_fold<S, 2>>(n), 256, 0, st>>>(
                            T, P, red, active, defi, z, mask0)));
  } else if (masks == 1) {
    JFC_DISPATCH(shape, (k_fold<S, 1><<<row_grid<&k_fold<S, 1>>(n), 256, 0, st>>>(
                            T, P, red, active, defi, z, mask0)));
  } else {
    JFC_DISPATCH(shape, (k_fold<S, 0><<<row_grid<&k_fold<S, 0>>(n), 256, 0, st>>>(
                            T, P, red, active, defi, z, mask0)));
  }
}

// graphcut_init_cut (graphcut_compute_cut.cl:19-49): label = 1 where
// excess > 0, tiles holding such a node seed the jump flooding (`seed`
// zeroed by the launcher).  The same pass sums the remaining deficit
// max(-e, 0) per instance into def[inst] (reduce_flow, reduce.cl:17-42, in
// int64): the cut search is the last step of a solve and reads the final
// excess.
template <class S>
__global__ __launch_bounds__(256) void k_cut_init(Tiles T, Planes P, unsigned char *seed,
                                                  long long *def) {
  __shared__ long long buf[4];
  long long *dst[1] = {def};
  const int stride[1] = {1};
  InstSums<1> sums;
  sums.reset();
  // two groups in flight per lane: 67 registers, 7 waves per SIMD (four: 95,
  // 5 waves, 0.013 ms slower on the 15000x10000 8-connected cut)
  constexpr int U = 2;
  const int nsteps = row_steps(T);
  for (int step = blockIdx.x; step < nsteps; step += gridDim.x) {
    const RowStep<S> RS(T, step);
    sums.enter(RS.R0.inst, dst, stride, buf);
    for (int k0 = threadIdx.x; k0 < RS.slots; k0 += U * 256) {
      Group g[U];
      RowRef<S> R[U];
      int4 e[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * 256;
        if (k < RS.slots) RS.at(T, k, R[u], g[u]);
        else { R[u] = RS.R0; g[u] = Group{0, 0}; }
        e[u] = load4(P.e, R[u].v0 + g[u].x, g[u].n);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (g[u].n == 0) continue;
        unsigned bits = 0;
        for (int i = 0; i < g[u].n; ++i) {
          const int x = comp(e[u], i);
          sums.a[0] += max(-x, 0);
          bits |= (unsigned)(x > 0) << i;
        }
        store4_u8(P.lab, R[u].v0 + g[u].x, g[u].n, bits);
        flag_group<S>(seed, R[u].tile0, g[u], bits);
      }
    }
  }
  sums.flush(dst, stride, buf);
}

void launch_cut_init(int shape, const Tiles &T, const Planes &P,
                     unsigned char *seed, long long *def, hipStream_t st, bool cleared) {
  if (!cleared) {
    launch_fill_u8(seed, T.ntiles, 0, st);
    launch_fill_i32((int *)def, 2LL * T.ninst, 0, st);
  }
  JFC_DISPATCH(shape,
               (k_cut_init<S><<<row_grid<&k_cut_init<S>>((T.nrows + T.row_step - 1) / T.row_step),
                                256, 0, st>>>(T, P, seed, def)));
}

// graphcut_init_global_relabel (graphcut_global_relabel.cl:19-48) as a pass
// over the nodes: h = 0 at deficit nodes, HINF elsewhere, written with the
// relabel's tag, and the exact deficit flags of the tiles (`defi`, zeroed by
// the launcher).  Used before the first tile pass of a 3D relabel; in 2D the
// first tile pass starts from e itself.
template <class S>
__global__ __launch_bounds__(256) void k_gr_init(Tiles T, Planes P, int tag,
                                                 unsigned char *defi) {
  const int zero = tag_word(0, tag), inf = tag_word(HINF, tag);
  const int nsteps = row_steps(T);
  for (int step = blockIdx.x; step < nsteps; step += gridDim.x) {
    const RowStep<S> RS(T, step);
    for (int k0 = threadIdx.x; k0 < RS.slots; k0 += RS_U * 256) {
      Group g[RS_U];
      RowRef<S> R[RS_U];
      int4 e[RS_U];
#pragma unroll
      for (int u = 0; u < RS_U; ++u) {
        const int k = k0 + u * 256;
        if (k < RS.slots) RS.at(T, k, R[u], g[u]);
        else { R[u] = RS.R0; g[u] = Group{0, 0}; }
        e[u] = load4(P.e, R[u].v0 + g[u].x, g[u].n);
      }
#pragma unroll
      for (int u = 0; u < RS_U; ++u) {
        if (g[u].n == 0) continue;
        int4 h;
        h.x = e[u].x < 0 ? zero : inf; h.y = e[u].y < 0 ? zero : inf;
        h.z = e[u].z < 0 ? zero : inf; h.w = e[u].w < 0 ? zero : inf;
        store4(P.h, R[u].v0 + g[u].x, g[u].n, h);
        unsigned bits = 0;
        for (int i = 0; i < g[u].n; ++i) bits |= (unsigned)(comp(e[u], i) < 0) << i;
        flag_group<S>(defi, R[u].tile0, g[u], bits);
      }
    }
  }
}

void launch_gr_init(int shape, const Tiles &T, const Planes &P, int tag, unsigned char *defi,
                    hipStream_t st, bool cleared) {
  if (!cleared) launch_fill_u8(defi, T.ntiles, 0, st);
  const int n = (T.nrows + T.row_step - 1) / T.row_step;
  JFC_DISPATCH(shape,
               (k_gr_init<S><<<row_grid<&k_gr_init<S>>(n), 256, 0, st>>>(T, P, tag, defi)));
}

// Cut energy on the input capacities (QGCWidget.cpp:2452-2520), summed per
// instance into energy[inst] (zeroed by the caller).
template <class S>
__global__ __launch_bounds__(256) void k_energy(Tiles T, Planes P, long long *energy) {
  __shared__ long long buf[4];
  const Lane<S> L(threadIdx.x);
  const Chunk C(T.ntiles);
  int cur = -1;
  long long a = 0;
  for (int t = C.t0; t < C.t1; ++t) {
    const int inst = T.tile_inst[t];
    if (inst != cur) {
      if (cur >= 0) {
        const long long A = block_sum_ll(a, buf);
        if (threadIdx.x == 0) atomicAdd((unsigned long long *)&energy[cur], (unsigned long long)A);
      }
      cur = inst;
      a = 0;
    }
    const TileRef r = tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    if (gx < I.nx && gy < I.ny && gz < I.nz) {
      const long long v = node_index(I, gx, gy, gz);
      if (P.lab[v]) {
        a += P.ct[v];
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) {
          if (d >= I.K) break;
          int dx, dy, dz;
          dir_of<S::D3>(d, dx, dy, dz);
          const int X = gx + dx, Y = gy + dy, Z = gz + dz;
          if (X < 0 || X >= I.nx || Y < 0 || Y >= I.ny || Z < 0 || Z >= I.nz) continue;
          if (!P.lab[node_index(I, X, Y, Z)]) a += P.c0[d * P.stride + v];
        }
      } else {
        a += P.cs[v];
      }
    }
  }
  if (cur >= 0) {
    const long long A = block_sum_ll(a, buf);
    if (threadIdx.x == 0) atomicAdd((unsigned long long *)&energy[cur], (unsigned long long)A);
  }
}

void launch_energy(int shape, const Tiles &T, const Planes &P, long long *energy,
                   hipStream_t st) {
  JFC_DISPATCH(shape, (k_energy<S><<<tile_grid(T.ntiles), 256, 0, st>>>(T, P, energy)));
}

// Subgradient of the cut energy for instance `inst` (outputs in the
// instance's own node numbering).
template <class S>
__global__ __launch_bounds__(256) void k_backward(Tiles T, Planes P, int t_begin,
                                                  int t_count, float *dcs,
                                                  float *dct, float *dnb,
                                                  long long user_total) {
  const Lane<S> L(threadIdx.x);
  for (int t = t_begin + blockIdx.x; t < t_begin + t_count; t += gridDim.x) {
    const TileRef r = tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    if (!(gx < I.nx && gy < I.ny && gz < I.nz)) continue;
    const long long v = node_index(I, gx, gy, gz);
    // outputs in the callers' numbering: one instance (user_total = 0) its
    // own, with plane stride = its node count; all instances the graph's
    // concatenation, with plane stride = all their nodes
    const long long own = gx + (long long)I.nx * (gy + (long long)I.ny * gz);
    const long long lv = user_total ? I.user_off + own : own;
    const long long n = user_total ? user_total : (long long)I.nx * I.ny * I.nz;
    const bool s = P.lab[v] != 0;
    if (dcs) dcs[lv] = s ? 0.f : 1.f;
    if (dct) dct[lv] = s ? 1.f : 0.f;
    if (dnb) {
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) {
        if (d >= I.K) break;
        int dx, dy, dz;
        dir_of<S::D3>(d, dx, dy, dz);
        const int X = gx + dx, Y = gy + dy, Z = gz + dz;
        float g = 0.f;
        if (s && X >= 0 && X < I.nx && Y >= 0 && Y < I.ny && Z >= 0 && Z < I.nz)
          g = P.lab[node_index(I, X, Y, Z)] ? 0.f : 1.f;
        dnb[d * n + lv] = g;
      }
    }
  }
}

void launch_backward(int shape, const Tiles &T, const Planes &P,
                     const InstDesc &I, float *dcs, float *dct, float *dnb,
                     hipStream_t st) {
  JFC_DISPATCH(shape, (k_backward<S><<<tile_grid(I.ntiles), 256, 0, st>>>(
                          T, P, I.tile_off, I.ntiles, dcs, dct, dnb, 0)));
}

void launch_backward_all(int shape, const Tiles &T, const Planes &P, float *dcs,
                         float *dct, float *dnb, long long user_total, hipStream_t st) {
  JFC_DISPATCH(shape, (k_backward<S><<<tile_grid(T.ntiles), 256, 0, st>>>(
                          T, P, 0, T.ntiles, dcs, dct, dnb, user_total)));
}

// ---- initialisation of relabel / cut searches --------------------------------

// graphcut_init_tag (graphcut_compute_tag.cl:19-49): the tag plane arrives
// holding the foreground seed bytes; it leaves as 0/1, and tiles with a seed
// seed the jump flooding.
template <class S>
__global__ __launch_bounds__(256) void k_tag_init(Tiles T, unsigned char *tag,
                                                  unsigned char *seed) {
  const Lane<S> L(threadIdx.x);
  const Chunk C(T.ntiles);
  __shared__ int s_w[8];
  int par = 0;
  for (int t = C.t0; t < C.t1; t += TU) {
    TileNode nd[TU];
    unsigned char f[TU];
#pragma unroll
    for (int u = 0; u < TU; ++u) nd[u] = tile_node<S>(T, t + u, C.t1, L);
#pragma unroll
    for (int u = 0; u < TU; ++u) f[u] = nd[u].valid ? tag[nd[u].v] : 0;
    int bits = 0;
#pragma unroll
    for (int u = 0; u < TU; ++u)
      if (nd[u].valid) {
        tag[nd[u].v] = (unsigned char)(f[u] != 0);
        bits |= (f[u] != 0) << u;
      }
    bits = block_or_bits(bits, s_w, par ^= 1);
    if (threadIdx.x < TU && t + (int)threadIdx.x < C.t1)
      seed[t + threadIdx.x] = (unsigned char)((bits >> threadIdx.x) & 1);
  }
}

void launch_tag_init(int shape, const Tiles &T, unsigned char *tag, unsigned char *seed,
                     hipStream_t st) {
  JFC_DISPATCH(shape, (k_tag_init<S><<<tile_grid(T.ntiles), 256, 0, st>>>(T, tag, seed)));
}

// ---- jump flooding over the tile grid + counting sort --------------------------

__device__ __forceinline__ void tile_xyz(const InstDesc &I, int lt, int &x, int &y,
                                         int &z) {
  x = lt % I.tx;
  const int q = lt / I.tx;
  y = q % I.ty;
  z = q / I.ty;
}

__global__ void k_jf_init(Tiles T, const unsigned char *seed, int *jf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.ntiles) return;
  const int i = T.tile_inst[t];
  jf[t] = seed[t] ? t - T.inst[i].tile_off : -1;
}

// One jump-flooding pass at jump `step`: keep the candidate seed with the
// smallest taxicab distance (ties: smallest tile index) among the 3^D
// neighbours at that jump (graphcut_bfs_block.cl:43-63).
__global__ void k_jf_pass(Tiles T, const int *in, int *out, int step, int d3) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.ntiles) return;
  const InstDesc I = T.inst[T.tile_inst[t]];
  int x, y, z;
  tile_xyz(I, t - I.tile_off, x, y, z);
  int best = in[t], bd = INT_MAX;
  if (best >= 0) {
    int sx, sy, sz;
    tile_xyz(I, best, sx, sy, sz);
    bd = abs(sx - x) + abs(sy - y) + abs(sz - z);
  }
  const int zr = d3 ? 1 : 0;
  for (int dz = -zr; dz <= zr; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (!dx && !dy && !dz) continue;
        const int X = x + dx * step, Y = y + dy * step, Z = z + dz * step;
        if (X < 0 || X >= I.tx || Y < 0 || Y >= I.ty || Z < 0 || Z >= I.tz) continue;
        const int c = in[I.tile_off + X + I.tx * (Y + I.ty * Z)];
        if (c < 0) continue;
        int sx, sy, sz;
        tile_xyz(I, c, sx, sy, sz);
        const int d = abs(sx - x) + abs(sy - y) + abs(sz - z);
        if (d < bd || (d == bd && c < best)) { bd = d; best = c; }
      }
  out[t] = best;
}

__global__ void k_jf_dist(Tiles T, const int *jf, int *tdist, int nbins) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.ntiles) return;
  const InstDesc I = T.inst[T.tile_inst[t]];
  const int c = jf[t];
  int d = nbins - 1;
  if (c >= 0) {
    int x, y, z, sx, sy, sz;
    tile_xyz(I, t - I.tile_off, x, y, z);
    tile_xyz(I, c, sx, sy, sz);
    d = min(abs(sx - x) + abs(sy - y) + abs(sz - z), nbins - 2);
  }
  tdist[t] = d;
}

// ---- the same distances by a separable taxicab distance transform ---------------
//
// The taxicab distance to the nearest seed tile separates by axis:
// d(p) = min over lines of the 1D transform g(p) = min_q (f(q) + |p - q|),
// applied along x, then y (then z).  On one line, g = min(F, B) with the
// prefix minimum F(p) = p + min_{q<=p} (f(q) - q) and the suffix minimum
// B(p) = -p + min_{q>=p} (f(q) + q).  One wavefront owns a line: each lane
// scans a contiguous chunk, and the chunk minima are combined across lanes
// by shuffles.  The result equals jump flooding's exactly, in D launches
// instead of 2 ceil(log2(max tile dim)) + 2.
constexpr int kDtBig = 1 << 28;   // "no seed yet" (far above any distance)

__device__ __forceinline__ int wave_excl_prefix_min(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v = min(v, u);
  }
  const int e = __shfl_up(v, 1);
  return lane == 0 ? INT_MAX : e;
}

__device__ __forceinline__ int wave_excl_suffix_min(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_down(v, o);
    if (lane + o < 64) v = min(v, u);
  }
  const int e = __shfl_down(v, 1);
  return lane == 63 ? INT_MAX : e;
}

// One axis of the transform over every line of every instance.  `lines`
// holds, for this axis, the prefix count of lines per instance.  FIRST reads
// the seed flags (0 / kDtBig); LAST writes the clamped distance bins of
// k_jf_dist (nbins - 1 for instances without a seed tile).  `out` also holds
// the prefix pass before the suffix pass combines with it.
template <int AXIS, bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_dt_line(Tiles T, const int *lines, int nlines,
                                                 const unsigned char *seed, const int *in,
                                                 int *out, int *tdist, int nbins) {
  const int line = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (line >= nlines) return;   // whole wave
  int lo = 0, hi = T.ninst - 1;  // instance: last with lines[i] <= line
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (lines[mid] <= line) lo = mid;
    else hi = mid - 1;
  }
  const InstDesc I = T.inst[lo];
  const int l = line - lines[lo];
  int L, step, base;
  if (AXIS == 0) {
    L = I.tx; step = 1; base = I.tx * l;                                  // l = y + ty z
  } else if (AXIS == 1) {
    L = I.ty; step = I.tx; base = (l % I.tx) + I.tx * I.ty * (l / I.tx);  // l = x + tx z
  } else {
    L = I.tz; step = I.tx * I.ty; base = l;                               // l = x + tx y
  }
  base += I.tile_off;
  const int lane = lane_id();
  const int c = (L + 63) >> 6;
  const int p0 = min(L, lane * c), p1 = min(L, p0 + c);
  auto val = [&](int p) -> int {
    const int t = base + p * step;
    return FIRST ? (seed[t] ? 0 : kDtBig) : in[t];
  };
  int m = INT_MAX;
  for (int p = p0; p < p1; ++p) m = min(m, val(p) - p);
  int run = wave_excl_prefix_min(m);
  for (int p = p0; p < p1; ++p) {
    run = min(run, val(p) - p);
    out[base + p * step] = run + p;
  }
  m = INT_MAX;
  for (int p = p0; p < p1; ++p) m = min(m, val(p) + p);
  run = wave_excl_suffix_min(m);
  for (int p = p1 - 1; p >= p0; --p) {
    run = min(run, val(p) + p);
    const int t = base + p * step;
    const int d = min(out[t], run - p);
    if (LAST) tdist[t] = d >= kDtBig / 2 ? nbins - 1 : min(d, nbins - 2);
    else out[t] = d;
  }
}

// Largest finite tile distance (reduce_histogram, reduce.cl:56-141).
__global__ void k_max_dist(const int *tdist, int n, int unreachable, int *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int d = (t < n && tdist[t] != unreachable) ? tdist[t] : 0;
  for (int o = 32; o > 0; o >>= 1) d = max(d, __shfl_xor(d, o));
  if (lane_id() == 0 && d > 0) atomicMax(out, d);
}

// Histogram of tile distances (graphcut_bfs_histogram): counted in LDS per
// workgroup, then one global atomic per non-empty bin -- most tiles share a
// few small distances, so direct global atomics would serialise on them.
__global__ __launch_bounds__(256) void k_hist(const int *tdist, int n, int nbins, int *hist) {
  extern __shared__ int sh[];
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int t0 = blockIdx.x * per, t1 = min(n, t0 + per);
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) atomicAdd(&sh[tdist[t]], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// In-place exclusive scan of n ints by one 1024-lane workgroup (the clogs
// Scan of QGCWidget.cpp:810, on a bin count small enough for one group).
__global__ __launch_bounds__(1024) void k_scan_1wg(int *a, int n) {
  __shared__ int s[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < n ? a[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int add = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    const int incl = s[threadIdx.x];
    const int c = carry;
    if (i < n) a[i] = c + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c + incl;
    __syncthreads();
  }
}

// Scatter tiles to their sorted position (graphcut_bfs_mapping): each
// workgroup reserves one run per non-empty bin, then places its tiles inside
// the runs by LDS rank.
__global__ __launch_bounds__(256) void k_scatter_order(const int *tdist, int n, int nbins,
                                                       int *cursor, int *order) {
  extern __shared__ int sh[];
  int *cnt = sh, *base = sh + nbins;
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int t0 = blockIdx.x * per, t1 = min(n, t0 + per);
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) atomicAdd(&cnt[tdist[t]], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
    base[i] = cnt[i] ? atomicAdd(&cursor[i], cnt[i]) : 0;
    cnt[i] = 0;
  }
  __syncthreads();
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int d = tdist[t];
    order[base[d] + atomicAdd(&cnt[d], 1)] = t;
  }
}

void launch_tile_order(int shape, const Tiles &T, const unsigned char *seed,
                       int *jf0, int *jf1, int *tdist, int *hist, int nbins,
                       int *order, int max_tile_dim, int *max_dist,
                       const TileLines &lines, int jump_flooding, hipStream_t st,
                       bool hist_cleared) {
  const int n = T.ntiles;
  const int g = (n + 255) / 256;
  const int d3 = shape_is_3d(shape) ? 1 : 0;
  if (jump_flooding) {
    k_jf_init<<<g, 256, 0, st>>>(T, seed, jf0);
    int *a = jf0, *b = jf1;
    int top = 1;
    while (top < max_tile_dim) top <<= 1;
    for (int step = top >> 1; step >= 1; step >>= 1)
      for (int rep = 0; rep < 2; ++rep) {
        k_jf_pass<<<g, 256, 0, st>>>(T, a, b, step, d3);
        int *tmp = a; a = b; b = tmp;
      }
    k_jf_dist<<<g, 256, 0, st>>>(T, a, tdist, nbins);
  } else {
    auto gl = [](int nl) { return (nl + 3) / 4; };
    k_dt_line<0, true, false><<<gl(lines.total[0]), 256, 0, st>>>(
        T, lines.prefix[0], lines.total[0], seed, nullptr, jf0, nullptr, nbins);
    if (d3) {
      k_dt_line<1, false, false><<<gl(lines.total[1]), 256, 0, st>>>(
          T, lines.prefix[1], lines.total[1], nullptr, jf0, jf1, nullptr, nbins);
      k_dt_line<2, false, true><<<gl(lines.total[2]), 256, 0, st>>>(
          T, lines.prefix[2], lines.total[2], nullptr, jf1, jf0, tdist, nbins);
    } else {
      k_dt_line<1, false, true><<<gl(lines.total[1]), 256, 0, st>>>(
          T, lines.prefix[1], lines.total[1], nullptr, jf0, jf1, tdist, nbins);
    }
  }
  if (max_dist) {
    launch_fill_i32(max_dist, 1, 0, st);
    k_max_dist<<<g, 256, 0, st>>>(tdist, n, nbins - 1, max_dist);
  }
  if (!hist_cleared) launch_fill_i32(hist, nbins, 0, st);
  const int gh = std::min(g, 512);
  k_hist<<<gh, 256, nbins * sizeof(int), st>>>(tdist, n, nbins, hist);
  k_scan_1wg<<<1, 1024, 0, st>>>(hist, nbins);
  k_scatter_order<<<gh, 256, 2 * nbins * sizeof(int), st>>>(tdist, n, nbins, hist, order);
}

// ---- list compaction (ballots + one atomic per 1024-lane block) ------------------
//
// Flagged entries of `src` (in order, or reversed) are appended to the queue
// list and their flags cleared; with `mat`, entries of tiles whose residual
// planes are already written carry kEntryMat, so the push kernel knows it
// without a dependent load.  Each wave counts its hits with a ballot; the
// block adds its total to the list count (and to the sweep and run totals)
// with one atomic each, so a pass over 10^5..10^6 tiles costs a few hundred
// same-address atomics.
constexpr int kCompactBlock = 1024;

__global__ __launch_bounds__(kCompactBlock) void k_compact(const int *src, int n, int reverse,
                                                           unsigned char *flag,
                                                           const unsigned char *mat,
                                                           const unsigned char *touched,
                                                           const unsigned char *hmat, Queue q,
                                                           int *sweep_cnt, long long *total,
                                                           const int *stop) {
  __shared__ int s_cnt[kCompactBlock / 64];
  __shared__ int s_base;
  // a sweep issued after the solve's decision to stop sweeping (a relabel is
  // due, no tile is active, or the sweep budget is spent) compacts nothing
  // and leaves the flags for the sweeps after the host has acted
  if (stop && *stop) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int t = -1;
  bool f = false;
  if (i < n) {
    t = src ? src[reverse ? n - 1 - i : i] : (reverse ? n - 1 - i : i);
    f = flag[t] != 0;
    if (f) flag[t] = 0;
  }
  const unsigned long long m = __ballot(f);
  const int w = threadIdx.x >> 6, lane = lane_id();
  if (lane == 0) s_cnt[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int k = 0; k < kCompactBlock / 64; ++k) sum += s_cnt[k];
    int base = 0;
    if (sum) {
      base = atomicAdd(q.ctr + QSLOT * q.slot, sum);
      if (sweep_cnt) atomicAdd(sweep_cnt, sum);
      if (total) atomicAdd((unsigned long long *)total, (unsigned long long)sum);
    }
    s_base = base;
  }
  __syncthreads();
  if (f) {
    int base = s_base;
    for (int k = 0; k < w; ++k) base += s_cnt[k];
    int ent = t;
    if (mat) ent |= mat[t] ? kEntryMat : (touched && !touched[t] ? kEntryClean : 0);
    if (hmat && hmat[t]) ent |= kEntryHeights;
    q.list[base + __popcll(m & ((1ull << lane) - 1ull))] = ent;
  }
}

// End of a sweep: file the sweep's active-tile count (accumulated by the
// compactions in c.state[0]) under the sweep index c.state[1], start the next
// sweep, and decide on the device what the host would decide after reading
// the count, so that the host can queue sweeps ahead of its checks:
//   stop = 1: no tile was active (converged, or the certificate decides);
//   stop = 2: the active set neither shrank nor grew by more than
//             (1 - stall_ratio) over the last sweep since the last relabel
//             (a global relabel is due, or the remaining excess is cut off
//             from every deficit);
//   stop = 3: the sweep budget is spent;
//   stop = 4: the tile visits since the last relabel exceed the budget (the
//             work of relabel_work visits per tile): a relabel is due.
// A sweep that starts with stop set compacts nothing and files nothing.
// Every sweep, no-op or not, posts its stop word to the host ring tagged with
// its sequence number in the solve, so that the host reads it with no copy
// and tells it from an older word in the same slot.
__device__ __forceinline__ void post_stop(const SweepCtl &c, int seq, int stop) {
  if (c.ring)
    __hip_atomic_store(c.ring + seq % c.ring_n, seq * 8 + stop, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_sweep_end(SweepCtl c) {
  int *st = c.state;
  const int seq = st[6];
  st[6] = seq + 1;
  if (st[2]) {
    post_stop(c, seq, st[2]);
    return;
  }
  const int i = st[1];
  const int cnt = st[0];
  if (i < c.cap) c.hist[i] = cnt;
  st[0] = 0;
  st[1] = i + 1;
  int stop = 0;
  if (cnt == 0) {
    stop = 1;
  } else if (i + 1 >= c.max_sweeps) {
    stop = 3;
  } else {
    const int since = i - st[3];  // sweeps since the last relabel, before this one
    // stalled: the active count repeats that of the sweep before, or two
    // sweeps running repeat the two before them (excess trading places
    // between two sets of tiles, sweep after sweep)
    auto near = [&](int a, int b) {
      return (double)a >= c.stall_ratio * b && a * c.stall_ratio <= (double)b;
    };
    bool stalled = false;
    if (c.stall_ratio > 0.0 && since >= 1 && i < c.cap) {
      stalled = near(cnt, c.hist[i - 1]);
      if (!stalled && since >= 3)
        stalled = near(cnt, c.hist[i - 2]) && near(c.hist[i - 1], c.hist[i - 3]);
    }
    const long long done = c.total[0] - *(const long long *)(st + 4);
    const bool budget = c.budget > 0 && (double)done > (double)c.budget;
    if (stalled) stop = 2;
    else if (budget) stop = 4;
  }
  st[2] = stop;
  post_stop(c, seq, stop);
}

void launch_sweep_end(const SweepCtl &c, hipStream_t st) { k_sweep_end<<<1, 1, 0, st>>>(c); }

// After a global relabel: sweeping resumes, and the relabel's sweep index
// and visit total are the base of the next stall and budget tests.
__global__ void k_sweep_resume(SweepCtl c) {
  int *st = c.state;
  st[2] = 0;
  st[3] = st[1];
  *(long long *)(st + 4) = c.total[0];
}

void launch_sweep_resume(const SweepCtl &c, hipStream_t st) {
  k_sweep_resume<<<1, 1, 0, st>>>(c);
}

void launch_compact(const int *src, int n, int reverse, unsigned char *flag,
                    const unsigned char *mat, const unsigned char *touched, const Queue &q,
                    int *sweep_cnt, long long *total, const int *stop, hipStream_t st,
                    const unsigned char *hmat) {
  if (n <= 0) return;
  k_compact<<<(n + kCompactBlock - 1) / kCompactBlock, kCompactBlock, 0, st>>>(
      src, n, reverse, flag, mat, touched, hmat, q, sweep_cnt, total, stop);
}

// ---- persistent tile queue ----------------------------------------------------------

// Workgroups claim runs of `chunk` consecutive list entries (16 for long
// lists, down to 1 for short ones).  Each workgroup's first run is static
// (group b takes run first(b) < G, G = grid size), so short or empty lists
// cost no atomics.  Later runs are claimed dynamically, which balances tiles
// of very different cost (a relabel visit iterates to its tile's fixpoint).
// Workgroups are dispatched to the 8 XCDs round-robin (b % 8), so with 8
// heads, head x serves the workgroups of XCD x: runs >= G form groups of 8
// consecutive runs, group q belongs to head q % nheads, and a head hands out
// its groups' runs in order.  Neighbouring tiles of a jump-flooding band
// thus stay on one XCD's L2, and claims spread over 8 addresses in 8
// different cache lines.  Every run
// belongs to one head and every head has workgroups, so the list is always
// drained.  When the static first runs already cover the list (G * chunk >=
// count, the short lists of late sweeps and relabel passes) no claim is
// issued at all: thousands of workgroups claiming on 8 addresses would
// serialise there.  Slot s of `ctr` holds the list count at [QSLOT*s] and
// head h at [QSLOT*s + QWORD*(1+h)]; the consumer of slot s zeroes the other
// slot's words for the next producer.
//
// VisitCursor walks a queue that way for the tile-visit kernels (push, relax,
// cut flood).  A workgroup learns its next list entry while it discharges
// the current tile, so that a visit starts with neither a global load nor a
// barrier of its own: thread 0 works out the next position during the local
// iterations (the rest of the current run, else the run its claim obtained,
// in which case it claims the run after that at once) and loads its entry;
// it publishes position and entry in LDS before the visit's last barrier,
// after which every thread reads them.  A direct list (Queue::pick) is
// walked in strides of the grid instead, workgroup b from position b, each
// workgroup taking the next position whose tile is picked.
struct VisitCursor {
  Queue q;
  int count, chunk, head_id, nheads;
  int *heads;
  const int *list;
  int ahead;             // (thread 0) claim for the run after the current one
  int nidx, nend, nent;  // (thread 0) the next position, its run's end, its entry
  int *s;                // LDS: next position, run end, entry
  int idx, end, ent;     // current position (uniform; idx < 0: drained)
  __device__ __forceinline__ VisitCursor(const Queue &q_, int *lds3) : q(q_) {
    if (blockIdx.x == 0 && threadIdx.x < 9) q.ctr[QSLOT * (q.slot ^ 1) + QWORD * threadIdx.x] = 0;
    count = q.fixed_count >= 0 ? q.fixed_count : q.ctr[QSLOT * q.slot];
    chunk = min(16, max(1, count / (int)(gridDim.x * 4)));
    heads = q.ctr + QSLOT * q.slot + QWORD;
    list = q.list;
    s = lds3;
    const int b = blockIdx.x;
    if (q.pick) {
      chunk = 1;
      ahead = -1;
      if (threadIdx.x == 0) {
        find(b);
        publish();
      }
      __syncthreads();
      advance();
      return;
    }
    nheads = (gridDim.x % 8 == 0) ? 8 : 1;
    head_id = b % nheads;
    int first = b;
    if (gridDim.x % 64 == 0) {
      const int w = b >> 3;
      first = (((w >> 3) << 3) + (b & 7)) * 8 + (w & 7);
    }
    idx = first * chunk;
    end = min(idx + chunk, count);
    ahead = -1;
    nidx = -1;
    nend = nent = 0;
    if (idx < count) {
      ent = __builtin_amdgcn_readfirstlane(list[idx]);
      if (threadIdx.x == 0 && (long long)gridDim.x * chunk < count)
        ahead = atomicAdd(&heads[QWORD * head_id], 1);
    } else {
      idx = -1;
      ent = 0;
    }
  }
  // thread 0, direct list: take position i's tile if it is picked (and the
  // solve has not stopped sweeping); its entry, or -1
  __device__ __forceinline__ int take(int i) const {
    const int t = list[i];
    if (!q.pick[t]) return -1;
    q.pick[t] = 0;
    int e = t;
    if (q.mat) e |= q.mat[t] ? kEntryMat : (q.touched && !q.touched[t] ? kEntryClean : 0);
    if (q.hmat && q.hmat[t]) e |= kEntryHeights;
    if (q.cnt) atomicAdd(q.cnt, 1);
    if (q.total) atomicAdd((unsigned long long *)q.total, 1ull);
    return e;
  }
  // thread 0, direct list: the first taken position from i on, in strides of
  // the grid
  __device__ __forceinline__ void find(int i) {
    nidx = -1;
    nend = nent = 0;
    if (q.stop && *q.stop) return;
    for (; i < count; i += (int)gridDim.x) {
      const int e = take(i);
      if (e >= 0) {
        nidx = i;
        nend = i + 1;
        nent = e;
        return;
      }
    }
  }
  // thread 0, during a visit: the position after the current one, and its entry
  __device__ __forceinline__ void prefetch() {
    if (q.pick) {
      find(idx + (int)gridDim.x);
      return;
    }
    nidx = -1;
    if (idx + 1 < end) {
      nidx = idx + 1;
      nend = end;
    } else if (ahead >= 0) {
      const int k = ahead;  // k-th run handed out by this head
      const int run = (int)gridDim.x + 8 * (head_id + nheads * (k >> 3)) + (k & 7);
      if ((long long)run * chunk < count) {
        nidx = run * chunk;
        nend = min(nidx + chunk, count);
        ahead = (long long)gridDim.x * chunk < count ? atomicAdd(&heads[QWORD * head_id], 1) : -1;
      }
    }
    if (nidx >= 0) nent = list[nidx];
  }
  // thread 0, after prefetch() and before the visit's last barrier
  __device__ __forceinline__ void publish() const {
    s[0] = nidx;
    s[1] = nend;
    s[2] = nent;
  }
  // every thread, after the visit's last barrier
  __device__ __forceinline__ void advance() {
    idx = __builtin_amdgcn_readfirstlane(s[0]);
    end = __builtin_amdgcn_readfirstlane(s[1]);
    ent = __builtin_amdgcn_readfirstlane(s[2]);
  }
};

// ---- residual-arc masks and lazy residual planes ----------------------------------
//
// rmask[v] bit d = (residual of arc d of v > 0); bit 8 + d = "fresh": the
// residual still equals the input c0[d][v] and cap[d][v] has not been
// written in this solve.  The relabel and cut searches only need the low
// bits, so they read 4 bytes per node instead of K planes.  mask0 holds the
// bits of the input planes (every plane fresh), built once per change of the
// inputs and copied to rmask when a solve starts, both inside the fold pass
// (k_fold); k_push reads fresh planes
// from c0 and writes them to cap, so no separate reset copy of the residual
// planes is needed.
// Cells of the padded tile a connectivity reads outside the tile: the sides
// and corners in 2D, the faces in 3D.  (The edges and corners of a 3D padded
// tile are never read: 6-connected arcs only cross faces.)
template <class S>
struct Ring {
  static constexpr int N = S::D3 ? 2 * (S::TW * S::TH + S::TW * S::TD + S::TH * S::TD)
                                 : 2 * S::PW + 2 * S::TH;
  static __device__ __forceinline__ int pos(int k) {
    if constexpr (!S::D3) {
      if (k < S::PW) return k;
      k -= S::PW;
      if (k < S::PW) return (S::PH - 1) * S::PW + k;
      k -= S::PW;
      if (k < S::TH) return (k + 1) * S::PW;
      k -= S::TH;
      return (k + 1) * S::PW + S::PW - 1;
    } else {
      constexpr int XY = S::TW * S::TH, XZ = S::TW * S::TD, YZ = S::TH * S::TD;
      int x, y, z;
      if (k < 2 * XY) {
        const int q = k % XY;
        x = q % S::TW + 1;
        y = q / S::TW + 1;
        z = k < XY ? 0 : S::PD - 1;
      } else if (k < 2 * XY + 2 * XZ) {
        k -= 2 * XY;
        const int q = k % XZ;
        x = q % S::TW + 1;
        z = q / S::TW + 1;
        y = k < XZ ? 0 : S::PH - 1;
      } else {
        k -= 2 * XY + 2 * XZ;
        const int q = k % YZ;
        y = q % S::TH + 1;
        z = q / S::TH + 1;
        x = k < YZ ? 0 : S::PW - 1;
      }
      return x + S::PW * (y + S::PH * z);
    }
  }
};

// The ring split over the four waves of a 256-lane workgroup: wave w owns the
// kRingQ consecutive cells from w * kRingQ (consecutive cells are neighbours
// along a row or face, so a wave's loads stay on few lines), its lanes
// taking kRingPerLane of them each.
template <class S>
constexpr int kRingQ = (Ring<S>::N + 3) / 4;
template <class S>
constexpr int kRingPerLane = (kRingQ<S> + 63) / 64;

// A 3D ring cell (a face cell of the padded tile) has one arc into the tile
// (6-connected arcs cross faces only): its direction from the cell, which is
// also the one plane of the cell that flow out of the tile changes.
template <class S>
__device__ __forceinline__ int inward_dir(int p) {
  int x, y, z, nout;
  bool inter;
  pad_coords<S>(p, x, y, z, inter, nout);
  return x < 0 ? 0 : x >= S::TW ? 1 : y < 0 ? 2 : y >= S::TH ? 3 : z < 0 ? 4 : 5;
}

// ---- block-wise push-relabel -------------------------------------------------------
//
// One workgroup discharges one tile at a time: e, h and the K residual planes
// of the tile go to LDS with a halo of neighbour heights.  Each local
// iteration relabels every active node (e > 0, finite h) to
// min over residual arcs of h_nb + 1 when it has no admissible arc, then
// pushes min(e, c) along arcs with h == h_nb + 1 in direction order.  Heights
// are not written during the push phase, so an arc is never pushed both ways
// at once and the reverse residual of an arc is touched by one lane only;
// only excess needs LDS atomics.  Flow pushed into halo cells accumulates in
// their LDS slots and is flushed to HBM once per visit.  The loop ends when a
// whole iteration pushes nothing (graphcut_push_relabel.cl:77-182) or after
// max_iters iterations -- sparse_iters when the launch's list is no longer
// than its grid and the solve has run kSparseAfter sweeps (every workgroup
// holds at most one tile, so a visit that keeps iterating delays no other
// tile, and the flow crosses the tile in one visit instead of one sweep per
// step).
// ONE: a one-instance graph, whose descriptor every visit reads from the
// kernel arguments instead of loading it behind the tile index.
template <class S, bool ONE>
// 7 waves per SIMD: with one ring cell per lane the visit fits 69 VGPRs
// without spills (the 15000x10000 8-connected cut: 4.67 ms at 5 waves, 4.54
// at 6, 4.48 at 7, 4.54 at 8, where it spills).  Two ring cells per lane
// (32x4x2, slim flush) fit 96: 5 waves.
__global__ __launch_bounds__(256, kRingPerLane<S> > 1 ? 5 : 7) void k_push(
    Tiles T, Planes P, Queue q, unsigned char *flag, unsigned char *mat, unsigned char *touched,
    unsigned char *defi, unsigned char *hmat, int max_iters, int sparse_iters,
    const int *sweep_state, long long *iter_total) {
  __shared__ int s_h[S::PADN];
  __shared__ int s_e[S::PADN];
  __shared__ int s_c[S::KMAX][S::PADN];
  // Two ring cells per lane (3D 32x4x2): the flush finds its cells again
  // from the lane index and their mask words in LDS, and takes flow back on
  // a cell's one inward plane, so no ring state is held in registers
  // through the local iterations (4 waves per SIMD otherwise).
  constexpr bool kSlim = S::D3 && kRingPerLane<S> > 1;
  __shared__ unsigned s_hm[kSlim ? S::PADN : 1];
  const int tid = threadIdx.x;
  const Lane<S> L(tid);
  __shared__ int s_next[3];
  __shared__ int s_w[8];
  int par = 0;
  VisitCursor Qc(q, s_next);
  // the sparse cap from the solve's kSparseAfter-th sweep on: a solve that
  // ends within a few sweeps has nothing to gain from long visits
  const int cap_iters =
      Qc.count <= (int)gridDim.x && sweep_state[1] >= kSparseAfter ? sparse_iters : max_iters;
  const long long st = P.stride;
  long long my_iters = 0;
  while (Qc.idx >= 0) {
    const int ent = Qc.ent;
    const int t = ent & kEntryTile;
    const TileRef r = ONE ? tile_ref_of<S>(T.one, t) : tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int K = I.K;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    const bool valid = gx < I.nx && gy < I.ny && gz < I.nz;
    // A valid label never exceeds the distance to a deficit node, which is
    // below the node count; a label that reaches it proves the node cut off
    // (the role of cl_max_height in graphcut_push_relabel.cl, made exact).
    const long long nn = (long long)I.nx * I.ny * I.nz;
    const int hcap = nn < HINF ? (int)nn : HINF;
    // A tile whose every plane is already written in this solve (kEntryMat,
    // from mat[t]) loads cap directly; a tile no visit or flush has written
    // yet (kEntryClean) loads c0 directly, and its mask words are those of
    // the inputs (every plane fresh, residual bits c0 > 0), so neither needs
    // its mask word before its plane loads; otherwise the per-plane fresh
    // bits pick c0 or cap.
    const bool fresh_tile = (ent & kEntryMat) == 0;
    const bool clean = (ent & kEntryClean) != 0;

    // ---- loads: every global load of the visit is issued before any is
    // used; lanes off the grid, planes d >= K and dead halo cells read a
    // valid dummy address and discard the value.
    const long long gv = valid ? node_index(I, gx, gy, gz) : 0;
    long long ca[S::KMAX];
#pragma unroll
    for (int d = 0; d < S::KMAX; ++d) ca[d] = (valid && d < K) ? d * st + gv : 0;
    // halo: the ring cells (sides and corners in 2D, faces in 3D), split
    // evenly over the four waves so that each wave loads, clears and flushes
    // the same number of cells; the workgroup barrier after the flush then
    // waits for no straggling wave
    HaloCell hc[kRingPerLane<S>];
    int hp[kRingPerLane<S>];
    int hh[kRingPerLane<S>];
    unsigned hm[kRingPerLane<S>];
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      const int j = c * 64 + (tid & 63);
      const int k = (tid >> 6) * kRingQ<S> + j;
      hp[c] = j < kRingQ<S> && k < Ring<S>::N ? Ring<S>::pos(k) : S::PADN;
      hc[c] = halo_cell<S>(I, r.ox, r.oy, r.oz, hp[c], K);
      // a neighbour tile without stored heights in this solve reads as 0
      int ht = t;
      if (hc[c].live) {
        int x, y, z, nout;
        bool inter;
        pad_coords<S>(hp[c], x, y, z, inter, nout);
        ht = tile_of<S>(I, r.ox + x, r.oy + y, r.oz + z);
      }
      const int hw = P.h[hc[c].g];
      hh[c] = hmat[ht] ? hw & HMASK : 0;
      hm[c] = P.rmask[hc[c].g];
    }
    // heights not yet stored in this solve read as 0 (the start labelling);
    // the visit then stores all of them.  The compaction read hmat[t] into
    // the entry, so no load stands between the entry and the plane loads.
    const bool h_stored = (ent & kEntryHeights) != 0;
    const int hw0 = P.h[gv];
    int e0 = P.e[gv], h0 = h_stored ? hw0 & HMASK : 0;
    unsigned rm = 0;
    int c0r[S::KMAX];
    if (!fresh_tile) {
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) c0r[d] = P.cap[ca[d]];
    } else if (clean) {
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) c0r[d] = P.c0[ca[d]];
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d)
        if (d < K) rm |= (1u << (8 + d)) | ((unsigned)(c0r[d] > 0) << d);
    } else {
      rm = P.rmask[gv];
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) c0r[d] = ((rm >> (8 + d)) & 1u ? P.c0 : P.cap)[ca[d]];
    }
    if (!valid) {
      e0 = 0;
      h0 = HINF;
      rm = 0;
    }
#pragma unroll
    for (int d = 0; d < S::KMAX; ++d)
      if (!(valid && d < K)) c0r[d] = 0;

    s_e[L.me] = e0;
    s_h[L.me] = h0;
#pragma unroll
    for (int d = 0; d < S::KMAX; ++d) s_c[d][L.me] = c0r[d];
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      if (!hc[c].halo) continue;
      const int p = hp[c];
      s_h[p] = hc[c].live ? hh[c] : HINF;
      s_e[p] = 0;
      if constexpr (kSlim) s_hm[p] = hm[c];
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) s_c[d][p] = 0;
    }
    __syncthreads();
    // the next entry loads while the tile iterates
    if (tid == 0) Qc.prefetch();

    int it = 0, any;
    do {
      // relabel (monotone; neighbours may be read before or after their own
      // relabel in this iteration, which keeps the labelling valid).  Every
      // LDS word a phase needs is read before any is used, so a phase waits
      // on LDS once instead of once per direction.
      if (valid) {
        const int e = s_e[L.me];
        const int h = s_h[L.me];
        if (e > 0 && h < HINF) {
          int cv[S::KMAX], hv[S::KMAX];
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d) {
            cv[d] = s_c[d][L.me];
            hv[d] = s_h[L.me + lds_off<S>(d)];
          }
          int m = HINF;
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d)
            if (d < K && cv[d] > 0) m = min(m, hv[d] + 1);
          if (m >= hcap) m = HINF;
          if (m > h) s_h[L.me] = m;
        }
      }
      __syncthreads();
      // push.  The reverse residual s_c[d^1][nb] is written only by this
      // lane during the phase (its owner nb cannot push back along an arc
      // this lane pushes on: heights are frozen), so it is read up front too.
      int pushed = 0;
      if (valid) {
        const int e = s_e[L.me];
        const int h = s_h[L.me];
        if (e > 0 && h < HINF) {
          int cv[S::KMAX], hv[S::KMAX], rv[S::KMAX];
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d) {
            const int nb = L.me + lds_off<S>(d);
            cv[d] = s_c[d][L.me];
            hv[d] = s_h[nb];
            rv[d] = s_c[d ^ 1][nb];
          }
          int rem = e;
#pragma unroll
          for (int d = 0; d < S::KMAX; ++d) {
            if (d < K && rem > 0) {
              const int nb = L.me + lds_off<S>(d);
              const int c = cv[d];
              if (c > 0 && h == hv[d] + 1) {
                const int f = min(rem, c);
                s_c[d][L.me] = c - f;
                s_c[d ^ 1][nb] = rv[d] + f;
                atomicAdd(&s_e[nb], f);
                rem -= f;
              }
            }
          }
          if (rem != e) {
            atomicAdd(&s_e[L.me], rem - e);
            pushed = 1;
          }
        }
      }
      any = block_any(pushed, s_w, par);
    } while (any && ++it < cap_iters);
    my_iters += it + (any ? 0 : 1);
    if (tid == 0) Qc.publish();

    // ---- the flush's reads first: the flow that crossed into each ring cell
    // and the reverse-residual increments (LDS), and for increments into a
    // plane that is still fresh (its value lives in c0) the c0 value they add
    // to, loaded before the interior write-back so that its stores cover the
    // load latency.
    int fl[kRingPerLane<S>];
    int dc[kRingPerLane<S>][kSlim ? 1 : S::KMAX], cv[kRingPerLane<S>][kSlim ? 1 : S::KMAX];
    unsigned fr[kRingPerLane<S>], fp[kRingPerLane<S>];
    if constexpr (kSlim) {
      // the ring cells found again (an opaque copy of the lane index keeps
      // the compiler from holding the load phase's values instead)
      int tid2 = tid;
      asm volatile("" : "+v"(tid2));
#pragma unroll
      for (int c = 0; c < kRingPerLane<S>; ++c) {
        const int j = c * 64 + (tid2 & 63);
        const int k = (tid2 >> 6) * kRingQ<S> + j;
        hp[c] = j < kRingQ<S> && k < Ring<S>::N ? Ring<S>::pos(k) : S::PADN;
        hc[c] = halo_cell<S>(I, r.ox, r.oy, r.oz, hp[c], K);
      }
    }
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      fl[c] = 0;
      fr[c] = fp[c] = 0;
      if constexpr (!kSlim) {
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) dc[c][d] = cv[c][d] = 0;
      } else {
        dc[c][0] = cv[c][0] = 0;
      }
      if (!hc[c].live) continue;
      const int p = hp[c];
      fl[c] = s_e[p];
      if constexpr (kSlim) {
        const int d = inward_dir<S>(p);
        const unsigned m = s_hm[p];
        dc[c][0] = s_c[d][p];
        if (dc[c][0] != 0) {
          if ((m >> (8 + d)) & 1u) fr[c] = 1u << d;
          if (!((m >> d) & 1u)) fp[c] = 1u << d;
        }
        if (fr[c]) cv[c][0] = P.c0[d * st + hc[c].g];
        continue;
      } else {
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) {
        dc[c][d] = d < K ? s_c[d][p] : 0;
        if (dc[c][d] != 0) {
          if ((hm[c] >> (8 + d)) & 1u) fr[c] |= 1u << d;
          if (!((hm[c] >> d) & 1u)) fp[c] |= 1u << d;
        }
      }
      if (fr[c]) {
        const long long g = hc[c].g;
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) cv[c][d] = P.c0[((fr[c] >> d) & 1u) ? d * st + g : 0];
      }
      }
    }

    // ---- write back what changed (stores only)
    int act = 0, dfc = 0;
    if (valid) {
      const int e1 = s_e[L.me], h1 = s_h[L.me];
      if (e1 != e0) P.e[gv] = e1;
      if (h1 != h0 || !h_stored) P.h[gv] = h1;
      // fresh planes are written whether or not they changed; the mask word
      // of a written tile holds residual bits only (exact, no fresh bits)
      unsigned m0 = 0, m1 = 0;
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d)
        if (d < K) {
          const int c1 = s_c[d][L.me];
          if (c1 != c0r[d] || ((rm >> (8 + d)) & 1u)) P.cap[d * st + gv] = c1;
          m0 |= (unsigned)(c0r[d] > 0) << d;
          m1 |= (unsigned)(c1 > 0) << d;
        }
      if (m1 != (fresh_tile ? rm : m0)) P.rmask[gv] = m1;
      act = e1 > 0 && h1 < HINF;
      dfc = e1 < 0;
    }
    // ---- flush flow that crossed into the halo: excess and the reverse
    // residual of each crossing arc (graphcut_push_relabel.cl:184-195).
    // Plane d of halo node g is the arc from g into this tile, so no other
    // tile of this colour class touches plane d of g or its two mask bits in
    // this launch; other tiles may update other bits of the same mask word,
    // hence atomics.  A plane that is still fresh is written as c0 + dc; the
    // others take a fire-and-forget atomic add.
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      if (!hc[c].live) continue;
      const int p = hp[c];
      const long long g = hc[c].g;
      if (fl[c] != 0) {
        // flow entered g, so some plane of g changes below: g's tile is no
        // longer clean
        atomicAdd(&P.e[g], fl[c]);
        const int ht = tile_of<S>(I, r.ox + (p % S::PW) - 1, r.oy + (p / S::PW) % S::PH - 1,
                                  S::D3 ? r.oz + p / (S::PW * S::PH) - 1 : 0);
        flag[ht] = 1;
        touched[ht] = 1;
      }
      if constexpr (kSlim) {
        const long long a = inward_dir<S>(p) * st + g;
        if (fr[c]) P.cap[a] = cv[c][0] + dc[c][0];
        else if (dc[c][0] != 0) atomicAdd(&P.cap[a], dc[c][0]);
      } else {
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) {
          if ((fr[c] >> d) & 1u) P.cap[d * st + g] = cv[c][d] + dc[c][d];
          else if (dc[c][d] != 0) atomicAdd(&P.cap[d * st + g], dc[c][d]);
        }
      }
      const unsigned flip = fp[c] | (fr[c] << 8);
      if (flip) atomicXor(&P.rmask[g], flip);
    }
    // the tile's active flag and deficit flag, one barrier for both
    const int ad = block_any2(act | (dfc << 1), s_w, par);
    if (tid == 0) {
      if (ad & 1) flag[t] = 1;
      defi[t] = (unsigned char)(ad >> 1);
    }
    if (fresh_tile && tid == 0) mat[t] = 1;
    if (!h_stored && tid == 0) hmat[t] = 1;
    Qc.advance();
  }
  // one atomic per workgroup, spread over 64 counters
  if (tid == 0 && iter_total && my_iters)
    atomicAdd((unsigned long long *)&iter_total[blockIdx.x & 63], (unsigned long long)my_iters);
}

void launch_push(int shape, int nwg, const Tiles &T, const Planes &P,
                 const Queue &q, unsigned char *flag, unsigned char *mat,
                 unsigned char *touched, unsigned char *defi, unsigned char *hmat,
                 int max_iters, int sparse_iters, const int *sweep_state,
                 long long *iter_total, hipStream_t st) {
  if (T.ninst == 1) {
    JFC_DISPATCH(shape, (k_push<S, true><<<resident_grid<&k_push<S, true>>(nwg), 256, 0, st>>>(
                            T, P, q, flag, mat, touched, defi, hmat, max_iters, sparse_iters,
                            sweep_state, iter_total)));
  } else {
    JFC_DISPATCH(shape, (k_push<S, false><<<resident_grid<&k_push<S, false>>(nwg), 256, 0, st>>>(
                            T, P, q, flag, mat, touched, defi, hmat, max_iters, sparse_iters,
                            sweep_state, iter_total)));
  }
}

// ---- exact global relabel: tile-local Bellman-Ford to a fixpoint ------------------
//
// h(u) = min(h(u), min over arcs with residual > 0 of h(v) + 1), iterated in
// LDS until the tile is stable; tiles whose halo changed are revisited.  The
// fixpoint over all tiles is the exact breadth-first distance to the nearest
// deficit node (graphcut_global_relabel.cl:50-124).
template <class S, bool FIRST, bool ONE>
__global__ __launch_bounds__(256) void k_relax(Tiles T, Planes P, Queue q,
                                               unsigned char *dirty,
                                               unsigned char *active, int tag,
                                               const InstDesc *pinst, int pdims) {
  constexpr bool first = FIRST;
  __shared__ int s_h[S::PADN];
  const int tid = threadIdx.x;
  const Lane<S> L(tid);
  __shared__ int s_next[3];
  __shared__ int s_w[8];
  int par = 0;
  VisitCursor Qc(q, s_next);
  while (Qc.idx >= 0) {
    const int t = Qc.ent;
    const TileRef r = ONE ? tile_ref_of<S>(T.one, t) : tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int K = I.K;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    const bool valid = gx < I.nx && gy < I.ny && gz < I.nz;
    // every load of the visit is issued before any is used (dead lanes and
    // halo cells read a valid dummy address)
    const long long gv = valid ? node_index(I, gx, gy, gz) : 0;
    // halo: the ring cells, split evenly over the four waves (as in k_push)
    HaloCell hc[kRingPerLane<S>];
    int hp[kRingPerLane<S>];
    int hh[kRingPerLane<S>], he[kRingPerLane<S>];
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      const int j = c * 64 + (tid & 63);
      const int k = (tid >> 6) * kRingQ<S> + j;
      hp[c] = j < kRingQ<S> && k < Ring<S>::N ? Ring<S>::pos(k) : S::PADN;
      hc[c] = halo_cell<S>(I, r.ox, r.oy, r.oz, hp[c], K);
      // device-coherent read: a neighbour tile settled earlier in this same
      // pass (possibly on another XCD) is seen at once, so one pass carries
      // the distance front across many tiles in list order.  Any tagged
      // value read is an upper bound of the distance, so stale is safe.
      hh[c] = __hip_atomic_load(&P.h[hc[c].g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      he[c] = first ? P.e[hc[c].g] : 0;
    }
    // the first pass starts from e (the relabel's initial labelling) and
    // needs no stored height
    int e0 = P.e[gv];
    int h0 = first ? 0 : P.h[gv] & HMASK;
    unsigned mask = P.rmask[gv] & ((1u << K) - 1u);
    if (first) h0 = e0 < 0 ? 0 : HINF;
    if (!valid) {
      h0 = HINF;
      e0 = 0;
      mask = 0;
    }
    s_h[L.me] = h0;
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      if (!hc[c].halo) continue;
      int v = HINF;
      if (hc[c].live) {
        // an untagged word is left from the pushes or the relabel before:
        // the neighbour still holds its start value
        v = ((unsigned)hh[c] >> 30) == (unsigned)tag ? hh[c] & HMASK : (he[c] < 0 ? 0 : HINF);
      }
      s_h[hp[c]] = v;
    }
    __syncthreads();
    // the next entry loads while the tile iterates
    if (tid == 0) Qc.prefetch();
    int any;
    do {
      int ch = 0;
      if (mask) {
        // all neighbour words read before any is used (one LDS wait)
        const int h = s_h[L.me];
        int hv[S::KMAX];
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) hv[d] = s_h[L.me + lds_off<S>(d)];
        int m = h;
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d)
          if (mask & (1u << d)) m = min(m, hv[d] + 1);
        if (m < h) {
          s_h[L.me] = m;
          ch = 1;
        }
      }
      any = block_any(ch, s_w, par);
    } while (any);
    if (valid) {
      const int h1 = s_h[L.me];
      // the first pass writes every word (with the tag); later passes the
      // changed ones
      if (first || h1 != h0)
        __hip_atomic_store(&P.h[gv], tag_word(h1, tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (h1 != h0) {
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) {
          if (d >= K) break;
          int dx, dy, dz;
          dir_of<S::D3>(d, dx, dy, dz);
          const int lx = L.lx + dx, ly = L.ly + dy, lz = L.lz + dz;
          if (lx >= 0 && lx < S::TW && ly >= 0 && ly < S::TH && lz >= 0 && lz < S::TD) continue;
          const int X = gx + dx, Y = gy + dy, Z = gz + dz;
          if (X < 0 || X >= I.nx || Y < 0 || Y >= I.ny || Z < 0 || Z >= I.nz) continue;
          dirty[tile_of<S>(I, X, Y, Z)] = 1;
        }
      }
    }
    // graphcut_count_node fused in: the tile's last visit leaves its
    // active flag (a node with excess and a finite height)
    if (tid == 0) Qc.publish();
    const bool on = valid && e0 > 0 && s_h[L.me] < HINF;
    const int act = block_any(on, s_w, par);
    if (pinst) {
      // flags of the push tiling: each active node sets its push tile's.
      // During a relabel e is fixed and heights only fall, so a node once
      // active stays active, and setting flags (never clearing) is exact.
      if (on) {
        const InstDesc &PI = pinst[ONE ? 0 : T.tile_inst[t]];
        active[PI.tile_off + (gx >> (pdims & 15)) +
               PI.tx * ((gy >> ((pdims >> 4) & 15)) + PI.ty * (gz >> (pdims >> 8)))] = 1;
      }
    } else if (tid == 0) {
      active[t] = (unsigned char)act;
    }
    Qc.advance();
  }
}

// One wavefront per tile.  A relax visit is bound by its load latency (the
// local fixpoint is a few cheap LDS iterations), so the tiles in flight per
// CU matter more than the lanes per tile: a 64-lane wave owning a tile keeps
// up to four times as many tiles in flight as a 256-lane workgroup.  Lane l
// holds the nodes l + 64 j; the iterations are ordered by program order in
// the wave, and a lane's later nodes already see its earlier nodes' new
// heights (Gauss-Seidel within the iteration, still monotone).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


template <class S, bool FIRST, bool ONE>
__global__ __launch_bounds__(64) void k_relax_w(Tiles T, Planes P, Queue q,
                                                unsigned char *dirty, unsigned char *active,
                                                int tag) {
  constexpr bool first = FIRST;
  constexpr int NJ = S::NT / 64;
  constexpr int NR = (Ring<S>::N + 63) / 64;
  __shared__ int s_h[S::PADN];
  __shared__ int s_next[3];
  const int lane = threadIdx.x;
  VisitCursor Qc(q, s_next);
  while (Qc.idx >= 0) {
    const int t = Qc.ent;
    const TileRef r = ONE ? tile_ref_of<S>(T.one, t) : tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int K = I.K;
    // loads: ring heights (device-coherent and tagged, as in k_relax) and
    // the lane's nodes
    int hh[NR], he[NR];
    bool hlive[NR];
    long long hg[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int k = lane + 64 * i;
      const HaloCell hc =
          halo_cell<S>(I, r.ox, r.oy, r.oz, k < Ring<S>::N ? Ring<S>::pos(k) : S::PADN, K);
      hlive[i] = hc.live;
      hg[i] = hc.g;
      hh[i] = __hip_atomic_load(&P.h[hc.g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      he[i] = 0;
    }
    int h0[NJ];
    unsigned mk[NJ];
    bool pos[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const Lane<S> L(lane + 64 * j);
      const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
      const bool valid = gx < I.nx && gy < I.ny && gz < I.nz;
      const long long gv = valid ? node_index(I, gx, gy, gz) : 0;
      const int e0 = P.e[gv];
      h0[j] = first ? 0 : P.h[gv] & HMASK;
      mk[j] = P.rmask[gv] & ((1u << K) - 1u);
      pos[j] = e0 > 0;
      if (first) h0[j] = e0 < 0 ? 0 : HINF;
      if (!valid) {
        h0[j] = HINF;
        mk[j] = 0;
        pos[j] = false;
      }
    }
    // first pass: the excess of a halo node is needed only where its height
    // word is stale, so it is loaded after the word and only there (most of
    // a 2D tile's halo loads are side-column cells, one 128-byte line each)
    if (first) {
#pragma unroll
      for (int i = 0; i < NR; ++i)
        if (hlive[i] && ((unsigned)hh[i] >> 30) != (unsigned)tag) he[i] = P.e[hg[i]];
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) s_h[Lane<S>(lane + 64 * j).me] = h0[j];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int k = lane + 64 * i;
      if (k >= Ring<S>::N) continue;
      int v = HINF;
      if (hlive[i])
        v = ((unsigned)hh[i] >> 30) == (unsigned)tag ? hh[i] & HMASK : (he[i] < 0 ? 0 : HINF);
      s_h[Ring<S>::pos(k)] = v;
    }
    wave_lds_sync();
    // the next entry loads while the tile iterates
    if (lane == 0) Qc.prefetch();
    bool any;
    do {
      bool ch = false;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        if (!mk[j]) continue;
        const int me = Lane<S>(lane + 64 * j).me;
        const int h = s_h[me];
        int hv[S::KMAX];
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) hv[d] = s_h[me + lds_off<S>(d)];
        int m = h;
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d)
          if (mk[j] & (1u << d)) m = min(m, hv[d] + 1);
        if (m < h) {
          s_h[me] = m;
          ch = true;
        }
      }
      any = __ballot(ch) != 0;
      wave_lds_sync();
    } while (any);
    bool act = false;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const Lane<S> L(lane + 64 * j);
      const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
      if (!(gx < I.nx && gy < I.ny && gz < I.nz)) continue;
      const int h1 = s_h[L.me];
      act |= pos[j] && h1 < HINF;
      if (first || h1 != h0[j])
        __hip_atomic_store(&P.h[node_index(I, gx, gy, gz)], tag_word(h1, tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (h1 == h0[j]) continue;
#pragma unroll
      for (int d = 0; d < S::KMAX; ++d) {
        if (d >= K) break;
        int dx, dy, dz;
        dir_of<S::D3>(d, dx, dy, dz);
        const int lx = L.lx + dx, ly = L.ly + dy, lz = L.lz + dz;
        if (lx >= 0 && lx < S::TW && ly >= 0 && ly < S::TH && lz >= 0 && lz < S::TD) continue;
        const int X = gx + dx, Y = gy + dy, Z = gz + dz;
        if (X < 0 || X >= I.nx || Y < 0 || Y >= I.ny || Z < 0 || Z >= I.nz) continue;
        dirty[tile_of<S>(I, X, Y, Z)] = 1;
      }
    }
    const bool a = __ballot(act) != 0;
    if (lane == 0) {
      active[t] = (unsigned char)a;
      Qc.publish();
    }
    wave_lds_sync();
    Qc.advance();
  }
}

// 2D tiles: the wave-per-tile kernel (a relabel of the 15000x10000
// 8-connected cut 1.67 -> 1.21 ms).
// 3D tiles iterate more per visit (fronts cross 8x8x4 tiles in z as well)
// and there the 256-lane visit's four parallel waves win (the 512^3
// volume's relabels 19 -> 15 ms), so 3D keeps the workgroup kernel.
template <bool FIRST, bool ONE>
static void relax_launch(int shape, int nwg, const Tiles &T, const Planes &P, const Queue &q,
                         unsigned char *dirty, unsigned char *active, int tag, hipStream_t st,
                         const InstDesc *pinst, int pdims) {
  if (shape_is_3d(shape)) {
    JFC_DISPATCH(shape, (k_relax<S, FIRST, ONE><<<resident_grid<&k_relax<S, FIRST, ONE>>(nwg),
                                                  256, 0, st>>>(T, P, q, dirty, active, tag,
                                                                pinst, pdims)));
  } else {
    JFC_DISPATCH(shape,
                 (k_relax_w<S, FIRST, ONE><<<resident_grid_b<&k_relax_w<S, FIRST, ONE>, 64>(
                                                 4 * nwg),
                                             64, 0, st>>>(T, P, q, dirty, active, tag)));
  }
}

void launch_relax(int shape, int nwg, const Tiles &T, const Planes &P,
                  const Queue &q, unsigned char *dirty, unsigned char *active,
                  bool first_pass, int tag, hipStream_t st, const InstDesc *pinst, int pdims) {
  if (first_pass) {
    if (T.ninst == 1)
      relax_launch<true, true>(shape, nwg, T, P, q, dirty, active, tag, st, pinst, pdims);
    else relax_launch<true, false>(shape, nwg, T, P, q, dirty, active, tag, st, pinst, pdims);
  } else {
    if (T.ninst == 1)
      relax_launch<false, true>(shape, nwg, T, P, q, dirty, active, tag, st, pinst, pdims);
    else relax_launch<false, false>(shape, nwg, T, P, q, dirty, active, tag, st, pinst, pdims);
  }
}

// ---- canonical cut: residual reachability from excess nodes -------------------------
//
// A labelled node labels every neighbour it has residual capacity toward
// (graphcut_compute_cut.cl:95-113, here from the tail of the arc so only the
// node's own planes are read).  Labels entering the halo are written to HBM
// and mark the neighbour tile for the next pass.  A labelled node with a
// deficit means an augmenting path still exists: *cert is set
// (graphcut_compute_cut.cl:135-139).
//
// TAG: the same flood for the tag post-process (graphcut_compute_tag.cl:51-140,
// :142-242): P.lab is the tag plane, a tagged node tags its axis neighbours
// (4 in 2D, 6 in 3D) that are on the source side of the cut (`object`).
// LDS codes: 0 open, 1 labelled on entry, 2 labelled here, 4 closed.
template <class S, bool TAG, bool ONE>
__global__ __launch_bounds__(256) void k_reach(Tiles T, Planes P, Queue q,
                                               unsigned char *dirty, int *cert,
                                               const unsigned char *object) {
  __shared__ unsigned char s_l[S::PADN];
  const int tid = threadIdx.x;
  const Lane<S> L(tid);
  __shared__ int s_next[3];
  __shared__ int s_w[8];
  int par = 0;
  VisitCursor Qc(q, s_next);
  constexpr unsigned kAxis = S::D3 ? 0x3fu : 0xfu;
  while (Qc.idx >= 0) {
    const int t = Qc.ent;
    const TileRef r = ONE ? tile_ref_of<S>(T.one, t) : tile_ref<S>(T, t);
    const InstDesc &I = r.I;
    const int K = TAG ? (S::D3 ? 6 : 4) : I.K;
    const int gx = r.ox + L.lx, gy = r.oy + L.ly, gz = r.oz + L.lz;
    const bool valid = gx < I.nx && gy < I.ny && gz < I.nz;
    // every load of the visit is issued before any is used (dead lanes and
    // halo cells read a valid dummy address)
    const long long gv = valid ? node_index(I, gx, gy, gz) : 0;
    // halo: the ring cells, split evenly over the four waves (as in k_push)
    HaloCell hc[kRingPerLane<S>];
    int hp[kRingPerLane<S>];
    unsigned char hl[kRingPerLane<S>], ho[kRingPerLane<S>];
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      const int j = c * 64 + (tid & 63);
      const int k = (tid >> 6) * kRingQ<S> + j;
      hp[c] = j < kRingQ<S> && k < Ring<S>::N ? Ring<S>::pos(k) : S::PADN;
      hc[c] = halo_cell<S>(I, r.ox, r.oy, r.oz, hp[c], K);
      hl[c] = P.lab[hc[c].g];
      if (TAG) ho[c] = object[hc[c].g];
    }
    unsigned char l0 = P.lab[gv];
    unsigned mask;
    bool deficit = false;
    if (TAG) {
      l0 = l0 ? 1 : object[gv] ? 0 : 4;
      mask = kAxis;
    } else {
      deficit = P.e[gv] < 0;
      mask = P.rmask[gv] & ((1u << K) - 1u);
    }
    if (!valid) {
      l0 = TAG ? 4 : 0;
      mask = 0;
      deficit = false;
    }
    if (!TAG && l0 && deficit) *cert = 1;
    s_l[L.me] = l0;
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      if (!hc[c].halo) continue;
      unsigned char lv = 1;
      if (hc[c].live) lv = TAG ? (hl[c] ? 1 : ho[c] ? 0 : 4) : hl[c];
      s_l[hp[c]] = lv;
    }
    __syncthreads();
    // the next entry loads while the tile iterates
    if (tid == 0) Qc.prefetch();
    int any;
    do {
      int ch = 0;
      const unsigned me = s_l[L.me];
      if (mask && (me == 1u || me == 2u)) {
        // neighbour labels read up front (one LDS wait); another lane may
        // label the same neighbour meanwhile, and both then write the same 2
        unsigned char lv[S::KMAX];
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d) lv[d] = s_l[L.me + lds_off<S>(d)];
#pragma unroll
        for (int d = 0; d < S::KMAX; ++d)
          if ((mask & (1u << d)) && lv[d] == 0) {
            s_l[L.me + lds_off<S>(d)] = 2;
            ch = 1;
          }
      }
      any = block_any(ch, s_w, par);
    } while (any);
    if (valid && s_l[L.me] == 2) {
      P.lab[gv] = 1;
      if (!TAG && deficit) *cert = 1;
    }
    // halo cells labelled by this visit (only live ring cells can be)
#pragma unroll
    for (int c = 0; c < kRingPerLane<S>; ++c) {
      if (!hc[c].live || s_l[hp[c]] != 2) continue;
      int x, y, z, nout;
      bool inter;
      pad_coords<S>(hp[c], x, y, z, inter, nout);
      const int X = r.ox + x, Y = r.oy + y, Z = r.oz + z;
      P.lab[hc[c].g] = 1;
      dirty[tile_of<S>(I, X, Y, Z)] = 1;
    }
    if (tid == 0) Qc.publish();
    __syncthreads();
    Qc.advance();
  }
}

template <bool TAG, bool ONE>
static void reach_launch(int shape, int nwg, const Tiles &T, const Planes &P, const Queue &q,
                         unsigned char *dirty, int *cert, const unsigned char *object,
                         hipStream_t st) {
  JFC_DISPATCH(shape, (k_reach<S, TAG, ONE><<<resident_grid<&k_reach<S, TAG, ONE>>(nwg), 256, 0,
                                              st>>>(T, P, q, dirty, cert, object)));
}

void launch_reach(int shape, int nwg, const Tiles &T, const Planes &P,
                  const Queue &q, unsigned char *dirty, int *cert,
                  hipStream_t st) {
  if (T.ninst == 1) reach_launch<false, true>(shape, nwg, T, P, q, dirty, cert, nullptr, st);
  else reach_launch<false, false>(shape, nwg, T, P, q, dirty, cert, nullptr, st);
}

void launch_tag(int shape, int nwg, const Tiles &T, const Planes &P,
                const Queue &q, unsigned char *dirty, const unsigned char *object,
                hipStream_t st) {
  if (T.ninst == 1) reach_launch<true, true>(shape, nwg, T, P, q, dirty, nullptr, object, st);
  else reach_launch<true, false>(shape, nwg, T, P, q, dirty, nullptr, object, st);
}

}  // namespace jfc
