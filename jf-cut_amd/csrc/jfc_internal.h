// jfc_internal.h -- structures shared by the host driver and the gfx950
// kernels of libjfcut.  Nothing here is part of the C-ABI (include/jfcut.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace jfc {

// Height of a node proven unable to reach any deficit node.
constexpr int HINF = 0x3fffffff;
// A height word holds the height in bits 0-29 and, in bits 30-31, the tag of
// the global relabel that wrote it (1 or 2, alternating; 0 from the push
// kernels and the zero start).  Every reader masks with HMASK.  A relabel's
// first pass takes a neighbour's word as a distance bound only when it
// carries the running relabel's tag; any other word is stale, and the
// neighbour's start value (0 at a deficit, HINF elsewhere) is used instead.
// The first pass rewrites every word with the tag, so a word left from the
// relabel before last can never carry the running tag.
constexpr int HMASK = 0x3fffffff;
__host__ __device__ inline int tag_word(int h, int tag) {
  return (int)((unsigned)h | ((unsigned)tag << 30));
}

// Per-instance reduction slots (int64): folded constant, initial deficit,
// energy, initial supply.  (The final deficit has an array of its own.)
constexpr int RED_CONST = 0, RED_DEF0 = 1, RED_ENERGY = 3, RED_SUP0 = 4,
              RED_SLOTS = 5;

// One grid instance inside a graph.  All instances share the plane arrays;
// instance i owns nodes [node_off, node_off + nx*ny*nz) and tiles
// [tile_off, tile_off + tx*ty*tz).
// Node v of an instance lives at node_off + x + px * (y + ny * z): rows are
// padded to px, a multiple of 32 nodes, so every 32-node tile row is one
// 128-byte line.  Pad nodes have no arcs and no terminal capacity.  Arrays
// exchanged with callers use the unpadded numbering x + nx * (y + ny * z)
// (user_off: the instance's first node in the graph's concatenation).
struct InstDesc {
  int nx, ny, nz;
  int conn, K;
  int tx, ty, tz;
  int tile_off;
  int ntiles;
  long long node_off;
  int px;
  long long user_off;
};

// The data term's second stream of a graph (null members: not created).
struct GraphAux {
  hipStream_t st;
  hipEvent_t fork, join, ready;
  void *pin;  // pinned host words
};

struct Planes {
  int *e;            // excess (> 0) / deficit (< 0), folded terminal arcs
  int *h;            // height word (HMASK bits); HINF = cannot reach a deficit
  int *cap;          // K residual planes, plane d at cap + d * stride
  const int *c0;     // input neighbour planes (same layout)
  const int *cs;     // input source capacities
  const int *ct;     // input sink capacities
  unsigned char *lab;// canonical labels (1 = source side)
  long long stride;  // nodes per plane (all instances)
  unsigned *rmask;   // bit d: residual of arc d > 0; bit 8+d: plane d still
                     // equals c0 (cap not yet written in this solve)
};

struct Tiles {
  const InstDesc *inst;
  const int *tile_inst;  // instance of every tile
  int ntiles;
  const int *row_off;    // first grid row (y, z) of every instance, + total
  int ninst;
  int nrows;             // rows of nx nodes over all instances
  int row_step;          // rows a workgroup streams at once in the whole-graph
                         // passes (several short rows of a one-instance graph)
  InstDesc one;          // the descriptor of instance 0 (kernel argument: a
                         // one-instance visit reads it with scalar loads)
};

// Tile shapes (256 nodes each, one node per lane).
enum ShapeId {
  SH_32x8 = 1,
  SH_64x4 = 2,
  SH_16x16 = 3,
  SH_8x8x4 = 4,
  SH_32x4x2 = 5
};
bool shape_is_3d(int shape);
void shape_dims(int shape, int *tw, int *th, int *td);

// Persistent work queues.  Slot s of ctr (QSLOT ints) holds the list count at
// [QSLOT*s] and claim head h at [QSLOT*s + QWORD*(1+h)], h < 8; every consumer
// zeroes the other slot's words, so producer/consumer launches alternate slots
// without any memset.  The words sit 256 bytes apart: device atomics on one
// cache line serialize, and with the eight heads in one line the claims of a
// launch queued behind each other (the 15000x10000 8-connected cut 5.64 ->
// 5.22 ms, the 2560x1440x50 video volume 16.9 -> 14.1 ms once they were
// split).
constexpr int QWORD = 64;
constexpr int QSLOT = 9 * QWORD;
struct Queue {
  int *list;
  int *ctr;
  int slot;
  int fixed_count;  // >= 0: consume list[0..fixed_count) instead of ctr count
  // Direct list (pick set): list[0..fixed_count) holds every tile of the
  // launch unfiltered, and the consumer does the compaction's work itself --
  // a tile is visited only if pick[t] is set (and pick[t] is cleared), its
  // entry bits come from mat / touched / hmat, and each visit adds 1 to
  // *cnt and *total; nothing is visited once *stop is nonzero.  For lists
  // no longer than the launch's grid, it saves the compaction launch.
  unsigned char *pick = nullptr;
  const unsigned char *mat = nullptr, *touched = nullptr, *hmat = nullptr;
  int *cnt = nullptr;
  long long *total = nullptr;
  const int *stop = nullptr;
};

// Sets the thread-local message returned by jfc_last_error().
void set_last_error(const std::string &msg);

// ---- launchers (all asynchronous on `st`) --------------------------------
void launch_fill_i32(int *p, long long n, int v, hipStream_t st);
void launch_fill_u8(unsigned char *p, long long n, unsigned char v, hipStream_t st);
// Up to 8 fills in one launch (bytes = 1 or 4 per element).
struct FillSeg {
  void *p;
  long long n;
  int v;
  int bytes;
};
struct FillSet {
  static constexpr int kMax = 12;
  FillSeg seg[kMax];
  int n = 0;
  void add(void *p, long long count, int v, int bytes) {
    if (count > 0 && n < kMax) seg[n++] = FillSeg{p, count, v, bytes};
  }
};
void launch_fill_multi(const FillSet &f, hipStream_t st);

// active (or null): tiles holding an excess node; defi: tiles holding a
// deficit node (the relabel's jump-flooding seeds).  Both zeroed here.
// zero_h: also write h = 0 over all nodes (the Fast-Cut kernels read every
// height; the push kernel instead treats the heights of tiles it has not
// visited in the solve as 0, see launch_push).
// masks: 0 = rmask untouched; 1 = rmask = mask0; 2 = build mask0 from the
// input planes and write it to rmask too (all in the same pass).
// cleared: the caller has zeroed active and defi already.
void launch_fold(int shape, const Tiles &T, const Planes &P, long long *red,
                 unsigned char *active, unsigned char *defi, bool zero_h, unsigned *mask0,
                 int masks, hipStream_t st, bool cleared = false);
void launch_validate(const int *p, long long n, int lo, int hi, int *bad,
                     hipStream_t st);
void launch_zero_border(int shape, const InstDesc &I, int *plane, int d,
                        hipStream_t st);
// fg / bg: rows of nx bytes (the caller's numbering); cs / ct: rows of px.
void launch_seeds(int *cs, int *ct, const unsigned char *fg, const unsigned char *bg, int nx,
                  int px, long long rows, hipStream_t st);
void launch_excess_split(const int *ef, int *cs, int *ct, long long n,
                         hipStream_t st);

// Also sums the remaining deficit max(-e, 0) per instance into def[inst].
// seed and def are zeroed here, unless the caller has (cleared).
void launch_cut_init(int shape, const Tiles &T, const Planes &P,
                     unsigned char *seed, long long *def, hipStream_t st, bool cleared = false);

// Jump-flooding tile distance + counting sort of tiles by distance.
// Lines of the tile grids along each axis: prefix[a][i] = lines along axis a
// in the instances before i, total[a] = all of them.
struct TileLines {
  const int *prefix[3];
  int total[3];
};
// hist_cleared: the caller has zeroed hist (nbins ints) already.
void launch_tile_order(int shape, const Tiles &T, const unsigned char *seed,
                       int *jf0, int *jf1, int *tdist, int *hist, int nbins,
                       int *order, int max_tile_dim, int *max_dist,
                       const TileLines &lines, int jump_flooding, hipStream_t st,
                       bool hist_cleared = false);

// Sweep bookkeeping on the device.  state (8 ints): [0] active tiles of the
// running sweep, [1] sweep index, [2] stop word (0 sweep on, 1 no active
// tile, 2 active set stalled, 3 sweep budget spent, 4 relabel budget
// spent), [3] sweep index of the last
// relabel, [4..5] tile visits (int64) at the last relabel, [6] sweeps ended
// (no-ops included) in the solve.
struct SweepCtl {
  int *hist;             // active tiles per sweep
  int *state;
  const long long *total;  // tile visits of the push kernel so far
  int cap;               // entries of hist
  int max_sweeps;
  double stall_ratio;
  long long budget;      // tile visits between relabels (0 = no budget)
  // Pinned host ring the end of sweep s posts s * 8 + stop into, at slot
  // s % ring_n (null: no posting).
  int *ring;
  int ring_n;
};
constexpr int kSweepState = 8;
void launch_sweep_end(const SweepCtl &c, hipStream_t st);
void launch_sweep_resume(const SweepCtl &c, hipStream_t st);

// Queue list entries: tile index, plus, when compacted with `mat`, kEntryMat
// if mat[t] is set, else kEntryClean if touched[t] is clear (no visit and no
// flush has written the tile's planes or mask words in this solve); when
// compacted with `hmat`, kEntryHeights if hmat[t] is set (the tile's heights
// are stored in this solve).
constexpr int kEntryMat = 1 << 30;
constexpr int kEntryClean = 1 << 29;
constexpr int kEntryHeights = 1 << 28;
constexpr int kEntryTile = kEntryHeights - 1;
// stop (or null): a nonzero word makes the compaction a no-op.
void launch_compact(const int *src, int n, int reverse, unsigned char *flag,
                    const unsigned char *mat, const unsigned char *touched, const Queue &q,
                    int *sweep_cnt, long long *total, const int *stop, hipStream_t st,
                    const unsigned char *hmat = nullptr);
// mat: per tile, 1 once every residual plane of the tile was written in
// this solve (cleared when a solve starts).  defi: a visit rewrites its
// tile's deficit flag (pushes never create deficits, so unvisited tiles'
// flags stay a superset).  hmat: per tile, 1 once its heights are stored in
// this solve (by a push visit, which then writes all of them, or by a global
// relabel); heights of a tile with hmat 0 read as 0, the labelling a solve
// starts from.  max_iters caps the local iterations of a visit, sparse_iters
// those of a launch whose list fits its grid once kSparseAfter sweeps ran.  iter_total: 64 int64 counters
// of local iterations (or null).
// touched: per tile, 1 once a flush wrote into the tile's planes or mask
// words (cleared when a solve starts).
// sweep_state: the solve's SweepCtl state ([1] = sweep index).
constexpr int kSparseAfter = 8;  // sweeps before the sparse cap applies
void launch_push(int shape, int nwg, const Tiles &T, const Planes &P,
                 const Queue &q, unsigned char *flag, unsigned char *mat,
                 unsigned char *touched, unsigned char *defi, unsigned char *hmat,
                 int max_iters, int sparse_iters, const int *sweep_state,
                 long long *iter_total, hipStream_t st);
// Fast-Cut comparator (jfc_fastcut.hip)
void launch_fc_materialize(const Planes &P, const unsigned *mask0, int K, hipStream_t st);
void launch_fc_sweep(int shape, int nwg, const Tiles &T, const Planes &P, const Queue &q,
                     unsigned char *flag, unsigned char *defi, hipStream_t st);
// Relabel start values written over all nodes with `tag`, and the exact
// deficit flags of the tiles (3D relabels).
void launch_gr_init(int shape, const Tiles &T, const Planes &P, int tag, unsigned char *defi,
                    hipStream_t st, bool cleared = false);
// One relabel pass.  The first pass (first_pass, every tile) starts each
// tile from e (0 at deficits, HINF elsewhere) and writes every height word
// with `tag`; later passes, and every pass after launch_gr_init, relax the
// queued tiles from their stored heights.  `active` holds the tiles' active
// flags; with `pinst` (3D shapes only: a relabel tiling other than the push tiling) it
// is indexed by the push tiling instead -- pinst its instance descriptors,
// pdims its tile dimensions as log2 (x | y << 4 | z << 8) -- and every
// visit sets the flags of the push tiles holding active nodes (the caller
// clears them before the relabel).
void launch_relax(int shape, int nwg, const Tiles &T, const Planes &P,
                  const Queue &q, unsigned char *dirty, unsigned char *active,
                  bool first_pass, int tag, hipStream_t st, const InstDesc *pinst = nullptr,
                  int pdims = 0);
void launch_reach(int shape, int nwg, const Tiles &T, const Planes &P,
                  const Queue &q, unsigned char *dirty, int *cert,
                  hipStream_t st);
// Tag post-process: init (tag plane holds foreground bytes on entry) and one
// flood pass over the queued tiles.
void launch_tag_init(int shape, const Tiles &T, unsigned char *tag,
                     unsigned char *seed, hipStream_t st);
void launch_tag(int shape, int nwg, const Tiles &T, const Planes &P,
                const Queue &q, unsigned char *dirty, const unsigned char *object,
                hipStream_t st);
// energy: one int64 per instance, zeroed by the caller.
void launch_energy(int shape, const Tiles &T, const Planes &P, long long *energy,
                   hipStream_t st);
void launch_backward(int shape, const Tiles &T, const Planes &P,
                     const InstDesc &I, float *dcs, float *dct, float *dnb,
                     hipStream_t st);
void launch_backward_all(int shape, const Tiles &T, const Planes &P, float *dcs,
                         float *dct, float *dnb, long long user_total, hipStream_t st);

}  // namespace jfc
