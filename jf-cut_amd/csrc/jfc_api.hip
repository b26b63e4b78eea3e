// jfc_api.hip -- host driver and C-ABI of libjfcut (include/jfcut.h).
//
// The solve driver replaces QGCWidget::slotCutClicked
// (src/QVisualizer/Graph Cut/QGCWidget.cpp:2364-2558).  Differences by design:
//   - the loop stays on the device: tile lists are rebuilt by ballot
//     compaction kernels and consumed by persistent workgroups, and the host
//     reads one counter every `check_every` sweeps instead of several
//     blocking readbacks per iteration;
//   - convergence is decided by the active-tile count, which is exact because
//     heights stay a valid distance labelling (global relabel is an exact
//     BFS, colour classes never touch each other's nodes); the canonical-cut
//     search then doubles as the certificate (QGCWidget.cpp:3237-3273);
//   - the max-flow value is an int64 reduction of the absorbed deficit
//     (reduce_flow in the reference accumulates int32, reduce.cl:17-42).
#include "jfcut.h"
#include "jfcut_debug.h"
#include "jfc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace jfc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

}  // namespace

// Host-only parts of the library (graph files) report through the same
// thread-local message.
void jfc::set_last_error(const std::string &msg) { g_err = msg; }

namespace {

#define JFC_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(JFC_ERR_DEVICE, std::string(#call) + ": " +               \
                                      hipGetErrorString(e_));                \
  } while (0)

template <class T>
int dalloc(T **p, size_t n, const char *what) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(JFC_ERR_OOM, std::string("device allocation failed: ") + what +
                                 " (" + std::to_string(n * sizeof(T)) + " bytes)");
  }
  return JFC_OK;
}

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

int default_shape(int dim) { return dim == 3 ? SH_8x8x4 : SH_32x8; }

int shape_from_option(int opt, int dim) {
  switch (opt) {
    case JFC_TILE_32x8: return dim == 2 ? SH_32x8 : 0;
    case JFC_TILE_64x4: return dim == 2 ? SH_64x4 : 0;
    case JFC_TILE_16x16: return dim == 2 ? SH_16x16 : 0;
    case JFC_TILE_8x8x4: return dim == 3 ? SH_8x8x4 : 0;
    case JFC_TILE_32x4x2: return dim == 3 ? SH_32x4x2 : 0;
    default: return default_shape(dim);
  }
}

// Process-wide autotuning table: size class -> JFC_TILE_* shape.
// Autotuned choices per size class: tile shape, local-iteration cap and the
// relabel before the first sweep.
struct Tuned {
  int shape = 0;
  int mli = 0;
  int ir = -1;  // initial relabel (-1: not tuned)
  int sc = -1;  // certify first on a stall (-1: not tuned)
  int sp = 0;   // cap of sparse launches (0: not tuned)
};

// Default cap on the local iterations of a launch whose list fits its grid,
// from the solve's kSparseAfter-th sweep on (the reference fixtures: 4.6 ->
// 2.2-3.0 ms against a cap of 2; from the first sweep on, larger caps slowed
// graphs whose few active tiles iterate without effect, the 512x512
// 4-connected cut 0.48 -> 0.71 ms at 64).
constexpr int kSparseIters = 32;
std::mutex g_tune_mu;
std::map<std::string, Tuned> g_tune;

int log2_round(double v) { return v <= 1.0 ? 0 : (int)std::lround(std::log2(v)); }

// Device counters (int64): [0] push tile visits, [1] relabel tile visits,
// [2] cut tile visits, [4] tag tile visits, [kIterSlots, +64) local
// iterations of the push kernel (workgroup b adds to slot b % 64, so no
// single address takes one atomic per workgroup).
constexpr int kIterSlots = 5, kTotals = kIterSlots + 64;
// Graphs of at most this many tiles: a search pass gives every tile a
// workgroup of its own or nearly (a device of 256 compute units keeps
// several tile visits per unit resident), so its passes skip the distance
// ordering, and its launches take their tiles from the flags (no compaction
// launches).  The reference's fixtures (1083-1200 tiles of 32x8 or 16x16)
// solve 5-15% faster at 2048 than at 1024.
constexpr int kSmallTiles = 2048;
// words of the results block besides the per-instance ones: the counters,
// the sweep state (kSweepState ints) and 4 flag words
constexpr size_t kResFixed = kTotals + (kSweepState + 4) / 2;

// Sweeps whose stop words can be in flight at once (pinned words and events).
constexpr int kStopRing = 8;

// algorithmic HBM bytes per node of one push-relabel visit: read and write
// e, h and K residual planes (int32)
long long push_bytes_per_node(int K) { return 2LL * 4 * (2 + K); }

}  // namespace

struct jfc_graph {
  int dim = 2;
  int tuned = 0;           // JFC_TILE_* chosen by jfc_autotune (0 = none)
  int tuned_mli = 0;       // local-iteration cap chosen by jfc_autotune (0 = none)
  int tuned_ir = -1;       // initial relabel chosen by jfc_autotune (-1 = none)
  int tuned_sc = -1;       // stall_certify chosen by jfc_autotune (-1 = none)
  int tuned_sp = 0;        // sparse_local_iters chosen by jfc_autotune (0 = none)
  std::string size_class;  // key of the process-wide autotuning table
  int Kg = 4;  // planes stored (max K over instances)
  int shape = 0;
  long long Ntot = 0;   // nodes stored, rows padded (plane stride)
  long long Nuser = 0;  // nodes of the instances as callers number them
  void *stage = nullptr;  // device staging of caller arrays (rows of nx)
  size_t stage_bytes = 0;
  void *model_scratch = nullptr;  // data-term scratch (histograms, centres, sums)
  size_t model_scratch_bytes = 0;
  std::vector<InstDesc> inst;
  std::vector<int> caps_set;  // per instance: bit 0 terminals, bit 1 + d neighbour planes
  std::vector<int> caps_failed;  // per instance: the same bits, for a setter that failed
                                 // (its planes are zeroed; solves refuse the instance)
  // tile tables for the current shape
  int ntiles = 0, ncolors = 2, max_tile_dim = 1, nbins = 2;
  std::vector<int> color_off;
  // device planes
  int *cs = nullptr, *ct = nullptr, *c0 = nullptr, *cap = nullptr;
  int *e = nullptr, *h = nullptr;
  unsigned char *lab = nullptr;
  unsigned char *tag = nullptr;  // tag post-process output (allocated on first use)
  // device tables / work space
  InstDesc *d_inst = nullptr;
  int *d_tile_inst = nullptr, *d_color_tiles = nullptr;
  int *d_list[2] = {nullptr, nullptr};
  int *d_ctr = nullptr;
  unsigned char *d_flag = nullptr, *d_dirty = nullptr, *d_seed = nullptr;
  unsigned char *d_dirty2 = nullptr;  // second dirty buffer of small-graph searches
  unsigned char *d_defi = nullptr;  // tiles that may hold a deficit node
  int gr_tag = 1;                   // tag of the last global relabel's heights
  unsigned char *d_mat = nullptr;  // per tile: residual planes written this solve
  unsigned char *d_hmat = nullptr; // per tile: heights stored this solve
  unsigned char *d_touched = nullptr;  // per tile: planes or masks written by a flush
  int *d_jf0 = nullptr, *d_jf1 = nullptr, *d_tdist = nullptr, *d_hist = nullptr,
      *d_order = nullptr;
  // The tiling of the 3D global relabel when it is not the push tiling:
  // 32x4x2 push tiles read whole 128-byte rows (an 8x8x4 tile's rows are 32
  // bytes), but a relabel front crosses about one tile per pass, and 32x4x2
  // tiles are half as deep in y and z, so their relabels run on 8x8x4 tiles
  // (the 2560x1440x50 video volume: push 1.48 ms on 32x4x2 tiles against
  // 2.81 on 8x8x4, relabel 2.95 ms on 8x8x4 against 4.56 on 32x4x2).  shape 0: the relabel runs on
  // the push tiling.  The relabel writes the push tiling's active flags.
  struct RelabelTiles {
    int shape = 0;
    std::vector<InstDesc> inst;
    InstDesc *d_inst = nullptr;
    int *d_tile_inst = nullptr;
    int ntiles = 0, max_tile_dim = 1, nbins = 2;
    unsigned char *d_dirty = nullptr, *d_defi = nullptr;
    int *d_jf0 = nullptr, *d_jf1 = nullptr, *d_tdist = nullptr, *d_hist = nullptr,
        *d_order = nullptr, *d_lines = nullptr;
    int lines_total[3] = {0, 0, 0};
  } rt;
  // One device block holds every word a solve reads back at its end (one
  // copy): the per-instance sums, the remaining deficits, the counters, the
  // sweep state and the flag words, in this order (res_words long longs).
  long long *d_res = nullptr;
  size_t res_words = 0;
  long long *d_red = nullptr, *d_total = nullptr;  // in d_res
  long long *d_def = nullptr;  // in d_res; per instance: remaining deficit after the cut
  long long *d_energy = nullptr;  // per instance: cut energy (jfc_get_energy)
  int *d_row_off = nullptr;    // per instance: first row; [count] = rows
  int *d_lines = nullptr;      // per axis: prefix count of tile lines per instance
  int lines_total[3] = {0, 0, 0};
  int nrows = 0;
  int *d_sweep = nullptr;
  int sweep_cap = 0;
  int last_sweeps = 0;
  int *d_flagword = nullptr;  // in d_res; [0] certificate, [1] validation
  unsigned *mask0 = nullptr;  // residual-arc bits of the input planes
  unsigned *rmask = nullptr;  // residual-arc bits of the current residual
  bool mask_dirty = true;     // inputs changed since the last solve started
  bool mask0_ok = false;      // mask0 holds the bits of the current inputs
  int slot = 0;
  bool solved = false;
  std::vector<long long> flows;
  std::vector<hipEvent_t> ev_pool;
  int nwg = 2048;
  int *d_sweep_state = nullptr;     // in d_res; {active tiles of the running sweep, sweep index}
  hipStream_t priv_stream = nullptr;  // solve stream when the caller passes NULL
  hipEvent_t priv_event = nullptr;
  // Second stream of the data term (its pair sums and pairwise weights run
  // beside the k-means and the unary terms), with the fork and join events.
  hipStream_t aux_stream = nullptr;
  hipEvent_t aux_fork = nullptr, aux_join = nullptr, aux_ready = nullptr;
  void *aux_pin = nullptr;  // pinned host words the second stream's copies use
  // The caller's stream for every call of the graph (jfc_set_stream): the
  // setters, the data term, the outputs and the backward issue their work
  // on it and wait for that stream alone.  NULL: the legacy default stream.
  hipStream_t ustream = nullptr;
  // Pinned host words the solve's decisions read (asynchronous copies, then
  // a wait on an event instead of a staged pageable copy).
  struct HostPin {
    long long total;  // tile visits so far (sweep checks)
    int cnt[2];       // last list count of two pass groups (searches)
    int recent[3];    // active tiles of the latest sweeps
    int flag;         // certificate word of a cut search
    int bad;          // validation word of a setter
    long long red[RED_SLOTS];  // one instance's sums (energy)
  } *pin = nullptr;
  // Pinned copy of d_res (the solve's results).
  long long *pin_res = nullptr;
  // Stop words the sweeps post (k_sweep_end), in fine-grained pinned memory
  // the device writes through to; d_ring is its device address.
  int *ring = nullptr, *d_ring = nullptr;
  hipEvent_t cnt_ev[2] = {nullptr, nullptr};
  hipEvent_t stop_ev[kStopRing] = {};
  int ensure_pin() {
    if (pin) return JFC_OK;
    for (hipEvent_t &ev : cnt_ev)
      if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
        return fail(JFC_ERR_DEVICE, "event creation failed");
      }
    for (hipEvent_t &ev : stop_ev)
      if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
        return fail(JFC_ERR_DEVICE, "event creation failed");
      }
    if (hipHostMalloc((void **)&pin, sizeof(HostPin), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&pin_res, res_words * sizeof(long long), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void **)&ring, kStopRing * sizeof(int), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void **)&d_ring, ring, 0) != hipSuccess) {
      (void)hipGetLastError();
      if (pin) (void)hipHostFree(pin);
      if (pin_res) (void)hipHostFree(pin_res);
      if (ring) (void)hipHostFree(ring);
      pin = nullptr;
      pin_res = nullptr;
      ring = d_ring = nullptr;
      return fail(JFC_ERR_OOM, "pinned host words");
    }
    return JFC_OK;
  }
  // Captured launch sequence of one sweep, per starting queue slot.
  struct SweepGraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t st = nullptr;
    int mli = 0;
    int sparse = 0;
    int max_sweeps = 0;
    double stall = 0.0, work = 0.0;
    const int *hist = nullptr;
  } sweep_graph[2];

  void drop_graphs() {
    for (auto &sg : sweep_graph) {
      if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
      sg = SweepGraph{};
    }
  }

  ~jfc_graph() { release(); }

  void release_tiles() {
    drop_graphs();
    dfree(d_tile_inst); dfree(d_color_tiles); dfree(d_list[0]); dfree(d_list[1]);
    dfree(d_flag); dfree(d_dirty); dfree(d_dirty2); dfree(d_seed); dfree(d_defi); dfree(d_jf0);
    dfree(d_jf1);
    dfree(d_tdist); dfree(d_hist); dfree(d_order); dfree(d_inst); dfree(d_mat);
    dfree(d_hmat);
    dfree(d_touched);
    dfree(d_lines);
    dfree(rt.d_inst); dfree(rt.d_tile_inst); dfree(rt.d_dirty); dfree(rt.d_defi);
    dfree(rt.d_jf0); dfree(rt.d_jf1); dfree(rt.d_tdist); dfree(rt.d_hist); dfree(rt.d_order);
    dfree(rt.d_lines);
    rt = RelabelTiles{};
  }
  void release_aux() {
    if (aux_fork) (void)hipEventDestroy(aux_fork);
    if (aux_join) (void)hipEventDestroy(aux_join);
    if (aux_ready) (void)hipEventDestroy(aux_ready);
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
    if (aux_pin) (void)hipHostFree(aux_pin);
    aux_fork = aux_join = aux_ready = nullptr;
    aux_stream = nullptr;
    aux_pin = nullptr;
  }
  void release() {
    release_tiles();
    dfree(cs); dfree(ct); dfree(c0); dfree(cap); dfree(e); dfree(h); dfree(lab); dfree(tag);
    dfree(d_ctr); dfree(d_res); dfree(d_sweep); dfree(d_energy); dfree(d_row_off);
    d_red = d_total = d_def = nullptr;
    d_sweep_state = d_flagword = nullptr;
    if (stage) (void)hipFree(stage);
    stage = nullptr;
    stage_bytes = 0;
    if (model_scratch) (void)hipFree(model_scratch);
    model_scratch = nullptr;
    model_scratch_bytes = 0;
    if (priv_event) (void)hipEventDestroy(priv_event);
    if (priv_stream) (void)hipStreamDestroy(priv_stream);
    priv_event = nullptr;
    priv_stream = nullptr;
    release_aux();
    for (hipEvent_t &ev : cnt_ev) {
      if (ev) (void)hipEventDestroy(ev);
      ev = nullptr;
    }
    for (hipEvent_t &ev : stop_ev) {
      if (ev) (void)hipEventDestroy(ev);
      ev = nullptr;
    }
    if (pin) (void)hipHostFree(pin);
    if (pin_res) (void)hipHostFree(pin_res);
    if (ring) (void)hipHostFree(ring);
    pin = nullptr;
    pin_res = nullptr;
    ring = d_ring = nullptr;
    dfree(mask0); dfree(rmask);
    for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
    ev_pool.clear();
  }

  Tiles tiles() const {
    // the whole-graph passes stream 4-node groups, 1024 per workgroup step:
    // a one-instance graph with short rows takes several rows per step
    int step = 1;
    if (inst.size() == 1) step = std::max(1, 1024 / ((inst[0].nx + 3) / 4));
    return Tiles{d_inst, d_tile_inst, ntiles, d_row_off, (int)inst.size(), nrows, step,
                 inst.empty() ? InstDesc{} : inst[0]};
  }
  Planes planes() const { return Planes{e, h, cap, c0, cs, ct, lab, Ntot, rmask}; }
  TileLines lines() const {
    const int n1 = (int)inst.size() + 1;
    return TileLines{{d_lines, d_lines + n1, d_lines + 2 * n1},
                     {lines_total[0], lines_total[1], lines_total[2]}};
  }
  long long inst_nodes(int i) const {
    return (long long)inst[i].nx * inst[i].ny * inst[i].nz;
  }
  long long inst_stored(int i) const {
    return (long long)inst[i].px * inst[i].ny * inst[i].nz;
  }

  // Build tile tables, colour classes and instance descriptors for `sh`.
  int set_shape(int sh) {
    if (sh == shape && d_tile_inst) return JFC_OK;
    release_tiles();
    shape = sh;
    int tw, th, td;
    shape_dims(sh, &tw, &th, &td);
    bool all4 = true;
    long long off = 0;
    max_tile_dim = 1;
    int maxsum = 0;
    for (auto &I : inst) {
      I.tx = (I.nx + tw - 1) / tw;
      I.ty = (I.ny + th - 1) / th;
      I.tz = (I.nz + td - 1) / td;
      const long long nt = (long long)I.tx * I.ty * I.tz;
      if (off + nt >= kEntryTile) return fail(JFC_ERR_INVALID, "graph too large for the tile queue");
      I.tile_off = (int)off;
      I.ntiles = (int)nt;
      off += nt;
      if (I.K != 4) all4 = false;
      max_tile_dim = std::max(max_tile_dim, std::max(I.tx, std::max(I.ty, I.tz)));
      maxsum = std::max(maxsum, I.tx + I.ty + I.tz);
    }
    ntiles = (int)off;
    // distance bins of the tile ordering (the last bin holds tiles of
    // instances without seed tiles); farther tiles share the last finite bin
    nbins = std::min(maxsum + 2, 4096);
    // colour classes: tiles of one class are never adjacent through an arc,
    // so no node is in one tile of the class and in the halo of another it
    // pushes across.  2D 4C and 3D 6C: checkerboard; 2D 8C: 2x2 pattern.
    ncolors = (dim == 2 && !all4) ? 4 : 2;
    std::vector<int> tile_inst(ntiles);
    std::vector<std::vector<int>> cls(ncolors);
    for (size_t i = 0; i < inst.size(); ++i) {
      const InstDesc &I = inst[i];
      for (int z = 0; z < I.tz; ++z)
        for (int y = 0; y < I.ty; ++y)
          for (int x = 0; x < I.tx; ++x) {
            const int t = I.tile_off + x + I.tx * (y + I.ty * z);
            tile_inst[t] = (int)i;
            const int c = ncolors == 4 ? ((x & 1) | ((y & 1) << 1)) : ((x + y + z) & 1);
            cls[c].push_back(t);
          }
    }
    // tile lines along x, y, z per instance (separable distance transform)
    const int n1 = (int)inst.size() + 1;
    std::vector<int> lines(3 * n1, 0);
    for (size_t i = 0; i < inst.size(); ++i) {
      const InstDesc &I = inst[i];
      const long long per[3] = {(long long)I.ty * I.tz, (long long)I.tx * I.tz,
                                (long long)I.tx * I.ty};
      for (int a = 0; a < 3; ++a) {
        const long long nx = lines[a * n1 + i] + per[a];
        if (nx >= (1LL << 31) / 4) return fail(JFC_ERR_INVALID, "too many tile lines");
        lines[a * n1 + i + 1] = (int)nx;
      }
    }
    for (int a = 0; a < 3; ++a) lines_total[a] = lines[a * n1 + n1 - 1];
    std::vector<int> flat;
    color_off.assign(ncolors + 1, 0);
    for (int c = 0; c < ncolors; ++c) {
      color_off[c] = (int)flat.size();
      flat.insert(flat.end(), cls[c].begin(), cls[c].end());
    }
    color_off[ncolors] = (int)flat.size();
    int rc;
    if ((rc = dalloc(&d_inst, inst.size(), "instances")) ||
        (rc = dalloc(&d_tile_inst, ntiles, "tile table")) ||
        (rc = dalloc(&d_color_tiles, ntiles, "colour classes")) ||
        (rc = dalloc(&d_list[0], list_cap(sh), "tile list")) ||
        (rc = dalloc(&d_list[1], list_cap(sh), "tile list")) ||
        (rc = dalloc(&d_flag, ntiles, "tile flags")) ||
        (rc = dalloc(&d_dirty, ntiles, "tile flags")) ||
        (rc = dalloc(&d_dirty2, ntiles, "tile flags")) ||
        (rc = dalloc(&d_seed, ntiles, "tile seeds")) ||
        (rc = dalloc(&d_defi, ntiles, "tile seeds")) ||
        (rc = dalloc(&d_jf0, ntiles, "jump flooding")) ||
        (rc = dalloc(&d_jf1, ntiles, "jump flooding")) ||
        (rc = dalloc(&d_tdist, ntiles, "tile distance")) ||
        (rc = dalloc(&d_hist, nbins, "tile histogram")) ||
        (rc = dalloc(&d_order, ntiles, "tile order")) ||
        (rc = dalloc(&d_mat, ntiles, "tile states")) ||
        (rc = dalloc(&d_hmat, ntiles, "tile states")) ||
        (rc = dalloc(&d_touched, ntiles, "tile states")) ||
        (rc = dalloc(&d_lines, 3 * n1, "tile lines")))
      return rc;
    JFC_HIP(hipMemcpy(d_lines, lines.data(), lines.size() * sizeof(int), hipMemcpyHostToDevice));
    JFC_HIP(hipMemcpy(d_inst, inst.data(), inst.size() * sizeof(InstDesc), hipMemcpyHostToDevice));
    JFC_HIP(hipMemcpy(d_tile_inst, tile_inst.data(), ntiles * sizeof(int), hipMemcpyHostToDevice));
    JFC_HIP(hipMemcpy(d_color_tiles, flat.data(), ntiles * sizeof(int), hipMemcpyHostToDevice));
    return sh == SH_32x4x2 ? build_relabel_tiles(SH_8x8x4) : JFC_OK;
  }

  // Tile count of shape sh over the instances.
  long long count_tiles(int sh) const {
    int tw, th, td;
    shape_dims(sh, &tw, &th, &td);
    long long n = 0;
    for (const auto &I : inst)
      n += (long long)((I.nx + tw - 1) / tw) * ((I.ny + th - 1) / th) * ((I.nz + td - 1) / td);
    return n;
  }
  // Capacity of the compaction lists: the tiles of the push tiling, or of
  // the relabel tiling where that is another.
  long long list_cap(int sh) const {
    return sh == SH_32x4x2 ? std::max(count_tiles(sh), count_tiles(SH_8x8x4)) : count_tiles(sh);
  }

  // The relabel tiling's tables (rt) for shape sh2 over the same instances:
  // descriptors, tile -> instance table, tile lines and the search buffers.
  int build_relabel_tiles(int sh2) {
    rt.shape = sh2;
    int tw, th, td;
    shape_dims(sh2, &tw, &th, &td);
    rt.inst = inst;
    long long off = 0;
    int maxsum = 0;
    rt.max_tile_dim = 1;
    for (auto &I : rt.inst) {
      I.tx = (I.nx + tw - 1) / tw;
      I.ty = (I.ny + th - 1) / th;
      I.tz = (I.nz + td - 1) / td;
      const long long nt = (long long)I.tx * I.ty * I.tz;
      if (off + nt >= kEntryTile) return fail(JFC_ERR_INVALID, "graph too large for the tile queue");
      I.tile_off = (int)off;
      I.ntiles = (int)nt;
      off += nt;
      rt.max_tile_dim = std::max(rt.max_tile_dim, std::max(I.tx, std::max(I.ty, I.tz)));
      maxsum = std::max(maxsum, I.tx + I.ty + I.tz);
    }
    rt.ntiles = (int)off;
    rt.nbins = std::min(maxsum + 2, 4096);
    std::vector<int> tile_inst(rt.ntiles);
    const int n1 = (int)rt.inst.size() + 1;
    std::vector<int> lines(3 * n1, 0);
    for (size_t i = 0; i < rt.inst.size(); ++i) {
      const InstDesc &I = rt.inst[i];
      std::fill(tile_inst.begin() + I.tile_off, tile_inst.begin() + I.tile_off + I.ntiles, (int)i);
      const long long per[3] = {(long long)I.ty * I.tz, (long long)I.tx * I.tz,
                                (long long)I.tx * I.ty};
      for (int a = 0; a < 3; ++a) {
        const long long nx = lines[a * n1 + i] + per[a];
        if (nx >= (1LL << 31) / 4) return fail(JFC_ERR_INVALID, "too many tile lines");
        lines[a * n1 + i + 1] = (int)nx;
      }
    }
    for (int a = 0; a < 3; ++a) rt.lines_total[a] = lines[a * n1 + n1 - 1];
    int rc;
    if ((rc = dalloc(&rt.d_inst, rt.inst.size(), "instances")) ||
        (rc = dalloc(&rt.d_tile_inst, rt.ntiles, "tile table")) ||
        (rc = dalloc(&rt.d_dirty, rt.ntiles, "tile flags")) ||
        (rc = dalloc(&rt.d_defi, rt.ntiles, "tile seeds")) ||
        (rc = dalloc(&rt.d_jf0, rt.ntiles, "jump flooding")) ||
        (rc = dalloc(&rt.d_jf1, rt.ntiles, "jump flooding")) ||
        (rc = dalloc(&rt.d_tdist, rt.ntiles, "tile distance")) ||
        (rc = dalloc(&rt.d_hist, rt.nbins, "tile histogram")) ||
        (rc = dalloc(&rt.d_order, rt.ntiles, "tile order")) ||
        (rc = dalloc(&rt.d_lines, 3 * n1, "tile lines")))
      return rc;
    JFC_HIP(hipMemcpy(rt.d_lines, lines.data(), lines.size() * sizeof(int), hipMemcpyHostToDevice));
    JFC_HIP(hipMemcpy(rt.d_inst, rt.inst.data(), rt.inst.size() * sizeof(InstDesc),
                      hipMemcpyHostToDevice));
    JFC_HIP(hipMemcpy(rt.d_tile_inst, tile_inst.data(), rt.ntiles * sizeof(int),
                      hipMemcpyHostToDevice));
    return JFC_OK;
  }

  // The tiling a search runs on (the relabel tiling for relabels, else the
  // push tiling).
  struct SearchTiling {
    int shape;
    Tiles T;
    int ntiles, nbins, max_tile_dim;
    unsigned char *dirty, *defi;
    int *jf0, *jf1, *tdist, *hist, *order;
    TileLines lines;
    const InstDesc *pinst;  // push tiling descriptors when the tiling is another
    int pdims;              // push tile dimensions, log2 (x | y << 4 | z << 8)
  };
  SearchTiling search_tiling(bool relabel) const {
    if (!relabel || !rt.shape)
      return SearchTiling{shape, tiles(), ntiles, nbins, max_tile_dim, d_dirty, d_defi, d_jf0,
                          d_jf1, d_tdist, d_hist, d_order, lines(), nullptr, 0};
    Tiles T = tiles();
    T.inst = rt.d_inst;
    T.tile_inst = rt.d_tile_inst;
    T.ntiles = rt.ntiles;
    T.one = rt.inst.empty() ? InstDesc{} : rt.inst[0];
    const int n1 = (int)rt.inst.size() + 1;
    const TileLines L{{rt.d_lines, rt.d_lines + n1, rt.d_lines + 2 * n1},
                      {rt.lines_total[0], rt.lines_total[1], rt.lines_total[2]}};
    int tw, th, td;
    shape_dims(shape, &tw, &th, &td);
    const int pdims = __builtin_ctz(tw) | __builtin_ctz(th) << 4 | __builtin_ctz(td) << 8;
    return SearchTiling{rt.shape, T, rt.ntiles, rt.nbins, rt.max_tile_dim, rt.d_dirty, rt.d_defi,
                        rt.d_jf0, rt.d_jf1, rt.d_tdist, rt.d_hist, rt.d_order, L, d_inst, pdims};
  }
};

// ---- accessors for the data-term file ---------------------------------------------

namespace jfc {
int model_fail(int code, const std::string &msg) { return fail(code, msg); }

int graph_instance(jfc_graph *g, int inst, InstDesc *I, int **cs, int **ct, int **c0,
                   long long *stride) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  if (inst < 0 || inst >= (int)g->inst.size())
    return fail(JFC_ERR_INVALID, "instance index out of range");
  *I = g->inst[inst];
  *cs = g->cs;
  *ct = g->ct;
  *c0 = g->c0;
  *stride = g->Ntot;
  return JFC_OK;
}

void graph_mark_inputs(jfc_graph *g, int inst) {
  g->caps_set[inst] = ~0;
  g->caps_failed[inst] = 0;
  g->mask_dirty = true;
  g->solved = false;
}

// A data-term call that failed after writing some of the instance's planes.
void graph_inputs_failed(jfc_graph *g, int inst) {
  g->caps_set[inst] = 0;
  g->caps_failed[inst] = ~0;
  g->mask_dirty = true;
  g->solved = false;
}

hipStream_t graph_stream(jfc_graph *g) { return g->ustream; }

int graph_stage(jfc_graph *g, size_t bytes, void **p);

// The graph's scratch block for the data-term paths, allocated on first use
// and grown when a call needs more.
int graph_scratch(jfc_graph *g, size_t bytes, void **p) {
  if (g->model_scratch_bytes < bytes) {
    if (g->model_scratch) (void)hipFree(g->model_scratch);
    g->model_scratch = nullptr;
    g->model_scratch_bytes = 0;
    if (hipMalloc(&g->model_scratch, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(JFC_ERR_OOM, "out of device memory: data-term scratch");
    }
    g->model_scratch_bytes = bytes;
  }
  *p = g->model_scratch;
  return JFC_OK;
}

// The data term's second stream, its fork / join / ready events and its
// pinned host words (`pin_bytes`): created on first use when `create` is
// set; otherwise null until then (the caller then runs on one stream).
int graph_aux(jfc_graph *g, bool create, size_t pin_bytes, GraphAux *a) {
  if (!g->aux_stream && create) {
    if (hipStreamCreateWithFlags(&g->aux_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g->aux_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->aux_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->aux_ready, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&g->aux_pin, pin_bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      g->release_aux();
      return fail(JFC_ERR_DEVICE, "cannot create the data term's second stream");
    }
  }
  a->st = g->aux_stream;
  a->fork = g->aux_fork;
  a->join = g->aux_join;
  a->ready = g->aux_ready;
  a->pin = g->aux_pin;
  return JFC_OK;
}

// The instance's slice of the excess plane (stored nodes), free between
// solves: every solve rewrites it from the inputs before reading it.
int graph_work_plane(jfc_graph *g, int inst, int **plane, long long *n) {
  *plane = g->e + g->inst[inst].node_off;
  *n = g->inst_stored(inst);
  return JFC_OK;
}
}  // namespace jfc

// ---- lifetime -------------------------------------------------------------------

extern "C" const char *jfc_last_error(void) { return g_err.c_str(); }

extern "C" void jfc_default_options(jfc_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_sweeps = 1 << 20;
  o->check_every = 2;
  o->relabel_work = 2.0;
  o->max_local_iters = 0;
  o->tile_shape = JFC_TILE_DEFAULT;
  o->profile = 0;
  o->stream = nullptr;
  o->initial_relabel = -1;
  o->stall_ratio = 0.99;
  o->graph_sweeps = 1;
  o->tile_order = 0;
  o->method = JFC_METHOD_JFCUT;
  o->sparse_local_iters = 0;
  o->stall_certify = -1;
}

static int create_impl(jfc_graph **out, int count, const int32_t *shapes) {
  if (!out) return fail(JFC_ERR_INVALID, "null graph pointer");
  *out = nullptr;
  if (count <= 0 || !shapes) return fail(JFC_ERR_INVALID, "empty batch");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(JFC_ERR_DEVICE, "no HIP device available");
  }
  jfc_graph *g = new jfc_graph();
  long long off = 0, uoff = 0;
  g->dim = shapes[2] > 1 ? 3 : 2;
  g->Kg = 0;
  for (int i = 0; i < count; ++i) {
    const int nx = shapes[4 * i], ny = shapes[4 * i + 1], nz = shapes[4 * i + 2],
              conn = shapes[4 * i + 3];
    const int dim = nz > 1 ? 3 : 2;
    bool ok = nx > 0 && ny > 0 && nz > 0 && dim == g->dim &&
              ((dim == 2 && (conn == 4 || conn == 8)) || (dim == 3 && conn == 6));
    if (dim == 3 && conn != 6) ok = false;
    if (!ok) {
      delete g;
      return fail(JFC_ERR_INVALID, "instance " + std::to_string(i) +
                                       ": unsupported shape/connectivity (2D: 4 or 8, 3D: 6; "
                                       "all instances of one graph share the dimensionality)");
    }
    InstDesc I{};
    I.nx = nx; I.ny = ny; I.nz = nz; I.conn = conn; I.K = conn;
    I.px = (int)(((long long)nx + 31) / 32 * 32);
    I.node_off = off;
    I.user_off = uoff;
    off += (long long)I.px * ny * nz;
    uoff += (long long)nx * ny * nz;
    g->inst.push_back(I);
    g->Kg = std::max(g->Kg, conn);
  }
  if (off >= (1LL << 40)) { delete g; return fail(JFC_ERR_INVALID, "graph too large"); }
  g->Ntot = off;
  g->Nuser = uoff;
  g->caps_set.assign(count, 0);
  g->caps_failed.assign(count, 0);
  g->flows.assign(count, 0);
  int rc;
  const size_t N = (size_t)off;
  if ((rc = dalloc(&g->cs, N, "source capacities")) ||
      (rc = dalloc(&g->ct, N, "sink capacities")) ||
      (rc = dalloc(&g->c0, N * g->Kg, "neighbour capacities")) ||
      (rc = dalloc(&g->cap, N * g->Kg, "residual capacities")) ||
      (rc = dalloc(&g->e, N, "excess")) || (rc = dalloc(&g->h, N, "heights")) ||
      (rc = dalloc(&g->lab, N, "labels")) || (rc = dalloc(&g->d_ctr, 2 * QSLOT, "queues")) ||
      (rc = dalloc(&g->mask0, N, "residual masks")) ||
      (rc = dalloc(&g->rmask, N, "residual masks")) ||
      (rc = dalloc(&g->d_res, kResFixed + (size_t)count * (RED_SLOTS + 1), "results")) ||
      (rc = dalloc(&g->d_energy, (size_t)count, "reductions")) ||
      (rc = dalloc(&g->d_row_off, (size_t)count + 1, "row table"))) {
    delete g;
    return rc;
  }
  g->res_words = kResFixed + (size_t)count * (RED_SLOTS + 1);
  g->d_red = g->d_res;
  g->d_def = g->d_red + (size_t)count * RED_SLOTS;
  g->d_total = g->d_def + count;
  g->d_sweep_state = (int *)(g->d_total + kTotals);
  g->d_flagword = g->d_sweep_state + kSweepState;
  {
    std::vector<int> row_off(count + 1, 0);
    long long rows = 0;
    for (int i = 0; i < count; ++i) {
      row_off[i] = (int)rows;
      rows += (long long)g->inst[i].ny * g->inst[i].nz;
    }
    if (rows >= (1LL << 31)) { delete g; return fail(JFC_ERR_INVALID, "graph too large"); }
    row_off[count] = (int)rows;
    g->nrows = (int)rows;
    if (hipMemcpy(g->d_row_off, row_off.data(), row_off.size() * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess) {
      delete g;
      return fail(JFC_ERR_DEVICE, "row table upload failed");
    }
  }
  // everything zero, so the pad nodes of the rows stay isolated and empty
  launch_fill_i32(g->cs, off, 0, 0);
  launch_fill_i32(g->ct, off, 0, 0);
  launch_fill_i32(g->c0, off * g->Kg, 0, 0);
  launch_fill_i32(g->cap, off * g->Kg, 0, 0);
  launch_fill_i32(g->e, off, 0, 0);
  launch_fill_i32(g->h, off, 0, 0);
  launch_fill_i32((int *)g->mask0, off, 0, 0);
  launch_fill_i32((int *)g->rmask, off, 0, 0);
  launch_fill_u8(g->lab, off, 0, 0);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    g->nwg = std::max(64, prop.multiProcessorCount * 8);
  {
    unsigned conns = 0;
    for (const auto &I : g->inst) conns |= 1u << I.conn;
    const double mean = (double)uoff / count;
    g->size_class = std::to_string(g->dim) + ":" + std::to_string(conns) + ":" +
                    std::to_string(log2_round(count)) + ":" + std::to_string(log2_round(mean));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(g->size_class);
    if (it != g_tune.end()) {
      g->tuned = it->second.shape;
      g->tuned_mli = it->second.mli;
      g->tuned_ir = it->second.ir;
      g->tuned_sc = it->second.sc;
      g->tuned_sp = it->second.sp;
    }
  }
  if ((rc = g->set_shape(default_shape(g->dim)))) { delete g; return rc; }
  if (hipDeviceSynchronize() != hipSuccess) {
    delete g;
    return fail(JFC_ERR_DEVICE, "device initialisation failed");
  }
  *out = g;
  return JFC_OK;
}

extern "C" int jfc_create(jfc_graph **g, int dims, int nx, int ny, int nz,
                          int connectivity) {
  if (dims == 2 && nz != 1) return fail(JFC_ERR_INVALID, "2D graph needs nz == 1");
  if (dims == 3 && nz < 2) return fail(JFC_ERR_INVALID, "3D graph needs nz >= 2");
  if (dims != 2 && dims != 3) return fail(JFC_ERR_INVALID, "dims must be 2 or 3");
  const int32_t s[4] = {nx, ny, nz, connectivity};
  return create_impl(g, 1, s);
}

extern "C" int jfc_create_batch(jfc_graph **g, int count, const int32_t *shapes) {
  return create_impl(g, count, shapes);
}

extern "C" void jfc_destroy(jfc_graph *g) {
  if (!g) return;
  (void)hipDeviceSynchronize();
  delete g;
}

extern "C" int jfc_num_instances(const jfc_graph *g) {
  return g ? (int)g->inst.size() : 0;
}

// ---- inputs ------------------------------------------------------------------------

static int check_inst(jfc_graph *g, int inst) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  if (inst < 0 || inst >= (int)g->inst.size())
    return fail(JFC_ERR_INVALID, "instance index out of range");
  return JFC_OK;
}

// Caller arrays (rows of nx) <-> stored planes (rows of px); pad entries of
// the stored side are left as they are (zero from creation).  Host arrays
// cross PCIe in one contiguous copy through the graph's device staging
// buffer; the row scatter / gather is a device-to-device 2D copy.  Every copy
// and kernel of an entry point is issued on the graph's stream (ustream), and
// the entry point waits for that stream alone before it returns (sync_user).
static int stage_for(jfc_graph *g, size_t bytes, void **p) {
  if (g->stage_bytes < bytes) {
    if (g->stage) (void)hipFree(g->stage);
    g->stage = nullptr;
    g->stage_bytes = 0;
    if (hipMalloc(&g->stage, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(JFC_ERR_OOM, "out of device memory: staging buffer");
    }
    g->stage_bytes = bytes;
  }
  *p = g->stage;
  return JFC_OK;
}
int jfc::graph_stage(jfc_graph *g, size_t bytes, void **p) { return stage_for(g, bytes, p); }

static int sync_user(jfc_graph *g) {
  JFC_HIP(hipStreamSynchronize(g->ustream));
  return JFC_OK;
}
static int put_rows(jfc_graph *g, const InstDesc &I, void *dst, const void *src, size_t elem,
                    int on_device) {
  const size_t row = (size_t)I.nx * elem, rows = (size_t)I.ny * I.nz;
  hipStream_t st = g->ustream;
  if (I.px == I.nx) {
    JFC_HIP(hipMemcpyAsync(dst, src, row * rows,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    return JFC_OK;
  }
  if (!on_device) {
    void *tmp = nullptr;
    if (int rc = stage_for(g, row * rows, &tmp)) return rc;
    JFC_HIP(hipMemcpyAsync(tmp, src, row * rows, hipMemcpyHostToDevice, st));
    src = tmp;
  }
  JFC_HIP(hipMemcpy2DAsync(dst, (size_t)I.px * elem, src, row, row, rows,
                           hipMemcpyDeviceToDevice, st));
  return JFC_OK;
}
static int get_rows(jfc_graph *g, const InstDesc &I, void *dst, const void *src, size_t elem,
                    int on_device) {
  const size_t row = (size_t)I.nx * elem, rows = (size_t)I.ny * I.nz;
  hipStream_t st = g->ustream;
  if (I.px == I.nx) {
    JFC_HIP(hipMemcpyAsync(dst, src, row * rows,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    return JFC_OK;
  }
  if (on_device) {
    JFC_HIP(hipMemcpy2DAsync(dst, row, src, (size_t)I.px * elem, row, rows,
                             hipMemcpyDeviceToDevice, st));
    return JFC_OK;
  }
  void *tmp = nullptr;
  if (int rc = stage_for(g, row * rows, &tmp)) return rc;
  JFC_HIP(hipMemcpy2DAsync(tmp, row, src, (size_t)I.px * elem, row, rows,
                           hipMemcpyDeviceToDevice, st));
  JFC_HIP(hipMemcpyAsync(dst, tmp, row * rows, hipMemcpyDeviceToHost, st));
  return JFC_OK;
}
static int put_plane(jfc_graph *g, const InstDesc &I, int *dst, const int32_t *src,
                     int on_device) {
  return put_rows(g, I, dst, src, sizeof(int32_t), on_device);
}

// Range check of up to two planes of n stored values on the graph's stream;
// waits for the stream.  Returns JFC_ERR_OVERFLOW naming the first plane
// with a value outside [lo, hi].
static int validate(jfc_graph *g, const int *p0, const int *p1, long long n, int lo, int hi,
                    const char *what0, const char *what1) {
  if (int rc = g->ensure_pin()) return rc;
  hipStream_t st = g->ustream;
  launch_fill_i32(g->d_flagword + 1, 2, 0, st);
  launch_validate(p0, n, lo, hi, g->d_flagword + 1, st);
  if (p1) launch_validate(p1, n, lo, hi, g->d_flagword + 2, st);
  JFC_HIP(hipMemcpyAsync(&g->pin->bad, g->d_flagword + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  JFC_HIP(hipMemcpyAsync(&g->pin->flag, g->d_flagword + 2, sizeof(int), hipMemcpyDeviceToHost, st));
  if (int rc = sync_user(g)) return rc;
  const char *bad = g->pin->bad ? what0 : (p1 && g->pin->flag) ? what1 : nullptr;
  if (bad)
    return fail(JFC_ERR_OVERFLOW, std::string(bad) + " outside [" + std::to_string(lo) + ", " +
                                      std::to_string(hi) + "]");
  return JFC_OK;
}

// A setter that failed leaves the plane(s) it wrote zeroed and the instance
// marked incomplete: the solve refuses it until the inputs are set again.
static void input_failed(jfc_graph *g, int inst, int bits) {
  g->caps_set[inst] &= ~bits;
  g->caps_failed[inst] |= bits;
  g->mask_dirty = true;
  g->solved = false;
}

extern "C" int jfc_set_stream(jfc_graph *g, void *stream) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  g->ustream = (hipStream_t)stream;
  return JFC_OK;
}

extern "C" int jfc_set_terminal_caps(jfc_graph *g, int inst, const int32_t *cap_source,
                                     const int32_t *cap_sink, int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!cap_source || !cap_sink) return fail(JFC_ERR_INVALID, "null capacity pointer");
  const long long n = g->inst_stored(inst);
  const long long o = g->inst[inst].node_off;
  if (!(rc = put_plane(g, g->inst[inst], g->cs + o, cap_source, on_device)) &&
      !(rc = put_plane(g, g->inst[inst], g->ct + o, cap_sink, on_device)))
    rc = validate(g, g->cs + o, g->ct + o, n, 0, JFC_MAX_TERMINAL_CAP, "source capacity",
                  "sink capacity");
  if (rc) {
    launch_fill_i32(g->cs + o, n, 0, g->ustream);
    launch_fill_i32(g->ct + o, n, 0, g->ustream);
    (void)hipStreamSynchronize(g->ustream);
    input_failed(g, inst, 1);
    return rc;
  }
  g->caps_set[inst] |= 1;
  g->caps_failed[inst] &= ~1;
  g->solved = false;
  return JFC_OK;
}

extern "C" int jfc_set_excess(jfc_graph *g, int inst, const int32_t *ef, int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!ef) return fail(JFC_ERR_INVALID, "null excess pointer");
  const long long n = g->inst_stored(inst);
  const long long o = g->inst[inst].node_off;
  // staged in the instance's slice of the excess plane, which every solve
  // rewrites from cs - ct before reading it
  rc = put_plane(g, g->inst[inst], g->e + o, ef, on_device);
  if (!rc)
    rc = validate(g, g->e + o, nullptr, n, -JFC_MAX_TERMINAL_CAP, JFC_MAX_TERMINAL_CAP,
                  "excess", nullptr);
  if (!rc) {
    launch_excess_split(g->e + o, g->cs + o, g->ct + o, n, g->ustream);
    rc = sync_user(g);
  }
  if (rc) {
    launch_fill_i32(g->cs + o, n, 0, g->ustream);
    launch_fill_i32(g->ct + o, n, 0, g->ustream);
    (void)hipStreamSynchronize(g->ustream);
    input_failed(g, inst, 1);
    return rc;
  }
  g->caps_set[inst] |= 1;
  g->caps_failed[inst] &= ~1;
  g->solved = false;
  return JFC_OK;
}

extern "C" int jfc_set_neighbor_caps(jfc_graph *g, int inst, int arc, const int32_t *cap,
                                     int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!cap) return fail(JFC_ERR_INVALID, "null capacity pointer");
  if (arc < 0 || arc >= g->inst[inst].K)
    return fail(JFC_ERR_INVALID, "arc index outside the instance's connectivity");
  const long long n = g->inst_stored(inst);
  int *dst = g->c0 + (long long)arc * g->Ntot + g->inst[inst].node_off;
  rc = put_plane(g, g->inst[inst], dst, cap, on_device);
  if (!rc) rc = validate(g, dst, nullptr, n, 0, JFC_MAX_NEIGHBOR_CAP, "neighbour capacity", nullptr);
  // on failure the plane is zeroed and the mask of the inputs is rebuilt
  // before the next solve, so no stale residual bit of the old arcs survives
  g->mask_dirty = true;
  g->solved = false;
  if (rc) {
    launch_fill_i32(dst, n, 0, g->ustream);
    (void)hipStreamSynchronize(g->ustream);
    input_failed(g, inst, 2 << arc);
    return rc;
  }
  launch_zero_border(g->shape, g->inst[inst], g->c0 + (long long)arc * g->Ntot, arc, g->ustream);
  if ((rc = sync_user(g))) return rc;
  g->caps_set[inst] |= 2 << arc;
  g->caps_failed[inst] &= ~(2 << arc);
  return JFC_OK;
}

extern "C" int jfc_set_seeds(jfc_graph *g, int inst, const uint8_t *fg, const uint8_t *bg,
                             int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  const InstDesc &I = g->inst[inst];
  const long long o = I.node_off;
  const long long rows = (long long)I.ny * I.nz, n = rows * I.nx;
  // the kernel reads the caller's unpadded seed bytes; host bytes are staged
  // in one contiguous copy each
  const unsigned char *dfg = fg, *dbg = bg;
  if (!on_device && (fg || bg)) {
    void *tmp = nullptr;
    if ((rc = stage_for(g, 2 * (size_t)n, &tmp))) return rc;
    unsigned char *s = (unsigned char *)tmp;
    if (fg) JFC_HIP(hipMemcpyAsync(s, fg, n, hipMemcpyHostToDevice, g->ustream));
    if (bg) JFC_HIP(hipMemcpyAsync(s + n, bg, n, hipMemcpyHostToDevice, g->ustream));
    dfg = fg ? s : nullptr;
    dbg = bg ? s + n : nullptr;
  }
  launch_seeds(g->cs + o, g->ct + o, dfg, dbg, I.nx, I.px, rows, g->ustream);
  if ((rc = sync_user(g))) return rc;
  g->caps_set[inst] |= 1;
  g->solved = false;
  return JFC_OK;
}

extern "C" int jfc_input_planes(jfc_graph *g, int inst, int32_t **cap_source,
                                int32_t **cap_sink, int32_t **neighbor, int64_t *row_pitch,
                                int64_t *plane_stride) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  const long long o = g->inst[inst].node_off;
  if (cap_source) *cap_source = g->cs + o;
  if (cap_sink) *cap_sink = g->ct + o;
  if (neighbor) *neighbor = g->c0 + o;
  if (row_pitch) *row_pitch = g->inst[inst].px;
  if (plane_stride) *plane_stride = g->Ntot;
  return JFC_OK;
}

extern "C" int jfc_inputs_written(jfc_graph *g, int inst) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  graph_mark_inputs(g, inst);
  return JFC_OK;
}

// ---- solve -------------------------------------------------------------------------------

namespace {

struct Timer {
  std::vector<hipEvent_t> *pool;
  size_t used = 0;
  hipStream_t st;
  hipEvent_t next() {
    if (used == pool->size()) {
      hipEvent_t ev;
      if (hipEventCreate(&ev) != hipSuccess) return nullptr;
      pool->push_back(ev);
    }
    return (*pool)[used++];
  }
  hipEvent_t mark() {
    hipEvent_t ev = next();
    if (ev) (void)hipEventRecord(ev, st);
    return ev;
  }
  static double ms(hipEvent_t a, hipEvent_t b) {
    float f = 0.f;
    if (a && b && hipEventElapsedTime(&f, a, b) == hipSuccess) return f;
    return 0.0;
  }
};

struct Solver {
  jfc_graph *g;
  jfc_options o;
  jfc_stats s;
  hipStream_t st;
  Timer tm;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> push_ev, gr_ev, cut_ev;

  SweepCtl sweep_ctl() const {
    SweepCtl c;
    c.hist = g->d_sweep;
    c.state = g->d_sweep_state;
    c.total = g->d_total;
    c.cap = g->sweep_cap;
    c.max_sweeps = o.max_sweeps;
    c.stall_ratio = o.stall_ratio;
    c.budget = o.relabel_work > 0 ? (long long)std::floor(o.relabel_work * g->ntiles) : 0;
    c.ring = g->d_ring;
    c.ring_n = kStopRing;
    return c;
  }

  Queue queue_from_list() {
    Queue q{g->d_list[g->slot], g->d_ctr, g->slot, -1};
    return q;
  }
  void advance() { g->slot ^= 1; }

  // relabel / cut passes issued between two host checks
  static constexpr int kPassGroup = 4;

  enum SearchMode { kRelabel = 0, kCut = 1, kTag = 2 };

  // Tile passes of a monotone search (relabel, cut or tag) until no tile is
  // dirty.  The first pass visits every tile (relabel, in jump-flooding
  // distance order) or the seed tiles (cut, tag); later passes visit the
  // dirty tiles, alternating direction (back-and-forth,
  // QGCWidget.cpp:2221-2240).  Tile visits are counted on the device
  // (d_total) and read once per solve.
  // cleared: the caller has zeroed the dirty flags and the distance
  // histogram already (search_fills).
  int search(int mode, int *passes, bool cleared = false) {
    const jfc_graph::SearchTiling ST = g->search_tiling(mode == kRelabel);
    const Tiles &T = ST.T;
    Planes P = g->planes();
    if (mode == kTag) P.lab = g->tag;
    // Relabel passes take the tiles in distance order from the tiles holding
    // a deficit node (d_defi, kept by the fold and the push visits): a relax
    // visit reads its halo heights device-coherently, so a tile settled
    // earlier in the same pass is seen and one pass carries the front across
    // many tiles.  The cut and tag floods read their halo labels with plain
    // loads and gain nothing from the order (the 15000x10000 8-connected
    // cut's: 0.58 ms ordered, 0.50 ms not, the same 9 passes; the 2560x1440x50
    // video volume's: 0.74 -> 0.61 ms), and a graph
    // of at most kSmallTiles tiles gives every tile a workgroup of its own in
    // a pass: those take the tiles in colour order and skip the five launches
    // of the distance ordering.
    // (A relabel tiling other than the push tiling has no colour lists: its
    // relabels always take the distance order.)
    int *order = ST.order;
    if ((ST.ntiles <= kSmallTiles && !ST.pinst) || mode != kRelabel) {
      order = g->d_color_tiles;
    } else {
      launch_tile_order(ST.shape, T, ST.defi, ST.jf0, ST.jf1, ST.tdist, ST.hist, ST.nbins,
                        ST.order, ST.max_tile_dim, nullptr, ST.lines, o.tile_order, st, cleared);
    }
    // A small graph's passes after the first take their tiles straight from
    // the flags (Queue::pick over the colour lists, no compaction launch).
    // The flags are then double-buffered: pass p takes the tiles flagged in
    // D[p & 1] and flags the next pass's in D[(p + 1) & 1], so a flag set
    // during a pass is read only after the launch boundary that makes the
    // labels or heights behind it visible.
    const bool direct = ST.ntiles <= kSmallTiles && !ST.pinst;
    unsigned char *const D[2] = {direct ? g->d_dirty2 : ST.dirty, ST.dirty};
    if (!cleared) {
      launch_fill_u8(ST.dirty, ST.ntiles, 0, st);
      if (direct) launch_fill_u8(g->d_dirty2, ST.ntiles, 0, st);
    }
    const int nwg = std::min(g->nwg, ST.ntiles);
    // totals: [1] relabel tiles, [2] cut tiles, [4] tag tiles
    long long *total = g->d_total + (mode == kTag ? 4 : 1 + mode);
    if (int rc = g->ensure_pin()) return rc;
    // dirty: the flags this pass sets for the next
    auto flood = [&](const Queue &q, bool full_pass, unsigned char *dirty) {
      if (mode == kCut) launch_reach(ST.shape, nwg, T, P, q, dirty, g->d_flagword, st);
      else if (mode == kTag) launch_tag(ST.shape, nwg, T, P, q, dirty, g->lab, st);
      else launch_relax(ST.shape, nwg, T, P, q, dirty, g->d_flag, full_pass && g->dim != 3,
                        g->gr_tag, st, ST.pinst, ST.pdims);
    };
    if (mode != kRelabel) {
      // tiles without a seed gain labels only through their halo, so the
      // first pass visits the seed tiles (in distance order)
      if (direct) {
        Queue qd{order, g->d_ctr, g->slot, ST.ntiles};
        qd.pick = g->d_seed;
        qd.cnt = g->d_ctr + QSLOT * g->slot;
        qd.total = total;
        flood(qd, false, D[1]);
      } else {
        Queue q = queue_from_list();
        launch_compact(order, ST.ntiles, 0, g->d_seed, nullptr, nullptr, q, nullptr, total,
                       nullptr, st);
        flood(q, false, D[1]);
      }
    } else {
      flood(Queue{order, g->d_ctr, g->slot, ST.ntiles}, true, D[1]);
    }
    advance();
    *passes += 1;
    // later passes over the dirty tiles, alternating direction, in groups;
    // an empty compaction ends the search.  Each group ends by copying its
    // last list count to pinned memory, and groups are issued one ahead of
    // the host check, so the device runs group k+1 while the host waits for
    // group k's count instead of idling through every check.  A group issued
    // after convergence compacts empty lists and changes nothing.
    int pass = 1;
    auto issue_group = [&](int k) -> int {
      // the first group is short: a cut flood has done its work within a
      // pass or two, and passes queued past the end each cost two launches
      const int size = k == 0 ? 2 : kPassGroup;
      for (int j = 0; j < size; ++j, ++pass) {
        if (direct) {
          // the pass's visit count goes to its slot's count word (zeroed by
          // the launch before, as a compaction's would be)
          Queue qd{order, g->d_ctr, g->slot, ST.ntiles};
          qd.pick = D[pass & 1];
          qd.cnt = g->d_ctr + QSLOT * g->slot;
          qd.total = total;
          flood(qd, false, D[(pass + 1) & 1]);
        } else {
          Queue qc = queue_from_list();
          launch_compact(order, ST.ntiles, pass & 1, ST.dirty, nullptr, nullptr, qc, nullptr,
                         total, nullptr, st);
          flood(qc, false, ST.dirty);
        }
        advance();
        *passes += 1;
      }
      // the last pass's slot is the one before the current
      JFC_HIP(hipMemcpyAsync(&g->pin->cnt[k & 1], g->d_ctr + QSLOT * (g->slot ^ 1), sizeof(int),
                             hipMemcpyDeviceToHost, st));
      JFC_HIP(hipEventRecord(g->cnt_ev[k & 1], st));
      return JFC_OK;
    };
    if (int rc = issue_group(0)) return rc;
    for (int k = 0;; ++k) {
      if (int rc = issue_group(k + 1)) return rc;
      JFC_HIP(hipEventSynchronize(g->cnt_ev[k & 1]));
      if (g->pin->cnt[k & 1] == 0) break;
      if (pass > 1000000) return fail(JFC_ERR_STATE, "search did not settle");
    }
    return JFC_OK;
  }

  // Tag post-process over the current labels; g->tag holds the foreground
  // seed bytes on entry.
  int tag_pass(int *passes) {
    launch_fill_i32(g->d_ctr, 2 * QSLOT, 0, st);
    g->slot = 0;
    launch_tag_init(g->shape, g->tiles(), g->tag, g->d_seed, st);
    return search(kTag, passes);
  }

  // Exact relabel (graphcut_init_global_relabel + graphcut_global_relabel,
  // graphcut_global_relabel.cl:19-124).  In 2D there is no separate init pass
  // over the nodes: the first tile pass starts every tile from e and marks
  // the heights it writes with a fresh tag, so a neighbour's stale height is
  // recognised and replaced by its start value (a relabel of the
  // 15000x10000 8-connected cut 1.03 -> 0.94 ms).  3D relabels keep the pass, which also makes the tile seeds
  // exact: 3D tiles have a halo as large as the tile, and without the pass
  // (halo excess read in the first tile pass, seeds from the push visits'
  // deficit flags) the 2560x1440x50 video volume took 14.7 ms instead of
  // 13.4.
  // The zeroing a search needs (dirty flags, distance histogram), to go into
  // its caller's fill launch.
  void search_fills(FillSet &f, bool relabel = false) const {
    const jfc_graph::SearchTiling ST = g->search_tiling(relabel);
    f.add(ST.hist, ST.nbins, 0, 4);
    f.add(ST.dirty, ST.ntiles, 0, 1);
    if (ST.ntiles <= kSmallTiles && !ST.pinst) f.add(g->d_dirty2, ST.ntiles, 0, 1);
  }

  int global_relabel() {
    hipEvent_t a = tm.mark();
    g->gr_tag = 3 - g->gr_tag;
    const jfc_graph::SearchTiling ST = g->search_tiling(true);
    {
      FillSet f;
      search_fills(f, true);
      if (g->dim == 3) f.add(ST.defi, ST.ntiles, 0, 1);
      // the relax passes set the push tiling's active flags, never clear them
      if (ST.pinst) f.add(g->d_flag, g->ntiles, 0, 1);
      launch_fill_multi(f, st);
    }
    if (g->dim == 3)
      launch_gr_init(ST.shape, ST.T, g->planes(), g->gr_tag, ST.defi, st, true);
    int rc = search(kRelabel, &s.relabel_passes, true);
    if (rc) return rc;
    launch_fill_u8(g->d_hmat, g->ntiles, 1, st);  // every height is stored now
    gr_ev.push_back({a, tm.mark()});
    s.global_relabels++;
    return JFC_OK;
  }

  int cuts = 0;  // canonical-cut searches run

  // Returns 1 in *found when an augmenting path still exists.
  hipEvent_t last_cut_end = nullptr;

  // Returns 1 in *found when an augmenting path still exists.  The search is
  // the solve's last step when it finds none, so the certificate word comes
  // back with every result of the solve, in one copy of the results block.
  int canonical_cut(int *found) {
    hipEvent_t a = tm.mark();
    ++cuts;
    {
      FillSet f;
      f.add(g->d_flagword, 1, 0, 4);
      f.add(g->d_seed, g->ntiles, 0, 1);
      f.add(g->d_def, 2LL * (long long)g->inst.size(), 0, 4);
      search_fills(f);
      launch_fill_multi(f, st);
    }
    launch_cut_init(g->shape, g->tiles(), g->planes(), g->d_seed, g->d_def, st, true);
    int rc = search(kCut, &s.cut_passes, true);
    if (rc) return rc;
    last_cut_end = tm.mark();
    cut_ev.push_back({a, last_cut_end});
    JFC_HIP(hipMemcpyAsync(g->pin_res, g->d_res, g->res_words * sizeof(long long),
                           hipMemcpyDeviceToHost, st));
    JFC_HIP(hipEventRecord(g->cnt_ev[0], st));
    JFC_HIP(hipEventSynchronize(g->cnt_ev[0]));
    *found = *((const int *)g->pin_res + ((const int *)g->d_flagword - (const int *)g->d_res));
    return JFC_OK;
  }

  // One push-relabel sweep: per colour class, compact the flagged tiles and
  // discharge them; then file the sweep's active-tile count.  The sequence is
  // the same for every sweep that starts on the same queue slot (the colour
  // count is even), which is what lets replay_sweep() capture it once.
  void issue_sweep(hipStream_t s_, bool timed) {
    const Tiles T = g->tiles();
    const Planes P = g->planes();
    for (int c = 0; c < g->ncolors; ++c) {
      Queue q = queue_from_list();
      const int n = g->color_off[c + 1] - g->color_off[c];
      if (n <= kSmallTiles) {
        // a colour class of a small graph: the push workgroups pick their
        // flagged tiles from the class's list themselves (one per workgroup)
        q = Queue{g->d_color_tiles + g->color_off[c], g->d_ctr, g->slot, n};
        q.pick = g->d_flag;
        q.mat = g->d_mat;
        q.touched = g->d_touched;
        q.hmat = g->d_hmat;
        q.cnt = g->d_sweep_state;
        q.total = g->d_total;
        q.stop = g->d_sweep_state + 2;
      } else {
        launch_compact(g->d_color_tiles + g->color_off[c], n, 0, g->d_flag, g->d_mat,
                       g->d_touched, q, g->d_sweep_state, g->d_total, g->d_sweep_state + 2, s_,
                       g->d_hmat);
      }
      hipEvent_t a = (timed && o.profile) ? tm.mark() : nullptr;
      launch_push(g->shape, std::min(g->nwg, std::max(1, n)), T, P, q, g->d_flag, g->d_mat,
                  g->d_touched, g->d_defi, g->d_hmat, o.max_local_iters, o.sparse_local_iters,
                  g->d_sweep_state, g->d_total + kIterSlots, s_);
      if (timed && o.profile) push_ev.push_back({a, tm.mark()});
      advance();
    }
    launch_sweep_end(sweep_ctl(), s_);
  }

  // One Fast-Cut sweep (QGCWidget::relabelPushFast, QGCW:2156-2205): every
  // active tile in one list, Jacobi relabel, line pushes, active flags.
  void issue_fast_sweep() {
    Queue q = queue_from_list();
    launch_compact(g->d_color_tiles, g->ntiles, 0, g->d_flag, nullptr, nullptr, q,
                   g->d_sweep_state, g->d_total, g->d_sweep_state + 2, st);
    hipEvent_t a = o.profile ? tm.mark() : nullptr;
    launch_fc_sweep(g->shape, std::min(g->nwg, g->ntiles), g->tiles(), g->planes(), q,
                    g->d_flag, g->d_defi, st);
    if (o.profile) push_ev.push_back({a, tm.mark()});
    advance();
    launch_sweep_end(sweep_ctl(), st);
  }

  // Launch one sweep as a hipGraph captured on first use (per starting
  // slot, stream and local-iteration cap; tile-table or counter changes drop
  // the captured graphs).  The sweep contains no memset node.
  int replay_sweep() {
    auto &sg = g->sweep_graph[g->slot];
    if (!sg.exec || sg.st != st || sg.mli != o.max_local_iters ||
        sg.sparse != o.sparse_local_iters || sg.hist != g->d_sweep ||
        sg.max_sweeps != o.max_sweeps || sg.stall != o.stall_ratio ||
        sg.work != o.relabel_work) {
      if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
      sg = jfc_graph::SweepGraph{};
      const int slot0 = g->slot;
      hipGraph_t graph = nullptr;
      JFC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      issue_sweep(st, false);
      const hipError_t ec = hipStreamEndCapture(st, &graph);
      g->slot = slot0;
      if (ec != hipSuccess) return fail(JFC_ERR_DEVICE, std::string("sweep capture: ") + hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        sg.exec = nullptr;
        return fail(JFC_ERR_DEVICE, std::string("sweep graph: ") + hipGetErrorString(ei));
      }
      sg.st = st;
      sg.mli = o.max_local_iters;
      sg.sparse = o.sparse_local_iters;
      sg.max_sweeps = o.max_sweeps;
      sg.stall = o.stall_ratio;
      sg.work = o.relabel_work;
      sg.hist = g->d_sweep;
    }
    JFC_HIP(hipGraphLaunch(sg.exec, st));
    for (int c = 0; c < g->ncolors; ++c) advance();
    return JFC_OK;
  }

  int run() {
    const Tiles T = g->tiles();
    const Planes P = g->planes();
    const int nI = (int)g->inst.size();
    // reset the residual state from the inputs
    if (g->sweep_cap < o.max_sweeps + 1) {
      g->drop_graphs();
      dfree(g->d_sweep);
      int rc = dalloc(&g->d_sweep, (size_t)o.max_sweeps + 1, "sweep counters");
      if (rc) return rc;
      g->sweep_cap = o.max_sweeps + 1;
    }
    hipEvent_t t0 = tm.mark();
    // counters and per-tile states in one launch (the sweep history needs no
    // clearing: each sweep files its entry before any read of it); no
    // heights stored yet, so the push kernel reads them as 0 (the Fast-Cut
    // kernels read every height, and the fold writes the zeros for them)
    {
      FillSet f;
      f.add(g->d_ctr, 2 * QSLOT, 0, 4);
      f.add(g->d_red, 2LL * nI * RED_SLOTS, 0, 4);
      f.add(g->d_total, 2 * kTotals, 0, 4);
      f.add(g->d_sweep_state, kSweepState, 0, 4);
      f.add(g->d_mat, g->ntiles, 0, 1);
      f.add(g->d_hmat, g->ntiles, 0, 1);
      f.add(g->d_touched, g->ntiles, 0, 1);
      // the fold's tile flags
      if (!o.initial_relabel) f.add(g->d_flag, g->ntiles, 0, 1);
      f.add(g->d_defi, g->ntiles, 0, 1);
      launch_fill_multi(f, st);
    }
    g->slot = 0;
    // the residual-arc masks restart from the inputs' with every plane
    // marked fresh: k_push reads fresh planes from c0 and writes them to cap
    const long long K = g->Kg, nodes = g->Nuser;
    // one pass over the nodes folds the terminal arcs (read cs, ct, write e)
    // and starts the masks: built from the input planes after an input
    // change (read K planes, write rmask), else copied from mask0 (read
    // mask0, write rmask).  mask0 is written by the first solve that can
    // reuse it -- the second solve of the same inputs -- so a solve of new
    // inputs, the common case, skips that write (Fast-Cut reads mask0 at
    // once and always has it written).
    const bool fast = o.method == JFC_METHOD_FASTCUT;
    int masks;
    if (g->mask_dirty) {
      g->mask_dirty = false;
      s.inputs_fresh = 1;
      masks = fast ? 2 : 3;
      g->mask0_ok = fast;
    } else {
      masks = g->mask0_ok ? 1 : 2;
      g->mask0_ok = true;
    }
    s.solve_bytes += nodes * (12 + (masks == 2 ? 4 * K + 8 : masks == 3 ? 4 * K + 4 : 8));
    if (int rc0 = g->ensure_pin()) return rc0;
    launch_fold(g->shape, T, P, g->d_red, o.initial_relabel ? nullptr : g->d_flag, g->d_defi,
                o.method == JFC_METHOD_FASTCUT && !o.initial_relabel, g->mask0, masks, st, true);
    // Fast-Cut reads and writes the residual planes directly
    if (o.method == JFC_METHOD_FASTCUT) launch_fc_materialize(P, g->mask0, g->Kg, st);
    // the fold's sums are read with the results (no host wait here): an
    // instance without terminal capacity has no excess, so the solve passes
    // over it, and the error is reported at the end
    std::vector<long long> red((size_t)nI * RED_SLOTS);
    int rc;
    if (o.initial_relabel) {
      if ((rc = global_relabel())) return rc;
    }  // else the zero heights (a valid labelling) and the fold's flags
    // Sweeps go out up to `ahead` beyond the last one whose stop word the
    // host has read.  The device decides after every sweep whether to go on
    // (k_sweep_end: no active tile, relabel due, budget spent), and a sweep
    // queued after a stop compacts nothing, so the host's wait for a stop
    // word overlaps the sweeps queued behind it.
    const int ahead = std::min(std::max(1, o.check_every), kStopRing - 1);
    int cert_rounds = 0;
    bool done = false;
    // sweeps issued in this solve (no-ops included; the device numbers them
    // the same way), and the stop words the host has read or skipped
    int sent = 0, known = 0;
    // no word of an earlier solve matches a sequence number of this one
    for (int k = 0; k < kStopRing; ++k) g->ring[k] = -1;
    for (;;) {
      int stopw = 0;
      for (;;) {
        while (sent < known + ahead) {
          // the first sweep of a solve runs eagerly (it also primes the
          // launchers' occupancy queries); later ones replay a captured graph
          if (o.method == JFC_METHOD_FASTCUT) {
            issue_fast_sweep();
            s.push_launches += 2 + 2 * g->dim;
          } else {
            if (o.graph_sweeps && !o.profile && sent > 0) {
              if ((rc = replay_sweep())) return rc;
            } else {
              issue_sweep(st, true);
            }
            s.push_launches += g->ncolors;
          }
          JFC_HIP(hipEventRecord(g->stop_ev[sent % kStopRing], st));
          ++sent;
        }
        const int k = known % kStopRing;
        JFC_HIP(hipEventSynchronize(g->stop_ev[k]));
        // the sweep's end kernel has completed, so its word is in the ring
        // (written through to host memory); re-read a few times before
        // giving up on a word that does not carry the expected number
        const volatile int *ring = g->ring;
        int w = ring[k];
        for (int spin = 0; (w >> 3) != known && spin < (1 << 20); ++spin) w = ring[k];
        if ((w >> 3) != known)
          return fail(JFC_ERR_DEVICE, "sweep " + std::to_string(known) +
                                          ": its stop word did not reach the host");
        stopw = w & 7;
        ++known;
        if (stopw) break;
      }
      // the sweeps queued behind the stop are no-ops: skip their words
      known = sent;
      // the sweeps queued behind the stop were no-ops
      if (stopw == 3) break;  // sweep budget spent
      if (stopw == 2 && o.stall_certify > 0) {
        // a stalled active set is often excess cut off from every deficit:
        // the canonical-cut search proves that (and is the solve's last
        // step) at less cost than the relabel that would show it
        int found = 0;
        if ((rc = canonical_cut(&found))) return rc;
        if (!found) {
          done = true;
          break;
        }
      }
      if (stopw == 2 || stopw == 4) {
        if ((rc = global_relabel())) return rc;
        launch_sweep_resume(sweep_ctl(), st);
        continue;
      }
      // no active tile: the canonical cut doubles as the certificate
      int found = 0;
      if ((rc = canonical_cut(&found))) return rc;
      if (!found) {
        done = true;
        break;
      }
      // an augmenting path survived: relabel exactly and continue
      s.certificate_failures++;
      if (++cert_rounds > 16) return fail(JFC_ERR_STATE, "certificate kept failing");
      if ((rc = global_relabel())) return rc;
      launch_sweep_resume(sweep_ctl(), st);
    }
    if (!done) {
      // labels of the current preflow are still useful to callers
      int found = 0;
      if ((rc = canonical_cut(&found))) return rc;
    }
    // the last cut search ran on the final excess, left the remaining deficit
    // of every instance in d_def and copied the results block back (the
    // solve's device work ends there)
    hipEvent_t t1 = last_cut_end;
    const long long *pred = g->pin_res, *pdef = pred + red.size(), *totals = pdef + nI;
    s.sweeps = ((const int *)(totals + kTotals))[1];
    std::copy(pred, pred + red.size(), red.begin());
    const std::vector<long long> def(pdef, pdef + nI);
    for (int i = 0; i < nI; ++i)
      if (red[i * RED_SLOTS + RED_DEF0] == 0 && red[i * RED_SLOTS + RED_SUP0] == 0 &&
          red[i * RED_SLOTS + RED_CONST] == 0)
        return fail(JFC_ERR_NO_SEEDS, "instance " + std::to_string(i) +
                                          " has no terminal capacity (no labels)");
    for (int i = 0; i < nI; ++i)
      g->flows[i] = red[i * RED_SLOTS + RED_CONST] + red[i * RED_SLOTS + RED_DEF0] - def[i];
    s.flow = g->flows[0];
    s.converged = done ? 1 : 0;
    s.push_tiles = totals[0];
    s.local_iters = 0;
    for (int k = 0; k < 64; ++k) s.local_iters += totals[kIterSlots + k];
    g->last_sweeps = s.sweeps;
    // a relabel's first pass visits every tile without a compaction
    s.relabel_tiles = totals[1] + (long long)g->ntiles * s.global_relabels;
    s.cut_tiles = totals[2];
    s.push_bytes = totals[0] * 256LL * push_bytes_per_node(g->Kg);
    // relabels: a first pass over every tile (2D: read e and mask, write h;
    // 3D: after a pass over the nodes that reads e and writes h), later
    // passes read h, e and mask and write h; cut searches: one pass over the
    // nodes (read e, write the label) and flood visits (read label, e and
    // mask, write label)
    const long long first = (long long)g->ntiles * s.global_relabels;
    s.solve_bytes += s.push_bytes + first * 256LL * (g->dim == 3 ? 16 : 12) +
                     totals[1] * 256LL * 16 + (g->dim == 3 ? 8LL * nodes * s.global_relabels : 0) +
                     5LL * nodes * cuts + totals[2] * 256LL * 10;
    s.ms_total = Timer::ms(t0, t1);
    for (auto &p : push_ev) s.ms_push += Timer::ms(p.first, p.second);
    for (auto &p : gr_ev) s.ms_relabel += Timer::ms(p.first, p.second);
    for (auto &p : cut_ev) s.ms_cut += Timer::ms(p.first, p.second);
    g->solved = true;
    return done ? JFC_OK : JFC_NOT_CONVERGED;
  }
};

}  // namespace

extern "C" int jfc_solve(jfc_graph *g, const jfc_options *opt, jfc_stats *stats) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  Solver S{};
  S.g = g;
  if (opt) S.o = *opt;
  else jfc_default_options(&S.o);
  if (S.o.max_sweeps <= 0) S.o.max_sweeps = 1 << 20;
  if (S.o.max_sweeps > (1 << 24)) S.o.max_sweeps = 1 << 24;
  if (S.o.check_every <= 0) S.o.check_every = 1;
  // 0: the cap jfc_autotune chose for the graph's size class, else 2
  if (S.o.max_local_iters <= 0) S.o.max_local_iters = g->tuned_mli > 0 ? g->tuned_mli : 2;
  // 0: the tuned cap, else kSparseIters; negative: the dense cap
  if (S.o.sparse_local_iters == 0) S.o.sparse_local_iters = g->tuned_sp ? g->tuned_sp : kSparseIters;
  if (S.o.sparse_local_iters < 0) S.o.sparse_local_iters = S.o.max_local_iters;
  // -1: the choice jfc_autotune made for the size class, else none
  if (S.o.initial_relabel < 0) S.o.initial_relabel = g->tuned_ir > 0 ? 1 : 0;
  if (S.o.stall_certify < 0) S.o.stall_certify = g->tuned_sc > 0 ? 1 : 0;
  if (S.o.tile_order != 0 && S.o.tile_order != 1)
    return fail(JFC_ERR_INVALID, "tile_order must be 0 (distance transform) or 1 (jump flooding)");
  if (S.o.method != JFC_METHOD_JFCUT && S.o.method != JFC_METHOD_FASTCUT)
    return fail(JFC_ERR_INVALID, "unknown method");
  if (S.o.method == JFC_METHOD_FASTCUT)
    for (const InstDesc &I : g->inst)
      if (I.K == 8)
        return fail(JFC_ERR_INVALID, "Fast-Cut runs on 4- and 6-connected graphs only");
  const int sh = shape_from_option(
      S.o.tile_shape == JFC_TILE_DEFAULT && g->tuned ? g->tuned : S.o.tile_shape, g->dim);
  if (!sh) return fail(JFC_ERR_INVALID, "tile shape does not match the graph's dimensionality");
  int rc = g->set_shape(sh);
  if (rc) return rc;
  for (size_t i = 0; i < g->inst.size(); ++i)
    if (g->caps_failed[i])
      return fail(JFC_ERR_STATE, "instance " + std::to_string(i) +
                                     ": an input setter failed and zeroed its plane; set the "
                                     "capacities again before solving");
  std::memset(&S.s, 0, sizeof(S.s));
  // the options' stream, else the graph's stream (jfc_set_stream)
  S.st = S.o.stream ? (hipStream_t)S.o.stream : g->ustream;
  if (!S.st) {
    // the legacy default stream cannot be captured: solve on a stream of the
    // graph's own, ordered after the work already queued on the default one
    if (!g->priv_stream) {
      if (hipStreamCreateWithFlags(&g->priv_stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&g->priv_event, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return fail(JFC_ERR_DEVICE, "cannot create the solve stream");
      }
    }
    JFC_HIP(hipEventRecord(g->priv_event, nullptr));
    JFC_HIP(hipStreamWaitEvent(g->priv_stream, g->priv_event, 0));
    S.st = g->priv_stream;
  }
  S.tm.pool = &g->ev_pool;
  S.tm.st = S.st;
  g->solved = false;
  rc = S.run();
  if (stats) *stats = S.s;
  return rc;
}

// ---- outputs -----------------------------------------------------------------------------

extern "C" int jfc_get_flow(jfc_graph *g, int inst, int64_t *flow) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!flow) return fail(JFC_ERR_INVALID, "null output");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  *flow = g->flows[inst];
  return JFC_OK;
}

extern "C" int jfc_get_flows(jfc_graph *g, int64_t *flows) {
  if (!g || !flows) return fail(JFC_ERR_INVALID, "null argument");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  for (size_t i = 0; i < g->flows.size(); ++i) flows[i] = g->flows[i];
  return JFC_OK;
}

extern "C" int jfc_get_labels(jfc_graph *g, int inst, uint8_t *labels, int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!labels) return fail(JFC_ERR_INVALID, "null output");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  if ((rc = get_rows(g, g->inst[inst], labels, g->lab + g->inst[inst].node_off, 1, on_device)))
    return rc;
  return sync_user(g);
}

extern "C" int jfc_get_tags(jfc_graph *g, int inst, const uint8_t *fg, uint8_t *tags,
                            int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!fg || !tags) return fail(JFC_ERR_INVALID, "null foreground or output");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  if (!g->tag && (rc = dalloc(&g->tag, (size_t)g->Ntot, "tag plane"))) return rc;
  const long long o = g->inst[inst].node_off;
  launch_fill_u8(g->tag, g->Ntot, 0, g->ustream);
  if ((rc = put_rows(g, g->inst[inst], g->tag + o, fg, 1, on_device))) return rc;
  Solver S{};
  S.g = g;
  jfc_default_options(&S.o);
  S.st = g->ustream;
  S.tm.pool = &g->ev_pool;
  S.tm.st = g->ustream;
  int passes = 0;
  if ((rc = S.tag_pass(&passes))) return rc;
  if ((rc = get_rows(g, g->inst[inst], tags, g->tag + o, 1, on_device))) return rc;
  return sync_user(g);
}

extern "C" int jfc_get_energy(jfc_graph *g, int inst, int64_t *energy) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!energy) return fail(JFC_ERR_INVALID, "null output");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  if ((rc = g->ensure_pin())) return rc;
  const int nI = (int)g->inst.size();
  launch_fill_i32((int *)g->d_energy, 2LL * nI, 0, g->ustream);
  launch_energy(g->shape, g->tiles(), g->planes(), g->d_energy, g->ustream);
  JFC_HIP(hipMemcpyAsync(&g->pin->red[0], g->d_energy + inst, sizeof(long long),
                         hipMemcpyDeviceToHost, g->ustream));
  if ((rc = sync_user(g))) return rc;
  *energy = g->pin->red[0];
  return JFC_OK;
}

extern "C" int jfc_get_terminal_caps(jfc_graph *g, int inst, int32_t *cap_source,
                                     int32_t *cap_sink, int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  const InstDesc &I = g->inst[inst];
  if (cap_source && (rc = get_rows(g, I, cap_source, g->cs + I.node_off, 4, on_device))) return rc;
  if (cap_sink && (rc = get_rows(g, I, cap_sink, g->ct + I.node_off, 4, on_device))) return rc;
  return sync_user(g);
}

extern "C" int jfc_get_neighbor_caps(jfc_graph *g, int inst, int arc, int32_t *cap,
                                     int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!cap) return fail(JFC_ERR_INVALID, "null output");
  if (arc < 0 || arc >= g->inst[inst].K)
    return fail(JFC_ERR_INVALID, "arc index outside the instance's connectivity");
  if ((rc = get_rows(g, g->inst[inst], cap,
                     g->c0 + (long long)arc * g->Ntot + g->inst[inst].node_off, 4, on_device)))
    return rc;
  return sync_user(g);
}

extern "C" int jfc_debug_tiles(jfc_graph *g, int inst, int32_t *ntiles, int32_t *tile_dims,
                               int32_t *grid_dims, int32_t *tile_off) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  int tw, th, td;
  shape_dims(g->shape, &tw, &th, &td);
  if (ntiles) *ntiles = g->ntiles;
  if (tile_dims) { tile_dims[0] = tw; tile_dims[1] = th; tile_dims[2] = td; }
  const InstDesc &I = g->inst[inst];
  if (grid_dims) { grid_dims[0] = I.tx; grid_dims[1] = I.ty; grid_dims[2] = I.tz; }
  if (tile_off) *tile_off = I.tile_off;
  return JFC_OK;
}

extern "C" int jfc_debug_tile_order(jfc_graph *g, int method, const uint8_t *tile_seed,
                                    int32_t *tile_dist, int32_t *order, int32_t *max_dist) {
  if (!g || !tile_seed) return fail(JFC_ERR_INVALID, "null argument");
  if (method != 0 && method != 1) return fail(JFC_ERR_INVALID, "method must be 0 or 1");
  JFC_HIP(hipMemcpy(g->d_seed, tile_seed, g->ntiles, hipMemcpyHostToDevice));
  launch_tile_order(g->shape, g->tiles(), g->d_seed, g->d_jf0, g->d_jf1, g->d_tdist,
                    g->d_hist, g->nbins, g->d_order, g->max_tile_dim, g->d_flagword + 2,
                    g->lines(), method, 0);
  JFC_HIP(hipDeviceSynchronize());
  if (tile_dist) JFC_HIP(hipMemcpy(tile_dist, g->d_tdist, g->ntiles * 4, hipMemcpyDeviceToHost));
  if (order) JFC_HIP(hipMemcpy(order, g->d_order, g->ntiles * 4, hipMemcpyDeviceToHost));
  if (max_dist) JFC_HIP(hipMemcpy(max_dist, g->d_flagword + 2, 4, hipMemcpyDeviceToHost));
  return JFC_OK;
}

extern "C" int jfc_debug_tile_flags(jfc_graph *g, int which, uint8_t *flags) {
  if (!g || !flags || which < 0 || which > 2) return fail(JFC_ERR_INVALID, "bad argument");
  const int nI = (int)g->inst.size();
  launch_fill_i32((int *)g->d_red, 2LL * nI * RED_SLOTS, 0, 0);
  launch_fold(g->shape, g->tiles(), g->planes(), g->d_red, which == 2 ? g->d_flag : nullptr,
              g->d_defi, false, nullptr, 0, 0);
  unsigned char *src = which == 0 ? g->d_defi : g->d_flag;
  if (which == 1) {
    launch_cut_init(g->shape, g->tiles(), g->planes(), g->d_seed, g->d_def, 0);
    src = g->d_seed;
  }
  JFC_HIP(hipDeviceSynchronize());
  JFC_HIP(hipMemcpy(flags, src, g->ntiles, hipMemcpyDeviceToHost));
  g->solved = false;
  return JFC_OK;
}

extern "C" int jfc_autotune(jfc_graph *g, const jfc_options *opt, int reps, int *chosen) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  if (reps < 1) reps = 1;
  jfc_options o;
  if (opt) o = *opt;
  else jfc_default_options(&o);
  const int cand2[3] = {JFC_TILE_32x8, JFC_TILE_64x4, JFC_TILE_16x16};
  const int cand3[2] = {JFC_TILE_8x8x4, JFC_TILE_32x4x2};
  const int *cand = g->dim == 3 ? cand3 : cand2;
  const int nc = g->dim == 3 ? 2 : 3;
  // the local-iteration cap is tuned too unless the caller fixed it: fronts
  // that cross tiles in few visits favour more iterations per visit
  // (the 512^3 volume: 21.6 -> 14.3 ms at 4), others favour 2 (the
  // 2560x1440x50 video volume)
  const int mli_fixed = o.max_local_iters > 0 ? o.max_local_iters : 0;
  const int mcand[3] = {2, 3, 4};
  const int nm = mli_fixed ? 1 : 3;
  // so is the relabel before the first sweep: it pays where zero heights
  // let flow slosh for sweeps (the 2560x1440x50 video volume: 12.2 -> 9.4
  // ms) and costs where the push sweeps settle most of the flow first (the
  // 15000x10000 8-connected cut: 4.8 -> 8.5 ms)
  const int ir_fixed = o.initial_relabel >= 0 ? o.initial_relabel : -1;
  const int ni = ir_fixed >= 0 ? 1 : 2;
  // and whether a stalled active set first tries the certificate (it pays
  // where the last relabel would only prove the remaining excess cut off):
  // it changes how many sweeps and relabels a shape needs, so it is tried
  // with every shape rather than after the shape is chosen.  Then, on the
  // best of those, the cap of sparse launches (large where flow must cross
  // whole tiles, small where visits of a few tiles would iterate without
  // effect).
  const int sc_fixed = o.stall_certify >= 0 ? o.stall_certify : -1;
  const int ns = sc_fixed >= 0 ? 1 : 2;
  const int sp_fixed = o.sparse_local_iters != 0 ? o.sparse_local_iters : 0;
  auto time_it = [&](const jfc_options &t, double *med) -> int {
    std::vector<double> ms;
    for (int r = 0; r <= reps; ++r) {
      jfc_stats st;
      const int rc = jfc_solve(g, &t, &st);
      if (rc < 0) return rc;
      if (r > 0) ms.push_back(st.ms_total);  // the first run warms the shape's tables
    }
    std::sort(ms.begin(), ms.end());
    *med = ms[ms.size() / 2];
    return JFC_OK;
  };
  jfc_options best_o = o;
  double best_ms = 0.0;
  bool have = false;
  auto consider = [&](const jfc_options &t) -> int {
    double med = 0.0;
    if (int rc = time_it(t, &med)) return rc;
    if (!have || med < best_ms) {
      best_o = t;
      best_ms = med;
      have = true;
    }
    return JFC_OK;
  };
  o.sparse_local_iters = sp_fixed ? sp_fixed : kSparseIters;
  for (int c = 0; c < nc; ++c)
    for (int m = 0; m < nm; ++m)
      for (int i = 0; i < ni; ++i)
        for (int k = 0; k < ns; ++k) {
          o.tile_shape = cand[c];
          o.max_local_iters = mli_fixed ? mli_fixed : mcand[m];
          o.initial_relabel = ir_fixed >= 0 ? ir_fixed : i;
          o.stall_certify = sc_fixed >= 0 ? sc_fixed : k;
          if (int rc = consider(o)) return rc;
        }
  if (!sp_fixed) {
    const int spc[3] = {-1, 8, 64};
    const jfc_options base = best_o;
    for (int v : spc) {
      jfc_options t = base;
      t.sparse_local_iters = v;
      if (int rc = consider(t)) return rc;
    }
  }
  g->tuned = best_o.tile_shape;
  g->tuned_mli = mli_fixed ? 0 : best_o.max_local_iters;
  g->tuned_ir = ir_fixed >= 0 ? -1 : best_o.initial_relabel;
  g->tuned_sc = sc_fixed >= 0 ? -1 : best_o.stall_certify;
  g->tuned_sp = sp_fixed ? 0 : best_o.sparse_local_iters;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[g->size_class] = Tuned{g->tuned, g->tuned_mli, g->tuned_ir, g->tuned_sc, g->tuned_sp};
  }
  if (chosen) *chosen = g->tuned;
  return JFC_OK;
}

extern "C" int jfc_get_tuning(const jfc_graph *g, jfc_options *opt) {
  if (!g || !opt) return fail(JFC_ERR_INVALID, "null argument");
  opt->tile_shape = g->tuned ? g->tuned : (g->dim == 3 ? JFC_TILE_8x8x4 : JFC_TILE_32x8);
  opt->max_local_iters = g->tuned_mli > 0 ? g->tuned_mli : 2;
  opt->initial_relabel = g->tuned_ir > 0 ? 1 : 0;
  opt->stall_certify = g->tuned_sc > 0 ? 1 : 0;
  opt->sparse_local_iters = g->tuned_sp ? g->tuned_sp : kSparseIters;
  return JFC_OK;
}

extern "C" int jfc_get_tile_shape(const jfc_graph *g, int *shape) {
  if (!g || !shape) return fail(JFC_ERR_INVALID, "null argument");
  *shape = g->tuned ? g->tuned : (g->dim == 3 ? JFC_TILE_8x8x4 : JFC_TILE_32x8);
  return JFC_OK;
}

extern "C" int jfc_get_local_iters(const jfc_graph *g, int *max_local_iters) {
  if (!g || !max_local_iters) return fail(JFC_ERR_INVALID, "null argument");
  *max_local_iters = g->tuned_mli > 0 ? g->tuned_mli : 2;
  return JFC_OK;
}

extern "C" int jfc_get_initial_relabel(const jfc_graph *g, int *initial_relabel) {
  if (!g || !initial_relabel) return fail(JFC_ERR_INVALID, "null argument");
  *initial_relabel = g->tuned_ir > 0 ? 1 : 0;
  return JFC_OK;
}

extern "C" int jfc_graph_size(const jfc_graph *g, int64_t *nodes, int32_t *kmax) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  if (nodes) *nodes = g->Nuser;
  if (kmax) *kmax = g->Kg;
  return JFC_OK;
}

// Gradients into device outputs directly, or into host outputs through the
// staging buffer (n floats per terminal output, K*n for the neighbour planes).
extern "C" int jfc_backward_all(jfc_graph *g, float *d_cap_source, float *d_cap_sink,
                                float *d_neighbor, int on_device) {
  if (!g) return fail(JFC_ERR_INVALID, "null graph");
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  const long long n = g->Nuser;
  hipStream_t st = g->ustream;
  if (on_device) {
    launch_backward_all(g->shape, g->tiles(), g->planes(), d_cap_source, d_cap_sink,
                        d_neighbor, n, st);
    return sync_user(g);
  }
  int rc;
  void *tmp = nullptr;
  if ((rc = stage_for(g, (size_t)n * (2 + g->Kg) * sizeof(float), &tmp))) return rc;
  float *a = d_cap_source ? (float *)tmp : nullptr;
  float *b = d_cap_sink ? (float *)tmp + n : nullptr;
  float *c = d_neighbor ? (float *)tmp + 2 * n : nullptr;
  // planes beyond an instance's connectivity keep the caller's values
  if (c) JFC_HIP(hipMemcpyAsync(c, d_neighbor, n * g->Kg * 4, hipMemcpyHostToDevice, st));
  launch_backward_all(g->shape, g->tiles(), g->planes(), a, b, c, n, st);
  if (a) JFC_HIP(hipMemcpyAsync(d_cap_source, a, n * 4, hipMemcpyDeviceToHost, st));
  if (b) JFC_HIP(hipMemcpyAsync(d_cap_sink, b, n * 4, hipMemcpyDeviceToHost, st));
  if (c) JFC_HIP(hipMemcpyAsync(d_neighbor, c, n * g->Kg * 4, hipMemcpyDeviceToHost, st));
  return sync_user(g);
}

extern "C" int jfc_debug_sweep_counts(jfc_graph *g, int32_t *counts, int32_t n) {
  if (!g || !counts) return fail(JFC_ERR_INVALID, "null argument");
  const int m = std::min(n, g->last_sweeps);
  if (m > 0) JFC_HIP(hipMemcpy(counts, g->d_sweep, m * 4, hipMemcpyDeviceToHost));
  return g->last_sweeps;
}

extern "C" int jfc_backward(jfc_graph *g, int inst, float *d_cap_source, float *d_cap_sink,
                            float *d_neighbor, int on_device) {
  int rc;
  if ((rc = check_inst(g, inst))) return rc;
  if (!g->solved) return fail(JFC_ERR_STATE, "graph not solved");
  const InstDesc &I = g->inst[inst];
  const long long n = g->inst_nodes(inst);
  hipStream_t st = g->ustream;
  if (on_device) {
    launch_backward(g->shape, g->tiles(), g->planes(), I, d_cap_source, d_cap_sink,
                    d_neighbor, st);
    return sync_user(g);
  }
  void *tmp = nullptr;
  if ((rc = stage_for(g, (size_t)n * (2 + I.K) * sizeof(float), &tmp))) return rc;
  float *a = d_cap_source ? (float *)tmp : nullptr;
  float *b = d_cap_sink ? (float *)tmp + n : nullptr;
  float *c = d_neighbor ? (float *)tmp + 2 * n : nullptr;
  launch_backward(g->shape, g->tiles(), g->planes(), I, a, b, c, st);
  if (a) JFC_HIP(hipMemcpyAsync(d_cap_source, a, n * 4, hipMemcpyDeviceToHost, st));
  if (b) JFC_HIP(hipMemcpyAsync(d_cap_sink, b, n * 4, hipMemcpyDeviceToHost, st));
  if (c) JFC_HIP(hipMemcpyAsync(d_neighbor, c, n * I.K * 4, hipMemcpyDeviceToHost, st));
  return sync_user(g);
}
