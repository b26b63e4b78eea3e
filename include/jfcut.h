/*
 * jfcut.h -- C-ABI of libjfcut, the MI355X (gfx950) grid max-flow / min-cut
 * engine that replaces JF-Cut's graph-cut hot path.
 *
 * The reference has no library boundary: its operator surface is implicit in
 * QGCWidget (src/QVisualizer/Graph Cut/QGCWidget.cpp, "QGCW" below).  Each
 * entry point names the reference code it stands in for.
 *
 *   graph  = one or more grid instances (2D 4- or 8-connected, 3D
 *            6-connected), int32 structure-of-arrays capacity planes in HBM.
 *   solve  = block-wise push-relabel on LDS tiles with exact global relabel,
 *            run on the device until no node with excess can reach a node
 *            with deficit; then the canonical cut (nodes reachable from a node
 *            with excess in the residual graph) and the int64 max-flow value.
 *
 * Conventions
 *   - Node v = x + nx*(y + ny*z).  Direction d of a connectivity:
 *       2D: 0:+x 1:-x 2:+y 3:-y 4:(+x,+y) 5:(-x,-y) 6:(-x,+y) 7:(+x,-y)
 *       3D: 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z
 *     The reverse of direction d is d^1.  Neighbour plane d holds, per node
 *     v, the capacity of the arc v -> v + off(d); arcs leaving the grid are
 *     ignored (stored as 0).
 *   - Terminal capacities: 0 <= cap <= JFC_MAX_TERMINAL_CAP (2^30, the hard
 *     seed capacity).  Neighbour capacities: 0 <= cap <= JFC_MAX_NEIGHBOR_CAP.
 *     Larger values are rejected with JFC_ERR_OVERFLOW, so every intermediate
 *     value of the solve fits int32 and the flow is exact.
 *   - Pointers named *_device (or passed with on_device = 1) are HIP device
 *     pointers; all others are host pointers.  Inputs are copied; the caller
 *     keeps ownership.
 *   - Every function returns JFC_OK (0) or a negative JFC_ERR_* code, except
 *     jfc_solve, which may also return JFC_NOT_CONVERGED (1).  The message of
 *     the last error of the calling thread is returned by jfc_last_error().
 *   - A graph may be used by one host thread at a time; distinct graphs are
 *     independent.  Streams: every call of a graph issues its device work on
 *     the graph's stream (jfc_set_stream; NULL = the legacy default stream),
 *     ordered after the work the caller queued there before the call, and
 *     returns after that work completed (it waits for that stream only, not
 *     for the whole device).  jfc_solve uses jfc_options.stream instead when
 *     that is set.  Device pointers passed in are read in that stream order:
 *     a caller producing inputs on another stream must order it first.
 */
#ifndef JFCUT_H
#define JFCUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JFC_OK 0
#define JFC_NOT_CONVERGED 1
#define JFC_ERR_INVALID (-1)
#define JFC_ERR_DEVICE (-2)
#define JFC_ERR_OOM (-3)
#define JFC_ERR_STATE (-4)
#define JFC_ERR_OVERFLOW (-5)
#define JFC_ERR_NO_SEEDS (-6)

#define JFC_MAX_TERMINAL_CAP (1 << 30)
#define JFC_MAX_NEIGHBOR_CAP (1 << 26)

typedef struct jfc_graph jfc_graph;

/* Solver options (the reference's per-tab UI values read in
 * QGCW:2364-2374 and QGCSetting.cpp:19-36, gathered in one struct). */
typedef struct {
  int max_sweeps;         /* totalIterations: cap on push-relabel sweeps (each sweep
                             visits every colour class of active tiles once) */
  int check_every;        /* countIterations: sweeps queued ahead of the host's
                             convergence check (default 2, at most 7).  The
                             device decides after every sweep whether to go on
                             (no active tile, relabel due, sweep budget
                             spent); sweeps queued past that point do nothing */
  double relabel_work;    /* run an exact global relabel once the tiles visited
                             since the last one exceed this many times the tile
                             count (0 = only at start and on certificate failure) */
  int max_local_iters;    /* cap on LDS discharge iterations per tile visit
                             (0 = the cap jfc_autotune chose for the graph's
                             size class, else 2) */
  int tile_shape;         /* JFC_TILE_* (0 = default for the dimensionality) */
  int profile;            /* 1 = time every push-relabel launch with HIP events */
  void *stream;           /* hipStream_t to issue work on (NULL = default stream) */
  int initial_relabel;    /* 1 = exact global relabel before the first sweep; 0 =
                             start from zero heights (the reference's default
                             "no global relabel", QGCW:2394-2401); -1 (default)
                             = the choice jfc_autotune made for the graph's
                             size class, else 0 */
  double stall_ratio;     /* run a global relabel when the active-tile count of a
                             sweep stays within [r, 1/r] of the previous sweep's,
                             or two sweeps running repeat the two before them
                             (the active set has stopped changing, or trades
                             places between two sets of tiles; 0 = off) */
  int graph_sweeps;       /* 1 = replay each sweep's launches (compaction and
                             push per colour class) as one captured hipGraph
                             (default); 0 = launch them one by one.  profile = 1
                             always launches them one by one. */
  int tile_order;         /* tile distance field that orders the relabel
                             passes of graphs of more than 2048 tiles: 0 =
                             separable taxicab distance transform (default, D
                             launches), 1 = jump flooding
                             (graphcut_bfs_block.cl:43-63, 2 ceil(log2) + 2
                             launches); the distances are identical.  Cut
                             and tag passes, and every pass of a smaller
                             graph, take the tiles in colour order */
  int method;             /* JFC_METHOD_JFCUT (0): tile push-relabel in LDS;
                             JFC_METHOD_FASTCUT (1): the widget's Fast-Cut
                             comparator (radioButtonCutMethodJFCut unchecked,
                             QGCW:2419-2427): Jacobi relabel plus line pushes
                             along each axis, 4- and 6-connected graphs only.
                             Same flow and labels either way. */
  int sparse_local_iters; /* cap on LDS discharge iterations per visit in a
                             push launch whose active-tile list is no longer
                             than its grid (each workgroup visits at most one
                             tile), from the solve's 8th sweep on: 0 (default)
                             = the cap jfc_autotune chose for the graph's size
                             class, else 32; negative = max_local_iters (the
                             reference loops each visit until no push,
                             graphcut_push_relabel.cl:77-182) */
  int stall_certify;      /* when the active set stalls (stall_ratio): 1 = run
                             the canonical-cut search first and finish if it
                             finds no augmenting path, else relabel; 0 =
                             relabel; -1 (default) = the choice jfc_autotune
                             made for the graph's size class, else 0 */
} jfc_options;

#define JFC_METHOD_JFCUT 0
#define JFC_METHOD_FASTCUT 1

#define JFC_TILE_DEFAULT 0
#define JFC_TILE_32x8 1
#define JFC_TILE_64x4 2
#define JFC_TILE_16x16 3
#define JFC_TILE_8x8x4 4
/* 32x4x2 push tiles (whole 128-byte rows) run their global relabels on
 * 8x8x4 tiles, whose relabel fronts cross half as many tiles in y and z. */
#define JFC_TILE_32x4x2 5

typedef struct {
  int64_t flow;             /* max-flow of instance 0 (all instances: jfc_get_flows) */
  int converged;            /* 1 if every instance reached the certificate */
  int sweeps;               /* push-relabel sweeps run */
  int global_relabels;      /* exact global relabels run */
  int relabel_passes;       /* tile passes inside the global relabels */
  int cut_passes;           /* tile passes of the canonical-cut search */
  int certificate_failures; /* cut searches that still found an augmenting path */
  int64_t push_tiles;       /* tile visits of the push-relabel kernel */
  int64_t push_launches;    /* push-relabel kernel launches */
  int64_t relabel_tiles;    /* tile visits of the relabel kernel */
  int64_t cut_tiles;        /* tile visits of the cut kernel */
  int64_t push_bytes;       /* algorithmic HBM bytes of the push-relabel visits */
  double ms_total;          /* device time of the solve (HIP events) */
  double ms_push;           /* summed push-relabel launch time (profile = 1) */
  double ms_relabel;        /* global relabel phases */
  double ms_cut;            /* canonical cut + certificate */
  int64_t local_iters;      /* LDS discharge iterations summed over tile visits */
  int64_t solve_bytes;      /* algorithmic HBM bytes of every kernel of the solve:
                               per-node passes (input masks, fold, cut start) at
                               their bytes per node, tile kernels at their bytes
                               per node of each visited tile */
  int inputs_fresh;         /* 1 if the inputs changed since the last solve (the
                               residual-arc masks were rebuilt in this solve) */
} jfc_stats;

/* ---- lifetime ------------------------------------------------------------ */

/* Fill *opt with the defaults. */
void jfc_default_options(jfc_options *opt);

/* Create a graph of one instance.  Replaces the buffer set-up of
 * QGCW:815-887 (initConfigurations) and QGCW:955-958 (devExcessFlow,
 * devHeight, devCapacity1/2).  dims is 2 or 3; connectivity 4 or 8 (2D) or 6
 * (3D); nz must be 1 for 2D. */
int jfc_create(jfc_graph **g, int dims, int nx, int ny, int nz, int connectivity);

/* Create a graph of `count` independent instances solved together (no
 * reference counterpart: the reference solves one image at a time).
 * shapes holds count rows of (nx, ny, nz, connectivity); nz == 1 means 2D. */
int jfc_create_batch(jfc_graph **g, int count, const int32_t *shapes);

/* Release every resource of the graph (the widget destructor's buffer
 * release, QGCWidget.h:103-131).  NULL is ignored. */
void jfc_destroy(jfc_graph *g);

int jfc_num_instances(const jfc_graph *g);

/* The stream (hipStream_t) every later call of the graph issues its work on;
 * NULL = the legacy default stream (the initial setting).  The reference's
 * single in-order command queue (QGCW:808) becomes the caller's stream, so a
 * torch or HIP pipeline that produces capacities on it needs no device-wide
 * synchronisation. */
int jfc_set_stream(jfc_graph *g, void *stream);

/* Thread-local message of the last error (QError / cl::Error text,
 * infrastructures/QError.h:37-74, QGCW:2533-2557). */
const char *jfc_last_error(void);

/* ---- inputs (QGCW:2631-2722 slotLoadGraphClicked, 2560-2606 slotModelClicked) */

/* Source / sink capacities of instance `inst` (N values each). */
int jfc_set_terminal_caps(jfc_graph *g, int inst, const int32_t *cap_source,
                          const int32_t *cap_sink, int on_device);

/* Terminal capacities in the reference's folded form: one signed int32 per
 * node, ef > 0 = source capacity, ef < 0 = sink capacity (the .ef file /
 * devExcessFlow, QGCW:2631-2722). */
int jfc_set_excess(jfc_graph *g, int inst, const int32_t *ef, int on_device);

/* Neighbour capacities of direction `arc` (N values). */
int jfc_set_neighbor_caps(jfc_graph *g, int inst, int arc, const int32_t *cap,
                          int on_device);

/* Hard seeds: fg[v] != 0 -> (source, sink) = (2^30, 0); bg[v] != 0 ->
 * (0, 2^30); applied on top of the terminal capacities already set (fg wins
 * where both are set).  Either pointer may be NULL.  Stands in for the
 * cl_cut foreground/background bytes and the seed rule of
 * graphcut_compute_excessflow.cl:71-74. */
int jfc_set_seeds(jfc_graph *g, int inst, const uint8_t *fg, const uint8_t *bg,
                  int on_device);

/* Zero-copy inputs: device pointers to instance `inst`'s stored input planes,
 * for callers whose own kernels produce capacities in place (the reference's
 * modeling kernels write devExcessFlow / devCapacity directly, QGCW:2560-2606).
 * Node (x, y, z) of the instance is element x + row_pitch * (y + ny * z) of
 * cap_source, cap_sink and of each neighbour plane d at
 * neighbor + d * plane_stride; row_pitch >= nx, and the elements of a row
 * past nx must stay 0.  After writing, call jfc_inputs_written: it marks the
 * instance's inputs as set and drops every cache derived from them (the
 * residual-arc masks), so the next solve starts from the new values.  The
 * library does not range-check values written this way: sources and sinks
 * must lie in [0, JFC_MAX_TERMINAL_CAP], neighbour capacities in
 * [0, JFC_MAX_NEIGHBOR_CAP], and arcs leaving the grid must be 0.  Any
 * output pointer may be NULL. */
int jfc_input_planes(jfc_graph *g, int inst, int32_t **cap_source, int32_t **cap_sink,
                     int32_t **neighbor, int64_t *row_pitch, int64_t *plane_stride);
int jfc_inputs_written(jfc_graph *g, int inst);

/* ---- solve (QGCW:2364-2558 slotCutClicked) --------------------------------- */

/* Run every instance to the max-flow certificate or opt->max_sweeps.
 * opt and stats may be NULL.  Returns JFC_OK, JFC_NOT_CONVERGED or an error. */
int jfc_solve(jfc_graph *g, const jfc_options *opt, jfc_stats *stats);

/* ---- outputs --------------------------------------------------------------- */

/* Max-flow value of one instance = min-cut energy (QGCW:2452-2520). */
int jfc_get_flow(jfc_graph *g, int inst, int64_t *flow);

/* Max-flow values of all instances (count values). */
int jfc_get_flows(jfc_graph *g, int64_t *flows);

/* Canonical labels: 1 = source side (reachable from the source in the
 * residual graph), 0 = sink side.  Equals cl_cut.object of the reference at
 * convergence (graphcut_compute_cut.cl:19-140). */
int jfc_get_labels(jfc_graph *g, int inst, uint8_t *labels, int on_device);

/* Tag post-process of the solved cut (graphcut_init_tag / graphcut_compute_tag,
 * cl/GraphCut/graphcut_compute_tag.cl:19-242; the cl_cut.tag byte):
 * tags[v] = 1 iff v is a foreground seed (fg[v] != 0) or a source-side node
 * joined to one through axis-adjacent source-side nodes (4-adjacency in 2D
 * whatever the graph's connectivity, 6 in 3D).  fg and tags hold the
 * instance's nodes; both host or both device (on_device). */
int jfc_get_tags(jfc_graph *g, int inst, const uint8_t *fg, uint8_t *tags,
                 int on_device);

/* Cut energy of the current labels on the input capacities, computed on the
 * device (QGCW:2452-2520); equals the flow once solved. */
int jfc_get_energy(jfc_graph *g, int inst, int64_t *energy);

/* Read back the input capacities of an instance, as set or as built by
 * jfc_model_image (the graph save of QGCW:2724-2736, slotSaveGraphClicked). */
int jfc_get_terminal_caps(jfc_graph *g, int inst, int32_t *cap_source,
                          int32_t *cap_sink, int on_device);
int jfc_get_neighbor_caps(jfc_graph *g, int inst, int arc, int32_t *cap,
                          int on_device);

/* ---- backward (no reference counterpart) --------------------------------- */

/* Exact subgradient of the min-cut energy E = sum_T cs + sum_S ct +
 * sum_{u in S, v in T} c(u->v) for one instance, as fp32:
 *   d_cap_source[v] = 1[v in T], d_cap_sink[v] = 1[v in S],
 *   d_neighbor[d*N + v] = 1[v in S and v+off(d) in T] (K planes).
 * Any output may be NULL.  Call after jfc_solve. */
int jfc_backward(jfc_graph *g, int inst, float *d_cap_source, float *d_cap_sink,
                 float *d_neighbor, int on_device);

/* The same for every instance of a graph in one launch: outputs use the
 * graph's concatenated node numbering (instance i starts after the nodes of
 * instances 0..i-1), d_neighbor holds Kmax planes of that many nodes (Kmax =
 * largest connectivity of the graph; planes beyond an instance's own
 * connectivity are left untouched). */
int jfc_backward_all(jfc_graph *g, float *d_cap_source, float *d_cap_sink,
                     float *d_neighbor, int on_device);

/* Total nodes of the graph (all instances) and the largest connectivity. */
int jfc_graph_size(const jfc_graph *g, int64_t *nodes, int32_t *kmax);

/* ---- tile-shape autotuning (no reference counterpart: the reference takes
 * the tile size from the UI, QGCW:826-835) ---------------------------------- */

/* Time `reps` solves of the graph (after one warm-up) with every tile shape of
 * its dimensionality (2D: 32x8, 64x4, 16x16; 3D: 8x8x4, 32x4x2), crossed with
 * every local-iteration cap of {2, 3, 4} unless opt->max_local_iters fixes
 * it, with initial_relabel 0 and 1 unless opt->initial_relabel fixes it and
 * with stall_certify 0 and 1 unless opt->stall_certify fixes it; then, on
 * the fastest of those, sparse_local_iters -1, 8 and 64 unless
 * opt->sparse_local_iters fixes it; all under the other options of opt (may
 * be NULL).  The fastest choice is what the graph's solves use when those
 * options are left at their defaults (JFC_TILE_DEFAULT, 0, -1, -1, 0).  The
 * choice is also stored for the graph's size
 * class (dimensionality, connectivities, instance count and mean instance
 * size, each rounded to a power of two): graphs of the class created later in
 * the process start with it.  Results are unchanged by either choice.
 * *chosen (may be NULL) receives the JFC_TILE_* value. */
int jfc_autotune(jfc_graph *g, const jfc_options *opt, int reps, int *chosen);

/* The JFC_TILE_* shape a solve with tile_shape = JFC_TILE_DEFAULT uses. */
int jfc_get_tile_shape(const jfc_graph *g, int *shape);

/* The local-iteration cap a solve with max_local_iters = 0 uses. */
int jfc_get_local_iters(const jfc_graph *g, int *max_local_iters);

/* The initial_relabel value (0 or 1) a solve with initial_relabel = -1 uses. */
int jfc_get_initial_relabel(const jfc_graph *g, int *initial_relabel);

/* Every tuned choice at once: writes tile_shape, max_local_iters,
 * initial_relabel, stall_certify and sparse_local_iters of *opt with the
 * values solves that leave them at their defaults use (other fields are
 * not touched). */
int jfc_get_tuning(const jfc_graph *g, jfc_options *opt);

/* ---- data term (QGCW:2560-2606 slotModelClicked, QGCW:1997-2107 kmeans) --- */

typedef struct {
  int k_fg, k_bg;        /* clusters per side (reference spinBoxKMeans, <= 16) */
  int kmeans_iters;      /* max Lloyd iterations (spinBoxKMeansIterations) */
  int64_t unary_scale;   /* unary = round(unary_scale * d2 / range^2) */
  int64_t range;         /* intensity range in the unary (255 for 8-bit) */
  double lambda_xy;      /* pairwise weight in x/y (and 2D diagonals / sqrt 2) */
  double lambda_z;       /* pairwise weight along z */
  int z_separate;        /* 1: z arcs use their own mean squared difference */
} jfc_model_params;

typedef struct {
  int iters_fg, iters_bg;      /* Lloyd iterations run per side */
  int64_t centroids_fg[16 * 3];/* fixed-point centroids (value * 256) */
  int64_t centroids_bg[16 * 3];
  int64_t seeds_fg, seeds_bg;  /* seeded node counts */
  int64_t pair_sum[2];         /* sum of squared neighbour differences per group */
  int64_t pair_count[2];       /* edges per group (0 = xy, 1 = z if separate) */
} jfc_model_stats;

/* Build the whole graph of instance `inst` from an image and seed map on the
 * device: integer k-means per side on the seeded pixels, unary terms from the
 * distance to the nearest centroid of the opposite model, hard seeds, and
 * pairwise weights round(lambda * exp(-D / (2 s2))) with s2 the mean squared
 * neighbour difference.  image: N * channels values of 1 or 2 bytes
 * (channels 1 or 3); seeds: N bytes, 1 = foreground, 2 = background. */
int jfc_model_image(jfc_graph *g, int inst, const void *image, int channels,
                    int bytes_per_channel, const uint8_t *seeds, int on_device,
                    const jfc_model_params *params, jfc_model_stats *stats);

/* jfc_model_image without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  image and
 * seeds must be device pointers (there is no staging buffer).  The scratch is
 * a static device array of JFC_MODEL_STATIC_BYTES bytes per GPU, held until
 * the process ends; it limits the seeded nodes (both sides together) to
 * JFC_MODEL_STATIC_MAX_SEEDS, beyond which the call returns JFC_ERR_INVALID
 * before writing any plane.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (from several threads or streams) race and
 * must be serialised by the caller.  Results are identical to
 * jfc_model_image's. */
#define JFC_MODEL_STATIC_MAX_SEEDS (1 << 22)
#define JFC_MODEL_STATIC_BYTES (666112 + 4 * JFC_MODEL_STATIC_MAX_SEEDS)
int jfc_model_image_static(jfc_graph *g, int inst, const void *image, int channels,
                           int bytes_per_channel, const uint8_t *seeds,
                           const jfc_model_params *params, jfc_model_stats *stats);

/* ---- the widget's own data terms (QGCW:2560-2606 with its "Modeling" and
 * "Feature" panels, QGCPanel.ui) ------------------------------------------- */

typedef struct {
  int lazy;              /* 1: lazy snapping (radioButtonLazySnapping), 0: BK */
  float weight[4];       /* feature channel weights (doubleSpinBoxWeightX..W),
                            normalised to unit length inside (glm::normalize) */
  float lazy_neighbor;   /* lambda of graphcut_compute_capacity_lazy */
  float lazy_terminal;   /* lambda of graphcut_compute_excessflow_lazy */
  float bk_neighbor;     /* sigma of graphcut_compute_capacity_bk */
  float bk_terminal;     /* lambda of graphcut_compute_excessflow_bk */
  int k;                 /* k-means clusters per side, 2..32 (spinBoxK) */
  int kmeans_iterations; /* 1..256 (spinBoxKMeansIterations) */
} jfc_widget_model_params;

typedef struct {
  int64_t seeds[2];          /* seeded nodes: [0] foreground, [1] background */
  int iterations[2];         /* k-means iterations run per side (lazy) */
  float centroids[2][32][4]; /* centroids the lazy unary used (weighted features) */
  uint32_t hist_sum[2];      /* seeded nodes counted by the BK histograms */
} jfc_widget_model_stats;

/* The panel's defaults (QGCPanel.ui): lazy, weights (0.2126, 0.7152, 0.0722, 0),
 * lambdas 1 / 1, sigma 0.002, BK terminal 0, k 8, 32 iterations. */
void jfc_default_widget_model_params(jfc_widget_model_params *p);

/* The "Model" button (slotModelClicked QGCW:2560-2606) on the device, for a
 * 4-connected 2D or 6-connected 3D instance.  rgba: N x 4 bytes, the
 * widget's CL_UNORM_INT8 feature image; fg, bg: N bytes, nonzero = seeded
 * (cl_cut bytes 0 and 1).
 * - pairwise: graphcut_compute_capacity_lazy / _bk (graphcut_compute_capacity.cl:21-254);
 * - lazy unary: k-means per side from the seeded levels (QGCW:1997-2107,
 *   graphcut_histogram.cl:41-60, graphcut_kmeans.cl:21-159), then
 *   graphcut_compute_excessflow_lazy (graphcut_compute_excessflow.cl:47-77);
 * - BK unary: 256-bin seeded histograms (graphcut_histogram.cl:19-39, QGCW:3168-3180),
 *   then graphcut_compute_excessflow_bk (graphcut_compute_excessflow.cl:19-45).
 * Seeds get ef = +-cl_max_flow (32767); ef > 0 becomes the source capacity and
 * ef < 0 the sink capacity of the node.  Arithmetic is float32 with every
 * operation IEEE-rounded in the reference's order (exp and log in double,
 * rounded once to float); a side with no seeds returns JFC_ERR_NO_SEEDS.
 * Where the reference is undefined or racy the library fixes one reading:
 * the BK unary clamps its histogram level to 0..255, non-finite unaries
 * become 0 (NaN) or +-2^30, k-means takes the highest seeded node index per
 * level as that level's representative, and centroid sums are exact before
 * the division. */
int jfc_model_widget(jfc_graph *g, int inst, const uint8_t *rgba, const uint8_t *fg,
                     const uint8_t *bg, int on_device, const jfc_widget_model_params *params,
                     jfc_widget_model_stats *stats);

/* ---- graph files (host only; no device work) ----------------------------- */

/* MFI max-flow instance, the format of the reference's CPU benchmarks
 * (benchmark/GridCut/mfi.h:4-77, mfi.cpp:13-193): "MFI", a packed header
 * (dimension, connectivity, width, height, depth, int64 maxflow,
 * symmetric_arcs, 4096 B metadata, active_count, terminal / neighbour
 * capacity types), source and sink planes, one plane per arc in MFI arc
 * order, and an active mask.  Capacities are read into / written from the
 * jfcut direction order (MFI LE GE EL EG LL GL LG GG -> 1 0 3 2 5 7 6 4 in 2D;
 * LEE GEE ELE EGE EEL EEG -> 1 0 3 2 5 4 in 3D).  Supported: 2D-4C, 2D-8C,
 * 3D-6C. */
int jfc_mfi_info(const char *path, int32_t *dimension, int32_t *connectivity,
                 int32_t *nx, int32_t *ny, int32_t *nz, int64_t *maxflow);
int jfc_mfi_read(const char *path, int32_t *cap_source, int32_t *cap_sink,
                 int32_t *caps);
/* Writes the smallest unsigned type that holds each capacity group; the
 * active mask is all ones. */
int jfc_mfi_write(const char *path, int nx, int ny, int nz, int connectivity,
                  const int32_t *cap_source, const int32_t *cap_sink,
                  const int32_t *caps, int64_t maxflow);

/* The reference's raw graph files <prefix>.ef / .c1 / .c2
 * (slotLoadGraphClicked / slotSaveGraphClicked, QGCW:2631-2736): int32
 * folded terminal capacities; 2D ushort2 c1 = (+x, +y), c2 = (-y, -x);
 * 3D uchar4 c1 = (+x, +y, +z, -), c2 = (-z, -y, -x, -).  caps holds the
 * 4 (2D) or 6 (3D) direction planes.  Writing rejects capacities that do not
 * fit the file's type. */
int jfc_raw_read(const char *prefix, int nx, int ny, int nz, int32_t *ef,
                 int32_t *caps);
int jfc_raw_write(const char *prefix, int nx, int ny, int nz, const int32_t *ef,
                  const int32_t *caps);

/* .cut files: 4 bytes per node (foreground, background, object, tag), the
 * cl_cut of QGCSetting.h:45-51 (QGCW:2971-2996).  Bytes are read as
 * 0 / non-zero; written as 0 / 127.  Any pointer may be NULL. */
int jfc_cut_read(const char *path, int64_t n, uint8_t *fg, uint8_t *bg,
                 uint8_t *object, uint8_t *tag);
int jfc_cut_write(const char *path, int64_t n, const uint8_t *fg,
                  const uint8_t *bg, const uint8_t *object, const uint8_t *tag);

/* .node files: one record per node of a grid padded up to whole tiles of
 * block_x x block_y x block_z (QGCW:2642-2717; records cl_node_2d / cl_node_3d
 * of QGCSetting.h:31-43).  2D record {int32 height, int32 excess, uint16
 * cap[4] = (+x, +y, -y, -x)}; 3D record {int32 height, int32 excess, uint8
 * cap[8] = (+x, +y, +z, -z, -y, -x, -, -)}.  height may be NULL on read and on
 * write (then written as 0).  Padding records are written as zeros. */
int jfc_node_read(const char *path, int nx, int ny, int nz, int block_x,
                  int block_y, int block_z, int32_t *ef, int32_t *height,
                  int32_t *caps);
int jfc_node_write(const char *path, int nx, int ny, int nz, int block_x,
                   int block_y, int block_z, const int32_t *ef,
                   const int32_t *height, const int32_t *caps);

/* The MFI export of the reference's graph-cut widget (saveMFI, QGCW:3339-3473):
 * folded excess ef -> (source, sink) = (max(ef, 0), max(-ef, 0)); 2D-4C with
 * uint16 or 3D-6C with uint8 neighbour capacities, uint32 terminals;
 * symmetric_arcs = 1; active_count = number of nodes with ef >= 0; active mask
 * = 1 foreground seed, 2 background seed, 3 object, 4 tag, else 0 (first match
 * in that order; any of fg/bg/object/tag may be NULL); maxflow as given. */
int jfc_mfi_export(const char *path, int nx, int ny, int nz, const int32_t *ef,
                   const int32_t *caps, const uint8_t *fg, const uint8_t *bg,
                   const uint8_t *object, const uint8_t *tag, int64_t maxflow);

/* ---- synthetic workloads (benchmark inputs; no reference counterpart) ---- */

/* 2D RGB image: fg ellipse at the centre (semi-axes 150x100 scaled by
 * min(W,H)/512, mean colour (200,80,60)) on background (40,90,160), nblobs
 * Gaussian colour blobs (sigma 20-60 px, amplitude +-40 per channel),
 * Gaussian noise sigma `noise`, clipped to 0..255; 5-px scribbles: 2 inside
 * the ellipse (seed 1) and 2 in the bands above / below it (seed 2). */
int jfc_synth_image(uint8_t *rgb, uint8_t *seeds, int W, int H, int nblobs,
                    double noise, uint64_t seed);

/* Video volume W x H x T RGB: the ellipse of jfc_synth_image translates
 * 4 px/frame along a random direction and rotates 1 degree/frame; noise
 * sigma `noise` per frame; scribbles on the first and last frame only. */
int jfc_synth_video(uint8_t *rgb, uint8_t *seeds, int W, int H, int T, int nblobs,
                    double noise, uint64_t seed);

/* uint16 volume: `nell` random ellipsoids (radii 20-120, intensity
 * 1200-3000) over 400, 3D Gaussian smoothing (sigma 2), noise sigma 80;
 * 3 seven-voxel-wide strokes in the largest ellipsoid (seed 1) and 3 in the
 * background (seed 2). */
int jfc_synth_volume(uint16_t *vol, uint8_t *seeds, int nx, int ny, int nz,
                     int nell, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
