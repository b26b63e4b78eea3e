/*
 * jfcut_debug.h -- probes of libjfcut's internal tile machinery, for tests.
 *
 * They expose steps the reference performs inside QGCWidget::computeBFS
 * (src/QVisualizer/Graph Cut/QGCWidget.cpp:3093-3127): jump flooding over the
 * tile grid (graphcut_bfs_block.cl:43-63), the maximum tile distance
 * (reduce_histogram, bin/x64/cl/Infrastructures/reduce.cl:56-141) and the
 * counting sort of tiles by distance (graphcut_bfs_histogram / clogs scan /
 * graphcut_bfs_mapping, graphcut_bfs_block.cl:65-91), so that they can be
 * checked against their closed forms.
 */
#ifndef JFCUT_DEBUG_H
#define JFCUT_DEBUG_H

#include "jfcut.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tile grid of the graph's current tile shape: total tiles, tile size and,
 * for instance `inst`, the tile counts along x, y, z and its first tile id. */
int jfc_debug_tiles(jfc_graph *g, int inst, int32_t *ntiles, int32_t *tile_dims /*3*/,
                    int32_t *grid_dims /*3*/, int32_t *tile_off);

/* Run the device tile ordering on the given per-tile seed flags (ntiles bytes,
 * host): tile_dist[t] = taxicab distance in tiles to the nearest seed tile of
 * the same instance (or the last bin if the instance has none); order = all
 * tiles sorted by that distance (ties in any order); *max_dist = largest
 * finite distance.  Host outputs of ntiles entries.  method: the
 * jfc_options.tile_order value (0 distance transform, 1 jump flooding). */
int jfc_debug_tile_order(jfc_graph *g, int method, const uint8_t *tile_seed,
                         int32_t *tile_dist, int32_t *order, int32_t *max_dist);

/* Per-tile flags computed by the device from the current input capacities
 * (folded e = cs - ct): which = 0 -> tiles holding a deficit node (relabel
 * seeds, graphcut_init_global_relabel), 1 -> tiles holding an excess node
 * (cut seeds, graphcut_init_cut), 2 -> active tiles at zero heights
 * (graphcut_count_node after the clear of h).  ntiles bytes, host output. */
int jfc_debug_tile_flags(jfc_graph *g, int which, uint8_t *flags);

/* Active tiles visited by each push-relabel sweep of the last solve (all
 * colour classes of the sweep together); returns the number of sweeps, writes
 * at most n counts. */
int jfc_debug_sweep_counts(jfc_graph *g, int32_t *counts, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
