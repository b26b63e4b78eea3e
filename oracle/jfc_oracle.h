/*
 * jfc_oracle.h -- CPU oracle for the JF-Cut grid max-flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * It restates, in plain sequential C, what the reference computes on the hot
 * path (file:line citations are relative to the JF-Cut source tree):
 *
 *   orc_maxflow          exact max-flow value of the folded grid graph
 *                        (push-relabel semantics of
 *                        src/QVisualizer/cl/GraphCut/graphcut_push_relabel.cl:53-180:
 *                        active = excess > 0 and height finite; relabel =
 *                        min over residual arcs of (h_nb + 1); push when
 *                        h - h_nb == 1; sequential FIFO order, exact global
 *                        relabel as in graphcut_global_relabel.cl:19-124).
 *   orc_canonical_cut    nodes reachable from a node with positive excess in
 *                        the residual graph (graphcut_compute_cut.cl:19-140),
 *                        plus the convergence certificate (no such node has a
 *                        deficit, graphcut_compute_cut.cl:135-139).
 *   orc_tag              foreground-connected part of the object
 *                        (graphcut_compute_tag.cl:19-242).
 *   orc_energy           Sum_T cs + Sum_S ct + Sum_{u in S, v in T} c(u->v) on
 *                        the initial capacities (QGCWidget.cpp:2452-2520).
 *   orc_kmeans_*         integer k-means data term (restating the intent of
 *                        graphcut_kmeans.cl:21-89 + QGCWidget.cpp:1997-2107 with
 *                        exact integer centroids; see jfc_oracle.c).
 *   orc_unary / orc_pairwise   integer unary / pairwise builders.
 *
 * Parity pins (tests/test_oracle.py): the reference fixtures' published
 * max-flows and flower's canonical labels, BK303 built from the reference
 * sources (oracle/_ref), scipy's maximum_flow, brute force over all labelings
 * of tiny grids, and the weak-duality identity flow == energy(labels).
 */
#ifndef JFC_ORACLE_H
#define JFC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Direction d of a connectivity: 2D 4C uses d=0..3, 2D 8C d=0..7, 3D 6C d=0..5.
 * The opposite of direction d is d^1. */
int orc_ndirs(int conn);
void orc_dir(int conn, int d, int *dx, int *dy, int *dz);

/* Solve the folded max-flow problem.
 *   cs, ct   : N terminal capacities (>= 0)
 *   caps     : K planes of N arc capacities, caps[d*N + v] = c(v -> v+off(d));
 *              arcs leaving the grid are ignored.
 * Outputs (any may be NULL): final excess e (N), final residual planes (K*N),
 * canonical labels (N bytes, 1 = source side), certificate flag.
 * Returns the max-flow value (int64), or -1 on allocation failure. */
int64_t orc_maxflow(int nx, int ny, int nz, int conn,
                    const int32_t *cs, const int32_t *ct, const int32_t *caps,
                    int32_t *e_out, int32_t *res_out, uint8_t *labels_out,
                    int *certificate_ok);

/* Tag post-process: tag[v] = 1 iff v is a foreground seed or an object node
 * joined to one through axis-adjacent (4 in 2D, 6 in 3D) object nodes.
 * Returns 0, or -1 on allocation failure. */
int orc_tag(int nx, int ny, int nz, const uint8_t *fg, const uint8_t *object,
            uint8_t *tag);

/* Canonical cut of a residual state: labels[v] = 1 iff v is reachable from a
 * node with e > 0 along arcs with residual > 0.  Returns 1 iff no reached node
 * has e < 0 (the preflow is maximum). */
int orc_canonical_cut(int nx, int ny, int nz, int conn, const int32_t *e,
                      const int32_t *res, uint8_t *labels);

/* Cut energy of labels (1 = source side) on the initial capacities. */
int64_t orc_energy(int nx, int ny, int nz, int conn, const int32_t *cs,
                   const int32_t *ct, const int32_t *caps, const uint8_t *labels);

/* ---- integer data term ------------------------------------------------- */

/* Fixed-point fraction bits of k-means centroids. */
#define ORC_KM_FRAC 8

/* Initial centroids: key = sum of channels (8-bit data) or value >> 4 (16-bit
 * single-channel data); the distinct keys present among the n samples are
 * listed in increasing order, key j*(m-1)/(K-1) (j = 0..K-1) is picked, and
 * the centroid is the rounded mean of the samples carrying that key. */
void orc_kmeans_init(const int32_t *feat, int n, int C, int wide, int K,
                     int64_t *centroids);

/* Lloyd iterations from the given centroids; stops when no centroid changes
 * or after max_iters.  Returns the number of iterations run. */
int orc_kmeans_iterate(const int32_t *feat, int n, int C, int K, int max_iters,
                       int64_t *centroids);

/* Unary terms: cs = round(100 * d2_bg / R^2), ct = round(100 * d2_fg / R^2)
 * with d2 the squared distance to the nearest centroid of that model; seeds
 * (1 = fg, 2 = bg) get (2^30, 0) / (0, 2^30). */
void orc_unary(const int32_t *feat, int64_t N, int C, const int64_t *cfg,
               int Kf, const int64_t *cbg, int Kb, int64_t R,
               const uint8_t *seeds, int32_t *cs, int32_t *ct);

/* Pairwise thresholds: T[k-1] = floor((2S/P) * ln(lambda/(k-0.5))) for
 * k = 1..round(lambda); returns the count.  w(D) = #{k : D <= T[k-1]}. */
int orc_pair_thresholds(double lambda, int64_t S, int64_t P, int64_t *T);

/* Pairwise capacities of every arc (both directions equal), 0 on arcs leaving
 * the grid.  Squared colour difference D = sum_c (I_u,c - I_v,c)^2.  Arcs in
 * the z direction use lambda_z and, if z_separate, their own mean S/P.  2D
 * diagonal arcs use lambda / sqrt(2).  Writes S and P used. */
void orc_pairwise(int nx, int ny, int nz, int conn, const int32_t *feat, int C,
                  double lambda_xy, double lambda_z, int z_separate,
                  int32_t *caps);

#ifdef __cplusplus
}
#endif
#endif
