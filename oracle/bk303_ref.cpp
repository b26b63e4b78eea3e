// bk303_ref.cpp -- C entry point around the reference's own CPU max-flow path
// (BK303, benchmark/GridCut/code/BK303), built into oracle/_ref/ by
// oracle/Makefile from the sources under /root/reference.  TEST/BASELINE
// INFRASTRUCTURE ONLY: the checker and the CPU baseline, never the product.
//
// Graph construction follows benchmark/GridCut/bench-BK303.cpp:39-46 (2D-4C)
// and :87-96 (3D-6C): add_tweights(cs, ct) per node and one add_edge per
// undirected grid edge with (forward, reverse) capacities.  8-connectivity
// adds the two diagonal edges the same way.  Labels use
// what_segment(i, SINK), so free nodes go to the sink side and the source set
// is the canonical one (reachable from the source in the residual graph).
#include <chrono>
#include <cstdint>
#include "graph.h"

typedef Graph<int, int, long long> BKGraph;

extern "C" int64_t bk303_grid_maxflow(int nx, int ny, int nz, int conn,
                                      const int32_t* cs, const int32_t* ct,
                                      const int32_t* caps, uint8_t* labels,
                                      double* t_init, double* t_maxflow,
                                      double* t_output) {
  static const int d2[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                               {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
  static const int d3[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                               {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  const int K = conn;
  if (!(conn == 4 || conn == 8 || conn == 6)) return -1;
  const int64_t N = (int64_t)nx * ny * nz;
  if (N <= 0 || N > 0x7fffffffLL) return -1;
  auto t0 = std::chrono::steady_clock::now();
  BKGraph* g = new BKGraph((int)N, (int)(N * (K / 2)));
  g->add_node((int)N);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t v = x + (int64_t)nx * (y + (int64_t)ny * z);
        g->add_tweights((int)v, cs[v], ct[v]);
        for (int d = 0; d < K; d += 2) {
          int dx, dy, dz;
          if (conn == 6) { dx = d3[d][0]; dy = d3[d][1]; dz = d3[d][2]; }
          else { dx = d2[d][0]; dy = d2[d][1]; dz = 0; }
          const int X = x + dx, Y = y + dy, Z = z + dz;
          if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
          const int64_t u = X + (int64_t)nx * (Y + (int64_t)ny * Z);
          g->add_edge((int)v, (int)u, caps[(int64_t)d * N + v],
                      caps[(int64_t)(d ^ 1) * N + u]);
        }
      }
  auto t1 = std::chrono::steady_clock::now();
  long long flow = g->maxflow();
  auto t2 = std::chrono::steady_clock::now();
  if (labels)
    for (int64_t v = 0; v < N; ++v)
      labels[v] = g->what_segment((int)v, BKGraph::SINK) == BKGraph::SOURCE;
  delete g;
  auto t3 = std::chrono::steady_clock::now();
  if (t_init) *t_init = std::chrono::duration<double>(t1 - t0).count();
  if (t_maxflow) *t_maxflow = std::chrono::duration<double>(t2 - t1).count();
  if (t_output) *t_output = std::chrono::duration<double>(t3 - t2).count();
  return (int64_t)flow;
}
