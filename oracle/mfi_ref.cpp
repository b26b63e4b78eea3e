// mfi_ref.cpp -- C entry points around the reference's own MFI reader and
// writer (benchmark/GridCut/mfi.cpp, compiled from the reference tree into
// oracle/_ref/ by oracle/Makefile).  TEST INFRASTRUCTURE ONLY: the checker of
// libjfcut's MFI files, never the product.
//
// Planes cross this boundary as int32 in MFI arc order (mfi.h:36-53).
// mfi_ref_bk303 solves a file the way benchmark/GridCut/bench-BK303.cpp:39-46
// (2D-4C) and :87-96 (3D-6C) do: add_edge(v, v + e, cap[G..][v],
// cap[L..][v + e]) for every +x / +y / +z grid edge.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include "graph.h"
#include "mfi.h"

typedef Graph<int, int, long long> BKGraph;

namespace {

int64_t cap_at(const MFI *m, const void *plane, bool terminal, int64_t i) {
  const MFI::Type t = terminal ? m->type_terminal_cap : m->type_neighbor_cap;
  if (t == MFI::TYPE_UINT8) return ((const uint8_t *)plane)[i];
  if (t == MFI::TYPE_UINT16) return ((const uint16_t *)plane)[i];
  return ((const uint32_t *)plane)[i];
}

void *typed_copy(const int32_t *src, int64_t n, int type) {
  const int sz = type == 0 ? 1 : type == 1 ? 2 : 4;
  uint8_t *dst = (uint8_t *)std::malloc((size_t)n * sz);
  for (int64_t i = 0; i < n; ++i) {
    if (sz == 1) dst[i] = (uint8_t)src[i];
    else if (sz == 2) ((uint16_t *)dst)[i] = (uint16_t)src[i];
    else ((uint32_t *)dst)[i] = (uint32_t)src[i];
  }
  return dst;
}

}  // namespace

// hdr = {dimension, connectivity, width, height, depth, symmetric_arcs,
//        active_count, type_terminal_cap, type_neighbor_cap}
extern "C" int mfi_ref_read(const char *path, int32_t *hdr, int64_t *maxflow,
                            int32_t *cs, int32_t *ct, int32_t *caps) {
  MFI *m = mfi_read(path);
  if (!m) return -1;
  hdr[0] = m->dimension;
  hdr[1] = m->connectivity;
  hdr[2] = m->width;
  hdr[3] = m->height;
  hdr[4] = m->depth;
  hdr[5] = m->symmetric_arcs;
  hdr[6] = m->active_count;
  hdr[7] = m->type_terminal_cap;
  hdr[8] = m->type_neighbor_cap;
  *maxflow = m->maxflow;
  const int64_t n = (int64_t)m->width * m->height * (m->dimension == 3 ? m->depth : 1);
  if (cs)
    for (int64_t i = 0; i < n; ++i) {
      cs[i] = (int32_t)cap_at(m, m->cap_source, true, i);
      ct[i] = (int32_t)cap_at(m, m->cap_sink, true, i);
    }
  if (caps)
    for (int a = 0; a < m->connectivity; ++a)
      for (int64_t i = 0; i < n; ++i)
        caps[a * n + i] = (int32_t)cap_at(m, m->cap_neighbor[a], false, i);
  mfi_free(m);
  return 0;
}

extern "C" int mfi_ref_write(const char *path, int dim, int conn, int w, int h, int d,
                             int64_t maxflow, int type_t, int type_n, const int32_t *cs,
                             const int32_t *ct, const int32_t *caps) {
  MFI m;
  std::memset(&m, 0, sizeof m);
  m.dimension = dim;
  m.connectivity = conn;
  m.width = w;
  m.height = h;
  m.depth = d;
  m.maxflow = maxflow;
  m.symmetric_arcs = false;
  const int64_t n = (int64_t)w * h * d;
  m.active_count = (int)n;
  m.type_terminal_cap = (MFI::Type)type_t;
  m.type_neighbor_cap = (MFI::Type)type_n;
  m.active_mask = (unsigned char *)std::malloc((size_t)n);
  std::memset(m.active_mask, 1, (size_t)n);
  m.cap_source = typed_copy(cs, n, type_t);
  m.cap_sink = typed_copy(ct, n, type_t);
  void *planes[26];
  for (int a = 0; a < conn; ++a) planes[a] = typed_copy(caps + a * n, n, type_n);
  m.cap_neighbor = planes;
  const bool ok = mfi_write(&m, path);
  for (int a = 0; a < conn; ++a) std::free(planes[a]);
  std::free(m.cap_source);
  std::free(m.cap_sink);
  std::free(m.active_mask);
  return ok ? 0 : -1;
}

extern "C" int64_t mfi_ref_bk303(const char *path) {
  MFI *m = mfi_read(path);
  if (!m) return -1;
  const bool d3 = m->dimension == 3;
  if (!((!d3 && m->connectivity == 4) || (d3 && m->connectivity == 6))) {
    mfi_free(m);
    return -2;
  }
  const int w = m->width, h = m->height, dd = d3 ? m->depth : 1;
  const int64_t n = (int64_t)w * h * dd;
  BKGraph *g = new BKGraph((int)n, (int)(n * (d3 ? 3 : 2)));
  g->add_node((int)n);
  // (G arc, L arc, offset) for the +x, +y, +z edges
  const int arcs3[3][2] = {{MFI::ARC_GEE, MFI::ARC_LEE},
                           {MFI::ARC_EGE, MFI::ARC_ELE},
                           {MFI::ARC_EEG, MFI::ARC_EEL}};
  const int arcs2[2][2] = {{MFI::ARC_GE, MFI::ARC_LE}, {MFI::ARC_EG, MFI::ARC_EL}};
  for (int z = 0; z < dd; ++z)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const int64_t v = x + (int64_t)w * (y + (int64_t)h * z);
        g->add_tweights((int)v, (int)cap_at(m, m->cap_source, true, v),
                        (int)cap_at(m, m->cap_sink, true, v));
        const int64_t off[3] = {1, w, (int64_t)w * h};
        const bool inside[3] = {x < w - 1, y < h - 1, z < dd - 1};
        for (int e = 0; e < (d3 ? 3 : 2); ++e) {
          if (!inside[e]) continue;
          const int ga = d3 ? arcs3[e][0] : arcs2[e][0];
          const int la = d3 ? arcs3[e][1] : arcs2[e][1];
          g->add_edge((int)v, (int)(v + off[e]), (int)cap_at(m, m->cap_neighbor[ga], false, v),
                      (int)cap_at(m, m->cap_neighbor[la], false, v + off[e]));
        }
      }
  const int64_t flow = g->maxflow();
  delete g;
  mfi_free(m);
  return flow;
}
