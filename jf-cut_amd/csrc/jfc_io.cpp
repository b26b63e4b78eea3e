// jfc_io.cpp -- graph files: MFI instances and the reference's raw
// .ef/.c1/.c2/.cut files.  Host code; see include/jfcut.h for the formats.
//
// The MFI layout is the one of benchmark/GridCut/mfi.cpp:13-193 (kept alike
// because it is an exchange format); the code is this library's own.
#include "jfcut.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

// MFI arc index -> jfcut direction index
const int kMfi2D[8] = {1, 0, 3, 2, 5, 7, 6, 4};
const int kMfi3D[6] = {1, 0, 3, 2, 5, 4};

#pragma pack(push, 1)
struct MfiHeader {
  int32_t dimension, connectivity, width, height, depth;
  int64_t maxflow;
  int32_t symmetric_arcs;
  int8_t metadata[4096];
  int32_t active_count;
  int32_t type_terminal_cap, type_neighbor_cap;
};
#pragma pack(pop)
static_assert(sizeof(MfiHeader) == 4140, "packed MFI header");

struct File {
  FILE *f;
  explicit File(const char *p, const char *mode) : f(std::fopen(p, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

int type_size(int32_t t) { return t == 0 ? 1 : t == 1 ? 2 : t == 2 ? 4 : 0; }

bool read_plane(FILE *f, int32_t type, int64_t n, int32_t *out) {
  const int sz = type_size(type);
  std::vector<uint8_t> buf((size_t)n * sz);
  if (std::fread(buf.data(), sz, (size_t)n, f) != (size_t)n) return false;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t v;
    if (sz == 1) v = buf[i];
    else if (sz == 2) { uint16_t x; std::memcpy(&x, &buf[2 * i], 2); v = x; }
    else std::memcpy(&v, &buf[4 * i], 4);
    if (v > 0x7fffffffu) return false;
    if (out) out[i] = (int32_t)v;
  }
  return true;
}

int32_t fit_type(const int32_t *const *planes, int count, int64_t n) {
  int32_t mx = 0;
  for (int p = 0; p < count; ++p)
    for (int64_t i = 0; i < n; ++i) mx = std::max(mx, planes[p][i]);
  return mx < 256 ? 0 : mx < 65536 ? 1 : 2;
}

bool write_plane(FILE *f, int32_t type, int64_t n, const int32_t *in) {
  const int sz = type_size(type);
  std::vector<uint8_t> buf((size_t)n * sz);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t v = (uint32_t)in[i];
    if (sz == 1) buf[i] = (uint8_t)v;
    else if (sz == 2) { const uint16_t x = (uint16_t)v; std::memcpy(&buf[2 * i], &x, 2); }
    else std::memcpy(&buf[4 * i], &v, 4);
  }
  return std::fwrite(buf.data(), sz, (size_t)n, f) == (size_t)n;
}

bool valid_conn(int dim, int conn) {
  return (dim == 2 && (conn == 4 || conn == 8)) || (dim == 3 && conn == 6);
}

bool read_header(FILE *f, MfiHeader &h) {
  char magic[3];
  if (std::fread(magic, 1, 3, f) != 3 || std::memcmp(magic, "MFI", 3) != 0) return false;
  if (std::fread(&h, sizeof h, 1, f) != 1) return false;
  if (h.dimension == 2) h.depth = 1;
  return valid_conn(h.dimension, h.connectivity) && h.width > 0 && h.height > 0 &&
         h.depth > 0 && type_size(h.type_terminal_cap) && type_size(h.type_neighbor_cap);
}

}  // namespace

extern "C" int jfc_mfi_info(const char *path, int32_t *dimension, int32_t *connectivity,
                            int32_t *nx, int32_t *ny, int32_t *nz, int64_t *maxflow) {
  if (!path) return JFC_ERR_INVALID;
  File F(path, "rb");
  MfiHeader h;
  if (!F.f || !read_header(F.f, h)) return JFC_ERR_INVALID;
  if (dimension) *dimension = h.dimension;
  if (connectivity) *connectivity = h.connectivity;
  if (nx) *nx = h.width;
  if (ny) *ny = h.height;
  if (nz) *nz = h.depth;
  if (maxflow) *maxflow = h.maxflow;
  return JFC_OK;
}

extern "C" int jfc_mfi_read(const char *path, int32_t *cap_source, int32_t *cap_sink,
                            int32_t *caps) {
  if (!path) return JFC_ERR_INVALID;
  File F(path, "rb");
  MfiHeader h;
  if (!F.f || !read_header(F.f, h)) return JFC_ERR_INVALID;
  const int64_t n = (int64_t)h.width * h.height * h.depth;
  if (!read_plane(F.f, h.type_terminal_cap, n, cap_source) ||
      !read_plane(F.f, h.type_terminal_cap, n, cap_sink))
    return JFC_ERR_INVALID;
  const int *map = h.dimension == 3 ? kMfi3D : kMfi2D;
  for (int a = 0; a < h.connectivity; ++a)
    if (!read_plane(F.f, h.type_neighbor_cap, n, caps ? caps + (int64_t)map[a] * n : nullptr))
      return JFC_ERR_INVALID;
  return JFC_OK;
}

extern "C" int jfc_mfi_write(const char *path, int nx, int ny, int nz, int connectivity,
                             const int32_t *cap_source, const int32_t *cap_sink,
                             const int32_t *caps, int64_t maxflow) {
  const int dim = nz > 1 ? 3 : 2;
  if (!path || !cap_source || !cap_sink || !caps || nx <= 0 || ny <= 0 || nz <= 0 ||
      !valid_conn(dim, connectivity))
    return JFC_ERR_INVALID;
  const int64_t n = (int64_t)nx * ny * nz;
  for (const int32_t *p : {cap_source, cap_sink})
    for (int64_t i = 0; i < n; ++i)
      if (p[i] < 0) return JFC_ERR_INVALID;
  for (int64_t i = 0; i < n * connectivity; ++i)
    if (caps[i] < 0) return JFC_ERR_INVALID;
  File F(path, "wb");
  if (!F.f) return JFC_ERR_INVALID;
  MfiHeader h;
  std::memset(&h, 0, sizeof h);
  h.dimension = dim;
  h.connectivity = connectivity;
  h.width = nx;
  h.height = ny;
  h.depth = nz;
  h.maxflow = maxflow;
  h.symmetric_arcs = 0;
  h.active_count = (int32_t)std::min<int64_t>(n, 0x7fffffff);
  const int32_t *term[2] = {cap_source, cap_sink};
  h.type_terminal_cap = fit_type(term, 2, n);
  std::vector<const int32_t *> nb(connectivity);
  for (int d = 0; d < connectivity; ++d) nb[d] = caps + (int64_t)d * n;
  h.type_neighbor_cap = fit_type(nb.data(), connectivity, n);
  if (std::fwrite("MFI", 1, 3, F.f) != 3 || std::fwrite(&h, sizeof h, 1, F.f) != 1)
    return JFC_ERR_INVALID;
  if (!write_plane(F.f, h.type_terminal_cap, n, cap_source) ||
      !write_plane(F.f, h.type_terminal_cap, n, cap_sink))
    return JFC_ERR_INVALID;
  const int *map = dim == 3 ? kMfi3D : kMfi2D;
  for (int a = 0; a < connectivity; ++a)
    if (!write_plane(F.f, h.type_neighbor_cap, n, caps + (int64_t)map[a] * n))
      return JFC_ERR_INVALID;
  std::vector<uint8_t> active((size_t)n, 1);
  if (std::fwrite(active.data(), 1, (size_t)n, F.f) != (size_t)n) return JFC_ERR_INVALID;
  return JFC_OK;
}

namespace {

// raw c1/c2 component -> direction plane
//   2D ushort2: c1 = (+x, +y), c2 = (-y, -x)
//   3D uchar4 : c1 = (+x, +y, +z, -), c2 = (-z, -y, -x, -)
const int kRaw2D[2][2] = {{0, 2}, {3, 1}};
const int kRaw3D[2][3] = {{0, 2, 4}, {5, 3, 1}};

std::string cat(const char *a, const char *b) { return std::string(a) + b; }

}  // namespace

extern "C" int jfc_raw_read(const char *prefix, int nx, int ny, int nz, int32_t *ef,
                            int32_t *caps) {
  if (!prefix || nx <= 0 || ny <= 0 || nz <= 0) return JFC_ERR_INVALID;
  const int64_t n = (int64_t)nx * ny * nz;
  const bool d3 = nz > 1;
  if (ef) {
    File F(cat(prefix, ".ef").c_str(), "rb");
    if (!F.f || std::fread(ef, 4, (size_t)n, F.f) != (size_t)n) return JFC_ERR_INVALID;
  }
  if (!caps) return JFC_OK;
  for (int k = 0; k < 2; ++k) {
    File F(cat(prefix, k ? ".c2" : ".c1").c_str(), "rb");
    if (!F.f) return JFC_ERR_INVALID;
    if (!d3) {
      std::vector<uint16_t> b((size_t)n * 2);
      if (std::fread(b.data(), 2, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
      for (int c = 0; c < 2; ++c)
        for (int64_t i = 0; i < n; ++i) caps[(int64_t)kRaw2D[k][c] * n + i] = b[2 * i + c];
    } else {
      std::vector<uint8_t> b((size_t)n * 4);
      if (std::fread(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
      for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < n; ++i) caps[(int64_t)kRaw3D[k][c] * n + i] = b[4 * i + c];
    }
  }
  return JFC_OK;
}

extern "C" int jfc_raw_write(const char *prefix, int nx, int ny, int nz, const int32_t *ef,
                             const int32_t *caps) {
  if (!prefix || !ef || !caps || nx <= 0 || ny <= 0 || nz <= 0) return JFC_ERR_INVALID;
  const int64_t n = (int64_t)nx * ny * nz;
  const bool d3 = nz > 1;
  const int K = d3 ? 6 : 4;
  const int32_t hi = d3 ? 255 : 65535;
  for (int64_t i = 0; i < n * K; ++i)
    if (caps[i] < 0 || caps[i] > hi) return JFC_ERR_OVERFLOW;
  {
    File F(cat(prefix, ".ef").c_str(), "wb");
    if (!F.f || std::fwrite(ef, 4, (size_t)n, F.f) != (size_t)n) return JFC_ERR_INVALID;
  }
  for (int k = 0; k < 2; ++k) {
    File F(cat(prefix, k ? ".c2" : ".c1").c_str(), "wb");
    if (!F.f) return JFC_ERR_INVALID;
    if (!d3) {
      std::vector<uint16_t> b((size_t)n * 2);
      for (int c = 0; c < 2; ++c)
        for (int64_t i = 0; i < n; ++i) b[2 * i + c] = (uint16_t)caps[(int64_t)kRaw2D[k][c] * n + i];
      if (std::fwrite(b.data(), 2, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
    } else {
      std::vector<uint8_t> b((size_t)n * 4, 0);
      for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < n; ++i) b[4 * i + c] = (uint8_t)caps[(int64_t)kRaw3D[k][c] * n + i];
      if (std::fwrite(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
    }
  }
  return JFC_OK;
}

extern "C" int jfc_cut_read(const char *path, int64_t n, uint8_t *fg, uint8_t *bg,
                            uint8_t *object) {
  if (!path || n <= 0) return JFC_ERR_INVALID;
  File F(path, "rb");
  std::vector<uint8_t> b((size_t)n * 4);
  if (!F.f || std::fread(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
  for (int64_t i = 0; i < n; ++i) {
    if (fg) fg[i] = b[4 * i] != 0;
    if (bg) bg[i] = b[4 * i + 1] != 0;
    if (object) object[i] = b[4 * i + 2] != 0;
  }
  return JFC_OK;
}

extern "C" int jfc_cut_write(const char *path, int64_t n, const uint8_t *fg,
                             const uint8_t *bg, const uint8_t *object) {
  if (!path || n <= 0) return JFC_ERR_INVALID;
  std::vector<uint8_t> b((size_t)n * 4, 0);
  for (int64_t i = 0; i < n; ++i) {
    b[4 * i] = fg && fg[i] ? 127 : 0;
    b[4 * i + 1] = bg && bg[i] ? 127 : 0;
    b[4 * i + 2] = object && object[i] ? 127 : 0;
  }
  File F(path, "wb");
  if (!F.f || std::fwrite(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
  return JFC_OK;
}
