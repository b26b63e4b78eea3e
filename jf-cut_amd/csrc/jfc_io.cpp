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

namespace jfc {
void set_last_error(const std::string &msg);

// Records a message for a failed graph-file call.
int io_result(int rc, const char *fn, const char *path) {
  if (rc < 0)
    set_last_error(std::string(fn) + (rc == JFC_ERR_OVERFLOW ? ": a capacity does not fit the file type: "
                                                               : ": invalid argument, unreadable file or malformed content: ") +
                   (path ? path : "(null)"));
  return rc;
}
}  // namespace jfc

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

static int mfi_info_impl(const char *path, int32_t *dimension, int32_t *connectivity,
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

static int mfi_read_impl(const char *path, int32_t *cap_source, int32_t *cap_sink,
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

namespace {

// Writes an MFI file from planes in jfcut direction order.  Header fields other
// than the grid and connectivity come from `h`.
int write_mfi_file(const char *path, MfiHeader h, const int32_t *cs, const int32_t *ct,
                   const int32_t *caps, const uint8_t *active) {
  const int64_t n = (int64_t)h.width * h.height * h.depth;
  File F(path, "wb");
  if (!F.f) return JFC_ERR_INVALID;
  if (std::fwrite("MFI", 1, 3, F.f) != 3 || std::fwrite(&h, sizeof h, 1, F.f) != 1)
    return JFC_ERR_INVALID;
  if (!write_plane(F.f, h.type_terminal_cap, n, cs) ||
      !write_plane(F.f, h.type_terminal_cap, n, ct))
    return JFC_ERR_INVALID;
  const int *map = h.dimension == 3 ? kMfi3D : kMfi2D;
  for (int a = 0; a < h.connectivity; ++a)
    if (!write_plane(F.f, h.type_neighbor_cap, n, caps + (int64_t)map[a] * n))
      return JFC_ERR_INVALID;
  if (std::fwrite(active, 1, (size_t)n, F.f) != (size_t)n) return JFC_ERR_INVALID;
  return JFC_OK;
}

MfiHeader grid_header(int nx, int ny, int nz, int conn, int64_t maxflow) {
  MfiHeader h;
  std::memset(&h, 0, sizeof h);
  h.dimension = nz > 1 ? 3 : 2;
  h.connectivity = conn;
  h.width = nx;
  h.height = ny;
  h.depth = nz;
  h.maxflow = maxflow;
  return h;
}

bool fits(const int32_t *p, int64_t n, int32_t hi) {
  for (int64_t i = 0; i < n; ++i)
    if (p[i] < 0 || p[i] > hi) return false;
  return true;
}

}  // namespace

static int mfi_write_impl(const char *path, int nx, int ny, int nz, int connectivity,
                             const int32_t *cap_source, const int32_t *cap_sink,
                             const int32_t *caps, int64_t maxflow) {
  const int dim = nz > 1 ? 3 : 2;
  if (!path || !cap_source || !cap_sink || !caps || nx <= 0 || ny <= 0 || nz <= 0 ||
      !valid_conn(dim, connectivity))
    return JFC_ERR_INVALID;
  const int64_t n = (int64_t)nx * ny * nz;
  if (!fits(cap_source, n, 0x7fffffff) || !fits(cap_sink, n, 0x7fffffff) ||
      !fits(caps, n * connectivity, 0x7fffffff))
    return JFC_ERR_INVALID;
  MfiHeader h = grid_header(nx, ny, nz, connectivity, maxflow);
  h.symmetric_arcs = 0;
  h.active_count = (int32_t)std::min<int64_t>(n, 0x7fffffff);
  const int32_t *term[2] = {cap_source, cap_sink};
  h.type_terminal_cap = fit_type(term, 2, n);
  std::vector<const int32_t *> nb(connectivity);
  for (int d = 0; d < connectivity; ++d) nb[d] = caps + (int64_t)d * n;
  h.type_neighbor_cap = fit_type(nb.data(), connectivity, n);
  std::vector<uint8_t> active((size_t)n, 1);
  return write_mfi_file(path, h, cap_source, cap_sink, caps, active.data());
}

static int mfi_export_impl(const char *path, int nx, int ny, int nz, const int32_t *ef,
                              const int32_t *caps, const uint8_t *fg, const uint8_t *bg,
                              const uint8_t *object, const uint8_t *tag, int64_t maxflow) {
  if (!path || !ef || !caps || nx <= 0 || ny <= 0 || nz <= 0) return JFC_ERR_INVALID;
  const bool d3 = nz > 1;
  const int K = d3 ? 6 : 4;
  const int64_t n = (int64_t)nx * ny * nz;
  if (!fits(caps, n * K, d3 ? 255 : 65535)) return JFC_ERR_OVERFLOW;
  std::vector<int32_t> cs((size_t)n), ct((size_t)n);
  std::vector<uint8_t> active((size_t)n);
  int64_t nactive = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (ef[i] == INT32_MIN) return JFC_ERR_OVERFLOW;
    cs[i] = ef[i] >= 0 ? ef[i] : 0;
    ct[i] = ef[i] >= 0 ? 0 : -ef[i];
    nactive += ef[i] >= 0;
    active[i] = fg && fg[i] ? 1 : bg && bg[i] ? 2 : object && object[i] ? 3 : tag && tag[i] ? 4 : 0;
  }
  MfiHeader h = grid_header(nx, ny, nz, K, maxflow);
  h.symmetric_arcs = 1;
  h.active_count = (int32_t)std::min<int64_t>(nactive, 0x7fffffff);
  h.type_terminal_cap = 2;
  h.type_neighbor_cap = d3 ? 0 : 1;
  return write_mfi_file(path, h, cs.data(), ct.data(), caps, active.data());
}

namespace {

// raw c1/c2 component -> direction plane
//   2D ushort2: c1 = (+x, +y), c2 = (-y, -x)
//   3D uchar4 : c1 = (+x, +y, +z, -), c2 = (-z, -y, -x, -)
const int kRaw2D[2][2] = {{0, 2}, {3, 1}};
const int kRaw3D[2][3] = {{0, 2, 4}, {5, 3, 1}};

std::string cat(const char *a, const char *b) { return std::string(a) + b; }

}  // namespace

static int raw_read_impl(const char *prefix, int nx, int ny, int nz, int32_t *ef,
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

static int raw_write_impl(const char *prefix, int nx, int ny, int nz, const int32_t *ef,
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

static int cut_read_impl(const char *path, int64_t n, uint8_t *fg, uint8_t *bg,
                            uint8_t *object, uint8_t *tag) {
  if (!path || n <= 0) return JFC_ERR_INVALID;
  File F(path, "rb");
  std::vector<uint8_t> b((size_t)n * 4);
  if (!F.f || std::fread(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
  for (int64_t i = 0; i < n; ++i) {
    if (fg) fg[i] = b[4 * i] != 0;
    if (bg) bg[i] = b[4 * i + 1] != 0;
    if (object) object[i] = b[4 * i + 2] != 0;
    if (tag) tag[i] = b[4 * i + 3] != 0;
  }
  return JFC_OK;
}

static int cut_write_impl(const char *path, int64_t n, const uint8_t *fg,
                             const uint8_t *bg, const uint8_t *object, const uint8_t *tag) {
  if (!path || n <= 0) return JFC_ERR_INVALID;
  std::vector<uint8_t> b((size_t)n * 4, 0);
  for (int64_t i = 0; i < n; ++i) {
    b[4 * i] = fg && fg[i] ? 127 : 0;
    b[4 * i + 1] = bg && bg[i] ? 127 : 0;
    b[4 * i + 2] = object && object[i] ? 127 : 0;
    b[4 * i + 3] = tag && tag[i] ? 127 : 0;
  }
  File F(path, "wb");
  if (!F.f || std::fwrite(b.data(), 1, b.size(), F.f) != b.size()) return JFC_ERR_INVALID;
  return JFC_OK;
}

namespace {

// .node record payload -> direction plane
//   2D uint16 cap[4] = (+x, +y, -y, -x); 3D uint8 cap[8] = (+x, +y, +z, -z, -y, -x, -, -)
const int kNode2D[4] = {0, 2, 3, 1};
const int kNode3D[6] = {0, 2, 4, 5, 3, 1};

int64_t round_up(int v, int b) { return ((int64_t)v + b - 1) / b * b; }

bool node_args(int nx, int ny, int nz, int bx, int by, int bz) {
  return nx > 0 && ny > 0 && nz > 0 && bx > 0 && by > 0 && bz > 0 && (nz > 1 || bz == 1);
}

}  // namespace

static int node_read_impl(const char *path, int nx, int ny, int nz, int block_x,
                             int block_y, int block_z, int32_t *ef, int32_t *height,
                             int32_t *caps) {
  if (!path || !node_args(nx, ny, nz, block_x, block_y, block_z)) return JFC_ERR_INVALID;
  const bool d3 = nz > 1;
  const int64_t px = round_up(nx, block_x), py = round_up(ny, block_y),
                pz = d3 ? round_up(nz, block_z) : 1;
  const int64_t n = (int64_t)nx * ny * nz, records = px * py * pz;
  std::vector<uint8_t> buf((size_t)records * 16);
  {
    File F(path, "rb");
    if (!F.f || std::fread(buf.data(), 16, (size_t)records, F.f) != (size_t)records)
      return JFC_ERR_INVALID;
  }
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t v = x + (int64_t)nx * (y + (int64_t)ny * z);
        const uint8_t *r = &buf[(size_t)(x + px * (y + py * z)) * 16];
        int32_t hv, ev;
        std::memcpy(&hv, r, 4);
        std::memcpy(&ev, r + 4, 4);
        if (height) height[v] = hv;
        if (ef) ef[v] = ev;
        if (!caps) continue;
        if (!d3) {
          for (int c = 0; c < 4; ++c) {
            uint16_t w;
            std::memcpy(&w, r + 8 + 2 * c, 2);
            caps[(int64_t)kNode2D[c] * n + v] = w;
          }
        } else {
          for (int c = 0; c < 6; ++c) caps[(int64_t)kNode3D[c] * n + v] = r[8 + c];
        }
      }
  return JFC_OK;
}

static int node_write_impl(const char *path, int nx, int ny, int nz, int block_x,
                              int block_y, int block_z, const int32_t *ef,
                              const int32_t *height, const int32_t *caps) {
  if (!path || !ef || !caps || !node_args(nx, ny, nz, block_x, block_y, block_z))
    return JFC_ERR_INVALID;
  const bool d3 = nz > 1;
  const int64_t px = round_up(nx, block_x), py = round_up(ny, block_y),
                pz = d3 ? round_up(nz, block_z) : 1;
  const int64_t n = (int64_t)nx * ny * nz, records = px * py * pz;
  if (!fits(caps, n * (d3 ? 6 : 4), d3 ? 255 : 65535)) return JFC_ERR_OVERFLOW;
  std::vector<uint8_t> buf((size_t)records * 16, 0);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t v = x + (int64_t)nx * (y + (int64_t)ny * z);
        uint8_t *r = &buf[(size_t)(x + px * (y + py * z)) * 16];
        const int32_t hv = height ? height[v] : 0;
        std::memcpy(r, &hv, 4);
        std::memcpy(r + 4, &ef[v], 4);
        if (!d3) {
          for (int c = 0; c < 4; ++c) {
            const uint16_t w = (uint16_t)caps[(int64_t)kNode2D[c] * n + v];
            std::memcpy(r + 8 + 2 * c, &w, 2);
          }
        } else {
          for (int c = 0; c < 6; ++c) r[8 + c] = (uint8_t)caps[(int64_t)kNode3D[c] * n + v];
        }
      }
  File F(path, "wb");
  if (!F.f || std::fwrite(buf.data(), 16, (size_t)records, F.f) != (size_t)records)
    return JFC_ERR_INVALID;
  return JFC_OK;
}

// ---- C-ABI entry points: every failure leaves a message for jfc_last_error ----

extern "C" int jfc_mfi_info(const char *path, int32_t *dimension, int32_t *connectivity,
                            int32_t *nx, int32_t *ny, int32_t *nz, int64_t *maxflow) {
  return jfc::io_result(mfi_info_impl(path, dimension, connectivity, nx, ny, nz, maxflow), "jfc_mfi_info", path);
}

extern "C" int jfc_mfi_read(const char *path, int32_t *cap_source, int32_t *cap_sink,
                            int32_t *caps) {
  return jfc::io_result(mfi_read_impl(path, cap_source, cap_sink, caps), "jfc_mfi_read", path);
}

extern "C" int jfc_mfi_write(const char *path, int nx, int ny, int nz, int connectivity,
                             const int32_t *cap_source, const int32_t *cap_sink,
                             const int32_t *caps, int64_t maxflow) {
  return jfc::io_result(mfi_write_impl(path, nx, ny, nz, connectivity, cap_source, cap_sink, caps, maxflow), "jfc_mfi_write", path);
}

extern "C" int jfc_mfi_export(const char *path, int nx, int ny, int nz, const int32_t *ef,
                              const int32_t *caps, const uint8_t *fg, const uint8_t *bg,
                              const uint8_t *object, const uint8_t *tag, int64_t maxflow) {
  return jfc::io_result(mfi_export_impl(path, nx, ny, nz, ef, caps, fg, bg, object, tag, maxflow), "jfc_mfi_export", path);
}

extern "C" int jfc_raw_read(const char *prefix, int nx, int ny, int nz, int32_t *ef,
                            int32_t *caps) {
  return jfc::io_result(raw_read_impl(prefix, nx, ny, nz, ef, caps), "jfc_raw_read", prefix);
}

extern "C" int jfc_raw_write(const char *prefix, int nx, int ny, int nz, const int32_t *ef,
                             const int32_t *caps) {
  return jfc::io_result(raw_write_impl(prefix, nx, ny, nz, ef, caps), "jfc_raw_write", prefix);
}

extern "C" int jfc_cut_read(const char *path, int64_t n, uint8_t *fg, uint8_t *bg,
                            uint8_t *object, uint8_t *tag) {
  return jfc::io_result(cut_read_impl(path, n, fg, bg, object, tag), "jfc_cut_read", path);
}

extern "C" int jfc_cut_write(const char *path, int64_t n, const uint8_t *fg,
                             const uint8_t *bg, const uint8_t *object, const uint8_t *tag) {
  return jfc::io_result(cut_write_impl(path, n, fg, bg, object, tag), "jfc_cut_write", path);
}

extern "C" int jfc_node_read(const char *path, int nx, int ny, int nz, int block_x,
                             int block_y, int block_z, int32_t *ef, int32_t *height,
                             int32_t *caps) {
  return jfc::io_result(node_read_impl(path, nx, ny, nz, block_x, block_y, block_z, ef, height, caps), "jfc_node_read", path);
}

extern "C" int jfc_node_write(const char *path, int nx, int ny, int nz, int block_x,
                              int block_y, int block_z, const int32_t *ef,
                              const int32_t *height, const int32_t *caps) {
  return jfc::io_result(node_write_impl(path, nx, ny, nz, block_x, block_y, block_z, ef, height, caps), "jfc_node_write", path);
}
