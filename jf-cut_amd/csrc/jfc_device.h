// jfc_device.h -- device-side helpers shared by the gfx950 kernels: tile
// shapes, direction tables, tile/node addressing and wave primitives.
#pragma once

#include "jfc_internal.h"

namespace jfc {

// A tile of TW x TH (x TD) nodes, one node per lane of a 256-lane
// workgroup (4 wave64s), staged in LDS with a one-node halo.
template <int TW_, int TH_, int TD_>
struct Shape {
  static constexpr int TW = TW_, TH = TH_, TD = TD_;
  static constexpr bool D3 = (TD_ > 1);
  static constexpr int PW = TW + 2, PH = TH + 2, PD = D3 ? TD + 2 : 1;
  static constexpr int PADN = PW * PH * PD;
  static constexpr int NT = TW * TH * TD;
  static constexpr int KMAX = D3 ? 6 : 8;
  static_assert(NT == 256, "one node per lane of a 256-lane workgroup");
};

using S32x8 = Shape<32, 8, 1>;
using S64x4 = Shape<64, 4, 1>;
using S16x16 = Shape<16, 16, 1>;
using S8x8x4 = Shape<8, 8, 4>;
using S32x4x2 = Shape<32, 4, 2>;

#define JFC_DISPATCH(shape, BODY)                           \
  switch (shape) {                                          \
    case SH_32x8: { using S = S32x8; BODY; } break;         \
    case SH_64x4: { using S = S64x4; BODY; } break;         \
    case SH_16x16: { using S = S16x16; BODY; } break;       \
    case SH_8x8x4: { using S = S8x8x4; BODY; } break;       \
    case SH_32x4x2: { using S = S32x4x2; BODY; } break;     \
    default: break;                                         \
  }

// Direction d -> offset.  2D: +x -x +y -y (+x+y) (-x-y) (-x+y) (+x-y);
// 3D: +x -x +y -y +z -z.  The reverse of d is d^1.
template <bool D3>
__device__ __forceinline__ void dir_of(int d, int &dx, int &dy, int &dz) {
  if (D3) {
    dx = (d == 0) - (d == 1);
    dy = (d == 2) - (d == 3);
    dz = (d == 4) - (d == 5);
  } else {
    dx = (d == 0 || d == 4 || d == 7) - (d == 1 || d == 5 || d == 6);
    dy = (d == 2 || d == 4 || d == 6) - (d == 3 || d == 5 || d == 7);
    dz = 0;
  }
}

template <class S>
__device__ __forceinline__ int lds_off(int d) {
  int dx, dy, dz;
  dir_of<S::D3>(d, dx, dy, dz);
  return dx + S::PW * dy + S::PW * S::PH * dz;
}

// Tile t resolved to its instance and origin.
struct TileRef {
  InstDesc I;
  int ltx, lty, ltz;
  int ox, oy, oz;
};

template <class S>
__device__ __forceinline__ TileRef tile_ref_of(const InstDesc &I, int t) {
  TileRef r;
  r.I = I;
  const int lt = t - r.I.tile_off;
  r.ltx = lt % r.I.tx;
  const int q = lt / r.I.tx;
  r.lty = q % r.I.ty;
  r.ltz = q / r.I.ty;
  r.ox = r.ltx * S::TW;
  r.oy = r.lty * S::TH;
  r.oz = r.ltz * S::TD;
  return r;
}

template <class S>
__device__ __forceinline__ TileRef tile_ref(const Tiles &T, int t) {
  // one instance: the descriptor does not depend on t (no dependent load)
  return tile_ref_of<S>(T.ninst == 1 ? T.inst[0] : T.inst[T.tile_inst[t]], t);
}


__device__ __forceinline__ long long node_index(const InstDesc &I, int x, int y,
                                                int z) {
  return I.node_off + x + (long long)I.px * (y + (long long)I.ny * z);
}

template <class S>
__device__ __forceinline__ int tile_of(const InstDesc &I, int x, int y, int z) {
  return I.tile_off + x / S::TW + I.tx * (y / S::TH + I.ty * (z / S::TD));
}

// Lane layout inside the tile.
template <class S>
struct Lane {
  int lx, ly, lz, me;
  __device__ __forceinline__ explicit Lane(int tid) {
    lx = tid % S::TW;
    ly = (tid / S::TW) % S::TH;
    lz = tid / (S::TW * S::TH);
    me = (lx + 1) + S::PW * (ly + 1) + (S::D3 ? S::PW * S::PH * (lz + 1) : 0);
  }
};

// Padded cell p -> local coordinates (may be -1 or TW etc. on the halo).
template <class S>
__device__ __forceinline__ void pad_coords(int p, int &x, int &y, int &z,
                                           bool &interior, int &nout) {
  const int px = p % S::PW;
  const int py = (p / S::PW) % S::PH;
  const int pz = p / (S::PW * S::PH);
  x = px - 1;
  y = py - 1;
  z = S::D3 ? pz - 1 : 0;
  nout = (x < 0 || x >= S::TW) + (y < 0 || y >= S::TH) +
         (S::D3 ? (z < 0 || z >= S::TD) : 0);
  interior = nout == 0;
}

// Halo cell needed by a connectivity: face cells always, 2D corners only for
// 8-connectivity, 3D edges/corners never.
template <class S>
__device__ __forceinline__ bool halo_needed(int nout, int K) {
  if (nout == 1) return true;
  return !S::D3 && nout == 2 && K == 8;
}

// Padded cell p of a tile as seen by one lane: halo (not interior), and
// live = a halo cell the connectivity needs that lies inside the instance,
// with its node index (0 otherwise, a valid dummy address for loads).
struct HaloCell {
  bool halo, live;
  long long g;
};

template <class S>
__device__ __forceinline__ HaloCell halo_cell(const InstDesc &I, int ox, int oy, int oz,
                                              int p, int K) {
  HaloCell c{false, false, 0};
  if (p >= S::PADN) return c;
  int x, y, z, nout;
  bool inter;
  pad_coords<S>(p, x, y, z, inter, nout);
  if (inter) return c;
  c.halo = true;
  if (!halo_needed<S>(nout, K)) return c;
  const int X = ox + x, Y = oy + y, Z = oz + z;
  if (X >= 0 && X < I.nx && Y >= 0 && Y < I.ny && Z >= 0 && Z < I.nz) {
    c.live = true;
    c.g = node_index(I, X, Y, Z);
  }
  return c;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Append `t` (if pred) to a list with one atomic per wave.
__device__ __forceinline__ void wave_append(bool pred, int t, int *list,
                                            int *count, int *sweep_cnt,
                                            long long *total) {
  const unsigned long long m = __ballot(pred);
  const int n = __popcll(m);
  if (n == 0) return;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) {
    base = atomicAdd(count, n);
    if (sweep_cnt) atomicAdd(sweep_cnt, n);
    if (total) atomicAdd((unsigned long long *)total, (unsigned long long)n);
  }
  base = __shfl(base, leader);
  if (pred) {
    const unsigned long long below = m & ((1ull << lane) - 1ull);
    list[base + __popcll(below)] = t;
  }
}

}  // namespace jfc
