// lbm_access.h -- device-side access to the resident population buffer.
//
// The resident buffer always holds post-collision (pre-streaming) values,
// laid out as Grid::lay says (PopLayout).  "Pulled" access reads the streamed
// value f_i(x), i.e. the post-collision value of direction i at x - c_i, with
// half-way bounce-back at the y faces when walls are on: beyond a wall the
// pull takes the node's own post-collision value of opp(i)
// (CHAN/streaming.f90:175-187, moved from the x faces to the y faces).
#pragma once
#include "lbm_common.h"

namespace mfc {

__device__ __forceinline__ void node_coords(uint32_t n, const Grid& g, int& x, int& y, int& z,
                                            uint32_t& row) {
  row = n / (uint32_t)g.nx;
  x = (int)(n - row * (uint32_t)g.nx);
  const uint32_t zz = row / (uint32_t)g.ny;
  y = (int)(row - zz * (uint32_t)g.ny);
  z = (int)zz;
}

// Offset of node (x, y, z) within a plane (the plane's own element is
// plane * pitch + this).
__device__ __forceinline__ uint64_t node_off(const Grid& g, int x, int y, int z) {
  return (uint64_t)x + g.rs * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
}

__device__ __forceinline__ uint64_t node_off_wrapped(const Grid& g, int x, int y, int z) {
  const int wx = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
  const int wy = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
  const int wz = z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z);
  return node_off(g, wx, wy, wz);
}

// Location (plane, in-plane offset) of the post-collision value of direction
// i at node (x,y,z) in layout `lay`.
template <bool WALLS>
__device__ __forceinline__ void post_src(int lay, int i, int x, int y, int z, const Grid& g,
                                         int& plane, uint64_t& slot) {
  if (lay == LAY_AB) {
    plane = i;
    slot = node_off(g, x, y, z);
    return;
  }
  if (lay == LAY_AA_EVEN) {
    plane = opp(i);
    slot = node_off(g, x, y, z);
    return;
  }
  const int ty = y + cy(i);
  if (WALLS && (ty < 0 || ty >= g.ny)) {
    plane = opp(i);
    slot = node_off(g, x, y, z);
    return;
  }
  plane = i;
  slot = node_off_wrapped(g, x + cx(i), ty, z + cz(i));
}

// Location of the streamed value f_i at node (x,y,z): the post-collision value
// of i at x - c_i, or beyond a wall that of opp(i) at the node itself.
template <bool WALLS>
__device__ __forceinline__ void pull_src_lay(int lay, int i, int x, int y, int z, const Grid& g,
                                             int& plane, uint64_t& slot) {
  if (lay == LAY_AA_ODD) {   // x - c_i + c_i = x, and the wall image lands there too
    plane = i;
    slot = node_off(g, x, y, z);
    return;
  }
  const int sy = y - cy(i);
  const bool wall = WALLS && (sy < 0 || sy >= g.ny);
  if (wall) {
    plane = lay == LAY_AB ? opp(i) : i;
    slot = node_off(g, x, y, z);
    return;
  }
  plane = lay == LAY_AB ? i : opp(i);
  slot = node_off_wrapped(g, x - cx(i), sy, z - cz(i));
}

template <bool WALLS>
__device__ __forceinline__ void pull_src(int i, int x, int y, int z, const Grid& g, int& plane,
                                         uint64_t& src) {
  pull_src_lay<WALLS>(g.lay, i, x, y, z, g, plane, src);
}

template <typename T, typename R, bool WALLS>
__device__ __forceinline__ void load_pulled(const T* A, const Grid& g, int x, int y, int z,
                                            uint32_t n, R (&f)[NPOP]) {
#pragma unroll
  for (int i = 0; i < NPOP; ++i) {
    int pl;
    uint64_t s;
    pull_src<WALLS>(i, x, y, z, g, pl, s);
    f[i] = (R)A[(uint64_t)pl * g.pitch + s];
  }
}

// The post-collision values of node n (the reference's f between collision
// and streaming).
template <typename T, typename R, bool WALLS>
__device__ __forceinline__ void load_direct(const T* A, const Grid& g, int x, int y, int z,
                                            uint32_t n, R (&f)[NPOP]) {
#pragma unroll
  for (int i = 0; i < NPOP; ++i) {
    int pl;
    uint64_t s;
    post_src<WALLS>(g.lay, i, x, y, z, g, pl, s);
    f[i] = (R)A[(uint64_t)pl * g.pitch + s];
  }
}

// Write population i of node n into the resident buffer so that the current
// phase reproduces it: as its post-collision value (pre-streaming) or at its
// pull source (post-streaming).
template <typename T, bool WALLS>
__device__ __forceinline__ void store_phase(T* A, const Grid& g, int i, int x, int y, int z,
                                            uint32_t n, bool pulled, double v) {
  int pl;
  uint64_t s;
  if (pulled)
    pull_src<WALLS>(i, x, y, z, g, pl, s);
  else
    post_src<WALLS>(g.lay, i, x, y, z, g, pl, s);
  A[(uint64_t)pl * g.pitch + s] = (T)v;
}

// Value of direction i at node n after collision, read by the bounce-back.
template <typename T>
__device__ __forceinline__ double post_value(const T* A, const Grid& g, int i, uint32_t n) {
  // plane-major two buffers: the node index is the in-plane offset
  if (g.lay == LAY_AB && g.rs == (uint64_t)g.nx) return (double)A[(uint64_t)i * g.pitch + n];
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  int pl;
  uint64_t s;
  if (g.walls_y)
    post_src<true>(g.lay, i, x, y, z, g, pl, s);
  else
    post_src<false>(g.lay, i, x, y, z, g, pl, s);
  return (double)A[(uint64_t)pl * g.pitch + s];
}

}  // namespace mfc
