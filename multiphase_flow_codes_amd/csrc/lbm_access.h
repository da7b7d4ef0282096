// lbm_access.h -- device-side access to the resident population buffer.
//
// The resident buffer always holds post-collision (pre-streaming) values;
// "pulled" access reads f_i(x) = buf[i][x - c_i], i.e. the post-streaming
// field, with half-way bounce-back at the y faces when walls are on
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

// Source of the pulled population i at node (x,y,z): plane index and node.
template <bool WALLS>
__device__ __forceinline__ void pull_src(int i, int x, int y, int z, uint32_t n, const Grid& g,
                                         int& plane, uint32_t& src) {
  const int sx = x - cx(i), sy = y - cy(i), sz = z - cz(i);
  const int wx = sx < 0 ? sx + g.nx : (sx >= g.nx ? sx - g.nx : sx);
  const int wy = sy < 0 ? sy + g.ny : (sy >= g.ny ? sy - g.ny : sy);
  const int wz = sz < 0 ? sz + g.nz : (sz >= g.nz ? sz - g.nz : sz);
  plane = i;
  src = (uint32_t)wx + (uint32_t)g.nx * ((uint32_t)wy + (uint32_t)g.ny * (uint32_t)wz);
  if (WALLS && (sy < 0 || sy >= g.ny)) {  // half-way bounce-back (CHAN/streaming.f90:175-187)
    plane = opp(i);
    src = n;
  }
}

template <typename T, typename R, bool WALLS>
__device__ __forceinline__ void load_pulled(const T* __restrict__ A, const Grid& g, int x, int y,
                                            int z, uint32_t n, R (&f)[NPOP]) {
#pragma unroll
  for (int i = 0; i < NPOP; ++i) {
    int pl;
    uint32_t s;
    pull_src<WALLS>(i, x, y, z, n, g, pl, s);
    f[i] = (R)A[(uint64_t)pl * g.ncell + s];
  }
}

template <typename T, typename R>
__device__ __forceinline__ void load_direct(const T* __restrict__ A, const Grid& g, uint32_t n,
                                            R (&f)[NPOP]) {
#pragma unroll
  for (int i = 0; i < NPOP; ++i) f[i] = (R)A[(uint64_t)i * g.ncell + n];
}

// Write population i of node n into the resident buffer so that the current
// phase reproduces it: directly (pre-streaming) or at its pull source
// (post-streaming).
template <typename T, bool WALLS>
__device__ __forceinline__ void store_phase(T* __restrict__ A, const Grid& g, int i, int x, int y,
                                            int z, uint32_t n, bool pulled, double v) {
  if (!pulled) {
    A[(uint64_t)i * g.ncell + n] = (T)v;
    return;
  }
  int pl;
  uint32_t s;
  pull_src<WALLS>(i, x, y, z, n, g, pl, s);
  A[(uint64_t)pl * g.ncell + s] = (T)v;
}

}  // namespace mfc
