// particles.hip -- finite-size moving spheres: node maps, boundary links,
// interpolated bounce-back with momentum exchange, lubrication, Newton-Euler
// motion and refilling of uncovered nodes (BOX/partlib.f90).
//
// Data layout.  Node maps are double-buffered (current / previous step):
// an int32 id map (particle id, 1-based, or 0) and a 1-bit solid mask read by
// the collide kernel.  Boundary links are a compact list grouped by particle
// (node index, direction, wall fraction), built in a deterministic order by
// one workgroup per particle scanning the particle's bounding box; no dense
// 18-per-node link arrays exist.  Maps are maintained incrementally: a build
// first clears the nodes its buffer marked two steps earlier, so no pass
// ever touches the whole lattice.  All particle state is FP64.
//
// Compiled with -ffp-contract=off: the geometry is evaluated exactly as the
// reference writes it, so maps and wall fractions match the FP64 oracle bit
// for bit.
#include "force_table.h"
#include "internal.h"
#include "lbm_access.h"
#include "lbm_particle_math.h"
#include "step_args.h"

namespace mfc {

constexpr int PB = 256;  // threads per particle workgroup
// A segment is one z-slab (G.slab planes, 4 or 8) of one particle's bounding
// box: the unit of work of every box-scan kernel, so a sphere spreads over
// several workgroups.  Segment s of particle p is block p * nseg + s.


struct AxisRange {
  int lo, cnt;
};

__device__ __forceinline__ int wrap_any(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ bool sane_coord(double c, int n) { return isfinite(c) && c >= 0.0 && c < (double)n; }

// Candidate node indices (unwrapped, 0-based) whose centre i+0.5 can lie
// within radp2 of c: a superset of the reference's test |pnt - c_img| <= radp2.
__device__ __forceinline__ AxisRange axis_range(double c, double radp2) {
  AxisRange r;
  r.lo = (int)floor(c - radp2 - 0.5) - 1;
  const int hi = (int)ceil(c + radp2 - 0.5) + 1;
  r.cnt = hi - r.lo + 1;
  return r;
}

// Per-particle constants of a box scan, computed by thread 0.
struct BoxInfo {
  int ok;
  double x0, y0, z0, xr, yr;   // original centre and sweep-rest bases
  int kfirst, jfirst;          // first k (index order) in the box, first j of that row
  AxisRange ax, ay, az;
};

__device__ double y_image(double ybase, double ypnt, const PGeom& G) {
  return G.walls_y ? ybase : near_image(ybase, ypnt, G.ny);
}

// one-period wrap (the box candidate ranges never span more than a period)
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

__device__ __forceinline__ bool bit_at(const uint32_t* __restrict__ mask, uint32_t n) {
  return (mask[n >> 5] >> (n & 31u)) & 1u;
}

// ---- pair-search cells ------------------------------------------------------
// Cell of coordinate v in [0, n) on an axis of nc cells (each at least as
// wide as the longest pair range, so partners sit in adjacent cells).
__device__ __forceinline__ int cell_of(double v, int n, int nc) {
  int c = (int)(v * (double)nc / (double)n);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__device__ __forceinline__ void cell_coords(const double* __restrict__ ypglb, int p, const PGeom& G,
                                            const CellArgs& C, int& cx_, int& cy_, int& cz_) {
  cx_ = cell_of(ypglb[3 * p], G.nx, C.ncx);
  cy_ = cell_of(ypglb[3 * p + 1], G.ny, C.ncy);
  cz_ = cell_of(ypglb[3 * p + 2], G.nz, C.ncz);
}

// Index of neighbour cell k (0..26) of (cx_, cy_, cz_), or -1 beyond a wall.
__device__ __forceinline__ int cell_neighbour(const CellArgs& C, const PGeom& G, int cx_, int cy_,
                                              int cz_, int k) {
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  int y = cy_ + dy;
  if (G.walls_y) {
    if (y < 0 || y >= C.ncy) return -1;
  } else {
    y = wrap1(y, C.ncy);
  }
  const int x = wrap1(cx_ + dx, C.ncx), z = wrap1(cz_ + dz, C.ncz);
  return x + C.ncx * (y + C.ncy * z);
}

__global__ void k_cell_zero(int32_t* __restrict__ cnt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = 0;
}

__global__ void __launch_bounds__(256)
    k_cell_build(const double* __restrict__ ypglb, int np, PGeom G, CellArgs C,
                 int32_t* __restrict__ status) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (!(sane_coord(ypglb[3 * p], G.nx) && sane_coord(ypglb[3 * p + 1], G.ny) &&
        sane_coord(ypglb[3 * p + 2], G.nz)))
    return;   // reported by the move / links kernels
  int a, b, c;
  cell_coords(ypglb, p, G, C, a, b, c);
  const int cell = a + C.ncx * (b + C.ncy * c);
  const int slot = atomicAdd(&C.cnt[cell], 1);
  if (slot < kCellCap)
    C.list[cell * kCellCap + slot] = p;
  else
    atomicAdd(&status[ST_CELL_OVERFLOW], 1);
}

// Overlap test of ovlpchk (partlib.f90:118-171) for the pair (p, q), q > p.
__device__ __forceinline__ bool pair_overlaps(const double* __restrict__ ypglb, int p, int q,
                                              const PGeom& G) {
  const double xc = ypglb[3 * p], yc = ypglb[3 * p + 1], zc = ypglb[3 * p + 2];
  const double xb = near_image(ypglb[3 * q], xc, G.nx);
  const double yb = G.walls_y ? ypglb[3 * q + 1] : near_image(ypglb[3 * q + 1], yc, G.ny);
  const double zb = near_image(ypglb[3 * q + 2], zc, G.nz);
  const double dist = sqrt((xc - xb) * (xc - xb) + (yc - yb) * (yc - yb) + (zc - zb) * (zc - zb));
  return dist - 2.0 * G.rad < 0.0;
}

// Does particle p overlap a later one?  Evaluated by the calling block: all
// later particles (no cells) or one neighbour cell per thread.
__device__ bool block_overlap(const double* __restrict__ ypglb, int p, int np, const PGeom& G,
                              const CellArgs& C) {
  __shared__ int hit;
  if (threadIdx.x == 0) hit = 0;
  __syncthreads();
  if (C.cnt) {
    if (threadIdx.x < 27) {
      int a, b, c;
      cell_coords(ypglb, p, G, C, a, b, c);
      const int cell = cell_neighbour(C, G, a, b, c, (int)threadIdx.x);
      if (cell >= 0) {
        const int n = min(C.cnt[cell], kCellCap);
        for (int k = 0; k < n; ++k) {
          const int q = C.list[cell * kCellCap + k];
          if (q > p && pair_overlaps(ypglb, p, q, G)) hit = 1;
        }
      }
    }
  } else {
    for (int q = p + 1 + threadIdx.x; q < np; q += PB)
      if (pair_overlaps(ypglb, p, q, G)) hit = 1;
  }
  __syncthreads();
  return hit != 0;
}

// Per-particle constants of a box scan (shared memory), computed by the whole
// block: thread 0 sets the centre and candidate ranges, then every thread
// tests one z or y candidate and the first passing index wins by atomicMin.
__device__ void box_info(const double* __restrict__ pos, int p, const PGeom& G, BoxInfo& B) {
  if (threadIdx.x == 0) {
    B.x0 = pos[3 * p];
    B.y0 = pos[3 * p + 1];
    B.z0 = pos[3 * p + 2];
    B.ok = sane_coord(B.x0, G.nx) && sane_coord(B.y0, G.ny) && sane_coord(B.z0, G.nz);
    if (B.ok) {
      B.xr = sweep_rest_base(B.x0, G.nx);
      B.yr = G.walls_y ? B.y0 : sweep_rest_base(B.y0, G.ny);
      B.ax = axis_range(B.x0, G.radp2);
      B.ay = axis_range(B.y0, G.radp2);
      B.az = axis_range(B.z0, G.radp2);
      if (B.ax.cnt > G.nx || B.ay.cnt > G.ny || B.az.cnt > G.nz) B.ok = 0;
    }
    B.kfirst = 1 << 30;
    B.jfirst = 1 << 30;
  }
  __syncthreads();
  if (B.ok) {
    for (int t = threadIdx.x; t < B.az.cnt; t += blockDim.x) {
      const int kw = wrap1(B.az.lo + t, G.nz);
      const double zp = (double)(kw + 1) - 0.5;
      const double zz0 = zp - near_image(B.z0, zp, G.nz);
      if (fabs(zz0) <= G.radp2) atomicMin(&B.kfirst, kw);
    }
    for (int t = threadIdx.x; t < B.ay.cnt; t += blockDim.x) {
      const int jw = G.walls_y ? B.ay.lo + t : wrap1(B.ay.lo + t, G.ny);
      if (jw < 0 || jw >= G.ny) continue;
      const double yp = (double)(jw + 1) - 0.5;
      const double yy0 = yp - y_image(B.y0, yp, G);
      if (fabs(yy0) <= G.radp2) atomicMin(&B.jfirst, jw);
    }
  }
  __syncthreads();
}

// Geometry of node (iw, jw, kw) of the box.  Returns false if the node is
// outside the reference's bounding box (partlib.f90:965-985) or beyond a wall.
__device__ bool node_geom(const BoxInfo& B, const PGeom& G, int iw, int jw, int kw, LinkGeom& L) {
  if (jw < 0 || jw >= G.ny) return false;
  const double zp = (double)(kw + 1) - 0.5;
  L.zimg = near_image(B.z0, zp, G.nz);
  L.zz0 = zp - L.zimg;
  if (!(fabs(L.zz0) <= G.radp2)) return false;
  const double yp = (double)(jw + 1) - 0.5;
  L.yimg = y_image(kw == B.kfirst ? B.y0 : B.yr, yp, G);
  L.yy0 = yp - L.yimg;
  if (!(fabs(L.yy0) <= G.radp2)) return false;
  const double xp = (double)(iw + 1) - 0.5;
  L.ximg = near_image((kw == B.kfirst && jw == B.jfirst) ? B.x0 : B.xr, xp, G.nx);
  L.xx0 = xp - L.ximg;
  if (!(fabs(L.xx0) <= G.radp2)) return false;
  L.rr0 = L.xx0 * L.xx0 + L.yy0 * L.yy0 + L.zz0 * L.zz0;
  return true;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total);

// Rows (y, z) of one segment's walk: the x run of each row that can hold
// nodes within r + 1.5 of the centre (a superset: the node test decides),
// and the runs' prefix sums.  A segment has at most PB rows (one thread per
// row): slab planes of at most PB / slab candidate rows each, which the
// choice of slab guarantees (segment_planes).
struct SegRows {
  int start[PB + 1];   // first flattened index of each row
  short xa[PB];        // first box x index of each row's run
  int nrows;
};

// Walk of one segment (z-slab s) of the box in passes of PB candidate nodes,
// the slab's row runs flattened x-fastest so every lane of a pass has a node.
// Every thread runs the same number of passes (barriers are allowed in fn);
// fn(active, iw, jw, kw) is called once per pass, in row-major, x-ascending
// node order.
template <typename F>
__device__ __forceinline__ void walk_segment(const BoxInfo& B, const PGeom& G, int s, SegRows& R,
                                             int* sh, F&& fn) {
  const int z0 = s * G.slab;
  const int nzs = max(0, min(G.slab, B.az.cnt - z0));
  const int nrows = nzs * B.ay.cnt;   // <= PB
  // x run of each row (a row at most PB nodes)
  int run = 0, xa = 0;
  const int r = (int)threadIdx.x;
  if (r < nrows) {
    const int tz = r / B.ay.cnt, ty = r - tz * B.ay.cnt;
    const int kw = wrap1(B.az.lo + z0 + tz, G.nz);
    const int jw = G.walls_y ? B.ay.lo + ty : wrap1(B.ay.lo + ty, G.ny);
    if (jw >= 0 && jw < G.ny) {
      const double zp = (double)(kw + 1) - 0.5;
      const double zz0 = zp - near_image(B.z0, zp, G.nz);
      const double yp = (double)(jw + 1) - 0.5;
      const double yy0 = yp - y_image(kw == B.kfirst ? B.y0 : B.yr, yp, G);
      const double rem = G.rlink2 - yy0 * yy0 - zz0 * zz0;
      if (fabs(zz0) <= G.radp2 && fabs(yy0) <= G.radp2 && rem >= 0.0) {
        // node i (unwrapped) sits at i + 0.5; keep |i + 0.5 - x0| <= sqrt(rem) + 1
        const double dx = sqrt(rem) + 1.0;
        const int ilo = max((int)ceil(B.x0 - dx - 0.5), B.ax.lo);
        const int ihi = min((int)floor(B.x0 + dx - 0.5), B.ax.lo + B.ax.cnt - 1);
        if (ihi >= ilo) {
          run = ihi - ilo + 1;
          xa = ilo - B.ax.lo;
        }
      }
    }
  }
  int total;
  const int st = block_exclusive_scan(run, sh, total);
  if (r < nrows) {
    R.start[r] = st;
    R.xa[r] = (short)xa;
  }
  if (r == 0) R.start[nrows] = total;
  __syncthreads();
  for (int t0 = 0; t0 < total; t0 += PB) {
    const int t = t0 + (int)threadIdx.x;
    const bool act = t < total;
    int iw = 0, jw = 0, kw = 0;
    if (act) {
      int lo = 0, hi = nrows;   // the row holding t: start[lo] <= t < start[lo + 1]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (R.start[mid] <= t) lo = mid; else hi = mid;
      }
      const int tz = lo / B.ay.cnt, ty = lo - tz * B.ay.cnt;
      const int tx = R.xa[lo] + (t - R.start[lo]);
      kw = wrap1(B.az.lo + z0 + tz, G.nz);
      jw = G.walls_y ? B.ay.lo + ty : wrap1(B.ay.lo + ty, G.ny);
      iw = wrap1(B.ax.lo + tx, G.nx);
    }
    fn(act, iw, jw, kw);
  }
}

// Neighbour test of direction q from node (iw,jw,kw); fills rr1 and offsets.
__device__ __forceinline__ bool link_hit(const LinkGeom& L, int iw, int jw, int kw, int q, double r2,
                                         double& xx1, double& yy1, double& zz1, double& rr1) {
  xx1 = (double)(iw + 1 + cx(q)) - 0.5 - L.ximg;
  yy1 = (double)(jw + 1 + cy(q)) - 0.5 - L.yimg + 0.0;
  zz1 = (double)(kw + 1 + cz(q)) - 0.5 - L.zimg + 0.0;
  rr1 = xx1 * xx1 + yy1 * yy1 + zz1 * zz1;
  return (rr1 - r2) <= 0.0;
}

// Wavefront-wide inclusive scans through DPP lane moves (no LDS round trip):
// within each row of 16 lanes by shifts of 1, 2, 4 and 8 (lanes shifted in
// from outside the row contribute 0), then lane 15 of each row into the next
// row and lane 31 into the upper half.  Every lane of the wavefront must be
// active.  DPP controls: row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143; rows left out by the row mask keep the old value 0.
__device__ __forceinline__ int wave_inclusive_add(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  return x;
}

// Bitwise OR of x over the wavefront (every lane active), in every lane.
__device__ __forceinline__ uint32_t wave_or(uint32_t x) {
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Exclusive prefix sum over the PB threads of a block: a DPP scan inside
// each 64-lane wavefront, then the wavefront totals through LDS.
__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int x = wave_inclusive_add(v);
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < PB / 64; ++q) {
    const int t = sh[q];
    before += (q < w) ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return before + x - v;
}


// Set (or clear) the solid bit of node n for every lane with `on`, issuing
// one atomic per distinct mask word per wavefront: lanes walking a box row
// share one or two words, and same-address atomics from one instruction
// would otherwise serialise.  Must be called by every lane of the wavefront.
__device__ __forceinline__ void mask_update(uint32_t* __restrict__ mask, bool on, uint32_t n, bool set) {
  uint64_t pending = __ballot(on);
  const uint32_t wd = n >> 5;
  const uint32_t bit = on ? (1u << (n & 31u)) : 0u;
  const int lane = threadIdx.x & 63;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t wl = (uint32_t)__builtin_amdgcn_readlane((int)wd, leader);
    const bool mem = on && wd == wl;
    const uint32_t v = wave_or(mem ? bit : 0u);
    if (lane == leader) {
      if (set)
        atomicOr(&mask[wl], v);
      else
        atomicAnd(&mask[wl], ~v);
    }
    pending &= ~__ballot(mem);
  }
}

// Position of each flagged lane in a deterministic block-wide order (pass
// order, then thread order) for two flags at once, with one barrier per call:
// wave totals go through LDS slots that alternate between calls.  base_* hold
// the running totals of earlier calls and are advanced by this one.
__device__ __forceinline__ void block_place2(bool fa, bool fb, int (*wt)[PB / 64][2], int& parity,
                                             int& base_a, int& base_b, int& at_a, int& at_b) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t ba = __ballot(fa), bb = __ballot(fb);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (lane == 0) {
    wt[parity][w][0] = __popcll(ba);
    wt[parity][w][1] = __popcll(bb);
  }
  __syncthreads();
  int pa = 0, pb = 0, ta = 0, tb = 0;
#pragma unroll
  for (int q = 0; q < PB / 64; ++q) {
    const int ca = wt[parity][q][0], cb = wt[parity][q][1];
    pa += (q < w) ? ca : 0;
    pb += (q < w) ? cb : 0;
    ta += ca;
    tb += cb;
  }
  at_a = base_a + pa + __popcll(ba & below);
  at_b = base_b + pb + __popcll(bb & below);
  base_a += ta;
  base_b += tb;
  parity ^= 1;
}

// beads_links, first half of the map rebuild (partlib.f90:945-949): reset the
// nodes the same map buffer marked two steps earlier, from that build's solid
// lists.  Every marked node of the buffer is to be reset before any new mark
// (the marks come in a later grid), so each listed node's id and its whole
// 32-node mask word are stored as zero: no read, no atomic; a word or a node
// listed by several spheres receives the same zero from each.
__device__ void clear_segment(int b, int nseg, int seg_cap, const uint32_t* __restrict__ sol_list,
                              const int32_t* __restrict__ sol_cnt, int32_t* __restrict__ isn,
                              uint32_t* __restrict__ mask) {
  const int cnt = sol_cnt[b];
  const uint32_t* lst = sol_list + (size_t)b * seg_cap;
  for (int e = (int)threadIdx.x; e < cnt; e += PB) {
    const uint32_t n = lst[e];
    isn[n] = 0;
    mask[n >> 5] = 0u;
  }
}

__global__ void __launch_bounds__(PB)
    k_map_clear(int nseg, int seg_cap, const uint32_t* __restrict__ sol_list,
                const int32_t* __restrict__ sol_cnt, int32_t* __restrict__ isn,
                uint32_t* __restrict__ mask) {
  clear_segment(blockIdx.x, nseg, seg_cap, sol_list, sol_cnt, isn, mask);
}

// beads_links (partlib.f90:934-1043) for one segment of a particle's box:
// classify every box node (solid: mark the id map and mask, list it; fluid
// within r + 1.5: keep it in LDS), then test the 18 directions of the kept
// nodes, one node per lane, and store node and direction bits per segment.
// With `ovl` set the first segment of each particle also performs the overlap
// check of beads_redistribute (partlib.f90:2631-2646) on the freshly wrapped
// positions.  Dynamic LDS: seg_cap node indices.
__device__ __forceinline__ void links_scan_body(int b_, const double* __restrict__ ypglb, int np, int nseg, int seg_cap, PGeom G,
                 int32_t* __restrict__ isn, uint32_t* __restrict__ mask,
                 uint32_t* __restrict__ sol_list, int32_t* __restrict__ sol_cnt,
                 uint32_t* __restrict__ sh_node, uint32_t* __restrict__ sh_hits,
                 int32_t* __restrict__ sh_cnt, int32_t* __restrict__ lk_cnt, int ovl, int istep,
                 int32_t* __restrict__ status, CellArgs cells) {
  extern __shared__ uint32_t shell[];
  __shared__ BoxInfo B;
  __shared__ int wt[2][PB / 64][2];
  __shared__ int sh[PB / 64];
  const int b = b_;
  const int p = b / nseg, s = b - p * nseg;
  if (p >= np) return;
  if (ovl && s == 0) {
    // ovlpchk: flag p if any later particle is closer than 2r
    const bool hit = block_overlap(ypglb, p, np, G, cells);
    if (threadIdx.x == 0 && hit) {
      atomicAdd(&status[ST_OVERLAP], 1);
      atomicCAS(&status[ST_OVERLAP_STEP], 0, istep);
    }
  }
  box_info(ypglb, p, G, B);
  if (!B.ok) {
    if (threadIdx.x == 0) {
      lk_cnt[b] = 0;
      sol_cnt[b] = 0;
      sh_cnt[b] = 0;
      if (s == 0) atomicAdd(&status[ST_BOX_FAULT], 1);
    }
    return;
  }
  uint32_t* sol = sol_list + (size_t)b * seg_cap;
  int parity = 0, nsol = 0, nshell = 0;
  __shared__ SegRows R;
  walk_segment(B, G, s, R, sh, [&](bool act, int iw, int jw, int kw) {
    LinkGeom L;
    const bool in = act && node_geom(B, G, iw, jw, kw, L) && L.rr0 <= G.rlink2;
    const bool solid = in && (L.rr0 - G.r2 <= 0.0);
    const bool near = in && !solid;
    const uint32_t n = (uint32_t)iw + (uint32_t)G.nx * ((uint32_t)jw + (uint32_t)G.ny * (uint32_t)kw);
    if (solid) atomicMax(&isn[n], p + 1);   // the highest id containing the node wins
    mask_update(mask, solid, n, true);
    int at_s, at_n;
    block_place2(solid, near, wt, parity, nsol, nshell, at_s, at_n);
    if (solid && at_s < seg_cap) sol[at_s] = n;
    if (near && at_n < seg_cap) shell[at_n] = n;
  });
  __syncthreads();
  const int nsh = min(nshell, seg_cap);
  uint32_t* shn = sh_node + (size_t)b * seg_cap;
  uint32_t* shh = sh_hits + (size_t)b * seg_cap;
  int nl = 0;
  for (int e0 = 0; e0 < nsh; e0 += PB) {
    const int e = e0 + (int)threadIdx.x;
    if (e < nsh) {
      const uint32_t n = shell[e];
      const int iw = (int)(n % (uint32_t)G.nx);
      const uint32_t rw = n / (uint32_t)G.nx;
      const int jw = (int)(rw % (uint32_t)G.ny), kw = (int)(rw / (uint32_t)G.ny);
      LinkGeom L;
      node_geom(B, G, iw, jw, kw, L);
      uint32_t hits = 0;
      for (int q = 1; q < NPOP; ++q) {
        double a1, a2, a3, a4;
        if (link_hit(L, iw, jw, kw, q, G.r2, a1, a2, a3, a4)) hits |= 1u << q;
      }
      shn[e] = n;
      shh[e] = hits;
      nl += __popc(hits);
    }
  }
  int tot;
  block_exclusive_scan(nl, sh, tot);
  if (threadIdx.x == 0) {
    lk_cnt[b] = tot;
    sol_cnt[b] = min(nsol, seg_cap);
    sh_cnt[b] = nsh;
  }
}

// Exclusive scan of the per-segment link counts (one workgroup of 1024):
// each thread sums a contiguous run, the run totals are scanned in LDS.
__device__ __forceinline__ void scan_counts_body(int b_, const int32_t* __restrict__ cnt, const int32_t* __restrict__ scnt, int nsegs,
                  int64_t* __restrict__ off, int64_t cap, int32_t* __restrict__ status,
                  int32_t* __restrict__ fill_count, int32_t* __restrict__ ct_count) {
  if (threadIdx.x == 0) {
    *fill_count = 0;   // the next beads_filling starts an empty list
    *ct_count = 0;     // the expansion lists this rebuild's contact links
  }
  __shared__ int64_t part[1024];
  __shared__ int64_t spart[1024];
  const int t = threadIdx.x;
  const int per = (nsegs + 1023) / 1024;
  const int a = min(nsegs, t * per), e = min(nsegs, a + per);
  int64_t s = 0, ss = 0;
  for (int i = a; i < e; ++i) {
    s += cnt[i];
    ss += scnt[i];
  }
  part[t] = s;
  spart[t] = ss;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int64_t v = (t >= d) ? part[t - d] : 0;
    const int64_t w = (t >= d) ? spart[t - d] : 0;
    __syncthreads();
    part[t] += v;
    spart[t] += w;
    __syncthreads();
  }
  int64_t acc = part[t] - s;
  for (int i = a; i < e; ++i) {
    off[i] = acc;
    acc += cnt[i];
  }
  if (t == 1023) {
    off[nsegs] = part[1023];
    status[ST_NLINKS] = (int32_t)min(part[1023], (int64_t)0x7fffffff);
    status[ST_NSOLID] = (int32_t)min(spart[1023], (int64_t)0x7fffffff);
    if (part[1023] > cap) atomicAdd(&status[ST_LINK_OVERFLOW], 1);
  }
}

__device__ void fill_find_segment(int b, const FillArgs& fa, int nseg, int32_t* __restrict__ status);

// Second half of beads_links: expand each segment's (node, direction bits)
// into links at the segment's offset, in (node, direction) order, with the
// wall fraction of every link (partlib.f90:1013-1024), one link per lane and
// the node geometry rebuilt exactly as the scan saw it.  Blocks past the link
// segments run the first pass of beads_filling (the new map is complete at
// this point).
__device__ __forceinline__ void links_expand_body(int b_, const double* __restrict__ ypglb, int np, int nseg, int seg_cap, PGeom G,
                   const uint32_t* __restrict__ sh_node, const uint32_t* __restrict__ sh_hits,
                   const int32_t* __restrict__ sh_cnt, const int64_t* __restrict__ off, int64_t cap,
                   uint32_t* __restrict__ lk_node, uint8_t* __restrict__ lk_dir,
                   double* __restrict__ lk_alpha, int32_t* __restrict__ status, int nsegs,
                   FillArgs fa, const uint32_t* __restrict__ mask, uint64_t* __restrict__ ct_list,
                   int32_t* __restrict__ ct_count) {
  __shared__ BoxInfo B;
  __shared__ int sh[PB / 64];
  const int b = b_;
  if (b >= nsegs) {
    fill_find_segment(b - nsegs, fa, nseg, status);
    return;
  }
  const int p = b / nseg;
  if (p >= np) return;
  box_info(ypglb, p, G, B);
  if (!B.ok) return;
  const int nsh = sh_cnt[b];
  const uint32_t* shn = sh_node + (size_t)b * seg_cap;
  const uint32_t* shh = sh_hits + (size_t)b * seg_cap;
  int64_t base = off[b];
  for (int e0 = 0; e0 < nsh; e0 += PB) {
    const int e = e0 + (int)threadIdx.x;
    const bool act = e < nsh;
    uint32_t hits = act ? shh[e] : 0u;
    const uint32_t n = act ? shn[e] : 0u;
    int tot;
    const int excl = block_exclusive_scan(__popc(hits), sh, tot);
    int64_t w = base + excl;
    while (hits) {
      const int q = __ffs(hits) - 1;
      hits &= hits - 1;
      if (w < cap) {
        lk_node[w] = n;
        lk_dir[w] = (uint8_t)q;
      }
      ++w;
    }
    base += tot;
  }
  __syncthreads();   // the segment's links, written above by this block, are read below
  const int64_t a = off[b], e = min(off[b + 1], cap);
  int alpha_bad = 0;
  for (int64_t l = a + threadIdx.x; l < e; l += PB) {
    const uint32_t n = lk_node[l];
    const int q = lk_dir[l];
    const int iw = (int)(n % (uint32_t)G.nx);
    const uint32_t rw = n / (uint32_t)G.nx;
    const int jw = (int)(rw % (uint32_t)G.ny), kw = (int)(rw / (uint32_t)G.ny);
    LinkGeom L;
    node_geom(B, G, iw, jw, kw, L);
    double xx1, yy1, zz1, rr1;
    link_hit(L, iw, jw, kw, q, G.r2, xx1, yy1, zz1, rr1);
    const double alpha = link_alpha(L, xx1, yy1, zz1, rr1, G.r2);
    if (alpha < 0.0 || alpha > 1.0) ++alpha_bad;
    lk_alpha[l] = alpha;
    if (bit_at(mask, n)) {   // the node is solid for another sphere: a contact link
      const int slot = atomicAdd(ct_count, 1);
      if (slot < cap) ct_list[slot] = ((uint64_t)(uint32_t)p << 40) | (uint64_t)l;
    }
  }
  if (alpha_bad) atomicAdd(&status[ST_ALPHA_FAULT], alpha_bad);
}


// beads_lubforce (partlib.f90:2211-2292): one workgroup sums the pair forces
// of particle a; each pair is evaluated as the reference does, from the lower
// id's side (so both partners see bitwise opposite pair forces).
// Sphere-wall lubrication (channel beads_lubforce, CHAN/partlib.f90:2316-2355,
// with its wall axis moved to y): the wall at y = 0 pushes up, the wall at
// y = ny down.  Added to the y pair sum.
__device__ __forceinline__ double wall_lub(double yd, const PGeom& G, const LubArgs& L) {
  const double dist0 = L.mingap_w + G.rad;
  double s = 0.0;
  double dist = yd, hgap = dist - G.rad;
  if (dist < dist0) {
    const double d1 = hgap < 0.0 ? -hgap : 0.0, t = (dist0 - dist) / L.mingap_w;
    s += (t * t / L.stf0_w + d1 / L.mingap_w / L.stf1_w) * L.fpc;
  }
  dist = (double)G.ny - yd;
  hgap = dist - G.rad;
  if (dist < dist0) {
    const double d1 = hgap < 0.0 ? -hgap : 0.0, t = (dist0 - dist) / L.mingap_w;
    s -= (t * t / L.stf0_w + d1 / L.mingap_w / L.stf1_w) * L.fpc;
  }
  return s;
}

__device__ void lub_block(int a, const double* __restrict__ ypglb, int np, const PGeom& G,
                          const LubArgs& L) {
  __shared__ double red[3][PB / 64];
  const double dist0 = 2.0 * G.rad + L.mingap;
  const double pre = (dist0 + 1e-6) * (dist0 + 1e-6);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int b = threadIdx.x; b < np; b += PB) {
    if (b == a) continue;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const double xc = ypglb[3 * lo], yc = ypglb[3 * lo + 1], zc = ypglb[3 * lo + 2];
    const double xb = near_image(ypglb[3 * hi], xc, G.nx);
    const double yb = G.walls_y ? ypglb[3 * hi + 1] : near_image(ypglb[3 * hi + 1], yc, G.ny);
    const double zb = near_image(ypglb[3 * hi + 2], zc, G.nz);
    const double dx = xc - xb, dy = yc - yb, dz = zc - zb;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < pre)) continue;
    const double dist = sqrt(d2);
    const double fpd = lub_fpd(dist, G.rad, L.mingap, L.fpc, L.stf0, L.stf1);
    const double sgn = (a == lo) ? 1.0 : -1.0;
    fx += sgn * (fpd * dx);
    fy += sgn * (fpd * dy);
    fz += sgn * (fpd * dz);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v[3] = {fx, fy, fz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_down(v[c], o, 64);
    if (lane == 0) red[c][w] = v[c];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < PB / 64; ++q) s += red[threadIdx.x][q];
    if (G.walls_y && threadIdx.x == 1) s += wall_lub(ypglb[3 * a + 1], G, L);
    L.flubp[3 * a + threadIdx.x] = s;
  }
}

// Pair force of the pair (a, b) on a (the reference evaluates each pair from
// the lower id's side, so both partners get bitwise opposite forces); false
// if the pair is out of range.
__device__ __forceinline__ bool lub_pair(const double* __restrict__ ypglb, int a, int b,
                                         const PGeom& G, const LubArgs& L, double pre, double& fx,
                                         double& fy, double& fz) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const double xc = ypglb[3 * lo], yc = ypglb[3 * lo + 1], zc = ypglb[3 * lo + 2];
  const double xb = near_image(ypglb[3 * hi], xc, G.nx);
  const double yb = G.walls_y ? ypglb[3 * hi + 1] : near_image(ypglb[3 * hi + 1], yc, G.ny);
  const double zb = near_image(ypglb[3 * hi + 2], zc, G.nz);
  const double dx = xc - xb, dy = yc - yb, dz = zc - zb;
  const double d2 = dx * dx + dy * dy + dz * dz;
  if (!(d2 < pre)) return false;
  const double fpd = lub_fpd(sqrt(d2), G.rad, L.mingap, L.fpc, L.stf0, L.stf1);
  const double sgn = (a == lo) ? 1.0 : -1.0;
  fx = sgn * (fpd * dx);
  fy = sgn * (fpd * dy);
  fz = sgn * (fpd * dz);
  return true;
}

// beads_lubforce of particle a from the cells: partners in the 27 cells
// around it, summed in increasing id (a fixed order); then the walls.  One
// pass over the cells lists the partners in range, sorted and each once (a
// cell grid of fewer than 3 cells on an axis meets a cell twice); only a
// sphere with more than kLubList partners falls back to finding each next
// id by a further pass.
constexpr int kLubList = 32;
__device__ void lub_cells(int a, const double* __restrict__ ypglb, int np, const PGeom& G,
                          const LubArgs& L, int32_t* __restrict__ status) {
  if (a >= np) return;
  const CellArgs& C = L.cells;
  const double dist0 = 2.0 * G.rad + L.mingap;
  const double pre = (dist0 + 1e-6) * (dist0 + 1e-6);
  int ca, cb, cc;
  cell_coords(ypglb, a, G, C, ca, cb, cc);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int ids[kLubList];
  int m = 0;
  bool full = false;
  for (int k = 0; k < 27; ++k) {
    const int cell = cell_neighbour(C, G, ca, cb, cc, k);
    if (cell < 0) continue;
    const int n = min(C.cnt[cell], kCellCap);
    for (int e = 0; e < n; ++e) {
      const int b = C.list[cell * kCellCap + e];
      if (b == a) continue;
      double fx, fy, fz;
      if (!lub_pair(ypglb, a, b, G, L, pre, fx, fy, fz)) continue;
      int pos = m;
      while (pos > 0 && ids[pos - 1] > b) --pos;
      if (pos > 0 && ids[pos - 1] == b) continue;   // met through a repeated cell
      if (m == kLubList) {
        full = true;
        continue;
      }
      for (int i = m; i > pos; --i) ids[i] = ids[i - 1];
      ids[pos] = b;
      ++m;
    }
  }
  if (!full) {
    for (int i = 0; i < m; ++i) {
      double fx, fy, fz;
      lub_pair(ypglb, a, ids[i], G, L, pre, fx, fy, fz);
      sx += fx;
      sy += fy;
      sz += fz;
    }
    if (G.walls_y) sy += wall_lub(ypglb[3 * a + 1], G, L);
    L.flubp[3 * a] = sx;
    L.flubp[3 * a + 1] = sy;
    L.flubp[3 * a + 2] = sz;
    return;
  }
  int last = -1;
  for (int guard = 0; guard < 27 * kCellCap; ++guard) {
    int next = 0x7fffffff;
    for (int k = 0; k < 27; ++k) {
      const int cell = cell_neighbour(C, G, ca, cb, cc, k);
      if (cell < 0) continue;
      const int n = min(C.cnt[cell], kCellCap);
      for (int e = 0; e < n; ++e) {
        const int b = C.list[cell * kCellCap + e];
        if (b <= last || b >= next || b == a) continue;
        double fx, fy, fz;
        if (lub_pair(ypglb, a, b, G, L, pre, fx, fy, fz)) next = b;
      }
    }
    if (next == 0x7fffffff) break;
    double fx, fy, fz;
    lub_pair(ypglb, a, next, G, L, pre, fx, fy, fz);
    sx += fx;
    sy += fy;
    sz += fz;
    last = next;
  }
  if (G.walls_y) sy += wall_lub(ypglb[3 * a + 1], G, L);
  L.flubp[3 * a] = sx;
  L.flubp[3 * a + 1] = sy;
  L.flubp[3 * a + 2] = sz;
}

__global__ void __launch_bounds__(PB)
    k_lubforce(const double* __restrict__ ypglb, int np, PGeom G, LubArgs L, int32_t* __restrict__ status) {
  if (L.cells.cnt)
    lub_cells((int)(blockIdx.x * PB + threadIdx.x), ypglb, np, G, L, status);
  else if ((int)blockIdx.x < np)
    lub_block(blockIdx.x, ypglb, np, G, L);
}


// Bounce-back of one link (partlib.f90:1363-1703): the wall point and its
// rigid-body velocity, then the interpolated value f9 sent back into the
// fluid and ff1 = f_q(x) post-collision; ff1 + f9 is the link's momentum
// exchange.  (xx0, yy0, zz0) is the wall point relative to the centre.
// With `gather` false only the geometry is evaluated.  With kSub the values
// read at x itself can be replaced: f_q(x) by sub_q (has_q) and, for
// alpha > 1/2, f_qbar(x) by sub_qbar (has_qbar).
template <typename T, bool kSub = false>
__device__ __forceinline__ void link_bounce(const T* A, const Grid& g, const PGeom& G,
                                            const uint32_t* __restrict__ mask, uint32_t n, int q,
                                            double alpha, double xcp, double ycp, double zcp,
                                            double w1, double w2, double w3, double o1, double o2,
                                            double o3, double ww1, double ww2, bool gather,
                                            double& ff1, double& f9, double& xx0, double& yy0,
                                            double& zz0, bool has_q = false, double sub_q = 0.0,
                                            bool has_qbar = false, double sub_qbar = 0.0) {
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  const int ix = cx(q), iy = cy(q), iz = cz(q);
  const double xp = (double)(x + 1) - 0.5, yp = (double)(y + 1) - 0.5, zp = (double)(z + 1) - 0.5;
  const double xc = near_image(xcp, xp, g.nx);
  const double yc = G.walls_y ? ycp : near_image(ycp, yp, g.ny);
  const double zc = near_image(zcp, zp, g.nz);
  xx0 = xp + (double)ix * alpha - xc;
  yy0 = yp + (double)iy * alpha - yc;
  zz0 = zp + (double)iz * alpha - zc;
  if (!gather) return;
  const double uwx = w1 + o2 * zz0 - o3 * yy0;
  const double uwy = w2 + o3 * xx0 - o1 * zz0;
  const double uwz = w3 + o1 * yy0 - o2 * xx0;
  const double uwpro = uwx * (double)ix + uwy * (double)iy + uwz * (double)iz;
  // x - c and x - 2c (periodic; beyond a wall counts as solid)
  const int y1 = y - iy, y2 = y - 2 * iy;
  const bool out1 = g.walls_y && (y1 < 0 || y1 >= g.ny);
  const bool out2 = g.walls_y && (y2 < 0 || y2 >= g.ny);
  const uint32_t m1 = out1 ? n : (uint32_t)wrap1(x - ix, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap1(y1, g.ny) + (uint32_t)g.ny * (uint32_t)wrap1(z - iz, g.nz));
  const uint32_t m2 = out2 ? n : (uint32_t)wrap1(x - 2 * ix, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap1(y2, g.ny) + (uint32_t)g.ny * (uint32_t)wrap1(z - 2 * iz, g.nz));
  // the values the interpolation may use are gathered together with the
  // solidity tests of x - c and x - 2c (one dependent load stage less; a
  // value at a solid node is read but not used): the alpha <= 1/2 branch
  // takes f_q at x, x - c and x - 2c, the alpha > 1/2 branch f_q at x and
  // f_qbar at x and x - c
  const int qo = opp(q);
  const double wt = q <= 6 ? ww1 : ww2;
  const bool lo = alpha <= 0.5;
  ff1 = post_value(A, g, q, n);
  double v1 = post_value(A, g, lo ? q : qo, lo ? m1 : n);
  const double v2 = post_value(A, g, lo ? q : qo, lo ? m2 : m1);
  if (kSub) {
    if (has_q) ff1 = sub_q;
    if (has_qbar && !lo) v1 = sub_qbar;
  }
  const bool sol1 = out1 || bit_at(mask, m1);
  const bool sol2 = out2 || bit_at(mask, m2);
  if (lo)
    f9 = bounce_back_value(alpha, ff1, sol1 ? 0.0 : v1, (sol1 || sol2) ? 0.0 : v2, sol1, sol1 || sol2,
                           0.0, 0.0, wt, uwpro);
  else
    f9 = bounce_back_value(alpha, ff1, 0.0, 0.0, sol1, true, v1, sol1 ? 0.0 : v2, wt, uwpro);
}

// Contact links in the reference's order (one workgroup).  The reference
// runs its links in (k, j, i, q) loop order and stores each value at once into
// the solid node x + c (partlib.f90:1363-1621).  A link reads f_q and, for
// alpha > 1/2, f_qbar at its own node x; every other value it uses is at a
// fluid node, which no link writes.  So only a contact link (x solid for
// another sphere) can read a stored value: the one its writer stored at x if
// the loop ran that writer earlier and the store reached x in place, not
// through the halo copy of a periodic y or z face, which the reference
// applies after the loop (partlib.f90:1605-1621, 2101-2171).  Non-contact
// links keep their snapshot values; the contact links' values are the fixed
// point of f9(C) = F(the values of C's earlier writers), found by Jacobi
// passes over the contact list until no value changes (writers run earlier,
// so the dependencies are acyclic and pass m settles chains of length m).
// The writer of f_q(x) is the link of x's owner at x + c_q along -c_q, the
// writer of f_qbar(x) its link at x - c_q along c_q.
constexpr int kOrderBatch = 16;
template <typename T>
__device__ void bb_contact_ordered(const BounceArgs& a) {
  __shared__ int changed;
  const T* A = (const T*)a.A;
  const Grid& g = a.g;
  const int64_t total = min((int64_t)*a.ct_count, a.cap);
  if (total > a.ord_cap) {
    if (threadIdx.x == 0) atomicAdd(&a.status[ST_ORDER_CAP], 1);
    return;
  }
  for (int64_t e = threadIdx.x; e < total; e += PB) {
    const uint64_t ent = a.ct_list[e];
    const int p = (int)(ent >> 40);
    const int64_t l = (int64_t)(ent & ((1ull << 40) - 1));
    OrderEntry o;
    o.dep[0] = o.dep[1] = -1;
    o.val[0] = o.val[1] = o.f9 = o.ff1 = 0.0;
    const uint32_t n = l < a.cap ? a.lk_node[l] : 0u;
    const int q = l < a.cap ? a.lk_dir[l] : 0;
    if (p < a.np && l < a.cap && n < g.ncell && q >= 1 && q < NPOP) {
      const double alpha = a.lk_alpha[l];
      double ff1, f9, xx0, yy0, zz0;
      link_bounce<T>(A, g, a.G, a.mask, n, q, alpha, a.ypglb[3 * p], a.ypglb[3 * p + 1],
                     a.ypglb[3 * p + 2], a.wp[3 * p], a.wp[3 * p + 1], a.wp[3 * p + 2], a.omgp[3 * p],
                     a.omgp[3 * p + 1], a.omgp[3 * p + 2], a.ww1, a.ww2, true, ff1, f9, xx0, yy0, zz0);
      a.lk_out[l] = make_double2(f9, ff1);
      int x, y, z;
      uint32_t row;
      node_coords(n, g, x, y, z, row);
      const int own = a.isn[n] - 1;
      for (int s = 0; s < 2 && own >= 0 && own < a.np; ++s) {
        if (s == 1 && !(alpha > 0.5)) break;
        const int sg = s == 0 ? 1 : -1;
        const int xw = x + sg * cx(q), yw = y + sg * cy(q), zw = z + sg * cz(q);
        if (yw < 0 || yw >= g.ny || zw < 0 || zw >= g.nz) continue;   // stored through a halo
        const uint32_t nw = (uint32_t)wrap1(xw, g.nx) + (uint32_t)g.nx * ((uint32_t)yw + (uint32_t)g.ny * (uint32_t)zw);
        if (nw > n) continue;   // the loop runs the writer later
        const int qw = s == 0 ? opp(q) : q;
        const int64_t lo = a.off[(int64_t)own * a.nseg];
        const int64_t hi = min(a.off[(int64_t)(own + 1) * a.nseg], a.cap);
        // the first match, kOrderBatch independent loads per round trip
        int64_t lw = -1;
        for (int64_t k0 = lo; k0 < hi && lw < 0; k0 += kOrderBatch) {
          uint32_t nn[kOrderBatch];
          int dd[kOrderBatch];
#pragma unroll
          for (int j = 0; j < kOrderBatch; ++j) {
            const bool in = k0 + j < hi;
            nn[j] = in ? a.lk_node[k0 + j] : ~0u;
            dd[j] = in ? (int)a.lk_dir[k0 + j] : -1;
          }
#pragma unroll
          for (int j = 0; j < kOrderBatch; ++j)
            if (lw < 0 && nn[j] == nw && dd[j] == qw) lw = k0 + j;
        }
        if (lw < 0) continue;
        if (bit_at(a.mask, nw)) {
          o.dep[s] = lw | kOrderContact;
        } else {
          double wf1, wf9, wx, wy, wz;
          link_bounce<T>(A, g, a.G, a.mask, nw, qw, a.lk_alpha[lw], a.ypglb[3 * own],
                         a.ypglb[3 * own + 1], a.ypglb[3 * own + 2], a.wp[3 * own], a.wp[3 * own + 1],
                         a.wp[3 * own + 2], a.omgp[3 * own], a.omgp[3 * own + 1], a.omgp[3 * own + 2],
                         a.ww1, a.ww2, true, wf1, wf9, wx, wy, wz);
          o.dep[s] = lw;
          o.val[s] = (double)(T)wf9;   // what the writer stores
        }
      }
    }
    a.ord[e] = o;
  }
  for (int64_t pass = 0; pass <= total + 1; ++pass) {
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();   // phase-1 entries and the previous pass's values are visible
    for (int64_t e = threadIdx.x; e < total; e += PB) {
      OrderEntry& o = a.ord[e];
      if (o.dep[0] < 0 && o.dep[1] < 0) continue;
      const uint64_t ent = a.ct_list[e];
      const int p = (int)(ent >> 40);
      const int64_t l = (int64_t)(ent & ((1ull << 40) - 1));
      double v[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int64_t d = o.dep[s];
        v[s] = d < 0 ? 0.0 : ((d & kOrderContact) ? (double)(T)a.lk_out[d & ~kOrderContact].x : o.val[s]);
      }
      double ff1, f9, xx0, yy0, zz0;
      link_bounce<T, true>(A, g, a.G, a.mask, a.lk_node[l], a.lk_dir[l], a.lk_alpha[l], a.ypglb[3 * p],
                           a.ypglb[3 * p + 1], a.ypglb[3 * p + 2], a.wp[3 * p], a.wp[3 * p + 1],
                           a.wp[3 * p + 2], a.omgp[3 * p], a.omgp[3 * p + 1], a.omgp[3 * p + 2],
                           a.ww1, a.ww2, true, ff1, f9, xx0, yy0, zz0, o.dep[0] >= 0, v[0],
                           o.dep[1] >= 0, v[1]);
      o.f9 = f9;
      o.ff1 = ff1;
    }
    __syncthreads();   // every next value is computed before any is stored
    for (int64_t e = threadIdx.x; e < total; e += PB) {
      const OrderEntry& o = a.ord[e];
      if (o.dep[0] < 0 && o.dep[1] < 0) continue;
      const int64_t l = (int64_t)(a.ct_list[e] & ((1ull << 40) - 1));
      const double2 cur = a.lk_out[l];
      if (__double_as_longlong(cur.x) != __double_as_longlong(o.f9) ||
          __double_as_longlong(cur.y) != __double_as_longlong(o.ff1)) {
        a.lk_out[l] = make_double2(o.f9, o.ff1);
        changed = 1;
      }
    }
    __syncthreads();
    const int more = changed;
    __syncthreads();   // every thread has read the flag before it is reset
    if (!more) break;
  }
}

// Pre-pass of the bounce-back.  Contact links (node x solid for another
// sphere, spheres almost touching) read post-collision values at a solid
// node, which the bounce-back grid overwrites: their values are taken first,
// from the untouched field, into lk_out (f9, ff1).  Every other link reads
// fluid nodes only and writes only solid ones, so the bounce-back grid can
// scatter as it goes; all links see the one post-collision snapshot.  The
// first nct blocks stride over the contact list; the blocks after them
// compute beads_lubforce (which the bounce-back does not read).
template <typename T>
__device__ __forceinline__ void bb_contact_body(int b_, const BounceArgs& a) {
  if (b_ >= a.nct) {
    const int r = b_ - a.nct;
    if (a.lub.cells.cnt)
      lub_cells(r * PB + (int)threadIdx.x, a.ypglb, a.np, a.G, a.lub, a.status);
    else
      lub_block(r, a.ypglb, a.np, a.G, a.lub);
    return;
  }
  if (a.ref_order) {
    if (b_ == 0) bb_contact_ordered<T>(a);
    return;
  }
  const int nblk = a.nct;
  const T* A = (const T*)a.A;
  const int64_t total = min((int64_t)*a.ct_count, a.cap);
  for (int64_t e = (int64_t)b_ * PB + threadIdx.x; e < total; e += (int64_t)nblk * PB) {
    const uint64_t ent = a.ct_list[e];
    const int p = (int)(ent >> 40);
    const int64_t l = (int64_t)(ent & ((1ull << 40) - 1));
    if (p >= a.np || l >= a.cap) continue;   // never true for a consistent list
    const uint32_t n = a.lk_node[l];
    const int q = a.lk_dir[l];
    if (n >= a.g.ncell || q < 1 || q >= NPOP) continue;
    double ff1 = 0.0, f9 = 0.0, xx0, yy0, zz0;
    link_bounce<T>(A, a.g, a.G, a.mask, n, q, a.lk_alpha[l], a.ypglb[3 * p], a.ypglb[3 * p + 1],
                   a.ypglb[3 * p + 2], a.wp[3 * p], a.wp[3 * p + 1], a.wp[3 * p + 2], a.omgp[3 * p],
                   a.omgp[3 * p + 1], a.omgp[3 * p + 2], a.ww1, a.ww2, true, ff1, f9, xx0, yy0, zz0);
    a.lk_out[l] = make_double2(f9, ff1);
  }
}

// beads_collision (partlib.f90:1363-1703), one unit (1/split of a segment's
// links) per block: the bounce-back value of every link, written at once into
// the opposite population of the solid neighbour x + c (partlib.f90:1605-1621),
// and the unit's momentum-exchange force and torque (FP64, fixed reduction
// order) into seg_ft.  Blocks past the units (fused step) reset one segment
// of the map buffer the coming beads_links rebuilds (built two steps ago).
template <typename T>
__device__ __forceinline__ void bounce_back_body(int b_, const BounceArgs& a) {
  __shared__ double red[6][PB / 64];
  T* A = (T*)a.A;
  const Grid& g = a.g;
  const int u = b_;
  const int nunits = a.nsegs * a.split;
  if (u >= nunits) {
    clear_segment(u - nunits, a.nseg, a.seg_cap, a.clr_list, a.clr_cnt, a.clr_isn, a.clr_mask);
    return;
  }
  const int b = u / a.split, h = u - b * a.split;
  const int p = b / a.nseg;
  const int64_t s0 = a.off[b], s1 = min(a.off[b + 1], a.cap);
  const int64_t per = (max(s1 - s0, (int64_t)0) + a.split - 1) / a.split;
  const int64_t l0 = s0 + h * per;
  const int64_t e = min(s1, l0 + per);
  const double xcp = a.ypglb[3 * p], ycp = a.ypglb[3 * p + 1], zcp = a.ypglb[3 * p + 2];
  const double w1 = a.wp[3 * p], w2 = a.wp[3 * p + 1], w3 = a.wp[3 * p + 2];
  const double o1 = a.omgp[3 * p], o2 = a.omgp[3 * p + 1], o3 = a.omgp[3 * p + 2];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int contact = 0;
  for (int64_t l = l0 + threadIdx.x; l < e; l += PB) {
    const uint32_t n = a.lk_node[l];
    const int q = a.lk_dir[l];
    const double alpha = a.lk_alpha[l];
    if (n >= g.ncell || q < 1 || q >= NPOP) continue;  // never true for a consistent list
    const bool cont = bit_at(a.mask, n);
    double ff1 = 0.0, f9 = 0.0, xx0, yy0, zz0;
    link_bounce<T>(A, g, a.G, a.mask, n, q, alpha, xcp, ycp, zcp, w1, w2, w3, o1, o2, o3, a.ww1,
                   a.ww2, !cont, ff1, f9, xx0, yy0, zz0);
    if (cont) {
      const double2 v = a.lk_out[l];
      f9 = v.x;
      ff1 = v.y;
      ++contact;
    }
    // scatter: the post-collision value of opp(q) at the solid node x + c
    int x, y, z;
    uint32_t row;
    node_coords(n, g, x, y, z, row);
    const int yt = y + cy(q);
    if (!(g.walls_y && (yt < 0 || yt >= g.ny))) {
      const int xt = wrap1(x + cx(q), g.nx), zt = wrap1(z + cz(q), g.nz), ytw = wrap1(yt, g.ny);
      int pl;
      uint64_t s;
      if (g.walls_y)
        post_src<true>(g.lay, opp(q), xt, ytw, zt, g, pl, s);
      else
        post_src<false>(g.lay, opp(q), xt, ytw, zt, g, pl, s);
      A[(uint64_t)pl * g.pitch + s] = (T)f9;
    }
    const double dff = ff1 + f9;
    const double dxm = dff * (double)cx(q), dym = dff * (double)cy(q), dzm = dff * (double)cz(q);
    acc[0] += dxm;
    acc[1] += dym;
    acc[2] += dzm;
    acc[3] += dzm * yy0 - dym * zz0;
    acc[4] += dxm * zz0 - dzm * xx0;
    acc[5] += dym * xx0 - dxm * yy0;
  }
  // wavefront tree reduction (fixed order), then the four wave partials
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double v = acc[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[c][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PB / 64; ++q) v += red[threadIdx.x][q];
    a.seg_ft[6ll * u + threadIdx.x] = v;
  }
  if (contact) atomicAdd(&a.status[ST_CONTACT_LINKS], contact);
}

// fHIp / torqp of each particle: its units' sums (nunit per particle), in
// unit order.
__device__ __forceinline__ void seg_sum(const double* __restrict__ seg_ft, int nunit, int p,
                                        double (&f)[3], double (&t)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = 0.0, u = 0.0;
    for (int s = 0; s < nunit; ++s) {
      v += seg_ft[6ll * ((int64_t)p * nunit + s) + c];
      u += seg_ft[6ll * ((int64_t)p * nunit + s) + 3 + c];
    }
    f[c] = v;
    t[c] = u;
  }
}

__global__ void __launch_bounds__(256)
    k_seg_reduce(const double* __restrict__ seg_ft, int nseg, int np, double* __restrict__ fHIp,
                 double* __restrict__ torqp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double f[3], t[3];
  seg_sum(seg_ft, nseg, p, f, t);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fHIp[3 * p + c] = f[c];
    torqp[3 * p + c] = t[c];
  }
}


__device__ __forceinline__ bool wrap_particle(double* yp, double* ypglb, double* ypglbp, int p,
                                              const PGeom& G) {
  const int ns[3] = {G.nx, G.ny, G.nz};
  bool bad = false;
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    const double v = yp[q], N = (double)ns[c];
    if (!(G.walls_y && c == 1)) {
      if (v >= N) yp[q] = yp[q] - N;
      if (v < 0.0) yp[q] = yp[q] + N;
    }
    if (!sane_coord(yp[q], ns[c])) bad = true;
    ypglbp[q] = ypglb[q];
    ypglb[q] = yp[q];
  }
  return bad;
}

// beads_move (partlib.f90:2295-2394): second-order Newton-Euler update.  With
// `wrap` set the periodic wrap of beads_redistribute (partlib.f90:2600-2626)
// follows in the same thread; with seg_ft set the hydrodynamic force and
// torque are first summed from the bounce-back segments (and stored).  Every
// input is loaded before the first store (the argument arrays are distinct,
// but the compiler cannot know it).
__device__ __forceinline__ void move_body(int b_, MoveArgs m, int np, int first, int wrap, PGeom G,
                                          int32_t* __restrict__ status,
                                          const double* __restrict__ seg_ft, int nseg) {
  const int p = b_ * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double dt = 1.0;
  const double bf[3] = {m.bfx, m.bfy, m.bfz};
  const double pf[3] = {m.pfx, m.pfy, m.pfz};
  double yp[3], wp[3], omgp[3], fhi[3], flub[3], torq[3], fpp[3], tpp[3], dwdt[3], domgdt[3], th[3];
  double yg[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    yp[c] = m.yp[q];
    yg[c] = m.ypglb[q];
    wp[c] = m.wp[q];
    omgp[c] = m.omgp[q];
    if (!seg_ft) {
      fhi[c] = m.fHIp[q];
      torq[c] = m.torqp[q];
    }
    flub[c] = m.flubp[q];
    fpp[c] = m.forcepp[q];
    tpp[c] = m.torqpp[q];
    dwdt[c] = m.dwdt[q];
    domgdt[c] = m.domgdt[q];
    th[c] = first ? 0.0 : m.thetap[q];
  }
  if (seg_ft) seg_sum(seg_ft, nseg, p, fhi, torq);
  double wpp[3], omgpp[3], fp[3], dwn[3], don[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    wpp[c] = wp[c];
    omgpp[c] = omgp[c];
    double F = fhi[c] + flub[c];
    if (c == 1) F = F - m.volp * m.rhog;
    if (bf[c] != 0.0) F = F + m.volp * bf[c];
    if (pf[c] != 0.0) F = F + m.volp * pf[c];
    fp[c] = F;
    if (first) fpp[c] = F;
    dwn[c] = 0.5 * (F + fpp[c]) / m.amp;
    if (first) dwdt[c] = dwn[c];
    wp[c] = wpp[c] + 0.5 * dt * (dwdt[c] + dwn[c]);
  }
  bool fast = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d = 0.5 * dt * (wp[c] + wpp[c]);
    yp[c] = yp[c] + d;
    fast = fast || !(fabs(d) < 1.0);
  }
  // the overlapped step assumes a centre moves less than one node per step
  if (fast) atomicAdd(&status[ST_FAST], 1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (first) tpp[c] = torq[c];
    don[c] = 0.5 * (torq[c] + tpp[c]) / m.aip;
    if (first) domgdt[c] = don[c];
    omgp[c] = omgpp[c] + 0.5 * dt * (domgdt[c] + don[c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) th[c] = th[c] + 0.5 * dt * (omgp[c] + omgpp[c]);
  bool bad = false;
  double ygp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ygp[c] = yg[c];
    if (wrap) {
      // beads_redistribute: wrap into the periodic box (not across a wall),
      // ypglbp <- ypglb, ypglb <- yp
      const int n = c == 0 ? G.nx : (c == 1 ? G.ny : G.nz);
      const double v = yp[c], N = (double)n;
      if (!(G.walls_y && c == 1)) {
        if (v >= N) yp[c] = yp[c] - N;
        if (v < 0.0) yp[c] = yp[c] + N;
      }
      if (!sane_coord(yp[c], n)) bad = true;
      ygp[c] = yg[c];
      yg[c] = yp[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    m.wpp[q] = wpp[c];
    m.omgpp[q] = omgpp[c];
    m.forcep[q] = fp[c];
    m.wp[q] = wp[c];
    m.yp[q] = yp[c];
    m.omgp[q] = omgp[c];
    m.thetap[q] = th[c];
    m.dwdt[q] = dwn[c];
    m.domgdt[q] = don[c];
    m.forcepp[q] = fp[c];
    m.torqpp[q] = torq[c];
    if (seg_ft) {
      m.fHIp[q] = fhi[c];
      m.torqp[q] = torq[c];
    }
    if (wrap) {
      m.ypglbp[q] = ygp[c];
      m.ypglb[q] = yg[c];
    }
  }
  if (bad) atomicAdd(&status[ST_NONFINITE], 1);
}

// beads_redistribute, one-rank content (partlib.f90:2600-2626): periodic wrap,
// ypglbp <- ypglb, ypglb <- yp.
__global__ void __launch_bounds__(256)
    k_wrap(double* __restrict__ yp, double* __restrict__ ypglb, double* __restrict__ ypglbp, int np,
           PGeom G, int32_t* __restrict__ status) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (wrap_particle(yp, ypglb, ypglbp, p, G)) atomicAdd(&status[ST_NONFINITE], 1);
}

// ovlpchk (partlib.f90:118-171): one workgroup per particle flags it if any
// later particle is closer than 2r.
__global__ void __launch_bounds__(PB)
    k_overlap(const double* __restrict__ ypglb, int np, PGeom G, int istep, CellArgs cells,
              int32_t* __restrict__ status) {
  const int a = blockIdx.x;
  if (a >= np) return;
  const bool hit = block_overlap(ypglb, a, np, G, cells);
  if (threadIdx.x == 0 && hit) {
    atomicAdd(&status[ST_OVERLAP], 1);
    atomicCAS(&status[ST_OVERLAP_STEP], 0, istep);
  }
}


// beads_filling, pass 1 (partlib.f90:2846-2849): the nodes that were solid
// for particle p in the previous map (p owns them there) and are fluid in the
// new one, from the previous build's solid lists.  One workgroup per segment.
__device__ void fill_find_segment(int b, const FillArgs& fa, int nseg, int32_t* __restrict__ status) {
  const int p = b / nseg;
  const int cnt = fa.sol_cnt_old[b];
  const uint32_t* lst = fa.sol_old + (size_t)b * fa.seg_cap;
  const int lane = threadIdx.x & 63;
  int nfill = 0;
  for (int e0 = 0; e0 < cnt; e0 += PB) {
    const int e = e0 + (int)threadIdx.x;
    const uint32_t n = e < cnt ? lst[e] : 0u;
    // the new map's bit first: nearly every node stays solid, and only the
    // uncovered ones need the owner test (a node of several spheres is
    // listed once, by its owner)
    const bool fill = e < cnt && !bit_at(fa.mask_new, n) && fa.isn_old[n] == p + 1;
    // one atomic per wavefront for the list slots
    const uint64_t bal = __ballot(fill);
    if (bal == 0ull) continue;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    int slot0 = 0;
    if (lane == leader) slot0 = atomicAdd(fa.count, __popcll(bal));
    slot0 = __builtin_amdgcn_readlane(slot0, leader);
    if (fill) {
      const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      const int slot = slot0 + __popcll(bal & below);
      if (slot < fa.cap) fa.list[slot] = ((uint64_t)(uint32_t)p << 32) | n;
      ++nfill;
    }
  }
  if (nfill) atomicAdd(&status[ST_REFILL], nfill);
}

__global__ void __launch_bounds__(PB)
    k_fill_find(FillArgs fa, int nseg, int32_t* __restrict__ status) {
  fill_find_segment(blockIdx.x, fa, nseg, status);
}

// beads_filling, pass 2 (partlib.f90:2851-3089): one wavefront per listed
// node; lane l examines neighbour l+1 (fluid in both maps?  its density; the
// non-equilibrium part if it lies along the back-traced normal), then lane i
// writes population i of the node: equilibrium at the rigid-body velocity
// (new w, omega; old centre) and the neighbours' mean density, plus the
// non-equilibrium part.  Written into the pull sources of the node.
template <typename T, bool WALLS>
__device__ __forceinline__ void fill_apply_body(int b_, T* __restrict__ A, Grid g, FillArgs fa, int nblk) {
  __shared__ double fneq[4][NPOP];
  __shared__ double srho[4][NPOP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t nwaves = (int64_t)nblk * 4;
  const int64_t total = min((int64_t)*fa.count, fa.cap);
  // the trip count is the same for every wavefront of the block (barriers inside)
  for (int64_t base = (int64_t)b_ * 4; base < total; base += nwaves) {
    const int64_t e = base + w;
    const bool active = e < total;
    const uint64_t ent = active ? fa.list[e] : 0ull;
    const int id = (int)(ent >> 32);
    const uint32_t n = (uint32_t)(ent & 0xffffffffu);
    int i, j, k;
    uint32_t row;
    node_coords(n, g, i, j, k, row);
    const double xp = (double)(i + 1) - 0.5, yp = (double)(j + 1) - 0.5, zp = (double)(k + 1) - 0.5;
    const double xc = near_image(fa.ypglbp[3 * id] * 1.0, xp, g.nx);
    const double yc = WALLS ? fa.ypglbp[3 * id + 1] : near_image(fa.ypglbp[3 * id + 1] * 1.0, yp, g.ny);
    const double zc = near_image(fa.ypglbp[3 * id + 2] * 1.0, zp, g.nz);
    const double w1 = -0.5 * (fa.wp[3 * id] * 1.0 + fa.wpp[3 * id] * 1.0);
    const double w2 = -0.5 * (fa.wp[3 * id + 1] * 1.0 + fa.wpp[3 * id + 1] * 1.0);
    const double w3 = -0.5 * (fa.wp[3 * id + 2] * 1.0 + fa.wpp[3 * id + 2] * 1.0);
    const double o1 = -0.5 * (fa.omgp[3 * id] * 1.0 + fa.omgpp[3 * id] * 1.0);
    const double o2 = -0.5 * (fa.omgp[3 * id + 1] * 1.0 + fa.omgpp[3 * id + 1] * 1.0);
    const double o3 = -0.5 * (fa.omgp[3 * id + 2] * 1.0 + fa.omgpp[3 * id + 2] * 1.0);
    const int ipmx = refill_direction(xp, yp, zp, xc, yc, zc, w1, w2, w3, o1, o2, o3, fa.rad);
    // lane l < 18: neighbour d = l + 1
    bool fluid = false;
    if (active && lane < NPOP - 1) {
      const int d = lane + 1;
      const int a = i + cx(d), bb = j + cy(d), c = k + cz(d);
      const bool out = WALLS && (bb < 0 || bb >= g.ny);
      if (!out) {
        const uint32_t q = (uint32_t)wrap1(a, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap1(bb, g.ny) + (uint32_t)g.ny * (uint32_t)wrap1(c, g.nz));
        fluid = !bit_at(fa.mask_new, q) && !bit_at(fa.mask_old, q);
        double s = 0.0;
        if (fluid) {
          int qx, qy, qz;
          uint32_t r2;
          node_coords(q, g, qx, qy, qz, r2);
          double f9[NPOP];
          load_pulled<T, double, WALLS>(A, g, qx, qy, qz, q, f9);
          for (int ii = 0; ii < NPOP; ++ii) s = s + f9[ii];
          if (d == ipmx) {
            double rho9 = 0.0, u9 = 0.0, v9 = 0.0, w9 = 0.0;
            for (int ip = 0; ip < NPOP; ++ip) {
              rho9 = rho9 + f9[ip];
              u9 = u9 + (double)cx(ip) * f9[ip];
              v9 = v9 + (double)cy(ip) * f9[ip];
              w9 = w9 + (double)cz(ip) * f9[ip];
            }
            // feq_point's expressions one direction at a time (no array of 19)
            const double usqr = 1.5 * (u9 * u9 + v9 * v9 + w9 * w9);
            fneq[w][0] = f9[0] - fa.ww0 * (rho9 - usqr);
            for (int ip = 1; ip < NPOP; ++ip) {
              const double G = (double)cx(ip) * u9 + (double)cy(ip) * v9 + (double)cz(ip) * w9;
              fneq[w][ip] = f9[ip] - (ip <= 6 ? fa.ww1 : fa.ww2) * (rho9 + 3.0 * G + 4.5 * G * G - usqr);
            }
          }
        }
        srho[w][lane] = s;
      } else {
        srho[w][lane] = 0.0;
      }
    }
    const uint64_t fl = __ballot(fluid);
    const bool have_neq = (fl >> (ipmx - 1)) & 1ull;
    __syncthreads();
    if (active && lane < NPOP) {
      const int nghb = __popcll(fl);
      double rho9 = 0.0;
      for (int l = 0; l < NPOP - 1; ++l)
        if ((fl >> l) & 1ull) rho9 = rho9 + srho[w][l];
      rho9 = (nghb > 0) ? rho9 / (double)nghb : fa.rho0;
      const double xx0 = xp - xc, yy0 = yp - yc, zz0 = zp - zc;
      const double W1 = fa.wp[3 * id], W2 = fa.wp[3 * id + 1], W3 = fa.wp[3 * id + 2];
      const double O1 = fa.omgp[3 * id], O2 = fa.omgp[3 * id + 1], O3 = fa.omgp[3 * id + 2];
      const double u9 = W1 + O2 * zz0 - O3 * yy0;
      const double v9 = W2 + O3 * xx0 - O1 * zz0;
      const double w9 = W3 + O1 * yy0 - O2 * xx0;
      const double usqr = 1.5 * (u9 * u9 + v9 * v9 + w9 * w9);
      const int ip = lane;
      const double G = (double)cx(ip) * u9 + (double)cy(ip) * v9 + (double)cz(ip) * w9;
      const double feq = ip == 0 ? fa.ww0 * (rho9 - usqr)
                                 : (ip <= 6 ? fa.ww1 : fa.ww2) * (rho9 + 3.0 * G + 4.5 * G * G - usqr);
      const double fn = have_neq ? fneq[w][ip] : 0.0;
      store_phase<T, WALLS>(A, g, ip, i, j, k, n, true, fn + feq);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Launch forms of the step kernels.  Each kernel body takes its block index;
// k_* launches one box (arguments by value), kb_* one grid over several boxes
// at once: block b belongs to box c with off[c] <= b < off[c+1] and reads that
// box's arguments from the table t (the ensemble engine's per-case tables).
__device__ __forceinline__ int case_of(const int* __restrict__ off, int nc, int b) {
  int lo = 0, hi = nc;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// 8 waves per SIMD (the compiler fits the scan in 64 VGPRs without spills):
// a box of a few hundred spheres then runs its segments in one round
__global__ void __launch_bounds__(PB) __attribute__((amdgpu_waves_per_eu(8))) k_links_scan(ScanArgs a) {
  links_scan_body((int)blockIdx.x, a.ypglb, a.np, a.nseg, a.seg_cap, a.G, a.isn, a.mask, a.sol_list, a.sol_cnt, a.sh_node, a.sh_hits, a.sh_cnt, a.lk_cnt, a.ovl, a.istep, a.status, a.cells);
}
__global__ void __launch_bounds__(PB) __attribute__((amdgpu_waves_per_eu(8)))
    kb_links_scan(const ScanArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const ScanArgs a = t[c];
  links_scan_body((int)blockIdx.x - off[c], a.ypglb, a.np, a.nseg, a.seg_cap, a.G, a.isn, a.mask, a.sol_list, a.sol_cnt, a.sh_node, a.sh_hits, a.sh_cnt, a.lk_cnt, a.ovl, a.istep, a.status, a.cells);
}

// The same scan over many segments, in two grids of kScanChunk-segment
// chunks: each chunk's sums, then each chunk scans itself after adding the
// chunks before it.  Integer sums: the result equals the one-block scan.
constexpr int kScanPer = kScanChunk / PB;
__global__ void __launch_bounds__(PB) k_count_partials(ScanCountArgs a) {
  __shared__ int sh[PB / 64];
  const int b = blockIdx.x;
  int s = 0, ss = 0;
  for (int k = 0; k < kScanPer; ++k) {
    const int i = b * kScanChunk + k * PB + (int)threadIdx.x;
    if (i < a.nsegs) {
      s += a.cnt[i];
      ss += a.scnt[i];
    }
  }
  int t1, t2;
  block_exclusive_scan(s, sh, t1);
  block_exclusive_scan(ss, sh, t2);
  if (threadIdx.x == 0) {
    a.part[2 * b] = t1;
    a.part[2 * b + 1] = t2;
  }
}

__global__ void __launch_bounds__(PB) k_count_offsets(ScanCountArgs a) {
  __shared__ int sh[PB / 64];
  __shared__ int64_t red[2][PB];
  const int b = blockIdx.x;
  if (b == 0 && threadIdx.x == 0) {
    *a.fill_count = 0;   // the next beads_filling starts an empty list
    *a.ct_count = 0;     // the expansion lists this rebuild's contact links
  }
  // chunks before this one (and all chunks, for the totals)
  const int nb = gridDim.x;
  int64_t before = 0, tot = 0, stot = 0;
  for (int k = threadIdx.x; k < nb; k += PB) {
    const int64_t v = a.part[2 * k];
    tot += v;
    stot += a.part[2 * k + 1];
    if (k < b) before += v;
  }
  red[0][threadIdx.x] = before;
  red[1][threadIdx.x] = tot;
  __syncthreads();
  for (int o = PB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  const int64_t base = red[0][0], total = red[1][0];
  __syncthreads();
  // this chunk: thread t owns segments [t * kScanPer, (t + 1) * kScanPer)
  const int i0 = b * kScanChunk + (int)threadIdx.x * kScanPer;
  int s = 0;
  for (int k = 0; k < kScanPer; ++k)
    if (i0 + k < a.nsegs) s += a.cnt[i0 + k];
  int ct;
  int64_t acc = base + block_exclusive_scan(s, sh, ct);
  for (int k = 0; k < kScanPer; ++k) {
    if (i0 + k < a.nsegs) {
      a.off[i0 + k] = acc;
      acc += a.cnt[i0 + k];
    }
  }
  if (b == nb - 1) {
    // every chunk's solid count, by the same reduction
    red[1][threadIdx.x] = stot;
    __syncthreads();
    for (int o = PB / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[1][threadIdx.x] += red[1][threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      a.off[a.nsegs] = total;
      a.status[ST_NLINKS] = (int32_t)min(total, (int64_t)0x7fffffff);
      a.status[ST_NSOLID] = (int32_t)min(red[1][0], (int64_t)0x7fffffff);
      if (total > a.cap) atomicAdd(&a.status[ST_LINK_OVERFLOW], 1);
    }
  }
}

__global__ void __launch_bounds__(1024) k_scan_counts(ScanCountArgs a) {
  scan_counts_body((int)blockIdx.x, a.cnt, a.scnt, a.nsegs, a.off, a.cap, a.status, a.fill_count, a.ct_count);
}
__global__ void __launch_bounds__(1024)
    kb_scan_counts(const ScanCountArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const ScanCountArgs a = t[c];
  scan_counts_body((int)blockIdx.x - off[c], a.cnt, a.scnt, a.nsegs, a.off, a.cap, a.status, a.fill_count, a.ct_count);
}

__global__ void __launch_bounds__(PB) k_links_expand(ExpandArgs a) {
  links_expand_body((int)blockIdx.x, a.ypglb, a.np, a.nseg, a.seg_cap, a.G, a.sh_node, a.sh_hits, a.sh_cnt, a.off, a.cap, a.lk_node, a.lk_dir, a.lk_alpha, a.status, a.nsegs, a.fa, a.mask, a.ct_list, a.ct_count);
}
__global__ void __launch_bounds__(PB)
    kb_links_expand(const ExpandArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const ExpandArgs a = t[c];
  links_expand_body((int)blockIdx.x - off[c], a.ypglb, a.np, a.nseg, a.seg_cap, a.G, a.sh_node, a.sh_hits, a.sh_cnt, a.off, a.cap, a.lk_node, a.lk_dir, a.lk_alpha, a.status, a.nsegs, a.fa, a.mask, a.ct_list, a.ct_count);
}

template <typename T>
__global__ void __launch_bounds__(PB) k_bb_contact(BounceArgs a) {
  bb_contact_body<T>((int)blockIdx.x, a);
}
template <typename T>
__global__ void __launch_bounds__(PB)
    kb_bb_contact(const BounceArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const BounceArgs a = t[c];
  bb_contact_body<T>((int)blockIdx.x - off[c], a);
}

// Workgroups are dealt to the 8 XCDs round-robin in launch order.  Launch
// index L of the first nsegs blocks is mapped so that all segments of a
// particle run on one XCD (particle p on XCD p % 8): neighbouring segments
// read overlapping stencils and that XCD's L2 then holds them once.
__device__ __forceinline__ int xcd_segment(int L, int nsegs, int nseg) {
  const int group = 8 * nseg;
  const int full = (nsegs / group) * group;
  if (L >= full) return L;
  const int q = L / group, r = L - q * group;
  return (q * 8 + r % 8) * nseg + r / 8;
}

template <typename T>
__global__ void __launch_bounds__(PB) k_bounce_back(BounceArgs a) {
  bounce_back_body<T>(xcd_segment((int)blockIdx.x, a.nsegs * a.split, a.nseg * a.split), a);
}
template <typename T>
__global__ void __launch_bounds__(PB)
    kb_bounce_back(const BounceArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const BounceArgs a = t[c];
  bounce_back_body<T>((int)blockIdx.x - off[c], a);
}

__global__ void __launch_bounds__(256) k_move(MoveKArgs a) {
  move_body((int)blockIdx.x, a.m, a.np, a.first, a.wrap, a.G, a.status, a.seg_ft, a.nseg);
}
__global__ void __launch_bounds__(256)
    kb_move(const MoveKArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const MoveKArgs a = t[c];
  move_body((int)blockIdx.x - off[c], a.m, a.np, a.first, a.wrap, a.G, a.status, a.seg_ft, a.nseg);
}

// 4 waves per SIMD (at most 128 VGPRs): with the 1024 blocks of the grid
// every wavefront is resident at once
template <typename T, bool WALLS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
    k_fill_apply(FillApplyArgs a, TableJob tj) {
  if ((int)blockIdx.x >= a.nblk) {   // blocks past the refill: the next step's forcing table
    force_table_block(tj.M, tj.phy, tj.phz, tj.ny, tj.nz, tj.G, blockIdx.x - (uint32_t)a.nblk);
    return;
  }
  fill_apply_body<T, WALLS>((int)blockIdx.x, (T*)a.A, a.g, a.fa, a.nblk);
}
static_assert(sizeof(FillApplyArgs) + sizeof(TableJob) <= 4096, "refill kernel arguments over 4 KB");
template <typename T, bool WALLS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
    kb_fill_apply(const FillApplyArgs* __restrict__ t, const int* __restrict__ off, int nc) {
  const int c = case_of(off, nc, (int)blockIdx.x);
  const FillApplyArgs a = t[c];
  fill_apply_body<T, WALLS>((int)blockIdx.x - off[c], (T*)a.A, a.g, a.fa, a.nblk);
}

// trilinear velocity at the particle centres (initpartvel,
// partlib.f90:177-263 with trilinear :320-831); node (i) sits at i - 0.5.
__global__ void k_initpartvel(const double* __restrict__ ypglb, int np, Grid g,
                              const double* __restrict__ ux, const double* __restrict__ uy,
                              const double* __restrict__ uz, double* __restrict__ wp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double xc = ypglb[3 * p], yc = ypglb[3 * p + 1], zc = ypglb[3 * p + 2];
  if (!(sane_coord(xc, g.nx) && sane_coord(yc, g.ny) && sane_coord(zc, g.nz))) return;
  const int i = (int)(xc + 0.5), j = (int)(yc + 0.5), k = (int)(zc + 0.5);
  const double tx = xc + 0.5 - (double)i, ty = yc + 0.5 - (double)j, tz = zc + 0.5 - (double)k;
  const double tx0 = 1.0 - tx, ty0 = 1.0 - ty, tz0 = 1.0 - tz;
  const int i0 = wrap_any(i - 1, g.nx), i1 = wrap_any(i, g.nx);
  const int j0 = wrap_any(j - 1, g.ny), j1 = wrap_any(j, g.ny);
  const int k0 = wrap_any(k - 1, g.nz), k1 = wrap_any(k, g.nz);
  const double* us[3] = {ux, uy, uz};
  for (int c = 0; c < 3; ++c) {
    const double* u = us[c];
    auto at = [&](int a, int b, int d) {
      return u[(uint64_t)a + (uint64_t)g.nx * ((uint64_t)b + (uint64_t)g.ny * (uint64_t)d)];
    };
    wp[3 * p + c] = tx0 * ty0 * tz0 * at(i0, j0, k0) + tx * ty0 * tz0 * at(i1, j0, k0) +
                    tx * ty * tz0 * at(i1, j1, k0) + tx0 * ty * tz0 * at(i0, j1, k0) +
                    tx0 * ty0 * tz * at(i0, j0, k1) + tx * ty0 * tz * at(i1, j0, k1) +
                    tx * ty * tz * at(i1, j1, k1) + tx0 * ty * tz * at(i0, j1, k1);
  }
}

__global__ void k_maps_dense(Grid g, const int32_t* __restrict__ isn, const uint32_t* __restrict__ mask,
                             int32_t* __restrict__ ib, int32_t* __restrict__ is) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= g.ncell) return;
  const bool solid = mask != nullptr && bit_at(mask, n);
  ib[n] = solid ? 1 : -1;
  is[n] = (isn != nullptr && isn[n] > 0) ? isn[n] : -1;
}

__global__ void k_links_dense(Grid g, int nseg, const int64_t* __restrict__ off, int64_t cap,
                              const uint32_t* __restrict__ lk_node, const uint8_t* __restrict__ lk_dir,
                              const double* __restrict__ lk_alpha, int32_t* __restrict__ ilink,
                              int32_t* __restrict__ mlink, double* __restrict__ alink) {
  const int b = blockIdx.x;
  const int p = b / nseg;
  const int64_t a = off[b], e = min(off[b + 1], cap);
  for (int64_t l = a + threadIdx.x; l < e; l += blockDim.x) {
    if (lk_node[l] >= g.ncell || lk_dir[l] < 1 || lk_dir[l] >= NPOP) continue;
    const uint64_t d = (uint64_t)(lk_dir[l] - 1) + 18ull * lk_node[l];
    ilink[d] = 1;
    mlink[d] = p + 1;
    alink[d] = lk_alpha[l];
  }
}

// ---------------------------------------------------------------------------
static PGeom pgeom(const Ctx& c) {
  PGeom G;
  G.nx = c.g.nx;
  G.ny = c.g.ny;
  G.nz = c.g.nz;
  G.walls_y = c.g.walls_y;
  G.rad = c.p.rad;
  G.radp2 = c.p.rad * 1.0 + 2.0;
  G.r2 = (c.p.rad * 1.0) * (c.p.rad * 1.0);
  G.rlink2 = (c.p.rad + 1.5) * (c.p.rad + 1.5);
  G.rsolid2 = (c.p.rad + 0.5) * (c.p.rad + 0.5);
  G.slab = c.pt.slab;
  return G;
}

// the widest candidate range axis_range() can return, in nodes
static int box_candidates(double rad) {
  const double radp2 = rad + 2.0;
  return (int)ceil(2.0 * radp2 + 1.0) + 4;
}

int segments_per_particle(double rad, int slab) {
  return (box_candidates(rad) + slab - 1) / slab;
}

int segment_capacity(double rad, int slab) {
  const int maxcnt = box_candidates(rad);
  return slab * maxcnt * maxcnt;
}

// Planes per segment.  8 when the spheres are many (the segment grids of the
// link rebuild and the bounce-back still fill the GPU several times over
// with half as many, twice as large workgroups: at 1024^3 with 20,000
// spheres of radius 6 the three kernels take 0.5 ms less per step), else 4
// (300 spheres of radius 8 at 256^3 need the parallelism: 0.06 ms more per
// step with 8).  8 also needs at most PB / 8 candidate rows per plane and
// keeps the scan's shell list within 20 KB of LDS, so only for
// box_candidates <= 25 (rad <= 8).  MFC_SEG_PLANES=4 or 8 forces one (8 only
// where it fits).
constexpr int64_t kSlab8MinSegments = 8192;
int segment_planes(double rad, int npart) {
  const bool fits8 = box_candidates(rad) <= 25;
  const char* e = getenv("MFC_SEG_PLANES");
  if (e && e[0] == '4') return 4;
  if (e && e[0] == '8') return fits8 ? 8 : 4;
  return (fits8 && (int64_t)npart * segments_per_particle(rad, 8) >= kSlab8MinSegments) ? 8 : 4;
}

// Reset map buffer `buf` to empty (no-op if it is empty already).
hipError_t launch_map_clear(Ctx& c, int buf) {
  Particles& P = c.pt;
  if (P.map_np[buf] > 0)
    hipLaunchKernelGGL(k_map_clear, dim3(P.map_np[buf] * P.nseg), dim3(PB), 0, c.stream, P.nseg, P.seg_cap,
                       P.sol_list[buf], P.sol_cnt[buf], P.isn[buf], P.mask[buf]);
  P.map_np[buf] = 0;
  return hipGetLastError();
}

// Arguments of beads_filling for the maps `old` (previous) and `neu` (new);
// false if there is nothing to refill.
static bool make_fill_args(Ctx& c, int old, int neu, FillArgs& fa) {
  Particles& P = c.pt;
  if (P.map_np[old] == 0 || old == neu) return false;
  fa.ypglbp = P.ypglbp; fa.wp = P.wp; fa.wpp = P.wpp; fa.omgp = P.omgp; fa.omgpp = P.omgpp;
  fa.sol_old = P.sol_list[old];
  fa.sol_cnt_old = P.sol_cnt[old];
  fa.seg_cap = P.seg_cap;
  fa.isn_old = P.isn[old];
  fa.mask_new = P.mask[neu];
  fa.mask_old = P.mask[old];
  fa.rad = c.p.rad; fa.rho0 = c.p.rho0; fa.ww0 = c.p.ww0; fa.ww1 = c.p.ww1; fa.ww2 = c.p.ww2;
  fa.list = P.fill_list;
  fa.count = P.fill_count;
  fa.cap = P.fill_cap;
  return true;
}

// The cell grid if it is on (else cnt = nullptr: all pairs).  The lists must
// be current (ensure_cells) before a kernel reads them.
static CellArgs cell_args(const Ctx& c) {
  const Particles& P = c.pt;
  CellArgs C;
  C.cnt = c.ens ? nullptr : P.cell_cnt;   // ensemble members search all pairs
  C.list = P.cell_list;
  C.ncx = P.cell_n[0];
  C.ncy = P.cell_n[1];
  C.ncz = P.cell_n[2];
  return C;
}

// Blocks of the lubrication part of a grid: one per particle (all pairs) or
// one thread per particle (cells).
static int lub_blocks(const Ctx& c) {
  return cell_args(c).cnt ? (c.pt.np + PB - 1) / PB : c.pt.np;
}

hipError_t ensure_cells(Ctx& c) {
  Particles& P = c.pt;
  if (!cell_args(c).cnt || P.cells_ok) return hipSuccess;
  const int ncell = P.cell_n[0] * P.cell_n[1] * P.cell_n[2];
  hipLaunchKernelGGL(k_cell_zero, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, P.cell_cnt, ncell);
  hipLaunchKernelGGL(k_cell_build, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.ypglb, P.np,
                     pgeom(c), cell_args(c), c.d_status);
  P.cells_ok = true;
  return hipGetLastError();
}

static LubArgs lub_args(Ctx& c) {
  LubArgs L;
  L.mingap = c.p.mingap;
  L.fpc = c.p.volp * c.p.rho0 * 9.8 / c.p.gscale;
  L.stf0 = c.p.stf0;
  L.stf1 = c.p.stf1;
  L.mingap_w = c.p.mingap_w;
  L.stf0_w = c.p.stf0_w;
  L.stf1_w = c.p.stf1_w;
  L.flubp = c.pt.flubp;
  L.cells = cell_args(c);
  return L;
}

// ---- argument builders: one box's arguments of each step kernel ------------
// (used by the single-box launchers below and by the ensemble engine)

ScanArgs scan_args(Ctx& c, int target, int ovl, int istep) {
  Particles& P = c.pt;
  ScanArgs a;
  a.ypglb = P.ypglb; a.np = P.np; a.nseg = P.nseg; a.seg_cap = P.seg_cap; a.G = pgeom(c);
  a.isn = P.isn[target]; a.mask = P.mask[target]; a.sol_list = P.sol_list[target];
  a.sol_cnt = P.sol_cnt[target]; a.sh_node = P.sh_node; a.sh_hits = P.sh_hits; a.sh_cnt = P.sh_cnt;
  a.lk_cnt = P.lk_cnt; a.ovl = ovl; a.istep = istep; a.status = c.d_status;
  a.cells = cell_args(c);
  return a;
}

ScanCountArgs count_args(Ctx& c, int target) {
  Particles& P = c.pt;
  ScanCountArgs a;
  a.cnt = P.lk_cnt; a.scnt = P.sol_cnt[target]; a.nsegs = P.np * P.nseg; a.off = P.lk_off;
  a.cap = P.lk_cap; a.status = c.d_status; a.fill_count = P.fill_count; a.ct_count = P.ct_count;
  a.part = P.lk_part;
  return a;
}

// with_find: the refill search (previous map P.map_old vs `target`) rides in
// the same grid; *blocks gets the grid size
ExpandArgs expand_args(Ctx& c, int target, bool with_find, int* blocks) {
  Particles& P = c.pt;
  ExpandArgs a;
  a.ypglb = P.ypglb; a.np = P.np; a.nseg = P.nseg; a.seg_cap = P.seg_cap; a.G = pgeom(c);
  a.sh_node = P.sh_node; a.sh_hits = P.sh_hits; a.sh_cnt = P.sh_cnt; a.off = P.lk_off; a.cap = P.lk_cap;
  a.lk_node = P.lk_node; a.lk_dir = P.lk_dir; a.lk_alpha = P.lk_alpha; a.status = c.d_status;
  a.nsegs = P.np * P.nseg;
  a.mask = P.mask[target]; a.ct_list = P.ct_list; a.ct_count = P.ct_count;
  a.fa = FillArgs{};
  const int nfind = (with_find && make_fill_args(c, P.map_old, target, a.fa)) ? P.map_np[P.map_old] * P.nseg : 0;
  *blocks = a.nsegs + nfind;
  return a;
}

// fused: the lubrication forces ride in the same grid
// fused (mfc_step): lubrication blocks, and the map buffer the coming links
// rebuild (map_cur ^ 1, built two steps ago) cleared in the same grid and
// marked empty
BounceArgs bounce_args(Ctx& c, bool fused, int nct, int* bb_blocks, int* pre_blocks) {
  Particles& P = c.pt;
  BounceArgs a;
  a.A = c.pop[c.cur]; a.g = c.g; a.G = pgeom(c); a.mask = P.mask[P.map_cur]; a.nseg = P.nseg;
  a.off = P.lk_off; a.cap = P.lk_cap; a.lk_node = P.lk_node; a.lk_dir = P.lk_dir;
  a.lk_alpha = P.lk_alpha; a.lk_out = P.lk_out; a.ct_list = P.ct_list; a.ct_count = P.ct_count;
  a.ypglb = P.ypglb; a.wp = P.wp; a.omgp = P.omgp;
  a.ww1 = c.p.ww1; a.ww2 = c.p.ww2; a.seg_ft = P.seg_ft; a.status = c.d_status;
  a.nsegs = P.np * P.nseg; a.np = P.np; a.lub = lub_args(c);
  a.split = P.bb_split;
  a.ref_order = P.bb_order == MFC_BB_REFERENCE;
  a.nct = a.ref_order ? 1 : nct;
  a.ord = P.ord; a.ord_cap = P.ord_cap; a.isn = P.isn[P.map_cur];
  a.nlub = fused ? lub_blocks(c) : 0;
  const int t = P.map_cur ^ 1;
  const int nclr = (fused && P.map_np[t] > 0) ? P.map_np[t] * P.nseg : 0;
  a.seg_cap = P.seg_cap;
  a.clr_list = P.sol_list[t]; a.clr_cnt = P.sol_cnt[t]; a.clr_isn = P.isn[t]; a.clr_mask = P.mask[t];
  if (nclr) P.map_np[t] = 0;
  *bb_blocks = a.nsegs * a.split + nclr;
  *pre_blocks = a.nct + a.nlub;
  return a;
}

MoveKArgs move_args(Ctx& c, int first, int wrap, bool from_seg, int* blocks) {
  Particles& P = c.pt;
  MoveKArgs a;
  MoveArgs& m = a.m;
  m.yp = P.yp; m.ypglb = P.ypglb; m.ypglbp = P.ypglbp;
  m.wp = P.wp; m.wpp = P.wpp; m.omgp = P.omgp; m.omgpp = P.omgpp;
  m.fHIp = P.fHIp; m.flubp = P.flubp; m.torqp = P.torqp; m.forcep = P.forcep;
  m.forcepp = P.forcepp; m.torqpp = P.torqpp; m.dwdt = P.dwdt; m.domgdt = P.domgdt;
  m.thetap = P.thetap;
  m.volp = c.p.volp; m.rhog = c.p.rhog; m.amp = c.p.amp; m.aip = c.p.aip;
  m.bfx = c.p.body_force[0]; m.bfy = c.p.body_force[1]; m.bfz = c.p.body_force[2];
  m.pfx = c.p.particle_force[0]; m.pfy = c.p.particle_force[1]; m.pfz = c.p.particle_force[2];
  a.np = P.np; a.first = first; a.wrap = wrap; a.G = pgeom(c); a.status = c.d_status;
  a.seg_ft = from_seg ? P.seg_ft : nullptr; a.nseg = P.nseg * P.bb_split;
  *blocks = (P.np + 255) / 256;
  return a;
}

// false if there is nothing to refill
bool fill_apply_args(Ctx& c, int nblk, FillApplyArgs& a) {
  Particles& P = c.pt;
  if (!make_fill_args(c, P.map_old, P.map_cur, a.fa)) return false;
  a.A = c.pop[c.cur]; a.g = c.g; a.nblk = nblk;
  return true;
}

// ---- far and near tiles of the overlapped step ----------------------------------
// A tile is kTileX consecutive x nodes of one row (tile t = tx + tpr * row).
// The particle phase of step n (bounce-back, lubrication, move, links,
// refill, with the clearing of the map built two steps back) reads and
// writes no node farther than r + 3 sqrt 2 + 2 from the centres it starts
// from, as long as no sphere moves a node or more in one step (k_move counts
// such steps); the next collide touches only nodes within sqrt 2 of its
// own.  Tiles with no node within reach = r + 3 sqrt 2 + 2 + sqrt 2 of a
// centre (periodic images; none across the y walls) are "far": their collide
// for step n + 1 can run while the particle phase of step n does.

// One workgroup per sphere marks the tiles of every row whose node centres
// come within `reach` of the centre.
__global__ void __launch_bounds__(PB)
    k_tile_mark(const double* __restrict__ ypglb, int np, PGeom G, double reach, int tpr,
                uint8_t* __restrict__ mark) {
  const int p = blockIdx.x;
  if (p >= np) return;
  const double cx = ypglb[3 * p], cy = ypglb[3 * p + 1], cz = ypglb[3 * p + 2];
  const int j0 = (int)floor(cy - reach - 0.5), j1 = (int)ceil(cy + reach - 0.5);
  const int k0 = (int)floor(cz - reach - 0.5), k1 = (int)ceil(cz + reach - 0.5);
  const int nj = j1 - j0 + 1, nk = k1 - k0 + 1;
  const double r2 = reach * reach;
  for (int t = threadIdx.x; t < nj * nk; t += PB) {
    const int jj = j0 + t % nj, kk = k0 + t / nj;
    const double dy = (double)jj + 0.5 - cy, dz = (double)kk + 0.5 - cz;
    const double d2 = dy * dy + dz * dz;
    if (d2 > r2) continue;
    if (G.walls_y && (jj < 0 || jj >= G.ny)) continue;
    const int j = ((jj % G.ny) + G.ny) % G.ny, k = ((kk % G.nz) + G.nz) % G.nz;
    uint8_t* row = mark + (size_t)tpr * ((size_t)j + (size_t)G.ny * k);
    const double e = sqrt(r2 - d2);
    const int i0 = (int)ceil(cx - e - 0.5), i1 = (int)floor(cx + e - 0.5);
    if (i1 - i0 + 1 >= G.nx) {
      for (int b = 0; b < tpr; ++b) row[b] = 1;
      continue;
    }
    for (int i = i0; i <= i1;) {
      const int iw = ((i % G.nx) + G.nx) % G.nx;
      const int b = iw / kTileX;
      row[b] = 1;
      // first node past this tile (the last tile of a row may be short)
      const int end = kTileX * (b + 1) < G.nx ? kTileX * (b + 1) : G.nx;
      i += end - iw;
    }
  }
}

// Near tiles per chunk of kTileChunk tiles.
__global__ void __launch_bounds__(PB)
    k_tile_count(const uint8_t* __restrict__ mark, uint64_t ntiles, int32_t* __restrict__ cnt) {
  __shared__ int red[PB / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTileChunk;
  int c = 0;
  for (int i = threadIdx.x; i < kTileChunk; i += PB) {
    const uint64_t t = base + (uint64_t)i;
    if (t < ntiles && mark[t]) ++c;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < PB / 64; ++w) s += red[w];
    cnt[blockIdx.x] = s;
  }
}

// Exclusive offsets of the near and far tiles of each chunk; totals in
// tot[0] (near) and tot[1] (far).  One workgroup.
__global__ void __launch_bounds__(1024)
    k_tile_offsets(const int32_t* __restrict__ cnt, int nchunk, uint64_t ntiles,
                   uint32_t* __restrict__ off_near, uint32_t* __restrict__ off_far,
                   uint32_t* __restrict__ tot) {
  __shared__ uint32_t part[1024];
  const int per = (nchunk + 1023) / 1024;
  const int c0 = threadIdx.x * per, c1 = min(nchunk, c0 + per);
  uint32_t s = 0;
  for (int c = c0; c < c1; ++c) s += (uint32_t)cnt[c];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t near = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int c = c0; c < c1; ++c) {
    const uint64_t first = (uint64_t)c * kTileChunk;
    off_near[c] = near;
    off_far[c] = (uint32_t)(first - near);
    near += (uint32_t)cnt[c];
  }
  if (threadIdx.x == 1023) {
    tot[0] = part[1023];
    tot[1] = (uint32_t)(ntiles - part[1023]);
  }
}

// Each chunk writes its tiles, in order, into the near and far lists, and
// clears its marks for the next step.
__global__ void __launch_bounds__(PB)
    k_tile_emit(uint8_t* __restrict__ mark, uint64_t ntiles, const uint32_t* __restrict__ off_near,
                const uint32_t* __restrict__ off_far, uint32_t* __restrict__ near_list,
                uint32_t* __restrict__ far_list) {
  __shared__ uint32_t wsum[PB / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTileChunk;
  uint32_t near = off_near[blockIdx.x], far = off_far[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i0 = 0; i0 < kTileChunk; i0 += PB) {
    const uint64_t t = base + (uint64_t)(i0 + threadIdx.x);
    const bool in = t < ntiles;
    const bool nr = in && mark[t] != 0;
    const uint64_t bal = __ballot(nr);
    const uint64_t bin = __ballot(in);
    const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    const uint32_t below_in = (uint32_t)__popcll(bin & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int q = 0; q < PB / 64; ++q) {
      if (q < w) pre += wsum[q];
      tot += wsum[q];
    }
    // tiles of earlier waves in this pass: 64 each (all in range until the end)
    const uint32_t pre_in = (uint32_t)(64 * w);
    if (in) {
      if (nr) {
        near_list[near + pre + below] = (uint32_t)t;
        mark[t] = 0;
      } else {
        far_list[far + (pre_in - pre) + (below_in - below)] = (uint32_t)t;
      }
    }
    near += tot;
    far += (uint32_t)PB - tot;
    __syncthreads();
  }
}

hipError_t launch_tiles(Ctx& c, TileLists& T) {
  Particles& P = c.pt;
  const PGeom G = pgeom(c);
  hipLaunchKernelGGL(k_tile_mark, dim3(P.np), dim3(PB), 0, c.stream, P.ypglb, P.np, G, T.reach, T.tpr,
                     T.mark);
  const int nchunk = (int)((T.ntiles + kTileChunk - 1) / kTileChunk);
  hipLaunchKernelGGL(k_tile_count, dim3(nchunk), dim3(PB), 0, c.stream, T.mark, T.ntiles, T.cnt);
  hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, c.stream, T.cnt, nchunk, T.ntiles, T.off_near,
                     T.off_far, T.tot);
  hipLaunchKernelGGL(k_tile_emit, dim3(nchunk), dim3(PB), 0, c.stream, T.mark, T.ntiles, T.off_near,
                     T.off_far, T.near_list, T.far_list);
  return hipGetLastError();
}

// ---- single-box launchers -------------------------------------------------------
hipError_t launch_links(Ctx& c, int target, int overlap_check, int istep, bool with_find) {
  Particles& P = c.pt;
  const int nsegs = P.np * P.nseg;
  hipError_t e = launch_map_clear(c, target);
  if (e != hipSuccess) return e;
  if (P.np <= 0) return hipGetLastError();
  int nexp = 0;
  if (overlap_check) {
    e = ensure_cells(c);
    if (e != hipSuccess) return e;
  }
  const ExpandArgs ea = expand_args(c, target, with_find, &nexp);
  hipLaunchKernelGGL(k_links_scan, dim3(nsegs), dim3(PB), sizeof(uint32_t) * (size_t)P.seg_cap,
                     c.stream, scan_args(c, target, overlap_check, istep));
  if (nsegs > kScanChunk) {
    const int nb = (nsegs + kScanChunk - 1) / kScanChunk;
    const ScanCountArgs ca = count_args(c, target);
    hipLaunchKernelGGL(k_count_partials, dim3(nb), dim3(PB), 0, c.stream, ca);
    hipLaunchKernelGGL(k_count_offsets, dim3(nb), dim3(PB), 0, c.stream, ca);
  } else {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, c.stream, count_args(c, target));
  }
  hipLaunchKernelGGL(k_links_expand, dim3(nexp), dim3(PB), 0, c.stream, ea);
  P.map_np[target] = P.np;
  return hipGetLastError();
}

constexpr int kContactBlocks = 16;   // grid of the contact pre-pass (contacts are rare)

hipError_t launch_bounce_back(Ctx& c, bool fused) {
  Particles& P = c.pt;
  if (fused) {
    const hipError_t e = ensure_cells(c);
    if (e != hipSuccess) return e;
  }
  int nb = 0, npre = 0;
  const BounceArgs ba = bounce_args(c, fused, kContactBlocks, &nb, &npre);
  if (c.prec == MFC_PREC_FP64) {
    hipLaunchKernelGGL(k_bb_contact<double>, dim3(npre), dim3(PB), 0, c.stream, ba);
    hipLaunchKernelGGL(k_bounce_back<double>, dim3(nb), dim3(PB), 0, c.stream, ba);
  } else {
    hipLaunchKernelGGL(k_bb_contact<float>, dim3(npre), dim3(PB), 0, c.stream, ba);
    hipLaunchKernelGGL(k_bounce_back<float>, dim3(nb), dim3(PB), 0, c.stream, ba);
  }
  if (!fused)
    hipLaunchKernelGGL(k_seg_reduce, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.seg_ft,
                       P.nseg * P.bb_split, P.np, P.fHIp, P.torqp);
  return hipGetLastError();
}

hipError_t launch_lubforce(Ctx& c) {
  Particles& P = c.pt;
  const hipError_t e = ensure_cells(c);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lubforce, dim3(lub_blocks(c)), dim3(PB), 0, c.stream, P.ypglb, P.np, pgeom(c),
                     lub_args(c), c.d_status);
  return hipGetLastError();
}

hipError_t launch_move(Ctx& c, int first, int wrap, bool from_seg) {
  int nb = 0;
  const MoveKArgs a = move_args(c, first, wrap, from_seg, &nb);
  hipLaunchKernelGGL(k_move, dim3(nb), dim3(256), 0, c.stream, a);
  if (wrap) c.pt.cells_ok = false;   // the wrapped positions are the new ypglb
  return hipGetLastError();
}

hipError_t launch_redistribute(Ctx& c, int istep) {
  Particles& P = c.pt;
  hipLaunchKernelGGL(k_wrap, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.yp, P.ypglb,
                     P.ypglbp, P.np, pgeom(c), c.d_status);
  P.cells_ok = false;
  const hipError_t e = ensure_cells(c);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_overlap, dim3(P.np), dim3(PB), 0, c.stream, P.ypglb, P.np, pgeom(c), istep,
                     cell_args(c), c.d_status);
  return hipGetLastError();
}

constexpr int kFillBlocks = 1024;

hipError_t launch_filling(Ctx& c, bool find, const TableJob* job) {
  Particles& P = c.pt;
  FillApplyArgs a;
  if (!fill_apply_args(c, kFillBlocks, a)) return job ? launch_force_table(c, job->slot, job->M) : hipSuccess;
  if (find)
    hipLaunchKernelGGL(k_fill_find, dim3(P.map_np[P.map_old] * P.nseg), dim3(PB), 0, c.stream, a.fa,
                       P.nseg, c.d_status);
  TableJob tj;
  if (job)
    tj = *job;
  else
    tj.nblk = 0;
  const dim3 grid(kFillBlocks + tj.nblk), blk(256);
  if (c.prec == MFC_PREC_FP64) {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_fill_apply<double, true>), grid, blk, 0, c.stream, a, tj);
    else
      hipLaunchKernelGGL((k_fill_apply<double, false>), grid, blk, 0, c.stream, a, tj);
  } else {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_fill_apply<float, true>), grid, blk, 0, c.stream, a, tj);
    else
      hipLaunchKernelGGL((k_fill_apply<float, false>), grid, blk, 0, c.stream, a, tj);
  }
  return hipGetLastError();
}

hipError_t launch_initpartvel(Ctx& c) {
  Particles& P = c.pt;
  hipLaunchKernelGGL(k_initpartvel, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.ypglb, P.np,
                     c.g, c.ux, c.uy, c.uz, P.wp);
  return hipGetLastError();
}

hipError_t launch_maps_to_host_layout(Ctx& c, int which, int32_t* ib, int32_t* is) {
  Particles& P = c.pt;
  const int b = which == 0 ? P.map_cur : P.map_old;
  const bool have = P.isn[b] != nullptr && P.map_np[b] > 0;
  hipLaunchKernelGGL(k_maps_dense, dim3((unsigned)((c.g.ncell + 255) / 256)), dim3(256), 0, c.stream,
                     c.g, have ? P.isn[b] : nullptr, have ? P.mask[b] : nullptr, ib, is);
  return hipGetLastError();
}

hipError_t launch_links_dense(Ctx& c, int32_t* ilink, int32_t* mlink, double* alink) {
  Particles& P = c.pt;
  hipLaunchKernelGGL(k_links_dense, dim3(P.np * P.nseg), dim3(256), 0, c.stream, c.g, P.nseg, P.lk_off,
                     P.lk_cap, P.lk_node, P.lk_dir, P.lk_alpha, ilink, mlink, alink);
  return hipGetLastError();
}

}  // namespace mfc

namespace mfc {

size_t stage_arg_bytes(Stage st) {
  switch (st) {
    case SG_BOUNCE: return sizeof(BounceArgs);
    case SG_CONTACT: return sizeof(BounceArgs);
    case SG_MOVE: return sizeof(MoveKArgs);
    case SG_SCAN: return sizeof(ScanArgs);
    case SG_COUNT: return sizeof(ScanCountArgs);
    case SG_EXPAND: return sizeof(ExpandArgs);
    case SG_FILL: return sizeof(FillApplyArgs);
  }
  return 0;
}

hipError_t launch_batched(Stage st, int prec, int walls, const void* t, const int* off, int nc,
                          int blocks, size_t lds, hipStream_t s) {
  if (nc <= 0 || blocks <= 0) return hipSuccess;
  const bool f64 = prec == MFC_PREC_FP64;
  switch (st) {
    case SG_BOUNCE:
      if (f64)
        hipLaunchKernelGGL(kb_bounce_back<double>, dim3(blocks), dim3(PB), 0, s, (const BounceArgs*)t, off, nc);
      else
        hipLaunchKernelGGL(kb_bounce_back<float>, dim3(blocks), dim3(PB), 0, s, (const BounceArgs*)t, off, nc);
      break;
    case SG_CONTACT:
      if (f64)
        hipLaunchKernelGGL(kb_bb_contact<double>, dim3(blocks), dim3(PB), 0, s, (const BounceArgs*)t, off, nc);
      else
        hipLaunchKernelGGL(kb_bb_contact<float>, dim3(blocks), dim3(PB), 0, s, (const BounceArgs*)t, off, nc);
      break;
    case SG_MOVE:
      hipLaunchKernelGGL(kb_move, dim3(blocks), dim3(256), 0, s, (const MoveKArgs*)t, off, nc);
      break;
    case SG_SCAN:
      hipLaunchKernelGGL(kb_links_scan, dim3(blocks), dim3(PB), lds, s, (const ScanArgs*)t, off, nc);
      break;
    case SG_COUNT:
      hipLaunchKernelGGL(kb_scan_counts, dim3(blocks), dim3(1024), 0, s, (const ScanCountArgs*)t, off, nc);
      break;
    case SG_EXPAND:
      hipLaunchKernelGGL(kb_links_expand, dim3(blocks), dim3(PB), 0, s, (const ExpandArgs*)t, off, nc);
      break;
    case SG_FILL:
      if (f64) {
        if (walls)
          hipLaunchKernelGGL((kb_fill_apply<double, true>), dim3(blocks), dim3(256), 0, s, (const FillApplyArgs*)t, off, nc);
        else
          hipLaunchKernelGGL((kb_fill_apply<double, false>), dim3(blocks), dim3(256), 0, s, (const FillApplyArgs*)t, off, nc);
      } else {
        if (walls)
          hipLaunchKernelGGL((kb_fill_apply<float, true>), dim3(blocks), dim3(256), 0, s, (const FillApplyArgs*)t, off, nc);
        else
          hipLaunchKernelGGL((kb_fill_apply<float, false>), dim3(blocks), dim3(256), 0, s, (const FillApplyArgs*)t, off, nc);
      }
      break;
  }
  return hipGetLastError();
}

}  // namespace mfc
