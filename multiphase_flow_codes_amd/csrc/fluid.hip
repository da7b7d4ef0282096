// fluid.hip -- lattice kernels of libmfc_lbm for gfx950.
//
// Storage: populations are SoA, direction i of node (x, y, z) at
// pop[i * pitch + x + rs * (y + ny*z)]: by default the 19 directions' rows of
// one (y, z) side by side (pitch = nx, rs = 19 nx), else whole planes
// (rs = nx, pitch = the node count plus 512 B); in FP64 or FP32 (deviations
// f_i - w_i*rho0).  The resident buffer holds the
// post-collision (pre-streaming) field; streaming is never a separate pass:
// every consumer pulls f_i(x) = buf[i][x - c_i] (half-way bounce-back at the
// y walls when they are on).  One kernel therefore streams, takes moments,
// evaluates the forcing field and collides, reading 19 and writing 19
// populations per node: the 152 B (FP32) / 304 B (FP64) per node that bound
// the step.  Each lane owns one node; consecutive lanes are consecutive x,
// so every plane access of a wavefront is a contiguous 256 B (FP32) or
// 512 B (FP64) run.
#include "force_table.h"
#include "internal.h"
#include "lbm_access.h"
#include "lbm_collide.h"
#include "step_args.h"

namespace mfc {


// Forcing field at node x of row `row` from a per-row table
//   F_c = Re G_c0 + 2 Re(e1 G_c1) + 2 Re(e2 G_c2),  e_k = exp(i 2 pi k x / nx),
// the direct form of the reference's inverse real FFT of its few modes
// (gaussian.f90:325-369, mpifft.f90:111-158).
__device__ __forceinline__ void force_at(const double* __restrict__ G, const double* __restrict__ phx,
                                         uint32_t row, int x, int nx, double& Fx, double& Fy,
                                         double& Fz) {
  const double* t = G + 18ull * row;
  const int x2 = (2 * x >= nx) ? 2 * x - nx : 2 * x;
  const double c1 = phx[2 * x], s1 = phx[2 * x + 1];
  const double c2 = phx[2 * x2], s2 = phx[2 * x2 + 1];
  double v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* q = t + 6 * c;
    v[c] = q[0] + 2.0 * (q[2] * c1 - q[3] * s1) + 2.0 * (q[4] * c2 - q[5] * s2);
  }
  Fx = v[0];
  Fy = v[1];
  Fz = v[2];
}

// Forcing field at node x from one row of a forcing table, in precision R.
template <typename R>
__device__ __forceinline__ void force_row(const double* __restrict__ t, R c1, R s1, R c2, R s2,
                                          R& Fx, R& Fy, R& Fz) {
  R v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* q = t + 6 * c;
    v[c] = (R)q[0] + (R)2 * ((R)q[2] * c1 - (R)q[3] * s1) + (R)2 * ((R)q[4] * c2 - (R)q[5] * s2);
  }
  Fx = v[0];
  Fy = v[1];
  Fz = v[2];
}


// Fused stream + macro + forcing + MRT collision.
//
// Launch: blockIdx.z = z, blockIdx.y = y, blockIdx.x * blockDim.x + lane = x.
// Everything that depends on (y, z) only -- the nine neighbour-row offsets,
// the wall test and the forcing-table rows -- is wave-uniform and lives in
// scalar registers; per lane only x +- 1 wraps.  Stores are non-temporal:
// the post-collision field is not re-read until the next step.
//
// ARR: conserved moments from the rho/u arrays (the reference's literal data
// flow, used while they are authoritative, e.g. the pre-relaxation);
// otherwise recomputed from the pulled populations plus half the forcing
// captured at the last macrovar, which is exactly what those arrays hold.
//
// LAY: layout of the input field.  LAY_AB reads buffer A and writes buffer B.
// The in-place (AA) layouts use one buffer: every node reads its 19 streamed
// values and writes post-collision value i where it read opp(i), so each
// node touches exactly the locations it read and nodes never race.  From
// LAY_AA_ODD that is all in the node's own slots; from LAY_AA_EVEN it is the
// neighbours' (the result is in LAY_AA_ODD, and the other way round).
// NTL: the two-buffer field is read with non-temporal loads (it is not read
// again this step); with rows split over several x-chunks plain loads let the
// chunks share the cache lines at their edges (1024^3: 1.2 % faster collide;
// whole rows at 256^3: 3-8 % slower with plain loads).
template <typename T, typename R, bool ARR, bool WALLS, int LAY, bool NTL = true>
__device__ __forceinline__ void collide_node(const CollideArgs& a, int x, int y, int z) {
  const T* __restrict__ A = (const T*)a.A;
  T* __restrict__ B = (T*)a.B;
  T* F = (T*)a.B;   // the one buffer of the in-place layouts (A == B)
  const uint32_t* __restrict__ mask = a.mask;
  const Grid& g = a.g;
  const CollideConsts& k = a.k;
  const ForceArgs& fa = a.fa;
  const double* __restrict__ rho_a = a.rho;
  const double* __restrict__ ux_a = a.ux;
  const double* __restrict__ uy_a = a.uy;
  const double* __restrict__ uz_a = a.uz;
  const uint32_t nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
  const uint32_t ym = y == 0 ? ny - 1 : y - 1, yp = y == g.ny - 1 ? 0 : y + 1;
  const uint32_t zm = z == 0 ? g.nz - 1 : z - 1, zp = z == g.nz - 1 ? 0 : z + 1;
  const uint32_t row = (uint32_t)y + ny * (uint32_t)z;
  // neighbour rows, indexed [dy+1][dz+1] for the source row (y - cy, z - cz)
  const uint32_t yy[3] = {yp, (uint32_t)y, ym};   // cy = -1, 0, +1
  const uint32_t zz[3] = {zp, (uint32_t)z, zm};   // cz = -1, 0, +1
  const uint32_t xm = x == 0 ? nx - 1 : x - 1, xpp = x == g.nx - 1 ? 0 : x + 1;
  const uint32_t xs[3] = {xpp, (uint32_t)x, xm};   // cx = -1, 0, +1
  const uint32_t n = (uint32_t)x + nx * row;            // node index (mask, moment arrays)
  const uint64_t own = (uint64_t)x + g.rs * row;        // in-plane offset of the node
  const bool wall_lo = WALLS && y == 0, wall_hi = WALLS && y == g.ny - 1;
  // location of the streamed value of direction i: plane offset (pl * pitch,
  // a select between two compile-time planes on the wave-uniform wall test,
  // so it stays a scalar) and in-plane offset
  auto src_of = [&](int i, uint64_t& pbase, uint64_t& src) {
    if (LAY == LAY_AA_ODD) {
      pbase = (uint64_t)i * g.pitch;
      src = own;
      return;
    }
    const int ccx = cx(i), ccy = cy(i), ccz = cz(i);
    const bool wall = (ccy > 0 && wall_lo) || (ccy < 0 && wall_hi);
    src = wall ? own : xs[ccx + 1] + g.rs * (yy[ccy + 1] + ny * zz[ccz + 1]);
    const int p_in = LAY == LAY_AB ? i : opp(i);       // plane of an interior pull
    const int p_wall = LAY == LAY_AB ? opp(i) : i;     // plane beyond a wall
    const uint64_t b_in = (uint64_t)p_in * g.pitch, b_wall = (uint64_t)p_wall * g.pitch;
    pbase = wall ? b_wall : b_in;
  };
  R f[NPOP];
#pragma unroll
  for (int i = 0; i < NPOP; ++i) {
    uint64_t pb;
    uint64_t src;
    src_of(i, pb, src);
    f[i] = LAY == LAY_AB ? (NTL ? (R)__builtin_nontemporal_load(&A[pb + src]) : (R)A[pb + src])
                         : (R)F[pb + src];
  }
  const bool solid = mask != nullptr && ((mask[n >> 5] >> (n & 31u)) & 1u);
  if (!solid) {
    const int x2 = (2 * x >= g.nx) ? 2 * x - g.nx : 2 * x;
    const R c1 = (R)fa.phx[2 * x], s1 = (R)fa.phx[2 * x + 1];
    const R c2 = (R)fa.phx[2 * x2], s2 = (R)fa.phx[2 * x2 + 1];
    R rho, ux, uy, uz;
    R mx = (R)fa.bfx, my = (R)fa.bfy, mz = (R)fa.bfz;   // force of the macrovar moments
    if (ARR) {
      rho = (R)rho_a[n];
      ux = (R)ux_a[n];
      uy = (R)uy_a[n];
      uz = (R)uz_a[n];
    } else {
      R jx, jy, jz;
      density_momentum<R>(f, rho, jx, jy, jz);
      if (fa.gmac) {
        R gx, gy, gz;
        force_row<R>(fa.gmac + 18ull * row, c1, s1, c2, s2, gx, gy, gz);
        mx += gx;
        my += gy;
        mz += gz;
      }
      ux = jx + (R)0.5 * mx;
      uy = jy + (R)0.5 * my;
      uz = jz + (R)0.5 * mz;
    }
    R Fx = (R)fa.bfx, Fy = (R)fa.bfy, Fz = (R)fa.bfz;
    if (fa.gcur) {
      R gx, gy, gz;
      force_row<R>(fa.gcur + 18ull * row, c1, s1, c2, s2, gx, gy, gz);
      Fx += gx;
      Fy += gy;
      Fz += gz;
    }
    // FP32 arithmetic takes the conserved corrections in closed form; FP64
    // keeps the reference's own arithmetic
    if (!ARR && sizeof(R) == 4)
      mrt_collide<R, true>(f, rho, ux, uy, uz, Fx, Fy, Fz, k, mx, my, mz);
    else
      mrt_collide<R, false>(f, rho, ux, uy, uz, Fx, Fy, Fz, k);
  }
  if (LAY == LAY_AB) {
    // one running pointer over the planes: no per-plane 64-bit base to keep
    T* q = B + own;
#pragma unroll
    for (int i = 0; i < NPOP; ++i) {
      __builtin_nontemporal_store((T)f[i], q);
      q += g.pitch;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NPOP; ++i) {
      uint64_t pb;
      uint64_t dst;
      src_of(opp(i), pb, dst);
      __builtin_nontemporal_store((T)f[i], &F[pb + dst]);
    }
  }
}

// Workgroups are dealt to the XCDs round-robin in launch order, and each XCD
// has its own L2.  The tile (x-chunk bx, row) of launch index L:
//   remap 0: L = bx + nbx * row (x-chunk fastest);
//   remap 1: each XCD takes a contiguous range of tiles (rows and their chunks);
//   remap 2: rows stay dealt round-robin (row r on XCD r % 8, so the XCDs
//            sweep neighbouring addresses together) but all x-chunks of a
//            row run on the row's XCD, whose L2 then holds the cache lines
//            the chunks share at their edges.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ void xcd_tile(uint32_t L, uint32_t total, uint32_t nbx, int remap,
                                         uint32_t& bx, uint32_t& row) {
  if (remap == 1) {
    const uint32_t per = total / kXcds;
    const uint32_t t = L >= per * kXcds ? L : (L % kXcds) * per + L / kXcds;
    bx = t % nbx;
    row = t / nbx;
    return;
  }
  if (remap == 2 && nbx > 1) {
    const uint32_t group = kXcds * nbx;          // 8 rows, all their chunks
    const uint32_t full = (total / group) * group;
    if (L < full) {
      const uint32_t q = L / group, r = L - q * group;
      bx = r / kXcds;
      row = q * kXcds + r % kXcds;
      return;
    }
  }
  bx = L % nbx;
  row = L / nbx;
}

// FP32 storage with FP64 arithmetic, two buffers: 4 waves per SIMD (128
// VGPRs, 20-28 B/lane of spills) instead of 3 (132 VGPRs) measured 6 % faster
// at 256^3; the same bound on the FP64 kernel made it 4 % slower.
template <typename T, typename R, bool ARR, bool WALLS, int LAY, bool NTL>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && sizeof(R) == 8 && LAY == LAY_AB ? 4 : 1)))
    k_collide(CollideArgs a, int nbx, int remap) {
  uint32_t bx, row;
  xcd_tile(blockIdx.x, gridDim.x, (uint32_t)nbx, remap, bx, row);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x >= a.g.nx) return;
  const uint32_t y = row % (uint32_t)a.g.ny, z = row / (uint32_t)a.g.ny;
  collide_node<T, R, ARR, WALLS, LAY, NTL>(a, x, (int)y, (int)z);
}

// The collide of the tiles of a list (overlapped step): one tile (kTileX
// nodes of a row) per wavefront, four per workgroup, the grid sized to the
// list.  Workgroups are dealt to the XCDs round-robin (workgroup b runs on
// XCD b % 8), so the launch index is remapped to give each XCD a contiguous
// range of the list, in order: neighbouring rows then meet in one L2.
// Occupancy as the plain collide reaches it: FP32 storage with FP64
// arithmetic held to 4 waves per SIMD with two buffers; FP64 arithmetic in
// place held to 3 (the in-place layouts with walls otherwise take 178 VGPRs,
// 2 waves).
template <typename T, typename R>
constexpr int tile_waves(int lay) {
  return sizeof(R) == 8 ? (lay == LAY_AB ? (sizeof(T) == 4 ? 4 : 1) : 3) : 1;
}

template <typename T, typename R, bool WALLS, int LAY>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(tile_waves<T, R>(LAY))))
    k_collide_tiles(CollideArgs a, const uint32_t* __restrict__ list, uint32_t n, uint32_t tpr) {
  uint32_t b = blockIdx.x;
  const uint32_t per = gridDim.x / kXcds;
  if (b < per * kXcds) b = (b % kXcds) * per + b / kXcds;
  // the wavefront's index is uniform: keep it (and the tile, row, y, z) scalar
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t L = b * 4u + w;
  if (L >= n) return;
  const uint32_t t = list[L];
  const uint32_t row = t / tpr;
  const int x = (int)((t - row * tpr) * (uint32_t)kTileX + (threadIdx.x & 63u));
  if (x >= a.g.nx) return;
  collide_node<T, R, false, WALLS, LAY>(a, x, (int)(row % (uint32_t)a.g.ny), (int)(row / (uint32_t)a.g.ny));
}

// Several boxes in one grid (ensemble engine): block b belongs to box c with
// off[c] <= b < off[c+1]; the box's nodes are numbered x fastest and taken
// BS per block, so consecutive lanes still read consecutive x (and rows
// continue into the next row without a gap).
__device__ __forceinline__ int case_of_block(const int* __restrict__ off, int nc, int b) {
  int lo = 0, hi = nc;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// ARR_ANY: some member takes its moments from the rho/u arrays (a.arr, its
// first step after joining); without it only the recomputing path is built.
// remap = 1 deals the grid to the 8 XCDs in contiguous ranges (workgroups are
// dispatched to the XCDs round-robin, so launch-order block b runs on XCD
// b % 8): consecutive node ranges, whose neighbour rows overlap, then share
// one L2.  The block-to-nodes map is a bijection either way, so results do
// not depend on it.
template <typename T, typename R, bool WALLS, int BS, bool ARR_ANY>
__global__ void __launch_bounds__(BS)
    kb_collide(const CollideArgs* __restrict__ t, const int* __restrict__ off, int nc, int remap) {
  uint32_t b = blockIdx.x;
  if (remap) {
    const uint32_t per = gridDim.x / kXcds;
    if (b < per * kXcds) b = (b % kXcds) * per + b / kXcds;
  }
  const int c = case_of_block(off, nc, (int)b);
  const CollideArgs a = t[c];
  const uint64_t n = (uint64_t)((int)b - off[c]) * BS + threadIdx.x;
  if (n >= a.g.ncell) return;
  const uint32_t nx = (uint32_t)a.g.nx, ny = (uint32_t)a.g.ny;
  const uint32_t x = (uint32_t)(n % nx);
  const uint32_t r = (uint32_t)(n / nx);
  if (ARR_ANY && a.arr)
    collide_node<T, R, true, WALLS, LAY_AB>(a, (int)x, (int)(r % ny), (int)(r / ny));
  else
    collide_node<T, R, false, WALLS, LAY_AB>(a, (int)x, (int)(r % ny), (int)(r / ny));
}

// Forcing table of one step (force_table.h): one thread per (row, component).
__global__ void __launch_bounds__(256)
    k_force_table(ModeSet M, const double* __restrict__ phy, const double* __restrict__ phz, int ny,
                  int nz, double* __restrict__ G) {
  force_table_block(M, phy, phz, ny, nz, G, blockIdx.x);
}

// macrovar (collision.f90:232-332): rho = sum f, u = sum c f + F/2, then the
// rigid-body velocity and rho = rhopart at solid nodes.
template <typename T, bool PULLED, bool WALLS>
__global__ void __launch_bounds__(256)
    k_macro(const T* __restrict__ A, Grid g, ForceArgs fa, const uint32_t* __restrict__ mask,
            const int32_t* __restrict__ isn, const double* __restrict__ ypglb,
            const double* __restrict__ wp, const double* __restrict__ omgp, double rhopart,
            double* __restrict__ rho, double* __restrict__ ux, double* __restrict__ uy,
            double* __restrict__ uz) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= g.ncell) return;
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  double f[NPOP];
  if (PULLED)
    load_pulled<T, double, WALLS>(A, g, x, y, z, n, f);
  else
    load_direct<T, double, WALLS>(A, g, x, y, z, n, f);
  double r, jx, jy, jz;
  density_momentum<double>(f, r, jx, jy, jz);
  double Fx = fa.bfx, Fy = fa.bfy, Fz = fa.bfz;
  if (fa.gmac) {
    double gx, gy, gz;
    force_at(fa.gmac, fa.phx, row, x, g.nx, gx, gy, gz);
    Fx += gx;
    Fy += gy;
    Fz += gz;
  }
  double u = jx + Fx / 2.0, v = jy + Fy / 2.0, w = jz + Fz / 2.0;
  if (mask != nullptr && ((mask[n >> 5] >> (n & 31u)) & 1u)) {
    const int id = isn[n] - 1;
    const double xp = (double)(x + 1) - 0.5, yp = (double)(y + 1) - 0.5, zp = (double)(z + 1) - 0.5;
    const double xc = near_image(ypglb[3 * id], xp, g.nx);
    const double yc = WALLS ? ypglb[3 * id + 1] : near_image(ypglb[3 * id + 1], yp, g.ny);
    const double zc = near_image(ypglb[3 * id + 2], zp, g.nz);
    const double xx = xp - xc, yy = yp - yc, zz = zp - zc;
    const double* W = wp + 3 * id;
    const double* O = omgp + 3 * id;
    u = W[0] + (O[1] * zz - O[2] * yy);
    v = W[1] + (O[2] * xx - O[0] * zz);
    w = W[2] + (O[0] * yy - O[1] * xx);
    r = rhopart;
  }
  rho[n] = r;
  ux[n] = u;
  uy[n] = v;
  uz[n] = w;
}

// rhoupdat (collision.f90:213-229) plus the pre-relaxation residual
// max|rho_new - rho_old| (main.f90:91-96).
template <typename T, bool PULLED, bool WALLS>
__global__ void __launch_bounds__(256)
    k_rhoupdat(const T* __restrict__ A, Grid g, double* __restrict__ rho,
               unsigned long long* __restrict__ maxbits) {
  __shared__ double red[256];
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  double d = 0.0;
  if (n < g.ncell) {
    int x, y, z;
    uint32_t row;
    node_coords(n, g, x, y, z, row);
    double f[NPOP];
    if (PULLED)
      load_pulled<T, double, WALLS>(A, g, x, y, z, n, f);
    else
      load_direct<T, double, WALLS>(A, g, x, y, z, n, f);
    double r = f[0];
#pragma unroll
    for (int i = 1; i < NPOP; ++i) r = r + f[i];
    d = fabs(r - rho[n]);
    rho[n] = r;
  }
  red[threadIdx.x] = d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && maxbits) atomicMax(maxbits, (unsigned long long)__double_as_longlong(red[0]));
}

// initpop (initial.f90:3-28): deviation equilibrium with rho = 0.
template <typename T, bool WALLS>
__global__ void __launch_bounds__(256)
    k_initpop(T* __restrict__ A, Grid g, CollideConsts k, bool pulled, double* __restrict__ rho,
              const double* __restrict__ ux, const double* __restrict__ uy,
              const double* __restrict__ uz) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= g.ncell) return;
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  rho[n] = 0.0;
  const double u = ux[n], v = uy[n], w = uz[n];
  const double usqr = 1.5 * (u * u + v * v + w * w);
#pragma unroll
  for (int i = 0; i < NPOP; ++i) {
    const double G = (double)cx(i) * u + (double)cy(i) * v + (double)cz(i) * w;
    const double wt = i == 0 ? k.ww0 : (i <= 6 ? k.ww1 : k.ww2);
    const double fe = i == 0 ? wt * (0.0 - usqr) : wt * (0.0 + 3.0 * G + 4.5 * G * G - usqr);
    store_phase<T, WALLS>(A, g, i, x, y, z, n, pulled, fe);
  }
}

// SoA (resident, current phase) -> AoS double f(0:18, ...) for z in [z0, z0+nzs)
template <typename T, bool WALLS>
__global__ void __launch_bounds__(256)
    k_pop_to_aos(const T* __restrict__ A, Grid g, double* __restrict__ dst, int z0, uint32_t count,
                 bool pulled) {
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q >= count) return;
  const uint32_t n = q + (uint32_t)z0 * (uint32_t)g.nx * (uint32_t)g.ny;
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  double f[NPOP];
  if (pulled)
    load_pulled<T, double, WALLS>(A, g, x, y, z, n, f);
  else
    load_direct<T, double, WALLS>(A, g, x, y, z, n, f);
#pragma unroll
  for (int i = 0; i < NPOP; ++i) dst[19ull * q + i] = f[i];
}

template <typename T, bool WALLS>
__global__ void __launch_bounds__(256)
    k_aos_to_pop(T* __restrict__ A, Grid g, const double* __restrict__ src, int z0, uint32_t count,
                 bool pulled) {
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q >= count) return;
  const uint32_t n = q + (uint32_t)z0 * (uint32_t)g.nx * (uint32_t)g.ny;
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
#pragma unroll
  for (int i = 0; i < NPOP; ++i) store_phase<T, WALLS>(A, g, i, x, y, z, n, pulled, src[19ull * q + i]);
}

__global__ void __launch_bounds__(256)
    k_force_field(Grid g, ForceArgs fa, double* __restrict__ fx, double* __restrict__ fy,
                  double* __restrict__ fz) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= g.ncell) return;
  int x, y, z;
  uint32_t row;
  node_coords(n, g, x, y, z, row);
  double Fx = fa.bfx, Fy = fa.bfy, Fz = fa.bfz;
  if (fa.gcur) {
    double gx, gy, gz;
    force_at(fa.gcur, fa.phx, row, x, g.nx, gx, gy, gz);
    Fx += gx;
    Fy += gy;
    Fz += gz;
  }
  fx[n] = Fx;
  fy[n] = Fy;
  fz[n] = Fz;
}

// ---------------------------------------------------------------------------
static inline unsigned nblocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

static ForceArgs force_args(const Ctx& c) {
  ForceArgs fa;
  fa.gcur = c.ft.live[c.ft.cur] ? c.ft.g[c.ft.cur] : nullptr;
  fa.gmac = c.ft.live[c.ft.mac] ? c.ft.g[c.ft.mac] : nullptr;
  fa.phx = c.ft.phx;
  fa.bfx = c.p.body_force[0];
  fa.bfy = c.p.body_force[1];
  fa.bfz = c.p.body_force[2];
  return fa;
}

hipError_t launch_force_table(Ctx& c, int slot, const ModeSet& m) {
  const uint32_t rows = (uint32_t)c.g.ny * (uint32_t)c.g.nz;
  hipLaunchKernelGGL(k_force_table, dim3((3 * rows + 255) / 256), dim3(256), 0, c.stream, m,
                     c.ft.phy, c.ft.phz, c.g.ny, c.g.nz, c.ft.g[slot]);
  return hipGetLastError();
}

// x-row blocks: 64..256 lanes covering a row, one block per (x-chunk, y, z),
// a one-dimensional grid of nbx * ny * nz tiles
static void row_launch(const Grid& g, dim3& grid, dim3& blk, int& nbx) {
  int bx = ((g.nx + 63) / 64) * 64;
  if (bx > 256) bx = 256;
  blk = dim3(bx);
  nbx = (g.nx + bx - 1) / bx;
  grid = dim3((unsigned)((uint64_t)nbx * g.ny * g.nz));
}

// Arguments of one box's collide launch (also used by the ensemble engine).
CollideArgs collide_args(Ctx& c) {
  CollideArgs a;
  a.A = c.pop[c.cur];
  a.B = c.g.lay == LAY_AB ? c.pop[c.cur ^ 1] : c.pop[c.cur];
  a.mask = (c.pt.np > 0 && c.pt.map_np[c.pt.map_cur] > 0) ? c.pt.mask[c.pt.map_cur] : nullptr;
  a.g = c.g;
  a.k = c.cc;
  a.fa = force_args(c);
  a.rho = c.rho; a.ux = c.ux; a.uy = c.uy; a.uz = c.uz;
  a.arr = 0;
  return a;
}

template <typename T, typename R, int LAY, bool NTL>
static void collide_launch(Ctx& c, bool arrays, const CollideArgs& a, dim3 grid, dim3 blk, int nbx,
                           int rm) {
  if (arrays) {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_collide<T, R, true, true, LAY, NTL>), grid, blk, 0, c.stream, a, nbx, rm);
    else
      hipLaunchKernelGGL((k_collide<T, R, true, false, LAY, NTL>), grid, blk, 0, c.stream, a, nbx, rm);
  } else {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_collide<T, R, false, true, LAY, NTL>), grid, blk, 0, c.stream, a, nbx, rm);
    else
      hipLaunchKernelGGL((k_collide<T, R, false, false, LAY, NTL>), grid, blk, 0, c.stream, a, nbx, rm);
  }
}

template <typename T, typename R, int LAY>
static void collide_dispatch_lay(Ctx& c, bool arrays, const CollideArgs& a) {
  dim3 grid, blk;
  int nbx;
  row_launch(c.g, grid, blk, nbx);
  // rows split over several chunks (nx > 256), or FP64 arithmetic on FP64
  // populations: contiguous tile ranges per XCD; otherwise launch order
  // (measured best for each: at 256^3 FP64 2.3 % faster step with ranges,
  // FP32 2.7 % slower)
  const int rm = c.xcd_remap >= 0 ? c.xcd_remap : ((nbx > 1 || sizeof(T) == 8) ? 1 : 0);
  if constexpr (LAY == LAY_AB) {
    if (nbx > 1) {   // rows in several x-chunks: plain loads
      collide_launch<T, R, LAY, false>(c, arrays, a, grid, blk, nbx, rm);
      return;
    }
  }
  collide_launch<T, R, LAY, true>(c, arrays, a, grid, blk, nbx, rm);
}

template <typename T, typename R>
static void collide_dispatch(Ctx& c, bool arrays, const CollideArgs& a) {
  if (c.g.lay == LAY_AA_ODD)
    collide_dispatch_lay<T, R, LAY_AA_ODD>(c, arrays, a);
  else if (c.g.lay == LAY_AA_EVEN)
    collide_dispatch_lay<T, R, LAY_AA_EVEN>(c, arrays, a);
  else
    collide_dispatch_lay<T, R, LAY_AB>(c, arrays, a);
}

template <typename T, typename R, int LAY>
static void collide_tiles_lay(Ctx& c, const CollideArgs& a, const uint32_t* list, uint32_t n) {
  const uint32_t tpr = (uint32_t)((c.g.nx + kTileX - 1) / kTileX);
  const dim3 grid((n + 3) / 4);
  if (c.g.walls_y)
    hipLaunchKernelGGL((k_collide_tiles<T, R, true, LAY>), grid, dim3(256), 0, c.stream, a, list, n, tpr);
  else
    hipLaunchKernelGGL((k_collide_tiles<T, R, false, LAY>), grid, dim3(256), 0, c.stream, a, list, n, tpr);
}

template <typename T, typename R>
static void collide_tiles_dispatch(Ctx& c, const CollideArgs& a, const uint32_t* list, uint32_t n) {
  if (c.g.lay == LAY_AA_ODD)
    collide_tiles_lay<T, R, LAY_AA_ODD>(c, a, list, n);
  else if (c.g.lay == LAY_AA_EVEN)
    collide_tiles_lay<T, R, LAY_AA_EVEN>(c, a, list, n);
  else
    collide_tiles_lay<T, R, LAY_AB>(c, a, list, n);
}

hipError_t launch_collide_tiles(Ctx& c, const uint32_t* list, uint32_t n, bool far) {
  if (n == 0) return hipSuccess;
  CollideArgs a = collide_args(c);
  if (far) a.mask = nullptr;   // no node of a far tile is solid in any map of the step
  if (c.prec == MFC_PREC_FP64)
    collide_tiles_dispatch<double, double>(c, a, list, n);
  else if (c.prec == MFC_PREC_FP32_F64MATH)
    collide_tiles_dispatch<float, double>(c, a, list, n);
  else
    collide_tiles_dispatch<float, float>(c, a, list, n);
  return hipGetLastError();
}

hipError_t launch_collide(Ctx& c, bool arrays) {
  const CollideArgs a = collide_args(c);
  if (c.prec == MFC_PREC_FP64)
    collide_dispatch<double, double>(c, arrays, a);
  else if (c.prec == MFC_PREC_FP32_F64MATH)
    collide_dispatch<float, double>(c, arrays, a);
  else
    collide_dispatch<float, float>(c, arrays, a);
  return hipGetLastError();
}

template <typename T>
static void macro_dispatch(Ctx& c, bool pulled, ForceArgs fa) {
  const T* A = (const T*)c.pop[c.cur];
  const bool parts = c.pt.np > 0 && c.pt.map_np[c.pt.map_cur] > 0;
  const uint32_t* mask = parts ? c.pt.mask[c.pt.map_cur] : nullptr;
  const int32_t* isn = parts ? c.pt.isn[c.pt.map_cur] : nullptr;
  const dim3 grid(nblocks(c.g.ncell)), blk(256);
#define MFC_MACRO(P, W)                                                                         \
  hipLaunchKernelGGL((k_macro<T, P, W>), grid, blk, 0, c.stream, A, c.g, fa, mask, isn,         \
                     c.pt.ypglb, c.pt.wp, c.pt.omgp, c.p.rhopart, c.rho, c.ux, c.uy, c.uz)
  if (pulled) {
    if (c.g.walls_y) MFC_MACRO(true, true); else MFC_MACRO(true, false);
  } else {
    if (c.g.walls_y) MFC_MACRO(false, true); else MFC_MACRO(false, false);
  }
#undef MFC_MACRO
}

hipError_t launch_macro(Ctx& c, bool pulled) {
  ForceArgs fa = force_args(c);
  if (c.prec == MFC_PREC_FP64)
    macro_dispatch<double>(c, pulled, fa);
  else
    macro_dispatch<float>(c, pulled, fa);
  return hipGetLastError();
}

template <typename T>
static void rhoupdat_dispatch(Ctx& c, bool pulled, unsigned long long* mb) {
  const T* A = (const T*)c.pop[c.cur];
  const dim3 grid(nblocks(c.g.ncell)), blk(256);
  if (pulled) {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_rhoupdat<T, true, true>), grid, blk, 0, c.stream, A, c.g, c.rho, mb);
    else
      hipLaunchKernelGGL((k_rhoupdat<T, true, false>), grid, blk, 0, c.stream, A, c.g, c.rho, mb);
  } else {
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_rhoupdat<T, false, true>), grid, blk, 0, c.stream, A, c.g, c.rho, mb);
    else
      hipLaunchKernelGGL((k_rhoupdat<T, false, false>), grid, blk, 0, c.stream, A, c.g, c.rho, mb);
  }
}

hipError_t launch_rhoupdat(Ctx& c, bool pulled, double* d_maxdiff) {
  unsigned long long* mb = (unsigned long long*)d_maxdiff;
  if (c.prec == MFC_PREC_FP64)
    rhoupdat_dispatch<double>(c, pulled, mb);
  else
    rhoupdat_dispatch<float>(c, pulled, mb);
  return hipGetLastError();
}

hipError_t launch_initpop(Ctx& c) {
  const dim3 grid(nblocks(c.g.ncell)), blk(256);
  const bool pulled = c.phase == PHASE_POST_STREAM;
  if (c.prec == MFC_PREC_FP64) {
    double* A = (double*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_initpop<double, true>), grid, blk, 0, c.stream, A, c.g, c.cc, pulled, c.rho, c.ux, c.uy, c.uz);
    else
      hipLaunchKernelGGL((k_initpop<double, false>), grid, blk, 0, c.stream, A, c.g, c.cc, pulled, c.rho, c.ux, c.uy, c.uz);
  } else {
    float* A = (float*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_initpop<float, true>), grid, blk, 0, c.stream, A, c.g, c.cc, pulled, c.rho, c.ux, c.uy, c.uz);
    else
      hipLaunchKernelGGL((k_initpop<float, false>), grid, blk, 0, c.stream, A, c.g, c.cc, pulled, c.rho, c.ux, c.uy, c.uz);
  }
  return hipGetLastError();
}

hipError_t launch_pop_to_aos(Ctx& c, double* dst, int z0, int nzs, bool pulled) {
  const uint32_t count = (uint32_t)c.g.nx * (uint32_t)c.g.ny * (uint32_t)nzs;
  const dim3 grid(nblocks(count)), blk(256);
  if (c.prec == MFC_PREC_FP64) {
    const double* A = (const double*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_pop_to_aos<double, true>), grid, blk, 0, c.stream, A, c.g, dst, z0, count, pulled);
    else
      hipLaunchKernelGGL((k_pop_to_aos<double, false>), grid, blk, 0, c.stream, A, c.g, dst, z0, count, pulled);
  } else {
    const float* A = (const float*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_pop_to_aos<float, true>), grid, blk, 0, c.stream, A, c.g, dst, z0, count, pulled);
    else
      hipLaunchKernelGGL((k_pop_to_aos<float, false>), grid, blk, 0, c.stream, A, c.g, dst, z0, count, pulled);
  }
  return hipGetLastError();
}

hipError_t launch_aos_to_pop(Ctx& c, const double* src, int z0, int nzs, bool pulled) {
  const uint32_t count = (uint32_t)c.g.nx * (uint32_t)c.g.ny * (uint32_t)nzs;
  const dim3 grid(nblocks(count)), blk(256);
  if (c.prec == MFC_PREC_FP64) {
    double* A = (double*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_aos_to_pop<double, true>), grid, blk, 0, c.stream, A, c.g, src, z0, count, pulled);
    else
      hipLaunchKernelGGL((k_aos_to_pop<double, false>), grid, blk, 0, c.stream, A, c.g, src, z0, count, pulled);
  } else {
    float* A = (float*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_aos_to_pop<float, true>), grid, blk, 0, c.stream, A, c.g, src, z0, count, pulled);
    else
      hipLaunchKernelGGL((k_aos_to_pop<float, false>), grid, blk, 0, c.stream, A, c.g, src, z0, count, pulled);
  }
  return hipGetLastError();
}

hipError_t launch_force_field(Ctx& c, int slot, double* fx, double* fy, double* fz) {
  ForceArgs fa = force_args(c);
  fa.gcur = c.ft.live[slot] ? c.ft.g[slot] : nullptr;
  hipLaunchKernelGGL(k_force_field, dim3(nblocks(c.g.ncell)), dim3(256), 0, c.stream, c.g, fa, fx, fy, fz);
  return hipGetLastError();
}

}  // namespace mfc

namespace mfc {

template <typename T, typename R, int BS, bool ARR_ANY>
static void batched_collide_k(int walls, const CollideArgs* tab, const int* off, int nc, int blocks,
                              int remap, hipStream_t s) {
  if (walls)
    hipLaunchKernelGGL((kb_collide<T, R, true, BS, ARR_ANY>), dim3(blocks), dim3(BS), 0, s, tab, off, nc,
                       remap);
  else
    hipLaunchKernelGGL((kb_collide<T, R, false, BS, ARR_ANY>), dim3(blocks), dim3(BS), 0, s, tab, off, nc,
                       remap);
}

template <int BS, bool ARR_ANY>
static void batched_collide_bs(int prec, int walls, const void* t, const int* off, int nc, int blocks,
                               int remap, hipStream_t s) {
  const CollideArgs* tab = (const CollideArgs*)t;
  if (prec == MFC_PREC_FP64)
    batched_collide_k<double, double, BS, ARR_ANY>(walls, tab, off, nc, blocks, remap, s);
  else if (prec == MFC_PREC_FP32_F64MATH)
    batched_collide_k<float, double, BS, ARR_ANY>(walls, tab, off, nc, blocks, remap, s);
  else
    batched_collide_k<float, float, BS, ARR_ANY>(walls, tab, off, nc, blocks, remap, s);
}

hipError_t launch_batched_collide(int prec, int walls, int bs, int remap, const void* t, const int* off,
                                  int nc, int blocks, bool arr_any, hipStream_t s) {
  if (nc <= 0 || blocks <= 0) return hipSuccess;
  switch (bs) {
    case 64:
      arr_any ? batched_collide_bs<64, true>(prec, walls, t, off, nc, blocks, remap, s)
              : batched_collide_bs<64, false>(prec, walls, t, off, nc, blocks, remap, s);
      break;
    case 128:
      arr_any ? batched_collide_bs<128, true>(prec, walls, t, off, nc, blocks, remap, s)
              : batched_collide_bs<128, false>(prec, walls, t, off, nc, blocks, remap, s);
      break;
    case 256:
      arr_any ? batched_collide_bs<256, true>(prec, walls, t, off, nc, blocks, remap, s)
              : batched_collide_bs<256, false>(prec, walls, t, off, nc, blocks, remap, s);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mfc
