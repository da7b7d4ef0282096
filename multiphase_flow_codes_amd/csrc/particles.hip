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
#include "internal.h"
#include "lbm_access.h"
#include "lbm_particle_math.h"

namespace mfc {

constexpr int PB = 256;  // threads per particle workgroup

struct PGeom {
  int nx, ny, nz, walls_y;
  double rad, radp2, r2;
};

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

__device__ void box_info(const double* __restrict__ pos, int p, const PGeom& G, BoxInfo& B) {
  B.x0 = pos[3 * p];
  B.y0 = pos[3 * p + 1];
  B.z0 = pos[3 * p + 2];
  B.ok = sane_coord(B.x0, G.nx) && sane_coord(B.y0, G.ny) && sane_coord(B.z0, G.nz);
  if (!B.ok) return;
  B.xr = sweep_rest_base(B.x0, G.nx);
  B.yr = G.walls_y ? B.y0 : sweep_rest_base(B.y0, G.ny);
  B.ax = axis_range(B.x0, G.radp2);
  B.ay = axis_range(B.y0, G.radp2);
  B.az = axis_range(B.z0, G.radp2);
  if (B.ax.cnt > G.nx || B.ay.cnt > G.ny || B.az.cnt > G.nz) {
    B.ok = 0;
    return;
  }
  B.kfirst = 1 << 30;
  for (int t = 0; t < B.az.cnt; ++t) {
    const int kw = wrap_any(B.az.lo + t, G.nz);
    const double zp = (double)(kw + 1) - 0.5;
    const double zz0 = zp - near_image(B.z0, zp, G.nz);
    if (fabs(zz0) <= G.radp2 && kw < B.kfirst) B.kfirst = kw;
  }
  B.jfirst = 1 << 30;
  for (int t = 0; t < B.ay.cnt; ++t) {
    const int jw = G.walls_y ? B.ay.lo + t : wrap_any(B.ay.lo + t, G.ny);
    if (jw < 0 || jw >= G.ny) continue;
    const double yp = (double)(jw + 1) - 0.5;
    const double yy0 = yp - y_image(B.y0, yp, G);
    if (fabs(yy0) <= G.radp2 && jw < B.jfirst) B.jfirst = jw;
  }
}

// Geometry of box node t.  Returns false if the node is outside the
// reference's bounding box (or outside the walls).
__device__ bool box_node(const BoxInfo& B, const PGeom& G, int t, int& iw, int& jw, int& kw,
                         LinkGeom& L) {
  const int tx = t % B.ax.cnt;
  const int rest = t / B.ax.cnt;
  const int ty = rest % B.ay.cnt;
  const int tz = rest / B.ay.cnt;
  kw = wrap_any(B.az.lo + tz, G.nz);
  jw = G.walls_y ? B.ay.lo + ty : wrap_any(B.ay.lo + ty, G.ny);
  iw = wrap_any(B.ax.lo + tx, G.nx);
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

// Neighbour test of direction q from node (iw,jw,kw); fills rr1 and offsets.
__device__ __forceinline__ bool link_hit(const LinkGeom& L, int iw, int jw, int kw, int q, double r2,
                                         double& xx1, double& yy1, double& zz1, double& rr1) {
  xx1 = (double)(iw + 1 + cx(q)) - 0.5 - L.ximg;
  yy1 = (double)(jw + 1 + cy(q)) - 0.5 - L.yimg + 0.0;
  zz1 = (double)(kw + 1 + cz(q)) - 0.5 - L.zimg + 0.0;
  rr1 = xx1 * xx1 + yy1 * yy1 + zz1 * zz1;
  return (rr1 - r2) <= 0.0;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total) {
  // Hillis-Steele inclusive scan over PB threads in shared memory
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < PB; off <<= 1) {
    const int a = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[PB - 1];
  const int excl = sh[t] - v;
  __syncthreads();
  return excl;
}

// Pass 1 of beads_links: count solid nodes and boundary links per particle.
__global__ void __launch_bounds__(PB)
    k_links_count(const double* __restrict__ ypglb, int np, PGeom G, int32_t* __restrict__ lk_cnt,
                  int32_t* __restrict__ sol_cnt, int32_t* __restrict__ status) {
  __shared__ BoxInfo B;
  __shared__ int sh[PB];
  const int p = blockIdx.x;
  if (p >= np) return;
  if (threadIdx.x == 0) box_info(ypglb, p, G, B);
  __syncthreads();
  if (!B.ok) {
    if (threadIdx.x == 0) {
      lk_cnt[p] = 0;
      sol_cnt[p] = 0;
      atomicAdd(&status[ST_BOX_FAULT], 1);
    }
    return;
  }
  const int nbox = B.ax.cnt * B.ay.cnt * B.az.cnt;
  int nl = 0, ns = 0;
  for (int t = threadIdx.x; t < nbox; t += PB) {
    int iw, jw, kw;
    LinkGeom L;
    if (!box_node(B, G, t, iw, jw, kw, L)) continue;
    if (L.rr0 - G.r2 <= 0.0) {
      ++ns;
    } else {
      for (int q = 1; q < NPOP; ++q) {
        double a, b, c, d;
        if (link_hit(L, iw, jw, kw, q, G.r2, a, b, c, d)) ++nl;
      }
    }
  }
  int tot;
  block_exclusive_scan(nl, sh, tot);
  int tots;
  block_exclusive_scan(ns, sh, tots);
  if (threadIdx.x == 0) {
    lk_cnt[p] = tot;
    sol_cnt[p] = tots;
  }
}

// Exclusive scan of the per-particle link counts (one workgroup).
__global__ void __launch_bounds__(1024)
    k_scan_counts(const int32_t* __restrict__ cnt, const int32_t* __restrict__ scnt, int np,
                  int64_t* __restrict__ off, int64_t cap, int32_t* __restrict__ status) {
  __shared__ int64_t part[1024];
  __shared__ int64_t spart[1024];
  const int t = threadIdx.x;
  const int per = (np + 1023) / 1024;
  const int a = t * per, b = min(np, a + per);
  int64_t s = 0, ss = 0;
  for (int i = a; i < b; ++i) {
    s += cnt[i];
    ss += scnt[i];
  }
  part[t] = s;
  spart[t] = ss;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0, acs = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t v = part[i];
      part[i] = acc;
      acc += v;
      acs += spart[i];
    }
    off[np] = acc;
    status[ST_NLINKS] = (int32_t)min(acc, (int64_t)0x7fffffff);
    status[ST_NSOLID] = (int32_t)min(acs, (int64_t)0x7fffffff);
    if (acc > cap) atomicAdd(&status[ST_LINK_OVERFLOW], 1);
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int i = a; i < b; ++i) {
    off[i] = acc;
    acc += cnt[i];
  }
}

// Pass 2 of beads_links: mark solid nodes (id map + mask) and write the links
// of particle p, in box-scan order, at its offset.
__global__ void __launch_bounds__(PB)
    k_links_write(const double* __restrict__ ypglb, int np, PGeom G, const int64_t* __restrict__ off,
                  int64_t cap, int32_t* __restrict__ isn, uint32_t* __restrict__ mask,
                  uint32_t* __restrict__ lk_node, uint8_t* __restrict__ lk_dir,
                  double* __restrict__ lk_alpha, int32_t* __restrict__ status) {
  __shared__ BoxInfo B;
  __shared__ int sh[PB];
  const int p = blockIdx.x;
  if (p >= np) return;
  if (threadIdx.x == 0) box_info(ypglb, p, G, B);
  __syncthreads();
  if (!B.ok) return;
  const int nbox = B.ax.cnt * B.ay.cnt * B.az.cnt;
  int64_t base = off[p];
  int alpha_bad = 0;
  for (int t0 = 0; t0 < nbox; t0 += PB) {
    const int t = t0 + threadIdx.x;
    int iw = 0, jw = 0, kw = 0;
    LinkGeom L;
    bool in = (t < nbox) && box_node(B, G, t, iw, jw, kw, L);
    bool solid = in && (L.rr0 - G.r2 <= 0.0);
    uint32_t hits = 0;
    if (in && !solid) {
      for (int q = 1; q < NPOP; ++q) {
        double a, b, c, d;
        if (link_hit(L, iw, jw, kw, q, G.r2, a, b, c, d)) hits |= 1u << q;
      }
    }
    const uint32_t n = (uint32_t)iw + (uint32_t)G.nx * ((uint32_t)jw + (uint32_t)G.ny * (uint32_t)kw);
    if (solid) {
      isn[n] = p + 1;
      atomicOr(&mask[n >> 5], 1u << (n & 31u));
    }
    int tot;
    const int excl = block_exclusive_scan(__popc(hits), sh, tot);
    int64_t w = base + excl;
    for (int q = 1; q < NPOP; ++q) {
      if (!(hits & (1u << q))) continue;
      double xx1, yy1, zz1, rr1;
      link_hit(L, iw, jw, kw, q, G.r2, xx1, yy1, zz1, rr1);
      const double alpha = link_alpha(L, xx1, yy1, zz1, rr1, G.r2);
      if (alpha < 0.0 || alpha > 1.0) ++alpha_bad;
      if (w < cap) {
        lk_node[w] = n;
        lk_dir[w] = (uint8_t)q;
        lk_alpha[w] = alpha;
      }
      ++w;
    }
    base += tot;
  }
  if (alpha_bad) atomicAdd(&status[ST_ALPHA_FAULT], alpha_bad);
}

// Clear the nodes a map buffer marked for particle p (by id, so exactly the
// nodes that were set are reset).
__global__ void __launch_bounds__(PB)
    k_clear_map(const double* __restrict__ pos, int np, PGeom G, int32_t* __restrict__ isn,
                uint32_t* __restrict__ mask) {
  __shared__ BoxInfo B;
  const int p = blockIdx.x;
  if (p >= np) return;
  if (threadIdx.x == 0) box_info(pos, p, G, B);
  __syncthreads();
  if (!B.ok) return;
  const int nbox = B.ax.cnt * B.ay.cnt * B.az.cnt;
  for (int t = threadIdx.x; t < nbox; t += PB) {
    const int tx = t % B.ax.cnt, rest = t / B.ax.cnt, ty = rest % B.ay.cnt, tz = rest / B.ay.cnt;
    const int kw = wrap_any(B.az.lo + tz, G.nz);
    const int jw = G.walls_y ? B.ay.lo + ty : wrap_any(B.ay.lo + ty, G.ny);
    const int iw = wrap_any(B.ax.lo + tx, G.nx);
    if (jw < 0 || jw >= G.ny) continue;
    const uint32_t n = (uint32_t)iw + (uint32_t)G.nx * ((uint32_t)jw + (uint32_t)G.ny * (uint32_t)kw);
    if (isn[n] == p + 1) {
      isn[n] = 0;
      atomicAnd(&mask[n >> 5], ~(1u << (n & 31u)));
    }
  }
}

__device__ __forceinline__ bool bit_at(const uint32_t* __restrict__ mask, uint32_t n) {
  return (mask[n >> 5] >> (n & 31u)) & 1u;
}

// beads_collision (partlib.f90:1363-1703) for the links of particle p:
// interpolated bounce-back value of every link into lk_out, and the
// momentum-exchange force and torque on the particle reduced in FP64 in a
// fixed order.  All links read one snapshot of the post-collision field; the
// values are scattered into the solid nodes by k_bb_scatter afterwards.
template <typename T>
__global__ void __launch_bounds__(PB)
    k_bounce_back(const T* __restrict__ A, Grid g, PGeom G, const uint32_t* __restrict__ mask,
                  const int64_t* __restrict__ off, int64_t cap, const uint32_t* __restrict__ lk_node,
                  const uint8_t* __restrict__ lk_dir, const double* __restrict__ lk_alpha,
                  double* __restrict__ lk_out, const double* __restrict__ ypglb,
                  const double* __restrict__ wp, const double* __restrict__ omgp, double ww1,
                  double ww2, double* __restrict__ fHIp, double* __restrict__ torqp,
                  int32_t* __restrict__ status) {
  __shared__ double red[6][PB];
  const int p = blockIdx.x;
  const int64_t a = off[p];
  const int64_t b = min(off[p + 1], cap);
  const double xcp = ypglb[3 * p], ycp = ypglb[3 * p + 1], zcp = ypglb[3 * p + 2];
  const double w1 = wp[3 * p], w2 = wp[3 * p + 1], w3 = wp[3 * p + 2];
  const double o1 = omgp[3 * p], o2 = omgp[3 * p + 1], o3 = omgp[3 * p + 2];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int contact = 0;
  for (int64_t l = a + threadIdx.x; l < b; l += PB) {
    const uint32_t n = lk_node[l];
    const int q = lk_dir[l];
    const double alpha = lk_alpha[l];
    if (n >= g.ncell || q < 1 || q >= NPOP) continue;  // never true for a consistent list
    int x, y, z;
    uint32_t row;
    node_coords(n, g, x, y, z, row);
    const int ix = cx(q), iy = cy(q), iz = cz(q);
    const double xp = (double)(x + 1) - 0.5, yp = (double)(y + 1) - 0.5, zp = (double)(z + 1) - 0.5;
    const double xc = near_image(xcp, xp, g.nx);
    const double yc = G.walls_y ? ycp : near_image(ycp, yp, g.ny);
    const double zc = near_image(zcp, zp, g.nz);
    const double xx0 = xp + (double)ix * alpha - xc;
    const double yy0 = yp + (double)iy * alpha - yc;
    const double zz0 = zp + (double)iz * alpha - zc;
    const double uwx = w1 + o2 * zz0 - o3 * yy0;
    const double uwy = w2 + o3 * xx0 - o1 * zz0;
    const double uwz = w3 + o1 * yy0 - o2 * xx0;
    const double uwpro = uwx * (double)ix + uwy * (double)iy + uwz * (double)iz;
    // x - c and x - 2c (periodic; outside the walls counts as solid)
    const int y1 = y - iy, y2 = y - 2 * iy;
    const bool out1 = g.walls_y && (y1 < 0 || y1 >= g.ny);
    const bool out2 = g.walls_y && (y2 < 0 || y2 >= g.ny);
    const uint32_t m1 = out1 ? n : (uint32_t)wrap_any(x - ix, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap_any(y1, g.ny) + (uint32_t)g.ny * (uint32_t)wrap_any(z - iz, g.nz));
    const uint32_t m2 = out2 ? n : (uint32_t)wrap_any(x - 2 * ix, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap_any(y2, g.ny) + (uint32_t)g.ny * (uint32_t)wrap_any(z - 2 * iz, g.nz));
    const bool sol1 = out1 || bit_at(mask, m1);
    const bool sol2 = out2 || bit_at(mask, m2);
    const int qo = opp(q);
    const double ff1 = (double)A[(uint64_t)q * g.ncell + n];
    const double ff2a = (double)A[(uint64_t)q * g.ncell + m1];
    const double ff3a = (double)A[(uint64_t)q * g.ncell + m2];
    const double ffo = (double)A[(uint64_t)qo * g.ncell + n];
    const double ffo1 = (double)A[(uint64_t)qo * g.ncell + m1];
    const double wt = q <= 6 ? ww1 : ww2;
    const double f9 = bounce_back_value(alpha, ff1, ff2a, ff3a, sol1, sol2, ffo, ffo1, wt, uwpro);
    lk_out[l] = f9;
    if (bit_at(mask, n)) ++contact;
    const double dff = ff1 + f9;
    const double dxm = dff * (double)ix, dym = dff * (double)iy, dzm = dff * (double)iz;
    acc[0] += dxm;
    acc[1] += dym;
    acc[2] += dzm;
    acc[3] += dzm * yy0 - dym * zz0;
    acc[4] += dxm * zz0 - dzm * xx0;
    acc[5] += dym * xx0 - dxm * yy0;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = PB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    fHIp[3 * p] = red[0][0];
    fHIp[3 * p + 1] = red[1][0];
    fHIp[3 * p + 2] = red[2][0];
    torqp[3 * p] = red[3][0];
    torqp[3 * p + 1] = red[4][0];
    torqp[3 * p + 2] = red[5][0];
  }
  if (contact) atomicAdd(&status[ST_CONTACT_LINKS], contact);
}

// Write each link's bounce-back value into the opposite population of the
// solid neighbour x + c (partlib.f90:1605-1621).
template <typename T>
__global__ void __launch_bounds__(PB)
    k_bb_scatter(T* __restrict__ A, Grid g, const int64_t* __restrict__ off, int64_t cap,
                 const uint32_t* __restrict__ lk_node, const uint8_t* __restrict__ lk_dir,
                 const double* __restrict__ lk_out) {
  const int p = blockIdx.x;
  const int64_t a = off[p];
  const int64_t b = min(off[p + 1], cap);
  for (int64_t l = a + threadIdx.x; l < b; l += PB) {
    const uint32_t n = lk_node[l];
    const int q = lk_dir[l];
    if (n >= g.ncell || q < 1 || q >= NPOP) continue;
    int x, y, z;
    uint32_t row;
    node_coords(n, g, x, y, z, row);
    const int yt = y + cy(q);
    if (g.walls_y && (yt < 0 || yt >= g.ny)) continue;
    const uint32_t t = (uint32_t)wrap_any(x + cx(q), g.nx) + (uint32_t)g.nx * ((uint32_t)wrap_any(yt, g.ny) + (uint32_t)g.ny * (uint32_t)wrap_any(z + cz(q), g.nz));
    A[(uint64_t)opp(q) * g.ncell + t] = (T)lk_out[l];
  }
}

// beads_lubforce (partlib.f90:2211-2292): one thread per particle sums its
// pair forces in the reference's order (pairs with lower ids first).  Each
// pair is evaluated as the reference does, from the lower id's side.
__global__ void __launch_bounds__(256)
    k_lubforce(const double* __restrict__ ypglb, int np, PGeom G, double mingap, double fpc,
               double stf0, double stf1, double* __restrict__ flubp) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= np) return;
  const double dist0 = 2.0 * G.rad + mingap;
  const double pre = (dist0 + 1e-6) * (dist0 + 1e-6);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int b = 0; b < np; ++b) {
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
    const double fpd = lub_fpd(dist, G.rad, mingap, fpc, stf0, stf1);
    if (a == lo) {
      fx += fpd * dx;
      fy += fpd * dy;
      fz += fpd * dz;
    } else {
      fx -= fpd * dx;
      fy -= fpd * dy;
      fz -= fpd * dz;
    }
  }
  flubp[3 * a] = fx;
  flubp[3 * a + 1] = fy;
  flubp[3 * a + 2] = fz;
}

struct MoveArgs {
  double *yp, *wp, *wpp, *omgp, *omgpp, *fHIp, *flubp, *torqp, *forcep, *forcepp, *torqpp, *dwdt,
      *domgdt, *thetap;
  double volp, rhog, amp, aip;
  double bfx, bfy, bfz;
};

// beads_move (partlib.f90:2295-2394): second-order Newton-Euler update.
__global__ void __launch_bounds__(256) k_move(MoveArgs m, int np, int first) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double dt = 1.0;
  const double bf[3] = {m.bfx, m.bfy, m.bfz};
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    if (first) m.thetap[q] = 0.0;
    m.wpp[q] = m.wp[q];
    m.omgpp[q] = m.omgp[q];
  }
  double dwn[3], don[3];
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    double F = m.fHIp[q] + m.flubp[q];
    if (c == 1) F = F - m.volp * m.rhog;
    if (bf[c] != 0.0) F = F + m.volp * bf[c];
    m.forcep[q] = F;
    if (first) m.forcepp[q] = F;
    dwn[c] = 0.5 * (F + m.forcepp[q]) / m.amp;
    if (first) m.dwdt[q] = dwn[c];
    m.wp[q] = m.wpp[q] + 0.5 * dt * (m.dwdt[q] + dwn[c]);
  }
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    m.yp[q] = m.yp[q] + 0.5 * dt * (m.wp[q] + m.wpp[q]);
  }
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    if (first) m.torqpp[q] = m.torqp[q];
    don[c] = 0.5 * (m.torqp[q] + m.torqpp[q]) / m.aip;
    if (first) m.domgdt[q] = don[c];
    m.omgp[q] = m.omgpp[q] + 0.5 * dt * (m.domgdt[q] + don[c]);
  }
  for (int c = 0; c < 3; ++c) {
    const int q = 3 * p + c;
    m.thetap[q] = m.thetap[q] + 0.5 * dt * (m.omgp[q] + m.omgpp[q]);
    m.dwdt[q] = dwn[c];
    m.domgdt[q] = don[c];
    m.forcepp[q] = m.forcep[q];
    m.torqpp[q] = m.torqp[q];
  }
}

// beads_redistribute, one-rank content (partlib.f90:2600-2626): periodic wrap,
// ypglbp <- ypglb, ypglb <- yp.
__global__ void __launch_bounds__(256)
    k_wrap(double* __restrict__ yp, double* __restrict__ ypglb, double* __restrict__ ypglbp, int np,
           PGeom G, int32_t* __restrict__ status) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
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
  if (bad) atomicAdd(&status[ST_NONFINITE], 1);
}

// ovlpchk (partlib.f90:118-171): flag the lower id of any pair closer than 2r.
__global__ void __launch_bounds__(256)
    k_overlap(const double* __restrict__ ypglb, int np, PGeom G, int istep,
              int32_t* __restrict__ status) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= np - 1) return;
  const double xc = ypglb[3 * a], yc = ypglb[3 * a + 1], zc = ypglb[3 * a + 2];
  for (int b = a + 1; b < np; ++b) {
    const double xb = near_image(ypglb[3 * b], xc, G.nx);
    const double yb = G.walls_y ? ypglb[3 * b + 1] : near_image(ypglb[3 * b + 1], yc, G.ny);
    const double zb = near_image(ypglb[3 * b + 2], zc, G.nz);
    const double dist = sqrt((xc - xb) * (xc - xb) + (yc - yb) * (yc - yb) + (zc - zb) * (zc - zb));
    const double hgap = dist - 2.0 * G.rad;
    if (hgap < 0.0) {
      atomicAdd(&status[ST_OVERLAP], 1);
      atomicCAS(&status[ST_OVERLAP_STEP], 0, istep);
      break;
    }
  }
}

struct FillArgs {
  const double *ypglbp, *wp, *wpp, *omgp, *omgpp, *mpos;
  const int32_t* isn_old;
  const uint32_t *mask_new, *mask_old;
  double rad, rho0, ww0, ww1, ww2;
};

// beads_filling (partlib.f90:2846-3094): nodes solid in the previous map and
// fluid in the new one get the equilibrium at the rigid-body velocity and the
// mean density of their fluid neighbours, plus the non-equilibrium part of
// the neighbour along the back-traced surface normal.  One workgroup scans
// the previous-step box of particle p.
template <typename T, bool WALLS>
__global__ void __launch_bounds__(PB)
    k_filling(T* __restrict__ A, Grid g, PGeom G, FillArgs fa, int np, int32_t* __restrict__ status) {
  __shared__ BoxInfo B;
  const int p = blockIdx.x;
  if (p >= np) return;
  if (threadIdx.x == 0) box_info(fa.mpos, p, G, B);
  __syncthreads();
  if (!B.ok) return;
  const int nbox = B.ax.cnt * B.ay.cnt * B.az.cnt;
  int nfill = 0;
  for (int t = threadIdx.x; t < nbox; t += PB) {
    const int tx = t % B.ax.cnt, rest = t / B.ax.cnt, ty = rest % B.ay.cnt, tz = rest / B.ay.cnt;
    const int k = wrap_any(B.az.lo + tz, G.nz);
    const int j = G.walls_y ? B.ay.lo + ty : wrap_any(B.ay.lo + ty, G.ny);
    const int i = wrap_any(B.ax.lo + tx, G.nx);
    if (j < 0 || j >= G.ny) continue;
    const uint32_t n = (uint32_t)i + (uint32_t)G.nx * ((uint32_t)j + (uint32_t)G.ny * (uint32_t)k);
    if (fa.isn_old[n] != p + 1 || bit_at(fa.mask_new, n)) continue;
    ++nfill;
    const int id = p;
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
    double fn[NPOP];
    {
      const int a = i + cx(ipmx), bb = j + cy(ipmx), c = k + cz(ipmx);
      const bool out = WALLS && (bb < 0 || bb >= g.ny);
      const uint32_t q = out ? n : (uint32_t)wrap_any(a, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap_any(bb, g.ny) + (uint32_t)g.ny * (uint32_t)wrap_any(c, g.nz));
      if (!out && !bit_at(fa.mask_new, q) && !bit_at(fa.mask_old, q)) {
        int qx, qy, qz;
        uint32_t row;
        node_coords(q, g, qx, qy, qz, row);
        double f9[NPOP];
        load_pulled<T, double, WALLS>(A, g, qx, qy, qz, q, f9);
        double rho9 = 0.0, u9 = 0.0, v9 = 0.0, w9 = 0.0;
        for (int ip = 0; ip < NPOP; ++ip) {
          rho9 = rho9 + f9[ip];
          u9 = u9 + (double)cx(ip) * f9[ip];
          v9 = v9 + (double)cy(ip) * f9[ip];
          w9 = w9 + (double)cz(ip) * f9[ip];
        }
        double feq[NPOP];
        feq_point(u9, v9, w9, rho9, fa.ww0, fa.ww1, fa.ww2, feq);
        for (int ip = 0; ip < NPOP; ++ip) fn[ip] = f9[ip] - feq[ip];
      } else {
        for (int ip = 0; ip < NPOP; ++ip) fn[ip] = 0.0;
      }
    }
    int nghb = 0;
    double rho9 = 0.0;
    for (int ip = 1; ip < NPOP; ++ip) {
      const int a = i + cx(ip), bb = j + cy(ip), c = k + cz(ip);
      if (WALLS && (bb < 0 || bb >= g.ny)) continue;
      const uint32_t q = (uint32_t)wrap_any(a, g.nx) + (uint32_t)g.nx * ((uint32_t)wrap_any(bb, g.ny) + (uint32_t)g.ny * (uint32_t)wrap_any(c, g.nz));
      if (!bit_at(fa.mask_new, q) && !bit_at(fa.mask_old, q)) {
        ++nghb;
        int qx, qy, qz;
        uint32_t row;
        node_coords(q, g, qx, qy, qz, row);
        double f9[NPOP];
        load_pulled<T, double, WALLS>(A, g, qx, qy, qz, q, f9);
        for (int ii = 0; ii < NPOP; ++ii) rho9 = rho9 + f9[ii];
      }
    }
    rho9 = (nghb > 0) ? rho9 / (double)nghb : fa.rho0;
    const double xx0 = xp - xc, yy0 = yp - yc, zz0 = zp - zc;
    const double W1 = fa.wp[3 * id], W2 = fa.wp[3 * id + 1], W3 = fa.wp[3 * id + 2];
    const double O1 = fa.omgp[3 * id], O2 = fa.omgp[3 * id + 1], O3 = fa.omgp[3 * id + 2];
    const double u9 = W1 + O2 * zz0 - O3 * yy0;
    const double v9 = W2 + O3 * xx0 - O1 * zz0;
    const double w9 = W3 + O1 * yy0 - O2 * xx0;
    double feq[NPOP];
    feq_point(u9, v9, w9, rho9, fa.ww0, fa.ww1, fa.ww2, feq);
    for (int ip = 0; ip < NPOP; ++ip) store_phase<T, WALLS>(A, g, ip, i, j, k, n, true, fn[ip] + feq[ip]);
  }
  if (nfill) atomicAdd(&status[ST_REFILL], nfill);
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

__global__ void k_links_dense(Grid g, const int64_t* __restrict__ off, int64_t cap,
                              const uint32_t* __restrict__ lk_node, const uint8_t* __restrict__ lk_dir,
                              const double* __restrict__ lk_alpha, int32_t* __restrict__ ilink,
                              int32_t* __restrict__ mlink, double* __restrict__ alink) {
  const int p = blockIdx.x;
  const int64_t a = off[p], b = min(off[p + 1], cap);
  for (int64_t l = a + threadIdx.x; l < b; l += blockDim.x) {
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
  return G;
}

hipError_t launch_clear_map(Ctx& c, int target) {
  Particles& P = c.pt;
  if (P.map_np[target] > 0)
    hipLaunchKernelGGL(k_clear_map, dim3(P.map_np[target]), dim3(PB), 0, c.stream, P.map_pos[target],
                       P.map_np[target], pgeom(c), P.isn[target], P.mask[target]);
  return hipGetLastError();
}

hipError_t launch_links(Ctx& c, int target) {
  Particles& P = c.pt;
  const PGeom G = pgeom(c);
  hipLaunchKernelGGL(k_links_count, dim3(P.np), dim3(PB), 0, c.stream, P.ypglb, P.np, G, P.lk_cnt,
                     P.sol_cnt, c.d_status);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, c.stream, P.lk_cnt, P.sol_cnt, P.np,
                     P.lk_off, P.lk_cap, c.d_status);
  hipLaunchKernelGGL(k_links_write, dim3(P.np), dim3(PB), 0, c.stream, P.ypglb, P.np, G, P.lk_off,
                     P.lk_cap, P.isn[target], P.mask[target], P.lk_node, P.lk_dir, P.lk_alpha,
                     c.d_status);
  hipMemcpyAsync(P.map_pos[target], P.ypglb, sizeof(double) * 3 * P.np, hipMemcpyDeviceToDevice, c.stream);
  P.map_np[target] = P.np;
  return hipGetLastError();
}

template <typename T>
static void bb_dispatch(Ctx& c) {
  Particles& P = c.pt;
  T* A = (T*)c.pop[c.cur];
  hipLaunchKernelGGL((k_bounce_back<T>), dim3(P.np), dim3(PB), 0, c.stream, A, c.g, pgeom(c),
                     P.mask[P.map_cur], P.lk_off, P.lk_cap, P.lk_node, P.lk_dir, P.lk_alpha, P.lk_out,
                     P.ypglb, P.wp, P.omgp, c.p.ww1, c.p.ww2, P.fHIp, P.torqp, c.d_status);
  hipLaunchKernelGGL((k_bb_scatter<T>), dim3(P.np), dim3(PB), 0, c.stream, A, c.g, P.lk_off, P.lk_cap,
                     P.lk_node, P.lk_dir, P.lk_out);
}

hipError_t launch_bounce_back(Ctx& c) {
  if (c.prec == MFC_PREC_FP64)
    bb_dispatch<double>(c);
  else
    bb_dispatch<float>(c);
  return hipGetLastError();
}

hipError_t launch_lubforce(Ctx& c) {
  Particles& P = c.pt;
  const double fpc = c.p.volp * c.p.rho0 * 9.8 / c.p.gscale;
  hipLaunchKernelGGL(k_lubforce, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.ypglb, P.np,
                     pgeom(c), c.p.mingap, fpc, c.p.stf0, c.p.stf1, P.flubp);
  return hipGetLastError();
}

hipError_t launch_move(Ctx& c, int first) {
  Particles& P = c.pt;
  MoveArgs m;
  m.yp = P.yp; m.wp = P.wp; m.wpp = P.wpp; m.omgp = P.omgp; m.omgpp = P.omgpp;
  m.fHIp = P.fHIp; m.flubp = P.flubp; m.torqp = P.torqp; m.forcep = P.forcep;
  m.forcepp = P.forcepp; m.torqpp = P.torqpp; m.dwdt = P.dwdt; m.domgdt = P.domgdt;
  m.thetap = P.thetap;
  m.volp = c.p.volp; m.rhog = c.p.rhog; m.amp = c.p.amp; m.aip = c.p.aip;
  m.bfx = c.p.body_force[0]; m.bfy = c.p.body_force[1]; m.bfz = c.p.body_force[2];
  hipLaunchKernelGGL(k_move, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, m, P.np, first);
  return hipGetLastError();
}

hipError_t launch_redistribute(Ctx& c, int istep) {
  Particles& P = c.pt;
  hipLaunchKernelGGL(k_wrap, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.yp, P.ypglb,
                     P.ypglbp, P.np, pgeom(c), c.d_status);
  hipLaunchKernelGGL(k_overlap, dim3((P.np + 255) / 256), dim3(256), 0, c.stream, P.ypglb, P.np,
                     pgeom(c), istep, c.d_status);
  return hipGetLastError();
}

hipError_t launch_filling(Ctx& c) {
  Particles& P = c.pt;
  const int old = P.map_old;
  if (P.map_np[old] == 0 || old == P.map_cur) return hipSuccess;
  FillArgs fa;
  fa.ypglbp = P.ypglbp; fa.wp = P.wp; fa.wpp = P.wpp; fa.omgp = P.omgp; fa.omgpp = P.omgpp;
  fa.mpos = P.map_pos[old];
  fa.isn_old = P.isn[old];
  fa.mask_new = P.mask[P.map_cur];
  fa.mask_old = P.mask[old];
  fa.rad = c.p.rad; fa.rho0 = c.p.rho0; fa.ww0 = c.p.ww0; fa.ww1 = c.p.ww1; fa.ww2 = c.p.ww2;
  const int np = P.map_np[old];
  if (c.prec == MFC_PREC_FP64) {
    double* A = (double*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_filling<double, true>), dim3(np), dim3(PB), 0, c.stream, A, c.g, pgeom(c), fa, np, c.d_status);
    else
      hipLaunchKernelGGL((k_filling<double, false>), dim3(np), dim3(PB), 0, c.stream, A, c.g, pgeom(c), fa, np, c.d_status);
  } else {
    float* A = (float*)c.pop[c.cur];
    if (c.g.walls_y)
      hipLaunchKernelGGL((k_filling<float, true>), dim3(np), dim3(PB), 0, c.stream, A, c.g, pgeom(c), fa, np, c.d_status);
    else
      hipLaunchKernelGGL((k_filling<float, false>), dim3(np), dim3(PB), 0, c.stream, A, c.g, pgeom(c), fa, np, c.d_status);
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
  hipLaunchKernelGGL(k_links_dense, dim3(P.np), dim3(256), 0, c.stream, c.g, P.lk_off, P.lk_cap,
                     P.lk_node, P.lk_dir, P.lk_alpha, ilink, mlink, alink);
  return hipGetLastError();
}

}  // namespace mfc
