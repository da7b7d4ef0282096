// lbm_particle_math.h -- per-link / per-node particle arithmetic shared by
// the HIP kernels and the CPU check build.  Compiled without floating-point
// contraction so geometric tests and wall fractions are evaluated exactly as
// written (node maps are compared bit for bit).
#pragma once
#include <math.h>

#include "lbm_common.h"

namespace mfc {

// Centre bases that the reference's beads_links effectively uses
// (partlib.f90:952-1043).  There the centre coordinates persist across the
// k/j/i loops and are re-imaged with two conditional shifts at every node.
// Working the state machine through: z is imaged from its original value at
// every node; y from its original value during the first j-sweep (first k in
// the box) and from `y_rest` afterwards; x from its original value during
// the first i-sweep (first (k,j) in the box) and from `x_rest` afterwards,
// where rest = c if no upward shift happened at the last node of a sweep,
// else fl(c + N) - N (the round trip is exact for c >= N/2, not below).
MFC_HD double sweep_rest_base(double c, int n) {
  const double s_end = near_image(c, (double)n - 0.5, n);
  return (s_end != c) ? s_end - (double)n : c;
}

struct LinkGeom {
  double xx0, yy0, zz0, rr0;   // node minus (imaged) centre and its square norm
  double ximg, yimg, zimg;     // imaged centre used for this node
};

// One boundary link's wall fraction (partlib.f90:1013-1024): alpha is the
// fraction of the link from the fluid node to the sphere surface.
MFC_HD double link_alpha(const LinkGeom& G, double xx1, double yy1, double zz1, double rr1,
                         double r2) {
  const double aa = G.rr0 + rr1 - 2.0 * (G.xx0 * xx1 + G.yy0 * yy1 + G.zz0 * zz1);
  const double bb = xx1 * (G.xx0 - xx1) + yy1 * (G.yy0 - yy1) + zz1 * (G.zz0 - zz1);
  const double cc = rr1 - r2;
  const double a0 = bb / aa;
  const double a1 = sqrt(a0 * a0 - cc / aa);
  double alpha = -a0 + a1;
  alpha = 1.0 - alpha;
  return alpha;
}

// Interpolated bounce-back value of one link (partlib.f90:1425-1536).
//   ff1 = f_q(x), ff2a = f_q(x - c), ff3a = f_q(x - 2c), sol1/sol2 = solidity
//   of x - c and x - 2c, ffo = f_qbar(x), ffo1 = f_qbar(x - c);
//   wt = lattice weight of q, uwpro = wall velocity . c_q.
MFC_HD double bounce_back_value(double alpha, double ff1, double ff2a, double ff3a, bool sol1,
                                bool sol2, double ffo, double ffo1, double wt, double uwpro) {
  double f9;
  if (alpha <= 0.5) {
    if (sol1) {
      f9 = ff1 - 6.0 * wt * uwpro;
    } else if (sol2) {
      f9 = 2.0 * alpha * (ff1 - ff2a) + ff2a - 6.0 * wt * uwpro;
    } else {
      const double c1 = alpha * (1.0 + 2.0 * alpha);
      const double c2 = 1.0 - 4.0 * alpha * alpha;
      const double c3 = -alpha * (1.0 - 2.0 * alpha);
      f9 = c1 * ff1 + c2 * ff2a + c3 * ff3a - 6.0 * wt * uwpro;
    }
  } else {
    if (sol1) {
      const double c1 = 0.5 / alpha;
      const double c2 = 1.0 - c1;
      f9 = c1 * ff1 + c2 * ffo - 6.0 * wt * c1 * uwpro;
    } else {
      const double c1 = 1.0 / alpha / (2.0 * alpha + 1.0);
      const double c2 = (2.0 * alpha - 1.0) / alpha;
      const double c3 = 1.0 - c1 - c2;
      f9 = c1 * ff1 + c2 * ffo + c3 * ffo1 - 6.0 * wt * c1 * uwpro;
    }
  }
  return f9;
}

// Deviation equilibrium at one node (feqpnt, partlib.f90:3222-3247).
MFC_HD void feq_point(double u, double v, double w, double rho, double ww0, double ww1, double ww2,
                      double (&feq)[NPOP]) {
  const double usqr = 1.5 * (u * u + v * v + w * w);
  feq[0] = ww0 * (rho - usqr);
#pragma unroll
  for (int i = 1; i < NPOP; ++i) {
    const double G = (double)cx(i) * u + (double)cy(i) * v + (double)cz(i) * w;
    feq[i] = (i <= 6 ? ww1 : ww2) * (rho + 3.0 * G + 4.5 * G * G - usqr);
  }
}

// Direction along which an uncovered node takes its non-equilibrium part
// (partlib.f90:2853-2925): back-trace the node to the old sphere surface with
// minus the mean particle velocity, apply the rotation, and pick the lattice
// direction best aligned with the outward normal there (axis directions
// weighted by sqrt(2), ties to the lowest index).
MFC_HD int refill_direction(double xp, double yp, double zp, double xc, double yc, double zc,
                            double w1, double w2, double w3, double o1, double o2, double o3,
                            double rad) {
  const double aa = w1 * w1 + w2 * w2 + w3 * w3;
  const double bb = (xp - xc) * w1 + (yp - yc) * w2 + (zp - zc) * w3;
  const double cc = (xp - xc) * (xp - xc) + (yp - yc) * (yp - yc) + (zp - zc) * (zp - zc) -
                    (rad * 1.0) * (rad * 1.0);
  const double d0 = bb / aa;
  const double d1 = sqrt(d0 * d0 - cc / aa);
  double ddt = -d0 + d1;
  if (ddt < 0.0) ddt = 0.0;
  if (ddt > 1.0) ddt = 1.0;
  const double x1 = xp + w1 * ddt, y1 = yp + w2 * ddt, z1 = zp + w3 * ddt;
  const double x2 = x1 + (o2 * (z1 - zc) - o3 * (y1 - yc)) * ddt;
  const double y2 = y1 + (o3 * (x1 - xc) - o1 * (z1 - zc)) * ddt;
  const double z2 = z1 + (o1 * (y1 - yc) - o2 * (x1 - xc)) * ddt;
  const double xx0 = x2 - xc, yy0 = y2 - yc, zz0 = z2 - zc;
  double prod0 = -100.0;
  int ipmx = 1;
  const double sq2 = sqrt(2.0);
  for (int i = 1; i < NPOP; ++i) {
    double prod = (double)cx(i) * xx0 + (double)cy(i) * yy0 + (double)cz(i) * zz0;
    if (i <= 6) prod = prod * sq2;
    if (prod > prod0) {
      ipmx = i;
      prod0 = prod;
    }
  }
  return ipmx;
}

// Lubrication force magnitude over distance for one pair (partlib.f90:2249-2267).
MFC_HD double lub_fpd(double dist, double rad, double mingap, double fpc, double stf0, double stf1) {
  const double dist0 = 2.0 * rad + mingap;
  const double hgap = dist - 2.0 * rad;
  double fp = 0.0;
  if (dist < dist0) {
    double d1 = 0.0;
    if (hgap < 0.0) d1 = -hgap;
    const double t = (dist0 - dist) / mingap;
    fp = (t * t / stf0 + d1 / mingap / stf1) * fpc;
  }
  return fp / dist;
}

}  // namespace mfc
