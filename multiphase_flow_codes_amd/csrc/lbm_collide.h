// lbm_collide.h -- per-node D3Q19 MRT collision with the merged forcing term,
// written once for device kernels and the CPU check build.
//
// Semantics (3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/collision.f90:8-210):
//   Fbar_0 = -u.F,  Fbar_i = w_i (3 c.F + 9 (c.F)(c.u) - 3 u.F);
//   g = f + Fbar/2;  non-conserved moments of g relax toward the equilibria
//   built from the stored density rho and momentum u; the post-collision
//   conserved moments ARE rho and u (not the moments of g); f <- f_post + Fbar/2.
//
// Formulation used here: the collision is applied as a correction,
//   f_out = g - M^T D^-1 dm + Fbar/2,   dm = m(g) - m_post,
// with dm_k = s_k (m_k(g) - m_k^eq) for the 15 non-conserved moments and
// dm = (sum g - rho, sum c g - u) for the four conserved ones.  This is
// algebraically the reference's back-transform M^-1 m_post + Fbar/2, but
// never rebuilds f from O(1) moment sums, so the FP32 path keeps the
// precision of the populations themselves.  Moment basis: d'Humieres et al.
// 2002 D3Q19 (rho, e, eps, j, q, 3p_xx, 3pi_xx, p_ww, pi_ww, p_xy, p_yz,
// p_xz, m), with squared row norms D = (19, 2394, 252, 10, 40, ..., 36, 72,
// 12, 24, 4, 4, 4, 8, 8, 8).
#pragma once
#include "lbm_common.h"

namespace mfc {

// Direction-class partial sums of a population vector.
template <typename R>
struct ClassSums {
  R s0;              // rest
  R sx, sy, sz;      // axis pair sums  f(+a) + f(-a)
  R dx, dy, dz;      // axis pair diffs f(+a) - f(-a)
  // edge quartets: even = same-sign pair sum, odd = opposite-sign pair sum,
  // and the two signed differences carrying the first axis' sign
  R xy_e, xy_o, xy_de, xy_do;
  R xz_e, xz_o, xz_de, xz_do;
  R yz_e, yz_o, yz_de, yz_do;
};

template <typename R>
MFC_HD ClassSums<R> class_sums(const R (&g)[NPOP]) {
  ClassSums<R> s;
  s.s0 = g[0];
  s.sx = g[1] + g[2]; s.dx = g[1] - g[2];
  s.sy = g[3] + g[4]; s.dy = g[3] - g[4];
  s.sz = g[5] + g[6]; s.dz = g[5] - g[6];
  // xy: 7(+,+) 8(-,+) 9(+,-) 10(-,-)
  s.xy_e = g[7] + g[10]; s.xy_o = g[8] + g[9];
  s.xy_de = g[7] - g[10]; s.xy_do = g[9] - g[8];
  // xz: 11(+,+) 12(-,+) 13(+,-) 14(-,-)
  s.xz_e = g[11] + g[14]; s.xz_o = g[12] + g[13];
  s.xz_de = g[11] - g[14]; s.xz_do = g[13] - g[12];
  // yz: 15(+,+) 16(-,+) 17(+,-) 18(-,-)
  s.yz_e = g[15] + g[18]; s.yz_o = g[16] + g[17];
  s.yz_de = g[15] - g[18]; s.yz_do = g[17] - g[16];
  return s;
}

// Density and momentum of a population vector.
template <typename R>
MFC_HD void density_momentum(const R (&f)[NPOP], R& rho, R& jx, R& jy, R& jz) {
  const ClassSums<R> s = class_sums(f);
  const R exy = s.xy_e + s.xy_o, exz = s.xz_e + s.xz_o, eyz = s.yz_e + s.yz_o;
  rho = s.s0 + (s.sx + s.sy + s.sz) + (exy + exz + eyz);
  jx = s.dx + (s.xy_de + s.xy_do) + (s.xz_de + s.xz_do);
  jy = s.dy + (s.xy_de - s.xy_do) + (s.yz_de + s.yz_do);
  jz = s.dz + (s.xz_de - s.xz_do) + (s.yz_de - s.yz_do);
}

// Forcing term Fbar_i of the merged collision (collision.f90:45-62).
template <typename R>
MFC_HD void forcing_term(R (&Fb)[NPOP], R ux, R uy, R uz, R Fx, R Fy, R Fz, R w1, R w2) {
  const R G3 = ux * Fx + uy * Fy + uz * Fz;
  Fb[0] = -G3;
#pragma unroll
  for (int i = 1; i < NPOP; ++i) {
    const R G1 = (R)cx(i) * Fx + (R)cy(i) * Fy + (R)cz(i) * Fz;
    const R G2 = (R)cx(i) * ux + (R)cy(i) * uy + (R)cz(i) * uz;
    const R w = (i <= 6) ? w1 : w2;
    Fb[i] = w * ((R)3 * G1 + (R)9 * G1 * G2 - (R)3 * G3);
  }
}

// Corrections corr = M^T D^-1 dm of the collision (see the header comment)
// from the class sums of g = f + Fbar/2.
//   rho, u*  the stored density and momentum the reference reads from its
//            rho/ux/uy/uz arrays
// EXACT_CONS: rho and u are known to be the macrovar moments of f itself,
// rho = sum f and u = sum c f + Fm/2 (Fm: the force macrovar used).  Then the
// conserved corrections are exactly dR = 0 and dJ = (F - Fm)/20, since
// sum Fbar = 0 and sum c Fbar = F; taking them in closed form avoids the
// cancellation of two rounded 19-term sums.
template <typename R, bool EXACT_CONS>
MFC_HD void mrt_corrections(const ClassSums<R>& s, R rho, R ux, R uy, R uz, R Fx, R Fy, R Fz,
                            const CollideConsts& k, R mFx, R mFy, R mFz, R (&corr)[NPOP]) {
  const R sa = s.sx + s.sy + s.sz;
  const R exy = s.xy_e + s.xy_o, exz = s.xz_e + s.xz_o, eyz = s.yz_e + s.yz_o;
  const R se = exy + exz + eyz;
  // odd moments of each edge class along its two axes
  const R ex_xy = s.xy_de + s.xy_do, ey_xy = s.xy_de - s.xy_do;
  const R ex_xz = s.xz_de + s.xz_do, ez_xz = s.xz_de - s.xz_do;
  const R ey_yz = s.yz_de + s.yz_do, ez_yz = s.yz_de - s.yz_do;

  // moments of g
  const R m_rho = s.s0 + sa + se;
  const R m_jx = s.dx + ex_xy + ex_xz;
  const R m_jy = s.dy + ey_xy + ey_yz;
  const R m_jz = s.dz + ez_xz + ez_yz;
  const R m_e = (R)-30 * s.s0 + (R)-11 * sa + (R)8 * se;
  const R m_eps = (R)12 * s.s0 + (R)-4 * sa + se;
  const R m_qx = (R)-4 * s.dx + ex_xy + ex_xz;
  const R m_qy = (R)-4 * s.dy + ey_xy + ey_yz;
  const R m_qz = (R)-4 * s.dz + ez_xz + ez_yz;
  const R m_pxx = (R)2 * s.sx - s.sy - s.sz + exy + exz - (R)2 * eyz;
  const R m_pixx = (R)-4 * s.sx + (R)2 * (s.sy + s.sz) + exy + exz - (R)2 * eyz;
  const R m_pww = s.sy - s.sz + exy - exz;
  const R m_piww = (R)-2 * (s.sy - s.sz) + exy - exz;
  const R m_pxy = s.xy_e - s.xy_o;
  const R m_pxz = s.xz_e - s.xz_o;
  const R m_pyz = s.yz_e - s.yz_o;
  const R m_mx = ex_xy - ex_xz;
  const R m_my = ey_yz - ey_xy;
  const R m_mz = ez_xz - ez_yz;

  // equilibria from the stored rho, u (collision.f90:66-81)
  const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  const R u2 = ux2 + uy2 + uz2;
  const R eq_e = (R)-11 * rho + (R)19 * u2;
  const R eq_eps = (R)k.omegepsl * rho + (R)k.omegepslj * u2;
  const R eq_q = (R)(-2.0 / 3.0);
  const R eq_pxx = (R)2 * ux2 - uy2 - uz2;
  const R eq_pww = uy2 - uz2;

  // moment corrections dm, already divided by the row norms
  R dR = (m_rho - rho) * (R)(1.0 / 19.0);
  R dJx = (m_jx - ux) * (R)0.1, dJy = (m_jy - uy) * (R)0.1, dJz = (m_jz - uz) * (R)0.1;
  if (EXACT_CONS) {
    dR = (R)0;
    dJx = (Fx - mFx) * (R)0.05;
    dJy = (Fy - mFy) * (R)0.05;
    dJz = (Fz - mFz) * (R)0.05;
  }
  const R dE = (R)k.s1 * (m_e - eq_e) * (R)(1.0 / 2394.0);
  const R dP = (R)k.s2 * (m_eps - eq_eps) * (R)(1.0 / 252.0);
  const R dQx = (R)k.s4 * (m_qx - eq_q * ux) * (R)(1.0 / 40.0);
  const R dQy = (R)k.s4 * (m_qy - eq_q * uy) * (R)(1.0 / 40.0);
  const R dQz = (R)k.s4 * (m_qz - eq_q * uz) * (R)(1.0 / 40.0);
  const R dXX = (R)k.s9 * (m_pxx - eq_pxx) * (R)(1.0 / 36.0);
  const R dPX = (R)k.s10 * (m_pixx - (R)k.omegxx * eq_pxx) * (R)(1.0 / 72.0);
  const R dWW = (R)k.s9 * (m_pww - eq_pww) * (R)(1.0 / 12.0);
  const R dPW = (R)k.s10 * (m_piww - (R)k.omegxx * eq_pww) * (R)(1.0 / 24.0);
  const R dSxy = (R)k.s13 * (m_pxy - ux * uy) * (R)0.25;
  const R dSxz = (R)k.s13 * (m_pxz - ux * uz) * (R)0.25;
  const R dSyz = (R)k.s13 * (m_pyz - uy * uz) * (R)0.25;
  const R dMx = (R)k.s16 * m_mx * (R)0.125;
  const R dMy = (R)k.s16 * m_my * (R)0.125;
  const R dMz = (R)k.s16 * m_mz * (R)0.125;

  // M^T applied to the scaled corrections, grouped by direction class
  const R base_a = dR - (R)11 * dE - (R)4 * dP;
  const R base_e = dR + (R)8 * dE + dP;
  const R ax = dJx - (R)4 * dQx, ay = dJy - (R)4 * dQy, az = dJz - (R)4 * dQz;
  const R bx = dJx + dQx, by = dJy + dQy, bz = dJz + dQz;
  const R cxx = base_a + (R)2 * dXX - (R)4 * dPX;
  const R cyy = base_a - dXX + (R)2 * dPX + dWW - (R)2 * dPW;
  const R czz = base_a - dXX + (R)2 * dPX - dWW + (R)2 * dPW;
  const R exy0 = base_e + dXX + dPX + dWW + dPW;
  const R exz0 = base_e + dXX + dPX - dWW - dPW;
  const R eyz0 = base_e - (R)2 * dXX - (R)2 * dPX;

  corr[0] = dR - (R)30 * dE + (R)12 * dP;
  corr[1] = cxx + ax; corr[2] = cxx - ax;
  corr[3] = cyy + ay; corr[4] = cyy - ay;
  corr[5] = czz + az; corr[6] = czz - az;
  // xy edges: + cx*bx + cy*by + cx*cy*dSxy + cx*dMx - cy*dMy
  corr[7] = exy0 + bx + by + dSxy + dMx - dMy;
  corr[8] = exy0 - bx + by - dSxy - dMx - dMy;
  corr[9] = exy0 + bx - by - dSxy + dMx + dMy;
  corr[10] = exy0 - bx - by + dSxy - dMx + dMy;
  // xz edges: + cx*bx + cz*bz + cx*cz*dSxz - cx*dMx + cz*dMz
  corr[11] = exz0 + bx + bz + dSxz - dMx + dMz;
  corr[12] = exz0 - bx + bz - dSxz + dMx + dMz;
  corr[13] = exz0 + bx - bz - dSxz - dMx - dMz;
  corr[14] = exz0 - bx - bz + dSxz + dMx - dMz;
  // yz edges: + cy*by + cz*bz + cy*cz*dSyz + cy*dMy - cz*dMz
  corr[15] = eyz0 + by + bz + dSyz + dMy - dMz;
  corr[16] = eyz0 - by + bz - dSyz - dMy - dMz;
  corr[17] = eyz0 + by - bz - dSyz + dMy + dMz;
  corr[18] = eyz0 - by - bz + dSyz - dMy + dMz;
}

// In-place MRT collision of one fluid node.
//   f        post-streaming populations (deviations), overwritten with the
//            post-collision populations
//   F*       the forcing field at this node for the current step
// (rho, u*, EXACT_CONS and mF* as for mrt_corrections)
template <typename R, bool EXACT_CONS = false>
MFC_HD void mrt_collide(R (&f)[NPOP], R rho, R ux, R uy, R uz, R Fx, R Fy, R Fz,
                        const CollideConsts& k, R mFx = 0, R mFy = 0, R mFz = 0) {
  R Fb[NPOP];
  forcing_term<R>(Fb, ux, uy, uz, Fx, Fy, Fz, (R)k.ww1, (R)k.ww2);
  R g[NPOP];
#pragma unroll
  for (int i = 0; i < NPOP; ++i) g[i] = f[i] + (R)0.5 * Fb[i];
  R corr[NPOP];
  mrt_corrections<R, EXACT_CONS>(class_sums(g), rho, ux, uy, uz, Fx, Fy, Fz, k, mFx, mFy, mFz, corr);
#pragma unroll
  for (int i = 0; i < NPOP; ++i) f[i] = (g[i] - corr[i]) + (R)0.5 * Fb[i];
}

}  // namespace mfc
