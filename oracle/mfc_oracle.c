/*
 * mfc_oracle.c -- CPU FP64 oracle (test infrastructure only) for the D3Q19
 * MRT lattice-Boltzmann step with finite-size moving spheres.
 *
 * Clean-room restatement of the reference algorithm, one rank, periodic box.
 * Each function names the reference routine (BOX/ =
 * 3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/) whose arithmetic and
 * ordering it reproduces.  Where the reference's behaviour depends on its MPI
 * halo copies (beads_collision) the single-rank behaviour is reproduced
 * exactly (bb_mode 0); bb_mode 1 evaluates every bounce-back link against one
 * snapshot of the post-collision field (the GPU library's default order; the
 * GPU's reference order follows bb_mode 0).  The two agree whenever no
 * boundary link sits on a node that is solid for another sphere (near
 * contact).
 *
 * What pins each function (tests/test_oracle_ref.py, tests/test_oracle_pins.py,
 * tests/test_oracle_pins_particles.py; their headers give the detail):
 *   orc_collision_mrt, orc_macrovar, orc_rhoupdat  the compiled reference
 *                      (collision.f90, para.f90), mass/momentum balance, BGK limit
 *   orc_streaming      numpy.roll, brute-force wall rule
 *   orc_initpop/feqpnt equilibrium moments
 *   orc_forcing_modes  solenoidal/Hermitian structure; OU increment variance,
 *                      kurtosis, autocorrelation and stationary variance
 *   orc_force_field    numpy.fft.irfftn of the same spectrum
 *   orc_ranf           Park-Miller known value; shuffle-slot replay
 *   orc_beads_links    brute-force solid and link sets, wall point on the sphere
 *   orc_beads_collision  Lagrange interpolation on quadratic fields (all five
 *                      branches), momentum exchange and torque sums, rotating-
 *                      sphere torque, moving-sphere force closed form; in
 *                      bb_mode 0 the slots that differ from bb_mode 1 are
 *                      exactly those the link map, the loop order and the
 *                      halo rule predict
 *   orc_beads_lubforce hand-computed magnitudes, symmetry, wall law
 *   orc_beads_move     ramped-force closed form, uniform acceleration
 *   orc_beads_filling  non-equilibrium copy from ipmx on a linear field, tie
 *                      rule, solid ipmx, the rho0 fallback, uniform equilibrium
 *   orc_initpartvel    trilinear exactness; orc_initpartpos no overlap, seeded
 *   orc_timestep       shear-wave decay rate, Poiseuille profile (the
 *                      composition itself has no compiled reference here)
 *
 * Built with -ffp-contract=off: no fused multiply-add, so geometric tests
 * (node maps) are evaluated exactly as written.
 */
#include "mfc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPOP 19

/* lattice (BOX/para.f90:171-175) */
static const int CX[NPOP] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[NPOP] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP[NPOP] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

int orc_nthreads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static inline long nid(const orc_params *P, int i, int j, int k) {
  return (long)i + (long)P->nx * ((long)j + (long)P->ny * (long)k);
}
static inline long ncells(const orc_params *P) {
  return (long)P->nx * (long)P->ny * (long)P->nz;
}
static inline int wrap(int a, int n) {
  if (a < 0) a += n;
  if (a >= n) a -= n;
  return a;
}
static inline double wwgt(const orc_params *P, int ip) {
  return ip == 0 ? P->ww0 : (ip <= 6 ? P->ww1 : P->ww2);
}

/* ranff / ranpp: Park-Miller minimal standard with Bays-Durham shuffle
 * (BOX/gaussian.f90:374-404, BOX/partlib.f90:83-115). */
double orc_ranf(int *idum, int *iy, int *iv) {
  const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
  const int NDIV = 1 + (IM - 1) / NTAB;
  const double AM = 1.0 / IM, EPS = 1.2e-7, RNMX = 1.0 - EPS;
  int j, k;
  if (*idum <= 0 || *iy == 0) {
    *idum = (-*idum > 1) ? -*idum : 1;
    for (j = NTAB + 8; j >= 1; --j) {
      k = *idum / IQ;
      *idum = IA * (*idum - k * IQ) - IR * k;
      if (*idum < 0) *idum += IM;
      if (j <= NTAB) iv[j - 1] = *idum;
    }
    *iy = iv[0];
  }
  k = *idum / IQ;
  *idum = IA * (*idum - k * IQ) - IR * k;
  if (*idum < 0) *idum += IM;
  j = 1 + *iy / NDIV;
  *iy = iv[j - 1];
  iv[j - 1] = *idum;
  {
    double r = AM * (double)(*iy);
    return r < RNMX ? r : RNMX;
  }
}

/* mode-array index: a(i,j,k) with dims (6,5,5), 1-based in the reference */
#define AIX(i, j, k) (((i) - 1) + 6 * (((j) - 1) + 5 * ((k) - 1)))

/* kxr of the (re, im) slot i = 1..6: both slots of a pair carry the pair's kx
 * (kxr(i) = int(i/2) set for odd i and copied to i+1, para.f90:260-263) */
static double kx_of(int i) { return (double)((i - 1) / 2); }
static double kyz_of(int j) { return j <= 3 ? (double)(j - 1) : (double)(j - 6); } /* kyr/kzr */

/* FORCING, spectral part: OU update of the low modes, solenoidal projection,
 * Hermitian symmetry at kx=0 (BOX/gaussian.f90:194-306). */
void orc_forcing_modes(const orc_params *P, orc_state *S) {
  double fr[6 * 5 * 5 * 3];
  const double smallx = 1.0e-18;
  const double var = 447.31 * (P->vscale * P->tscale) * (P->vscale * P->tscale);
  const double tf = 0.038 / P->tscale, dt = 1.0;
  int mc, i, j, k;
  for (mc = 0; mc < 3; ++mc)
    for (k = 1; k <= 5; ++k)
      for (j = 1; j <= 5; ++j)
        for (i = 1; i <= 6; i += 2) {
          double c1 = orc_ranf(&S->iseedf, &S->iyf, S->ivf), rmag, rph;
          if (c1 < 1.0e-6) c1 = 1.0e-6;
          rmag = sqrt(-(4.0 * var * dt * log(c1) / tf));
          rph = orc_ranf(&S->iseedf, &S->iyf, S->ivf);
          fr[AIX(i, j, k) + 150 * mc] = rmag * cos(P->pi2 * rph);
          fr[AIX(i + 1, j, k) + 150 * mc] = rmag * sin(P->pi2 * rph);
        }
  /* real-valued forcing: kx=0 plane Hermitian, zero mean (gaussian.f90:220-236) */
  for (mc = 0; mc < 3; ++mc) {
    double *q = fr + 150 * mc;
    for (k = 1; k <= 5; ++k) {
      int kk = (k == 1) ? 1 : 7 - k;
      for (j = 2; j <= 3; ++j) {
        int jj = 7 - j;
        q[AIX(1, jj, kk)] = q[AIX(1, j, k)];
        q[AIX(2, jj, kk)] = -q[AIX(2, j, k)];
      }
    }
    for (k = 2; k <= 3; ++k) {
      int kk = 7 - k;
      q[AIX(1, 1, kk)] = q[AIX(1, 1, k)];
      q[AIX(2, 1, kk)] = -q[AIX(2, 1, k)];
    }
    q[AIX(1, 1, 1)] = 0.0;
    q[AIX(2, 1, 1)] = 0.0;
  }
  /* Euler step of the Langevin process + incompressible projection (:243-274) */
  for (k = 1; k <= 5; ++k)
    for (j = 1; j <= 5; ++j)
      for (i = 1; i <= 6; ++i) {
        int a = AIX(i, j, k);
        double kx = kx_of(i), ky = kyz_of(j), kz = kyz_of(k);
        double radm = sqrt(kx * kx + ky * ky + kz * kz);
        if (radm < P->xkf) {
          double t1, t2, t3, divb;
          S->b1r[a] = S->b1r[a] * (1.0 - dt / tf) + fr[a];
          S->b2r[a] = S->b2r[a] * (1.0 - dt / tf) + fr[a + 150];
          S->b3r[a] = S->b3r[a] * (1.0 - dt / tf) + fr[a + 300];
          t1 = kx * S->b1r[a];
          t2 = ky * S->b2r[a];
          t3 = kz * S->b3r[a];
          divb = (t1 + t2 + t3) / ((radm + smallx) * (radm + smallx));
          if (radm == 0.0) divb = 0.0;
          S->a1r[a] = S->b1r[a] - kx * divb;
          S->a2r[a] = S->b2r[a] - ky * divb;
          S->a3r[a] = S->b3r[a] - kz * divb;
        } else {
          S->a1r[a] = 0.0;
          S->a2r[a] = 0.0;
          S->a3r[a] = 0.0;
        }
      }
  S->a1r[AIX(1, 1, 1)] = S->a1r[AIX(2, 1, 1)] = 0.0;
  S->a2r[AIX(1, 1, 1)] = S->a2r[AIX(2, 1, 1)] = 0.0;
  S->a3r[AIX(1, 1, 1)] = S->a3r[AIX(2, 1, 1)] = 0.0;
  {
    double *as[3] = {S->a1r, S->a2r, S->a3r};
    for (mc = 0; mc < 3; ++mc) {
      double *q = as[mc];
      for (k = 1; k <= 5; ++k) {
        int kk = (k == 1) ? 1 : 7 - k;
        for (j = 2; j <= 3; ++j) {
          int jj = 7 - j;
          q[AIX(1, jj, kk)] = q[AIX(1, j, k)];
          q[AIX(2, jj, kk)] = -q[AIX(2, j, k)];
        }
      }
      for (k = 2; k <= 3; ++k) {
        int kk = 7 - k;
        q[AIX(1, 1, kk)] = q[AIX(1, 1, k)];
        q[AIX(2, 1, kk)] = -q[AIX(2, 1, k)];
      }
    }
  }
}

/* FORCING, physical part: the unnormalised inverse real 3-D transform of the
 * sparse spectrum (BOX/gaussian.f90:325-369 with the FFTW-2 conventions of
 * BOX/mpifft.f90:111-158: backward = exp(+i), c2r keeps Re of the kx=0 line
 * and doubles kx>0).  Evaluated directly on the few modes:
 *   F(x,y,z) = Re G0(y,z) + 2 sum_{kx=1,2} Re[e^{i 2pi kx x/Nx} G_kx(y,z)],
 *   G_kx(y,z) = sum_{ky,kz} A(kx,ky,kz) e^{i 2pi (ky y/Ny + kz z/Nz)},
 * with node coordinates x = i-1 etc.  The phase factors come from one table
 * per axis, E_N(m) = exp(i 2pi m/N). */
void orc_force_field(const orc_params *P, const double *a1r, const double *a2r,
                     const double *a3r, double *fx, double *fy, double *fz) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const double *as[3] = {a1r, a2r, a3r};
  double *out[3] = {fx, fy, fz};
  double *cxt = (double *)malloc(sizeof(double) * 2 * (nx + ny + nz));
  double *cyt = cxt + 2 * nx, *czt = cyt + 2 * ny;
  double *G = (double *)malloc(sizeof(double) * 18 * (size_t)ny * nz);
  int m;
  for (m = 0; m < nx; ++m) { cxt[2 * m] = cos(P->pi2 * m / nx); cxt[2 * m + 1] = sin(P->pi2 * m / nx); }
  for (m = 0; m < ny; ++m) { cyt[2 * m] = cos(P->pi2 * m / ny); cyt[2 * m + 1] = sin(P->pi2 * m / ny); }
  for (m = 0; m < nz; ++m) { czt[2 * m] = cos(P->pi2 * m / nz); czt[2 * m + 1] = sin(P->pi2 * m / nz); }
  {
    long yz;
#pragma omp parallel for
    for (yz = 0; yz < (long)ny * nz; ++yz) {
      int y = (int)(yz % ny), z = (int)(yz / ny);
      int c, kx, j, k;
      for (c = 0; c < 3; ++c)
        for (kx = 0; kx < 3; ++kx) {
          double gr = 0.0, gi = 0.0;
          for (k = 1; k <= 5; ++k) {
            int mz = wrap((int)(((long)(k <= 3 ? k - 1 : k - 6) * z) % nz), nz);
            double ezr = czt[2 * mz], ezi = czt[2 * mz + 1];
            for (j = 1; j <= 5; ++j) {
              int my = wrap((int)(((long)(j <= 3 ? j - 1 : j - 6) * y) % ny), ny);
              double eyr = cyt[2 * my], eyi = cyt[2 * my + 1];
              double ar = as[c][AIX(2 * kx + 1, j, k)], ai = as[c][AIX(2 * kx + 2, j, k)];
              double er = eyr * ezr - eyi * ezi, ei = eyr * ezi + eyi * ezr;
              gr += ar * er - ai * ei;
              gi += ar * ei + ai * er;
            }
          }
          G[18 * yz + 6 * c + 2 * kx] = gr;
          G[18 * yz + 6 * c + 2 * kx + 1] = gi;
        }
    }
  }
  {
    long n, nt = ncells(P);
#pragma omp parallel for
    for (n = 0; n < nt; ++n) {
      int x = (int)(n % nx);
      long yz = n / nx;
      int c;
      int m1 = x % nx, m2 = (2 * x) % nx;
      for (c = 0; c < 3; ++c) {
        const double *g = G + 18 * yz + 6 * c;
        double v = g[0];
        v += 2.0 * (g[2] * cxt[2 * m1] - g[3] * cxt[2 * m1 + 1]);
        v += 2.0 * (g[4] * cxt[2 * m2] - g[5] * cxt[2 * m2 + 1]);
        out[c][n] = v;
      }
    }
  }
  free(G);
  free(cxt);
}

void orc_forcing(const orc_params *P, orc_state *S) {
  orc_forcing_modes(P, S);
  orc_force_field(P, S->a1r, S->a2r, S->a3r, S->fx, S->fy, S->fz);
}

/* initpop (BOX/initial.f90:3-28): deviation equilibrium with rho = 0 */
void orc_initpop(const orc_params *P, orc_state *S) {
  long n, nt = ncells(P);
  for (n = 0; n < nt; ++n) S->rho[n] = 0.0;
#pragma omp parallel for
  for (n = 0; n < nt; ++n) {
    double u = S->ux[n], v = S->uy[n], w = S->uz[n];
    double usqr = 1.5 * (u * u + v * v + w * w);
    int ip;
    S->f[19 * n] = P->ww0 * (S->rho[n] - usqr);
    for (ip = 1; ip < NPOP; ++ip) {
      double G = CX[ip] * u + CY[ip] * v + CZ[ip] * w;
      S->f[19 * n + ip] = wwgt(P, ip) * (S->rho[n] + 3.0 * G + 4.5 * G * G - usqr);
    }
  }
}

/* collision_MRT (BOX/collision.f90:8-210): forcing merged into the MRT
 * collision, fluid nodes only; conserved moments taken from rho/ux/uy/uz. */
void orc_collision_mrt(const orc_params *P, orc_state *S) {
  long n, nt = ncells(P);
  const double c1 = -2.0 / 3.0, c2 = -11.0, c3 = 8.0, c4 = -4.0, c5 = 2.0;
  const double c3i = 1.0 / c3, c4i = 1.0 / c4;
  const double v1i = 1.0 / 19.0, v2i = 1.0 / 2394.0, v3i = 1.0 / 252.0, v4i = 1.0 / 10.0,
               v5i = 1.0 / 40.0, v6i = 1.0 / 36.0, v7i = 1.0 / 72.0, v8i = 1.0 / 12.0,
               v9i = 1.0 / 24.0, v8 = 12.0;
#pragma omp parallel for
  for (n = 0; n < nt; ++n) {
    double g[NPOP], Fb[NPOP];
    double r, u, v, w, us, vs, ws, Fx, Fy, Fz, G3, t1;
    double q1, q2, q3, q4, q5, q6, q7, q8, q9, q10, q11;
    double m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15;
    double e1, e2, e3, e4, e5, e6, e7, e8, e9, e10, e11, e12, e13, e14, e15;
    double p1, p2, p3, p4, p5, p6, p7, p8, p9, p10, p11, p12, p13, p14, p15, p16,
        p17, p18, p19, p20, p21;
    double A, B, C, D, E, Fq, Gq, H, I, K, Pq, s67, s89, s1011;
    int ip;
    double *f = S->f + 19 * n;
    if (S->ibnodes[n] >= 0) continue;
    r = S->rho[n]; u = S->ux[n]; v = S->uy[n]; w = S->uz[n];
    us = u * u; vs = v * v; ws = w * w;
    Fx = S->fx[n] + P->body_force[0];
    Fy = S->fy[n] + P->body_force[1];
    Fz = S->fz[n] + P->body_force[2];
    G3 = u * Fx + v * Fy + w * Fz;
    Fb[0] = -G3;
    for (ip = 1; ip < NPOP; ++ip) {
      double G1 = CX[ip] * Fx + CY[ip] * Fy + CZ[ip] * Fz;
      double G2 = CX[ip] * u + CY[ip] * v + CZ[ip] * w;
      Fb[ip] = wwgt(P, ip) * (3.0 * G1 + 9.0 * G1 * G2 - 3.0 * G3);
    }
    for (ip = 0; ip < NPOP; ++ip) g[ip] = f[ip] + Fb[ip] / 2.0;

    t1 = us + vs + ws;
    e1 = -11.0 * r + 19.0 * t1;
    e2 = P->omegepsl * r + P->omegepslj * t1;
    e3 = c1 * u; e4 = c1 * v; e5 = c1 * w;
    e6 = 2.0 * us - vs - ws;
    e7 = P->omegxx * e6;
    e8 = vs - ws;
    e9 = P->omegxx * e8;
    e10 = u * v; e11 = v * w; e12 = u * w;
    e13 = 0.0; e14 = 0.0; e15 = 0.0;

    q1 = g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    q2 = g[7] + g[8] + g[9] + g[10] + g[11] + g[12] + g[13] + g[14] + g[15] + g[16] + g[17] + g[18];
    q3 = g[7] - g[8] + g[9] - g[10] + g[11] - g[12] + g[13] - g[14];
    q4 = g[7] + g[8] - g[9] - g[10] + g[15] - g[16] + g[17] - g[18];
    q5 = g[11] + g[12] - g[13] - g[14] + g[15] + g[16] - g[17] - g[18];
    q6 = g[1] + g[2];
    q7 = g[3] + g[4] + g[5] + g[6];
    q8 = g[7] + g[8] + g[9] + g[10] + g[11] + g[12] + g[13] + g[14];
    q9 = g[15] + g[16] + g[17] + g[18];
    q10 = g[3] + g[4] - g[5] - g[6];
    q11 = g[7] + g[8] + g[9] + g[10] - g[11] - g[12] - g[13] - g[14];

    m1 = -30.0 * g[0] + c2 * q1 + c3 * q2;
    m2 = 12.0 * g[0] + c4 * q1 + q2;
    m3 = c4 * (g[1] - g[2]) + q3;
    m4 = c4 * (g[3] - g[4]) + q4;
    m5 = c4 * (g[5] - g[6]) + q5;
    m6 = c5 * q6 - q7 + q8 - c5 * q9;
    m7 = c4 * q6 + c5 * q7 + q8 - c5 * q9;
    m8 = q10 + q11;
    m9 = -c5 * q10 + q11;
    m10 = g[7] - g[8] - g[9] + g[10];
    m11 = g[15] - g[16] - g[17] + g[18];
    m12 = g[11] - g[12] - g[13] + g[14];
    m13 = g[7] - g[8] + g[9] - g[10] - g[11] + g[12] - g[13] + g[14];
    m14 = -g[7] - g[8] + g[9] + g[10] + g[15] - g[16] + g[17] - g[18];
    m15 = g[11] + g[12] - g[13] - g[14] - g[15] - g[16] + g[17] + g[18];

    m1 = m1 - P->s1 * (m1 - e1);
    m2 = m2 - P->s2 * (m2 - e2);
    m3 = m3 - P->s4 * (m3 - e3);
    m4 = m4 - P->s4 * (m4 - e4);
    m5 = m5 - P->s4 * (m5 - e5);
    m6 = m6 - P->s9 * (m6 - e6);
    m7 = m7 - P->s10 * (m7 - e7);
    m8 = m8 - P->s9 * (m8 - e8);
    m9 = m9 - P->s10 * (m9 - e9);
    m10 = m10 - P->s13 * (m10 - e10);
    m11 = m11 - P->s13 * (m11 - e11);
    m12 = m12 - P->s13 * (m12 - e12);
    m13 = m13 - P->s16 * (m13 - e13);
    m14 = m14 - P->s16 * (m14 - e14);
    m15 = m15 - P->s16 * (m15 - e15);

    p1 = v1i * r;
    p2 = c2 * v2i * m1;
    p3 = c3 * v2i * m1;
    p4 = c4 * v3i * m2;
    p5 = v3i * m2;
    p6 = v4i * u;
    p7 = v5i * m3;
    p8 = v4i * v;
    p9 = v5i * m4;
    p10 = v4i * w;
    p11 = v5i * m5;
    p12 = v6i * m6;
    p13 = v7i * m7;
    p14 = v8i * m8;
    p15 = v9i * m9;
    p16 = -c4i * m10;
    p17 = -c4i * m11;
    p18 = -c4i * m12;
    p19 = c3i * m13;
    p20 = c3i * m14;
    p21 = c3i * m15;

    g[0] = p1 - 30.0 * v2i * m1 + v8 * v3i * m2;
    A = p1 + p2 + p4;
    B = p1 + p3 + p5;
    C = p6 + c4 * p7;
    D = c5 * p12 + c4 * p13;
    E = p8 + c4 * p9;
    Fq = -p12 + c5 * p13 + p14 - c5 * p15;
    Gq = p10 + c4 * p11;
    H = -p12 + c5 * p13 - p14 + c5 * p15;
    I = p12 + p13 + p14 + p15;
    K = p12 + p13 - p14 - p15;
    Pq = -c5 * p12 - c5 * p13;
    s67 = p6 + p7;
    s89 = p8 + p9;
    s1011 = p10 + p11;

    g[1] = A + C + D;
    g[2] = A - C + D;
    g[3] = A + E + Fq;
    g[4] = A - E + Fq;
    g[5] = A + Gq + H;
    g[6] = A - Gq + H;
    g[7] = B + s67 + s89 + I + p16 + p19 - p20;
    g[8] = B - s67 + s89 + I - p16 - p19 - p20;
    g[9] = B + s67 - s89 + I - p16 + p19 + p20;
    g[10] = B - s67 - s89 + I + p16 - p19 + p20;
    g[11] = B + s67 + s1011 + K + p18 - p19 + p21;
    g[12] = B - s67 + s1011 + K - p18 + p19 + p21;
    g[13] = B + s67 - s1011 + K - p18 - p19 - p21;
    g[14] = B - s67 - s1011 + K + p18 + p19 - p21;
    g[15] = B + s89 + s1011 + Pq + p17 + p20 - p21;
    g[16] = B - s89 + s1011 + Pq - p17 - p20 - p21;
    g[17] = B + s89 - s1011 + Pq - p17 + p20 + p21;
    g[18] = B - s89 - s1011 + Pq + p17 - p20 + p21;

    for (ip = 0; ip < NPOP; ++ip) f[ip] = g[ip] + 0.5 * Fb[ip];
  }
}

/* streaming (BOX/streaming.f90:2-171): periodic pull f_i(x) <- f_i(x - c_i).
 * With walls on the y faces, a population whose source lies beyond a wall
 * takes the pre-streaming opposite population of the same node (half-way
 * bounce-back as CHAN/streaming.f90:175-187 applies it at its x walls). */
void orc_streaming(const orc_params *P, orc_state *S) {
  long nt = ncells(P), n;
  double *tmp = (double *)malloc(sizeof(double) * 19 * (size_t)nt);
  memcpy(tmp, S->f, sizeof(double) * 19 * (size_t)nt);
#pragma omp parallel for
  for (n = 0; n < nt; ++n) {
    int i = (int)(n % P->nx), j = (int)((n / P->nx) % P->ny), k = (int)(n / ((long)P->nx * P->ny));
    int ip;
    for (ip = 1; ip < NPOP; ++ip) {
      int js = j - CY[ip];
      if (P->walls_y && (js < 0 || js >= P->ny)) {
        S->f[19 * n + ip] = tmp[19 * n + OPP[ip]];
      } else {
        long s = nid(P, wrap(i - CX[ip], P->nx), wrap(js, P->ny), wrap(k - CZ[ip], P->nz));
        S->f[19 * n + ip] = tmp[19 * s + ip];
      }
    }
  }
  free(tmp);
}

/* rhoupdat (BOX/collision.f90:213-229) */
void orc_rhoupdat(const orc_params *P, orc_state *S) {
  long n, nt = ncells(P);
#pragma omp parallel for
  for (n = 0; n < nt; ++n) {
    double r = S->f[19 * n];
    int ip;
    for (ip = 1; ip < NPOP; ++ip) r = r + S->f[19 * n + ip];
    S->rho[n] = r;
  }
}

/* nearest periodic image of a centre coordinate relative to a node coordinate
 * (the two conditional shifts written out in BOX/collision.f90:295-305) */
static inline double near_img(double c, double pnt, int n) {
  double h = (double)(n / 2);
  if ((c - pnt) > h) c = c - (double)n;
  if ((c - pnt) < -h) c = c + (double)n;
  return c;
}

/* y image of a centre: none across walls */
static inline double y_img(const orc_params *P, double c, double pnt) {
  return P->walls_y ? c : near_img(c, pnt, P->ny);
}

/* macrovar (BOX/collision.f90:232-332) */
void orc_macrovar(const orc_params *P, orc_state *S) {
  long n, nt = ncells(P);
#pragma omp parallel for
  for (n = 0; n < nt; ++n) {
    const double *g = S->f + 19 * n;
    double a1 = g[7] - g[10], a2 = g[9] - g[8], a3 = g[11] - g[14], a4 = g[13] - g[12];
    double a5 = g[15] - g[18], a6 = g[17] - g[16];
    double u = g[1] - g[2] + a1 + a2 + a3 + a4;
    double v = g[3] - g[4] + a1 - a2 + a5 + a6;
    double w = g[5] - g[6] + a3 - a4 + a5 - a6;
    double r = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8] + g[9] + g[10] +
               g[11] + g[12] + g[13] + g[14] + g[15] + g[16] + g[17] + g[18];
    S->ux[n] = u + (S->fx[n] + P->body_force[0]) / 2.0;
    S->uy[n] = v + (S->fy[n] + P->body_force[1]) / 2.0;
    S->uz[n] = w + (S->fz[n] + P->body_force[2]) / 2.0;
    S->rho[n] = r;
  }
  if (!P->ipart) return;
  {
    int i, j, k;
    for (k = 0; k < P->nz; ++k)
      for (j = 0; j < P->ny; ++j)
        for (i = 0; i < P->nx; ++i) {
          long m = nid(P, i, j, k);
          int id;
          double xp, yp, zp, xc, yc, zc, xx, yy, zz;
          const double *W, *O;
          if (S->ibnodes[m] <= 0) continue;
          id = S->isnodes[m] - 1;
          xp = (double)(i + 1) - 0.5;
          yp = (double)(j + 1) - 0.5;
          zp = (double)(k + 1) - 0.5;
          xc = near_img(S->ypglb[3 * id], xp, P->nx);
          yc = y_img(P, S->ypglb[3 * id + 1], yp);
          zc = near_img(S->ypglb[3 * id + 2], zp, P->nz);
          xx = xp - xc; yy = yp - yc; zz = zp - zc;
          W = S->wp + 3 * id; O = S->omgp + 3 * id;
          S->ux[m] = W[0] + (O[1] * zz - O[2] * yy);
          S->uy[m] = W[1] + (O[2] * xx - O[0] * zz);
          S->uz[m] = W[2] + (O[0] * yy - O[1] * xx);
          S->rho[m] = P->rhopart;
        }
  }
}

/* ovlpchk (BOX/partlib.f90:118-171), one rank: local order == id order.
 * Flags the lower id of every pair closer than 2r (+ mingap at release).
 * Returns the number of flagged particles. */
int orc_ovlpchk(const orc_params *P, orc_state *S, int *flags) {
  int a, b, cnt = 0;
  for (a = 0; a < P->npart; ++a) flags[a] = 0;
  for (a = 0; a < P->npart - 1; ++a) {
    double xc = S->yp[3 * a], yc = S->yp[3 * a + 1], zc = S->yp[3 * a + 2];
    for (b = a + 1; b < P->npart; ++b) {
      double xb = near_img(S->ypglb[3 * b], xc, P->nx);
      double yb = y_img(P, S->ypglb[3 * b + 1], yc);
      double zb = near_img(S->ypglb[3 * b + 2], zc, P->nz);
      double dist = sqrt((xc - xb) * (xc - xb) + (yc - yb) * (yc - yb) + (zc - zb) * (zc - zb));
      double hgap = dist - 2.0 * P->rad;
      if (S->istep00 == 0) hgap = hgap - P->mingap;
      if (hgap < 0.0) { flags[a] = 1; break; }
    }
  }
  for (a = 0; a < P->npart; ++a) cnt += flags[a] != 0;
  return cnt;
}

/* initpartpos (BOX/partlib.f90:1-80): seeded rejection sampling; only the
 * flagged particles are redrawn.  Returns the number of draw rounds. */
int orc_initpartpos(const orc_params *P, orc_state *S) {
  int *flag = (int *)malloc(sizeof(int) * (size_t)P->npart);
  int p, rounds = 0;
  S->iseedp = -S->iseedp;
  S->iyp = 0;
  for (p = 0; p < 32; ++p) S->ivp[p] = 0;
  for (p = 0; p < P->npart; ++p) flag[p] = 1;
  for (;;) {
    ++rounds;
    for (p = 0; p < P->npart; ++p)
      if (flag[p] > 0) {
        S->ypglb[3 * p] = (double)P->nx * orc_ranf(&S->iseedp, &S->iyp, S->ivp);
        if (P->walls_y) {
          /* clear of the walls by rad + mingap_w (channel initpartpos, wall axis y) */
          const double rg = P->rad + P->mingap_w;
          S->ypglb[3 * p + 1] = rg + ((double)P->ny - 2.0 * rg) * orc_ranf(&S->iseedp, &S->iyp, S->ivp);
        } else {
          S->ypglb[3 * p + 1] = (double)P->ny * orc_ranf(&S->iseedp, &S->iyp, S->ivp);
        }
        S->ypglb[3 * p + 2] = (double)P->nz * orc_ranf(&S->iseedp, &S->iyp, S->ivp);
      }
    for (p = 0; p < 3 * P->npart; ++p) S->yp[p] = S->ypglb[p];
    if (orc_ovlpchk(P, S, flag) == 0) break;
    if (rounds > 100000) break;
  }
  free(flag);
  return rounds;
}

/* trilinear (BOX/partlib.f90:320-831) on one rank: the halo cases reduce to
 * periodic indexing; node (i) sits at x = i - 0.5. */
static double trilin(const orc_params *P, const double *u, double xc, double yc, double zc) {
  int i = (int)(xc + 0.5), j = (int)(yc + 0.5), k = (int)(zc + 0.5);
  double tx = xc + 0.5 - (double)i, ty = yc + 0.5 - (double)j, tz = zc + 0.5 - (double)k;
  double tx0 = 1.0 - tx, ty0 = 1.0 - ty, tz0 = 1.0 - tz;
  int i0 = wrap(i - 1, P->nx), i1 = wrap(i, P->nx);
  int j0 = wrap(j - 1, P->ny), j1 = wrap(j, P->ny);
  int k0 = wrap(k - 1, P->nz), k1 = wrap(k, P->nz);
  double v1 = u[nid(P, i0, j0, k0)], v2 = u[nid(P, i1, j0, k0)], v3 = u[nid(P, i1, j1, k0)],
         v4 = u[nid(P, i0, j1, k0)], v5 = u[nid(P, i0, j0, k1)], v6 = u[nid(P, i1, j0, k1)],
         v7 = u[nid(P, i1, j1, k1)], v8 = u[nid(P, i0, j1, k1)];
  return tx0 * ty0 * tz0 * v1 + tx * ty0 * tz0 * v2 + tx * ty * tz0 * v3 + tx0 * ty * tz0 * v4 +
         tx0 * ty0 * tz * v5 + tx * ty0 * tz * v6 + tx * ty * tz * v7 + tx0 * ty * tz * v8;
}

/* initpartvel (BOX/partlib.f90:177-263) */
void orc_initpartvel(const orc_params *P, orc_state *S) {
  int p;
  for (p = 0; p < P->npart; ++p) {
    const double *y = S->ypglb + 3 * p;
    S->wp[3 * p] = trilin(P, S->ux, y[0], y[1], y[2]);
    S->wp[3 * p + 1] = trilin(P, S->uy, y[0], y[1], y[2]);
    S->wp[3 * p + 2] = trilin(P, S->uz, y[0], y[1], y[2]);
  }
}

/* initpartomg (BOX/partlib.f90:837-931): the interpolated spin is discarded
 * and omgp is set to zero (:929). */
void orc_initpartomg(const orc_params *P, orc_state *S) {
  int p;
  for (p = 0; p < 3 * P->npart; ++p) S->omgp[p] = 0.0;
}

/* beads_links (BOX/partlib.f90:934-1058).  The loop structure, including the
 * centre coordinates that persist (and are re-imaged) across loop
 * iterations, is reproduced as written. */
void orc_beads_links(const orc_params *P, orc_state *S) {
  const long nt = ncells(P);
  const double radp2 = P->rad * 1.0 + 2.0, r2 = (P->rad * 1.0) * (P->rad * 1.0);
  const double hx = (double)(P->nx / 2), hy = (double)(P->ny / 2), hz = (double)(P->nz / 2);
  long n;
  int p, i, j, k, ip;
  for (n = 0; n < nt; ++n) {
    S->ibnodes[n] = -1;
    S->isnodes[n] = -1;
  }
  for (n = 0; n < 18 * nt; ++n) {
    S->ilink[n] = -1;
    S->mlink[n] = -1;
    S->alink[n] = 0.0;
  }
  S->alpha_faults = 0;
  for (p = 0; p < P->npart; ++p) {
    double xc = S->ypglb[3 * p], yc = S->ypglb[3 * p + 1], zc = S->ypglb[3 * p + 2];
    for (k = 1; k <= P->nz; ++k) {
      double zp = (double)k - 0.5, zz0;
      if ((zc - zp) > hz) zc = zc - (double)P->nz;
      if ((zc - zp) < -hz) zc = zc + (double)P->nz;
      zz0 = zp - zc;
      if (!(fabs(zz0) <= radp2)) continue;
      for (j = 1; j <= P->ny; ++j) {
        double yp = (double)j - 0.5, yy0;
        if (!P->walls_y) {
          if ((yc - yp) > hy) yc = yc - (double)P->ny;
          if ((yc - yp) < -hy) yc = yc + (double)P->ny;
        }
        yy0 = yp - yc;
        if (!(fabs(yy0) <= radp2)) continue;
        for (i = 1; i <= P->nx; ++i) {
          double xp = (double)i - 0.5, xx0, rr0, rr01;
          long m;
          if ((xc - xp) > hx) xc = xc - (double)P->nx;
          if ((xc - xp) < -hx) xc = xc + (double)P->nx;
          xx0 = xp - xc;
          if (!(fabs(xx0) <= radp2)) continue;
          m = nid(P, i - 1, j - 1, k - 1);
          rr0 = xx0 * xx0 + yy0 * yy0 + zz0 * zz0;
          rr01 = rr0 - r2;
          if (rr01 <= 0.0) {
            S->ibnodes[m] = 1;
            S->isnodes[m] = p + 1;
          } else {
            for (ip = 1; ip < NPOP; ++ip) {
              int im = i + CX[ip], jm = j + CY[ip], km = k + CZ[ip];
              double xx1 = (double)im - 0.5 - xc;
              double yy1 = (double)jm - 0.5 - yc + 0.0;
              double zz1 = (double)km - 0.5 - zc + 0.0;
              double rr1 = xx1 * xx1 + yy1 * yy1 + zz1 * zz1;
              double rr11 = rr1 - r2;
              if (rr11 <= 0.0) {
                double aa = rr0 + rr1 - 2.0 * (xx0 * xx1 + yy0 * yy1 + zz0 * zz1);
                double bb = xx1 * (xx0 - xx1) + yy1 * (yy0 - yy1) + zz1 * (zz0 - zz1);
                double cc = rr11, a0, a1, alpha;
                a0 = bb / aa;
                a1 = sqrt(a0 * a0 - cc / aa);
                alpha = -a0 + a1;
                alpha = 1.0 - alpha;
                S->ilink[(ip - 1) + 18 * m] = 1;
                S->mlink[(ip - 1) + 18 * m] = p + 1;
                S->alink[(ip - 1) + 18 * m] = alpha;
                if (alpha < 0.0 || alpha > 1.0) S->alpha_faults++;
              }
            }
          }
        }
      }
    }
  }
}

/* ---- beads_collision (BOX/partlib.f90:1064-1717) ------------------------ */

/* Single-rank view of the reference's population storage during the
 * bounce-back loop: the interior array f plus four halo copies filled before
 * the loop (exchng2, partlib.f90:1229) -- U: j in {ly+1,ly+2}, k in
 * [-1,lz+2]; D: j in {-1,0}, k in [-1,lz+2]; R: j in [1,ly], k in
 * {lz+1,lz+2}; L: j in [1,ly], k in {-1,0}.  Reads and writes are routed
 * exactly as partlib.f90:1442-1621 route them; halo writes are applied to f
 * after the loop (exchng3, partlib.f90:2101-2171). */
typedef struct {
  const orc_params *P;
  double *f;
  double *U, *D, *R, *L;     /* 19 * nx * (2 * (nz+4)) or 19 * nx * ny * 2 */
  unsigned char *fU, *fD, *fR, *fL;
} bbview;

static double *bb_slot(bbview *V, int ip, int i, int j, int k, unsigned char **flag) {
  const orc_params *P = V->P;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  long idx;
  /* i is 1-based and already wrapped; j, k 1-based, possibly outside */
  if (j > ny) {
    idx = ip + 19L * ((i - 1) + (long)nx * ((j - ny - 1) + 2L * (k + 1)));
    if (flag) *flag = V->fU + idx;
    return V->U + idx;
  } else if (j < 1) {
    idx = ip + 19L * ((i - 1) + (long)nx * ((j + 1) + 2L * (k + 1)));
    if (flag) *flag = V->fD + idx;
    return V->D + idx;
  } else if (k > nz) {
    idx = ip + 19L * ((i - 1) + (long)nx * ((j - 1) + (long)ny * (k - nz - 1)));
    if (flag) *flag = V->fR + idx;
    return V->R + idx;
  } else if (k < 1) {
    idx = ip + 19L * ((i - 1) + (long)nx * ((j - 1) + (long)ny * (k + 1)));
    if (flag) *flag = V->fL + idx;
    return V->L + idx;
  }
  if (flag) *flag = NULL;
  return V->f + ip + 19L * nid(P, i - 1, j - 1, k - 1);
}

static void bb_fill_halos(bbview *V) {
  const orc_params *P = V->P;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  int i, j, k, ip;
  for (k = -1; k <= nz + 2; ++k)
    for (j = 0; j < 2; ++j)
      for (i = 1; i <= nx; ++i)
        for (ip = 0; ip < NPOP; ++ip) {
          long o = ip + 19L * ((i - 1) + (long)nx * (j + 2L * (k + 1)));
          int kw = wrap(k - 1, nz);
          V->U[o] = V->f[ip + 19L * nid(P, i - 1, wrap(ny + j, ny), kw)];
          V->D[o] = V->f[ip + 19L * nid(P, i - 1, wrap(j - 2, ny), kw)];
        }
  for (k = 0; k < 2; ++k)
    for (j = 1; j <= ny; ++j)
      for (i = 1; i <= nx; ++i)
        for (ip = 0; ip < NPOP; ++ip) {
          long o = ip + 19L * ((i - 1) + (long)nx * ((j - 1) + (long)ny * k));
          V->R[o] = V->f[ip + 19L * nid(P, i - 1, j - 1, wrap(nz + k, nz))];
          V->L[o] = V->f[ip + 19L * nid(P, i - 1, j - 1, wrap(k - 2, nz))];
        }
}

static void bb_apply_halos(bbview *V) {
  const orc_params *P = V->P;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  int i, j, k, ip;
  for (k = -1; k <= nz + 2; ++k)
    for (j = 0; j < 2; ++j)
      for (i = 1; i <= nx; ++i)
        for (ip = 0; ip < NPOP; ++ip) {
          long o = ip + 19L * ((i - 1) + (long)nx * (j + 2L * (k + 1)));
          int kw = wrap(k - 1, nz);
          if (V->fU[o]) V->f[ip + 19L * nid(P, i - 1, wrap(ny + j, ny), kw)] = V->U[o];
          if (V->fD[o]) V->f[ip + 19L * nid(P, i - 1, wrap(j - 2, ny), kw)] = V->D[o];
        }
  for (k = 0; k < 2; ++k)
    for (j = 1; j <= ny; ++j)
      for (i = 1; i <= nx; ++i)
        for (ip = 0; ip < NPOP; ++ip) {
          long o = ip + 19L * ((i - 1) + (long)nx * ((j - 1) + (long)ny * k));
          if (V->fR[o]) V->f[ip + 19L * nid(P, i - 1, j - 1, wrap(nz + k, nz))] = V->R[o];
          if (V->fL[o]) V->f[ip + 19L * nid(P, i - 1, j - 1, wrap(k - 2, nz))] = V->L[o];
        }
}

void orc_beads_collision(const orc_params *P, orc_state *S) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const long nt = ncells(P);
  const long nyz = 19L * nx * 2 * (nz + 4), nyk = 19L * nx * ny * 2;
  double *fsnap = NULL;
  bbview V;
  int i, j, k, ip, p;
  V.P = P;
  V.f = S->f;
  V.U = V.D = V.R = V.L = NULL;
  V.fU = V.fD = V.fR = V.fL = NULL;
  if (P->bb_mode == 0) {
    V.U = (double *)malloc(sizeof(double) * nyz);
    V.D = (double *)malloc(sizeof(double) * nyz);
    V.R = (double *)malloc(sizeof(double) * nyk);
    V.L = (double *)malloc(sizeof(double) * nyk);
    V.fU = (unsigned char *)calloc(nyz, 1);
    V.fD = (unsigned char *)calloc(nyz, 1);
    V.fR = (unsigned char *)calloc(nyk, 1);
    V.fL = (unsigned char *)calloc(nyk, 1);
    bb_fill_halos(&V);
  } else {
    fsnap = (double *)malloc(sizeof(double) * 19 * (size_t)nt);
    memcpy(fsnap, S->f, sizeof(double) * 19 * (size_t)nt);
  }
  for (p = 0; p < 3 * P->npart; ++p) {
    S->fHIp[p] = 0.0;
    S->torqp[p] = 0.0;
  }
  S->nconflict = 0;
  for (k = 1; k <= nz; ++k)
    for (j = 1; j <= ny; ++j)
      for (i = 1; i <= nx; ++i)
        for (ip = 1; ip < NPOP; ++ip) {
          long m = nid(P, i - 1, j - 1, k - 1);
          int id, ipp, ix, iy, iz, im1, im2, ip1, jm1, jm2, jp1, km1, km2, kp1, ibm1, ibm2;
          double alpha, xc, yc, zc, xp, yp, zp, xx0, yy0, zz0, uwx, uwy, uwz, uwpro;
          double ff1, ff2, ff3, f9, c1, c2, c3, dff, dxm, dym, dzm, wt;
          const double *W, *O;
          if (S->ilink[(ip - 1) + 18 * m] <= 0) continue;
          if (S->ibnodes[m] > 0) S->nconflict++;
          id = S->mlink[(ip - 1) + 18 * m] - 1;
          alpha = S->alink[(ip - 1) + 18 * m];
          ipp = OPP[ip];
          xp = (double)i - 0.5;
          yp = (double)j - 0.5;
          zp = (double)k - 0.5;
          xc = near_img(S->ypglb[3 * id], xp, nx);
          yc = y_img(P, S->ypglb[3 * id + 1], yp);
          zc = near_img(S->ypglb[3 * id + 2], zp, nz);
          W = S->wp + 3 * id;
          O = S->omgp + 3 * id;
          ix = CX[ip]; iy = CY[ip]; iz = CZ[ip];
          im1 = i - ix; jm1 = j - iy; km1 = k - iz;
          im2 = i - 2 * ix; jm2 = j - 2 * iy; km2 = k - 2 * iz;
          ip1 = i + ix; jp1 = j + iy; kp1 = k + iz;
          if (im1 < 1) im1 += nx;
          if (im1 > nx) im1 -= nx;
          if (im2 < 1) im2 += nx;
          if (im2 > nx) im2 -= nx;
          if (ip1 < 1) ip1 += nx;
          if (ip1 > nx) ip1 -= nx;
          xx0 = xp + (double)ix * alpha - xc;
          yy0 = yp + (double)iy * alpha - yc;
          zz0 = zp + (double)iz * alpha - zc;
          wt = ip <= 6 ? P->ww1 : P->ww2;
          uwx = W[0] + O[1] * zz0 - O[2] * yy0;
          uwy = W[1] + O[2] * xx0 - O[0] * zz0;
          uwz = W[2] + O[0] * yy0 - O[1] * xx0;
          uwpro = uwx * (double)ix + uwy * (double)iy + uwz * (double)iz;
          if (P->bb_mode == 0) {
            ff1 = *bb_slot(&V, ip, i, j, k, NULL);
          } else {
            ff1 = fsnap[ip + 19 * m];
          }
#define RD(q, a, b, c) (P->bb_mode == 0 ? *bb_slot(&V, (q), (a), (b), (c), NULL) \
      : fsnap[(q) + 19L * nid(P, (a) - 1, wrap((b) - 1, ny), wrap((c) - 1, nz))])
#define IB(a, b, c) (S->ibnodes[nid(P, (a) - 1, wrap((b) - 1, ny), wrap((c) - 1, nz))])
          if (alpha <= 0.5) {
            ff2 = RD(ip, im1, jm1, km1);
            ibm1 = (P->walls_y && (jm1 < 1 || jm1 > ny)) ? 2 : IB(im1, jm1, km1);
            ff3 = RD(ip, im2, jm2, km2);
            ibm2 = (P->walls_y && (jm2 < 1 || jm2 > ny)) ? 2 : IB(im2, jm2, km2);
            if (ibm1 > 0) {
              f9 = ff1 - 6.0 * wt * uwpro;
            } else if (ibm2 > 0) {
              f9 = 2.0 * alpha * (ff1 - ff2) + ff2 - 6.0 * wt * uwpro;
            } else {
              c1 = alpha * (1.0 + 2.0 * alpha);
              c2 = 1.0 - 4.0 * alpha * alpha;
              c3 = -alpha * (1.0 - 2.0 * alpha);
              f9 = c1 * ff1 + c2 * ff2 + c3 * ff3 - 6.0 * wt * uwpro;
            }
          } else {
            ff2 = (P->bb_mode == 0) ? *bb_slot(&V, ipp, i, j, k, NULL) : fsnap[ipp + 19 * m];
            ff3 = RD(ipp, im1, jm1, km1);
            ibm1 = (P->walls_y && (jm1 < 1 || jm1 > ny)) ? 2 : IB(im1, jm1, km1);
            if (ibm1 > 0) {
              c1 = 0.5 / alpha;
              c2 = 1.0 - c1;
              f9 = c1 * ff1 + c2 * ff2 - 6.0 * wt * c1 * uwpro;
            } else {
              c1 = 1.0 / alpha / (2.0 * alpha + 1.0);
              c2 = (2.0 * alpha - 1.0) / alpha;
              c3 = 1.0 - c1 - c2;
              f9 = c1 * ff1 + c2 * ff2 + c3 * ff3 - 6.0 * wt * c1 * uwpro;
            }
          }
#undef RD
#undef IB
          if (P->walls_y && (jp1 < 1 || jp1 > ny)) {
            /* the solid neighbour lies beyond a wall: nothing to write */
          } else if (P->bb_mode == 0) {
            unsigned char *fl = NULL;
            double *dst = bb_slot(&V, ipp, ip1, jp1, kp1, &fl);
            *dst = f9;
            if (fl) *fl = 1;
          } else {
            S->f[ipp + 19L * nid(P, ip1 - 1, wrap(jp1 - 1, ny), wrap(kp1 - 1, nz))] = f9;
          }
          dff = ff1 + f9;
          dxm = dff * (double)ix;
          dym = dff * (double)iy;
          dzm = dff * (double)iz;
          S->fHIp[3 * id] += dxm;
          S->fHIp[3 * id + 1] += dym;
          S->fHIp[3 * id + 2] += dzm;
          S->torqp[3 * id] += dzm * yy0 - dym * zz0;
          S->torqp[3 * id + 1] += dxm * zz0 - dzm * xx0;
          S->torqp[3 * id + 2] += dym * xx0 - dxm * yy0;
        }
  if (P->bb_mode == 0) {
    bb_apply_halos(&V);
    free(V.U); free(V.D); free(V.R); free(V.L);
    free(V.fU); free(V.fD); free(V.fR); free(V.fL);
  } else {
    free(fsnap);
  }
}

/* beads_lubforce (BOX/partlib.f90:2211-2292), one rank, id order */
void orc_beads_lubforce(const orc_params *P, orc_state *S) {
  const double dist0 = 2.0 * P->rad + P->mingap;
  int a, b;
  for (a = 0; a < 3 * P->npart; ++a) S->flubp[a] = 0.0;
  for (a = 0; a < P->npart - 1; ++a) {
    double xc = S->yp[3 * a], yc = S->yp[3 * a + 1], zc = S->yp[3 * a + 2];
    for (b = a + 1; b < P->npart; ++b) {
      double xb = near_img(S->ypglb[3 * b], xc, P->nx);
      double yb = y_img(P, S->ypglb[3 * b + 1], yc);
      double zb = near_img(S->ypglb[3 * b + 2], zc, P->nz);
      double dx = xc - xb, dy = yc - yb, dz = zc - zb;
      double dist = sqrt(dx * dx + dy * dy + dz * dz);
      double hgap = dist - 2.0 * P->rad, fp = 0.0, fpd;
      if (dist < dist0) {
        double fpc = P->volp * P->rho0 * 9.8 / P->gscale, d1 = 0.0, t;
        if (hgap < 0.0) d1 = -hgap;
        t = (dist0 - dist) / P->mingap;
        fp = (t * t / P->stf0 + d1 / P->mingap / P->stf1) * fpc;
      }
      fpd = fp / dist;
      S->flubp[3 * a] += fpd * dx;
      S->flubp[3 * a + 1] += fpd * dy;
      S->flubp[3 * a + 2] += fpd * dz;
      S->flubp[3 * b] -= fpd * dx;
      S->flubp[3 * b + 1] -= fpd * dy;
      S->flubp[3 * b + 2] -= fpd * dz;
    }
  }
  if (P->walls_y) {
    /* sphere-wall lubrication (channel beads_lubforce, wall axis y): the
       bottom wall at y = 0 pushes up, the top wall at y = ny down */
    const double dist0 = P->mingap_w + P->rad;
    const double fpc = P->volp * P->rho0 * 9.8 / P->gscale;
    for (a = 0; a < P->npart; ++a) {
      const double yd = S->yp[3 * a + 1];
      double dist = yd, hgap = dist - P->rad, d1, t;
      if (dist < dist0) {
        d1 = hgap < 0.0 ? -hgap : 0.0;
        t = (dist0 - dist) / P->mingap_w;
        S->flubp[3 * a + 1] += (t * t / P->stf0_w + d1 / P->mingap_w / P->stf1_w) * fpc;
      }
      dist = (double)P->ny - yd;
      hgap = dist - P->rad;
      if (dist < dist0) {
        d1 = hgap < 0.0 ? -hgap : 0.0;
        t = (dist0 - dist) / P->mingap_w;
        S->flubp[3 * a + 1] -= (t * t / P->stf0_w + d1 / P->mingap_w / P->stf1_w) * fpc;
      }
    }
  }
}

/* beads_move (BOX/partlib.f90:2295-2394), one rank.  At the release step the
 * previous force, torque and accelerations equal the current ones. */
void orc_beads_move(const orc_params *P, orc_state *S) {
  const double dt = 1.0;
  const int first = (S->istep == P->irelease);
  int p, c;
  long n, nt = ncells(P);
  if (first)
    for (p = 0; p < 3 * P->npart; ++p) S->thetap[p] = 0.0;
  for (p = 0; p < 3 * P->npart; ++p) {
    S->wpp[p] = S->wp[p];
    S->omgpp[p] = S->omgp[p];
  }
  for (p = 0; p < P->npart; ++p) {
    double dwn[3], don[3];
    for (c = 0; c < 3; ++c) S->forcep[3 * p + c] = S->fHIp[3 * p + c] + S->flubp[3 * p + c];
    S->forcep[3 * p + 1] = S->forcep[3 * p + 1] - P->volp * P->rhog;
    /* uniform forces: the fluid's body force acts on the displaced volume, and
       the spheres' own force (e.g. buoyancy-corrected weight) */
    for (c = 0; c < 3; ++c) {
      if (P->body_force[c] != 0.0) S->forcep[3 * p + c] = S->forcep[3 * p + c] + P->volp * P->body_force[c];
      if (P->particle_force[c] != 0.0)
        S->forcep[3 * p + c] = S->forcep[3 * p + c] + P->volp * P->particle_force[c];
    }
    for (c = 0; c < 3; ++c) {
      int q = 3 * p + c;
      if (first) S->forcepp[q] = S->forcep[q];
      dwn[c] = 0.5 * (S->forcep[q] + S->forcepp[q]) / P->amp;
      if (first) S->dwdt[q] = dwn[c];
      S->wp[q] = S->wpp[q] + 0.5 * dt * (S->dwdt[q] + dwn[c]);
    }
    for (c = 0; c < 3; ++c) {
      int q = 3 * p + c;
      S->yp[q] = S->yp[q] + 0.5 * dt * (S->wp[q] + S->wpp[q]);
    }
    for (c = 0; c < 3; ++c) {
      int q = 3 * p + c;
      if (first) S->torqpp[q] = S->torqp[q];
      don[c] = 0.5 * (S->torqp[q] + S->torqpp[q]) / P->aip;
      if (first) S->domgdt[q] = don[c];
      S->omgp[q] = S->omgpp[q] + 0.5 * dt * (S->domgdt[q] + don[c]);
    }
    for (c = 0; c < 3; ++c) {
      int q = 3 * p + c;
      S->thetap[q] = S->thetap[q] + 0.5 * dt * (S->omgp[q] + S->omgpp[q]);
      S->dwdt[q] = dwn[c];
      S->domgdt[q] = don[c];
      S->forcepp[q] = S->forcep[q];
      S->torqpp[q] = S->torqp[q];
    }
  }
  for (n = 0; n < nt; ++n) {
    S->ibnodes0[n] = S->ibnodes[n];
    S->isnodes0[n] = S->isnodes[n];
  }
}

/* beads_redistribute, single-rank content (BOX/partlib.f90:2600-2646):
 * periodic wrap of yp, ypglbp <- ypglb, ypglb <- yp, overlap check.
 * Returns the overlap count (nonzero is fatal in the reference). */
int orc_beads_redistribute(const orc_params *P, orc_state *S) {
  int p, cnt;
  int *flags = (int *)malloc(sizeof(int) * (size_t)P->npart);
  for (p = 0; p < P->npart; ++p) {
    double *y = S->yp + 3 * p;
    double x0 = y[0], y0 = y[1], z0 = y[2];
    if (x0 >= (double)P->nx) y[0] = y[0] - (double)P->nx;
    if (x0 < 0.0) y[0] = y[0] + (double)P->nx;
    if (!P->walls_y) {
      if (y0 >= (double)P->ny) y[1] = y[1] - (double)P->ny;
      if (y0 < 0.0) y[1] = y[1] + (double)P->ny;
    }
    if (z0 >= (double)P->nz) y[2] = y[2] - (double)P->nz;
    if (z0 < 0.0) y[2] = y[2] + (double)P->nz;
  }
  for (p = 0; p < 3 * P->npart; ++p) {
    S->ypglbp[p] = S->ypglb[p];
    S->ypglb[p] = S->yp[p];
  }
  cnt = orc_ovlpchk(P, S, flags);
  S->novlp = cnt;
  free(flags);
  return cnt;
}

/* feqpnt (BOX/partlib.f90:3222-3247) */
void orc_feqpnt(const orc_params *P, double u, double v, double w, double rho, double *feq) {
  double usqr = 1.5 * (u * u + v * v + w * w);
  int ip;
  feq[0] = P->ww0 * (rho - usqr);
  for (ip = 1; ip < NPOP; ++ip) {
    double G = CX[ip] * u + CY[ip] * v + CZ[ip] * w;
    feq[ip] = wwgt(P, ip) * (rho + 3.0 * G + 4.5 * G * G - usqr);
  }
}

/* beads_filling (BOX/partlib.f90:2793-3111): refill nodes that were solid in
 * the old map and are fluid in the new one. */
void orc_beads_filling(const orc_params *P, orc_state *S) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  int i, j, k, ip;
  for (k = 0; k < nz; ++k)
    for (j = 0; j < ny; ++j)
      for (i = 0; i < nx; ++i) {
        long m = nid(P, i, j, k);
        int id, ipmx = 1, nghb;
        double xp, yp, zp, xc, yc, zc, w1, w2, w3, o1, o2, o3, aa, bb, cc, d0, d1, ddt;
        double x1, y1, z1, x2, y2, z2, xx0, yy0, zz0, prod0, rho9, u9, v9, w9;
        double feq[NPOP];
        const double *g;
        if (!(S->ibnodes0[m] > 0 && S->ibnodes[m] < 0)) continue;
        id = S->isnodes0[m] - 1;
        xp = (double)(i + 1) - 0.5;
        yp = (double)(j + 1) - 0.5;
        zp = (double)(k + 1) - 0.5;
        xc = near_img(S->ypglbp[3 * id] * 1.0, xp, nx);
        yc = y_img(P, S->ypglbp[3 * id + 1] * 1.0, yp);
        zc = near_img(S->ypglbp[3 * id + 2] * 1.0, zp, nz);
        w1 = -0.5 * (S->wp[3 * id] * 1.0 + S->wpp[3 * id] * 1.0);
        w2 = -0.5 * (S->wp[3 * id + 1] * 1.0 + S->wpp[3 * id + 1] * 1.0);
        w3 = -0.5 * (S->wp[3 * id + 2] * 1.0 + S->wpp[3 * id + 2] * 1.0);
        o1 = -0.5 * (S->omgp[3 * id] * 1.0 + S->omgpp[3 * id] * 1.0);
        o2 = -0.5 * (S->omgp[3 * id + 1] * 1.0 + S->omgpp[3 * id + 1] * 1.0);
        o3 = -0.5 * (S->omgp[3 * id + 2] * 1.0 + S->omgpp[3 * id + 2] * 1.0);
        aa = w1 * w1 + w2 * w2 + w3 * w3;
        bb = (xp - xc) * w1 + (yp - yc) * w2 + (zp - zc) * w3;
        cc = (xp - xc) * (xp - xc) + (yp - yc) * (yp - yc) + (zp - zc) * (zp - zc) -
             (P->rad * 1.0) * (P->rad * 1.0);
        d0 = bb / aa;
        d1 = sqrt(d0 * d0 - cc / aa);
        ddt = -d0 + d1;
        if (ddt < 0.0) ddt = 0.0;
        if (ddt > 1.0) ddt = 1.0;
        x1 = xp + w1 * ddt;
        y1 = yp + w2 * ddt;
        z1 = zp + w3 * ddt;
        x2 = x1 + (o2 * (z1 - zc) - o3 * (y1 - yc)) * ddt;
        y2 = y1 + (o3 * (x1 - xc) - o1 * (z1 - zc)) * ddt;
        z2 = z1 + (o1 * (y1 - yc) - o2 * (x1 - xc)) * ddt;
        xx0 = x2 - xc;
        yy0 = y2 - yc;
        zz0 = z2 - zc;
        prod0 = -100.0;
        for (ip = 1; ip < NPOP; ++ip) {
          double prod = (double)CX[ip] * xx0 + (double)CY[ip] * yy0 + (double)CZ[ip] * zz0;
          if (ip <= 6) prod = prod * sqrt(2.0);
          if (prod > prod0) {
            ipmx = ip;
            prod0 = prod;
          }
        }
        {
          int a = wrap(i + CX[ipmx], nx), b = wrap(j + CY[ipmx], ny), c = wrap(k + CZ[ipmx], nz);
          long q = nid(P, a, b, c);
          const int beyond = P->walls_y && (j + CY[ipmx] < 0 || j + CY[ipmx] >= ny);
          if (!beyond && S->ibnodes[q] < 0 && S->ibnodes0[q] < 0) {
            rho9 = 0.0; u9 = 0.0; v9 = 0.0; w9 = 0.0;
            g = S->f + 19 * q;
            for (ip = 0; ip < NPOP; ++ip) {
              rho9 = rho9 + g[ip];
              u9 = u9 + (double)CX[ip] * g[ip];
              v9 = v9 + (double)CY[ip] * g[ip];
              w9 = w9 + (double)CZ[ip] * g[ip];
            }
            orc_feqpnt(P, u9, v9, w9, rho9, feq);
            for (ip = 0; ip < NPOP; ++ip) S->f[ip + 19 * m] = g[ip] - feq[ip];
          } else {
            for (ip = 0; ip < NPOP; ++ip) S->f[ip + 19 * m] = 0.0;
          }
        }
        nghb = 0;
        rho9 = 0.0;
        for (ip = 1; ip < NPOP; ++ip) {
          int a = wrap(i + CX[ip], nx), b = wrap(j + CY[ip], ny), c = wrap(k + CZ[ip], nz);
          long q = nid(P, a, b, c);
          if (P->walls_y && (j + CY[ip] < 0 || j + CY[ip] >= ny)) continue;
          if (S->ibnodes[q] < 0 && S->ibnodes0[q] < 0) {
            int ii;
            nghb++;
            g = S->f + 19 * q;
            for (ii = 0; ii < NPOP; ++ii) rho9 = rho9 + g[ii];
          }
        }
        if (nghb > 0)
          rho9 = rho9 / (double)nghb;
        else
          rho9 = P->rho0;
        xx0 = xp - xc;
        yy0 = yp - yc;
        zz0 = zp - zc;
        w1 = S->wp[3 * id]; w2 = S->wp[3 * id + 1]; w3 = S->wp[3 * id + 2];
        o1 = S->omgp[3 * id]; o2 = S->omgp[3 * id + 1]; o3 = S->omgp[3 * id + 2];
        u9 = w1 + o2 * zz0 - o3 * yy0;
        v9 = w2 + o3 * xx0 - o1 * zz0;
        w9 = w3 + o1 * yy0 - o2 * xx0;
        orc_feqpnt(P, u9, v9, w9, rho9, feq);
        for (ip = 0; ip < NPOP; ++ip) S->f[ip + 19 * m] = S->f[ip + 19 * m] + feq[ip];
      }
}

/* particle release (BOX/main.f90:149-165) */
void orc_release(const orc_params *P, orc_state *S) {
  S->istep00 = 0;
  orc_initpartpos(P, S);
  orc_initpartvel(P, S);
  orc_initpartomg(P, S);
  orc_beads_links(P, S);
  S->istep00 = 1;
}

/* one pass of the time loop body (BOX/main.f90:172-188) for step S->istep.
 * The particle release of BOX/main.f90:149-165 is the caller's (orc_release).
 * Returns the overlap count after redistribution (nonzero is fatal). */
int orc_timestep(const orc_params *P, orc_state *S) {
  int ovl = 0;
  const int on = P->ipart && S->istep >= P->irelease;
  if (P->forcing_on) orc_forcing(P, S);
  orc_collision_mrt(P, S);
  if (on) orc_beads_collision(P, S);
  orc_streaming(P, S);
  if (on) {
    orc_beads_lubforce(P, S);
    orc_beads_move(P, S);
    ovl = orc_beads_redistribute(P, S);
    orc_beads_links(P, S);
    orc_beads_filling(P, S);
  }
  orc_macrovar(P, S);
  return ovl;
}
