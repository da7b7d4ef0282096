/*
 * mfc_oracle.h -- CPU FP64 oracle for the D3Q19 MRT lattice-Boltzmann step
 * with finite-size moving spheres.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room C restatement of the
 * reference algorithm (3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/,
 * abbreviated BOX/ below) used as the checker for the HIP library.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it; the product never calls into it.
 *
 * Conventions (all follow the reference):
 *   - populations are deviations f_i - w_i*rho0 (BOX/initial.f90:14-25);
 *   - f is stored as f(0:18,lx,ly,lz) column-major, i.e. the flat index of
 *     population ip at node (i,j,k) (0-based) is ip + 19*(i + nx*(j + ny*k));
 *   - per-node scalar fields (rho, ux, ..., ibnodes) use i + nx*(j + ny*k);
 *   - link arrays ilink/mlink/alink(1:18,lx,ly,lz) use (ip-1) + 18*node;
 *   - particle arrays (3,npart) use c + 3*p (p 0-based, ids in maps 1-based);
 *   - one rank, periodic in x and z; y periodic too unless walls_y, in which
 *     case half-way bounce-back walls bound y (node j sits at j + 1/2 between
 *     walls at 0 and ny), centres are never imaged across them, links and
 *     interpolation stencils beyond them count as solid, and spheres feel the
 *     walls' lubrication force (the channel code's wall semantics,
 *     3DChannel_Decaying_HIT_Finite_Size_Moving_Particles/partlib.f90, with
 *     its wall axis moved from x to y).
 */
#ifndef MFC_ORACLE_H
#define MFC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int nx, ny, nz;
  int ipart;          /* particles on (BOX/para.f90:226) */
  int npart;
  int irelease;       /* step at which particles are released (BOX/var_inc.f90:24) */
  int bb_mode;        /* 0: reference single-rank order, 1: snapshot order */
  int forcing_on;     /* 1: call FORCING each step (BOX/main.f90:172) */
  int walls_y;        /* 1: half-way bounce-back walls on the y faces */
  double rad, rhopart, mingap, rho0;
  double s1, s2, s4, s9, s10, s13, s16;
  double omegepsl, omegepslj, omegxx;
  double ww0, ww1, ww2;
  double volp, amp, aip, rhog, stf0, stf1, gscale;
  double vscale, tscale, xkf, pi2;
  double body_force[3];      /* uniform force density on the fluid, added to the forcing
                                field in collision and macrovar, and V_p times it on spheres */
  double particle_force[3];  /* uniform force per unit sphere volume on the spheres only */
  double mingap_w, stf0_w, stf1_w;  /* sphere-wall lubrication and placement clearance
                                       (3DChannel_.../var_inc.f90:31, para.f90:241-242) */
} orc_params;

typedef struct orc_state {
  double *f;                            /* 19*ncell */
  double *rho, *ux, *uy, *uz;           /* ncell */
  double *fx, *fy, *fz;                 /* force_realx/y/z, ncell */
  int *ibnodes, *isnodes, *ibnodes0, *isnodes0;   /* ncell */
  int *ilink, *mlink;                   /* 18*ncell */
  double *alink;                        /* 18*ncell */
  double *yp;                           /* rank-local positions (one rank: by id) */
  double *ypglb, *ypglbp, *wp, *wpp, *omgp, *omgpp;
  double *fHIp, *flubp, *torqp;
  double *dwdt, *domgdt, *forcep, *forcepp, *torqpp, *thetap;  /* 3*npart */
  double a1r[150], a2r[150], a3r[150], b1r[150], b2r[150], b3r[150];
  int ivf[32], ivp[32];
  int iseedf, iyf, iseedp, iyp;
  int istep, istep00;
  int nconflict;      /* last beads_collision: links whose node is solid */
  int alpha_faults;   /* last beads_links: alpha outside [0,1] */
  int novlp;          /* last overlap check: flagged particles */
} orc_state;

double orc_ranf(int *idum, int *iy, int *iv);
void orc_forcing_modes(const orc_params *P, orc_state *S);
void orc_force_field(const orc_params *P, const double *a1r, const double *a2r,
                     const double *a3r, double *fx, double *fy, double *fz);
void orc_forcing(const orc_params *P, orc_state *S);
void orc_initpop(const orc_params *P, orc_state *S);
void orc_collision_mrt(const orc_params *P, orc_state *S);
void orc_streaming(const orc_params *P, orc_state *S);
void orc_rhoupdat(const orc_params *P, orc_state *S);
void orc_macrovar(const orc_params *P, orc_state *S);
int orc_ovlpchk(const orc_params *P, orc_state *S, int *flags);
int orc_initpartpos(const orc_params *P, orc_state *S);
void orc_initpartvel(const orc_params *P, orc_state *S);
void orc_initpartomg(const orc_params *P, orc_state *S);
void orc_beads_links(const orc_params *P, orc_state *S);
void orc_beads_collision(const orc_params *P, orc_state *S);
void orc_beads_lubforce(const orc_params *P, orc_state *S);
void orc_beads_move(const orc_params *P, orc_state *S);
int orc_beads_redistribute(const orc_params *P, orc_state *S);
void orc_beads_filling(const orc_params *P, orc_state *S);
void orc_feqpnt(const orc_params *P, double u, double v, double w, double rho,
                double *feq);
void orc_release(const orc_params *P, orc_state *S);
int orc_timestep(const orc_params *P, orc_state *S);
int orc_nthreads(void);

#ifdef __cplusplus
}
#endif
#endif
