/*
 * mfc_lbm.h -- C ABI of libmfc_lbm, the MI355X (gfx950) implementation of the
 * D3Q19 MRT lattice-Boltzmann timestep with finite-size moving spheres.
 *
 * Each entry point replaces one argument-less subroutine of the reference
 * driver (3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/, "BOX/" below);
 * the cited file:line is the routine it stands in for.  The Fortran side
 * binds them with ISO_C_BINDING (multiphase_flow_codes_amd/fortran/).
 *
 * Layout conventions at the boundary (reference layout, one rank, whole box):
 *   f        double f(0:18,lx,ly,lz)   flat index ip + 19*(i + nx*(j + ny*k))
 *   fields   double a(lx,ly,lz)         flat index i + nx*(j + ny*k)
 *   maps     int32  ibnodes/isnodes(lx,ly,lz); ibnodes = +1 solid, -1 fluid;
 *            isnodes = particle id (1-based) or -1
 *   links    int32 ilink/mlink(1:18,lx,ly,lz), double alink(1:18,lx,ly,lz)
 *   particle double a(3,npart)          flat index c + 3*p
 *   modes    double a1r(6,5,5) etc.     forcing-mode coefficients
 * Populations are deviations f_i - w_i*rho0, as in the reference.
 *
 * Every call returns MFC_OK (0) or an MFC_ERR_* code; mfc_last_error() gives
 * the message.  Work is queued on the context's HIP stream; calls that hand
 * data back to the host synchronise.  Faults detected on the device
 * (overlap, capacity) are reported by the next synchronising call.
 * One host thread per context.
 */
#ifndef MFC_LBM_H
#define MFC_LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFC_OK 0
#define MFC_ERR_ARG 1        /* invalid argument or size */
#define MFC_ERR_HIP 2        /* HIP runtime error */
#define MFC_ERR_STATE 3      /* call not valid in the current state */
#define MFC_ERR_OVERLAP 4    /* particles overlap (BOX/partlib.f90:2641-2646) */
#define MFC_ERR_CAPACITY 5   /* particle or link capacity exceeded (BOX/partlib.f90:2582-2590) */
#define MFC_ERR_NONFINITE 6  /* particle state not finite or outside the box */
#define MFC_ERR_NODEVICE 7   /* no HIP device */

#define MFC_PREC_FP64 0          /* populations and arithmetic in FP64 (the drop-in) */
#define MFC_PREC_FP32 1          /* populations stored and collided in FP32 */
#define MFC_PREC_FP32_F64MATH 2  /* populations stored in FP32, collided in FP64 */

#define MFC_STORE_AB 0  /* two population buffers (2 x 19 values per node) */
#define MFC_STORE_AA 1  /* one buffer updated in place (AA pattern): 19 values per node,
                           e.g. a 1024^3 FP64 box in 163 GB; not for ensembles */

/* Run constants: what the reference holds in module var_inc after `para`
 * (BOX/para.f90:1-288, BOX/var_inc.f90:14-29). */
typedef struct mfc_params {
  int32_t nx, ny, nz;          /* lattice size (var_inc.f90:14) */
  int32_t precision;           /* MFC_PREC_FP64 (the drop-in), MFC_PREC_FP32 or MFC_PREC_FP32_F64MATH */
  int32_t npart_max;           /* particle capacity (npart, var_inc.f90:18) */
  int32_t walls_y;             /* 1: half-way bounce-back walls on the y faces */
  int32_t iseedf;              /* forcing RNG state after para (negative: reseed), para.f90:281 */
  int32_t irelease;            /* release step (var_inc.f90:24) */
  int32_t storage;             /* MFC_STORE_AB or MFC_STORE_AA */
  double s1, s2, s4, s9, s10, s13, s16;   /* MRT rates (para.f90:113-135) */
  double omegepsl, omegepslj, omegxx;     /* MRT equilibrium weights (para.f90:102-111) */
  double ww0, ww1, ww2;                   /* lattice weights (para.f90:167-169) */
  double rho0;                            /* var_inc.f90:27 */
  double rad, rhopart, mingap;            /* var_inc.f90:27-28 */
  double volp, amp, aip, rhog;            /* para.f90:230-239 */
  double stf0, stf1, gscale;              /* para.f90:237-245 */
  double vscale, tscale, xkf, pi2;        /* forcing scales (para.f90:37,60; gaussian.f90:195) */
  double body_force[3];                   /* uniform force added to the forcing field */
  double particle_force[3];               /* uniform force per unit sphere volume on the spheres
                                             only (e.g. the buoyancy-corrected weight) */
  double mingap_w, stf0_w, stf1_w;        /* walls_y: sphere-wall lubrication range (also the
                                             placement clearance) and stiffness
                                             (3DChannel_.../var_inc.f90:31, para.f90:241-242) */
} mfc_params;

typedef struct mfc_ctx mfc_ctx;

/* ---- life cycle ---------------------------------------------------------- */
/* Allocates the device state of one box (replaces allocarray, BOX/para.f90:291-355). */
int mfc_create(const mfc_params* p, mfc_ctx** out);
int mfc_destroy(mfc_ctx* ctx);
const char* mfc_last_error(const mfc_ctx* ctx);
/* Waits for queued work; returns the first pending device fault, if any. */
int mfc_sync(mfc_ctx* ctx);
/* Device status words (16 int32): overlap, capacity, contact-link counters. */
int mfc_status(mfc_ctx* ctx, int32_t* out16);
/* Queue work on an existing HIP stream instead of the context's own. */
int mfc_set_stream(mfc_ctx* ctx, void* hip_stream);
void* mfc_get_stream(mfc_ctx* ctx);

/* ---- state transfer (reference layout) ----------------------------------- */
int mfc_upload_f(mfc_ctx* ctx, const double* f);
int mfc_download_f(mfc_ctx* ctx, double* f);
/* rho/ux/uy/uz arrays of var_inc (BOX/var_inc.f90:91-92) */
int mfc_upload_macro(mfc_ctx* ctx, const double* rho, const double* ux, const double* uy,
                     const double* uz);
int mfc_download_macro(mfc_ctx* ctx, double* rho, double* ux, double* uy, double* uz);
/* force_realx/y/z of the last FORCING (BOX/var_inc.f90:79-80) */
int mfc_download_force(mfc_ctx* ctx, double* fx, double* fy, double* fz);
/* Forcing-process state, the content of the reference's force.<step> file
 * (input_outputf, BOX/saveload.f90:217-274): RNG iseedf/iyf/ivf(32), OU
 * modes b1r..b3r(6,5,5) and projected modes a1r..a3r(6,5,5). */
int mfc_get_forcing_state(mfc_ctx* ctx, int32_t* iseedf, int32_t* iyf, int32_t* ivf,
                          double* b1r, double* b2r, double* b3r, double* a1r, double* a2r,
                          double* a3r);
int mfc_set_forcing_state(mfc_ctx* ctx, int32_t iseedf, int32_t iyf, const int32_t* ivf,
                          const double* b1r, const double* b2r, const double* b3r,
                          const double* a1r, const double* a2r, const double* a3r);

/* ---- fluid step (BOX/main.f90:172-188) ----------------------------------- */
/* Synthetic initial velocity (no counterpart in the reference, which starts
 * from rest, BOX/main.f90:73-75; it replaces the upload of a host-generated
 * field before initpop): ux/uy/uz = scale * sum over the table of
 * 2 Re(a_c(k) exp(2 pi i (kx x/nx + ky y/ny + kz z/nz))) at the node indices
 * x, y, z = 0..n-1, rho = 0.  amp holds (re, im) of a_c for 0 <= kx <= 16 and
 * -16 <= ky, kz <= 16 at double index
 * 2 * (kx + 17 * ((ky + 16) + 33 * ((kz + 16) + 33 * c))) (+1 for im), c = 0..2. */
int mfc_init_velocity(mfc_ctx* ctx, const double* amp, double scale);
int mfc_initpop(mfc_ctx* ctx);          /* BOX/initial.f90:3-28 */
int mfc_forcing(mfc_ctx* ctx);          /* BOX/gaussian.f90:182-372 */
int mfc_collision_mrt(mfc_ctx* ctx);    /* BOX/collision.f90:8-210 */
int mfc_streaming(mfc_ctx* ctx);        /* BOX/streaming.f90:2-171 */
/* rhoupdat (BOX/collision.f90:213-229); optionally also returns
 * max|rho_new - rho_old|, the pre-relaxation residual of BOX/main.f90:91-96. */
int mfc_rhoupdat(mfc_ctx* ctx, double* max_abs_change);
int mfc_macrovar(mfc_ctx* ctx);         /* BOX/collision.f90:232-332 */

/* ---- particles (BOX/partlib.f90) ----------------------------------------- */
/* Sets npart particles; field order as in mfc_prt_get.  NULL fields are zeroed. */
#define MFC_PF_YPGLB 0
#define MFC_PF_YPGLBP 1
#define MFC_PF_WP 2
#define MFC_PF_WPP 3
#define MFC_PF_OMGP 4
#define MFC_PF_OMGPP 5
#define MFC_PF_FHIP 6
#define MFC_PF_FLUBP 7
#define MFC_PF_TORQP 8
#define MFC_PF_FORCEPP 9
#define MFC_PF_TORQPP 10
#define MFC_PF_DWDT 11
#define MFC_PF_DOMGDT 12
#define MFC_PF_THETAP 13
#define MFC_PF_COUNT 14
int mfc_prt_init(mfc_ctx* ctx, int32_t npart);
int mfc_prt_set(mfc_ctx* ctx, int32_t field, const double* a);
int mfc_prt_get(mfc_ctx* ctx, int32_t field, double* a);
int mfc_initpartpos(mfc_ctx* ctx, int32_t* iseedp, int32_t mingap_check); /* partlib.f90:1-80 */
int mfc_initpartvel(mfc_ctx* ctx);      /* BOX/partlib.f90:177-263 */
int mfc_initpartomg(mfc_ctx* ctx);      /* BOX/partlib.f90:837-931 (spin set to 0, :929) */
int mfc_beads_links(mfc_ctx* ctx);      /* BOX/partlib.f90:934-1058 */
int mfc_beads_collision(mfc_ctx* ctx);  /* BOX/partlib.f90:1064-1717 */
/* Order of the bounce-back links (BOX/partlib.f90:1363-1621), for
 * mfc_beads_collision and mfc_step.  MFC_BB_SNAPSHOT (the default) evaluates
 * every link against one post-collision field.  MFC_BB_REFERENCE reproduces
 * the reference's one-rank loop, which stores each link's value into the
 * solid node at once: a link whose own node is solid for another sphere
 * (spheres closer than about a node) then reads the value an earlier link
 * stored there.  The two differ only for such contact links (counted in
 * status word 4); the reference order evaluates them in one workgroup and
 * holds at most 65536 of them (MFC_ERR_CAPACITY beyond). */
#define MFC_BB_SNAPSHOT 0
#define MFC_BB_REFERENCE 1
int mfc_set_bb_order(mfc_ctx* ctx, int32_t order);
int mfc_beads_lubforce(mfc_ctx* ctx);   /* BOX/partlib.f90:2211-2292 */
int mfc_beads_move(mfc_ctx* ctx, int32_t istep);         /* BOX/partlib.f90:2295-2394 */
int mfc_beads_redistribute(mfc_ctx* ctx, int32_t istep); /* BOX/partlib.f90:2397-2648 */
int mfc_beads_filling(mfc_ctx* ctx);    /* BOX/partlib.f90:2793-3111 */
/* ibnodes/isnodes (current, which=0) or ibnodes0/isnodes0 (which=1). */
int mfc_get_maps(mfc_ctx* ctx, int32_t which, int32_t* ibnodes, int32_t* isnodes);
/* Dense ilink/mlink/alink in the reference layout (small boxes only). */
int mfc_get_links(mfc_ctx* ctx, int32_t* ilink, int32_t* mlink, double* alink);

/* ---- whole step ---------------------------------------------------------- */
/* nsteps passes of the time-loop body (BOX/main.f90:172-188) starting at step
 * number istep (the release at istep==irelease is the caller's, as in main).
 * With MFC_OVERLAP=1 or 2 in the environment at mfc_create, the particle
 * phase of each step runs on a second stream beside the next collide of the
 * lattice tiles far from every sphere; results are bitwise the same, and a
 * sphere moving a node or more in one step is then reported as a fault. */
int mfc_step(mfc_ctx* ctx, int32_t istep, int32_t nsteps);
/* Accumulated device time of the collide-stream kernel while timing is on. */
/* ---- FP64 diagnostics on the device ---------------------------------------- */
/* Sums over all nodes of the fields macrovar defines (BOX/collision.f90:232-332,
 * as statistc and rmsstat read them, BOX/saveload.f90:1111-1300):
 * out[0] = sum rho, out[1..3] = sum ux, uy, uz, out[4] = sum |u|^2,
 * out[5] = max |u|^2.  Fixed reduction order: bitwise repeatable. */
int mfc_diag_sums(mfc_ctx* ctx, double* out6);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that must not call
 * hipMalloc): mfc_diag_sums with its first-level partials and results in a
 * static device array of the library (6 x 1024 + 8 doubles per GPU,
 * resolved when a context is created) instead of buffers allocated on first
 * use.  out6 is DEVICE memory for the six values (same order as
 * mfc_diag_sums); the work is queued on the context's stream without
 * synchronising.  The rho/u arrays must exist already (MFC_ERR_STATE
 * otherwise; mfc_download_macro, mfc_diag_sums or an upload create them);
 * rho/u of a virtual macrovar are written into them as mfc_diag_sums does.
 * Nothing is allocated or freed.  Calls that overlap on one GPU share the
 * array and race. */
int mfc_diag_sums_noalloc(mfc_ctx* ctx, double* out6);
/* Kinetic-energy spectrum of statistc (BOX/saveload.f90:1273-1300): u
 * transformed with forward sign and 1/n per axis (BOX/mpifft.f90:52-108);
 * ek[ik-1] = sum over modes 0 <= kx < nx/2 (weight 1/2 at kx = 0), all ky, kz,
 * with int(|k| + 0.5) = ik, for ik = 1..nek.  Uses a complex FP64 scratch
 * field of 16 bytes per node, allocated on first use. */
int mfc_diag_spectrum(mfc_ctx* ctx, int32_t nek, double* ek);

/* ---- ensemble engine ------------------------------------------------------- */
/* No counterpart in the reference, which runs one box per job.  For batched
 * and serving use: many independent boxes, each a context created and
 * initialised with the calls above (different sizes, spheres, rates and body
 * forces; one precision and wall setting per ensemble; body-force driving
 * only, xkf = 0), advanced together with one grid per step stage through
 * per-box argument and block-offset tables.  A member steps bitwise as it
 * would alone (mfc_step).  While a box is a member its calls are queued on the
 * ensemble's stream. */
typedef struct mfc_ens mfc_ens;
int mfc_ens_create(mfc_ens** out);
int mfc_ens_destroy(mfc_ens* e);      /* members are released, not destroyed */
const char* mfc_ens_last_error(const mfc_ens* e);
/* Join at the end of a step (after macrovar, and beads_links if it has
 * spheres); istep = the number of the member's next step. */
int mfc_ens_add(mfc_ens* e, mfc_ctx* box, int32_t istep);
int mfc_ens_remove(mfc_ens* e, mfc_ctx* box);
int mfc_ens_size(const mfc_ens* e);
/* Steps go in chunks of up to 8: the tables of a chunk's steps are built on
 * the host and uploaded with one copy, then its stage grids run. */
int mfc_ens_step(mfc_ens* e, int32_t nsteps);
/* Waits for queued work; returns the first pending fault of any member. */
int mfc_ens_sync(mfc_ens* e);
/* Threads per collide block (64, 128 or 256): the pencil of consecutive x
 * nodes one workgroup collides. */
int mfc_ens_set_block(mfc_ens* e, int32_t bs);
/* Collide tile order: 0 launch order, 1 blocks dealt to the 8 XCDs in
 * contiguous ranges (results do not depend on it). */
int mfc_ens_set_tile_order(mfc_ens* e, int32_t order);
/* on = 1: each chunk's stage grids (kernel nodes only; the table upload
 * stays a plain copy) are recorded once as a HIP graph and replayed for
 * every chunk of the same launch shapes (a few recordings are kept); on = 0:
 * plain launches. */
int mfc_ens_set_graph(mfc_ens* e, int32_t on);
/* Times `steps` steps of every (block size, tile order) pair after `steps`
 * untimed ones, on the current members (they advance 12 * steps steps), and
 * keeps the fastest: chosen[0] = block size, chosen[1] = tile order;
 * ms_per_step6[k] = time of block {64,128,256}[k % 3] with order k / 3.
 * chosen and ms_per_step6 may be null. */
int mfc_ens_autotune(mfc_ens* e, int32_t steps, int32_t* chosen, double* ms_per_step6);
/* Host time spent building and enqueueing steps since the last call (waits
 * for table buffers excluded) and the number of steps it covers; resets both. */
int mfc_ens_host_time(mfc_ens* e, double* ms, int64_t* steps);
void* mfc_ens_get_stream(mfc_ens* e);

/* Collide-kernel timing with HIP events on the library's stream: every = N
 * brackets every N-th collide launch from now on (the first one included)
 * with an event pair, 0 turns it off.  Each event marker idles the GPU for a
 * few microseconds, so a sample (N > 1) keeps that out of the step time.
 * mfc_get_timing returns the summed duration of the timed launches and their
 * number. */
int mfc_set_timing(mfc_ctx* ctx, int32_t every);
int mfc_get_timing(mfc_ctx* ctx, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* MFC_LBM_H */
