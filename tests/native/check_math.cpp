// CPU build of the library's per-node math headers (the same source the HIP
// kernels compile), exported for the CPU tests.
#include "../../multiphase_flow_codes_amd/csrc/lbm_collide.h"
#include "../../multiphase_flow_codes_amd/csrc/lbm_particle_math.h"

extern "C" {

// f: n x 19 (in/out); per-node rho, ux, uy, uz, Fx, Fy, Fz; k: 16 constants
void chk_collide(int n, double* f, const double* rho, const double* ux, const double* uy,
                 const double* uz, const double* Fx, const double* Fy, const double* Fz,
                 const double* k, int fp32) {
  mfc::CollideConsts c{k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8], k[9],
                       k[10], k[11], k[12], 0.0, 0.0, 0.0};
  for (int q = 0; q < n; ++q) {
    if (fp32) {
      float g[19];
      for (int i = 0; i < 19; ++i) g[i] = (float)f[19 * q + i];
      mfc::mrt_collide<float>(g, (float)rho[q], (float)ux[q], (float)uy[q], (float)uz[q],
                              (float)Fx[q], (float)Fy[q], (float)Fz[q], c);
      for (int i = 0; i < 19; ++i) f[19 * q + i] = g[i];
    } else {
      double g[19];
      for (int i = 0; i < 19; ++i) g[i] = f[19 * q + i];
      mfc::mrt_collide<double>(g, rho[q], ux[q], uy[q], uz[q], Fx[q], Fy[q], Fz[q], c);
      for (int i = 0; i < 19; ++i) f[19 * q + i] = g[i];
    }
  }
}

void chk_moments(int n, const double* f, double* out4) {
  for (int q = 0; q < n; ++q) {
    double g[19];
    for (int i = 0; i < 19; ++i) g[i] = f[19 * q + i];
    double r, jx, jy, jz;
    mfc::density_momentum<double>(g, r, jx, jy, jz);
    out4[4 * q] = r; out4[4 * q + 1] = jx; out4[4 * q + 2] = jy; out4[4 * q + 3] = jz;
  }
}

int chk_dirs(int* out) {  // cx, cy, cz, opp for 19 directions
  for (int i = 0; i < 19; ++i) {
    out[4 * i] = mfc::cx(i); out[4 * i + 1] = mfc::cy(i); out[4 * i + 2] = mfc::cz(i);
    out[4 * i + 3] = mfc::opp(i);
  }
  return 19;
}

double chk_bb(double alpha, double ff1, double ff2a, double ff3a, int sol1, int sol2, double ffo,
              double ffo1, double wt, double uwpro) {
  return mfc::bounce_back_value(alpha, ff1, ff2a, ff3a, sol1, sol2, ffo, ffo1, wt, uwpro);
}

double chk_sweep_rest(double c, int n) { return mfc::sweep_rest_base(c, n); }

}
