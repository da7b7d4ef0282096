// Forcing table of one step: per (y,z) row the nine partial sums
//   G_{c,kx}(y,z) = sum_{ky,kz} A_c(kx,ky,kz) exp(i 2 pi (ky y/ny + kz z/nz))
// over the live modes of the reference's 3 x 5 x 5 set (gaussian.f90:243-359).
// Evaluated by k_force_table and, inside a multi-step call, by extra blocks
// of the refill grid. a*b+c is contracted within one expression here
// whatever the including file's setting, so both give the same table bit
// for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace mfc {

// A table evaluation riding in another kernel's grid: `nblk` blocks of 256
// threads write slot `slot` (array G).
struct TableJob {
  ModeSet M;
  const double* phy;
  const double* phz;
  double* G;
  int ny, nz;
  int slot;
  int nblk;
};

__device__ __forceinline__ void pick5(const double (&r)[5], const double (&i)[5], int d, double& a,
                                      double& b) {
  switch (d) {
    case -2: a = r[0]; b = i[0]; break;
    case -1: a = r[1]; b = i[1]; break;
    case 0: a = r[2]; b = i[2]; break;
    case 1: a = r[3]; b = i[3]; break;
    default: a = r[4]; b = i[4]; break;
  }
}

// Block `block` of the table (256 threads). The mode list is first copied
// into LDS by all threads at once: read from the kernel arguments inside the
// mode loop it is a chain of dependent scalar loads (each mode's ky, kz
// select the next phases), which set the kernel's time.
__device__ __forceinline__ void force_table_block(const ModeSet& M, const double* __restrict__ phy,
                                                  const double* __restrict__ phz, int ny, int nz,
                                                  double* __restrict__ G, uint32_t block) {
#pragma clang fp contract(on)
  __shared__ double sa[kMaxModes * 6];
  __shared__ int sk[kMaxModes];   // ky + 8 * kz, both in -2..2
  const int nm = M.n[0] + M.n[1] + M.n[2];
  for (int t = threadIdx.x; t < nm * 6; t += 256) sa[t] = (&M.a[0][0])[t];
  for (int t = threadIdx.x; t < nm; t += 256) sk[t] = (int)M.ky[t] + 8 * (int)M.kz[t];
  __syncthreads();
  // one thread per (row, component): a row's 18 doubles come from three
  // adjacent lanes, 6 each, so a wavefront stores one contiguous run
  const uint32_t t = block * 256u + threadIdx.x;
  const uint32_t row = t / 3u;
  const int cc = (int)(t - 3u * row);
  if (row >= (uint32_t)ny * (uint32_t)nz) return;
  const int y = (int)(row % (uint32_t)ny), z = (int)(row / (uint32_t)ny);
  double eyr[5], eyi[5], ezr[5], ezi[5];
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    const int my = wrapi((d * y) % ny, ny), mz = wrapi((d * z) % nz, nz);
    eyr[d + 2] = phy[2 * my];
    eyi[d + 2] = phy[2 * my + 1];
    ezr[d + 2] = phz[2 * mz];
    ezi[d + 2] = phz[2 * mz + 1];
  }
  double g[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) g[q] = 0.0;
  int m = 0;
#pragma unroll
  for (int kx = 0; kx < 3; ++kx) {
    for (int e = 0; e < M.n[kx]; ++e, ++m) {
      // the same mode for every lane: uniform selection of the phases
      const int k = __builtin_amdgcn_readfirstlane(sk[m]);
      const int kz = (k + 2 * 8 + 2) / 8 - 2, ky = k - 8 * kz;
      double a, b, c, d;
      pick5(eyr, eyi, ky, a, b);
      pick5(ezr, ezi, kz, c, d);
      const double er = a * c - b * d, ei = a * d + b * c;
      const double ar = sa[6 * m + 2 * cc], ai = sa[6 * m + 2 * cc + 1];
      g[2 * kx] += ar * er - ai * ei;
      g[2 * kx + 1] += ar * ei + ai * er;
    }
  }
  double* out = G + 18ull * row + 6 * cc;
#pragma unroll
  for (int q = 0; q < 6; ++q) out[q] = g[q];
}

}  // namespace mfc
