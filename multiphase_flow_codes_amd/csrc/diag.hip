// diag.hip -- FP64 diagnostics on the device: global sums of the macroscopic
// fields and the kinetic-energy spectrum of the reference's statistc
// (BOX/saveload.f90:1111-1300), with its own FFT.
//
// Sums use a fixed two-level tree (per-block partials, then one block), so
// they are bitwise repeatable.  The spectrum transforms one velocity
// component at a time through a complex FP64 scratch field: radix-2 in LDS
// for power-of-two lengths, a direct DFT otherwise; each axis is normalised
// by 1/n as mpifft3DRC does (BOX/mpifft.f90:52-108), forward sign e^{-i}.
#include <mutex>

#include "internal.h"

namespace mfc {

constexpr int DB = 256;          // threads per reduction block
constexpr int kSumBlocks = 1024; // partial sums of the first level

// Scratch of the allocation-free field sums (mfc_diag_sums_noalloc): the
// kSumBlocks x 6 first-level partials, then the 6 results it leaves before
// they are copied out.  Every GPU the library runs on holds its own copy
// until the process ends.  The host resolves its address once per device
// (sums_scratch) and passes that pointer to the kernels; the array's name in
// host code is only a stand-in.  Each call overwrites every word it reads,
// so nothing is cleared between calls; calls that overlap on one GPU share
// the array and race.
constexpr size_t kSumsScratchDoubles = 6 * (size_t)kSumBlocks + 8;
static __device__ __align__(256) double scratch[kSumsScratchDoubles];

// Per-block partial sums of rho, ux, uy, uz and u.u, and the block's max |u|^2.
__global__ void __launch_bounds__(DB)
    k_field_sums(const double* __restrict__ rho, const double* __restrict__ ux,
                 const double* __restrict__ uy, const double* __restrict__ uz, uint64_t n,
                 double* __restrict__ part) {
  __shared__ double red[6][DB / 64];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * DB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * DB) {
    const double a = ux[i], b = uy[i], c = uz[i];
    const double q = a * a + b * b + c * c;
    s[0] += rho[i];
    s[1] += a;
    s[2] += b;
    s[3] += c;
    s[4] += q;
    s[5] = fmax(s[5], q);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = s[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double t = __shfl_down(v, o, 64);
      v = (k == 5) ? fmax(v, t) : v + t;
    }
    if (lane == 0) red[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double v = red[k][0];
    for (int q = 1; q < DB / 64; ++q) v = (k == 5) ? fmax(v, red[k][q]) : v + red[k][q];
    part[6 * blockIdx.x + k] = v;
  }
}

__global__ void k_sums_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  if (threadIdx.x >= 6) return;
  const int k = threadIdx.x;
  double v = part[k];
  for (int b = 1; b < nblk; ++b) v = (k == 5) ? fmax(v, part[6 * b + k]) : v + part[6 * b + k];
  out[k] = v;
}

// One length-n FFT (or DFT) per block along `axis` of the complex field Z
// (x fastest).  With `re` set the line is read from the real field instead.
// tw[j] = exp(-2 pi i j / n).  Dynamic LDS: 2 n double2.
__global__ void __launch_bounds__(256)
    k_fft_axis(double2* __restrict__ Z, const double* __restrict__ re, Grid g, int axis,
               const double2* __restrict__ tw, int n, int log2n) {
  extern __shared__ double2 lbuf[];
  double2* a = lbuf;
  double2* o = lbuf + n;
  const uint64_t line = blockIdx.x;
  uint64_t base, es;
  if (axis == 0) {
    base = line * (uint64_t)g.nx;
    es = 1;
  } else if (axis == 1) {
    const uint64_t x = line % (uint64_t)g.nx, z = line / (uint64_t)g.nx;
    base = x + (uint64_t)g.nx * (uint64_t)g.ny * z;
    es = (uint64_t)g.nx;
  } else {
    base = line;
    es = (uint64_t)g.nx * (uint64_t)g.ny;
  }
  const bool pow2 = log2n >= 0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const uint64_t idx = base + (uint64_t)j * es;
    const double2 v = re ? make_double2(re[idx], 0.0) : Z[idx];
    const int d = pow2 ? (int)(__brev((unsigned)j) >> (32 - log2n)) : j;
    a[d] = v;
  }
  __syncthreads();
  double2* res = a;
  if (pow2) {
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int t = threadIdx.x; t < n / 2; t += blockDim.x) {
        const int grp = t / half, pos = t - grp * half;
        const int i0 = grp * len + pos, i1 = i0 + half;
        const double2 w = tw[pos * step];
        const double2 u = a[i0], v = a[i1];
        const double br = v.x * w.x - v.y * w.y, bi = v.x * w.y + v.y * w.x;
        a[i0] = make_double2(u.x + br, u.y + bi);
        a[i1] = make_double2(u.x - br, u.y - bi);
      }
      __syncthreads();
    }
  } else {
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      double sr = 0.0, si = 0.0;
      int m = 0;   // (j * k) mod n, advanced incrementally
      for (int j = 0; j < n; ++j) {
        const double2 w = tw[m], v = a[j];
        sr += v.x * w.x - v.y * w.y;
        si += v.x * w.y + v.y * w.x;
        m += k;
        if (m >= n) m -= n;
      }
      o[k] = make_double2(sr, si);
    }
    __syncthreads();
    res = o;
  }
  const double inv = 1.0 / (double)n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double2 v = res[j];
    Z[base + (uint64_t)j * es] = make_double2(v.x * inv, v.y * inv);
  }
}

__device__ __forceinline__ int signed_k(int i, int n) { return i < n / 2 + 1 ? i : i - n; }

// Kinetic-energy shells of the transformed component Z (statistc,
// BOX/saveload.f90:1273-1300): modes with 0 <= kx < nx/2, all ky, kz; weight
// 1/2 on the kx = 0 plane; shell ik = int(|k| + 0.5), k = 0 excluded.
// Per-block partial shells (LDS, FP64 atomics), summed in block order later.
__global__ void __launch_bounds__(256)
    k_spectrum_bins(const double2* __restrict__ Z, Grid g, int nek, double* __restrict__ part) {
  extern __shared__ double bins[];
  for (int i = threadIdx.x; i < nek; i += blockDim.x) bins[i] = 0.0;
  __syncthreads();
  const int hx = g.nx / 2;
  const uint64_t nmodes = (uint64_t)hx * g.ny * g.nz;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nmodes;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const int kx = (int)(t % (uint64_t)hx);
    const uint64_t r = t / (uint64_t)hx;
    const int j = (int)(r % (uint64_t)g.ny), k = (int)(r / (uint64_t)g.ny);
    const int ky = signed_k(j, g.ny), kz = signed_k(k, g.nz);
    const double k2 = (double)(kx * kx + ky * ky + kz * kz);
    const int ik = (int)(sqrt(k2) + 0.5);
    if (ik < 1 || ik > nek) continue;
    const double2 v = Z[(uint64_t)kx + (uint64_t)g.nx * ((uint64_t)j + (uint64_t)g.ny * k)];
    double e = v.x * v.x + v.y * v.y;
    if (kx == 0) e *= 0.5;
    atomicAdd(&bins[ik - 1], e);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nek; i += blockDim.x) part[(size_t)blockIdx.x * nek + i] = bins[i];
}

__global__ void k_fill_spectrum_zero(double* __restrict__ acc, int nek) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nek) acc[i] = 0.0;
}

__global__ void k_spectrum_final(const double* __restrict__ part, int nblk, int nek,
                                 double* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nek) return;
  double v = 0.0;
  for (int b = 0; b < nblk; ++b) v += part[(size_t)b * nek + i];
  acc[i] += v;
}

// ---------------------------------------------------------------------------
hipError_t launch_field_sums(Ctx& c, double* part, double* out) {
  hipLaunchKernelGGL(k_field_sums, dim3(kSumBlocks), dim3(DB), 0, c.stream, c.rho, c.ux, c.uy, c.uz,
                     (uint64_t)c.g.ncell, part);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, c.stream, part, kSumBlocks, out);
  return hipGetLastError();
}

static int ilog2_exact(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

int diag_blocks() { return kSumBlocks; }

size_t sums_scratch_doubles() { return kSumsScratchDoubles; }

// Device address of `scratch` on the current device, resolved with
// hipGetSymbolAddress the first time a context is created on that device
// and kept per device ordinal.
hipError_t sums_scratch(double** out) {
  static std::mutex mu;
  static double* ptr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[dev] = (double*)p;
  }
  *out = ptr[dev];
  return hipSuccess;
}

hipError_t launch_spectrum(Ctx& c, int nek, double2* Z, const double2* const tw[3], double* part,
                           double* acc) {
  const int ns[3] = {c.g.nx, c.g.ny, c.g.nz};
  const double* comps[3] = {c.ux, c.uy, c.uz};
  hipLaunchKernelGGL(k_fill_spectrum_zero, dim3((nek + 255) / 256), dim3(256), 0, c.stream, acc, nek);
  for (int comp = 0; comp < 3; ++comp) {
    for (int axis = 0; axis < 3; ++axis) {
      const int n = ns[axis];
      const uint64_t lines = c.g.ncell / (uint64_t)n;
      const int threads = n >= 512 ? 256 : (n >= 128 ? 128 : 64);
      hipLaunchKernelGGL(k_fft_axis, dim3((unsigned)lines), dim3(threads), sizeof(double2) * 2 * n,
                         c.stream, Z, axis == 0 ? comps[comp] : nullptr, c.g, axis, tw[axis], n,
                         ilog2_exact(n));
    }
    hipLaunchKernelGGL(k_spectrum_bins, dim3(kSumBlocks), dim3(256), sizeof(double) * nek, c.stream,
                       (const double2*)Z, c.g, nek, part);
    hipLaunchKernelGGL(k_spectrum_final, dim3((nek + 255) / 256), dim3(256), 0, c.stream, part,
                       kSumBlocks, nek, acc);
  }
  return hipGetLastError();
}

}  // namespace mfc
