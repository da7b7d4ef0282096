// initial.hip -- synthetic initial velocity fields on the device.
//
// A band-limited field u_c(x) = scale * sum_k 2 Re(a_c(k) e^{2 pi i k.x/n})
// over 0 <= kx <= 16, |ky|, |kz| <= 16 (x at the node indices), evaluated
// separably, one component at a time:
//   B1(kz, ky, x) = sum_kx a(kx, ky, kz) e^{2 pi i kx x/nx}   (33 x 33 x nx)
//   B2(kz, y, x)  = sum_ky B1(kz, ky, x) e^{2 pi i ky y/ny}    (33 x ny x nx)
//   u(x, y, z)    = 2 scale Re sum_kz B2(kz, y, x) e^{2 pi i kz z/nz}
// The last pass keeps a node column's 33 B2 values in registers and walks z,
// so it reads B2 once and writes the field once: a 1024^3 FP64 field costs a
// few tens of milliseconds instead of a host synthesis and a 26 GB upload.
// Phases come from the context's (cos, sin)(2 pi m / n) tables.
#include "internal.h"

namespace mfc {

constexpr int KM = 16;            // largest |k| component
constexpr int NK = 2 * KM + 1;    // ky, kz values
constexpr int NKX = KM + 1;       // kx values

__device__ __forceinline__ int phase_index(int k, int x, int n) {
  int m = (int)(((int64_t)k * x) % n);
  return m < 0 ? m + n : m;
}

// B1[(iz * NK + iy) * nx + x]
__global__ void __launch_bounds__(256)
    k_modes_x(const double2* __restrict__ a, const double* __restrict__ phx, int nx,
              double2* __restrict__ B1) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)(NK * NK) * (uint32_t)nx) return;
  const int x = (int)(t % (uint32_t)nx);
  const int row = (int)(t / (uint32_t)nx);   // iz * NK + iy
  double sr = 0.0, si = 0.0;
  for (int kx = 0; kx < NKX; ++kx) {
    const double2 v = a[row * NKX + kx];
    const int m = phase_index(kx, x, nx);
    const double c = phx[2 * m], s = phx[2 * m + 1];
    sr += v.x * c - v.y * s;
    si += v.x * s + v.y * c;
  }
  B1[t] = make_double2(sr, si);
}

// B2[(iz * ny + y) * nx + x]
__global__ void __launch_bounds__(256)
    k_modes_y(const double2* __restrict__ B1, const double* __restrict__ phy, int nx, int ny,
              double2* __restrict__ B2) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t plane = (uint64_t)nx * (uint64_t)ny;
  if (t >= (uint64_t)NK * plane) return;
  const int iz = (int)(t / plane);
  const uint32_t r = (uint32_t)(t - (uint64_t)iz * plane);
  const int y = (int)(r / (uint32_t)nx), x = (int)(r - (uint32_t)y * (uint32_t)nx);
  double sr = 0.0, si = 0.0;
  for (int iy = 0; iy < NK; ++iy) {
    const double2 v = B1[((size_t)iz * NK + iy) * nx + x];
    const int m = phase_index(iy - KM, y, ny);
    const double c = phy[2 * m], s = phy[2 * m + 1];
    sr += v.x * c - v.y * s;
    si += v.x * s + v.y * c;
  }
  B2[t] = make_double2(sr, si);
}

__global__ void __launch_bounds__(256)
    k_modes_z(const double2* __restrict__ B2, const double* __restrict__ phz, int nx, int ny,
              int nz, double scale, double* __restrict__ u) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint64_t plane = (uint64_t)nx * (uint64_t)ny;
  if (t >= plane) return;
  double2 b[NK];
#pragma unroll
  for (int iz = 0; iz < NK; ++iz) b[iz] = B2[(uint64_t)iz * plane + t];
  const double s2 = 2.0 * scale;
  for (int z = 0; z < nz; ++z) {
    double acc = 0.0;
#pragma unroll
    for (int iz = 0; iz < NK; ++iz) {
      const int m = phase_index(iz - KM, z, nz);
      acc += b[iz].x * phz[2 * m] - b[iz].y * phz[2 * m + 1];
    }
    u[(uint64_t)z * plane + t] = s2 * acc;
  }
}

__global__ void k_zero_f64(double* __restrict__ a, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

int init_mode_kx() { return NKX; }
int init_mode_k() { return NK; }

// amp: (re, im) of a_c(kx, ky, kz) at flat index kx + NKX * ((ky + KM) + NK * ((kz + KM) + NK * c)),
// already on the device; B1, B2 scratch of NK*NK*nx and NK*nx*ny double2.
hipError_t launch_init_velocity(Ctx& c, const double2* amp, double scale, double2* B1, double2* B2) {
  const Grid& g = c.g;
  double* comps[3] = {c.ux, c.uy, c.uz};
  const uint64_t plane = (uint64_t)g.nx * (uint64_t)g.ny;
  for (int comp = 0; comp < 3; ++comp) {
    const double2* a = amp + (size_t)comp * NKX * NK * NK;
    const uint32_t n1 = (uint32_t)(NK * NK) * (uint32_t)g.nx;
    hipLaunchKernelGGL(k_modes_x, dim3((n1 + 255) / 256), dim3(256), 0, c.stream, a, c.ft.phx, g.nx, B1);
    const uint64_t n2 = (uint64_t)NK * plane;
    hipLaunchKernelGGL(k_modes_y, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, c.stream, B1,
                       c.ft.phy, g.nx, g.ny, B2);
    hipLaunchKernelGGL(k_modes_z, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c.stream, B2,
                       c.ft.phz, g.nx, g.ny, g.nz, scale, comps[comp]);
  }
  hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((g.ncell + 255) / 256)), dim3(256), 0, c.stream, c.rho,
                     (uint64_t)g.ncell);
  return hipGetLastError();
}

}  // namespace mfc
