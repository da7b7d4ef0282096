// lbm_common.h -- lattice tables and small helpers shared by the host code,
// the HIP kernels and the CPU check build.
//
// D3Q19 ordering and opposite table follow the reference
// (3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/para.f90:171-175):
// 0 rest; 1..6 = +x,-x,+y,-y,+z,-z; 7..10 = (+,+,0),(-,+,0),(+,-,0),(-,-,0);
// 11..14 = (+,0,+),(-,0,+),(+,0,-),(-,0,-); 15..18 = (0,+,+),(0,-,+),(0,+,-),(0,-,-).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MFC_HD __host__ __device__ __forceinline__
#else
#define MFC_HD inline
#endif

namespace mfc {

constexpr int NPOP = 19;

// Direction components packed two bits per direction ((c + 1) in bits 2i,2i+1)
// so runtime-indexed lookups in device code need no memory table.
constexpr uint64_t pack_dirs(const int (&t)[NPOP]) {
  uint64_t r = 0;
  for (int i = 0; i < NPOP; ++i) r |= (uint64_t)(t[i] + 1) << (2 * i);
  return r;
}
constexpr int kCX[NPOP] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCY[NPOP] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCZ[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr uint64_t kPackX = pack_dirs(kCX);
constexpr uint64_t kPackY = pack_dirs(kCY);
constexpr uint64_t kPackZ = pack_dirs(kCZ);

MFC_HD int cx(int i) { return (int)((kPackX >> (2 * i)) & 3u) - 1; }
MFC_HD int cy(int i) { return (int)((kPackY >> (2 * i)) & 3u) - 1; }
MFC_HD int cz(int i) { return (int)((kPackZ >> (2 * i)) & 3u) - 1; }
// opposite direction: axis pairs (1,2),(3,4),(5,6) are adjacent; each edge
// quartet b..b+3 (b = 7, 11, 15) pairs b <-> b+3 and b+1 <-> b+2.
MFC_HD int opp(int i) {
  if (i == 0) return 0;
  if (i <= 6) return (i & 1) ? i + 1 : i - 1;
  const int b = 7 + 4 * ((i - 7) >> 2);
  return 2 * b + 3 - i;
}

// Collision constants handed to the kernels by value.
struct CollideConsts {
  double s1, s2, s4, s9, s10, s13, s16;   // MRT rates (para.f90:102-135)
  double omegepsl, omegepslj, omegxx;     // equilibrium weights (para.f90:102-111)
  double ww0, ww1, ww2;                   // lattice weights (para.f90:167-169)
  double bfx, bfy, bfz;                   // uniform body force added to the forcing field
};

// Population layouts of the resident buffer (where the post-collision value
// of direction i at node n lives).
//   LAY_AB       two buffers: the current one, plane i, slot n
//   LAY_AA_EVEN  one buffer (in-place AA storage) after an even collide:
//                plane opp(i), slot n
//   LAY_AA_ODD   one buffer after an odd collide: plane i at the node
//                n + c_i (plane opp(i), slot n where n + c_i is beyond a wall)
// In LAY_AA_ODD the streamed value f_i(n) is plane i, slot n: the natural
// layout an upload writes and the first collide reads in place.
enum PopLayout { LAY_AB = 0, LAY_AA_EVEN = 1, LAY_AA_ODD = 2 };

// Grid geometry handed to the kernels.
struct Grid {
  int nx, ny, nz;
  int walls_y;        // 1: half-way bounce-back walls on the y faces
  int lay;            // PopLayout of the resident buffer
  uint64_t ncell;     // nx*ny*nz
  // Population element of (plane q, node x + nx*(y + ny*z)):
  //   q * pitch + x + rs * (y + ny*z).
  // Plane-major layout: pitch >= ncell, rs = nx (each plane a contiguous
  // field).  Row-interleaved layout: pitch = nx, rs = 19*nx (the 19 planes'
  // rows of one (y, z) side by side).
  uint64_t pitch;     // elements from one population plane to the next
  uint64_t rs;        // elements from one row of a plane to the next row
};

MFC_HD int wrapi(int a, int n) {
  a = (a < 0) ? a + n : a;
  return (a >= n) ? a - n : a;
}

// Nearest periodic image of a centre coordinate c relative to a node
// coordinate pnt: the two conditional shifts used throughout the reference
// (collision.f90:295-305, partlib.f90:1384-1391, 2862-2869).
MFC_HD double near_image(double c, double pnt, int n) {
  const double h = (double)(n / 2);
  if ((c - pnt) > h) c = c - (double)n;
  if ((c - pnt) < -h) c = c + (double)n;
  return c;
}

}  // namespace mfc
