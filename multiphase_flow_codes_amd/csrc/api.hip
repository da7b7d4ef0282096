// api.hip -- C ABI of libmfc_lbm: context life cycle, host/device transfers,
// the state machine that lets the reference's separate subroutine calls run
// as the fused pull-scheme kernels, and the host half of FORCING.
//
// State machine.  The resident buffer holds post-collision values; the
// reference's `f` is either that buffer itself (phase PRE_STREAM, between
// collision and streaming) or its pulled view (phase POST_STREAM, after
// streaming).  `streaming` only flips the phase.  The reference's rho/ux/
// uy/uz arrays are either materialised (MOM_ARRAYS: uploaded, initpop,
// rhoupdat) or virtual (MOM_VIRTUAL: set by macrovar, recomputed inside the
// collide kernel from the pulled field and the forcing captured at
// macrovar), valid while the field is unchanged since that macrovar.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "force_table.h"
#include "internal.h"

using namespace mfc;


namespace {

constexpr int kTimingPool = 8192;
// contact links the reference-order bounce-back can hold (48 B each)
constexpr int64_t kOrderCap = int64_t(1) << 16;

int fail(mfc_ctx* ctx, int code, const char* what, hipError_t e = hipSuccess) {
  if (ctx) {
    char buf[512];
    if (e != hipSuccess)
      snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
      snprintf(buf, sizeof buf, "%s", what);
    ctx->c.err = buf;
  }
  return code;
}

#define HIPCHK(expr)                                        \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return fail(ctx, MFC_ERR_HIP, #expr, e_); \
  } while (0)

#define NEEDCTX \
  if (!ctx) return MFC_ERR_ARG

__global__ void k_set_status(int32_t* st, int idx, int val) { st[idx] = val; }
__global__ void k_set_double(double* d, double v) { d[0] = v; }
__global__ void k_fill_i32(int32_t* a, uint64_t n, int32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_f64(double* a, uint64_t n, double v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

hipError_t set_status(Ctx& c, int idx, int val) {
  hipLaunchKernelGGL(k_set_status, dim3(1), dim3(1), 0, c.stream, c.d_status, idx, val);
  return hipGetLastError();
}

// ---- FORCING, host half (gaussian.f90:182-306) ----------------------------
// Park-Miller minimal standard generator with Bays-Durham shuffle
// (Numerical Recipes ran1), as ranff/ranpp (gaussian.f90:374-404).
double ran1(int32_t* idum, int32_t* iy, int32_t* iv) {
  const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
  const int32_t NDIV = 1 + (IM - 1) / NTAB;
  const double AM = 1.0 / IM, RNMX = 1.0 - 1.2e-7;
  if (*idum <= 0 || *iy == 0) {
    *idum = std::max(-*idum, (int32_t)1);
    for (int j = NTAB + 8; j >= 1; --j) {
      const int32_t k = *idum / IQ;
      *idum = IA * (*idum - k * IQ) - IR * k;
      if (*idum < 0) *idum += IM;
      if (j <= NTAB) iv[j - 1] = *idum;
    }
    *iy = iv[0];
  }
  const int32_t k = *idum / IQ;
  *idum = IA * (*idum - k * IQ) - IR * k;
  if (*idum < 0) *idum += IM;
  const int32_t j = 1 + *iy / NDIV;
  *iy = iv[j - 1];
  iv[j - 1] = *idum;
  const double r = AM * (double)*iy;
  return r < RNMX ? r : RNMX;
}

inline int mix(int i, int j, int k) { return (i - 1) + 6 * ((j - 1) + 5 * (k - 1)); }
// kx of the (re, im) slot i = 1..6: both slots of a pair carry the pair's kx
inline double kx_of(int i) { return (double)((i - 1) / 2); }
inline double kyz_of(int j) { return j <= 3 ? (double)(j - 1) : (double)(j - 6); }

// kx = 0 plane made Hermitian so the physical field is real
void hermitian_kx0(double* q) {
  for (int k = 1; k <= 5; ++k) {
    const int kk = (k == 1) ? 1 : 7 - k;
    for (int j = 2; j <= 3; ++j) {
      const int jj = 7 - j;
      q[mix(1, jj, kk)] = q[mix(1, j, k)];
      q[mix(2, jj, kk)] = -q[mix(2, j, k)];
    }
  }
  for (int k = 2; k <= 3; ++k) {
    const int kk = 7 - k;
    q[mix(1, 1, kk)] = q[mix(1, 1, k)];
    q[mix(2, 1, kk)] = -q[mix(2, 1, k)];
  }
}

// One Euler step of the Ornstein-Uhlenbeck mode process with white-noise
// increments, solenoidal projection and Hermitian symmetry.
void forcing_modes_host(Ctx& c) {
  double fr[3][150];
  const double var = 447.31 * (c.p.vscale * c.p.tscale) * (c.p.vscale * c.p.tscale);
  const double tf = 0.038 / c.p.tscale, dt = 1.0, smallx = 1.0e-18;
  for (int mc = 0; mc < 3; ++mc)
    for (int k = 1; k <= 5; ++k)
      for (int j = 1; j <= 5; ++j)
        for (int i = 1; i <= 6; i += 2) {
          double c1 = ran1(&c.iseedf, &c.iyf, c.ivf);
          if (c1 < 1.0e-6) c1 = 1.0e-6;
          const double rmag = sqrt(-(4.0 * var * dt * log(c1) / tf));
          const double rph = ran1(&c.iseedf, &c.iyf, c.ivf);
          fr[mc][mix(i, j, k)] = rmag * cos(c.p.pi2 * rph);
          fr[mc][mix(i + 1, j, k)] = rmag * sin(c.p.pi2 * rph);
        }
  for (int mc = 0; mc < 3; ++mc) {
    hermitian_kx0(fr[mc]);
    fr[mc][mix(1, 1, 1)] = 0.0;
    fr[mc][mix(2, 1, 1)] = 0.0;
  }
  for (int k = 1; k <= 5; ++k)
    for (int j = 1; j <= 5; ++j)
      for (int i = 1; i <= 6; ++i) {
        const int a = mix(i, j, k);
        const double kx = kx_of(i), ky = kyz_of(j), kz = kyz_of(k);
        const double radm = sqrt(kx * kx + ky * ky + kz * kz);
        if (radm < c.p.xkf) {
          for (int mc = 0; mc < 3; ++mc) c.bmod[mc][a] = c.bmod[mc][a] * (1.0 - dt / tf) + fr[mc][a];
          const double t1 = kx * c.bmod[0][a], t2 = ky * c.bmod[1][a], t3 = kz * c.bmod[2][a];
          double divb = (t1 + t2 + t3) / ((radm + smallx) * (radm + smallx));
          if (radm == 0.0) divb = 0.0;
          c.amod[0][a] = c.bmod[0][a] - kx * divb;
          c.amod[1][a] = c.bmod[1][a] - ky * divb;
          c.amod[2][a] = c.bmod[2][a] - kz * divb;
        } else {
          c.amod[0][a] = c.amod[1][a] = c.amod[2][a] = 0.0;
        }
      }
  for (int mc = 0; mc < 3; ++mc) {
    c.amod[mc][mix(1, 1, 1)] = 0.0;
    c.amod[mc][mix(2, 1, 1)] = 0.0;
    hermitian_kx0(c.amod[mc]);
  }
}

// The table job of the modes in c.amod, into a slot no queued collide or
// macrovar reads; job.nblk = 0 when no mode is live.
int prepare_modes(mfc_ctx* ctx, TableJob& job) {
  Ctx& c = ctx->c;
  int slot = 0;
  while (slot == c.ft.cur || slot == c.ft.mac || slot == c.ft.prev) ++slot;
  ModeSet& m = job.M;
  int total = 0;
  for (int kx = 0; kx < 3; ++kx) {
    m.n[kx] = 0;
    for (int k = 1; k <= 5; ++k)
      for (int j = 1; j <= 5; ++j) {
        const int re = mix(2 * kx + 1, j, k), im = re + 1;
        bool live = false;
        for (int mc = 0; mc < 3; ++mc) live = live || c.amod[mc][re] != 0.0 || c.amod[mc][im] != 0.0;
        if (!live) continue;
        if (total == kMaxModes) return fail(ctx, MFC_ERR_ARG, "FORCING: more live modes than the table kernel takes (xkf > sqrt(8))");
        m.ky[total] = (int8_t)kyz_of(j);
        m.kz[total] = (int8_t)kyz_of(k);
        for (int mc = 0; mc < 3; ++mc) {
          m.a[total][2 * mc] = c.amod[mc][re];
          m.a[total][2 * mc + 1] = c.amod[mc][im];
        }
        ++m.n[kx];
        ++total;
      }
  }
  const uint32_t rows = (uint32_t)c.g.ny * (uint32_t)c.g.nz;
  job.phy = c.ft.phy;
  job.phz = c.ft.phz;
  job.G = c.ft.g[slot];
  job.ny = c.g.ny;
  job.nz = c.g.nz;
  job.slot = slot;
  job.nblk = total > 0 ? (int)((3 * rows + 255) / 256) : 0;
  return MFC_OK;
}

// The job's table becomes the one the next collide reads.
void make_current(Ctx& c, const TableJob& job) {
  c.ft.live[job.slot] = job.nblk > 0 ? 1 : 0;
  c.ft.prev = c.ft.cur;
  c.ft.cur = job.slot;
}

// Build the forcing table of the modes in c.amod and make it current.
int publish_modes(mfc_ctx* ctx) {
  Ctx& c = ctx->c;
  TableJob job;
  int rc = prepare_modes(ctx, job);
  if (rc) return rc;
  if (job.nblk > 0) HIPCHK(launch_force_table(c, job.slot, job.M));
  make_current(c, job);
  return MFC_OK;
}

int ensure_macro(mfc_ctx* ctx) {
  Ctx& c = ctx->c;
  if (c.rho) return MFC_OK;
  const size_t b = sizeof(double) * c.g.ncell;
  HIPCHK(hipMalloc(&c.rho, b));
  HIPCHK(hipMalloc(&c.ux, b));
  HIPCHK(hipMalloc(&c.uy, b));
  HIPCHK(hipMalloc(&c.uz, b));
  const unsigned nb = (unsigned)((c.g.ncell + 255) / 256);
  double* arrs[4] = {c.rho, c.ux, c.uy, c.uz};
  for (double* a : arrs) hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, c.stream, a, (uint64_t)c.g.ncell, 0.0);
  HIPCHK(hipGetLastError());
  return MFC_OK;
}

// Make the rho/u arrays hold the reference's values (materialising a
// virtual macrovar if needed).  A materialised macrovar stays virtual: the
// arrays are a copy for the caller, and the next collide still recomputes
// the moments itself, so a step's result (in FP32 arithmetic too) does not
// depend on whether the host looked at rho/u before it.
int need_arrays(mfc_ctx* ctx) {
  Ctx& c = ctx->c;
  if (c.mom == MOM_ARRAYS) return MFC_OK;
  if (c.mom == MOM_VIRTUAL) {
    if (c.mom_version != c.f_version)
      return fail(ctx, MFC_ERR_STATE, "rho/u requested after the field changed since macrovar");
    if (c.rho && c.arrays_version == c.mom_version) return MFC_OK;
    int rc = ensure_macro(ctx);
    if (rc) return rc;
    HIPCHK(launch_macro(c, c.phase == PHASE_POST_STREAM));
    c.arrays_version = c.mom_version;
    return MFC_OK;
  }
  int rc = ensure_macro(ctx);
  if (rc) return rc;
  c.mom = MOM_ARRAYS;
  return MFC_OK;
}

int ensure_staging(mfc_ctx* ctx, size_t bytes) {
  Ctx& c = ctx->c;
  if (c.staging_bytes >= bytes) return MFC_OK;
  if (c.staging) HIPCHK(hipFree(c.staging));
  c.staging = nullptr;
  c.staging_bytes = 0;
  HIPCHK(hipMalloc(&c.staging, bytes));
  c.staging_bytes = bytes;
  return MFC_OK;
}

int slab_planes(const Ctx& c) {
  const size_t plane = (size_t)c.g.nx * c.g.ny * 19 * sizeof(double);
  const size_t cap = (size_t)256 << 20;
  int n = (int)std::max<size_t>(1, cap / plane);
  return std::min(n, c.g.nz);
}

bool particles_on(const Ctx& c) { return c.pt.np > 0; }

int check_faults(mfc_ctx* ctx) {
  Ctx& c = ctx->c;
  HIPCHK(hipMemcpyAsync(c.h_status, c.d_status, sizeof(int32_t) * ST_COUNT, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  const int32_t* s = c.h_status;
  char buf[256];
  if (s[ST_OVERLAP] > 0) {
    snprintf(buf, sizeof buf, "fault: %d particles overlap at istep = %d", s[ST_OVERLAP], s[ST_OVERLAP_STEP]);
    return fail(ctx, MFC_ERR_OVERLAP, buf);
  }
  if (s[ST_LINK_OVERFLOW] > 0) return fail(ctx, MFC_ERR_CAPACITY, "fault: boundary-link pool exceeded");
  if (s[ST_ORDER_CAP] > 0) {
    snprintf(buf, sizeof buf, "fault: more contact links than the reference-order bounce-back holds (%lld)",
             (long long)kOrderCap);
    return fail(ctx, MFC_ERR_CAPACITY, buf);
  }
  if (s[ST_CELL_OVERFLOW] > 0)
    return fail(ctx, MFC_ERR_CAPACITY, "fault: too many particles in one pair-search cell (overlapping particles)");
  if (s[ST_NONFINITE] > 0 || s[ST_BOX_FAULT] > 0)
    return fail(ctx, MFC_ERR_NONFINITE, "fault: particle position not finite or outside the box");
  if (c.overlap_used && s[ST_FAST] > 0)
    return fail(ctx, MFC_ERR_NONFINITE, "fault: a sphere moved a node or more in one step (the overlapped "
                                        "step needs less; leave MFC_OVERLAP unset to run without it)");
  return MFC_OK;
}

// beads_links into the map buffer that does not hold the previous step's map
// (the one beads_filling still needs), clearing what that buffer marked
// before.  With `overlap_check` the links pass also runs ovlpchk.
int links_now(mfc_ctx* ctx, int overlap_check, int istep, bool with_find = false) {
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  const int target = (P.map_cur != P.map_old) ? P.map_cur : (P.map_cur ^ 1);
  HIPCHK(launch_links(c, target, overlap_check, istep, with_find));
  P.map_cur = target;
  return MFC_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

int mfc_create(const mfc_params* p, mfc_ctx** out) {
  if (!p || !out) return MFC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MFC_ERR_NODEVICE;
  if (p->nx < 4 || p->ny < 4 || p->nz < 4) return MFC_ERR_ARG;
  const uint64_t ncell = (uint64_t)p->nx * p->ny * p->nz;
  if (ncell >= (1ull << 31)) return MFC_ERR_ARG;
  if (p->precision != MFC_PREC_FP64 && p->precision != MFC_PREC_FP32 &&
      p->precision != MFC_PREC_FP32_F64MATH)
    return MFC_ERR_ARG;
  if (p->storage != MFC_STORE_AB && p->storage != MFC_STORE_AA) return MFC_ERR_ARG;
  mfc_ctx* ctx = new (std::nothrow) mfc_ctx();
  if (!ctx) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  c.p = *p;
  c.g.nx = p->nx;
  c.g.ny = p->ny;
  c.g.nz = p->nz;
  c.g.walls_y = p->walls_y ? 1 : 0;
  c.g.ncell = ncell;
  c.g.pitch = ncell;
  // in-place storage starts with the streamed field in each node's own slots
  c.g.lay = p->storage == MFC_STORE_AA ? LAY_AA_ODD : LAY_AB;
  c.cc = CollideConsts{p->s1, p->s2, p->s4, p->s9, p->s10, p->s13, p->s16, p->omegepsl, p->omegepslj,
                       p->omegxx, p->ww0, p->ww1, p->ww2, p->body_force[0], p->body_force[1],
                       p->body_force[2]};
  c.prec = p->precision;
  c.tsize = c.prec == MFC_PREC_FP64 ? 8 : 4;
  {
    const char* s = getenv("MFC_BB_SPLIT");
    if (s && s[0] >= '1' && s[0] <= '0' + kMaxBBSplit) c.pt.bb_split = s[0] - '0';
  }
  {
    const char* e = getenv("MFC_XCD_REMAP");
    c.xcd_remap = (e && e[0] >= '0' && e[0] <= '2') ? e[0] - '0' : -1;
  }
  {
    // The overlapped step is opt-in: at 1024^3 with 20,000 spheres the
    // bounce-back beside the far-tile collide gets too small a share of
    // the CUs and the step ends later than the plain sequence.
    // MFC_OVERLAP=1: the far collide beside the whole particle phase;
    // MFC_OVERLAP=2: beside the phase after the bounce-back, which then runs
    // with the GPU to itself
    const char* e = getenv("MFC_OVERLAP");
    c.overlap = (e && (e[0] == '1' || e[0] == '2')) ? e[0] - '0' : 0;
  }
  {
    // Population layout.  By default the 19 planes' rows of one (y, z) lie
    // side by side (pitch = nx, rs = 19 nx): a collide workgroup's 38
    // streams then fall in a few neighbouring pages instead of 38 planes
    // 4 GB apart.  Measured against the plane-major layout, alternating in
    // one box: FP32 collide 2.6 % faster at 1024^3 and 8 % at 256^3 (step
    // 2.5 % and 5 %), FP64 in place 2.3 % faster at 1024^3, FP64 and FP32
    // storage with FP64 collision unchanged at 256^3; padding the plane
    // stride or the row group was slower in a probe of the access pattern.
    // MFC_POP_LAYOUT=planes selects plane-major planes instead, 512 B
    // further apart than the node count (a power-of-two plane size would
    // start every stream at the same offset of each page; 512 B measured
    // 1.5-2 % faster than no padding there); MFC_PLANE_PAD overrides that
    // pad in 256-B units.
    const char* e = getenv("MFC_POP_LAYOUT");
    if (e && e[0] == 'p') {
      const char* d = getenv("MFC_PLANE_PAD");
      const long pad = d ? atol(d) : 2;
      if (pad >= 0 && pad < (1L << 20)) c.g.pitch = ncell + (uint64_t)pad * (256 / c.tsize);
      c.g.rs = (uint64_t)c.g.nx;
    } else {
      c.g.pitch = (uint64_t)c.g.nx;
      c.g.rs = (uint64_t)NPOP * c.g.nx;
    }
  }
  c.iseedf = p->iseedf;
  int rc = MFC_OK;
  auto bail = [&](int code) {
    mfc_destroy(ctx);
    return code;
  };
#define CK(expr)                                                         \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) { rc = fail(ctx, MFC_ERR_HIP, #expr, e_); fprintf(stderr, "mfc_create: %s\n", ctx->c.err.c_str()); return bail(rc); } \
  } while (0)
  CK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  c.own_stream = true;
  // one past the last element: plane 18, the last row, x = nx - 1
  const size_t pelems = (size_t)(NPOP - 1) * c.g.pitch + (size_t)(ncell / c.g.nx - 1) * c.g.rs + c.g.nx;
  const size_t pb = std::max(pelems, (size_t)NPOP * (c.g.rs == (uint64_t)c.g.nx ? c.g.pitch : 0)) * c.tsize;
  CK(hipMalloc(&c.pop[0], pb));
  CK(hipMemsetAsync(c.pop[0], 0, pb, c.stream));
  if (c.g.lay == LAY_AB) {
    CK(hipMalloc(&c.pop[1], pb));
    CK(hipMemsetAsync(c.pop[1], 0, pb, c.stream));
  }
  const size_t gb = sizeof(double) * 18 * (size_t)p->ny * p->nz;
  for (int s = 0; s < kForceSlots; ++s) CK(hipMalloc(&c.ft.g[s], gb));
  c.ft.cur = 0;
  c.ft.mac = 0;
  {
    std::vector<double> ph(2 * (p->nx + p->ny + p->nz));
    double* px = ph.data();
    double* py = px + 2 * p->nx;
    double* pz = py + 2 * p->ny;
    for (int m = 0; m < p->nx; ++m) { px[2 * m] = cos(p->pi2 * m / p->nx); px[2 * m + 1] = sin(p->pi2 * m / p->nx); }
    for (int m = 0; m < p->ny; ++m) { py[2 * m] = cos(p->pi2 * m / p->ny); py[2 * m + 1] = sin(p->pi2 * m / p->ny); }
    for (int m = 0; m < p->nz; ++m) { pz[2 * m] = cos(p->pi2 * m / p->nz); pz[2 * m + 1] = sin(p->pi2 * m / p->nz); }
    CK(hipMalloc(&c.ft.phx, sizeof(double) * ph.size()));
    c.ft.phy = c.ft.phx + 2 * p->nx;
    c.ft.phz = c.ft.phy + 2 * p->ny;
    CK(hipMemcpy(c.ft.phx, ph.data(), sizeof(double) * ph.size(), hipMemcpyHostToDevice));
  }
  CK(hipMalloc(&c.d_status, sizeof(int32_t) * ST_COUNT));
  CK(hipMemsetAsync(c.d_status, 0, sizeof(int32_t) * ST_COUNT, c.stream));
  CK(hipHostMalloc((void**)&c.h_status, sizeof(int32_t) * ST_COUNT, hipHostMallocDefault));
  CK(hipMalloc(&c.d_red, sizeof(double) * 8));
  CK(hipHostMalloc((void**)&c.h_red, sizeof(double) * 8, hipHostMallocDefault));
  // this device's static scratch of the allocation-free sums
  CK(sums_scratch(&c.sums_scratch));
  CK(hipStreamSynchronize(c.stream));
#undef CK
  *out = ctx;
  return MFC_OK;
}

int mfc_destroy(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.ens) return fail(ctx, MFC_ERR_STATE, "box is an ensemble member: remove it first");
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(c.pop[0]); F(c.pop[1]);
  F(c.rho); F(c.ux); F(c.uy); F(c.uz);
  for (int s = 0; s < kForceSlots; ++s) F(c.ft.g[s]);
  F(c.ft.phx);
  Particles& P = c.pt;
  double* pd[] = {P.yp, P.ypglb, P.ypglbp, P.wp, P.wpp, P.omgp, P.omgpp, P.fHIp, P.flubp, P.torqp,
                  P.forcep, P.forcepp, P.torqpp, P.dwdt, P.domgdt, P.thetap, P.lk_alpha, P.seg_ft};
  for (double* a : pd) F(a);
  F(P.lk_out); F(P.ct_list); F(P.ct_count);
  F(P.isn[0]); F(P.isn[1]); F(P.mask[0]); F(P.mask[1]);
  F(P.lk_node); F(P.lk_dir); F(P.lk_cnt); F(P.lk_off); F(P.lk_part);
  F(P.sol_list[0]); F(P.sol_list[1]); F(P.sol_cnt[0]); F(P.sol_cnt[1]);
  F(P.sh_node); F(P.sh_hits); F(P.sh_cnt);
  F(P.fill_list); F(P.fill_count); F(P.cell_cnt); F(P.cell_list); F(P.ord);
  F(c.d_status); F(c.d_red); F(c.staging);
  F(c.diag_part); F(c.diag_out); F(c.diag_Z);
  for (int a = 0; a < 3; ++a) F(c.diag_tw[a]);
  {
    TileLists& T = c.tiles;
    F(T.mark); F(T.cnt); F(T.off_near); F(T.off_far); F(T.tot); F(T.near_list); F(T.far_list);
    if (T.h_tot) (void)hipHostFree(T.h_tot);
    if (c.ev_coll) (void)hipEventDestroy(c.ev_coll);
    if (c.ev_tiles) (void)hipEventDestroy(c.ev_tiles);
    if (c.ev_part) (void)hipEventDestroy(c.ev_part);
    if (c.ev_bb) (void)hipEventDestroy(c.ev_bb);
    if (c.pstream) (void)hipStreamDestroy(c.pstream);
  }
  if (c.h_status) (void)hipHostFree(c.h_status);
  if (c.h_red) (void)hipHostFree(c.h_red);
  for (auto e : c.tpool) (void)hipEventDestroy(e);
  c.tpool.clear();
  if (c.stream && c.own_stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
  return MFC_OK;
}

const char* mfc_last_error(const mfc_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int mfc_sync(mfc_ctx* ctx) {
  NEEDCTX;
  return check_faults(ctx);
}

int mfc_status(mfc_ctx* ctx, int32_t* out16) {
  NEEDCTX;
  if (!out16) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  HIPCHK(hipMemcpyAsync(c.h_status, c.d_status, sizeof(int32_t) * ST_COUNT, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  memcpy(out16, c.h_status, sizeof(int32_t) * ST_COUNT);
  return MFC_OK;
}

int mfc_set_stream(mfc_ctx* ctx, void* s) {
  NEEDCTX;
  Ctx& c = ctx->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  if (c.own_stream) HIPCHK(hipStreamDestroy(c.stream));
  c.stream = (hipStream_t)s;
  c.own_stream = false;
  return MFC_OK;
}

void* mfc_get_stream(mfc_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

// ---- transfers -------------------------------------------------------------
int mfc_upload_f(mfc_ctx* ctx, const double* f) {
  NEEDCTX;
  if (!f) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  const int ns = slab_planes(c);
  const size_t plane = (size_t)c.g.nx * c.g.ny * 19;
  int rc = ensure_staging(ctx, plane * ns * sizeof(double));
  if (rc) return rc;
  for (int z0 = 0; z0 < c.g.nz; z0 += ns) {
    const int nz = std::min(ns, c.g.nz - z0);
    HIPCHK(hipMemcpyAsync(c.staging, f + plane * z0, plane * nz * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIPCHK(launch_aos_to_pop(c, (const double*)c.staging, z0, nz, c.phase == PHASE_POST_STREAM));
    HIPCHK(hipStreamSynchronize(c.stream));
  }
  c.f_version++;
  return MFC_OK;
}

int mfc_download_f(mfc_ctx* ctx, double* f) {
  NEEDCTX;
  if (!f) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  const int ns = slab_planes(c);
  const size_t plane = (size_t)c.g.nx * c.g.ny * 19;
  int rc = ensure_staging(ctx, plane * ns * sizeof(double));
  if (rc) return rc;
  for (int z0 = 0; z0 < c.g.nz; z0 += ns) {
    const int nz = std::min(ns, c.g.nz - z0);
    HIPCHK(launch_pop_to_aos(c, (double*)c.staging, z0, nz, c.phase == PHASE_POST_STREAM));
    HIPCHK(hipMemcpyAsync(f + plane * z0, c.staging, plane * nz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
  }
  return MFC_OK;
}

int mfc_upload_macro(mfc_ctx* ctx, const double* rho, const double* ux, const double* uy, const double* uz) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.mom == MOM_VIRTUAL) {
    int rc = need_arrays(ctx);
    if (rc) return rc;
  }
  int rc = ensure_macro(ctx);
  if (rc) return rc;
  const size_t b = sizeof(double) * c.g.ncell;
  if (rho) HIPCHK(hipMemcpyAsync(c.rho, rho, b, hipMemcpyHostToDevice, c.stream));
  if (ux) HIPCHK(hipMemcpyAsync(c.ux, ux, b, hipMemcpyHostToDevice, c.stream));
  if (uy) HIPCHK(hipMemcpyAsync(c.uy, uy, b, hipMemcpyHostToDevice, c.stream));
  if (uz) HIPCHK(hipMemcpyAsync(c.uz, uz, b, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  c.mom = MOM_ARRAYS;
  return MFC_OK;
}

int mfc_download_macro(mfc_ctx* ctx, double* rho, double* ux, double* uy, double* uz) {
  NEEDCTX;
  Ctx& c = ctx->c;
  int rc = need_arrays(ctx);
  if (rc) return rc;
  const size_t b = sizeof(double) * c.g.ncell;
  if (rho) HIPCHK(hipMemcpyAsync(rho, c.rho, b, hipMemcpyDeviceToHost, c.stream));
  if (ux) HIPCHK(hipMemcpyAsync(ux, c.ux, b, hipMemcpyDeviceToHost, c.stream));
  if (uy) HIPCHK(hipMemcpyAsync(uy, c.uy, b, hipMemcpyDeviceToHost, c.stream));
  if (uz) HIPCHK(hipMemcpyAsync(uz, c.uz, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_download_force(mfc_ctx* ctx, double* fx, double* fy, double* fz) {
  NEEDCTX;
  if (!fx || !fy || !fz) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  const size_t b = sizeof(double) * c.g.ncell;
  int rc = ensure_staging(ctx, 3 * b);
  if (rc) return rc;
  double* s = (double*)c.staging;
  HIPCHK(launch_force_field(c, c.ft.cur, s, s + c.g.ncell, s + 2 * c.g.ncell));
  HIPCHK(hipMemcpyAsync(fx, s, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(fy, s + c.g.ncell, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(fz, s + 2 * c.g.ncell, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_get_forcing_state(mfc_ctx* ctx, int32_t* iseedf, int32_t* iyf, int32_t* ivf, double* b1r,
                          double* b2r, double* b3r, double* a1r, double* a2r, double* a3r) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (iseedf) *iseedf = c.iseedf;
  if (iyf) *iyf = c.iyf;
  if (ivf) memcpy(ivf, c.ivf, sizeof c.ivf);
  double* bs[3] = {b1r, b2r, b3r};
  double* as[3] = {a1r, a2r, a3r};
  for (int m = 0; m < 3; ++m) {
    if (bs[m]) memcpy(bs[m], c.bmod[m], sizeof(double) * 150);
    if (as[m]) memcpy(as[m], c.amod[m], sizeof(double) * 150);
  }
  return MFC_OK;
}

int mfc_set_forcing_state(mfc_ctx* ctx, int32_t iseedf, int32_t iyf, const int32_t* ivf, const double* b1r,
                          const double* b2r, const double* b3r, const double* a1r, const double* a2r,
                          const double* a3r) {
  NEEDCTX;
  Ctx& c = ctx->c;
  c.iseedf = iseedf;
  c.iyf = iyf;
  if (ivf) memcpy(c.ivf, ivf, sizeof c.ivf);
  const double* bs[3] = {b1r, b2r, b3r};
  const double* as[3] = {a1r, a2r, a3r};
  for (int m = 0; m < 3; ++m) {
    if (bs[m]) memcpy(c.bmod[m], bs[m], sizeof(double) * 150);
    if (as[m]) memcpy(c.amod[m], as[m], sizeof(double) * 150);
  }
  if (a1r || a2r || a3r) return publish_modes(ctx);
  return MFC_OK;
}

// ---- fluid step ----------------------------------------------------------------
int mfc_initpop(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  int rc = need_arrays(ctx);
  if (rc) return rc;
  HIPCHK(launch_initpop(c));
  c.f_version++;
  c.mom = MOM_ARRAYS;
  return MFC_OK;
}

int mfc_init_velocity(mfc_ctx* ctx, const double* amp, double scale) {
  NEEDCTX;
  if (!amp) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  int rc = ensure_macro(ctx);
  if (rc) return rc;
  const size_t nk = (size_t)init_mode_k(), nkx = (size_t)init_mode_kx();
  const size_t na = 3 * nkx * nk * nk;
  const size_t n1 = nk * nk * (size_t)c.g.nx, n2 = nk * (size_t)c.g.nx * (size_t)c.g.ny;
  double2 *d_amp = nullptr, *B1 = nullptr, *B2 = nullptr;
  hipError_t e = hipMalloc(&d_amp, sizeof(double2) * na);
  if (e == hipSuccess) e = hipMalloc(&B1, sizeof(double2) * n1);
  if (e == hipSuccess) e = hipMalloc(&B2, sizeof(double2) * n2);
  if (e == hipSuccess) e = hipMemcpyAsync(d_amp, amp, sizeof(double2) * na, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = launch_init_velocity(c, d_amp, scale, B1, B2);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (d_amp) (void)hipFree(d_amp);
  if (B1) (void)hipFree(B1);
  if (B2) (void)hipFree(B2);
  if (e != hipSuccess) return fail(ctx, MFC_ERR_HIP, "mfc_init_velocity", e);
  c.mom = MOM_ARRAYS;
  return MFC_OK;
}

int mfc_forcing(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (!(c.p.xkf > 0.0)) return MFC_OK;   // no forced shell: the field stays the body force
  forcing_modes_host(c);
  return publish_modes(ctx);
}

int mfc_collision_mrt(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.phase != PHASE_POST_STREAM) return fail(ctx, MFC_ERR_STATE, "collision_MRT called twice without streaming");
  bool arrays;
  if (c.mom == MOM_VIRTUAL && c.mom_version == c.f_version) {
    arrays = false;
  } else if (c.mom == MOM_ARRAYS) {
    arrays = true;
  } else if (c.mom == MOM_NONE) {
    int rc = ensure_macro(ctx);
    if (rc) return rc;
    c.mom = MOM_ARRAYS;
    arrays = true;
  } else {
    return fail(ctx, MFC_ERR_STATE, "collision_MRT: rho/u from macrovar no longer match the field");
  }
  if (c.timing && c.tcount++ % c.timing == 0 && c.tpool_used + 2 <= (int)c.tpool.size()) {
    HIPCHK(hipEventRecord(c.tpool[c.tpool_used], c.stream));
    HIPCHK(launch_collide(c, arrays));
    HIPCHK(hipEventRecord(c.tpool[c.tpool_used + 1], c.stream));
    c.tpool_used += 2;
    c.ttimed++;
  } else {
    HIPCHK(launch_collide(c, arrays));
  }
  if (c.g.lay == LAY_AB)
    c.cur ^= 1;
  else
    c.g.lay = c.g.lay == LAY_AA_ODD ? LAY_AA_EVEN : LAY_AA_ODD;
  c.phase = PHASE_PRE_STREAM;
  c.f_version++;
  return MFC_OK;
}

int mfc_streaming(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.phase != PHASE_PRE_STREAM) return fail(ctx, MFC_ERR_STATE, "streaming called twice without collision");
  c.phase = PHASE_POST_STREAM;
  c.f_version++;
  return MFC_OK;
}

int mfc_rhoupdat(mfc_ctx* ctx, double* maxdiff) {
  NEEDCTX;
  Ctx& c = ctx->c;
  int rc = need_arrays(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, c.stream, c.d_red, 0.0);
  HIPCHK(hipGetLastError());
  HIPCHK(launch_rhoupdat(c, c.phase == PHASE_POST_STREAM, c.d_red));
  c.mom = MOM_ARRAYS;
  if (maxdiff) {
    HIPCHK(hipMemcpyAsync(c.h_red, c.d_red, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    *maxdiff = c.h_red[0];
  }
  return MFC_OK;
}

int mfc_macrovar(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.phase != PHASE_POST_STREAM) return fail(ctx, MFC_ERR_STATE, "macrovar expects the streamed field");
  c.mom = MOM_VIRTUAL;
  c.mom_version = c.f_version;
  c.ft.mac = c.ft.cur;
  return MFC_OK;
}

// ---- particles -------------------------------------------------------------------
int mfc_prt_init(mfc_ctx* ctx, int32_t npart) {
  NEEDCTX;
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  if (npart < 0 || npart >= (1 << 24)) return MFC_ERR_ARG;   // contact-list entries hold 24-bit ids
  if (npart > 0) {
    const int need = 2 * (int)ceil(c.p.rad + 2.0) + 6;
    if (c.g.nx < need || c.g.nz < need || (!c.g.walls_y && c.g.ny < need))
      return fail(ctx, MFC_ERR_ARG, "box too small for the particle radius (need 2*ceil(rad+2)+6 nodes along x, z and periodic y)");
    if (!(c.p.rad > 0.0)) return fail(ctx, MFC_ERR_ARG, "particle radius must be positive");
    if ((size_t)segment_capacity(c.p.rad, 4) * sizeof(uint32_t) > 65536)
      return fail(ctx, MFC_ERR_ARG, "particle radius too large for the box-scan kernels (rad <= 27)");
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  if (npart > P.np_max) {
    double** pd[] = {&P.yp, &P.ypglb, &P.ypglbp, &P.wp, &P.wpp, &P.omgp, &P.omgpp, &P.fHIp, &P.flubp,
                     &P.torqp, &P.forcep, &P.forcepp, &P.torqpp, &P.dwdt, &P.domgdt, &P.thetap};
    for (double** a : pd) {
      if (*a) HIPCHK(hipFree(*a));
      *a = nullptr;
      HIPCHK(hipMalloc(a, sizeof(double) * 3 * npart));
    }
    P.slab = segment_planes(c.p.rad, npart);
    P.nseg = segments_per_particle(c.p.rad, P.slab);
    P.seg_cap = segment_capacity(c.p.rad, P.slab);
    const size_t nsegs = (size_t)npart * P.nseg;
    const size_t nodes = nsegs * (size_t)P.seg_cap;
    void** segbufs[] = {(void**)&P.lk_cnt, (void**)&P.lk_off, (void**)&P.seg_ft,
                        (void**)&P.sol_list[0], (void**)&P.sol_list[1], (void**)&P.sol_cnt[0],
                        (void**)&P.sol_cnt[1], (void**)&P.sh_node, (void**)&P.sh_hits,
                        (void**)&P.sh_cnt};
    for (void** a : segbufs) {
      if (*a) HIPCHK(hipFree(*a));
      *a = nullptr;
    }
    HIPCHK(hipMalloc(&P.lk_cnt, sizeof(int32_t) * nsegs));
    HIPCHK(hipMalloc(&P.lk_off, sizeof(int64_t) * (nsegs + 1)));
    if (P.lk_part) HIPCHK(hipFree(P.lk_part));
    HIPCHK(hipMalloc(&P.lk_part, sizeof(int64_t) * 2 * (nsegs / kScanChunk + 1)));
    for (int b = 0; b < 2; ++b) {
      HIPCHK(hipMalloc(&P.sol_list[b], sizeof(uint32_t) * nodes));
      HIPCHK(hipMalloc(&P.sol_cnt[b], sizeof(int32_t) * nsegs));
    }
    HIPCHK(hipMalloc(&P.sh_node, sizeof(uint32_t) * nodes));
    HIPCHK(hipMalloc(&P.sh_hits, sizeof(uint32_t) * nodes));
    HIPCHK(hipMalloc(&P.sh_cnt, sizeof(int32_t) * nsegs));
    HIPCHK(hipMalloc(&P.seg_ft, sizeof(double) * 6 * nsegs * kMaxBBSplit));
    // every refilled node was solid: at most the solid nodes of all spheres
    const double rs = c.p.rad + 1.0;
    P.fill_cap = (int64_t)ceil(4.0 / 3.0 * M_PI * rs * rs * rs) * npart + 64;
    if (P.fill_list) HIPCHK(hipFree(P.fill_list));
    if (P.fill_count) HIPCHK(hipFree(P.fill_count));
    HIPCHK(hipMalloc(&P.fill_list, sizeof(uint64_t) * P.fill_cap));
    HIPCHK(hipMalloc(&P.fill_count, sizeof(int32_t)));
    // link pool: every link's node lies in the shell r < d <= r + sqrt(2)
    // around its centre, and a node outside a convex sphere sees it along at
    // most 9 of the 18 directions (never along both c and -c).
    const double r = c.p.rad, h = sqrt(3.0) / 2.0;
    const double ro = r + sqrt(2.0) + h, ri = std::max(0.0, r - h);
    const double shell = 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri);
    const int64_t per = (int64_t)ceil(shell) * 9 + 64;
    const int64_t cap = per * npart;
    if (P.lk_node) HIPCHK(hipFree(P.lk_node));
    if (P.lk_dir) HIPCHK(hipFree(P.lk_dir));
    if (P.lk_alpha) HIPCHK(hipFree(P.lk_alpha));
    if (P.lk_out) HIPCHK(hipFree(P.lk_out));
    if (P.ct_list) HIPCHK(hipFree(P.ct_list));
    if (!P.ct_count) {
      HIPCHK(hipMalloc(&P.ct_count, sizeof(int32_t)));
      HIPCHK(hipMemsetAsync(P.ct_count, 0, sizeof(int32_t), c.stream));
    }
    HIPCHK(hipMalloc(&P.lk_node, sizeof(uint32_t) * cap));
    HIPCHK(hipMalloc(&P.lk_dir, sizeof(uint8_t) * cap));
    HIPCHK(hipMalloc(&P.lk_alpha, sizeof(double) * cap));
    HIPCHK(hipMalloc(&P.lk_out, sizeof(double2) * cap));
    HIPCHK(hipMalloc(&P.ct_list, sizeof(uint64_t) * cap));
    P.lk_cap = cap;
    P.np_max = npart;
  }
  if (npart > 0 && !P.isn[0]) {
    const size_t words = (size_t)((c.g.ncell + 31) / 32);
    for (int b = 0; b < 2; ++b) {
      HIPCHK(hipMalloc(&P.isn[b], sizeof(int32_t) * c.g.ncell));
      HIPCHK(hipMalloc(&P.mask[b], sizeof(uint32_t) * words));
    }
  }
  if (P.isn[0]) {
    const size_t words = (size_t)((c.g.ncell + 31) / 32);
    for (int b = 0; b < 2; ++b) {
      HIPCHK(hipMemsetAsync(P.isn[b], 0, sizeof(int32_t) * c.g.ncell, c.stream));
      HIPCHK(hipMemsetAsync(P.mask[b], 0, sizeof(uint32_t) * words, c.stream));
      P.map_np[b] = 0;
    }
  }
  P.map_cur = P.map_old = 0;
  P.np = npart;
  // pair-search cells for many particles: cells at least as wide as the
  // lubrication range 2r + mingap (which covers the overlap range 2r), three
  // or more along every axis (distinct neighbours, bounded cell width)
  if (P.cell_cnt) HIPCHK(hipFree(P.cell_cnt));
  if (P.cell_list) HIPCHK(hipFree(P.cell_list));
  P.cell_cnt = nullptr;
  P.cell_list = nullptr;
  P.cells_ok = false;
  if (npart > kCellMinParticles) {
    const double h = 2.0 * c.p.rad + std::max(c.p.mingap, 0.0) + 1e-3;
    const int ns[3] = {c.g.nx, c.g.ny, c.g.nz};
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      P.cell_n[a] = std::max(1, (int)floor((double)ns[a] / h));
      if (P.cell_n[a] < 3) ok = false;
    }
    if (ok) {
      const size_t ncell = (size_t)P.cell_n[0] * P.cell_n[1] * P.cell_n[2];
      HIPCHK(hipMalloc(&P.cell_cnt, sizeof(int32_t) * ncell));
      HIPCHK(hipMalloc(&P.cell_list, sizeof(int32_t) * ncell * kCellCap));
    }
  }
  if (npart > 0) {
    double* pd[] = {P.yp, P.ypglb, P.ypglbp, P.wp, P.wpp, P.omgp, P.omgpp, P.fHIp, P.flubp, P.torqp,
                    P.forcep, P.forcepp, P.torqpp, P.dwdt, P.domgdt, P.thetap};
    for (double* a : pd) HIPCHK(hipMemsetAsync(a, 0, sizeof(double) * 3 * npart, c.stream));
    HIPCHK(hipMemsetAsync(P.lk_off, 0, sizeof(int64_t) * ((size_t)npart * P.nseg + 1), c.stream));
    HIPCHK(hipMemsetAsync(P.fill_count, 0, sizeof(int32_t), c.stream));
  }
  HIPCHK(hipMemsetAsync(c.d_status, 0, sizeof(int32_t) * ST_COUNT, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

static double* pfield(Particles& P, int f) {
  switch (f) {
    case MFC_PF_YPGLB: return P.ypglb;
    case MFC_PF_YPGLBP: return P.ypglbp;
    case MFC_PF_WP: return P.wp;
    case MFC_PF_WPP: return P.wpp;
    case MFC_PF_OMGP: return P.omgp;
    case MFC_PF_OMGPP: return P.omgpp;
    case MFC_PF_FHIP: return P.fHIp;
    case MFC_PF_FLUBP: return P.flubp;
    case MFC_PF_TORQP: return P.torqp;
    case MFC_PF_FORCEPP: return P.forcepp;
    case MFC_PF_TORQPP: return P.torqpp;
    case MFC_PF_DWDT: return P.dwdt;
    case MFC_PF_DOMGDT: return P.domgdt;
    case MFC_PF_THETAP: return P.thetap;
    default: return nullptr;
  }
}

int mfc_prt_set(mfc_ctx* ctx, int32_t field, const double* a) {
  NEEDCTX;
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  double* d = pfield(P, field);
  if (!d || !a || P.np <= 0) return MFC_ERR_ARG;
  const size_t b = sizeof(double) * 3 * P.np;
  HIPCHK(hipMemcpyAsync(d, a, b, hipMemcpyHostToDevice, c.stream));
  if (field == MFC_PF_YPGLB) {
    HIPCHK(hipMemcpyAsync(P.yp, a, b, hipMemcpyHostToDevice, c.stream));
    P.cells_ok = false;
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_prt_get(mfc_ctx* ctx, int32_t field, double* a) {
  NEEDCTX;
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  double* d = pfield(P, field);
  if (!d || !a || P.np <= 0) return MFC_ERR_ARG;
  HIPCHK(hipMemcpyAsync(a, d, sizeof(double) * 3 * P.np, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

// initpartpos (partlib.f90:1-80): seeded rejection sampling on the host; only
// the flagged particles (lower id of each close pair) are redrawn.
int mfc_initpartpos(mfc_ctx* ctx, int32_t* iseedp, int32_t mingap_check) {
  NEEDCTX;
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  if (!iseedp || P.np <= 0) return MFC_ERR_ARG;
  const int np = P.np;
  std::vector<double> y(3 * (size_t)np);
  std::vector<int> flag(np, 1);
  int32_t idum = -*iseedp, iy = 0, iv[32] = {0};
  const int ns[3] = {c.g.nx, c.g.ny, c.g.nz};
  for (int round = 0;; ++round) {
    for (int p = 0; p < np; ++p)
      if (flag[p] > 0)
        for (int k = 0; k < 3; ++k) {
          if (k == 1 && c.g.walls_y) {
            // clear of the walls by rad + mingap_w (channel initpartpos, wall axis y)
            const double rg = c.p.rad + c.p.mingap_w;
            y[3 * p + 1] = rg + ((double)c.g.ny - 2.0 * rg) * ran1(&idum, &iy, iv);
          } else {
            y[3 * p + k] = (double)ns[k] * ran1(&idum, &iy, iv);
          }
        }
    int cnt = 0;
    for (int a = 0; a < np; ++a) {
      flag[a] = 0;
      if (a == np - 1) break;
      const double xc = y[3 * a], yc = y[3 * a + 1], zc = y[3 * a + 2];
      for (int b = a + 1; b < np; ++b) {
        const double xb = near_image(y[3 * b], xc, c.g.nx);
        const double yb = c.g.walls_y ? y[3 * b + 1] : near_image(y[3 * b + 1], yc, c.g.ny);
        const double zb = near_image(y[3 * b + 2], zc, c.g.nz);
        const double dist = sqrt((xc - xb) * (xc - xb) + (yc - yb) * (yc - yb) + (zc - zb) * (zc - zb));
        double hgap = dist - 2.0 * c.p.rad;
        if (mingap_check) hgap = hgap - c.p.mingap;
        if (hgap < 0.0) {
          flag[a] = 1;
          break;
        }
      }
      cnt += flag[a];
    }
    if (cnt == 0) break;
    if (round > 100000) return fail(ctx, MFC_ERR_CAPACITY, "initpartpos: no non-overlapping placement found");
  }
  *iseedp = idum;
  P.cells_ok = false;
  HIPCHK(hipMemcpyAsync(P.ypglb, y.data(), sizeof(double) * y.size(), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(P.yp, y.data(), sizeof(double) * y.size(), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_initpartvel(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  int rc = need_arrays(ctx);
  if (rc) return rc;
  HIPCHK(launch_initpartvel(c));
  return MFC_OK;
}

int mfc_initpartomg(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  HIPCHK(hipMemsetAsync(c.pt.omgp, 0, sizeof(double) * 3 * c.pt.np, c.stream));
  return MFC_OK;
}

int mfc_beads_links(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  if (P.np <= 0) return MFC_ERR_ARG;
  return links_now(ctx, 0, 0);
}

int mfc_beads_collision(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0 || c.pt.map_np[c.pt.map_cur] == 0) return fail(ctx, MFC_ERR_STATE, "beads_collision before beads_links");
  if (c.phase != PHASE_PRE_STREAM) return fail(ctx, MFC_ERR_STATE, "beads_collision expects the post-collision field");
  HIPCHK(launch_bounce_back(c));
  c.f_version++;
  return MFC_OK;
}

int mfc_beads_lubforce(mfc_ctx* ctx) {
  NEEDCTX;
  if (ctx->c.pt.np <= 0) return MFC_ERR_ARG;
  HIPCHK(launch_lubforce(ctx->c));
  return MFC_OK;
}

int mfc_beads_move(mfc_ctx* ctx, int32_t istep) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  HIPCHK(launch_move(c, istep == c.p.irelease ? 1 : 0, 0));
  c.pt.map_old = c.pt.map_cur;   // ibnodes0 = ibnodes (partlib.f90:2376-2377)
  return MFC_OK;
}

int mfc_beads_redistribute(mfc_ctx* ctx, int32_t istep) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  HIPCHK(launch_redistribute(c, istep));
  return check_faults(ctx);
}

int mfc_beads_filling(mfc_ctx* ctx) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  if (c.phase != PHASE_POST_STREAM) return fail(ctx, MFC_ERR_STATE, "beads_filling expects the streamed field");
  HIPCHK(launch_filling(c));
  c.f_version++;
  return MFC_OK;
}

int mfc_get_maps(mfc_ctx* ctx, int32_t which, int32_t* ib, int32_t* is) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (!ib || !is) return MFC_ERR_ARG;
  const size_t b = sizeof(int32_t) * c.g.ncell;
  int rc = ensure_staging(ctx, 2 * b);
  if (rc) return rc;
  int32_t* s = (int32_t*)c.staging;
  HIPCHK(launch_maps_to_host_layout(c, which, s, s + c.g.ncell));
  HIPCHK(hipMemcpyAsync(ib, s, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(is, s + c.g.ncell, b, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_get_links(mfc_ctx* ctx, int32_t* ilink, int32_t* mlink, double* alink) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (!ilink || !mlink || !alink) return MFC_ERR_ARG;
  if (c.g.ncell > (1u << 21)) return fail(ctx, MFC_ERR_ARG, "mfc_get_links: dense link arrays limited to 2^21 nodes");
  if (c.pt.np <= 0) return MFC_ERR_ARG;
  const uint64_t n18 = 18ull * c.g.ncell;
  int rc = ensure_staging(ctx, n18 * (4 + 4 + 8));
  if (rc) return rc;
  int32_t* il = (int32_t*)c.staging;
  int32_t* ml = il + n18;
  double* al = (double*)(ml + n18);
  const unsigned nb = (unsigned)((n18 + 255) / 256);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, c.stream, il, n18, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, c.stream, ml, n18, -1);
  hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, c.stream, al, n18, 0.0);
  HIPCHK(hipGetLastError());
  HIPCHK(launch_links_dense(c, il, ml, al));
  HIPCHK(hipMemcpyAsync(ilink, il, n18 * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(mlink, ml, n18 * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(alink, al, n18 * 8, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

// ---- whole step ---------------------------------------------------------------
}  // extern "C"

namespace {

// The particle stream, its events and the tile lists of the overlapped step.
int ensure_overlap(mfc_ctx* ctx) {
  Ctx& c = ctx->c;
  if (c.pstream) return MFC_OK;
  TileLists& T = c.tiles;
  T.tpr = (c.g.nx + kTileX - 1) / kTileX;
  T.ntiles = (uint64_t)T.tpr * (uint64_t)c.g.ny * (uint64_t)c.g.nz;
  // the particle phase's own reach (stencils 3 sqrt 2 beyond the sphere,
  // 2 nodes of displacement) and the collide's pull (sqrt 2), with a margin
  T.reach = c.p.rad + 4.0 * sqrt(2.0) + 2.0 + 1e-6;
  const size_t nchunk = (size_t)((T.ntiles + kTileChunk - 1) / kTileChunk);
  if (nchunk > (size_t)1024 * 1024) return fail(ctx, MFC_ERR_ARG, "overlapped step: lattice too large for the tile lists");
  HIPCHK(hipMalloc(&T.mark, T.ntiles));
  HIPCHK(hipMemsetAsync(T.mark, 0, T.ntiles, c.stream));
  HIPCHK(hipMalloc(&T.cnt, sizeof(int32_t) * nchunk));
  HIPCHK(hipMalloc(&T.off_near, sizeof(uint32_t) * nchunk));
  HIPCHK(hipMalloc(&T.off_far, sizeof(uint32_t) * nchunk));
  HIPCHK(hipMalloc(&T.tot, sizeof(uint32_t) * 4));
  HIPCHK(hipHostMalloc((void**)&T.h_tot, sizeof(uint32_t) * 4, hipHostMallocDefault));
  HIPCHK(hipMalloc(&T.near_list, sizeof(uint32_t) * 2 * T.ntiles));
  HIPCHK(hipMalloc(&T.far_list, sizeof(uint32_t) * 2 * T.ntiles));
  HIPCHK(hipStreamSynchronize(c.stream));
  HIPCHK(hipStreamCreateWithFlags(&c.pstream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c.ev_coll, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c.ev_tiles, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c.ev_part, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c.ev_bb, hipEventDisableTiming));
  return MFC_OK;
}

// The collide of the far or the near tiles of list set k (their number read
// back into T.h_tot), timed like mfc_collision_mrt's launch when timing is on.
int collide_tiles(mfc_ctx* ctx, int k, bool far, bool timed) {
  Ctx& c = ctx->c;
  TileLists& T = c.tiles;
  const uint32_t* list = (far ? T.far_list : T.near_list) + (size_t)k * T.ntiles;
  const uint32_t n = T.h_tot[2 * k + (far ? 1 : 0)];
  if (n > T.ntiles) return fail(ctx, MFC_ERR_STATE, "overlapped step: tile count out of range");
  if (timed) HIPCHK(hipEventRecord(c.tpool[c.tpool_used], c.stream));
  HIPCHK(launch_collide_tiles(c, list, n, far));
  if (timed) {
    HIPCHK(hipEventRecord(c.tpool[c.tpool_used + 1], c.stream));
    c.tpool_used += 2;
  }
  return MFC_OK;
}

// mfc_step with the particle phase of step n on the particle stream, beside
// the collide of step n + 1 on the tiles far from every sphere; the near
// tiles follow once that phase is done.  Every kernel computes exactly what
// it computes in the plain sequence (the tiles only split the collide's
// grid), so results are bitwise those of the plain step.
int step_overlapped(mfc_ctx* ctx, int32_t istep, int32_t nsteps) {
  Ctx& c = ctx->c;
  Particles& P = c.pt;
  int rc = ensure_overlap(ctx);
  if (rc) return rc;
  c.overlap_used = true;
  const hipStream_t sA = c.stream, sB = c.pstream;
  TileLists& T = c.tiles;
  // step istep: its FORCING and the collide of the whole lattice
  if ((rc = mfc_forcing(ctx))) return rc;
  if ((rc = mfc_collision_mrt(ctx))) return rc;
  for (int s = 0; s < nsteps; ++s) {
    const int it = istep + s;
    const bool more = s + 1 < nsteps;
    const int k = (it + 1) & 1;   // list set of step it + 1's collide
    HIPCHK(hipEventRecord(c.ev_coll, sA));
    c.stream = sB;
    int prc = MFC_OK;
    hipError_t he = hipSuccess;
    do {
      if (more) {
        // tiles of step it + 1 from the positions this particle phase starts from
        TileLists Tk = T;
        Tk.tot = T.tot + 2 * k;
        Tk.near_list = T.near_list + (size_t)k * T.ntiles;
        Tk.far_list = T.far_list + (size_t)k * T.ntiles;
        if ((he = launch_tiles(c, Tk)) != hipSuccess) break;
        if ((he = hipMemcpyAsync(T.h_tot + 2 * k, Tk.tot, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, sB)) !=
            hipSuccess)
          break;
        if ((he = hipEventRecord(c.ev_tiles, sB)) != hipSuccess) break;
      }
      if ((he = hipStreamWaitEvent(sB, c.ev_coll, 0)) != hipSuccess) break;
      // beads_collision with beads_lubforce and the map clearing in its grid
      if ((he = launch_bounce_back(c, true)) != hipSuccess) break;
      if ((he = hipEventRecord(c.ev_bb, sB)) != hipSuccess) break;
      c.f_version++;
      if ((prc = mfc_streaming(ctx))) break;
      if ((he = launch_move(c, it == c.p.irelease ? 1 : 0, 1, true)) != hipSuccess) break;
      P.map_old = P.map_cur;
      if ((prc = links_now(ctx, 1, it, true))) break;
      if ((he = launch_filling(c, false, nullptr)) != hipSuccess) break;
      c.f_version++;
      if ((prc = mfc_macrovar(ctx))) break;
      he = hipEventRecord(c.ev_part, sB);
    } while (false);
    c.stream = sA;
    if (he != hipSuccess) return fail(ctx, MFC_ERR_HIP, "overlapped particle phase", he);
    if (prc) return prc;
    if (!more) {
      HIPCHK(hipStreamWaitEvent(sA, c.ev_part, 0));
      break;
    }
    // step it + 1: FORCING, then the collide of the far tiles beside the
    // particle phase, then of the near tiles after it
    if ((rc = mfc_forcing(ctx))) return rc;
    if (c.phase != PHASE_POST_STREAM || c.mom != MOM_VIRTUAL || c.mom_version != c.f_version)
      return fail(ctx, MFC_ERR_STATE, "overlapped step: moments not from the last macrovar");
    const bool timed = c.timing && c.tcount++ % c.timing == 0 && c.tpool_used + 4 <= (int)c.tpool.size();
    // the list lengths size the grids; the tiles finish early in the near
    // collide of step it, which stream A still runs while the host waits
    HIPCHK(hipEventSynchronize(c.ev_tiles));
    HIPCHK(hipStreamWaitEvent(sA, c.ev_tiles, 0));
    if (c.overlap == 2) HIPCHK(hipStreamWaitEvent(sA, c.ev_bb, 0));
    if ((rc = collide_tiles(ctx, k, true, timed))) return rc;
    HIPCHK(hipStreamWaitEvent(sA, c.ev_part, 0));
    if ((rc = collide_tiles(ctx, k, false, timed))) return rc;
    if (timed) c.ttimed++;
    if (c.g.lay == LAY_AB)
      c.cur ^= 1;
    else
      c.g.lay = c.g.lay == LAY_AA_ODD ? LAY_AA_EVEN : LAY_AA_ODD;
    c.phase = PHASE_PRE_STREAM;
    c.f_version++;
  }
  return MFC_OK;
}

}  // namespace

extern "C" {

int mfc_step(mfc_ctx* ctx, int32_t istep, int32_t nsteps) {
  NEEDCTX;
  Ctx& c = ctx->c;
  int rc;
  Particles& P = c.pt;
  if (nsteps > 0 && c.overlap && particles_on(c) && istep >= c.p.irelease && !c.ens &&
      P.map_np[P.map_cur] > 0 && c.phase == PHASE_POST_STREAM)
    return step_overlapped(ctx, istep, nsteps);
  // FORCING of step s + 1 drawn during step s (spheres on, more steps in
  // this call): its table is evaluated by extra blocks of step s's refill
  // grid and made current after step s's macrovar, which keeps step s's
  // table.  The draw is the same call the next step would make first, so the
  // forcing state after any return is the one of the routine-by-routine
  // sequence.
  bool drawn = false;
  int drawn_rc = MFC_OK;
  TableJob job;
  for (int s = 0; s < nsteps; ++s) {
    const int it = istep + s;
    const bool on = particles_on(c) && it >= c.p.irelease;
    if (drawn_rc) return drawn_rc;
    if (!drawn && (rc = mfc_forcing(ctx))) return rc;
    drawn = false;
    if ((rc = mfc_collision_mrt(ctx))) return rc;
    if (on) {
      if (P.map_np[P.map_cur] == 0) return fail(ctx, MFC_ERR_STATE, "time step with particles before beads_links");
      // bounce-back with beads_lubforce in the same grid, the scatter with the
      // clearing of the map buffer this step's links will rebuild
      HIPCHK(launch_bounce_back(c, true));
      c.f_version++;
    }
    if ((rc = mfc_streaming(ctx))) return rc;
    if (on) {
      // move with the periodic wrap of beads_redistribute fused in; links
      // with its overlap check fused in (reported by mfc_sync)
      HIPCHK(launch_move(c, it == c.p.irelease ? 1 : 0, 1, true));
      P.map_old = P.map_cur;
      // links with the refill search in the expansion grid, then the refill
      if ((rc = links_now(ctx, 1, it, true))) return rc;
      if (s + 1 < nsteps && c.p.xkf > 0.0) {
        forcing_modes_host(c);
        drawn = true;
        drawn_rc = prepare_modes(ctx, job);
      }
      HIPCHK(launch_filling(c, false, drawn && !drawn_rc && job.nblk > 0 ? &job : nullptr));
      c.f_version++;
    }
    if ((rc = mfc_macrovar(ctx))) return rc;
    if (drawn && !drawn_rc) make_current(c, job);
  }
  return MFC_OK;
}

// ---- diagnostics -------------------------------------------------------------------
static int ensure_diag(mfc_ctx* ctx, int nek) {
  Ctx& c = ctx->c;
  const int nb = diag_blocks();
  const int need = nek > 6 ? nek : 6;
  if (c.diag_part && need > c.diag_nek) {
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipFree(c.diag_part));
    HIPCHK(hipFree(c.diag_out));
    c.diag_part = nullptr;
    c.diag_out = nullptr;
  }
  if (!c.diag_part) {
    HIPCHK(hipMalloc(&c.diag_part, sizeof(double) * (size_t)nb * need));
    HIPCHK(hipMalloc(&c.diag_out, sizeof(double) * need));
    c.diag_nek = need;
  }
  return MFC_OK;
}

int mfc_diag_sums(mfc_ctx* ctx, double* out6) {
  NEEDCTX;
  if (!out6) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  int rc = need_arrays(ctx);
  if (rc) return rc;
  if ((rc = ensure_diag(ctx, 6))) return rc;
  HIPCHK(launch_field_sums(c, c.diag_part, c.diag_out));
  HIPCHK(hipMemcpyAsync(out6, c.diag_out, sizeof(double) * 6, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_diag_sums_noalloc(mfc_ctx* ctx, double* out6) {
  NEEDCTX;
  if (!out6) return MFC_ERR_ARG;
  Ctx& c = ctx->c;
  if (!c.rho) return fail(ctx, MFC_ERR_STATE, "diag_sums_noalloc: the rho/u arrays do not exist yet "
                                               "(a call that returns them allocates them first)");
  if (!c.sums_scratch) return fail(ctx, MFC_ERR_STATE, "diag_sums_noalloc: device scratch not resolved");
  const size_t need = 6 * (size_t)diag_blocks() + 6;
  if (need > sums_scratch_doubles())
    return fail(ctx, MFC_ERR_CAPACITY, "diag_sums_noalloc: the partial sums exceed the static scratch");
  if (c.mom != MOM_ARRAYS) {
    if (c.mom != MOM_VIRTUAL || c.mom_version != c.f_version)
      return fail(ctx, MFC_ERR_STATE, "rho/u requested after the field changed since macrovar");
    if (c.arrays_version != c.mom_version) {
      HIPCHK(launch_macro(c, c.phase == PHASE_POST_STREAM));
      c.arrays_version = c.mom_version;
    }
  }
  double* part = c.sums_scratch;
  double* res = part + 6 * (size_t)diag_blocks();
  HIPCHK(launch_field_sums(c, part, res));
  HIPCHK(hipMemcpyAsync(out6, res, sizeof(double) * 6, hipMemcpyDeviceToDevice, c.stream));
  return MFC_OK;
}

int mfc_diag_spectrum(mfc_ctx* ctx, int32_t nek, double* ek) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (!ek || nek < 1 || nek > 4096) return MFC_ERR_ARG;
  const int ns[3] = {c.g.nx, c.g.ny, c.g.nz};
  for (int a = 0; a < 3; ++a)
    if (ns[a] > 2048) return fail(ctx, MFC_ERR_ARG, "diag_spectrum: an axis longer than 2048 nodes");
  int rc = need_arrays(ctx);
  if (rc) return rc;
  if ((rc = ensure_diag(ctx, nek))) return rc;
  if (!c.diag_Z) {
    HIPCHK(hipMalloc(&c.diag_Z, sizeof(double2) * c.g.ncell));
    for (int a = 0; a < 3; ++a) {
      std::vector<double2> tw(ns[a]);
      for (int j = 0; j < ns[a]; ++j) {
        const double ang = -c.p.pi2 * (double)j / (double)ns[a];
        tw[j] = make_double2(cos(ang), sin(ang));
      }
      HIPCHK(hipMalloc(&c.diag_tw[a], sizeof(double2) * ns[a]));
      HIPCHK(hipMemcpy(c.diag_tw[a], tw.data(), sizeof(double2) * ns[a], hipMemcpyHostToDevice));
    }
  }
  const double2* tw[3] = {c.diag_tw[0], c.diag_tw[1], c.diag_tw[2]};
  HIPCHK(launch_spectrum(c, nek, c.diag_Z, tw, c.diag_part, c.diag_out));
  HIPCHK(hipMemcpyAsync(ek, c.diag_out, sizeof(double) * nek, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return MFC_OK;
}

int mfc_set_bb_order(mfc_ctx* ctx, int32_t order) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (order != MFC_BB_SNAPSHOT && order != MFC_BB_REFERENCE)
    return fail(ctx, MFC_ERR_ARG, "bounce-back order must be MFC_BB_SNAPSHOT or MFC_BB_REFERENCE");
  if (c.ens) return fail(ctx, MFC_ERR_STATE, "box is an ensemble member: remove it first");
  Particles& P = c.pt;
  if (order == MFC_BB_REFERENCE && !P.ord) {
    HIPCHK(hipMalloc(&P.ord, sizeof(OrderEntry) * kOrderCap));
    P.ord_cap = kOrderCap;
  }
  P.bb_order = order;
  return MFC_OK;
}

int mfc_set_timing(mfc_ctx* ctx, int32_t on) {
  NEEDCTX;
  Ctx& c = ctx->c;
  if (on && c.tpool.empty()) {
    c.tpool.resize(kTimingPool);
    for (auto& e : c.tpool) HIPCHK(hipEventCreate(&e));
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  c.tpool_used = 0;
  c.tcount = 0;
  c.ttimed = 0;
  c.timing = on > 0 ? on : 0;
  return MFC_OK;
}

int mfc_get_timing(mfc_ctx* ctx, double* total_ms, int64_t* launches) {
  NEEDCTX;
  Ctx& c = ctx->c;
  HIPCHK(hipStreamSynchronize(c.stream));
  double t = 0.0;
  for (int i = 0; i + 1 < c.tpool_used; i += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c.tpool[i], c.tpool[i + 1]));
    t += ms;
  }
  if (total_ms) *total_ms = t;
  if (launches) *launches = c.ttimed;
  return MFC_OK;
}

}  // extern "C"
