// internal.h -- device-state context and kernel launchers of libmfc_lbm.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/mfc_lbm.h"
#include "lbm_common.h"

namespace mfc {

enum Phase { PHASE_PRE_STREAM = 0, PHASE_POST_STREAM = 1 };
enum MomState { MOM_NONE = 0, MOM_ARRAYS = 1, MOM_VIRTUAL = 2 };

// device status words (int32), read back by mfc_sync / mfc_status.  Counters
// accumulate from mfc_prt_init on; ST_NLINKS / ST_NSOLID describe the last
// beads_links.
enum StatusWord {
  ST_OVERLAP = 0,        // overlap flags raised (fatal in the reference)
  ST_OVERLAP_STEP = 1,   // first step number at which an overlap was seen (0: none)
  ST_LINK_OVERFLOW = 2,  // link builds that did not fit the link pool
  ST_ALPHA_FAULT = 3,    // wall fractions outside [0,1] (reported, not fatal)
  ST_CONTACT_LINKS = 4,  // boundary links whose own node is solid (near contact)
  ST_NONFINITE = 5,      // particles with non-finite or out-of-box state
  ST_REFILL = 6,         // nodes refilled
  ST_NLINKS = 7,         // links built by the last beads_links
  ST_NSOLID = 8,         // solid nodes marked by the last beads_links
  ST_BOX_FAULT = 9,      // particle boxes that did not fit the lattice
  ST_CELL_OVERFLOW = 10, // pair-search cells or neighbour lists over capacity
  ST_FAST = 11,          // sphere steps that moved a centre a node or more
  ST_ORDER_CAP = 12,     // contact links beyond the reference-order table (not evaluated)
  ST_COUNT = 16
};

constexpr int kForceSlots = 4;
constexpr int kMaxDevices = 64;   // device ordinals the per-device tables cover
// particle pair searches use a cell grid above this many particles; a cell
// (between 1 and 4/3 times 2r + mingap wide) holds at most kCellCap centres
// (non-overlapping spheres never come close to it; more is reported)
constexpr int kCellMinParticles = 512;
// segment-count scans of more than kScanChunk segments run over several
// workgroups, kScanChunk segments each
constexpr int kScanChunk = 4096;
// a segment's links are split over up to this many bounce-back workgroups
constexpr int kMaxBBSplit = 4;
constexpr int kCellCap = 16;
// overlapped step: tiles of kTileX consecutive x nodes of a row, classified
// near/far in chunks of kTileChunk tiles
constexpr int kTileX = 64;
constexpr int kTileChunk = 4096;

// Near and far tile lists of the overlapped step (device), rebuilt every step
// from the sphere positions at the start of its particle phase.
struct TileLists {
  int tpr = 0;                 // tiles per row
  uint64_t ntiles = 0;
  double reach = 0.0;          // a tile within this distance of a centre is near
  uint8_t* mark = nullptr;     // per tile: 1 near (cleared again by the compaction)
  int32_t* cnt = nullptr;      // near tiles per chunk
  uint32_t* off_near = nullptr;
  uint32_t* off_far = nullptr;
  uint32_t* tot = nullptr;     // [0] near tiles, [1] far tiles (two list sets)
  uint32_t* h_tot = nullptr;   // pinned host copy of tot
  uint32_t* near_list = nullptr;
  uint32_t* far_list = nullptr;
};
// The live forcing modes of one FORCING call, grouped by kx (n[0] modes with
// kx = 0 first, then kx = 1, then kx = 2), each in (kz, ky) order, with
// (re, im) of the three components.  Handed to the table kernel by value.
constexpr int kMaxModes = 52;
struct ModeSet {
  int n[3];
  int8_t ky[kMaxModes], kz[kMaxModes];
  double a[kMaxModes][6];
};

struct ForceTables {
  double* g[kForceSlots] = {};                 // per (y,z) row: 3 comps x 3 kx x (re,im)
  int live[kForceSlots] = {};                  // 1: table holds nonzero modes
  int cur = 0;                                 // modes of the last FORCING
  int mac = 1;                                 // modes captured by the last macrovar
  int prev = 2;                                // modes of the FORCING before the last
  double* phx = nullptr;                       // (cos,sin)(2 pi m / nx)
  double* phy = nullptr;
  double* phz = nullptr;
};

// One contact link of the reference-order bounce-back.  dep[0] is the link
// that writes f_q at this link's node x (from x + c_q) and dep[1] the one that
// writes f_qbar there (from x - c_q, read only when alpha > 1/2), each kept
// only if the reference's loop runs it earlier and writes x in place: its
// link index, with kOrderContact set when that link is itself a contact link
// (its value is then resolved alongside), or -1.  val holds the value a
// non-contact writer stores.
constexpr int64_t kOrderContact = int64_t(1) << 62;
struct OrderEntry {
  int64_t dep[2];
  double val[2];
  double f9, ff1;   // this link's next values during resolution
};

struct Particles {
  int np = 0, np_max = 0;
  // (3, np) column-major, c + 3*p
  double *yp = nullptr, *ypglb = nullptr, *ypglbp = nullptr;
  double *wp = nullptr, *wpp = nullptr, *omgp = nullptr, *omgpp = nullptr;
  double *fHIp = nullptr, *flubp = nullptr, *torqp = nullptr;
  double *forcep = nullptr, *forcepp = nullptr, *torqpp = nullptr;
  double *dwdt = nullptr, *domgdt = nullptr, *thetap = nullptr;
  // node maps, double-buffered: isn = particle id (1-based) or 0, mask = solid bit
  int32_t* isn[2] = {nullptr, nullptr};
  uint32_t* mask[2] = {nullptr, nullptr};
  int map_np[2] = {0, 0};                    // particle count of that build (0: map empty)
  int map_cur = 0, map_old = 0;
  // Box segments: one z-slab of a particle's bounding box, the work unit of
  // the box kernels (segment s of particle p is p * nseg + s).  seg_cap bounds
  // the nodes of one segment, so per-segment lists never overflow.
  int nseg = 1;
  int seg_cap = 0;
  int slab = 4;   // z planes per segment, chosen with the allocation (segment_planes)
  // solid nodes each map build marked, per segment: what the build two steps
  // later clears and what beads_filling scans for uncovered nodes
  uint32_t* sol_list[2] = {nullptr, nullptr};
  int32_t* sol_cnt[2] = {nullptr, nullptr};
  // fluid nodes within r + 1.5 of a centre and the link directions they see
  // (bit q set: direction q crosses the surface), per segment
  uint32_t* sh_node = nullptr;
  uint32_t* sh_hits = nullptr;
  int32_t* sh_cnt = nullptr;
  // boundary links, grouped by segment
  int64_t lk_cap = 0;
  uint32_t* lk_node = nullptr;
  uint8_t* lk_dir = nullptr;
  double* lk_alpha = nullptr;
  double2* lk_out = nullptr;   // (f9, ff1) of contact links, taken before the scatter
  uint64_t* ct_list = nullptr; // contact links of the last build: (particle << 40) | link
  int32_t* ct_count = nullptr;
  // bounce-back order (MFC_BB_SNAPSHOT or MFC_BB_REFERENCE) and, for the
  // reference order, the per-contact-link table of its evaluation
  int bb_order = 0;
  OrderEntry* ord = nullptr;
  int64_t ord_cap = 0;
  int32_t* lk_cnt = nullptr;   // links per segment
  int64_t* lk_part = nullptr;  // per-chunk sums of the segment counts (multi-block scan)
  int64_t* lk_off = nullptr;   // link offset per segment (np * nseg + 1)
  double* seg_ft = nullptr;    // force and torque per bounce-back unit (6 each)
  // bounce-back units per segment (1..kMaxBBSplit, MFC_BB_SPLIT); one
  // measured best on MI355X at 256^3 and 1024^3 (more units: more blocks,
  // no better cache reuse)
  int bb_split = 1;
  // cell grid of the pair searches (on for many particles): per-cell counts
  // and particle lists, valid for the positions at the last build
  int cell_n[3] = {0, 0, 0};
  int32_t* cell_cnt = nullptr;
  int32_t* cell_list = nullptr;
  bool cells_ok = false;
  // nodes to refill, found by beads_filling's first pass
  uint64_t* fill_list = nullptr;
  int32_t* fill_count = nullptr;
  int64_t fill_cap = 0;
};

struct Ctx {
  mfc_params p{};
  Grid g{};
  CollideConsts cc{};
  int prec = 0;                 // MFC_PREC_*: FP64, FP32, or FP32 storage with FP64 collision
  size_t tsize = 8;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* pop[2] = {nullptr, nullptr};
  int cur = 0;
  int phase = PHASE_POST_STREAM;
  uint64_t f_version = 1;
  // macroscopic arrays (reference rho/ux/uy/uz), FP64
  double *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr;
  int mom = MOM_NONE;
  uint64_t mom_version = 0;
  uint64_t arrays_version = 0;      // MOM_VIRTUAL: the macrovar whose rho/u the arrays hold (0: none)
  ForceTables ft;
  // host forcing state (gaussian.f90 FORCING)
  int32_t iseedf = 0, iyf = 0, ivf[32] = {0};
  double amod[3][150] = {{0}}, bmod[3][150] = {{0}};
  Particles pt;
  int32_t* d_status = nullptr;
  int32_t* h_status = nullptr;      // pinned
  double* d_red = nullptr;          // small reduction buffer
  double* h_red = nullptr;          // pinned
  void* staging = nullptr;          // device staging for host transfers
  size_t staging_bytes = 0;
  std::string err;
  // optional timing of the collide kernel: event pairs recorded around every
  // `timing`-th launch (0: off), summed by mfc_get_timing
  int timing = 0;
  int64_t tcount = 0;
  int64_t ttimed = 0;               // collides timed (an overlapped collide is two launches)
  bool overlap_used = false;        // an overlapped step ran (its displacement bound is checked)
  std::vector<hipEvent_t> tpool;
  int tpool_used = 0;
  // diagnostics scratch, allocated on first use
  double* diag_part = nullptr;      // per-block partials
  double* diag_out = nullptr;       // results
  int diag_nek = 0;                 // shells the partials were sized for
  double2* diag_Z = nullptr;        // complex FP64 field
  double2* diag_tw[3] = {nullptr, nullptr, nullptr};   // exp(-2 pi i j / n) per axis
  void* ens = nullptr;              // the ensemble this box is a member of (its stream is on loan)
  int xcd_remap = -1;               // collide tile order over the XCDs (-1: by shape; MFC_XCD_REMAP=0/1/2)
  double* sums_scratch = nullptr;   // this device's static scratch of mfc_diag_sums_noalloc
  // overlapped step (mfc_step with spheres): the particle phase of step n runs
  // on pstream while the collide of step n + 1 covers the tiles far from every
  // sphere; opt-in with MFC_OVERLAP=1 (beside the whole phase) or 2 (beside
  // the phase after the bounce-back)
  int overlap = 0;
  hipStream_t pstream = nullptr;
  hipEvent_t ev_coll = nullptr, ev_tiles = nullptr, ev_part = nullptr, ev_bb = nullptr;
  TileLists tiles;
};

// ---- launchers (fluid.hip) ----
hipError_t launch_force_table(Ctx& c, int slot, const ModeSet& m);
hipError_t launch_collide(Ctx& c, bool arrays);
// the collide (moments recomputed, never the arrays) of the n tiles in
// `list`; with far set the solid mask is not read
hipError_t launch_collide_tiles(Ctx& c, const uint32_t* list, uint32_t n, bool far);
hipError_t launch_macro(Ctx& c, bool pulled);
hipError_t launch_rhoupdat(Ctx& c, bool pulled, double* d_maxdiff);
hipError_t launch_initpop(Ctx& c);
hipError_t launch_pop_to_aos(Ctx& c, double* dst, int z0, int nzs, bool pulled);
hipError_t launch_aos_to_pop(Ctx& c, const double* src, int z0, int nzs, bool pulled);
hipError_t launch_force_field(Ctx& c, int slot, double* fx, double* fy, double* fz);

// ---- launchers (particles.hip) ----
int segments_per_particle(double rad, int slab);
int segment_capacity(double rad, int slab);
int segment_planes(double rad, int npart);
// with_find: the first pass of beads_filling (previous map vs the new one)
// runs in the same grid as the link expansion
hipError_t launch_links(Ctx& c, int target, int overlap_check, int istep, bool with_find = false);
// with `fused` set (mfc_step) the bounce-back grid also computes the
// lubrication forces and clears the next map target, and fHIp / torqp are
// left for the move kernel to sum from the segments
hipError_t launch_bounce_back(Ctx& c, bool fused = false);
hipError_t launch_lubforce(Ctx& c);
hipError_t launch_map_clear(Ctx& c, int buf);
hipError_t launch_move(Ctx& c, int first, int wrap, bool from_seg = false);
hipError_t launch_redistribute(Ctx& c, int istep);
struct TableJob;
// job: a forcing table evaluated by extra blocks of the refill grid (or by
// its own launch when there is nothing to refill)
hipError_t launch_filling(Ctx& c, bool find = true, const TableJob* job = nullptr);
hipError_t launch_initpartvel(Ctx& c);
// (re)build the pair-search cells if they are on and stale
hipError_t ensure_cells(Ctx& c);
hipError_t launch_maps_to_host_layout(Ctx& c, int which, int32_t* ib, int32_t* is);
// near/far tile lists from the current positions (c.tiles allocated)
hipError_t launch_tiles(Ctx& c, TileLists& T);
hipError_t launch_links_dense(Ctx& c, int32_t* ilink, int32_t* mlink, double* alink);

// ---- launchers (initial.hip) ----
int init_mode_kx();   // kx values of the mode table (0..16)
int init_mode_k();    // ky, kz values (-16..16)
hipError_t launch_init_velocity(Ctx& c, const double2* amp, double scale, double2* B1, double2* B2);

// ---- launchers (diag.hip) ----
int diag_blocks();
hipError_t launch_field_sums(Ctx& c, double* part, double* out);
// the static device scratch of the allocation-free sums: its size in
// doubles and its address on the current device
size_t sums_scratch_doubles();
hipError_t sums_scratch(double** out);
hipError_t launch_spectrum(Ctx& c, int nek, double2* Z, const double2* const tw[3], double* part,
                           double* acc);

}  // namespace mfc

// the opaque handle of the C ABI
struct mfc_ctx {
  mfc::Ctx c;
};

