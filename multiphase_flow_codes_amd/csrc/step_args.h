// step_args.h -- argument blocks of the time-step kernels.  One block per
// box: passed by value for a single box, or gathered into per-case tables in
// device memory when one grid advances several boxes (ensemble engine).
#pragma once
#include "internal.h"

namespace mfc {

struct ForceArgs {
  const double* gcur;   // forcing table of the current step (may be null)
  const double* gmac;   // forcing table captured by the last macrovar (may be null)
  const double* phx;    // (cos,sin)(2 pi m / nx)
  double bfx, bfy, bfz; // uniform body force
};

// Arguments of the collide kernel (fluid.hip).
struct CollideArgs {
  const void* A;
  void* B;
  const uint32_t* mask;
  Grid g;
  CollideConsts k;
  ForceArgs fa;
  const double *rho, *ux, *uy, *uz;
  int arr;   // batched launches: conserved moments from the arrays (else recomputed)
};

struct PGeom {
  int nx, ny, nz, walls_y;
  double rad, radp2, r2;
  // (r + 1.5)^2: a node farther than r + 1.5 from the centre is neither
  // solid nor has a link, since every lattice vector is shorter than 1.5
  // (the margin over sqrt(2) dwarfs any rounding of the squared distances)
  double rlink2;
  // (r + 0.5)^2: no solid node lies farther out than this
  double rsolid2;
  // z planes per segment (4 or 8; Particles::slab)
  int slab;
};

struct FillArgs {
  const double *ypglbp, *wp, *wpp, *omgp, *omgpp;
  const int32_t* isn_old;
  const uint32_t *mask_new, *mask_old;
  const uint32_t* sol_old;    // solid lists of the previous map build
  const int32_t* sol_cnt_old;
  int seg_cap;
  double rad, rho0, ww0, ww1, ww2;
  uint64_t* list;       // (particle << 32) | node of every node to refill
  int32_t* count;       // list length (reset by k_scan_counts of the links pass)
  int64_t cap;
};

// Uniform cell grid of the particle pair searches (lubrication, overlap
// check): cells at least as wide as the longest pair range, at most
// kCellCap particles each.  cnt == nullptr: all pairs are searched.
struct CellArgs {
  int32_t* cnt;
  int32_t* list;   // kCellCap slots per cell
  int ncx, ncy, ncz;
};

struct LubArgs {
  CellArgs cells;
  double mingap, fpc, stf0, stf1;
  double mingap_w, stf0_w, stf1_w;   // sphere-wall lubrication (walls_y)
  double* flubp;
};

struct MoveArgs {
  double *yp, *ypglb, *ypglbp, *wp, *wpp, *omgp, *omgpp, *fHIp, *flubp, *torqp, *forcep, *forcepp,
      *torqpp, *dwdt, *domgdt, *thetap;
  double volp, rhog, amp, aip;
  double bfx, bfy, bfz;   // fluid body force (acts on the displaced volume)
  double pfx, pfy, pfz;   // force per unit sphere volume on the spheres only
};

struct ScanArgs {
  const double* ypglb;
  int np;
  int nseg;
  int seg_cap;
  PGeom G;
  int32_t* isn;
  uint32_t* mask;
  uint32_t* sol_list;
  int32_t* sol_cnt;
  uint32_t* sh_node;
  uint32_t* sh_hits;
  int32_t* sh_cnt;
  int32_t* lk_cnt;
  int ovl;
  int istep;
  int32_t* status;
  CellArgs cells;
};

struct ScanCountArgs {
  const int32_t* cnt;
  const int32_t* scnt;
  int nsegs;
  int64_t* off;
  int64_t cap;
  int32_t* status;
  int32_t* fill_count;
  int32_t* ct_count;
  int64_t* part;          // per-chunk sums of the multi-block form (2 per chunk)
};

struct ExpandArgs {
  const double* ypglb;
  int np;
  int nseg;
  int seg_cap;
  PGeom G;
  const uint32_t* sh_node;
  const uint32_t* sh_hits;
  const int32_t* sh_cnt;
  const int64_t* off;
  int64_t cap;
  uint32_t* lk_node;
  uint8_t* lk_dir;
  double* lk_alpha;
  int32_t* status;
  int nsegs;
  FillArgs fa;
  const uint32_t* mask;   // the map being built: links whose node is solid are contact links
  uint64_t* ct_list;
  int32_t* ct_count;
};

// Arguments of the bounce-back grid (link segments, then the map-clearing
// blocks) and of its pre-pass grid (nct blocks for the contact links, then
// nlub lubrication blocks).
struct BounceArgs {
  void* A;
  Grid g;
  PGeom G;
  const uint32_t* mask;
  int nseg;
  const int64_t* off;
  int64_t cap;
  const uint32_t* lk_node;
  const uint8_t* lk_dir;
  const double* lk_alpha;
  double2* lk_out;              // (f9, ff1) of the contact links
  const uint64_t* ct_list;      // contact links: (particle << 40) | link
  const int32_t* ct_count;
  const double* ypglb;
  const double* wp;
  const double* omgp;
  double ww1;
  double ww2;
  double* seg_ft;
  int32_t* status;
  int nsegs;
  int np;
  int split;                    // bounce-back work units per segment (seg_ft has nsegs * split)
  int nct;                      // contact blocks of the pre-pass grid
  int ref_order;                // contact links in the reference's in-place order (one block)
  OrderEntry* ord;              // its table (ord_cap entries) and the current owner map
  int64_t ord_cap;
  const int32_t* isn;
  int nlub;                     // lubrication blocks of the pre-pass grid (fused step, else 0)
  LubArgs lub;
  int seg_cap;                  // map clearing (clr_* of map buffer map_cur ^ 1)
  const uint32_t* clr_list;
  const int32_t* clr_cnt;
  int32_t* clr_isn;
  uint32_t* clr_mask;
};

struct MoveKArgs {
  MoveArgs m;
  int np;
  int first;
  int wrap;
  PGeom G;
  int32_t* status;
  const double* seg_ft;   // non-null: fHIp / torqp are first summed from the segments
  int nseg;
};

struct FillApplyArgs {
  void* A;
  Grid g;
  FillArgs fa;
  int nblk;
};

// ---- argument builders (fluid.hip, particles.hip) ----
CollideArgs collide_args(Ctx& c);
ScanArgs scan_args(Ctx& c, int target, int ovl, int istep);
ScanCountArgs count_args(Ctx& c, int target);
ExpandArgs expand_args(Ctx& c, int target, bool with_find, int* blocks);
// bb_blocks: the bounce-back grid; pre_blocks: the pre-pass grid (nct
// contact blocks + lubrication)
BounceArgs bounce_args(Ctx& c, bool fused, int nct, int* bb_blocks, int* pre_blocks);
MoveKArgs move_args(Ctx& c, int first, int wrap, bool from_seg, int* blocks);
bool fill_apply_args(Ctx& c, int nblk, FillApplyArgs& a);

// ---- launches over several boxes: t = device table of the stage's argument
// blocks (one per box), off = device block offsets (nc + 1), `blocks` = off[nc]
enum Stage { SG_CONTACT, SG_BOUNCE, SG_MOVE, SG_SCAN, SG_COUNT, SG_EXPAND, SG_FILL };
hipError_t launch_batched(Stage st, int prec, int walls, const void* t, const int* off, int nc,
                          int blocks, size_t lds, hipStream_t s);
// arr_any: some member's CollideArgs::arr is set
hipError_t launch_batched_collide(int prec, int walls, int bs, int remap, const void* t, const int* off,
                                  int nc, int blocks, bool arr_any, hipStream_t s);
size_t stage_arg_bytes(Stage st);

}  // namespace mfc
