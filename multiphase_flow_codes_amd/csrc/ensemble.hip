// ensemble.hip -- many independent boxes advanced together.
//
// Each member is an ordinary context (its own size, spheres, rates and body
// force; state created and initialised through the single-box API).  One
// ensemble step runs the time-loop body of every member with one grid per
// stage for all of them: collide, bounce-back (+ lubrication), scatter
// (+ map clearing), move, link scan, link offsets, link expansion (+ refill
// search), refill.  Each grid reads its members' argument blocks from a
// per-stage table and finds its member from a block-offset table
// (off[c] <= blockIdx.x < off[c+1]); both tables for a step are built on the
// host and uploaded with one copy.  The sequence and the per-member state
// transitions are exactly those of mfc_step, so a member advanced in the
// ensemble is bitwise identical to the same box advanced alone.
//
// The collide grid numbers a member's nodes x fastest in blocks of BS
// threads; BS (64, 128 or 256) is picked by mfc_ens_autotune, which times the
// candidates on the current membership.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mfc_lbm.h"
#include "internal.h"
#include "step_args.h"

using namespace mfc;

namespace {

constexpr int kStages = 8;          // collide + the seven particle stages
constexpr int kRing = 3;            // pinned table buffers in flight
constexpr int kFillBlocksPerBox = 32;
constexpr int kContactBlocksPerBox = 1;   // contact pre-pass: contacts are rare

struct Member {
  mfc_ctx* box;
  hipStream_t own_stream;   // the box's stream while outside the ensemble
  int istep;                // step number of the member's next step
};

}  // namespace

struct mfc_ens {
  std::vector<Member> m;
  hipStream_t stream = nullptr;
  int prec = -1, walls = -1;
  int collide_bs = 256;
  // device copy of one step's tables, and the pinned ring they come from
  char* d_tab = nullptr;
  size_t d_cap = 0;
  char* h_tab[kRing] = {};
  size_t h_cap[kRing] = {};
  hipEvent_t h_ev[kRing] = {};
  int ring = 0;
  // graph mode: each ring slot's step (table copy + the stage grids) is
  // captured once and replayed while the grid shapes stay the same
  int use_graph = 0;
  hipGraphExec_t gexec[kRing] = {};
  std::vector<long long> gsig[kRing];
  std::string err;
};

namespace {

int efail(mfc_ens* e, int code, const char* what, hipError_t h = hipSuccess) {
  char buf[512];
  if (h != hipSuccess)
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(h));
  else
    snprintf(buf, sizeof buf, "%s", what);
  e->err = buf;
  return code;
}

#define EHIP(expr)                                                   \
  do {                                                               \
    hipError_t h_ = (expr);                                          \
    if (h_ != hipSuccess) return efail(e, MFC_ERR_HIP, #expr, h_); \
  } while (0)

inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// One stage's part of a step's tables: argument blocks and block offsets.
struct StageTab {
  std::vector<char> args;
  std::vector<int> off{0};
  size_t arg_bytes = 0;
  size_t lds = 0;
  int nc = 0;
  void add(const void* a, int blocks) {
    const size_t o = args.size();
    args.resize(o + arg_bytes);
    memcpy(args.data() + o, a, arg_bytes);
    off.push_back(off.back() + blocks);
    ++nc;
  }
};

// Is the box ready to be stepped by the ensemble?
int check_member(mfc_ens* e, Ctx& c) {
  if (c.p.xkf > 0.0) return efail(e, MFC_ERR_ARG, "ensemble members are driven by a body force only (xkf = 0)");
  if (c.phase != PHASE_POST_STREAM) return efail(e, MFC_ERR_STATE, "ensemble member not at the end of a step");
  if (c.g.lay != LAY_AB) return efail(e, MFC_ERR_ARG, "ensemble members use two-buffer storage (MFC_STORE_AB)");
  if (!(c.mom == MOM_ARRAYS || (c.mom == MOM_VIRTUAL && c.mom_version == c.f_version)))
    return efail(e, MFC_ERR_STATE, "ensemble member needs macrovar before joining");
  if (c.pt.np > 0 && c.pt.map_np[c.pt.map_cur] == 0)
    return efail(e, MFC_ERR_STATE, "ensemble member with spheres needs beads_links before joining");
  return MFC_OK;
}

// Build the tables of one step for every member (advancing each member's
// host-side state exactly as mfc_step would), upload them, launch the stages.
int ens_one_step(mfc_ens* e) {
  StageTab coll, st[7];
  coll.arg_bytes = sizeof(CollideArgs);
  bool arr_any = false;
  const Stage order[7] = {SG_CONTACT, SG_BOUNCE, SG_MOVE, SG_SCAN, SG_COUNT, SG_EXPAND, SG_FILL};
  for (int k = 0; k < 7; ++k) st[k].arg_bytes = stage_arg_bytes(order[k]);
  for (Member& mb : e->m) {
    Ctx& c = mb.box->c;
    Particles& P = c.pt;
    const int it = mb.istep;
    const bool on = P.np > 0 && it >= c.p.irelease;
    // collision_MRT (body force only); moments from the arrays while they are
    // the authoritative copy, as mfc_collision_mrt decides
    {
      if (!(c.mom == MOM_ARRAYS || (c.mom == MOM_VIRTUAL && c.mom_version == c.f_version)))
        return efail(e, MFC_ERR_STATE, "ensemble: member moments out of date");
      CollideArgs a = collide_args(c);
      a.arr = c.mom == MOM_ARRAYS ? 1 : 0;
      arr_any = arr_any || a.arr;
      coll.add(&a, (int)((c.g.ncell + e->collide_bs - 1) / e->collide_bs));
      c.cur ^= 1;
      c.phase = PHASE_PRE_STREAM;
      c.f_version++;
    }
    if (on) {
      int nb = 0, npre = 0;
      const BounceArgs ba = bounce_args(c, true, kContactBlocksPerBox, &nb, &npre);
      st[0].add(&ba, npre);
      st[1].add(&ba, nb);
      c.f_version++;
    }
    c.phase = PHASE_POST_STREAM;   // streaming
    c.f_version++;
    if (on) {
      int nb = 0;
      const MoveKArgs ma = move_args(c, it == c.p.irelease ? 1 : 0, 1, true, &nb);
      st[2].add(&ma, nb);
      P.map_old = P.map_cur;
      const int target = P.map_cur ^ 1;
      if (P.map_np[target] != 0) return efail(e, MFC_ERR_STATE, "ensemble: link target map not cleared");
      const ScanArgs sc = scan_args(c, target, 1, it);
      st[3].add(&sc, P.np * P.nseg);
      st[3].lds = std::max(st[3].lds, sizeof(uint32_t) * (size_t)P.seg_cap);
      const ScanCountArgs cn = count_args(c, target);
      st[4].add(&cn, 1);
      const ExpandArgs ex = expand_args(c, target, true, &nb);
      st[5].add(&ex, nb);
      P.map_np[target] = P.np;
      P.map_cur = target;
      FillApplyArgs fa;
      if (fill_apply_args(c, kFillBlocksPerBox, fa)) st[6].add(&fa, kFillBlocksPerBox);
      c.f_version++;
    }
    // macrovar
    c.mom = MOM_VIRTUAL;
    c.mom_version = c.f_version;
    c.ft.mac = c.ft.cur;
    mb.istep = it + 1;
  }
  // lay the tables out in one buffer: [args][off] per stage
  StageTab* tabs[kStages] = {&coll, &st[0], &st[1], &st[2], &st[3], &st[4], &st[5], &st[6]};
  size_t pos[kStages][2], total = 0;
  for (int k = 0; k < kStages; ++k) {
    pos[k][0] = total;
    total = align_up(total + tabs[k]->args.size());
    pos[k][1] = total;
    total = align_up(total + sizeof(int) * tabs[k]->off.size());
  }
  const int r = e->ring;
  e->ring = (e->ring + 1) % kRing;
  EHIP(hipEventSynchronize(e->h_ev[r]));
  if (e->h_cap[r] < total) {
    if (e->h_tab[r]) EHIP(hipHostFree(e->h_tab[r]));
    e->h_tab[r] = nullptr;
    EHIP(hipHostMalloc((void**)&e->h_tab[r], total * 2, hipHostMallocDefault));
    e->h_cap[r] = total * 2;
  }
  if (e->d_cap < total) {
    EHIP(hipStreamSynchronize(e->stream));
    if (e->d_tab) EHIP(hipFree(e->d_tab));
    e->d_tab = nullptr;
    EHIP(hipMalloc((void**)&e->d_tab, total * 2));
    e->d_cap = total * 2;
  }
  char* h = e->h_tab[r];
  for (int k = 0; k < kStages; ++k) {
    if (!tabs[k]->args.empty()) memcpy(h + pos[k][0], tabs[k]->args.data(), tabs[k]->args.size());
    memcpy(h + pos[k][1], tabs[k]->off.data(), sizeof(int) * tabs[k]->off.size());
  }
  char* d = e->d_tab;
  auto enqueue = [&]() -> int {
    EHIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, e->stream));
    EHIP(launch_batched_collide(e->prec, e->walls, e->collide_bs, d + pos[0][0], (const int*)(d + pos[0][1]),
                                coll.nc, coll.off.back(), arr_any, e->stream));
    for (int k = 0; k < 7; ++k) {
      StageTab& t = st[k];
      EHIP(launch_batched(order[k], e->prec, e->walls, d + pos[k + 1][0], (const int*)(d + pos[k + 1][1]),
                          t.nc, t.off.back(), t.lds, e->stream));
    }
    return MFC_OK;
  };
  if (!e->use_graph) {
    const int rc = enqueue();
    if (rc) return rc;
    EHIP(hipEventRecord(e->h_ev[r], e->stream));
    return MFC_OK;
  }
  // what the captured work depends on besides the table contents: buffers,
  // sizes, launch shapes and the collide kernel (block size, array moments)
  std::vector<long long> sig = {(long long)(uintptr_t)d, (long long)(uintptr_t)h, (long long)total,
                                e->collide_bs, e->prec, e->walls, coll.nc, coll.off.back(),
                                arr_any ? 1LL : 0LL};
  for (int k = 0; k < 7; ++k) {
    sig.push_back(st[k].nc);
    sig.push_back(st[k].off.back());
    sig.push_back((long long)st[k].lds);
  }
  for (int k = 0; k < kStages; ++k) {
    sig.push_back((long long)pos[k][0]);
    sig.push_back((long long)pos[k][1]);
  }
  if (!e->gexec[r] || e->gsig[r] != sig) {
    if (e->gexec[r]) EHIP(hipGraphExecDestroy(e->gexec[r]));
    e->gexec[r] = nullptr;
    hipGraph_t g = nullptr;
    EHIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    const hipError_t end = hipStreamEndCapture(e->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (end != hipSuccess) return efail(e, MFC_ERR_HIP, "hipStreamEndCapture", end);
    const hipError_t inst = hipGraphInstantiate(&e->gexec[r], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (inst != hipSuccess) {
      e->gexec[r] = nullptr;
      return efail(e, MFC_ERR_HIP, "hipGraphInstantiate", inst);
    }
    e->gsig[r] = sig;
  }
  EHIP(hipGraphLaunch(e->gexec[r], e->stream));
  EHIP(hipEventRecord(e->h_ev[r], e->stream));
  return MFC_OK;
}

}  // namespace

extern "C" {

int mfc_ens_create(mfc_ens** out) {
  if (!out) return MFC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MFC_ERR_NODEVICE;
  mfc_ens* e = new (std::nothrow) mfc_ens();
  if (!e) return MFC_ERR_ARG;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return MFC_ERR_HIP;
  }
  for (int r = 0; r < kRing; ++r) {
    if (hipEventCreateWithFlags(&e->h_ev[r], hipEventDisableTiming) != hipSuccess) {
      mfc_ens_destroy(e);
      return MFC_ERR_HIP;
    }
  }
  *out = e;
  return MFC_OK;
}

int mfc_ens_destroy(mfc_ens* e) {
  if (!e) return MFC_ERR_ARG;
  if (e->stream) hipStreamSynchronize(e->stream);
  for (Member& mb : e->m) {
    mb.box->c.stream = mb.own_stream;
    mb.box->c.ens = nullptr;
  }
  e->m.clear();
  for (int r = 0; r < kRing; ++r) {
    if (e->h_tab[r]) hipHostFree(e->h_tab[r]);
    if (e->h_ev[r]) hipEventDestroy(e->h_ev[r]);
  }
  for (int r = 0; r < kRing; ++r)
    if (e->gexec[r]) hipGraphExecDestroy(e->gexec[r]);
  if (e->d_tab) hipFree(e->d_tab);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return MFC_OK;
}

const char* mfc_ens_last_error(const mfc_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int mfc_ens_add(mfc_ens* e, mfc_ctx* box, int32_t istep) {
  if (!e || !box) return MFC_ERR_ARG;
  Ctx& c = box->c;
  for (const Member& mb : e->m)
    if (mb.box == box) return efail(e, MFC_ERR_ARG, "box is already a member");
  int rc = check_member(e, c);
  if (rc) return rc;
  if (e->m.empty()) {
    e->prec = c.prec;
    e->walls = c.g.walls_y;
  } else if (c.prec != e->prec || c.g.walls_y != e->walls) {
    return efail(e, MFC_ERR_ARG, "ensemble members share one precision and one wall setting");
  }
  EHIP(hipStreamSynchronize(c.stream));
  if (c.ens) return efail(e, MFC_ERR_ARG, "box is a member of another ensemble");
  Member mb{box, c.stream, istep};
  c.stream = e->stream;   // the box's own calls now order with the ensemble's work
  c.ens = e;
  e->m.push_back(mb);
  return MFC_OK;
}

int mfc_ens_remove(mfc_ens* e, mfc_ctx* box) {
  if (!e || !box) return MFC_ERR_ARG;
  for (size_t i = 0; i < e->m.size(); ++i) {
    if (e->m[i].box != box) continue;
    EHIP(hipStreamSynchronize(e->stream));
    box->c.stream = e->m[i].own_stream;
    box->c.ens = nullptr;
    box->c.pt.cells_ok = false;   // the members' particles moved
    e->m.erase(e->m.begin() + (long)i);
    if (e->m.empty()) e->prec = e->walls = -1;
    return MFC_OK;
  }
  return efail(e, MFC_ERR_ARG, "box is not a member");
}

int mfc_ens_size(const mfc_ens* e) { return e ? (int)e->m.size() : -1; }

int mfc_ens_step(mfc_ens* e, int32_t nsteps) {
  if (!e || nsteps < 0) return MFC_ERR_ARG;
  if (e->m.empty()) return MFC_OK;
  for (int s = 0; s < nsteps; ++s) {
    int rc = ens_one_step(e);
    if (rc) return rc;
  }
  return MFC_OK;
}

int mfc_ens_sync(mfc_ens* e) {
  if (!e) return MFC_ERR_ARG;
  EHIP(hipStreamSynchronize(e->stream));
  for (Member& mb : e->m) {
    const int rc = mfc_sync(mb.box);
    if (rc) return efail(e, rc, mfc_last_error(mb.box));
  }
  return MFC_OK;
}

int mfc_ens_set_graph(mfc_ens* e, int32_t on) {
  if (!e) return MFC_ERR_ARG;
  EHIP(hipStreamSynchronize(e->stream));
  e->use_graph = on ? 1 : 0;
  return MFC_OK;
}

int mfc_ens_set_block(mfc_ens* e, int32_t bs) {
  if (!e || (bs != 64 && bs != 128 && bs != 256)) return MFC_ERR_ARG;
  e->collide_bs = bs;
  return MFC_OK;
}

// Time `steps` ensemble steps with each collide block size and keep the
// fastest.  The members advance by 3 * steps steps.
int mfc_ens_autotune(mfc_ens* e, int32_t steps, int32_t* chosen, double* ms_per_step) {
  if (!e || steps < 1) return MFC_ERR_ARG;
  if (e->m.empty()) return efail(e, MFC_ERR_STATE, "autotune of an empty ensemble");
  const int cand[3] = {64, 128, 256};
  hipEvent_t t0, t1;
  EHIP(hipEventCreate(&t0));
  EHIP(hipEventCreate(&t1));
  int best = 256;
  float best_ms = 1e30f;
  for (int k = 0; k < 3; ++k) {
    e->collide_bs = cand[k];
    EHIP(hipEventRecord(t0, e->stream));
    int rc = mfc_ens_step(e, steps);
    if (rc) return rc;
    EHIP(hipEventRecord(t1, e->stream));
    EHIP(hipEventSynchronize(t1));
    float ms = 0.f;
    EHIP(hipEventElapsedTime(&ms, t0, t1));
    if (ms_per_step) ms_per_step[k] = ms / steps;
    if (ms < best_ms) {
      best_ms = ms;
      best = cand[k];
    }
  }
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  e->collide_bs = best;
  if (chosen) *chosen = best;
  return MFC_OK;
}

void* mfc_ens_get_stream(mfc_ens* e) { return e ? (void*)e->stream : nullptr; }

}  // extern "C"
