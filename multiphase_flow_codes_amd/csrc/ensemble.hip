// ensemble.hip -- many independent boxes advanced together.
//
// Each member is an ordinary context (its own size, spheres, rates and body
// force; state created and initialised through the single-box API).  One
// ensemble step runs the time-loop body of every member with one grid per
// stage for all of them: collide, bounce-back (+ lubrication), scatter
// (+ map clearing), move, link scan, link offsets, link expansion (+ refill
// search), refill.  Each grid reads its members' argument blocks from a
// per-stage table and finds its member from a block-offset table
// (off[c] <= blockIdx.x < off[c+1]).  The sequence and the per-member state
// transitions are exactly those of mfc_step, so a member advanced in the
// ensemble is bitwise identical to the same box advanced alone.
//
// Steps go in chunks of up to kChunk: the host builds the tables of all the
// chunk's steps into one pinned buffer and uploads them with ONE plain copy;
// then the chunk's stage grids (8 per step) run, in graph mode as one
// recorded HIP graph replayed for every chunk of the same launch shapes.
// The graphs hold kernel nodes only (their table pointers point into the
// device table buffer); a graph is destroyed before any buffer its nodes
// reference is freed.
//
// The collide grid numbers a member's nodes x fastest in blocks of BS
// threads (the pencil), optionally dealt to the XCDs in contiguous ranges;
// mfc_ens_autotune picks BS (64, 128, 256) and the tile order by timing
// every pair on the current membership.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/mfc_lbm.h"
#include "internal.h"
#include "step_args.h"

using namespace mfc;

namespace {

constexpr int kStages = 8;          // collide + the seven particle stages
constexpr int kRing = 3;            // pinned table buffers in flight
constexpr int kChunk = 8;           // steps per table upload and per recorded graph
constexpr int kGraphCache = 4;      // recorded chunk graphs kept (by launch shapes)
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
  int collide_remap = 0;    // 1: collide blocks dealt to the XCDs in contiguous ranges
  // device copy of one chunk's tables, and the pinned ring they come from
  char* d_tab = nullptr;
  size_t d_cap = 0;
  char* h_tab[kRing] = {};
  size_t h_cap[kRing] = {};
  hipEvent_t h_ev[kRing] = {};
  int ring = 0;
  // graph mode: recorded chunks, keyed by what they depend on besides the
  // table contents (buffers, sizes, launch shapes, kernel variants)
  int use_graph = 0;
  struct Rec {
    std::vector<long long> sig;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Rec> graphs;
  // host time spent building and enqueueing steps (waits for the ring excluded)
  double host_ms = 0.0;
  long long host_steps = 0;
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

inline double host_ms_now() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

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

// Drops every recorded graph (after the stream has drained).
int drop_graphs(mfc_ens* e) {
  if (e->graphs.empty()) return MFC_OK;
  EHIP(hipStreamSynchronize(e->stream));
  for (auto& g : e->graphs)
    if (g.exec) EHIP(hipGraphExecDestroy(g.exec));
  e->graphs.clear();
  return MFC_OK;
}

// One step's tables for every member, advancing each member's host-side
// state exactly as mfc_step would.
struct StepTabs {
  StageTab coll, st[7];
  bool arr_any = false;
};
const Stage kOrder[7] = {SG_CONTACT, SG_BOUNCE, SG_MOVE, SG_SCAN, SG_COUNT, SG_EXPAND, SG_FILL};

int build_step(mfc_ens* e, StepTabs& T) {
  T.coll.arg_bytes = sizeof(CollideArgs);
  for (int k = 0; k < 7; ++k) T.st[k].arg_bytes = stage_arg_bytes(kOrder[k]);
  StageTab* st = T.st;
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
      T.arr_any = T.arr_any || a.arr;
      T.coll.add(&a, (int)((c.g.ncell + e->collide_bs - 1) / e->collide_bs));
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
  return MFC_OK;
}

// `n` (1..kChunk) steps of every member: tables built and uploaded at once,
// then the stage grids of all n steps (one recorded graph in graph mode).
int ens_chunk(mfc_ens* e, int n) {
  const double t0 = host_ms_now();
  StepTabs T[kChunk];
  for (int s = 0; s < n; ++s) {
    const int rc = build_step(e, T[s]);
    if (rc) return rc;
  }
  // lay the tables out in one buffer: per step, per stage [args][off]
  size_t pos[kChunk][kStages][2], total = 0;
  for (int s = 0; s < n; ++s) {
    StageTab* tabs[kStages] = {&T[s].coll, &T[s].st[0], &T[s].st[1], &T[s].st[2], &T[s].st[3], &T[s].st[4],
                               &T[s].st[5], &T[s].st[6]};
    for (int k = 0; k < kStages; ++k) {
      pos[s][k][0] = total;
      total = align_up(total + tabs[k]->args.size());
      pos[s][k][1] = total;
      total = align_up(total + sizeof(int) * tabs[k]->off.size());
    }
  }
  double waited = 0.0;
  const int r = e->ring;
  e->ring = (e->ring + 1) % kRing;
  {
    const double w0 = host_ms_now();
    EHIP(hipEventSynchronize(e->h_ev[r]));   // the slot's last upload has been read
    waited += host_ms_now() - w0;
  }
  if (e->h_cap[r] < total) {
    // no graph references host memory: the old buffer's copy has completed
    if (e->h_tab[r]) EHIP(hipHostFree(e->h_tab[r]));
    e->h_tab[r] = nullptr;
    EHIP(hipHostMalloc((void**)&e->h_tab[r], total * 2, hipHostMallocDefault));
    e->h_cap[r] = total * 2;
  }
  if (e->d_cap < total) {
    // the recorded graphs point into the old device buffer: drop them first
    const int rc = drop_graphs(e);
    if (rc) return rc;
    EHIP(hipStreamSynchronize(e->stream));
    if (e->d_tab) EHIP(hipFree(e->d_tab));
    e->d_tab = nullptr;
    EHIP(hipMalloc((void**)&e->d_tab, total * 2));
    e->d_cap = total * 2;
  }
  char* h = e->h_tab[r];
  for (int s = 0; s < n; ++s) {
    StageTab* tabs[kStages] = {&T[s].coll, &T[s].st[0], &T[s].st[1], &T[s].st[2], &T[s].st[3], &T[s].st[4],
                               &T[s].st[5], &T[s].st[6]};
    for (int k = 0; k < kStages; ++k) {
      if (!tabs[k]->args.empty()) memcpy(h + pos[s][k][0], tabs[k]->args.data(), tabs[k]->args.size());
      memcpy(h + pos[s][k][1], tabs[k]->off.data(), sizeof(int) * tabs[k]->off.size());
    }
  }
  char* d = e->d_tab;
  EHIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, e->stream));
  auto enqueue = [&]() -> int {
    for (int s = 0; s < n; ++s) {
      const StepTabs& U = T[s];
      EHIP(launch_batched_collide(e->prec, e->walls, e->collide_bs, e->collide_remap, d + pos[s][0][0],
                                  (const int*)(d + pos[s][0][1]), U.coll.nc, U.coll.off.back(), U.arr_any,
                                  e->stream));
      for (int k = 0; k < 7; ++k) {
        const StageTab& t = U.st[k];
        EHIP(launch_batched(kOrder[k], e->prec, e->walls, d + pos[s][k + 1][0],
                            (const int*)(d + pos[s][k + 1][1]), t.nc, t.off.back(), t.lds, e->stream));
      }
    }
    return MFC_OK;
  };
  if (!e->use_graph) {
    const int rc = enqueue();
    if (rc) return rc;
  } else {
    std::vector<long long> sig = {(long long)(uintptr_t)d, (long long)total, n, e->collide_bs,
                                  e->collide_remap, e->prec, e->walls};
    for (int s = 0; s < n; ++s) {
      sig.push_back(T[s].coll.nc);
      sig.push_back(T[s].coll.off.back());
      sig.push_back(T[s].arr_any ? 1LL : 0LL);
      for (int k = 0; k < 7; ++k) {
        sig.push_back(T[s].st[k].nc);
        sig.push_back(T[s].st[k].off.back());
        sig.push_back((long long)T[s].st[k].lds);
      }
      for (int k = 0; k < kStages; ++k) {
        sig.push_back((long long)pos[s][k][0]);
        sig.push_back((long long)pos[s][k][1]);
      }
    }
    mfc_ens::Rec* rec = nullptr;
    for (auto& g : e->graphs)
      if (g.sig == sig) rec = &g;
    if (!rec) {
      if ((int)e->graphs.size() >= kGraphCache) {
        // the oldest recording may still be queued: let the stream drain
        EHIP(hipStreamSynchronize(e->stream));
        EHIP(hipGraphExecDestroy(e->graphs.front().exec));
        e->graphs.erase(e->graphs.begin());
      }
      hipGraph_t g = nullptr;
      EHIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue();
      const hipError_t end = hipStreamEndCapture(e->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (end != hipSuccess) return efail(e, MFC_ERR_HIP, "hipStreamEndCapture", end);
      hipGraphExec_t ex = nullptr;
      const hipError_t inst = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (inst != hipSuccess) return efail(e, MFC_ERR_HIP, "hipGraphInstantiate", inst);
      e->graphs.push_back({sig, ex});
      rec = &e->graphs.back();
    }
    EHIP(hipGraphLaunch(rec->exec, e->stream));
  }
  EHIP(hipEventRecord(e->h_ev[r], e->stream));
  e->host_ms += host_ms_now() - t0 - waited;
  e->host_steps += n;
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
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (Member& mb : e->m) {
    mb.box->c.stream = mb.own_stream;
    mb.box->c.ens = nullptr;
  }
  e->m.clear();
  // graphs first: their nodes point into the table buffers freed below
  for (auto& g : e->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  e->graphs.clear();
  for (int r = 0; r < kRing; ++r) {
    if (e->h_tab[r]) (void)hipHostFree(e->h_tab[r]);
    if (e->h_ev[r]) (void)hipEventDestroy(e->h_ev[r]);
  }
  if (e->d_tab) (void)hipFree(e->d_tab);
  if (e->stream) (void)hipStreamDestroy(e->stream);
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
  for (int s = 0; s < nsteps; s += kChunk) {
    int rc = ens_chunk(e, std::min(kChunk, nsteps - s));
    if (rc) return rc;
  }
  return MFC_OK;
}

int mfc_ens_host_time(mfc_ens* e, double* ms, int64_t* steps) {
  if (!e) return MFC_ERR_ARG;
  if (ms) *ms = e->host_ms;
  if (steps) *steps = e->host_steps;
  e->host_ms = 0.0;
  e->host_steps = 0;
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

int mfc_ens_set_tile_order(mfc_ens* e, int32_t order) {
  if (!e || (order != 0 && order != 1)) return MFC_ERR_ARG;
  e->collide_remap = order;
  return MFC_OK;
}

// Time `steps` ensemble steps with each pair (collide block size, tile
// order), after `steps` untimed steps that record the pair's graphs, and
// keep the fastest.  The members advance by 12 * steps steps.
int mfc_ens_autotune(mfc_ens* e, int32_t steps, int32_t* chosen, double* ms_per_step) {
  if (!e || steps < 1) return MFC_ERR_ARG;
  if (e->m.empty()) return efail(e, MFC_ERR_STATE, "autotune of an empty ensemble");
  const int bsz[3] = {64, 128, 256};
  // the timing events, destroyed on every return
  struct Events {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    ~Events() {
      if (t0) (void)hipEventDestroy(t0);
      if (t1) (void)hipEventDestroy(t1);
    }
  } ev;
  EHIP(hipEventCreate(&ev.t0));
  EHIP(hipEventCreate(&ev.t1));
  hipEvent_t t0 = ev.t0, t1 = ev.t1;
  int best_bs = 256, best_order = 0;
  float best_ms = 1e30f;
  for (int k = 0; k < 6; ++k) {
    e->collide_bs = bsz[k % 3];
    e->collide_remap = k / 3;
    int rc = mfc_ens_step(e, steps);
    if (rc) return rc;
    EHIP(hipEventRecord(t0, e->stream));
    rc = mfc_ens_step(e, steps);
    if (rc) return rc;
    EHIP(hipEventRecord(t1, e->stream));
    EHIP(hipEventSynchronize(t1));
    float ms = 0.f;
    EHIP(hipEventElapsedTime(&ms, t0, t1));
    ms /= (float)steps;
    if (ms_per_step) ms_per_step[k] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best_bs = bsz[k % 3];
      best_order = k / 3;
    }
  }
  e->collide_bs = best_bs;
  e->collide_remap = best_order;
  if (chosen) {
    chosen[0] = best_bs;
    chosen[1] = best_order;
  }
  return MFC_OK;
}

void* mfc_ens_get_stream(mfc_ens* e) { return e ? (void*)e->stream : nullptr; }

}  // extern "C"
