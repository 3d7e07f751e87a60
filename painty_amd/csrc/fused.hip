// PNT_EXEC_FUSED: a stamp list planned once on the host, then one kernel
// launch per imprint (FootprintBrush::imprint, renderer/FootprintBrush.hxx:
// 73-143) in which the snapshot refresh and the visits run side by side.
//
// Blocks [0, ntiles) take one kT x kT tile of the window each (window_tile,
// stampcore.hpp); the blocks after them scan the dirty bytes of the
// snapshot ring (ring_scan).  The two sets touch disjoint pixels: the scan
// skips the window interior and the border pixels that a visit refreshes
// itself (a border visit reads the canvas value the refresh would have
// copied and writes the snapshot back), so they need no ordering.  A stamp
// that straddles canvas column/row 0 with a fractional centre visits some
// pixels twice; it runs as ring | window | shared-pixel levels, three
// launches in stream order.
//
// Between launches the stream orders everything: no grid barrier, no
// cooperative launch.  Consecutive stamps of a stroke overlap almost
// entirely, and tile t of every launch runs on the same XCD
// (tile_of_block), so a tile's pixels and map cells stay in that XCD's L2.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"

namespace pnt {

namespace {

struct FArgs {
  const PStamp* st;
  Layer7 canvas, snap;  // snap == canvas planes when the snapshot is off
  Plane7 pick;
  const float* fp;
  FpBox fb;  // nonzero footprint cells
  uint8_t* dirty;
  const int2* dl;
  const int* dl_off;
  int use_snap, rows, cols, hr, size_map, tiles_x, ntiles;
  double pickup_rate, dep_rate, cap;
  float pk[3], ps[3];  // brush paint (the last dip)
  DryClock clk;        // the canvas' drying clock; clk.tmap == nullptr: none
};

// The clock of stamp `st` (its own time point); tmap == nullptr without one.
__device__ __forceinline__ DryClock stamp_clock(const DryClock& base, const PStamp& st) {
  DryClock c = base;
  c.now_us = st.now_us;
  return c;
}

__device__ __forceinline__ StampGeom geom_of(const FArgs& A, const PStamp& st) {
  StampGeom g;
  g.cx = st.cx;
  g.cy = st.cy;
  g.c = st.c;
  g.s = st.s;
  g.hr = A.hr;
  g.size_map = A.size_map;
  g.rows = A.rows;
  g.cols = A.cols;
  return g;
}

// Stamp j: window tiles on blocks [0, ntiles), ring scan on the next
// ring_blocks blocks, and for a stamp on a duplicated canvas column/row 0
// its shared-pixel visits by level on one more block.  Visits on the window
// border refresh their own snapshot pixels (the ring scan leaves them).
template <int kT>
__global__ __launch_bounds__(kT* kT) void k_stamp_fused(FArgs A, int j, int ring_blocks) {
  __shared__ int sm[kT + 2][kT + 2];
  const PStamp& st = A.st[j];
  const StampGeom g = geom_of(A, st);
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  const bool pull = A.use_snap != 0;
  const DryClock clk = stamp_clock(A.clk, st);
  const VisitTarget T{&A.canvas, &A.snap, A.dirty, A.use_snap, A.pickup_rate, A.dep_rate, A.cap};
  const int b = (int)blockIdx.x;
  if (b < A.ntiles) {
    window_tile<kT>(T, A.pick.p, A.fp, A.fb, g, st, dup, pull, tile_of_block(b, A.ntiles),
                    A.tiles_x, sm, A.pk, A.ps, clk);
  } else if (b < A.ntiles + ring_blocks) {
    ring_scan<kRingGroups>(A.canvas, A.snap, A.dirty, A.fp, g, st.box, pull,
                 (b - A.ntiles) * kT * kT + (int)threadIdx.x, ring_blocks * kT * kT,
                 clk.tmap != nullptr);
  } else if (dup) {
    level_visits<kT * kT>(T, A.pick.p, A.fp, g, st, A.dl, A.dl_off, A.pk, A.ps, clk, pull);
  }
}

}  // namespace

// 4-pixel groups of box b, and the blocks that scan them (each thread
// takes kRingGroups groups per pass).
int ring_blocks_for(const RingBox& b, int threads, int groups) {
  if (b.a_x1 < b.a_x0 || b.a_y1 < b.a_y0) return 0;
  const int64_t span = b.a_x1 - b.a_x0 + 1;
  const int64_t ngroups = ((span + 3) / 4 + 1) * (int64_t)(b.a_y1 - b.a_y0 + 1);
  const int64_t blocks = (ngroups + groups * threads - 1) / (groups * threads);
  return (int)std::min<int64_t>(std::max<int64_t>(blocks, 1), 1024);
}


// ---------------------------------------------------- concurrent strokes --
// A dip cleans the pickup map (FootprintBrush.hxx:150-160), so the stamps
// after it depend on earlier stamps only through the canvas, the snapshot
// and their dirty bytes -- and a stamp touches those only inside its
// snapshot's allowed box (the window lies inside it).  The list is cut into
// units: a run of stamps of one stroke (from a dip to the next dip) within
// one drying phase (up to a PNT_STAMP_DRY_AFTER).  A unit depends on the
// earlier units of its phase whose boxes (unions of their stamps' allowed
// boxes) intersect its own, and a dryCanvas on everything before it.
// Independent units run at once on up to kMaxChains streams ("chains"),
// each chain with its own pickup map; every stamp still runs exactly the
// kernels of the sequential order, on pixels no concurrent unit touches,
// so the result is the same to the bit.  At the end the pickup map of the
// list's last stroke becomes the brush's.
constexpr int kMaxChains = 8;

struct FusedState {
  hipStream_t streams[kMaxChains] = {};  // [0] unused: chain 0 is the context's stream
  float* pick[kMaxChains] = {};          // [0] unused: chain 0 is the brush's own map
  int64_t pick_cap[kMaxChains] = {};
  std::vector<hipEvent_t> events;
};

void ktimer_release(KernelTimer*& t) {
  if (!t) return;
  for (hipEvent_t e : t->ev) (void)hipEventDestroy(e);
  delete t;
  t = nullptr;
}

double ktimer_total_ms(KernelTimer& t) {
  if (t.used == 0) return 0.0;
  PNT_HIP_CHECK(hipEventSynchronize(t.ev[t.used - 1]));
  double ms = 0.0;
  for (size_t k = 0; k + 1 < t.used; k += 2) {
    float x = 0.0f;
    PNT_HIP_CHECK(hipEventElapsedTime(&x, t.ev[k], t.ev[k + 1]));
    ms += x;
  }
  return ms;
}

// Records the next event of a timed launch on s (a start before the launch,
// an end after it); nothing when timing is off.
static void ktimer_mark(pnt_brush_s* b, hipStream_t s) {
  KernelTimer* t = b->ktimer;
  if (!t || !t->on) return;
  if (t->used == t->ev.size()) {
    hipEvent_t e;
    PNT_HIP_CHECK(hipEventCreate(&e));
    t->ev.push_back(e);
  }
  PNT_HIP_CHECK(hipEventRecord(t->ev[t->used++], s));
}

void fused_release(FusedState*& F) {
  if (!F) return;
  for (int c = 1; c < kMaxChains; ++c) {
    if (F->streams[c]) (void)hipStreamSynchronize(F->streams[c]);
    if (F->streams[c]) (void)hipStreamDestroy(F->streams[c]);
    if (F->pick[c]) (void)hipFree(F->pick[c]);
  }
  for (hipEvent_t e : F->events) (void)hipEventDestroy(e);
  delete F;
  F = nullptr;
}

static bool boxes_meet(const Unit& a, const Unit& b) {
  if (a.x1 < a.x0 || b.x1 < b.x0) return false;
  return !(a.x1 < b.x0 || b.x1 < a.x0 || a.y1 < b.y0 || b.y1 < a.y0);
}

// Cuts the list into units, finds their dependencies, and assigns chains by
// list scheduling on estimated durations (stamp counts): a unit goes to the
// chain where it can start first; a unit that continues a stroke stays on
// the stroke's chain, and one that continues the brush's current stroke
// (no dip or geometry change before it) runs on chain 0.  clear[k]: the
// pickup map is cleaned before stamp k (a new brush geometry).
std::vector<Unit> plan_units(const pnt_stamp* st, const PStamp* h, const uint8_t* clear,
                             int64_t n, int nchains) {
  std::vector<Unit> U;
  for (int64_t k = 0; k < n; ++k) {
    const bool dip = (st[k].flags & PNT_STAMP_DIP) != 0 || clear[k];  // a clean map
    if (U.empty() || dip || U.back().dry_after) {
      Unit u{k, k, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, -1, false, {}};
      u.chain = dip ? -1 : (U.empty() ? 0 : -2);  // -2: the previous unit's chain
      U.push_back(u);
    }
    Unit& u = U.back();
    u.s1 = k + 1;
    const RingBox& b = h[k].box;
    if (b.a_x1 >= b.a_x0 && b.a_y1 >= b.a_y0) {
      u.x0 = std::min(u.x0, b.a_x0);
      u.y0 = std::min(u.y0, b.a_y0);
      u.x1 = std::max(u.x1, b.a_x1);
      u.y1 = std::max(u.y1, b.a_y1);
    }
    if (st[k].flags & PNT_STAMP_DRY_AFTER) u.dry_after = true;
  }
  std::vector<double> avail(nchains, 0.0), finish(U.size(), 0.0);
  double phase_start = 0.0;
  size_t phase0 = 0;
  for (size_t i = 0; i < U.size(); ++i) {
    Unit& u = U[i];
    double dep_t = phase_start;
    for (size_t j = phase0; j < i; ++j)
      if (boxes_meet(U[j], u)) {
        u.deps.push_back((int)j);
        dep_t = std::max(dep_t, finish[j]);
      }
    if (u.chain == -2) u.chain = U[i - 1].chain;
    if (u.chain < 0) {
      int best = 0;
      for (int c = 1; c < nchains; ++c)
        if (std::max(avail[c], dep_t) < std::max(avail[best], dep_t)) best = c;
      u.chain = best;
    }
    finish[i] = std::max(avail[u.chain], dep_t) + (double)(u.s1 - u.s0);
    avail[u.chain] = finish[i];
    if (u.dry_after) {  // dryCanvas: everything after waits for everything before
      phase_start = *std::max_element(finish.begin(), finish.begin() + (std::ptrdiff_t)i + 1);
      std::fill(avail.begin(), avail.end(), phase_start);
      phase0 = i + 1;
    }
  }
  return U;
}

// For every stamp k of a unit, the units it depends on and how many of
// their leading stamps must have run first: one past the last of that
// unit's stamps whose allowed box meets stamp k's (none meeting: no entry).
// A stamp touches nothing outside its allowed box, so a unit may start
// while the units it follows still run, as long as every pair of stamps
// whose boxes meet keeps the list's order.
StampNeeds plan_stamp_needs(const std::vector<Unit>& U, const PStamp* h) {
  StampNeeds N;
  const int64_t n = U.empty() ? 0 : U.back().s1;
  N.off.assign((size_t)n + 1, 0);
  auto meet = [&](int64_t a, int64_t b) {
    const RingBox &x = h[a].box, &y = h[b].box;
    if (x.a_x1 < x.a_x0 || y.a_x1 < y.a_x0) return false;
    return !(x.a_x1 < y.a_x0 || y.a_x1 < x.a_x0 || x.a_y1 < y.a_y0 || y.a_y1 < x.a_y0);
  };
  for (size_t i = 0; i < U.size(); ++i)
    for (int64_t k = U[i].s0; k < U[i].s1; ++k) {
      N.off[(size_t)k] = (int)N.need.size();
      for (int d : U[i].deps) {
        int64_t last = -1;
        for (int64_t q = U[(size_t)d].s1 - 1; q >= U[(size_t)d].s0 && last < 0; --q)
          if (meet(q, k)) last = q;
        if (last >= 0) N.need.push_back(make_int2(d, (int)(last - U[(size_t)d].s0 + 1)));
      }
    }
  N.off[(size_t)n] = (int)N.need.size();
  return N;
}

namespace {

// Concurrent units run as waves on one stream (the default), or on
// concurrent streams with PAINTY_FUSED_SCHED=streams.
bool fused_waves() {
  const char* e = std::getenv("PAINTY_FUSED_SCHED");
  return !(e && std::strcmp(e, "streams") == 0);
}

// Concurrent chains of one call: `want` if given (> 0), else
// PAINTY_FUSED_CHAINS, else 8 for waves (the slots of a wave) or 4 for
// streams (the hardware queues of a process).
int fused_chains(int want) {
  int c = want > 0 ? want : (fused_waves() ? kMaxChains : 4);
  if (want <= 0)
    if (const char* e = std::getenv("PAINTY_FUSED_CHAINS")) c = std::atoi(e);
  return std::min(std::max(c, 1), kMaxChains);
}

// ----------------------------------------------------------------- waves --
// The concurrent strokes again, on one stream: each launch ("wave") runs
// the next phase of every unit that is ready -- a stamp (window tiles, ring
// scan and, for a stamp that straddles canvas column/row 0, its
// shared-pixel levels side by side), or the pickup-map clear of a dip or a
// geometry change -- each on its own range of blocks.  Units of
// one wave touch disjoint pixels and their own pickup maps, exactly as on
// concurrent streams, and the stream orders the waves: no events, and no
// limit from the hardware queues.
constexpr int kWaveSlots = 8;
enum : int { kPhClear = 0, kPhStamp = 1 };

struct WSlot {
  int j;       // stamp index (not used by kPhClear)
  int kind;    // kPh*
  int start;   // first block of the slot (a multiple of 8: tiles keep their XCD)
  int nblk;    // blocks with work (the slot's range may be padded past them)
  int ntiles, rb;
  int hr, size_map, tiles_x;
  const float* fp;
  FpBox fb;
  float* pick;         // the unit's pickup map (7 planes of mstride)
  int64_t mstride;
  float pk[3], ps[3];  // the brush paint of this stamp
  PStamp stamp;        // the stamp's plan, passed with the launch: its blocks start
                       // without a dependent load of it
};

struct WArgs {
  const PStamp* st;  // (the slots carry their stamps' plans; this is the whole list)
  Layer7 canvas, snap;
  uint8_t* dirty;
  const int2* dl;
  const int* dl_off;
  int use_snap, rows, cols;
  double pickup_rate, dep_rate, cap;
  DryClock clk;  // the canvas' drying clock; clk.tmap == nullptr: none
  int nslots;
  WSlot slot[kWaveSlots];
};
static_assert(sizeof(WArgs) <= 4000, "wave arguments must fit the kernel argument segment");


// 16x16 tiles: five workgroups per CU (at most 102 VGPRs; it needs 102
// without the bound and spills nothing under it).
template <int kT>
__global__ __launch_bounds__(kT* kT, kT == 16 ? 5 : 1) void k_stamp_wave(WArgs A) {
  __shared__ int sm[kT + 2][kT + 2];
  int si = 0;
  while (si + 1 < A.nslots && (int)blockIdx.x >= A.slot[si + 1].start) ++si;
  const WSlot& S = A.slot[si];
  const int b = (int)blockIdx.x - S.start;
  if (b >= S.nblk) return;
  if (S.kind == kPhClear) {  // dip() / setRadius: a clean pickup map
    const int64_t total = 7 * S.mstride;
    for (int64_t i = (int64_t)b * kT * kT + threadIdx.x; i < total; i += (int64_t)S.nblk * kT * kT)
      S.pick[i] = 0.0f;
    return;
  }
  const PStamp& st = S.stamp;
  StampGeom g;
  g.cx = st.cx;
  g.cy = st.cy;
  g.c = st.c;
  g.s = st.s;
  g.hr = S.hr;
  g.size_map = S.size_map;
  g.rows = A.rows;
  g.cols = A.cols;
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  const bool pull = A.use_snap != 0;  // border visits refresh their own snapshot pixels
  const VisitTarget T{&A.canvas, &A.snap, A.dirty, A.use_snap, A.pickup_rate, A.dep_rate, A.cap};
  float* pick[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) pick[c] = S.pick + c * S.mstride;
  const DryClock clk = stamp_clock(A.clk, st);
  if (b < S.ntiles) {
    window_tile<kT>(T, pick, S.fp, S.fb, g, st, dup, pull, tile_of_block(b, S.ntiles), S.tiles_x,
                    sm, S.pk, S.ps, clk);
  } else if (b < S.ntiles + S.rb) {
    ring_scan<kRingGroups>(A.canvas, A.snap, A.dirty, S.fp, g, st.box, pull, (b - S.ntiles) * kT * kT +
                 (int)threadIdx.x, S.rb * kT * kT, clk.tmap != nullptr);
  } else if (dup) {  // the shared-pixel visits, beside the tiles
    level_visits<kT * kT>(T, pick, S.fp, g, st, A.dl, A.dl_off, S.pk, S.ps, clk, pull);
  }
}

// One brush geometry of a list: its padded footprint on the device, its side
// and radius, and (for the brush's state after the list) the unpadded
// footprint it came from; core == nullptr: the brush's current geometry.
struct ListGeo {
  const float* fp;
  int size_map;
  double radius;
  const float* core;
  int width;
  FpBox fb;  // nonzero cells of fp
};

// Stamps [0, n) of a list on the fused path.  Stamp k uses geometry
// geo[k]; clear[k] cleans the pickup map before it (a geometry change).
template <int kT>
void run_fused_list(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n,
                    const PlanDev& D, const std::vector<ListGeo>& geos, const std::vector<int>& geo,
                    const std::vector<uint8_t>& clear, int want_chains) {
  hipStream_t s = b->ctx->stream;
  b->last_launches = 0;
  if (b->ktimer) b->ktimer->used = 0;
  FArgs A0;
  A0.st = D.st;
  A0.canvas = cv->layer();
  A0.snap = b->use_snap ? b->snap_planes() : cv->layer();
  A0.dirty = b->dirty;
  A0.dl = D.dl;
  A0.dl_off = D.dl_off;
  A0.use_snap = b->use_snap ? 1 : 0;
  A0.rows = cv->rows;
  A0.cols = cv->cols;
  A0.pickup_rate = b->pickup_rate;
  A0.dep_rate = b->dep_rate;
  A0.cap = b->cap;
  // Canvas::checkDry on every visited pixel when the canvas has a drying
  // clock; each stamp carries its own time point (PStamp::now_us)
  A0.clk = cv->clock != PNT_CLOCK_OFF && cv->tmap
               ? DryClock{cv->tmap, 0, cv->drying_ms, cv->R0(), cv->h()}
               : DryClock{};
  std::vector<FArgs> GA(geos.size(), A0);
  std::vector<int64_t> gstride(geos.size());
  int64_t max_stride = 0;
  for (size_t i = 0; i < geos.size(); ++i) {
    FArgs& a = GA[i];
    a.fp = geos[i].fp;
    a.fb = geos[i].fb;
    a.size_map = geos[i].size_map;
    a.hr = (a.size_map - 1) / 2;
    const int side = 2 * a.hr + 1;
    a.tiles_x = (side + kT - 1) / kT;
    a.ntiles = a.tiles_x * a.tiles_x;
    gstride[i] = round_up((int64_t)a.size_map * a.size_map, 64);
    max_stride = std::max(max_stride, gstride[i]);
  }

  // one stamp on stream `q` with pickup map `pick` (7 planes of its stride)
  auto launch_stamp = [&](int64_t k, hipStream_t q, float* pick) {
    FArgs a = GA[(size_t)geo[k]];
    const int64_t ms = gstride[(size_t)geo[k]];
    for (int c = 0; c < 7; ++c) a.pick.p[c] = pick + c * ms;
    if (clear[k] || (st[k].flags & PNT_STAMP_DIP))  // setRadius / dip(): a clean pickup map
      launch_fill(pick, 7 * ms, 0.0f, q);
    if (st[k].flags & PNT_STAMP_DIP)
      for (int c = 0; c < 3; ++c) {
        b->paintK[c] = st[k].K[c];
        b->paintS[c] = st[k].S[c];
      }
    for (int c = 0; c < 3; ++c) {
      a.pk[c] = b->paintK[c];
      a.ps[c] = b->paintS[c];
    }
    const PStamp& p = D.host[k];
    const int rb = a.use_snap ? ring_blocks_for(p.box, kT * kT) : 0;
    const int jk = (int)k;
    if (q == s) ktimer_mark(b, q);
    k_stamp_fused<kT><<<a.ntiles + rb + (dup_levels(p) ? 1 : 0), kT * kT, 0, q>>>(a, jk, rb);
    if (q == s) ktimer_mark(b, q);
    b->last_launches += 1;
    PNT_HIP_CHECK(hipGetLastError());
  };
  auto dry = [&](hipStream_t q) {
    if (b->use_snap)
      launch_dry_dirty(cv->n, cv->layer(), cv->R0(), cv->h(), b->dirty, q);
    else
      launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), q);
    if (A0.clk.tmap)  // dryCanvas stamps every pixel's time (Canvas.hxx:105-119)
      launch_fill_i64(cv->tmap, cv->n, canvas_now_us(cv), q);
  };
  // the brush ends with the geometry of the list's last stamp
  auto finish_geometry = [&]() {
    const ListGeo& g = geos[(size_t)geo[(size_t)n - 1]];
    if (g.core) brush_set_geometry(b, g.radius, g.core, g.width);  // clears b->pick
  };

  // ---- waves: the units' phases, scheduled on s ----------------------------
  auto run_waves = [&](const std::vector<Unit>& U, int used, float* const* pick) {
    const int T2 = kT * kT;
    struct Ph {
      int64_t k;
      int kind;
    };
    std::vector<std::vector<Ph>> ph(U.size());
    for (size_t i = 0; i < U.size(); ++i)
      for (int64_t k = U[i].s0; k < U[i].s1; ++k) {
        if (clear[k] || (st[k].flags & PNT_STAMP_DIP)) ph[i].push_back(Ph{k, kPhClear});
        ph[i].push_back(Ph{k, kPhStamp});
      }
    // the brush paint of every stamp: that of the last dip up to it
    std::vector<float> paint(6 * (size_t)n);
    float cur[6] = {b->paintK[0], b->paintK[1], b->paintK[2],
                    b->paintS[0], b->paintS[1], b->paintS[2]};
    for (int64_t k = 0; k < n; ++k) {
      if (st[k].flags & PNT_STAMP_DIP)
        for (int c = 0; c < 3; ++c) {
          cur[c] = st[k].K[c];
          cur[3 + c] = st[k].S[c];
        }
      std::memcpy(&paint[6 * (size_t)k], cur, sizeof cur);
    }
    WArgs W{};
    W.st = A0.st;
    W.canvas = A0.canvas;
    W.snap = A0.snap;
    W.dirty = A0.dirty;
    W.dl = A0.dl;
    W.dl_off = A0.dl_off;
    W.use_snap = A0.use_snap;
    W.rows = A0.rows;
    W.cols = A0.cols;
    W.pickup_rate = A0.pickup_rate;
    W.dep_rate = A0.dep_rate;
    W.cap = A0.cap;
    W.clk = A0.clk;
    std::vector<size_t> next(U.size(), 0);
    std::vector<char> done(U.size(), 0);
    std::vector<int64_t> sdone(U.size(), 0);  // stamps of each unit issued
    const StampNeeds NS = plan_stamp_needs(U, D.host);
    std::vector<int> prev_on_chain(U.size(), -1);
    {
      std::vector<int> last(used, -1);
      for (size_t i = 0; i < U.size(); ++i) {
        prev_on_chain[i] = last[U[i].chain];
        last[U[i].chain] = (int)i;
        if (ph[i].empty()) done[i] = 1;
      }
    }
    size_t g0 = 0;  // first unit of the current drying phase
    while (g0 < U.size()) {
      size_t g1 = g0;  // units [g0, g1] form the phase (up to a dry)
      while (g1 + 1 < U.size() && !U[g1].dry_after) ++g1;
      for (;;) {
        std::vector<int> ready;
        for (size_t i = g0; i <= g1 && (int)ready.size() < kWaveSlots; ++i) {
          if (done[i]) continue;
          bool ok = prev_on_chain[i] < 0 || done[(size_t)prev_on_chain[i]];
          const Ph& f = ph[i][next[i]];
          if (ok && f.kind == kPhStamp)  // the stamps it follows have run
            for (int j = NS.off[(size_t)f.k]; j < NS.off[(size_t)f.k + 1] && ok; ++j) {
              const int2 nd = NS.need[(size_t)j];
              ok = done[(size_t)nd.x] || sdone[(size_t)nd.x] >= nd.y;
            }
          if (ok) ready.push_back((int)i);
        }
        if (ready.empty()) break;  // the phase is done
        W.nslots = (int)ready.size();
        int start = 0;
        for (int q = 0; q < W.nslots; ++q) {
          const int i = ready[(size_t)q];
          const Ph& f = ph[(size_t)i][next[(size_t)i]];
          WSlot& S = W.slot[q];
          const FArgs& a = GA[(size_t)geo[(size_t)f.k]];
          S.j = (int)f.k;
          S.stamp = D.host[f.k];
          S.kind = f.kind;
          S.start = start;
          S.ntiles = a.ntiles;
          S.rb = a.use_snap ? ring_blocks_for(D.host[f.k].box, T2) : 0;
          S.hr = a.hr;
          S.size_map = a.size_map;
          S.tiles_x = a.tiles_x;
          S.fp = a.fp;
          S.fb = a.fb;
          S.mstride = gstride[(size_t)geo[(size_t)f.k]];
          S.pick = pick[U[(size_t)i].chain];
          for (int c = 0; c < 3; ++c) {
            S.pk[c] = paint[6 * (size_t)f.k + c];
            S.ps[c] = paint[6 * (size_t)f.k + 3 + c];
          }
          if (f.kind == kPhClear)
            S.nblk = (int)std::min<int64_t>(64, (7 * S.mstride + 4 * T2 - 1) / (4 * T2));
          else  // + the shared-pixel levels' block of a dup stamp
            S.nblk = S.ntiles + S.rb + (dup_levels(D.host[f.k]) ? 1 : 0);
          start += (std::max(S.nblk, 1) + 7) / 8 * 8;
        }
        ktimer_mark(b, s);
        k_stamp_wave<kT><<<(unsigned)start, T2, 0, s>>>(W);
        PNT_HIP_CHECK(hipGetLastError());
        ktimer_mark(b, s);
        b->last_launches += 1;
        for (int i : ready) {
          const Ph& f = ph[(size_t)i][next[(size_t)i]];
          if (f.kind == kPhStamp) sdone[(size_t)i] = f.k - U[(size_t)i].s0 + 1;
          if (++next[(size_t)i] == ph[(size_t)i].size()) done[(size_t)i] = 1;
        }
      }
      if (U[g1].dry_after) dry(s);
      g0 = g1 + 1;
    }
    std::memcpy(b->paintK, cur, 3 * sizeof(float));
    std::memcpy(b->paintS, cur + 3, 3 * sizeof(float));
  };

  const bool one_geo = geos.size() == 1 && !geos[0].core;
  const int nchains = fused_chains(want_chains);
  std::vector<Unit> U;
  if (nchains > 1 || !one_geo) U = plan_units(st, D.host, clear.data(), n, nchains);
  int used = 1;
  for (const Unit& u : U) used = std::max(used, u.chain + 1);
  if (used <= 1 && one_geo) {  // one chain, the brush's own map: the list in stream order
    for (int64_t k = 0; k < n; ++k) {
      launch_stamp(k, s, b->pick);
      if (st[k].flags & PNT_STAMP_DRY_AFTER) dry(s);
    }
  } else {
    if (!b->fused) b->fused = new FusedState;
    FusedState& F = *b->fused;
    hipStream_t S[kMaxChains];
    float* pick[kMaxChains];
    for (int c = 0; c < used; ++c) {
      if (c > 0 && !F.streams[c])
        PNT_HIP_CHECK(hipStreamCreateWithFlags(&F.streams[c], hipStreamNonBlocking));
      if (F.pick_cap[c] < 7 * max_stride) {
        if (c > 0) PNT_HIP_CHECK(hipStreamSynchronize(F.streams[c]));
        PNT_HIP_CHECK(hipStreamSynchronize(s));
        dev_free(F.pick[c]);
        F.pick[c] = nullptr;
        F.pick[c] = dev_alloc(7 * max_stride);
        F.pick_cap[c] = 7 * max_stride;
      }
      S[c] = c == 0 ? s : F.streams[c];
      pick[c] = F.pick[c];
    }
    if (fused_waves()) {
      // a first unit that continues the brush's stroke starts from its map
      if (!clear[0] && !(st[0].flags & PNT_STAMP_DIP))
        PNT_HIP_CHECK(hipMemcpyAsync(pick[0], b->pick,
                                     7 * gstride[(size_t)geo[0]] * sizeof(float),
                                     hipMemcpyDeviceToDevice, s));
      run_waves(U, used, pick);
      finish_geometry();
      PNT_HIP_CHECK(hipMemcpyAsync(b->pick, pick[U.back().chain],
                                   7 * gstride[(size_t)geo[(size_t)n - 1]] * sizeof(float),
                                   hipMemcpyDeviceToDevice, s));
      cv->epoch++;
      if (b->use_snap) {
        b->dirty_canvas = cv;
        b->dirty_epoch = cv->epoch;
      }
      return;
    }
    const size_t nev = 2 * U.size() + 2;  // units, dries, start
    while (F.events.size() < nev) {
      hipEvent_t e;
      PNT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      F.events.push_back(e);
    }
    size_t next_ev = 0;
    auto record = [&](hipStream_t q) {
      hipEvent_t e = F.events[next_ev++];
      PNT_HIP_CHECK(hipEventRecord(e, q));
      return e;
    };
    // a first unit that continues the brush's stroke starts from its map
    if (!clear[0] && !(st[0].flags & PNT_STAMP_DIP))
      PNT_HIP_CHECK(hipMemcpyAsync(pick[0], b->pick, 7 * gstride[(size_t)geo[0]] * sizeof(float),
                                   hipMemcpyDeviceToDevice, s));
    // the plan and everything before this call are on s
    const hipEvent_t ready = record(s);
    for (int c = 1; c < used; ++c) PNT_HIP_CHECK(hipStreamWaitEvent(S[c], ready, 0));
    std::vector<hipEvent_t> done(U.size(), nullptr);
    std::vector<int> last_unit(used, -1);  // per chain, in this phase
    for (size_t i = 0; i < U.size(); ++i) {
      const Unit& u = U[i];
      hipStream_t q = S[u.chain];
      for (int d : u.deps)
        if (U[d].chain != u.chain) PNT_HIP_CHECK(hipStreamWaitEvent(q, done[d], 0));
      for (int64_t k = u.s0; k < u.s1; ++k) launch_stamp(k, q, pick[u.chain]);
      done[i] = record(q);
      last_unit[u.chain] = (int)i;
      if (u.dry_after) {  // join every chain on s, dry, fork again
        for (int c = 1; c < used; ++c)
          if (last_unit[c] >= 0) PNT_HIP_CHECK(hipStreamWaitEvent(s, done[last_unit[c]], 0));
        dry(s);
        const hipEvent_t dried = record(s);
        for (int c = 1; c < used; ++c) PNT_HIP_CHECK(hipStreamWaitEvent(S[c], dried, 0));
        std::fill(last_unit.begin(), last_unit.end(), -1);
      }
    }
    for (int c = 1; c < used; ++c)
      if (last_unit[c] >= 0) PNT_HIP_CHECK(hipStreamWaitEvent(s, done[last_unit[c]], 0));
    finish_geometry();
    // the pickup map of the list's last stroke is the brush's from now on
    PNT_HIP_CHECK(hipMemcpyAsync(b->pick, pick[U.back().chain],
                                 7 * gstride[(size_t)geo[(size_t)n - 1]] * sizeof(float),
                                 hipMemcpyDeviceToDevice, s));
  }
  cv->epoch++;
  if (b->use_snap) {
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  }
}

template <int kT>
void run_fused_t(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n, int chains) {
  const PlanDev D = upload_plan(b, cv, st, n);
  const std::vector<ListGeo> geos{ListGeo{b->fp, b->size_map, b->radius, nullptr, b->width,
                                          FpBox{b->fbox[0], b->fbox[1], b->fbox[2], b->fbox[3]}}};
  run_fused_list<kT>(b, cv, st, n, D, geos, std::vector<int>((size_t)n, 0),
                     std::vector<uint8_t>((size_t)n, 0), chains);
}

// A stroke list: each stroke is FootprintBrush::setRadius(radius) with its
// footprint (a no-op within 0.5 of the current radius, FootprintBrush.hxx:
// 46-50), then its stamps.
template <int kT>
void run_fused_strokes_t(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stroke* S, int64_t ns) {
  std::vector<ListGeo> geos{ListGeo{b->fp, b->size_map, b->radius, nullptr, b->width,
                                    FpBox{b->fbox[0], b->fbox[1], b->fbox[2], b->fbox[3]}}};
  int64_t fp_floats = 0;
  std::vector<int64_t> fp_off{0};
  std::vector<PlanPart> parts;
  std::vector<pnt_stamp> all;
  std::vector<int> geo;
  std::vector<uint8_t> clear;
  int cur = 0;
  double cur_r = b->radius;
  bool pending_clear = false;  // setRadius cleaned the map; the next stamp starts from it
  for (int64_t i = 0; i < ns; ++i) {
    if (!(std::fabs(cur_r - S[i].radius) < 0.5)) {  // a new geometry, a clean map
      const int sm = footprint_size_map(S[i].radius, S[i].footprint, S[i].width);
      fp_off.push_back(fp_floats);
      fp_floats = round_up(fp_floats + (int64_t)sm * sm, 64);
      geos.push_back(ListGeo{nullptr, sm, S[i].radius, S[i].footprint, S[i].width, FpBox{}});
      cur = (int)geos.size() - 1;
      cur_r = S[i].radius;
      pending_clear = true;
    }
    if (S[i].n_stamps == 0) continue;
    clear.push_back(pending_clear ? 1 : 0);
    clear.insert(clear.end(), (size_t)S[i].n_stamps - 1, 0);
    pending_clear = false;
    parts.push_back(PlanPart{geos[(size_t)cur].size_map, geos[(size_t)cur].radius, S[i].stamps,
                             S[i].n_stamps});
    all.insert(all.end(), S[i].stamps, S[i].stamps + S[i].n_stamps);
    geo.insert(geo.end(), (size_t)S[i].n_stamps, cur);
  }
  const int64_t n = (int64_t)all.size();
  if (n > 0) {
    const float* d_fp = nullptr;
    // the padded footprints go straight into the plan's pinned staging
    // buffer, one host thread per few geometries
    const auto fill = [&](float* dst) {
      const int ng = (int)geos.size() - 1;
      const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
      const int nt = std::min({hw, 8, (int)(1 + fp_floats / (1 << 18)), ng});
      std::atomic<int> next{1};
      std::vector<std::exception_ptr> errs((size_t)std::max(nt, 1));
      auto work = [&](int t) {
        try {
          for (int g; (g = next.fetch_add(1)) <= ng;) {
            ListGeo& G = geos[(size_t)g];
            int box[4];
            pad_footprint_into(G.core, G.width, G.size_map, dst + fp_off[(size_t)g], box);
            G.fb = FpBox{box[0], box[1], box[2], box[3]};
          }
        } catch (...) {
          errs[(size_t)t] = std::current_exception();
          next.store(ng + 1);
        }
      };
      if (nt <= 1) {
        work(0);
      } else {
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
      }
      for (auto& e : errs)
        if (e) std::rethrow_exception(e);
    };
    const PlanDev D =
        upload_plan_parts(b, cv, parts.data(), (int)parts.size(), fp_floats, fill, &d_fp);
    for (size_t g = 1; g < geos.size(); ++g) geos[g].fp = d_fp + fp_off[g];
    run_fused_list<kT>(b, cv, all.data(), n, D, geos, geo, clear, 0);
  }
  // a geometry change after the last stamp (strokes without stamps)
  if (cur != (n > 0 ? geo.back() : 0))
    brush_set_geometry(b, geos[(size_t)cur].radius, geos[(size_t)cur].core, geos[(size_t)cur].width);
}

// ------------------------------------------------------------- chained --
// PNT_EXEC_CHAINED: a run of stamps (up to a dryCanvas) in ONE launch.  The
// blocks of stamp k, in order, are: the pickup-map clear of a dip, the ring
// scan, the window tiles, and the shared-pixel levels of a duplicated
// column/row 0.  A block takes a ticket from a counter when it starts and
// becomes block `ticket` of that sequence; it waits (bounded) only on
// counters that blocks with smaller tickets fill:
//   every block of stamp k   : all blocks of stamp k-1 are done
//   tiles of a dip stamp     : the clear blocks of k are done
//   tiles of a dup stamp     : the ring blocks of k are done (no pull)
//   the level block          : the tiles of k are done
// A block with a smaller ticket has started and itself waits only on
// smaller tickets, so the smallest unfinished block always runs: no
// deadlock, without a cooperative launch.  Consecutive stamps follow each
// other at the latency of one counter, not of a kernel launch, and the host
// issues one launch per run.
struct CArgs {
  FArgs f;              // f.st = this run's first stamp
  int n;                // stamps in the run
  const int* start;     // n+1: first ticket of each stamp
  const int4* roles;    // per stamp: clear, ring, tile, level block counts
  const float* paint;   // per stamp: K0 K1 K2 S0 S1 S2 of the brush
  unsigned* cnt;        // 4 per stamp: done, cleared, ring done, tiles done
  unsigned* ticket;
  unsigned* abort;
  unsigned* host_err;
  float* pick;          // pickup map base (7 planes of mstride)
  int64_t pick_floats;
  int long_sleep;       // poll every ~1 us instead of ~0.1 us
};

// Thread 0: bounded wait until *c >= target (relaxed polls, one acquire).
__device__ __forceinline__ bool chain_wait(const CArgs& A, unsigned* c, unsigned target) {
  unsigned spins = 0;
  while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (A.long_sleep)
      __builtin_amdgcn_s_sleep(20);
    else
      __builtin_amdgcn_s_sleep(2);
    if ((++spins & 63) == 0) {
      if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
      if (spins > kSpinLimit) {
        __hip_atomic_store(A.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(A.host_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return false;
      }
    }
  }
  return true;
}

template <int kT>
__global__ __launch_bounds__(kT* kT) void k_stamp_chain(CArgs A) {
  __shared__ int sm[kT + 2][kT + 2];
  __shared__ int s_v, s_ok;
  const int t = threadIdx.x;
  if (t == 0)
    s_v = (int)__hip_atomic_fetch_add(A.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int v = s_v;
  int lo = 0, hi = A.n;  // stamp k: start[k] <= v < start[k+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (A.start[mid] <= v) lo = mid; else hi = mid;
  }
  const int k = lo;
  const int off = v - A.start[k];
  const int4 r = A.roles[k];
  // role: 0 clear, 1 ring, 2 tile, 3 level
  const int role = off < r.x ? 0 : (off < r.x + r.y ? 1 : (off < r.x + r.y + r.z ? 2 : 3));
  unsigned* ck = A.cnt + 4 * k;
  const PStamp& st = A.f.st[k];
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  if (t == 0) {
    bool ok = true;
    if (k > 0) ok = chain_wait(A, A.cnt + 4 * (k - 1), (unsigned)(A.start[k] - A.start[k - 1]));
    if (ok && role == 2 && r.x) ok = chain_wait(A, ck + 1, (unsigned)r.x);
    if (ok && role == 2 && dup && r.y) ok = chain_wait(A, ck + 2, (unsigned)r.y);
    if (ok && role == 3) ok = chain_wait(A, ck + 3, (unsigned)r.z);
    if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  if (!s_ok) return;
  const StampGeom g = geom_of(A.f, st);
  const bool pull = A.f.use_snap && !dup;
  const float* pk = A.paint + 6 * k;
  const float* ps = pk + 3;
  if (role == 0) {  // dip(): clean() the pickup map
    for (int64_t i = (int64_t)off * kT * kT + t; i < A.pick_floats; i += (int64_t)r.x * kT * kT)
      A.pick[i] = 0.0f;
  } else if (role == 1) {
    ring_scan<kRingGroups>(A.f.canvas, A.f.snap, A.f.dirty, A.f.fp, g, st.box, pull,
                 (off - r.x) * kT * kT + t, r.y * kT * kT);
  } else if (role == 2) {
    const VisitTarget T{&A.f.canvas, &A.f.snap, A.f.dirty, A.f.use_snap, A.f.pickup_rate,
                        A.f.dep_rate, A.f.cap};
    float pkr[3] = {pk[0], pk[1], pk[2]}, psr[3] = {ps[0], ps[1], ps[2]};
    window_tile<kT>(T, A.f.pick.p, A.f.fp, A.f.fb, g, st, dup, pull, off - r.x - r.y, A.f.tiles_x, sm,
                    pkr, psr);
  } else {  // visits sharing canvas column/row 0, by dependency level
    const VisitTarget T{&A.f.canvas, &A.f.snap, A.f.dirty, A.f.use_snap, A.f.pickup_rate,
                        A.f.dep_rate, A.f.cap};
    float pkr[3] = {pk[0], pk[1], pk[2]}, psr[3] = {ps[0], ps[1], ps[2]};
    for (int L = 0; L < st.dl_nlev; ++L) {
      const int e0 = A.f.dl_off[st.dl_lev + L], e1 = A.f.dl_off[st.dl_lev + L + 1];
      for (int e = e0 + t; e < e1; e += kT * kT) {
        const int2 vv = A.f.dl[e];
        int64_t ic;
        const int m = stamp_visit(g, vv.x, vv.y, &ic);
        if (m >= 0 && A.f.fp[m] > 0.0f)
          cached_visit(T, GlobalPick{A.f.pick.p, m}, A.f.fp[m], ic, false, pkr, psr);
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
  // signal: every wave's stores drained, then one release and the counters
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (role == 0) __hip_atomic_fetch_add(ck + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (role == 1) __hip_atomic_fetch_add(ck + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (role == 2) __hip_atomic_fetch_add(ck + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(ck, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

// Device and pinned buffers of a brush's chained runs.
struct ChainState {
  char* stage = nullptr;
  int64_t stage_cap = 0;
  char* dev = nullptr;
  int64_t dev_cap = 0;
  hipEvent_t uploaded = nullptr;
};

static ChainState& chain_state(pnt_brush_s* b) {
  if (!b->chain) {
    auto* c = new ChainState;
    PNT_HIP_CHECK(hipEventCreateWithFlags(&c->uploaded, hipEventDisableTiming));
    b->chain = c;
  }
  return *b->chain;
}

void chain_release(ChainState*& c) {
  if (!c) return;
  if (c->stage) (void)hipHostFree(c->stage);
  if (c->dev) (void)hipFree(c->dev);
  if (c->uploaded) (void)hipEventDestroy(c->uploaded);
  delete c;
  c = nullptr;
}

template <int kT>
static void run_chained_t(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  hipStream_t s = b->ctx->stream;
  if (*(volatile unsigned*)b->ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier stamp kernel aborted");
  const PlanDev D = upload_plan(b, cv, st, n);
  FArgs F;
  F.st = D.st;
  F.canvas = cv->layer();
  F.snap = b->use_snap ? b->snap_planes() : cv->layer();
  F.pick = b->pick_planes();
  F.fp = b->fp;
  F.fb = FpBox{b->fbox[0], b->fbox[1], b->fbox[2], b->fbox[3]};
  F.dirty = b->dirty;
  F.dl = D.dl;
  F.dl_off = D.dl_off;
  F.use_snap = b->use_snap ? 1 : 0;
  F.rows = cv->rows;
  F.cols = cv->cols;
  F.hr = (b->size_map - 1) / 2;
  F.size_map = b->size_map;
  const int side = 2 * F.hr + 1;
  F.tiles_x = (side + kT - 1) / kT;
  F.ntiles = F.tiles_x * F.tiles_x;
  F.pickup_rate = b->pickup_rate;
  F.dep_rate = b->dep_rate;
  F.cap = b->cap;
  F.clk = DryClock{};  // a canvas with a drying clock stamps on the fused path
  // host plan of the whole list: roles, tickets, paints
  std::vector<int4> roles(n);
  std::vector<int> start(n + 1);
  std::vector<float> paint(6 * (size_t)n);
  const int64_t pick_floats = 7 * b->mstride;
  const int nclear = (int)std::min<int64_t>((pick_floats + kT * kT * 4 - 1) / (kT * kT * 4), 256);
  float pk[3] = {b->paintK[0], b->paintK[1], b->paintK[2]};
  float ps[3] = {b->paintS[0], b->paintS[1], b->paintS[2]};
  for (int64_t k = 0; k < n; ++k) {
    const PStamp& p = D.host[k];
    if (st[k].flags & PNT_STAMP_DIP)
      for (int c = 0; c < 3; ++c) {
        pk[c] = st[k].K[c];
        ps[c] = st[k].S[c];
      }
    for (int c = 0; c < 3; ++c) {
      paint[6 * k + c] = pk[c];
      paint[6 * k + 3 + c] = ps[c];
    }
    int4 r;
    r.x = (st[k].flags & PNT_STAMP_DIP) ? nclear : 0;
    r.y = F.use_snap ? ring_blocks_for(p.box, kT * kT) : 0;
    r.z = F.ntiles;
    r.w = (p.flags & (kStDupX | kStDupY)) && p.dl_nlev > 0 ? 1 : 0;
    roles[k] = r;
  }
  ChainState& B = chain_state(b);
  // runs up to each dryCanvas
  int64_t i = 0;
  while (i < n) {
    int64_t e = i;
    while (e < n - 1 && !(st[e].flags & PNT_STAMP_DRY_AFTER)) ++e;
    const int64_t cnt = e - i + 1;
    start[0] = 0;
    for (int64_t k = 0; k < cnt; ++k) {
      const int4 r = roles[i + k];
      start[k + 1] = start[k] + r.x + r.y + r.z + r.w;
    }
    const int64_t total_blocks = start[cnt];
    auto align = [](int64_t x) { return (x + 255) / 256 * 256; };
    const int64_t o_roles = 0;
    const int64_t o_start = align(o_roles + cnt * (int64_t)sizeof(int4));
    const int64_t o_paint = align(o_start + (cnt + 1) * (int64_t)sizeof(int));
    const int64_t o_cnt = align(o_paint + 6 * cnt * (int64_t)sizeof(float));
    const int64_t n_cnt = 4 * cnt + 2;
    const int64_t bytes = align(o_cnt + n_cnt * (int64_t)sizeof(unsigned));
    PNT_HIP_CHECK(hipEventSynchronize(B.uploaded));
    if (bytes > B.stage_cap) {
      if (B.stage) PNT_HIP_CHECK(hipHostFree(B.stage));
      B.stage = nullptr;
      B.stage_cap = std::max<int64_t>(bytes, 2 * B.stage_cap);
      PNT_HIP_CHECK(hipHostMalloc((void**)&B.stage, (size_t)B.stage_cap));
    }
    if (bytes > B.dev_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      if (B.dev) PNT_HIP_CHECK(hipFree(B.dev));
      B.dev = nullptr;
      B.dev_cap = std::max<int64_t>(bytes, 2 * B.dev_cap);
      PNT_HIP_CHECK(hipMalloc((void**)&B.dev, (size_t)B.dev_cap));
    }
    std::memcpy(B.stage + o_roles, roles.data() + i, cnt * sizeof(int4));
    std::memcpy(B.stage + o_start, start.data(), (cnt + 1) * sizeof(int));
    std::memcpy(B.stage + o_paint, paint.data() + 6 * i, 6 * cnt * sizeof(float));
    PNT_HIP_CHECK(hipMemcpyAsync(B.dev, B.stage, (size_t)o_cnt, hipMemcpyHostToDevice, s));
    PNT_HIP_CHECK(hipEventRecord(B.uploaded, s));
    unsigned* counters = reinterpret_cast<unsigned*>(B.dev + o_cnt);
    launch_fill_u32(counters, n_cnt, 0u, s);
    CArgs A;
    A.f = F;
    A.f.st = D.st + i;
    A.n = (int)cnt;
    A.start = reinterpret_cast<const int*>(B.dev + o_start);
    A.roles = reinterpret_cast<const int4*>(B.dev + o_roles);
    A.paint = reinterpret_cast<const float*>(B.dev + o_paint);
    A.cnt = counters;
    A.ticket = counters + 4 * cnt;
    A.abort = counters + 4 * cnt + 1;
    A.pick = b->pick;
    A.pick_floats = pick_floats;
    A.long_sleep = std::getenv("PAINTY_CHAIN_SLEEP") != nullptr;
    PNT_HIP_CHECK(hipHostGetDevicePointer((void**)&A.host_err, b->ctx->kernel_error, 0));
    PNT_REQUIRE(total_blocks < ((int64_t)1 << 31), "stamp run too long for one launch");
    k_stamp_chain<kT><<<(unsigned)total_blocks, kT * kT, 0, s>>>(A);
    PNT_HIP_CHECK(hipGetLastError());
    if (st[e].flags & PNT_STAMP_DRY_AFTER) {
      if (b->use_snap)
        launch_dry_dirty(cv->n, cv->layer(), cv->R0(), cv->h(), b->dirty, s);
      else
        launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), s);
    }
    i = e + 1;
  }
  for (int c = 0; c < 3; ++c) {  // brush paint after the list: its last dip
    b->paintK[c] = pk[c];
    b->paintS[c] = ps[c];
  }
  cv->epoch++;
  if (b->use_snap) {
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  }
}

void run_chained(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  if (n == 0) return;
  // 32x32 tiles: every block pays one agent-scope release and acquire (an
  // L2 writeback / invalidate per XCD), so fewer, larger blocks are cheaper
  const char* e = std::getenv("PAINTY_CHAIN_TILE");
  if (e && std::atoi(e) == 16)
    run_chained_t<16>(b, cv, st, n);
  else
    run_chained_t<32>(b, cv, st, n);
}

// ------------------------------------------------------ no-allocation call --
// pnt_imprint_stamps_noalloc: the stamps of a list in stream order on the
// brush's own pickup map (PNT_EXEC_FUSED, one chain), with the plan -- the
// stamp descriptors and shared-pixel visit lists the kernels read -- in a
// file-scope device array instead of library-allocated buffers, so the call
// allocates no device or pinned memory.  Each GPU holds its own copy of the
// array; its address is looked up once per device with hipGetSymbolAddress
// and kept here, and only that address is handed to kernels and copies.
constexpr int64_t kScratchBytes = 4 << 20;
constexpr int64_t kScratchStampBytes = 3 << 20;  // stamp descriptors; visit lists after
__device__ __align__(256) char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
char* g_scratch_ptr[64] = {};  // per device: the array's device address
}  // namespace

// The device address of `scratch` on the current device (looked up on the
// first call there).
static char* scratch_on(int device) {
  PNT_REQUIRE(device >= 0 && device < 64, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    PNT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  return g_scratch_ptr[device];
}

void run_fused_noalloc(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  hipStream_t s = b->ctx->stream;
  char* dev = scratch_on(b->ctx->device);
  // host plan (thread-local buffers keep their capacity between calls)
  thread_local std::vector<PStamp> h_st;
  thread_local std::vector<int2> h_dl;
  thread_local std::vector<int> h_off;
  h_st.clear();
  h_dl.clear();
  h_off.clear();
  plan_stamps(b, cv, st, n, h_st, h_dl, h_off);
  const int64_t bst = (int64_t)(h_st.size() * sizeof(PStamp));
  const int64_t bdl = (int64_t)(h_dl.size() * sizeof(int2));
  const int64_t boff = (int64_t)(h_off.size() * sizeof(int));
  PNT_REQUIRE(bst <= kScratchStampBytes,
              "stamp list too long for the no-allocation call (at most 24576 stamps)");
  PNT_REQUIRE(round_up(bdl, 256) + boff <= kScratchBytes - kScratchStampBytes,
              "shared-pixel visit lists too long for the no-allocation call");
  // the brush's snapshot and dirty bytes must already fit this canvas
  if (b->use_snap) {
    PNT_REQUIRE(b->dirty && round_up(cv->n, 16) + 16 <= b->dirty_cap &&
                    b->snap && kLayerStrides * round_up(std::max<int64_t>(cv->n, 1), 64) <= b->snap_cap,
                "the brush's snapshot buffers are not sized for this canvas (no-allocation "
                "call); stamp this canvas once with pnt_imprint_stamps first");
    brush_prepare_snapshot(b, cv, s);  // copies or marks; allocates nothing at these sizes
  }
  // the plan into the array, then the kernels read it (the copy is waited for
  // before returning control of the host buffers)
  if (bst) PNT_HIP_CHECK(hipMemcpyAsync(dev, h_st.data(), (size_t)bst, hipMemcpyHostToDevice, s));
  char* ldl = dev + kScratchStampBytes;
  char* loff = ldl + round_up(bdl, 256);
  if (bdl) PNT_HIP_CHECK(hipMemcpyAsync(ldl, h_dl.data(), (size_t)bdl, hipMemcpyHostToDevice, s));
  if (boff)
    PNT_HIP_CHECK(hipMemcpyAsync(loff, h_off.data(), (size_t)boff, hipMemcpyHostToDevice, s));
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  FArgs a;
  a.st = reinterpret_cast<const PStamp*>(dev);
  a.canvas = cv->layer();
  a.snap = b->use_snap ? b->snap_planes() : cv->layer();
  a.pick = b->pick_planes();
  a.fp = b->fp;
  a.fb = FpBox{b->fbox[0], b->fbox[1], b->fbox[2], b->fbox[3]};
  a.dirty = b->dirty;
  a.dl = reinterpret_cast<const int2*>(ldl);
  a.dl_off = reinterpret_cast<const int*>(loff);
  a.use_snap = b->use_snap ? 1 : 0;
  a.rows = cv->rows;
  a.cols = cv->cols;
  a.hr = (b->size_map - 1) / 2;
  a.size_map = b->size_map;
  const int side = 2 * a.hr + 1;
  a.tiles_x = (side + 15) / 16;
  a.ntiles = a.tiles_x * a.tiles_x;
  a.pickup_rate = b->pickup_rate;
  a.dep_rate = b->dep_rate;
  a.cap = b->cap;
  a.clk = cv->clock != PNT_CLOCK_OFF && cv->tmap
              ? DryClock{cv->tmap, 0, cv->drying_ms, cv->R0(), cv->h()}
              : DryClock{};
  for (int64_t k = 0; k < n; ++k) {
    if (st[k].flags & PNT_STAMP_DIP) {  // dip(): clean() and the new paint
      launch_fill(b->pick, 7 * b->mstride, 0.0f, s);
      for (int c = 0; c < 3; ++c) {
        b->paintK[c] = st[k].K[c];
        b->paintS[c] = st[k].S[c];
      }
    }
    for (int c = 0; c < 3; ++c) {
      a.pk[c] = b->paintK[c];
      a.ps[c] = b->paintS[c];
    }
    const PStamp& p = h_st[(size_t)k];
    const int rb = a.use_snap ? ring_blocks_for(p.box, 256) : 0;
    const int jk = (int)k;
    k_stamp_fused<16><<<a.ntiles + rb + (dup_levels(p) ? 1 : 0), 256, 0, s>>>(a, jk, rb);
    PNT_HIP_CHECK(hipGetLastError());
    if (st[k].flags & PNT_STAMP_DRY_AFTER) {
      if (b->use_snap)
        launch_dry_dirty(cv->n, cv->layer(), cv->R0(), cv->h(), b->dirty, s);
      else
        launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), s);
      if (a.clk.tmap) launch_fill_i64(cv->tmap, cv->n, canvas_now_us(cv), s);
    }
  }
  cv->epoch++;
  if (b->use_snap) {
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  }
}

void run_fused_strokes(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stroke* strokes, int64_t ns) {
  int kT = 16;
  if (const char* e = std::getenv("PAINTY_FUSED_TILE")) kT = std::atoi(e) == 32 ? 32 : 16;
  if (kT == 32)
    run_fused_strokes_t<32>(b, cv, strokes, ns);
  else
    run_fused_strokes_t<16>(b, cv, strokes, ns);
}

void run_fused(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n, int chains) {
  if (n == 0) return;
  // window tile side: 16 (256 threads), or 32 from PAINTY_FUSED_TILE
  int kT = 16;
  if (const char* e = std::getenv("PAINTY_FUSED_TILE")) kT = std::atoi(e) == 32 ? 32 : 16;
  if (kT == 32)
    run_fused_t<32>(b, cv, st, n, chains);
  else
    run_fused_t<16>(b, cv, st, n, chains);
}

}  // namespace pnt
