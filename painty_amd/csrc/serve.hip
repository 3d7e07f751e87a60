// A request server for stroke + composite work on resident canvases, with
// continuous batching on the device.
//
// Requests (a stamp list and a composite on a canvas identified by an id)
// are submitted from any thread and return a ticket.  Each request runs as
// the reference's fresh FootprintBrush would (renderer/FootprintBrush.hxx:
// 29-63): pickup map clean, snapshot copied whole from the canvas before its
// first imprint (:281-284), its stamps (FootprintBrush::imprint, :73-143)
// with dryCanvas() where a stamp asks for it (Canvas.hxx:105-121), then
// Renderer::compose (Renderer.hxx:48-53) into the canvas' resident composite
// planes, copied to the caller's host buffer when one is given.
//
// Execution.  Planner threads plan each request on arrival (FP64 stamp
// geometry, snapshot ring boxes, shared-pixel levels; its stamps cut into
// units -- runs of one stroke -- with the units they depend on; see
// plan_units in fused.hip).  One engine thread admits planned requests into
// up to n_slots request slots and drives the device with "waves": each wave
// is ONE launch of k_serve_wave, carrying the next phase of every ready unit
// of every admitted request of its lane (up to kPhases phases: a part of a
// canvas initialisation or of a full snapshot copy, a pickup-map clear, a
// stamp -- window tiles, ring scan and any shared-pixel levels side by side
// -- a part of a dryCanvas, or a part of the composite), oldest request
// first.  Requests that arrive while others run join the next wave: their
// work overlaps on the device, in the same launches.  Two lanes, each a
// stream of waves: the main one, and an express one for a few requests
// whose stamps form a long dependent chain, so that such a chain advances in
// small, short waves.  Phases in one wave touch disjoint memory (different
// canvases, or units of one canvas with disjoint snapshot boxes and their
// own pickup maps); a request's waves are all on its lane's stream, in
// order, and one canvas has one request running at a time, so each request
// computes exactly the arithmetic of running alone: the same to the bit.  At
// most kAhead waves are queued per lane ahead of the host, so a new request
// waits for few.
//
// Admission: requests on one canvas run in submission order, one at a time;
// a request whose canvas is not resident is admitted when its canvas fits
// the HBM budget (canvas + composite planes, 60 B/px), after evicting
// canvases that no queued or running request needs, least recently used
// first.  An evicted canvas is spilled to host memory by the copier thread,
// and restored from there when a later request names it.  A request slot's
// working memory (snapshot 32 B/px, dirty bytes 1 B/px, pickup maps, plan)
// is kept by the slot for the next request and is not counted in the budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"
#include "wave.hpp"

namespace pnt {
int guard_call(const std::function<void()>& f);  // api.hip: status + pnt_last_error
}

using namespace pnt;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kPhases = 160;       // phases per wave
constexpr int kChainsPerReq = 8;   // pickup maps of one request (concurrent units)
constexpr int kAhead = 3;          // waves queued ahead of the engine
constexpr int kT = kWT;            // window tile side
constexpr int kT2 = kWT2;          // threads per block

// What every phase of one admitted request shares: written to the device
// once, when the request is set up in its slot.
struct SReq {
  const PStamp* st;
  float* canvas;
  float* snap;      // the canvas itself when the snapshot is off
  uint8_t* dirty;
  float* pick;      // chain c's map at pick + c * 7 * mstride
  const float* fp;
  const int2* dl;
  const int* dl_off;
  const float* paint;  // 6 floats per stamp: the brush paint at that stamp
  float* out[3];
  int cstride, mstride, dirty_bytes, rows, cols, hr, size_map, tiles_x, ntiles, use_snap;
  int fb[4];
  float bg[3];
  int pad;
  double pickup_rate, dep_rate, cap;
};

// SReq as the kernel reads it: the same layout, device pointers global.
struct SReqG {
  gptr<const PStamp> st;
  gptr<float> canvas;
  gptr<float> snap;
  gptr<uint8_t> dirty;
  gptr<float> pick;
  gptr<const float> fp;
  gptr<const int2> dl;
  gptr<const int> dl_off;
  gptr<const float> paint;
  gptr<float> out[3];
  int cstride, mstride, dirty_bytes, rows, cols, hr, size_map, tiles_x, ntiles, use_snap;
  int fb[4];
  float bg[3];
  int pad;
  double pickup_rate, dep_rate, cap;
};
static_assert(sizeof(SReqG) == sizeof(SReq) && offsetof(SReqG, out) == offsetof(SReq, out) &&
                  offsetof(SReqG, cstride) == offsetof(SReq, cstride) &&
                  offsetof(SReqG, pickup_rate) == offsetof(SReq, pickup_rate),
              "SReqG mirrors SReq");

// One phase of a wave (a kernel argument): its request slot, its kind, the
// stamp it runs and the blocks it takes.
struct SPhase {
  int16_t slot;
  int8_t kind;
  int8_t chain;
  int k;      // stamp index (stamp kinds)
  int start;  // first block
  int nblk;
  int rb;     // ring-scan blocks (stamp kinds); part | nparts << 16 (streaming kinds)
};

struct SArgs {
  const SReqG* reqs;
  int nphases;
  int pad;
  SPhase ph[kPhases];
};
static_assert(sizeof(SArgs) <= 4000, "wave arguments must fit the kernel argument segment");

// One wave of the server: its phases arrive as kernel arguments, each
// request's shared fields from its device descriptor.
__global__ __launch_bounds__(kWT2, kWaveOccupancy) void k_serve_wave(SArgs A) {
  __shared__ int sm[kWT + 2][kWT + 2];
  int lo = 0, hi = A.nphases - 1;  // the last phase with start <= blockIdx.x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.ph[mid].start <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SPhase& P = A.ph[lo];
  const int b = (int)blockIdx.x - P.start;
  if (b >= P.nblk) return;
  const SReqG& Q = A.reqs[P.slot];
  SSlot S;
  S.st = (const PStamp*)(Q.st + P.k);
  S.canvas = (float*)Q.canvas;
  S.snap = (float*)Q.snap;
  S.dirty = (uint8_t*)Q.dirty;
  S.pick = (float*)(Q.pick + (int64_t)P.chain * 7 * Q.mstride);
  S.pick_dst = S.pick;
  S.fp = (const float*)Q.fp;
  S.dl = (const int2*)Q.dl;
  S.dl_off = (const int*)Q.dl_off;
#pragma unroll
  for (int c = 0; c < 3; ++c) S.out[c] = (float*)Q.out[c];
  S.cstride = Q.cstride;
  S.mstride = Q.mstride;
  S.dirty_bytes = Q.dirty_bytes;
  S.rows = Q.rows;
  S.cols = Q.cols;
  S.kind = P.kind;
  S.start = P.start;
  S.nblk = P.nblk;
  S.ntiles = Q.ntiles;
  // a streaming phase carries its part in rb: part | nparts << 16
  const bool streaming =
      P.kind == kSInit || P.kind == kSSnap || P.kind == kSDry || P.kind == kSCompose;
  S.rb = streaming ? 0 : P.rb;
  S.part = streaming ? (P.rb & 0xffff) : 0;
  S.nparts = streaming ? (P.rb >> 16) : 1;
  S.hr = Q.hr;
  S.size_map = Q.size_map;
  S.tiles_x = Q.tiles_x;
#pragma unroll
  for (int c = 0; c < 4; ++c) S.fb[c] = Q.fb[c];
  const bool stamp_kind = P.kind == kSStamp;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    S.pk[c] = stamp_kind ? Q.paint[6 * (int64_t)P.k + c] : 0.0f;
    S.ps[c] = stamp_kind ? Q.paint[6 * (int64_t)P.k + 3 + c] : 0.0f;
  }
  S.pickup_rate = Q.pickup_rate;
  S.dep_rate = Q.dep_rate;
  S.cap = Q.cap;
  S.use_snap = Q.use_snap;
#pragma unroll
  for (int c = 0; c < 3; ++c) S.bg[c] = Q.bg[c];
  S.tmap = nullptr;  // served canvases have no drying clock
  S.dry_ms = 0;
  run_wave_slot(S, b, sm);
}

// Blocks of one part of a streaming phase: at most 512, each thread taking
// several items, so that a part leaves room in its wave.
int part_blocks(int64_t items) {
  const int64_t b = (items + kT2 - 1) / kT2;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 512);
}

int64_t canvas_floats(int rows, int cols) {
  return (kCanvasStrides + 3) * round_up((int64_t)rows * cols, 64);  // canvas + composite planes
}
int64_t canvas_bytes(int rows, int cols) { return canvas_floats(rows, cols) * 4; }
int64_t dirty_bytes_for(int64_t n) { return round_up(n, 16) + 16; }

struct Phase {
  int64_t k;
  int kind;
};

struct Resident;

enum Stage : int { kPrologue = 0, kUnits, kDrying, kComposite, kIssued };

struct Req {
  int64_t ticket = 0;
  pnt_request r{};
  std::vector<pnt_stamp> stamps;
  std::vector<float> footprint;
  Clock::time_point t_submit;
  double latency_ms = 0.0;
  int status = PNT_OK;
  std::string error;
  bool done = false;
  // plan (planner thread)
  bool planned = false;
  int size_map = 0, hr = 0, tiles_x = 0, ntiles = 0;
  int64_t mstride = 0;
  int fbox[4] = {0, 0, -1, -1};
  std::vector<float> fp;  // padded footprint
  std::vector<PStamp> st;
  std::vector<int2> dl;
  std::vector<int> off;
  std::vector<Unit> U;
  std::vector<std::vector<Phase>> ph;  // per unit
  std::vector<int> prev_on_chain;      // per unit
  std::vector<float> paint;            // 6 per stamp: the brush paint at that stamp
  std::vector<int> rb;                 // ring-scan blocks per stamp
  int64_t plan_bytes = 0;
  // engine
  int slot = -1;
  Resident* res = nullptr;
  bool fresh = false;  // a new canvas: initialised by the prologue
  int stage = kPrologue;
  int lane = 0;              // 0: the main wave stream, 1: the express one
  int64_t alone_waves = 0;   // waves the request's stamps take when it runs alone
  int part = 0;           // next part of the current streaming phase
  size_t g0 = 0, g1 = 0;  // units of the current drying phase
  std::vector<size_t> next;
  std::vector<char> udone;
  std::vector<int64_t> sdone;  // stamps of each unit issued
  StampNeeds needs;
  const PStamp* d_st = nullptr;
  const float* d_paint = nullptr;
  const int2* d_dl = nullptr;
  const int* d_off = nullptr;
  const float* d_fp = nullptr;
};

enum Where : int { kResident = 0, kSpilling, kSpilled, kRestoring };

struct Resident {
  float* block = nullptr;  // canvas planes (kCanvasStrides x stride) then composite (3 x stride)
  int64_t cap = 0;         // floats of block
  int rows = 0, cols = 0;
  int64_t stride = 0;
  float bg[3] = {1.f, 1.f, 1.f};
  bool initialised = false;
  int active = 0;  // admitted requests not yet complete
  int where = kResident;
  std::vector<float> host;     // spilled canvas planes (kCanvasStrides x stride)
  pnt_canvas view = nullptr;   // handle on the query context, made on demand
  Clock::time_point last_use;
};

struct SlotRes {
  float* snap = nullptr;
  int64_t snap_cap = 0;  // floats
  uint8_t* dirty = nullptr;
  int64_t dirty_cap = 0;
  float* pick = nullptr;
  int64_t pick_cap = 0;  // floats
  char* plan = nullptr;
  int64_t plan_cap = 0;
  char* stage = nullptr;  // pinned
  int64_t stage_cap = 0;
  hipEvent_t fin = nullptr;  // after the request's composite
  Req* req = nullptr;
};

struct CopyJob {
  int kind;  // 0: composite to the caller, 1: spill, 2: restore
  Req* q = nullptr;
  int64_t canvas_id = 0;
};

}  // namespace

struct pnt_server_s {
  int device = 0;
  int64_t budget = 0;  // bytes
  int64_t held = 0;    // bytes of canvas blocks held: resident + pooled + in transit
  // A streaming phase (canvas initialisation, snapshot copy, dryCanvas,
  // composite) is cut into parts of about part_bytes of traffic, one part
  // per wave, so that no wave -- and no other request's next stamp -- waits
  // on the whole of a large canvas.
  int64_t part_bytes = 64ll << 20;
  std::multimap<int64_t, float*> pool;  // idle blocks by capacity (floats)
  std::mutex mu;
  std::condition_variable cv_engine, cv_done, cv_plan, cv_copy;
  std::deque<Req*> planq;   // to plan
  std::deque<Req*> queue;   // submitted, in order, not yet admitted
  std::unordered_map<int64_t, Req*> tickets;
  std::unordered_map<int64_t, Resident> canvases;
  std::unordered_map<int64_t, int> pending;  // queued requests per canvas
  std::unordered_set<int64_t> closing;       // released once idle (pnt_server_release)
  std::deque<CopyJob> copyq;
  std::vector<SlotRes> slots;
  std::vector<int> free_slots;
  bool stopping = false;
  int64_t completed = 0;
  int64_t ticket_seq = 1;
  int spills = 0;  // spills in progress
  // engine-thread state
  std::vector<Req*> active;  // admitted, in admission order
  pnt_ctx ctx = nullptr;     // engine stream
  pnt_ctx qctx = nullptr;    // views of resident canvases (pnt_server_canvas)
  pnt_ctx cctx = nullptr;    // copier stream
  SReq* d_reqs = nullptr;    // one descriptor per request slot (device)
  // Two wave streams ("lanes"): the main one carries the ready phases of
  // most requests; the express one those of up to express_max requests
  // whose stamps form a long dependent chain (at least express_waves waves
  // alone).  A long chain advances one stamp per wave it is in; in the
  // express lane its waves are small and short, instead of each waiting on
  // everything else in flight.  Each lane keeps at most kAhead waves queued.
  struct Lane {
    pnt_ctx ctx = nullptr;
    hipEvent_t ev[kAhead + 1] = {};
    int64_t waves = 0;
  } lanes[2];
  int express_max = 4, express_active = 0;
  int64_t express_waves = 200;
  std::thread engine, copier;
  std::vector<std::thread> planners;

  void plan(Req* q);
  void planner();
  bool reserve(int64_t need, float** block, int64_t* cap);
  bool evict_one();
  std::vector<Req*> admit();
  void setup(Req* q);
  bool collect(Req* q, SArgs& A, int& start, std::vector<std::pair<Req*, int>>& used);
  void write_descriptor(Req* q, int64_t& o);
  void after_wave(const std::vector<std::pair<Req*, int>>& used);
  void complete(Req* q);
  void finish(Req* q, int status, const std::string& err);
  void release_if_closing(int64_t id);
  void run_engine();
  void run_copier();
  void copy_job(const CopyJob& j);
};

// ------------------------------------------------------------- planning --
static int64_t alone_waves(const Req& q);

void pnt_server_s::plan(Req* q) {
  const pnt_request& r = q->r;
  q->size_map = footprint_size_map(r.radius, q->footprint.data(), r.width);
  q->hr = (q->size_map - 1) / 2;
  const int side = 2 * q->hr + 1;
  q->tiles_x = (side + kT - 1) / kT;
  q->ntiles = q->tiles_x * q->tiles_x;
  q->mstride = round_up((int64_t)q->size_map * q->size_map, 64);
  q->fp.assign((size_t)q->size_map * q->size_map, 0.0f);
  pad_footprint_into(q->footprint.data(), r.width, q->size_map, q->fp.data(), q->fbox);
  const int64_t n = (int64_t)q->stamps.size();
  plan_stamps_rc(q->size_map, r.radius, r.rows, r.cols, q->stamps.data(), n, q->st, q->dl,
                 q->off);
  std::vector<uint8_t> clear((size_t)n, 0);
  if (n) clear[0] = 1;  // a fresh brush starts from a clean pickup map
  q->U = n ? plan_units(q->stamps.data(), q->st.data(), clear.data(), n, kChainsPerReq)
           : std::vector<Unit>{};
  const bool snap = r.use_snapshot != 0;
  q->ph.assign(q->U.size(), {});
  q->rb.assign((size_t)n, 0);
  for (size_t i = 0; i < q->U.size(); ++i)
    for (int64_t k = q->U[i].s0; k < q->U[i].s1; ++k) {
      if (clear[(size_t)k] || (q->stamps[(size_t)k].flags & PNT_STAMP_DIP))
        q->ph[i].push_back(Phase{k, kSClear});
      const PStamp& p = q->st[(size_t)k];
      q->rb[(size_t)k] = snap ? ring_blocks_for(p.box, kT2) : 0;
      q->ph[i].push_back(Phase{k, kSStamp});
    }
  q->prev_on_chain.assign(q->U.size(), -1);
  {
    std::vector<int> last(kChainsPerReq, -1);
    for (size_t i = 0; i < q->U.size(); ++i) {
      q->prev_on_chain[i] = last[(size_t)q->U[i].chain];
      last[(size_t)q->U[i].chain] = (int)i;
    }
  }
  q->paint.assign(6 * (size_t)n, 0.0f);
  float cur[6] = {0, 0, 0, 0, 0, 0};  // a fresh brush holds no paint
  for (int64_t k = 0; k < n; ++k) {
    if (q->stamps[(size_t)k].flags & PNT_STAMP_DIP)
      for (int c = 0; c < 3; ++c) {
        cur[c] = q->stamps[(size_t)k].K[c];
        cur[3 + c] = q->stamps[(size_t)k].S[c];
      }
    std::memcpy(&q->paint[6 * (size_t)k], cur, sizeof cur);
  }
  auto al = [](int64_t x) { return round_up(x, 256); };
  q->plan_bytes = al((int64_t)(q->st.size() * sizeof(PStamp))) +
                  al((int64_t)(q->dl.size() * sizeof(int2))) +
                  al((int64_t)(q->off.size() * sizeof(int))) +
                  al((int64_t)(q->fp.size() * sizeof(float))) +
                  al((int64_t)(q->paint.size() * sizeof(float))) + al((int64_t)sizeof(SReq));
  q->next.assign(q->U.size(), 0);
  q->udone.assign(q->U.size(), 0);
  q->sdone.assign(q->U.size(), 0);
  q->needs = plan_stamp_needs(q->U, q->st.data());
  for (size_t i = 0; i < q->U.size(); ++i)
    if (q->ph[i].empty()) q->udone[i] = 1;
  q->g0 = q->g1 = 0;  // the first drying phase: units [0, g1]
  while (q->g1 + 1 < q->U.size() && !q->U[q->g1].dry_after) ++q->g1;
  q->alone_waves = alone_waves(*q);
}

// Waves the request's units take when it runs alone: every wave takes the
// next phase of every ready unit, as the engine does, and a dry is one more.
static int64_t alone_waves(const Req& q) {
  const size_t nu = q.U.size();
  std::vector<size_t> next(nu, 0);
  std::vector<char> done(q.udone);
  std::vector<int64_t> sdone(nu, 0);
  std::vector<int> ready;
  int64_t waves = 0;
  for (size_t g0 = 0; g0 < nu;) {
    size_t g1 = g0;
    while (g1 + 1 < nu && !q.U[g1].dry_after) ++g1;
    for (;;) {
      ready.clear();
      for (size_t i = g0; i <= g1; ++i) {
        if (done[i]) continue;
        bool ok = q.prev_on_chain[i] < 0 || done[(size_t)q.prev_on_chain[i]];
        const Phase& f = q.ph[i][next[i]];
        if (ok && f.kind == kSStamp)
          for (int e = q.needs.off[(size_t)f.k]; e < q.needs.off[(size_t)f.k + 1] && ok; ++e) {
            const int2 nd = q.needs.need[(size_t)e];
            ok = done[(size_t)nd.x] || sdone[(size_t)nd.x] >= nd.y;
          }
        if (ok) ready.push_back((int)i);
      }
      if (ready.empty()) break;
      ++waves;
      for (int i : ready) {
        const Phase& f = q.ph[(size_t)i][next[(size_t)i]];
        if (f.kind == kSStamp) sdone[(size_t)i] = f.k - q.U[(size_t)i].s0 + 1;
        if (++next[(size_t)i] == q.ph[(size_t)i].size()) done[(size_t)i] = 1;
      }
    }
    if (q.U[g1].dry_after) ++waves;
    g0 = g1 + 1;
  }
  return waves;
}

void pnt_server_s::planner() {
  (void)hipSetDevice(device);
  for (;;) {
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_plan.wait(lk, [&] { return stopping || !planq.empty(); });
      if (planq.empty()) return;
      q = planq.front();
      planq.pop_front();
    }
    const int st = guard_call([&] { plan(q); });
    std::lock_guard<std::mutex> lk(mu);
    if (st != PNT_OK) {
      q->status = st;
      q->error = pnt_last_error();
    }
    q->planned = true;
    cv_engine.notify_all();
  }
}

// ------------------------------------------------------------ admission --
// A block of at least `need` floats for a canvas (caller holds mu): an idle
// pooled block of close size, or room for a new one within the budget --
// spilling idle canvases, least recently used first, and releasing pooled
// blocks that fit nothing.  False when nothing can give way now (a spill in
// progress returns its block to the pool when its copy is done).
bool pnt_server_s::reserve(int64_t need, float** block, int64_t* cap) {
  for (;;) {
    auto it = pool.lower_bound(need);
    if (it != pool.end() && it->first <= 2 * need) {
      *cap = it->first;
      *block = it->second;
      pool.erase(it);
      return true;
    }
    if (held + need * 4 <= budget) {
      *cap = need;
      *block = nullptr;  // allocated by the engine
      held += need * 4;
      return true;
    }
    if (!pool.empty()) {  // release the largest idle block
      auto big = std::prev(pool.end());
      dev_free(big->second);
      held -= big->first * 4;
      pool.erase(big);
      continue;
    }
    if (!evict_one()) return false;
  }
}

// Starts spilling the least recently used idle canvas (no queued or running
// request) to host memory; its block goes to the pool when the copy is done.
// Caller holds mu.  False when no canvas can be evicted now.
bool pnt_server_s::evict_one() {
  if (spills > 0) return false;  // one at a time: its block may be enough
  int64_t victim = -1;
  Clock::time_point oldest = Clock::time_point::max();
  for (auto& kv : canvases) {
    const Resident& R = kv.second;
    if (R.where != kResident || R.active || pending.count(kv.first) || !R.block) continue;
    if (R.last_use < oldest) {
      oldest = R.last_use;
      victim = kv.first;
    }
  }
  if (victim < 0) return false;
  canvases[victim].where = kSpilling;
  ++spills;
  copyq.push_back(CopyJob{1, nullptr, victim});
  cv_copy.notify_all();
  return false;  // the block arrives in the pool later
}

// Admits queued requests, in submission order, while slots and memory allow
// (caller holds mu).  A request waits while an earlier request on its canvas
// is queued or running.
std::vector<Req*> pnt_server_s::admit() {
  std::vector<Req*> out;
  std::vector<int64_t> seen;
  for (auto it = queue.begin(); it != queue.end();) {
    Req* q = *it;
    const int64_t id = q->r.canvas_id;
    if (std::find(seen.begin(), seen.end(), id) != seen.end()) {
      ++it;
      continue;
    }
    seen.push_back(id);
    if (!q->planned) {
      ++it;
      continue;
    }
    if (q->status != PNT_OK) {  // planning failed
      it = queue.erase(it);
      if (--pending[id] == 0) pending.erase(id);
      finish(q, q->status, q->error);
      continue;
    }
    if (free_slots.empty()) break;
    auto c = canvases.find(id);
    if (c != canvases.end()) {
      Resident& R = c->second;
      if (R.rows != q->r.rows || R.cols != q->r.cols) {
        it = queue.erase(it);
        if (--pending[id] == 0) pending.erase(id);
        finish(q, PNT_EINVAL, "request size differs from its resident canvas");
        continue;
      }
      if (R.active || R.where == kSpilling || R.where == kRestoring) {
        ++it;
        continue;
      }
      if (R.where == kSpilled) {  // bring it back first
        float* blk = nullptr;
        int64_t cap = 0;
        if (reserve(canvas_floats(R.rows, R.cols), &blk, &cap)) {
          if (!blk) blk = dev_alloc(cap);
          R.block = blk;
          R.cap = cap;
          R.where = kRestoring;
          copyq.push_back(CopyJob{2, nullptr, id});
          cv_copy.notify_all();
        }
        ++it;
        continue;
      }
    } else {
      float* blk = nullptr;
      int64_t cap = 0;
      if (!reserve(canvas_floats(q->r.rows, q->r.cols), &blk, &cap)) {
        ++it;
        continue;
      }
      Resident nr;
      nr.block = blk;
      nr.cap = cap;
      nr.rows = q->r.rows;
      nr.cols = q->r.cols;
      nr.stride = round_up((int64_t)q->r.rows * q->r.cols, 64);
      for (int k = 0; k < 3; ++k) nr.bg[k] = q->r.background[k];
      c = canvases.emplace(id, std::move(nr)).first;
    }
    Resident& R = c->second;
    R.active++;
    q->res = &R;
    q->fresh = !R.initialised;
    R.initialised = true;
    q->slot = free_slots.back();
    free_slots.pop_back();
    slots[(size_t)q->slot].req = q;
    if (--pending[id] == 0) pending.erase(id);
    it = queue.erase(it);
    out.push_back(q);
  }
  return out;
}

// Device memory and the plan upload of an admitted request (engine thread).
void pnt_server_s::setup(Req* q) {
  SlotRes& S = slots[(size_t)q->slot];
  Resident& R = *q->res;
  hipStream_t s = lanes[q->lane].ctx->stream;
  if (!R.block) R.block = dev_alloc(R.cap);
  auto grow_f = [&](float*& p, int64_t& cap, int64_t need) {
    if (need <= cap) return;
    dev_free(p);
    p = nullptr;
    cap = std::max(need, 2 * cap);
    p = dev_alloc(cap);
  };
  const bool snap = q->r.use_snapshot != 0;
  if (snap) {
    grow_f(S.snap, S.snap_cap, kLayerStrides * R.stride);
    const int64_t need = dirty_bytes_for((int64_t)R.rows * R.cols);
    if (need > S.dirty_cap) {
      dev_free(S.dirty);
      S.dirty = nullptr;
      S.dirty_cap = std::max(need, 2 * S.dirty_cap);
      PNT_HIP_CHECK(hipMalloc((void**)&S.dirty, (size_t)S.dirty_cap));
    }
  }
  grow_f(S.pick, S.pick_cap, kChainsPerReq * 7 * q->mstride);
  if (q->plan_bytes > S.plan_cap) {
    dev_free(S.plan);
    S.plan = nullptr;
    S.plan_cap = std::max(q->plan_bytes, 2 * S.plan_cap);
    PNT_HIP_CHECK(hipMalloc((void**)&S.plan, (size_t)S.plan_cap));
  }
  if (q->plan_bytes > S.stage_cap) {
    if (S.stage) PNT_HIP_CHECK(hipHostFree(S.stage));
    S.stage = nullptr;
    S.stage_cap = std::max(q->plan_bytes, 2 * S.stage_cap);
    PNT_HIP_CHECK(hipHostMalloc((void**)&S.stage, (size_t)S.stage_cap));
  }
  // the slot's previous request completed (its fin event passed), so the
  // staging buffer and device plan are free to rewrite
  int64_t o = 0;
  auto put = [&](const void* src, size_t bytes) {
    const int64_t at = o;
    if (bytes) std::memcpy(S.stage + at, src, bytes);
    o += round_up((int64_t)bytes, 256);
    return at;
  };
  const int64_t o_st = put(q->st.data(), q->st.size() * sizeof(PStamp));
  const int64_t o_dl = put(q->dl.data(), q->dl.size() * sizeof(int2));
  const int64_t o_off = put(q->off.data(), q->off.size() * sizeof(int));
  const int64_t o_fp = put(q->fp.data(), q->fp.size() * sizeof(float));
  const int64_t o_paint = put(q->paint.data(), q->paint.size() * sizeof(float));
  const int64_t o_plan = o;
  if (o_plan) PNT_HIP_CHECK(hipMemcpyAsync(S.plan, S.stage, (size_t)o_plan, hipMemcpyHostToDevice, s));
  q->d_st = reinterpret_cast<const PStamp*>(S.plan + o_st);
  q->d_dl = reinterpret_cast<const int2*>(S.plan + o_dl);
  q->d_off = reinterpret_cast<const int*>(S.plan + o_off);
  q->d_fp = reinterpret_cast<const float*>(S.plan + o_fp);
  q->d_paint = reinterpret_cast<const float*>(S.plan + o_paint);
  write_descriptor(q, o);
}

// The request's shared phase fields into its slot's device descriptor, on
// the engine stream (after the waves of the slot's previous request).
void pnt_server_s::write_descriptor(Req* q, int64_t& o) {
  SlotRes& S = slots[(size_t)q->slot];
  const Resident& R = *q->res;
  const bool snap = q->r.use_snapshot != 0;
  SReq D{};
  D.st = q->d_st;
  D.canvas = R.block;
  D.snap = snap ? S.snap : R.block;
  D.dirty = snap ? S.dirty : nullptr;
  D.pick = S.pick;
  D.fp = q->d_fp;
  D.dl = q->d_dl;
  D.dl_off = q->d_off;
  D.paint = q->d_paint;
  for (int k = 0; k < 3; ++k) D.out[k] = R.block + (kCanvasStrides + k) * R.stride;
  D.cstride = (int)R.stride;
  D.mstride = (int)q->mstride;
  D.dirty_bytes = snap ? (int)dirty_bytes_for((int64_t)R.rows * R.cols) : 0;
  D.rows = R.rows;
  D.cols = R.cols;
  D.hr = q->hr;
  D.size_map = q->size_map;
  D.tiles_x = q->tiles_x;
  D.ntiles = q->ntiles;
  D.use_snap = snap ? 1 : 0;
  for (int k = 0; k < 4; ++k) D.fb[k] = q->fbox[k];
  for (int k = 0; k < 3; ++k) D.bg[k] = R.bg[k];
  D.pickup_rate = q->r.pickup_rate;
  D.dep_rate = q->r.deposition_rate;
  D.cap = 1.0;  // FootprintBrush.hxx:483
  std::memcpy(S.stage + o, &D, sizeof D);
  PNT_HIP_CHECK(hipMemcpyAsync(d_reqs + q->slot, S.stage + o, sizeof D, hipMemcpyHostToDevice,
                               lanes[q->lane].ctx->stream));
  o += round_up((int64_t)sizeof D, 256);
}

// ---------------------------------------------------------------- waves --
// Adds the request's ready phases to the wave being built (engine thread).
// A unit's phase is ready when its previous phase, its chain's previous unit
// and the units it depends on were issued in earlier waves.
bool pnt_server_s::collect(Req* q, SArgs& A, int& start,
                           std::vector<std::pair<Req*, int>>& used) {
  const Resident& R = *q->res;
  const bool snap = q->r.use_snapshot != 0;
  const int64_t n = (int64_t)R.rows * R.cols;
  auto base = [&](int kind) -> SPhase& {
    SPhase& P = A.ph[A.nphases];
    P = SPhase{};
    P.slot = (int16_t)q->slot;
    P.kind = (int8_t)kind;
    P.start = start;
    return P;
  };
  auto push = [&](SPhase& P, int unit) {
    start += (std::max(P.nblk, 1) + 7) / 8 * 8;  // phases start on a multiple of 8 blocks
    A.nphases++;
    used.push_back({q, unit});
  };
  bool any = false;
  for (;;) {  // stage changes that need no phase of their own
    if (q->stage == kPrologue && !q->fresh && !snap) {  // an existing canvas, no snapshot
      q->stage = q->U.empty() ? kComposite : kUnits;
      continue;
    }
    if (q->stage == kUnits) {
      bool all = true;
      for (size_t i = q->g0; i <= q->g1 && all; ++i) all = q->udone[i] != 0;
      if (!all) break;
      if (q->U[q->g1].dry_after) {
        q->stage = kDrying;
        break;
      }
      q->g0 = q->g1 + 1;
      if (q->g0 >= q->U.size()) {
        q->stage = kComposite;
        break;
      }
      q->g1 = q->g0;
      while (q->g1 + 1 < q->U.size() && !q->U[q->g1].dry_after) ++q->g1;
      continue;
    }
    break;
  }
  if (A.nphases == kPhases) return false;
  // part q->part of a streaming phase of `bytes` traffic over `items` per
  // part's range; `done` marks its last part
  auto streaming = [&](int kind, int64_t bytes, int64_t items, int done) {
    const int np = (int)std::min<int64_t>(std::max<int64_t>((bytes + part_bytes - 1) / part_bytes, 1),
                                          4096);
    SPhase& P = base(kind);
    P.rb = q->part | (np << 16);
    P.nblk = part_blocks(items / np);
    push(P, q->part + 1 == np ? done : -4);
  };
  const int64_t cs4 = R.stride / 4;  // float4s per plane
  switch (q->stage) {
    case kPrologue: {
      if (q->fresh)
        streaming(kSInit, (int64_t)(kCanvasStrides + (snap ? kLayerStrides : 0)) * R.stride * 4,
                  cs4, -1);
      else
        streaming(kSSnap, (int64_t)2 * kLayerStrides * R.stride * 4, cs4, -1);
      return true;
    }
    case kUnits: {
      for (size_t i = q->g0; i <= q->g1 && A.nphases < kPhases; ++i) {
        if (q->udone[i]) continue;
        bool ok = q->prev_on_chain[i] < 0 || q->udone[(size_t)q->prev_on_chain[i]];
        const Phase& f = q->ph[i][q->next[i]];
        if (ok && f.kind == kSStamp)  // the stamps it follows have run
          for (int e = q->needs.off[(size_t)f.k]; e < q->needs.off[(size_t)f.k + 1] && ok; ++e) {
            const int2 nd = q->needs.need[(size_t)e];
            ok = q->udone[(size_t)nd.x] || q->sdone[(size_t)nd.x] >= nd.y;
          }
        if (!ok) continue;
        SPhase& P = base(f.kind);
        P.k = (int)f.k;
        P.chain = (int8_t)q->U[i].chain;
        const int rb = q->rb[(size_t)f.k];
        switch (f.kind) {
          case kSClear:
            P.nblk = (int)std::min<int64_t>(64, (7 * q->mstride + 4 * kT2 - 1) / (4 * kT2));
            break;
          default:  // kSStamp, + the shared-pixel levels' block of a dup stamp
            P.rb = rb;
            P.nblk = q->ntiles + rb + (dup_levels(q->st[(size_t)f.k]) ? 1 : 0);
            break;
        }
        if (P.nblk == 0) P.nblk = 1;  // an empty ring: the block returns at once
        push(P, (int)i);
        any = true;
      }
      return any;
    }
    case kDrying:
      streaming(kSDry, 96 * n, n, -2);
      return true;
    case kComposite:
      streaming(kSCompose, 56 * n, n, -3);
      return true;
    default:
      return false;
  }
}

// Advances the requests that took part in the wave just launched.
void pnt_server_s::after_wave(const std::vector<std::pair<Req*, int>>& used) {
  for (const auto& u : used) {
    Req* q = u.first;
    if (u.second >= 0) {
      const size_t i = (size_t)u.second;
      if (q->ph[i][q->next[i]].kind == kSStamp) q->sdone[i] = q->ph[i][q->next[i]].k - q->U[i].s0 + 1;
      if (++q->next[i] == q->ph[i].size()) q->udone[i] = 1;
    } else if (u.second == -4) {  // a part of a streaming phase issued
      q->part++;
    } else if (u.second == -1) {  // prologue issued
      q->part = 0;
      q->stage = q->U.empty() ? kComposite : kUnits;
    } else if (u.second == -2) {  // dry issued: the next drying phase
      q->part = 0;
      q->g0 = q->g1 + 1;
      if (q->g0 >= q->U.size()) {
        q->stage = kComposite;
      } else {
        q->g1 = q->g0;
        while (q->g1 + 1 < q->U.size() && !q->U[q->g1].dry_after) ++q->g1;
        q->stage = kUnits;
      }
    } else if (u.second == -3) {  // composite issued: the request's last phase
      q->part = 0;
      q->stage = kIssued;
      PNT_HIP_CHECK(hipEventRecord(slots[(size_t)q->slot].fin, lanes[q->lane].ctx->stream));
    }
  }
}

// Marks a request done (caller holds mu).
void pnt_server_s::finish(Req* q, int status, const std::string& err) {
  q->status = status;
  q->error = err;
  q->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q->t_submit).count();
  q->done = true;
  ++completed;
  cv_done.notify_all();
}

// A canvas named by pnt_server_release goes once nothing needs it: no queued
// or running request and no copy in flight (caller holds mu).  Its device
// block returns to the pool; its host spill, if any, is freed.
void pnt_server_s::release_if_closing(int64_t id) {
  if (!closing.count(id) || pending.count(id)) return;
  auto c = canvases.find(id);
  if (c == canvases.end()) {
    closing.erase(id);
    return;
  }
  Resident& R = c->second;
  if (R.active || R.where == kSpilling || R.where == kRestoring) return;
  if (R.view) pnt_canvas_destroy(R.view);
  if (R.where == kResident) {
    if (R.block)
      pool.emplace(R.cap, R.block);
    else
      held -= R.cap * 4;  // reserved, never allocated
  }
  canvases.erase(c);
  closing.erase(id);
}

// The device finished a request (engine thread): its slot is free; the
// composite goes to the caller through the copier, or the request is done.
void pnt_server_s::complete(Req* q) {
  if (q->lane == 1) --express_active;  // engine-thread state
  std::lock_guard<std::mutex> lk(mu);
  free_slots.push_back(q->slot);
  slots[(size_t)q->slot].req = nullptr;
  if (q->r.out) {
    copyq.push_back(CopyJob{0, q, q->r.canvas_id});
    cv_copy.notify_all();
    return;
  }
  q->res->active--;
  q->res->last_use = Clock::now();
  finish(q, PNT_OK, "");
  release_if_closing(q->r.canvas_id);
}

void pnt_server_s::run_engine() {
  (void)hipSetDevice(device);
  SArgs A;
  std::vector<std::pair<Req*, int>> used;
  for (;;) {
    std::vector<Req*> fresh;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (stopping && queue.empty() && active.empty() && planq.empty()) return;
        fresh = admit();
        if (!fresh.empty() || !active.empty()) break;
        cv_engine.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
    for (Req* q : fresh) {
      // a long dependent chain goes to the express lane while it has room
      q->lane = lanes[1].ctx && express_active < express_max && q->alone_waves >= express_waves;
      if (q->lane == 1) ++express_active;
      const int st = guard_call([&] { setup(q); });
      if (st != PNT_OK) {
        if (q->lane == 1) --express_active;
        const std::string e = pnt_last_error();
        std::lock_guard<std::mutex> lk(mu);
        free_slots.push_back(q->slot);
        slots[(size_t)q->slot].req = nullptr;
        q->res->active--;
        finish(q, st, e);
        release_if_closing(q->r.canvas_id);
        continue;
      }
      active.push_back(q);
    }
    // one wave per lane from its requests' ready phases, oldest request
    // first; a lane with kAhead waves queued waits (blocking only when the
    // express lane is idle, so that it is never held up behind the main one)
    bool progressed = false;
    for (int ln = 0; ln < 2; ++ln) {
      Lane& L = lanes[ln];
      if (!L.ctx) continue;
      if (L.waves >= kAhead) {
        hipEvent_t oldest = L.ev[(L.waves - kAhead) % (kAhead + 1)];
        if (express_active == 0)
          PNT_HIP_CHECK(hipEventSynchronize(oldest));
        else if (hipEventQuery(oldest) != hipSuccess)
          continue;
      }
      A.reqs = reinterpret_cast<const SReqG*>(d_reqs);
      A.nphases = 0;
      int start = 0;
      used.clear();
      for (Req* q : active) {
        if (A.nphases == kPhases) break;
        if (q->lane != ln || q->stage == kIssued) continue;
        collect(q, A, start, used);
      }
      if (A.nphases == 0) continue;
      hipStream_t s = L.ctx->stream;
      k_serve_wave<<<(unsigned)start, kT2, 0, s>>>(A);
      PNT_HIP_CHECK(hipGetLastError());
      PNT_HIP_CHECK(hipEventRecord(L.ev[L.waves % (kAhead + 1)], s));
      ++L.waves;
      after_wave(used);
      progressed = true;
    }
    // requests whose device work is done
    for (size_t i = 0; i < active.size();) {
      Req* q = active[i];
      if (q->stage == kIssued && hipEventQuery(slots[(size_t)q->slot].fin) == hipSuccess) {
        complete(q);
        active.erase(active.begin() + (std::ptrdiff_t)i);
        progressed = true;
        continue;
      }
      ++i;
    }
    if (!progressed) {  // everything issued is still running
      Lane& X = lanes[1];
      if (express_active == 0) {  // wait a little
        std::unique_lock<std::mutex> lk(mu);
        cv_engine.wait_for(lk, std::chrono::microseconds(50));
      } else if (X.waves >= kAhead) {
        // the express lane's waves are a few microseconds each: wait for
        // its oldest, not for a timer, so that its queue never runs dry
        PNT_HIP_CHECK(hipEventSynchronize(X.ev[(X.waves - kAhead) % (kAhead + 1)]));
      } else {
        std::this_thread::yield();
      }
    }
  }
}

// ---------------------------------------------------------------- copier --
void pnt_server_s::copy_job(const CopyJob& j) {
  hipStream_t s = cctx->stream;
  if (j.kind == 0) {  // the composite to the caller's buffer
    Req* q = j.q;
    const Resident& R = *q->res;
    const int64_t n = (int64_t)R.rows * R.cols;
    for (int c = 0; c < 3; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(q->r.out + c * n, R.block + (kCanvasStrides + c) * R.stride,
                                   (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  Resident* R = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    R = &canvases.find(j.canvas_id)->second;
  }
  const int64_t floats = kCanvasStrides * R->stride;
  if (j.kind == 1) {  // spill to host memory
    R->host.resize((size_t)floats);
    PNT_HIP_CHECK(hipMemcpyAsync(R->host.data(), R->block, (size_t)floats * sizeof(float),
                                 hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  } else {  // restore from host memory
    PNT_HIP_CHECK(hipMemcpyAsync(R->block, R->host.data(), (size_t)floats * sizeof(float),
                                 hipMemcpyHostToDevice, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void pnt_server_s::run_copier() {
  (void)hipSetDevice(device);
  for (;;) {
    CopyJob j;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_copy.wait(lk, [&] { return stopping || !copyq.empty(); });
      if (copyq.empty()) return;
      j = copyq.front();
      copyq.pop_front();
    }
    const int st = guard_call([&] { copy_job(j); });
    const std::string err = st == PNT_OK ? std::string() : std::string(pnt_last_error());
    std::lock_guard<std::mutex> lk(mu);
    if (j.kind == 0) {
      j.q->res->active--;
      j.q->res->last_use = Clock::now();
      const int64_t id = j.q->r.canvas_id;
      finish(j.q, st, err);
      release_if_closing(id);
    } else {
      Resident& R = canvases.find(j.canvas_id)->second;
      if (j.kind == 1) {
        --spills;
        if (R.view) {
          pnt_canvas_destroy(R.view);
          R.view = nullptr;
        }
        if (st == PNT_OK) {
          pool.emplace(R.cap, R.block);
          R.block = nullptr;
          R.where = kSpilled;
        } else {
          R.where = kResident;  // keep it; admission may try another
          std::vector<float>().swap(R.host);
        }
      } else {
        R.where = kResident;
        if (st == PNT_OK) {
          std::vector<float>().swap(R.host);
        } else {  // the state could not come back: the block stays empty
          pool.emplace(R.cap, R.block);
          R.block = nullptr;
          R.where = kSpilled;
        }
      }
      release_if_closing(j.canvas_id);
    }
    cv_engine.notify_all();
  }
}

extern "C" {

int pnt_server_create(int device, int n_slots, int64_t hbm_budget_bytes, pnt_server* out) {
  return guard_call([&] {
    PNT_REQUIRE(out, "NULL argument");
    PNT_REQUIRE(n_slots >= 1 && n_slots <= 64, "n_slots must be in [1, 64]");
    PNT_REQUIRE(hbm_budget_bytes > 0, "HBM budget must be positive");
    auto* S = new pnt_server_s;
    S->device = device;
    S->budget = hbm_budget_bytes;
    if (const char* e = std::getenv("PAINTY_SERVE_PART_MB"))
      S->part_bytes = std::max<int64_t>(1, std::atoll(e)) << 20;
    try {
      if (const char* e = std::getenv("PAINTY_SERVE_EXPRESS")) S->express_max = std::max(0, std::atoi(e));
      if (const char* e = std::getenv("PAINTY_SERVE_EXPRESS_WAVES"))
        S->express_waves = std::max<int64_t>(1, std::atoll(e));
      for (pnt_ctx* c : {&S->ctx, &S->qctx, &S->cctx}) {
        const int st = pnt_ctx_create(device, c);
        if (st != PNT_OK) throw Error(st, pnt_last_error());
      }
      // the two lanes run side by side only on hardware queues of their own
      S->lanes[0].ctx = S->ctx;
      if (S->express_max > 0) {
        pnt_ctx_destroy(S->ctx);
        S->ctx = S->lanes[0].ctx = ctx_create_own_queue(device);
        S->lanes[1].ctx = ctx_create_own_queue(device);
      }
      for (auto& L : S->lanes)
        if (L.ctx)
          for (hipEvent_t& e : L.ev) PNT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      S->slots.resize((size_t)n_slots);
      for (int i = n_slots - 1; i >= 0; --i) S->free_slots.push_back(i);
      for (SlotRes& r : S->slots)
        PNT_HIP_CHECK(hipEventCreateWithFlags(&r.fin, hipEventDisableTiming));
      PNT_HIP_CHECK(hipMalloc((void**)&S->d_reqs, (size_t)n_slots * sizeof(SReq)));
    } catch (...) {
      if (S->d_reqs) (void)hipFree(S->d_reqs);
      for (SlotRes& r : S->slots)
        if (r.fin) (void)hipEventDestroy(r.fin);
      for (auto& L : S->lanes)
        for (hipEvent_t e : L.ev)
          if (e) (void)hipEventDestroy(e);
      for (pnt_ctx c : {S->ctx, S->qctx, S->cctx, S->lanes[1].ctx})
        if (c) pnt_ctx_destroy(c);
      delete S;
      throw;
    }
    S->engine = std::thread([S] { S->run_engine(); });
    S->copier = std::thread([S] { S->run_copier(); });
    const int np = (int)std::min(4u, std::max(1u, std::thread::hardware_concurrency() / 4));
    for (int i = 0; i < np; ++i) S->planners.emplace_back([S] { S->planner(); });
    *out = S;
  });
}

int pnt_server_submit(pnt_server S, const pnt_request* r, int64_t* ticket) {
  return guard_call([&] {
    PNT_REQUIRE(S && r && ticket, "NULL argument");
    PNT_REQUIRE(r->rows > 0 && r->cols > 0 && (int64_t)r->rows * r->cols < ((int64_t)1 << 31),
                "rows/cols out of range");
    PNT_REQUIRE(r->n_stamps >= 0 && (r->stamps || r->n_stamps == 0), "invalid stamp list");
    PNT_REQUIRE(r->footprint && r->width == (int)(2.0 * std::ceil(r->radius) + 1.0),
                "footprint width must be 2*ceil(radius)+1");
    PNT_REQUIRE(canvas_bytes(r->rows, r->cols) <= S->budget, "canvas exceeds the HBM budget");
    PNT_REQUIRE(r->radius > 0.0 && r->radius < 4096.0, "radius must be in (0, 4096)");
    auto* q = new Req;
    q->r = *r;
    q->stamps.assign(r->stamps, r->stamps + r->n_stamps);
    q->footprint.assign(r->footprint, r->footprint + (int64_t)r->width * r->width);
    q->r.stamps = nullptr;
    q->r.footprint = nullptr;
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stopping) {
      delete q;
      throw Error(PNT_ESTATE, "server is shutting down");
    }
    q->ticket = S->ticket_seq++;
    S->tickets[q->ticket] = q;
    S->pending[r->canvas_id]++;
    S->closing.erase(r->canvas_id);  // a new request on the id cancels its release
    S->queue.push_back(q);
    S->planq.push_back(q);
    *ticket = q->ticket;
    S->cv_plan.notify_one();
  });
}

}  // extern "C"

extern "C" {

int pnt_server_wait(pnt_server S, int64_t ticket, double* latency_ms) {
  return guard_call([&] {
    PNT_REQUIRE(S, "NULL server");
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto it = S->tickets.find(ticket);
      PNT_REQUIRE(it != S->tickets.end(), "unknown ticket");
      q = it->second;
      S->cv_done.wait(lk, [&] { return q->done; });
      S->tickets.erase(it);
    }
    if (latency_ms) *latency_ms = q->latency_ms;
    const int st = q->status;
    const std::string msg = q->error;
    delete q;
    if (st != PNT_OK) throw Error(st, msg);
  });
}

int pnt_server_canvas(pnt_server S, int64_t canvas_id, pnt_canvas* canvas, pnt_dev_rgb* composite) {
  return guard_call([&] {
    PNT_REQUIRE(S && canvas, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto c = S->canvases.find(canvas_id);
    PNT_REQUIRE(c != S->canvases.end(), "unknown canvas");
    Resident& R = c->second;
    PNT_REQUIRE(R.where == kResident && R.block, "canvas is not resident (spilled to the host)");
    PNT_REQUIRE(!R.active && !S->pending.count(canvas_id), "canvas has a queued or running request");
    if (!R.view) R.view = canvas_over(S->qctx, R.rows, R.cols, R.block, false);
    *canvas = R.view;
    if (composite)
      for (int k = 0; k < 3; ++k) composite->p[k] = R.block + (kCanvasStrides + k) * R.stride;
  });
}

int pnt_server_release(pnt_server S, int64_t canvas_id) {
  return guard_call([&] {
    PNT_REQUIRE(S, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    PNT_REQUIRE(S->canvases.count(canvas_id) || S->pending.count(canvas_id), "unknown canvas");
    S->closing.insert(canvas_id);
    S->release_if_closing(canvas_id);
    S->cv_engine.notify_all();
  });
}

int pnt_server_stats(pnt_server S, int64_t* resident_bytes, int64_t* resident_canvases,
                     int64_t* completed) {
  return guard_call([&] {
    PNT_REQUIRE(S, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    int64_t b = 0, c = 0;
    for (auto& kv : S->canvases)
      if (kv.second.block) {
        b += kv.second.cap * 4;
        ++c;
      }
    if (resident_bytes) *resident_bytes = b;
    if (resident_canvases) *resident_canvases = c;
    if (completed) *completed = S->completed;
  });
}

int pnt_server_destroy(pnt_server S) {
  return guard_call([&] {
    if (!S) return;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_done.wait(lk, [&] {  // let submitted work finish
        for (auto& kv : S->tickets)
          if (!kv.second->done) return false;
        return S->queue.empty();
      });
      S->stopping = true;
    }
    S->cv_engine.notify_all();
    S->cv_plan.notify_all();
    S->cv_copy.notify_all();
    S->engine.join();
    for (auto& t : S->planners) t.join();
    S->copier.join();
    (void)hipSetDevice(S->device);
    (void)hipStreamSynchronize(S->ctx->stream);
    if (S->lanes[1].ctx) (void)hipStreamSynchronize(S->lanes[1].ctx->stream);
    for (auto& kv : S->tickets) delete kv.second;
    for (auto& kv : S->canvases) {
      if (kv.second.view) pnt_canvas_destroy(kv.second.view);
      dev_free(kv.second.block);
    }
    for (auto& kv : S->pool) dev_free(kv.second);
    if (S->d_reqs) (void)hipFree(S->d_reqs);
    for (SlotRes& r : S->slots) {
      dev_free(r.snap);
      dev_free(r.dirty);
      dev_free(r.pick);
      dev_free(r.plan);
      if (r.stage) (void)hipHostFree(r.stage);
      if (r.fin) (void)hipEventDestroy(r.fin);
    }
    for (auto& L : S->lanes)
      for (hipEvent_t e : L.ev)
        if (e) (void)hipEventDestroy(e);
    if (S->lanes[1].ctx) pnt_ctx_destroy(S->lanes[1].ctx);
    pnt_ctx_destroy(S->ctx);
    pnt_ctx_destroy(S->qctx);
    pnt_ctx_destroy(S->cctx);
    delete S;
  });
}

}  // extern "C"
