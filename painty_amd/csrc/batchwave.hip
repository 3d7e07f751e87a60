// A batch of canvases as waves: the jobs' stamp lists are planned on the
// host into a sequence of launches, each carrying the next phase of every
// ready stroke unit of every job (wave.hpp), so hundreds of canvases advance
// together and the launch count is set by the longest job's dependent
// chain, not by the sum over jobs.  Selected by a batch configuration with
// `waves` set; pnt_batch_tune compares it with the one-launch persistent
// kernel (batch.hip) and keeps the faster.
//
// Per job (a canvas, its brush and stamps): the stamps are cut into units
// (runs of one stroke, plan_units in fused.hip) whose snapshot boxes decide
// which may run at once; `chunk` bounds how many units of the longest job
// are in flight (each has its own pickup map; shorter jobs get
// proportionally fewer).  The brush's own map takes the chain
// of the job's last unit, so it ends holding the last stroke's map, as it
// would after the same stamps run alone; a job whose first unit continues
// the brush's current stroke keeps that map on its first chain and copies
// the last chain's map back at the end.  dryCanvas and the composite are
// phases of their own.
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <exception>
#include <numeric>
#include <thread>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"
#include "wave.hpp"

namespace pnt {

namespace {

constexpr int kMaxChunk = 4;

struct BWJob {
  float* canvas;
  float* snap;
  uint8_t* dirty;
  float* out[3];
  float* pick[kMaxChunk];
  const float* fp;
  int64_t o_st, o_dl, o_off, o_paint;  // offsets into the plan image
  int cstride, mstride, dirty_bytes, rows, cols, hr, size_map, tiles_x, ntiles, use_snap;
  int fb[4];
  double pickup_rate, dep_rate, cap;
  int pad;
  int64_t* tmap;  // drying-clock time map of the canvas (nullptr: no clock)
  int64_t dry_ms;
};

// BWJob as the kernel reads it: the same layout, device pointers global.
struct BWJobG {
  gptr<float> canvas;
  gptr<float> snap;
  gptr<uint8_t> dirty;
  gptr<float> out[3];
  gptr<float> pick[kMaxChunk];
  gptr<const float> fp;
  int64_t o_st, o_dl, o_off, o_paint;
  int cstride, mstride, dirty_bytes, rows, cols, hr, size_map, tiles_x, ntiles, use_snap;
  int fb[4];
  double pickup_rate, dep_rate, cap;
  int pad;
  gptr<int64_t> tmap;
  int64_t dry_ms;
};
static_assert(sizeof(BWJobG) == sizeof(BWJob) && offsetof(BWJobG, fp) == offsetof(BWJob, fp) &&
                  offsetof(BWJobG, fb) == offsetof(BWJob, fb) &&
                  offsetof(BWJobG, tmap) == offsetof(BWJob, tmap),
              "BWJobG mirrors BWJob");

// 4-pixel groups per ring-scan thread in the batch's waves: twice the
// single-canvas count, so half the ring blocks.  The batch is bound by how
// many workgroups it runs, not by its chain of launches, and measured 2%
// faster with 4 (1 and 2 were slower); the single canvas, bound by its
// chain, is faster with 2.
constexpr int kBatchRingGroups = 4;

struct BWSlot {
  int job, k, kind, start, nblk, rb, chain, pad;
  // image offsets of the phase's stamp descriptor and paint: with them the
  // block loads its stamp, paint and job record at once, not one after
  // another
  int64_t o_stamp, o_paint;
};

// One wave: the group table gives the slot of every 8 blocks (slots start on
// multiples of 8 blocks).
__global__ __launch_bounds__(kWT2, kWaveOccupancy) void k_batch_wave(const char* base, const BWJobG* jobs,
                                                     const BWSlot* slots, const int* grp) {
  __shared__ int sm[kWT + 2][kWT + 2];
  const BWSlot d = slots[grp[blockIdx.x >> 3]];
  const int b = (int)blockIdx.x - d.start;
  if (b >= d.nblk) return;
  const BWJobG& J = jobs[d.job];
  SSlot S;
  S.st = reinterpret_cast<const PStamp*>(base + d.o_stamp);
  S.canvas = (float*)J.canvas;
  S.snap = (float*)J.snap;
  S.dirty = (uint8_t*)J.dirty;
  S.pick = (float*)J.pick[d.chain];
  S.pick_dst = (float*)J.pick[0];  // kSCopyPick: this chain's map -> the brush's own
  S.fp = (const float*)J.fp;
  S.dl = reinterpret_cast<const int2*>(base + J.o_dl);
  S.dl_off = reinterpret_cast<const int*>(base + J.o_off);
#pragma unroll
  for (int c = 0; c < 3; ++c) S.out[c] = (float*)J.out[c];
  S.cstride = J.cstride;
  S.mstride = J.mstride;
  S.dirty_bytes = J.dirty_bytes;
  S.rows = J.rows;
  S.cols = J.cols;
  S.kind = d.kind;
  S.start = d.start;
  S.nblk = d.nblk;
  S.ntiles = J.ntiles;
  S.rb = d.rb;
  S.hr = J.hr;
  S.size_map = J.size_map;
  S.tiles_x = J.tiles_x;
#pragma unroll
  for (int c = 0; c < 4; ++c) S.fb[c] = J.fb[c];
  const float* paint = reinterpret_cast<const float*>(base + d.o_paint);
  const bool stamp_kind = d.kind == kSStamp;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    S.pk[c] = stamp_kind ? paint[c] : 0.0f;
    S.ps[c] = stamp_kind ? paint[3 + c] : 0.0f;
  }
  S.pickup_rate = J.pickup_rate;
  S.dep_rate = J.dep_rate;
  S.cap = J.cap;
  S.use_snap = J.use_snap;
  S.bg[0] = S.bg[1] = S.bg[2] = 0.0f;
  S.tmap = (int64_t*)J.tmap;
  S.dry_ms = J.dry_ms;
  S.part = 0;  // streaming phases run whole
  S.nparts = 1;
  run_wave_slot<kBatchRingGroups>(S, b, sm);
}

struct Phase {
  int64_t k;
  int kind;
};

enum : int { kJUnits = 0, kJDrying, kJCopy, kJCompose, kJDone };

// Host plan and wave-simulation state of one job.
struct JobPlan {
  std::vector<PStamp> st;
  std::vector<int2> dl;
  std::vector<int> off;
  std::vector<Unit> U;
  std::vector<std::vector<Phase>> ph;
  std::vector<int> prev_on_chain;
  std::vector<float> paint;
  std::vector<int> rb;
  int chains = 1;       // chains used
  int perm[kMaxChunk];  // chain -> map slot in BWJob.pick (0: the brush's own map)
  bool copy_back = false;
  int stage = kJUnits;
  size_t g0 = 0, g1 = 0;
  std::vector<size_t> next;
  std::vector<char> udone;
  std::vector<int64_t> sdone;  // stamps of each unit issued
  StampNeeds needs;
};

void plan_job(const pnt_batch_job& job, int chunk, JobPlan& P) {
  pnt_brush b = job.brush;
  pnt_canvas cv = job.canvas;
  const int64_t n = job.n_stamps;
  plan_stamps(b, cv, job.stamps, n, P.st, P.dl, P.off);
  std::vector<uint8_t> clear((size_t)n, 0);
  P.U = n ? plan_units(job.stamps, P.st.data(), clear.data(), n, chunk) : std::vector<Unit>{};
  P.ph.assign(P.U.size(), {});
  P.rb.assign((size_t)n, 0);
  for (size_t i = 0; i < P.U.size(); ++i)
    for (int64_t k = P.U[i].s0; k < P.U[i].s1; ++k) {
      if (job.stamps[k].flags & PNT_STAMP_DIP) P.ph[i].push_back(Phase{k, kSClear});
      const PStamp& p = P.st[(size_t)k];
      P.rb[(size_t)k] = b->use_snap ? ring_blocks_for(p.box, kWT2, kBatchRingGroups) : 0;
      P.ph[i].push_back(Phase{k, kSStamp});
    }
  P.prev_on_chain.assign(P.U.size(), -1);
  std::vector<int> last(kMaxChunk, -1);
  P.chains = 1;
  for (size_t i = 0; i < P.U.size(); ++i) {
    P.prev_on_chain[i] = last[(size_t)P.U[i].chain];
    last[(size_t)P.U[i].chain] = (int)i;
    P.chains = std::max(P.chains, P.U[i].chain + 1);
  }
  // the brush's own map: the chain of the last unit, unless the first unit
  // continues the brush's stroke (no dip before it: it must start from that
  // map, on chain 0); then the last chain's map is copied back at the end
  const bool cont = n > 0 && !(job.stamps[0].flags & PNT_STAMP_DIP);
  const int lastc = P.U.empty() ? 0 : P.U.back().chain;
  const int own = cont ? 0 : lastc;
  P.copy_back = cont && lastc != 0;
  int nxt = 1;
  for (int c = 0; c < kMaxChunk; ++c) P.perm[c] = c == own ? 0 : (c < P.chains ? nxt++ : 0);
  P.paint.assign(6 * (size_t)n, 0.0f);
  float cur[6] = {b->paintK[0], b->paintK[1], b->paintK[2],
                  b->paintS[0], b->paintS[1], b->paintS[2]};
  for (int64_t k = 0; k < n; ++k) {
    if (job.stamps[k].flags & PNT_STAMP_DIP)
      for (int c = 0; c < 3; ++c) {
        cur[c] = job.stamps[k].K[c];
        cur[3 + c] = job.stamps[k].S[c];
      }
    std::memcpy(&P.paint[6 * (size_t)k], cur, sizeof cur);
  }
  P.next.assign(P.U.size(), 0);
  P.udone.assign(P.U.size(), 0);
  P.sdone.assign(P.U.size(), 0);
  P.needs = plan_stamp_needs(P.U, P.st.data());
  for (size_t i = 0; i < P.U.size(); ++i)
    if (P.ph[i].empty()) P.udone[i] = 1;
  P.g0 = P.g1 = 0;
  while (P.g1 + 1 < P.U.size() && !P.U[P.g1].dry_after) ++P.g1;
  P.stage = P.U.empty() ? kJCopy : kJUnits;
}

template <class F>
void parallel_for(int n, int64_t work, F&& f) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<int64_t>({(int64_t)std::min(hw, 16), (int64_t)n, 1 + work / 2048});
  if (nt <= 1) {
    for (int j = 0; j < n; ++j) f(j);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::exception_ptr> errs((size_t)nt);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      try {
        for (int j; (j = next.fetch_add(1)) < n;) f(j);
      } catch (...) {
        errs[(size_t)t] = std::current_exception();
        next.store(n);
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
}

}  // namespace

// Units in flight for each job: `chunk` for the job with the most stamps,
// proportionally fewer (at least 1) for shorter ones.  The batch ends with
// its longest job, so extra concurrency goes where it shortens that chain;
// short jobs keep one pickup map and add no map traffic.
static std::vector<int> job_chunks(const pnt_batch_job* jobs, int n, int chunk) {
  chunk = std::max(1, std::min(chunk, kMaxChunk));
  int64_t most = 1;
  for (int j = 0; j < n; ++j) most = std::max<int64_t>(most, jobs[j].n_stamps);
  std::vector<int> c((size_t)n, 1);
  for (int j = 0; j < n; ++j)
    c[(size_t)j] = std::max(1, (int)std::lround((double)chunk * (double)jobs[j].n_stamps /
                                                (double)most));
  return c;
}

// Plans a batch as waves (chunk: units of the longest job in flight at once).
// The brushes' snapshots are readied on stream s first (as for one launch).
// `pool` receives the extra pickup maps; plan_waves_size gives its size.
int64_t plan_waves_pool(const pnt_batch_job* jobs, int n, int chunk) {
  const std::vector<int> c = job_chunks(jobs, n, chunk);
  int64_t need = 0;
  for (int j = 0; j < n; ++j) need += (int64_t)(c[(size_t)j] - 1) * 7 * jobs[j].brush->mstride;
  return need;
}

BWPlan plan_waves(const pnt_batch_job* jobs, int n, int chunk, float* pool, hipStream_t s) {
  const std::vector<int> chunks = job_chunks(jobs, n, chunk);
  std::vector<JobPlan> P((size_t)n);
  int64_t work = 0;
  for (int j = 0; j < n; ++j) work += jobs[j].n_stamps;
  parallel_for(n, work, [&](int j) { plan_job(jobs[j], chunks[(size_t)j], P[(size_t)j]); });
  for (int j = 0; j < n; ++j) brush_prepare_snapshot(jobs[j].brush, jobs[j].canvas, s);

  // simulate: every wave takes every ready phase of every job
  std::vector<BWSlot> slots;
  std::vector<int> grp;
  std::vector<int64_t> ws, wg;
  std::vector<int> wb;
  std::vector<std::pair<int, int>> took;  // (job, unit or -1 dry / -2 copy / -3 compose)
  std::vector<int> jfirst((size_t)n, -1), jlast((size_t)n, -1);
  for (;;) {
    const int64_t s0 = (int64_t)slots.size(), g0 = (int64_t)grp.size();
    int start = 0;
    took.clear();
    auto add = [&](int j, int64_t k, int kind, int nblk, int rb, int chain, int tag) {
      BWSlot d{j, (int)std::max<int64_t>(k, 0), kind, start, std::max(nblk, 1), rb, chain, 0, 0, 0};
      const int span = (d.nblk + 7) / 8 * 8;
      for (int q = 0; q < span / 8; ++q) grp.push_back((int)(slots.size() - (size_t)s0));
      slots.push_back(d);
      start += span;
      took.push_back({j, tag});
    };
    for (int j = 0; j < n; ++j) {
      JobPlan& J = P[(size_t)j];
      const pnt_batch_job& job = jobs[j];
      const pnt_brush b = job.brush;
      const int64_t npx = job.canvas->n;
      // stage changes that need no phase of their own
      if (J.stage == kJUnits) {
        bool all = true;
        for (size_t i = J.g0; i <= J.g1 && all; ++i) all = J.udone[i] != 0;
        if (all) {
          if (J.U[J.g1].dry_after) {
            J.stage = kJDrying;
          } else {
            J.g0 = J.g1 + 1;
            if (J.g0 >= J.U.size()) {
              J.stage = kJCopy;
            } else {
              J.g1 = J.g0;
              while (J.g1 + 1 < J.U.size() && !J.U[J.g1].dry_after) ++J.g1;
            }
          }
        }
      }
      if (J.stage == kJCopy && !J.copy_back) J.stage = kJCompose;
      if (J.stage == kJCompose && !job.out[0]) J.stage = kJDone;
      switch (J.stage) {
        case kJUnits:
          for (size_t i = J.g0; i <= J.g1; ++i) {
            if (J.udone[i]) continue;
            bool ok = J.prev_on_chain[i] < 0 || J.udone[(size_t)J.prev_on_chain[i]];
            const Phase& f = J.ph[i][J.next[i]];
            if (ok && f.kind == kSStamp)  // the stamps it follows have run
              for (int e = J.needs.off[(size_t)f.k]; e < J.needs.off[(size_t)f.k + 1] && ok; ++e) {
                const int2 nd = J.needs.need[(size_t)e];
                ok = J.udone[(size_t)nd.x] || J.sdone[(size_t)nd.x] >= nd.y;
              }
            if (!ok) continue;
            const int rb = J.rb[(size_t)f.k];
            const int ntl = ((2 * ((b->size_map - 1) / 2) + 1 + kWT - 1) / kWT);
            const int ntiles = ntl * ntl;
            int nblk = 1;
            switch (f.kind) {
              case kSClear:
                nblk = (int)std::min<int64_t>(64, (7 * b->mstride + 4 * kWT2 - 1) / (4 * kWT2));
                break;
              case kSStamp:  // + the shared-pixel levels' block of a dup stamp
                nblk = ntiles + rb + (dup_levels(J.st[(size_t)f.k]) ? 1 : 0);
                break;
              default: nblk = 1; break;
            }
            add(j, f.k, f.kind, nblk, rb, J.perm[J.U[i].chain], (int)i);
          }
          break;
        case kJDrying:  // its stamp: the drying phase's last (its time point)
          add(j, J.U[J.g1].s1 - 1, kSDry, wave_stream_blocks(npx), 0, 0, -1);
          break;
        case kJCopy:
          add(j, 0, kSCopyPick, 64, 0, J.perm[J.U.back().chain], -2);
          break;
        case kJCompose:
          add(j, 0, kSCompose, wave_stream_blocks(npx), 0, 0, -3);
          break;
        default:
          break;
      }
    }
    if (took.empty()) break;
    ws.push_back(s0);
    wg.push_back(g0);
    wb.push_back(start);
    for (const auto& t : took) {
      JobPlan& J = P[(size_t)t.first];
      if (jfirst[(size_t)t.first] < 0) jfirst[(size_t)t.first] = (int)ws.size() - 1;
      jlast[(size_t)t.first] = (int)ws.size() - 1;
      if (t.second >= 0) {
        const size_t i = (size_t)t.second;
        if (J.ph[i][J.next[i]].kind == kSStamp) J.sdone[i] = J.ph[i][J.next[i]].k - J.U[i].s0 + 1;
        if (++J.next[i] == J.ph[i].size()) J.udone[i] = 1;
      } else if (t.second == -1) {
        J.g0 = J.g1 + 1;
        if (J.g0 >= J.U.size()) {
          J.stage = kJCopy;
        } else {
          J.g1 = J.g0;
          while (J.g1 + 1 < J.U.size() && !J.U[J.g1].dry_after) ++J.g1;
          J.stage = kJUnits;
        }
      } else if (t.second == -2) {
        J.stage = kJCompose;
      } else {
        J.stage = kJDone;
      }
    }
  }

  // the device image
  BWPlan W;
  W.njobs = n;
  auto al = [](int64_t x) { return round_up(x, 256); };
  std::vector<int64_t> o_st(n), o_dl(n), o_off(n), o_paint(n);
  int64_t o = al((int64_t)n * (int64_t)sizeof(BWJob));
  const int64_t o_slots = o;
  o = al(o + (int64_t)slots.size() * (int64_t)sizeof(BWSlot));
  const int64_t o_grp = o;
  o = al(o + (int64_t)grp.size() * (int64_t)sizeof(int));
  for (int j = 0; j < n; ++j) {
    const JobPlan& J = P[(size_t)j];
    o_st[j] = o;
    o = al(o + (int64_t)(J.st.size() * sizeof(PStamp)));
    o_dl[j] = o;
    o = al(o + (int64_t)(J.dl.size() * sizeof(int2)));
    o_off[j] = o;
    o = al(o + (int64_t)(J.off.size() * sizeof(int)));
    o_paint[j] = o;
    o = al(o + (int64_t)(J.paint.size() * sizeof(float)));
  }
  W.total = o;
  W.blob.assign((size_t)o, 0);
  W.o_jobs = 0;
  int64_t pool_at = 0;
  for (int j = 0; j < n; ++j) {
    const JobPlan& J = P[(size_t)j];
    const pnt_batch_job& job = jobs[j];
    pnt_brush b = job.brush;
    pnt_canvas cv = job.canvas;
    BWJob d{};
    d.canvas = cv->base;
    d.snap = b->use_snap ? b->snap : cv->base;
    d.dirty = b->use_snap ? b->dirty : nullptr;
    for (int c = 0; c < 3; ++c) d.out[c] = job.out[c];
    d.pick[0] = b->pick;
    for (int c = 1; c < kMaxChunk; ++c) {
      if (c < J.chains) {
        d.pick[c] = pool + pool_at;
        pool_at += 7 * b->mstride;
      } else {
        d.pick[c] = b->pick;
      }
    }
    d.fp = b->fp;
    d.o_st = o_st[j];
    d.o_dl = o_dl[j];
    d.o_off = o_off[j];
    d.o_paint = o_paint[j];
    d.cstride = (int)cv->stride;
    d.mstride = (int)b->mstride;
    d.dirty_bytes = 0;
    d.rows = cv->rows;
    d.cols = cv->cols;
    d.hr = (b->size_map - 1) / 2;
    d.size_map = b->size_map;
    d.tiles_x = (2 * d.hr + 1 + kWT - 1) / kWT;
    d.ntiles = d.tiles_x * d.tiles_x;
    d.use_snap = b->use_snap ? 1 : 0;
    for (int c = 0; c < 4; ++c) d.fb[c] = b->fbox[c];
    d.pickup_rate = b->pickup_rate;
    d.dep_rate = b->dep_rate;
    d.cap = b->cap;
    d.tmap = cv->clock != PNT_CLOCK_OFF ? cv->tmap : nullptr;
    d.dry_ms = cv->drying_ms;
    std::memcpy(W.blob.data() + (size_t)j * sizeof(BWJob), &d, sizeof d);
    if (!J.st.empty()) std::memcpy(W.blob.data() + o_st[j], J.st.data(), J.st.size() * sizeof(PStamp));
    if (!J.dl.empty()) std::memcpy(W.blob.data() + o_dl[j], J.dl.data(), J.dl.size() * sizeof(int2));
    if (!J.off.empty()) std::memcpy(W.blob.data() + o_off[j], J.off.data(), J.off.size() * sizeof(int));
    if (!J.paint.empty())
      std::memcpy(W.blob.data() + o_paint[j], J.paint.data(), J.paint.size() * sizeof(float));
  }
  W.pool_floats = pool_at;
  for (BWSlot& d : slots) {
    d.o_stamp = o_st[(size_t)d.job] + (int64_t)d.k * (int64_t)sizeof(PStamp);
    d.o_paint = o_paint[(size_t)d.job] + 6 * (int64_t)d.k * (int64_t)sizeof(float);
  }
  if (!slots.empty()) std::memcpy(W.blob.data() + o_slots, slots.data(), slots.size() * sizeof(BWSlot));
  if (!grp.empty()) std::memcpy(W.blob.data() + o_grp, grp.data(), grp.size() * sizeof(int));
  for (size_t w = 0; w < ws.size(); ++w) {
    W.wave_slots.push_back(o_slots + ws[w] * (int64_t)sizeof(BWSlot));
    W.wave_grp.push_back(o_grp + wg[w] * (int64_t)sizeof(int));
  }
  W.wave_blocks = wb;
  W.job_first = std::move(jfirst);
  W.job_last = std::move(jlast);
  return W;
}

// Copies the image (pinned `stage` -> `dev`) and launches every wave on s.
void enqueue_waves(const BWPlan& W, const char* stage, char* dev, hipStream_t s,
                   const hipEvent_t* ev) {
  if (W.total) PNT_HIP_CHECK(hipMemcpyAsync(dev, stage, (size_t)W.total, hipMemcpyHostToDevice, s));
  if (ev) PNT_HIP_CHECK(hipEventRecord(ev[0], s));
  const BWJobG* jobs = reinterpret_cast<const BWJobG*>(dev + W.o_jobs);
  for (size_t w = 0; w < W.wave_blocks.size(); ++w) {
    k_batch_wave<<<(unsigned)W.wave_blocks[w], kWT2, 0, s>>>(
        dev, jobs, reinterpret_cast<const BWSlot*>(dev + W.wave_slots[w]),
        reinterpret_cast<const int*>(dev + W.wave_grp[w]));
    PNT_HIP_CHECK(hipGetLastError());
    if (ev) PNT_HIP_CHECK(hipEventRecord(ev[w + 1], s));
  }
}

}  // namespace pnt
