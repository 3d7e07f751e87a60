// Many canvases, each with its own brush and stamp list, in one cooperative
// launch, followed by their composites.
//
// Work split: jobs (canvas + brush + stamps) are packed into groups of
// blocks by estimated cost.  A group runs its jobs one after another; each
// stamp of a job is FootprintBrush::imprint (renderer/FootprintBrush.hxx:
// 73-143) spread over the group's blocks in 32x32 window tiles.  A one-block
// group needs only workgroup barriers; a larger group uses the bounded group
// barrier of stampcore.hpp.  Blocks of one group sit in one XCD lane
// (block index = lane mod 8), so a job's window stays in one L2.
//
// Visit order within a stamp is kept as in k_stamp_main (kernels.hip): a map
// cell has at most two visitors, which are 4-neighbours in the window; the
// earlier one runs both, found through the map indices of the tile and a
// one-pixel halo in LDS.  Visits on canvas column/row 0 that share a pixel
// run after the others by dependency level.
//
// Snapshot ring: per-pixel dirty bytes (persistent.hip) scanned 16 at a time;
// visits on the window border refresh their own pixel.
//
// Composite: after its jobs, every block takes composite tiles from a shared
// counter; a tile waits for its job's "stamped" flag (set with release
// semantics by the job's group), so blocks that finish early composite other
// canvases.  Every wait is bounded.
#include <algorithm>
#include <atomic>
#include <exception>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"
#include "wave.hpp"

namespace pnt {

enum : uint32_t { kStDry = 8u };  // dryCanvas() after this stamp

namespace {

constexpr int64_t kComposeItem = 1 << 18;  // pixels per composite work item

struct BJob {
  Layer7 canvas, snap;
  Plane7 pick;
  Rgb3 R0, out;
  float* h;
  const float* fp;
  FpBox fb;  // nonzero footprint cells
  uint8_t* dirty;
  int compose, use_snap;
  int rows, cols, hr, size_map;
  double pickup_rate, dep_rate, cap;
  float paintK[3], paintS[3];
  int s0, s1;  // stamps [s0, s1) of the plan
  int pad[2];
};

struct BGroup {
  int job0, njobs;  // jobs job_list[job0 .. job0+njobs)
  int lane, k0;     // blocks lane + 8*(k0 .. k0+nblocks)
  int nblocks;
  int pad[3];
};

struct CItem {
  int job;
  int pad;
  int64_t begin, end;
};

struct BArgs {
  const BJob* jobs;
  const BGroup* groups;
  const int* block_group;
  const int* job_list;
  const PStamp* st;
  const int2* dl;
  const int* dl_off;
  const CItem* items;
  int n_items;
  unsigned* item_next;  // composite work counter
  unsigned* done;       // per job: stamping finished
  unsigned* bars;       // kBarWords per group
  unsigned* host_err;
};

// Bounded wait for a job's "stamped" flag (thread 0 only), then acquire.
__device__ __forceinline__ bool wait_done(unsigned* flag, unsigned* host_err, unsigned* abort) {
  unsigned spins = 0;
  while (!__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    __builtin_amdgcn_s_sleep(2);
    if ((++spins & 63) == 0) {
      if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
      if (spins > kSpinLimit) {
        __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(host_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return false;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// One launch of a batch plan; kWT x kWT window tiles, one thread per tile
// pixel.
template <int kWT>
__global__ __launch_bounds__(kWT* kWT) void k_batch(BArgs A) {
  constexpr int kBT = kWT * kWT;
  __shared__ int sm[kWT + 2][kWT + 2];
  __shared__ int s_item;
  __shared__ int s_ok;
  const int t = threadIdx.x;
  const int gid = A.block_group[blockIdx.x];
  const BGroup G = A.groups[gid];
  const int rank = (int)(blockIdx.x >> 3) - G.k0;
  const unsigned nb = (unsigned)G.nblocks;
  unsigned* bar = A.bars + (size_t)gid * kBarWords;
  unsigned* abort = bar + kBarAbort;
  bool alive = true;

  for (int jj = 0; jj < G.njobs && alive; ++jj) {
    const int jid = A.job_list[G.job0 + jj];
    const BJob& J = A.jobs[jid];
    StampGeom g;
    g.hr = J.hr;
    g.size_map = J.size_map;
    g.rows = J.rows;
    g.cols = J.cols;
    VisitTarget T{&J.canvas, &J.snap, J.dirty, J.use_snap, J.pickup_rate, J.dep_rate, J.cap};
    float pk[3], ps[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pk[c] = J.paintK[c];
      ps[c] = J.paintS[c];
    }
    const int side = 2 * J.hr + 1;
    const int tiles_x = (side + kWT - 1) / kWT;
    const int ntiles = tiles_x * tiles_x;
    const int64_t m = (int64_t)J.size_map * J.size_map;
    const int gtid = rank * kBT + t;
    const int gthreads = (int)nb * kBT;
    for (int j = J.s0; j < J.s1 && alive; ++j) {
      const PStamp& st = A.st[j];
      g.cx = st.cx;
      g.cy = st.cy;
      g.c = st.c;
      g.s = st.s;
      const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
      const bool pull = J.use_snap && !dup;
      if (st.flags & kStDip) {  // dip(): clean() the pickup map, new paint
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pk[c] = st.K[c];
          ps[c] = st.S[c];
        }
        for (int64_t i = gtid; i < m; i += gthreads)
#pragma unroll
          for (int c = 0; c < 7; ++c) J.pick.p[c][i] = 0.0f;
        if (!group_sync(bar, nb, rank, A.host_err)) {
          alive = false;
          break;
        }
      }
      if (J.use_snap) {
        ring_scan<1>(J.canvas, J.snap, J.dirty, J.fp, g, st.box, pull, gtid, gthreads);
        if (dup && !group_sync(bar, nb, rank, A.host_err)) {
          alive = false;
          break;
        }
      }
      // ---- visits: 32x32 window tiles, round-robin over the group's blocks
      for (int tile = rank; tile < ntiles; tile += (int)nb)
        window_tile<kWT>(T, J.pick.p, J.fp, J.fb, g, st, dup, pull, tile, tiles_x, sm, pk, ps);
      // ---- visits sharing canvas column/row 0, level by level
      for (int L = 0; L < st.dl_nlev; ++L) {
        if (!group_sync(bar, nb, rank, A.host_err)) {
          alive = false;
          break;
        }
        const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
        for (int e = e0 + gtid; e < e1; e += gthreads) {
          const int2 v = A.dl[e];
          int64_t ic;
          const int mm2 = stamp_visit(g, v.x, v.y, &ic);
          if (mm2 >= 0 && J.fp[mm2] > 0.0f)
            cached_visit(T, GlobalPick{J.pick.p, mm2}, J.fp[mm2], ic, false, pk, ps);
        }
      }
      if (!alive) break;
      if (!group_sync(bar, nb, rank, A.host_err)) {
        alive = false;
        break;
      }
      if (st.flags & kStDry) {  // Canvas::dryCanvas (Canvas.hxx:105-121)
        const int64_t n = (int64_t)J.rows * J.cols;
        for (int64_t i = gtid; i < n; i += gthreads) {
          float K[3], S[3], R[3];
          bool nz = false;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            K[c] = J.canvas.ks[c][i];
            S[c] = J.canvas.ks[3 + c][i];
            R[c] = J.R0.p[c][i];
            nz |= (K[c] != 0.0f) | (S[c] != 0.0f);
          }
          const double v = J.canvas.v[i];
          nz |= v != 0.0;
          if (!nz) continue;
          J.h[i] = h_plus(J.h[i], v);
          km_pixel(K, S, R, (float)v);
#pragma unroll
          for (int c = 0; c < 3; ++c) J.R0.p[c][i] = R[c];
#pragma unroll
          for (int c = 0; c < 6; ++c) J.canvas.ks[c][i] = 0.0f;
          J.canvas.v[i] = 0.0;
          if (J.use_snap) J.dirty[i] = 1;
        }
        if (!group_sync(bar, nb, rank, A.host_err)) {
          alive = false;
          break;
        }
      }
    }
    if (alive) {  // job stamped: every wave's stores drained, then one release
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (rank == 0 && t == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&A.done[jid], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  if (!alive) return;
  // ---- composites, taken from the shared counter
  for (;;) {
    __syncthreads();
    if (t == 0) {
      s_item = (int)__hip_atomic_fetch_add(A.item_next, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      s_ok = 1;
      if (s_item < A.n_items && !wait_done(&A.done[A.items[s_item].job], A.host_err, abort))
        s_ok = 0;
    }
    __syncthreads();
    const int item = s_item;
    if (item >= A.n_items || !s_ok) break;
    const CItem it = A.items[item];
    const BJob& J = A.jobs[it.job];
    // Renderer::compose (Renderer.hxx:26-41): four pixels per thread
    const int64_t b4 = it.begin / 4, e4 = it.end / 4;
    for (int64_t q = b4 + t; q < e4; q += kBT) {
      float4 r[3], k[3], s[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] = reinterpret_cast<const float4*>(J.R0.p[c])[q];
        k[c] = reinterpret_cast<const float4*>(J.canvas.ks[c])[q];
        s[c] = reinterpret_cast<const float4*>(J.canvas.ks[3 + c])[q];
      }
      const double2 v01 = reinterpret_cast<const double2*>(J.canvas.v)[2 * q];
      const double2 v23 = reinterpret_cast<const double2*>(J.canvas.v)[2 * q + 1];
      const float vv[4] = {(float)v01.x, (float)v01.y, (float)v23.x, (float)v23.y};
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        float K[3], S[3], R[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          K[c] = reinterpret_cast<const float*>(&k[c])[l];
          S[c] = reinterpret_cast<const float*>(&s[c])[l];
          R[c] = reinterpret_cast<const float*>(&r[c])[l];
        }
        km_pixel(K, S, R, vv[l]);
#pragma unroll
        for (int c = 0; c < 3; ++c) reinterpret_cast<float*>(&r[c])[l] = R[c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) reinterpret_cast<float4*>(J.out.p[c])[q] = r[c];
    }
    for (int64_t i = e4 * 4 + t; i < it.end; i += kBT) {  // tail of the canvas
      float K[3], S[3], R[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        K[c] = J.canvas.ks[c][i];
        S[c] = J.canvas.ks[3 + c][i];
        R[c] = J.R0.p[c][i];
      }
      km_pixel(K, S, R, (float)J.canvas.v[i]);
#pragma unroll
      for (int c = 0; c < 3; ++c) J.out.p[c][i] = R[c];
    }
  }
}

}  // namespace

// ------------------------------------------------------------------- host --
struct BatchState {
  int max_blocks16 = 0, max_blocks32 = 0;
  pnt_batch_config cfg{32, 16, 0, 1};  // the configuration runs use (pnt_batch_tune sets it)
  float* pool = nullptr;  // pickup maps of the waves' extra stroke chains
  int64_t pool_cap = 0;   // floats
  char* stage = nullptr;
  int64_t stage_cap = 0;
  char* dev = nullptr;
  int64_t dev_cap = 0;
  hipEvent_t uploaded = nullptr;
  hipEvent_t finished = nullptr;  // the last batch on the context's stream
};

static BatchState* batch_state(pnt_ctx ctx);

// A batch runs on the context's stream; a job whose canvas or brush belongs
// to another context (same device) is ordered after the batch on that
// context's stream too, so its next call there sees the batch's results.
static void order_job_streams(pnt_ctx ctx, const pnt_batch_job* jobs, int n) {
  BatchState& B = *batch_state(ctx);
  bool recorded = false;
  std::vector<hipStream_t> done;
  for (int j = 0; j < n; ++j)
    for (hipStream_t q : {jobs[j].canvas->ctx->stream, jobs[j].brush->ctx->stream}) {
      if (q == ctx->stream || std::find(done.begin(), done.end(), q) != done.end()) continue;
      if (!recorded) {
        PNT_HIP_CHECK(hipEventRecord(B.finished, ctx->stream));
        recorded = true;
      }
      PNT_HIP_CHECK(hipStreamWaitEvent(q, B.finished, 0));
      done.push_back(q);
    }
}

// cost of one stamp for a group of g blocks, in tile-iteration units
static double stamp_cost(int hr, int g, int tile) {
  const int side = 2 * hr + 1;
  const int tiles = ((side + tile - 1) / tile) * ((side + tile - 1) / tile);
  const double iters = std::ceil((double)tiles / g);
  return iters + (g > 1 ? 3.0 : 0.0) + 1.0;  // + group barrier, + ring scan / bookkeeping
}

struct Placement {
  std::vector<BGroup> groups;
  std::vector<int> job_list;
  std::vector<int> block_group;
  int nblocks = 0;
};

// Packs jobs into groups of blocks (within XCD lanes) to balance the
// makespan: each job starts with one block; blocks are then added to the
// job whose finish time is latest while that lowers it (up to max_group
// blocks a job); jobs beyond the block count share blocks, longest first
// onto the least loaded block.
static Placement place_jobs(const std::vector<int>& hr, const std::vector<int64_t>& nst,
                            int max_blocks, int max_group, int tile) {
  const int nj = (int)hr.size();
  Placement P;
  const int lanes = 8;
  int slots = std::max(lanes, (max_blocks / lanes) * lanes);
  std::vector<int> order(nj);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int j, int g) { return (double)nst[j] * stamp_cost(hr[j], g, tile); };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return cost(a, 1) > cost(b, 1); });
  std::vector<int> gsize;                 // blocks per group
  std::vector<std::vector<int>> gjobs;    // jobs per group
  std::vector<double> gload;
  if (nj <= slots) {
    for (int j : order) {
      gsize.push_back(1);
      gjobs.push_back({j});
      gload.push_back(cost(j, 1));
    }
    int used = nj;
    while (used < slots) {
      int worst = (int)(std::max_element(gload.begin(), gload.end()) - gload.begin());
      const int j = gjobs[worst][0];
      const int g = gsize[worst];
      if (g >= max_group) break;
      const double c1 = cost(j, g + 1);
      if (c1 >= gload[worst] * 0.97) break;
      gsize[worst] = g + 1;
      gload[worst] = c1;
      ++used;
    }
  } else {
    gsize.assign(slots, 1);
    gjobs.assign(slots, {});
    gload.assign(slots, 0.0);
    for (int j : order) {
      const int k = (int)(std::min_element(gload.begin(), gload.end()) - gload.begin());
      gjobs[k].push_back(j);
      gload[k] += cost(j, 1);
    }
  }
  // place groups into lanes: largest first into the lane with most room
  const int ng = (int)gsize.size();
  std::vector<int> gorder(ng);
  std::iota(gorder.begin(), gorder.end(), 0);
  std::sort(gorder.begin(), gorder.end(), [&](int a, int b) { return gsize[a] > gsize[b]; });
  std::vector<int> lane_fill(lanes, 0);
  const int per_lane = slots / lanes;
  P.groups.resize(ng);
  for (int gi : gorder) {
    int best = 0;
    for (int l = 1; l < lanes; ++l)
      if (lane_fill[l] < lane_fill[best]) best = l;
    if (lane_fill[best] + gsize[gi] > per_lane) gsize[gi] = per_lane - lane_fill[best];
    if (gsize[gi] < 1) throw Error(PNT_ESTATE, "batch placement overflow");
    BGroup& G = P.groups[gi];
    G.lane = best;
    G.k0 = lane_fill[best];
    G.nblocks = gsize[gi];
    G.job0 = (int)P.job_list.size();
    G.njobs = (int)gjobs[gi].size();
    for (int j : gjobs[gi]) P.job_list.push_back(j);
    lane_fill[best] += gsize[gi];
  }
  int maxk = *std::max_element(lane_fill.begin(), lane_fill.end());
  P.nblocks = maxk * lanes;
  P.block_group.assign(P.nblocks, -1);
  for (int gi = 0; gi < ng; ++gi) {
    const BGroup& G = P.groups[gi];
    for (int k = 0; k < G.nblocks; ++k) P.block_group[(G.k0 + k) * lanes + G.lane] = gi;
  }
  // padding blocks (lanes with fewer slots) get an empty group
  for (int b = 0; b < P.nblocks; ++b)
    if (P.block_group[b] < 0) {
      BGroup E{};
      E.lane = b % lanes;
      E.k0 = b / lanes;
      E.nblocks = 1;
      E.job0 = 0;
      E.njobs = 0;
      P.block_group[b] = (int)P.groups.size();
      P.groups.push_back(E);
    }
  return P;
}

// A batch, planned: the packed device image of jobs, groups, stamps, shared-
// pixel lists and composite items (counters follow it on the device).
struct BatchPlan {
  std::vector<char> blob;
  int64_t o_jobs, o_groups, o_bg, o_jl, o_st, o_dl, o_off, o_items, o_counters;
  int64_t n_counters, total;
  int njobs = 0, n_items = 0, nblocks = 0, tile = 32;
};

// Runs f(j) for every job, spread over up to 16 host threads by stamp count
// (planning is pure host work on each job's own brush and canvas).
template <class F>
static void parallel_jobs(const pnt_batch_job* jobs, int n, F&& f) {
  int64_t total = 0;
  for (int j = 0; j < n; ++j) total += jobs[j].n_stamps;
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<int64_t>({(int64_t)std::min(hw, 16), (int64_t)n, 1 + total / 2048});
  if (nt <= 1) {
    for (int j = 0; j < n; ++j) f(j);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::exception_ptr> errs(nt);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      try {
        for (int j; (j = next.fetch_add(1)) < n;) f(j);
      } catch (...) {
        errs[t] = std::current_exception();
        next.store(n);
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
}

// Plans a batch; readies each brush's snapshot and dirty bytes on the
// context's stream first (that work is done here, once).
static BatchPlan plan_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n,
                            const pnt_batch_config& cfg) {
  hipStream_t s = ctx->stream;
  BatchState& B = *batch_state(ctx);
  std::vector<PStamp> h_st;
  std::vector<int2> h_dl;
  std::vector<int> h_off;
  std::vector<BJob> h_jobs(n);
  std::vector<int> hr(n);
  std::vector<int64_t> nst(n);
  std::vector<CItem> items;
  // each job's stamps planned on its own (in parallel), then concatenated
  std::vector<std::vector<PStamp>> j_st(n);
  std::vector<std::vector<int2>> j_dl(n);
  std::vector<std::vector<int>> j_off(n);
  parallel_jobs(jobs, n, [&](int j) {
    plan_stamps(jobs[j].brush, jobs[j].canvas, jobs[j].stamps, jobs[j].n_stamps, j_st[j], j_dl[j],
                j_off[j]);
    for (int64_t k = 0; k < jobs[j].n_stamps; ++k)
      if (jobs[j].stamps[k].flags & PNT_STAMP_DRY_AFTER) j_st[j][k].flags |= kStDry;
  });
  for (int j = 0; j < n; ++j) {
    pnt_brush b = jobs[j].brush;
    pnt_canvas cv = jobs[j].canvas;
    const int s0 = (int)h_st.size();
    const int off0 = (int)h_off.size(), dl0 = (int)h_dl.size();
    for (PStamp p : j_st[j]) {
      if (p.dl_nlev > 0) p.dl_lev += off0;
      h_st.push_back(p);
    }
    for (int o : j_off[j]) h_off.push_back(o + dl0);
    h_dl.insert(h_dl.end(), j_dl[j].begin(), j_dl[j].end());
    std::vector<PStamp>().swap(j_st[j]);
    std::vector<int2>().swap(j_dl[j]);
    brush_prepare_snapshot(b, cv, s);
    BJob& J = h_jobs[j];
    J = BJob{};
    J.canvas = cv->layer();
    J.pick = b->pick_planes();
    J.snap = b->use_snap ? b->snap_planes() : cv->layer();
    J.R0 = cv->R0();
    J.h = cv->h();
    J.fp = b->fp;
    J.fb = FpBox{b->fbox[0], b->fbox[1], b->fbox[2], b->fbox[3]};
    J.dirty = b->dirty;
    J.use_snap = b->use_snap ? 1 : 0;
    J.rows = cv->rows;
    J.cols = cv->cols;
    J.hr = (b->size_map - 1) / 2;
    J.size_map = b->size_map;
    J.pickup_rate = b->pickup_rate;
    J.dep_rate = b->dep_rate;
    J.cap = b->cap;
    for (int c = 0; c < 3; ++c) {
      J.paintK[c] = b->paintK[c];
      J.paintS[c] = b->paintS[c];
      J.out.p[c] = jobs[j].out[c];
    }
    J.compose = jobs[j].out[0] != nullptr;
    J.s0 = s0;
    J.s1 = (int)h_st.size();
    hr[j] = J.hr;
    nst[j] = jobs[j].n_stamps;
  }
  const int tile = cfg.tile == 16 ? 16 : 32;
  const int max_blocks = tile == 16 ? B.max_blocks16 : B.max_blocks32;
  const Placement PL = place_jobs(hr, nst, max_blocks, std::max(1, cfg.max_group), tile);
  // composite items, jobs that finish stamping early first
  std::vector<int> by_cost(n);
  std::iota(by_cost.begin(), by_cost.end(), 0);
  std::vector<double> finish(n, 0.0);
  for (const BGroup& G : PL.groups) {
    double tacc = 0.0;
    for (int k = 0; k < G.njobs; ++k) {
      const int j = PL.job_list[G.job0 + k];
      tacc += (double)nst[j] * stamp_cost(hr[j], G.nblocks, tile);
      finish[j] = tacc;
    }
  }
  std::sort(by_cost.begin(), by_cost.end(), [&](int a, int b) { return finish[a] < finish[b]; });
  for (int j : by_cost) {
    if (!h_jobs[j].compose) continue;
    const int64_t npx = (int64_t)h_jobs[j].rows * h_jobs[j].cols;
    for (int64_t b0 = 0; b0 < npx; b0 += kComposeItem)
      items.push_back(CItem{j, 0, b0, std::min(npx, b0 + kComposeItem)});
  }
  BatchPlan P;
  auto align = [](int64_t x) { return (x + 255) / 256 * 256; };
  P.o_jobs = 0;
  P.o_groups = align(P.o_jobs + (int64_t)h_jobs.size() * sizeof(BJob));
  P.o_bg = align(P.o_groups + (int64_t)PL.groups.size() * sizeof(BGroup));
  P.o_jl = align(P.o_bg + (int64_t)PL.block_group.size() * sizeof(int));
  P.o_st = align(P.o_jl + (int64_t)PL.job_list.size() * sizeof(int));
  P.o_dl = align(P.o_st + (int64_t)h_st.size() * sizeof(PStamp));
  P.o_off = align(P.o_dl + (int64_t)h_dl.size() * sizeof(int2));
  P.o_items = align(P.o_off + (int64_t)h_off.size() * sizeof(int));
  P.o_counters = align(P.o_items + (int64_t)items.size() * sizeof(CItem));
  P.n_counters = 1 + n + (int64_t)PL.groups.size() * kBarWords;
  P.total = align(P.o_counters + P.n_counters * (int64_t)sizeof(unsigned));
  P.blob.assign((size_t)P.o_counters, 0);
  auto put = [&](int64_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(P.blob.data() + off, src, bytes);
  };
  put(P.o_jobs, h_jobs.data(), h_jobs.size() * sizeof(BJob));
  put(P.o_groups, PL.groups.data(), PL.groups.size() * sizeof(BGroup));
  put(P.o_bg, PL.block_group.data(), PL.block_group.size() * sizeof(int));
  put(P.o_jl, PL.job_list.data(), PL.job_list.size() * sizeof(int));
  put(P.o_st, h_st.data(), h_st.size() * sizeof(PStamp));
  put(P.o_dl, h_dl.data(), h_dl.size() * sizeof(int2));
  put(P.o_off, h_off.data(), h_off.size() * sizeof(int));
  put(P.o_items, items.data(), items.size() * sizeof(CItem));
  P.njobs = n;
  P.n_items = (int)items.size();
  P.nblocks = PL.nblocks;
  P.tile = tile;
  return P;
}

// Enqueues a planned batch on `s`: plan copy (staging -> dev), counters
// cleared by a kernel, then k_batch -- cooperative, or as a plain launch of
// at most the co-resident block count (inside a captured graph).
static void enqueue_plan(const BatchPlan& P, const char* stage, char* dev, hipStream_t s,
                         unsigned* host_err, bool cooperative) {
  PNT_HIP_CHECK(hipMemcpyAsync(dev, stage, (size_t)P.o_counters, hipMemcpyHostToDevice, s));
  unsigned* counters = reinterpret_cast<unsigned*>(dev + P.o_counters);
  launch_fill_u32(counters, P.n_counters, 0u, s);
  BArgs A;
  A.jobs = reinterpret_cast<const BJob*>(dev + P.o_jobs);
  A.groups = reinterpret_cast<const BGroup*>(dev + P.o_groups);
  A.block_group = reinterpret_cast<const int*>(dev + P.o_bg);
  A.job_list = reinterpret_cast<const int*>(dev + P.o_jl);
  A.st = reinterpret_cast<const PStamp*>(dev + P.o_st);
  A.dl = reinterpret_cast<const int2*>(dev + P.o_dl);
  A.dl_off = reinterpret_cast<const int*>(dev + P.o_off);
  A.items = reinterpret_cast<const CItem*>(dev + P.o_items);
  A.n_items = P.n_items;
  A.item_next = counters;
  A.done = counters + 1;
  A.bars = counters + 1 + P.njobs;
  A.host_err = host_err;
  const void* fn = P.tile == 16 ? reinterpret_cast<const void*>(k_batch<16>)
                                : reinterpret_cast<const void*>(k_batch<32>);
  const unsigned threads = (unsigned)(P.tile * P.tile);
  if (cooperative) {
    void* kargs[] = {&A};
    PNT_HIP_CHECK(hipLaunchCooperativeKernel(fn, dim3(P.nblocks), dim3(threads), kargs, 0, s));
  } else if (P.tile == 16) {
    k_batch<16><<<P.nblocks, threads, 0, s>>>(A);
  } else {
    k_batch<32><<<P.nblocks, threads, 0, s>>>(A);
  }
  PNT_HIP_CHECK(hipGetLastError());
}

// Host mirrors of brush and canvas state after a batch of these jobs.
static void finish_jobs(const pnt_batch_job* jobs, int n) {
  for (int j = 0; j < n; ++j) {
    pnt_brush b = jobs[j].brush;
    pnt_canvas cv = jobs[j].canvas;
    for (int64_t k = 0; k < jobs[j].n_stamps; ++k)
      if (jobs[j].stamps[k].flags & PNT_STAMP_DIP)
        for (int c = 0; c < 3; ++c) {
          b->paintK[c] = jobs[j].stamps[k].K[c];
          b->paintS[c] = jobs[j].stamps[k].S[c];
        }
    cv->epoch++;
    if (b->use_snap) {
      b->dirty_canvas = cv;
      b->dirty_epoch = cv->epoch;
    }
  }
}

// Grows a device buffer of the batch state (after the stream's earlier work).
static void grow_dev(char*& p, int64_t& cap, int64_t need, hipStream_t s) {
  if (need <= cap) return;
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  if (p) PNT_HIP_CHECK(hipFree(p));
  p = nullptr;
  cap = std::max<int64_t>(need, 2 * cap);
  PNT_HIP_CHECK(hipMalloc((void**)&p, (size_t)cap));
}

static void grow_stage(char*& p, int64_t& cap, int64_t need) {
  if (need <= cap) return;
  if (p) PNT_HIP_CHECK(hipHostFree(p));
  p = nullptr;
  cap = std::max<int64_t>(need, 2 * cap);
  PNT_HIP_CHECK(hipHostMalloc((void**)&p, (size_t)cap));
}

// The batch as waves (batchwave.hip) on the context's stream.  With `prof`,
// a timing event after every wave, read back before returning.
static void run_batch_waves(pnt_ctx ctx, const pnt_batch_job* jobs, int n, int chunk,
                            BatchWaveProfile* prof = nullptr) {
  hipStream_t s = ctx->stream;
  BatchState& B = *batch_state(ctx);
  const int64_t need = plan_waves_pool(jobs, n, chunk);
  if (need > B.pool_cap) {
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    dev_free(B.pool);
    B.pool = nullptr;
    B.pool_cap = std::max<int64_t>(need, 2 * B.pool_cap);
    B.pool = dev_alloc(B.pool_cap);
  }
  const BWPlan W = plan_waves(jobs, n, chunk, B.pool, s);
  PNT_HIP_CHECK(hipEventSynchronize(B.uploaded));  // the previous plan's copy has read the stage
  grow_stage(B.stage, B.stage_cap, W.total);
  grow_dev(B.dev, B.dev_cap, W.total, s);
  std::memcpy(B.stage, W.blob.data(), W.blob.size());
  struct Events {  // destroyed on every way out, a failed call's included
    std::vector<hipEvent_t> e;
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } events;
  std::vector<hipEvent_t>& ev = events.e;
  if (prof) {
    ev.assign(W.wave_blocks.size() + 1, nullptr);
    for (hipEvent_t& e : ev) PNT_HIP_CHECK(hipEventCreate(&e));
  }
  enqueue_waves(W, B.stage, B.dev, s, prof ? ev.data() : nullptr);
  PNT_HIP_CHECK(hipEventRecord(B.uploaded, s));
  order_job_streams(ctx, jobs, n);
  finish_jobs(jobs, n);
  if (prof) {
    PNT_HIP_CHECK(hipEventSynchronize(ev.back()));
    prof->wave_end_ms.assign(W.wave_blocks.size(), 0.0f);
    for (size_t w = 0; w < W.wave_blocks.size(); ++w)
      PNT_HIP_CHECK(hipEventElapsedTime(&prof->wave_end_ms[w], ev[0], ev[w + 1]));
    prof->wave_blocks = W.wave_blocks;
    prof->job_first = W.job_first;
    prof->job_last = W.job_last;
  }
}

void profile_batch_waves(pnt_ctx ctx, const pnt_batch_job* jobs, int n, BatchWaveProfile* prof) {
  if (*(volatile unsigned*)ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  const pnt_batch_config C = batch_state(ctx)->cfg;
  run_batch_waves(ctx, jobs, n, C.waves ? C.chunk : 4, prof);
}

void run_batch_one_launch(pnt_ctx ctx, const pnt_batch_job* jobs, int n,
                          const pnt_batch_config* cfg) {
  hipStream_t s = ctx->stream;
  if (*(volatile unsigned*)ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  BatchState& B = *batch_state(ctx);
  const pnt_batch_config C = cfg ? *cfg : B.cfg;
  bool clocked = false;  // checkDry runs in the waves, not in the persistent kernel
  for (int j = 0; j < n; ++j) clocked = clocked || jobs[j].canvas->clock != PNT_CLOCK_OFF;
  if (C.waves || clocked) {
    run_batch_waves(ctx, jobs, n, C.waves ? C.chunk : 4);
    return;
  }
  const BatchPlan P = plan_batch(ctx, jobs, n, C);
  PNT_HIP_CHECK(hipEventSynchronize(B.uploaded));
  if (P.total > B.stage_cap) {
    if (B.stage) PNT_HIP_CHECK(hipHostFree(B.stage));
    B.stage = nullptr;
    B.stage_cap = std::max<int64_t>(P.total, 2 * B.stage_cap);
    PNT_HIP_CHECK(hipHostMalloc((void**)&B.stage, (size_t)B.stage_cap));
  }
  if (P.total > B.dev_cap) {
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    if (B.dev) PNT_HIP_CHECK(hipFree(B.dev));
    B.dev = nullptr;
    B.dev_cap = std::max<int64_t>(P.total, 2 * B.dev_cap);
    PNT_HIP_CHECK(hipMalloc((void**)&B.dev, (size_t)B.dev_cap));
  }
  std::memcpy(B.stage, P.blob.data(), P.blob.size());
  unsigned* host_err = nullptr;
  PNT_HIP_CHECK(hipHostGetDevicePointer((void**)&host_err, ctx->kernel_error, 0));
  enqueue_plan(P, B.stage, B.dev, s, host_err, true);
  PNT_HIP_CHECK(hipEventRecord(B.uploaded, s));
  order_job_streams(ctx, jobs, n);
  finish_jobs(jobs, n);
}

void set_batch_config(pnt_ctx ctx, const pnt_batch_config& c) { batch_state(ctx)->cfg = c; }
pnt_batch_config get_batch_config(pnt_ctx ctx) { return batch_state(ctx)->cfg; }

// ----------------------------------------------------------- autotuner --
// Times candidate (tile, max group) configurations on the batch itself and
// keeps the fastest for the context.  The state every candidate changes --
// canvas planes, pickup maps, snapshots, dirty bytes, host mirrors -- is
// saved first and restored after each candidate, so tuning leaves the jobs
// as it found them.
struct SavedJob {
  float* canvas = nullptr;  // 11 planes
  float* pick = nullptr;    // 7 planes
  float* snap = nullptr;    // 7 planes
  uint8_t* dirty = nullptr;
  int64_t canvas_n = 0, pick_n = 0, snap_n = 0, dirty_n = 0;
  float paintK[3], paintS[3];
  uint64_t epoch = 0, dirty_epoch = 0;
  const pnt_canvas_s* dirty_canvas = nullptr;
};

static void save_or_restore(const pnt_batch_job* jobs, int n, std::vector<SavedJob>& sv,
                            bool save, hipStream_t s) {
  auto copy = [&](void* dst, const void* src, int64_t bytes) {
    if (bytes > 0) PNT_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, s));
  };
  for (int j = 0; j < n; ++j) {
    pnt_brush b = jobs[j].brush;
    pnt_canvas cv = jobs[j].canvas;
    SavedJob& S = sv[j];
    if (save) {
      S.canvas_n = kCanvasStrides * cv->stride;
      S.pick_n = 7 * b->mstride;
      S.snap_n = b->snap ? kLayerStrides * b->snap_stride : 0;
      S.dirty_n = b->dirty ? b->dirty_cap : 0;
      S.canvas = dev_alloc(S.canvas_n);
      S.pick = dev_alloc(S.pick_n);
      if (S.snap_n) S.snap = dev_alloc(S.snap_n);
      if (S.dirty_n) PNT_HIP_CHECK(hipMalloc((void**)&S.dirty, (size_t)S.dirty_n));
      copy(S.canvas, cv->base, S.canvas_n * 4);
      copy(S.pick, b->pick, S.pick_n * 4);
      copy(S.snap, b->snap, S.snap_n * 4);
      copy(S.dirty, b->dirty, S.dirty_n);
      for (int c = 0; c < 3; ++c) {
        S.paintK[c] = b->paintK[c];
        S.paintS[c] = b->paintS[c];
      }
      S.epoch = cv->epoch;
      S.dirty_epoch = b->dirty_epoch;
      S.dirty_canvas = b->dirty_canvas;
    } else {
      copy(cv->base, S.canvas, S.canvas_n * 4);
      copy(b->pick, S.pick, S.pick_n * 4);
      copy(b->snap, S.snap, S.snap_n * 4);
      copy(b->dirty, S.dirty, S.dirty_n);
      for (int c = 0; c < 3; ++c) {
        b->paintK[c] = S.paintK[c];
        b->paintS[c] = S.paintS[c];
      }
      cv->epoch = S.epoch;
      b->dirty_epoch = S.dirty_epoch;
      b->dirty_canvas = S.dirty_canvas;
    }
  }
}

void tune_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n, pnt_batch_config* best,
                double* best_ms) {
  hipStream_t s = ctx->stream;
  BatchState& B = *batch_state(ctx);
  // the snapshot and dirty bytes the runs will use exist before the save
  for (int j = 0; j < n; ++j) brush_prepare_snapshot(jobs[j].brush, jobs[j].canvas, s);
  std::vector<SavedJob> sv(n);
  save_or_restore(jobs, n, sv, true, s);
  // one persistent launch: tile x workgroups per job; waves: strokes of one
  // canvas in flight (the stamp-chunk of a wave)
  const pnt_batch_config cands[] = {{32, 16, 0, 1}, {32, 4, 0, 1}, {32, 1, 0, 1},
                                    {16, 64, 0, 1}, {16, 16, 0, 1}, {16, 4, 0, 1},
                                    {16, 1, 1, 1},  {16, 1, 1, 2},  {16, 1, 1, 4}};
  hipEvent_t e0, e1;
  PNT_HIP_CHECK(hipEventCreate(&e0));
  PNT_HIP_CHECK(hipEventCreate(&e1));
  double best_t = 1e30;
  pnt_batch_config best_c = B.cfg;
  try {
    // untimed first run: grows the plan buffers, warms the code and caches
    run_batch_one_launch(ctx, jobs, n, &cands[0]);
    save_or_restore(jobs, n, sv, false, s);
    for (const pnt_batch_config& c : cands) {
      PNT_HIP_CHECK(hipEventRecord(e0, s));
      run_batch_one_launch(ctx, jobs, n, &c);
      PNT_HIP_CHECK(hipEventRecord(e1, s));
      PNT_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      PNT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (*(volatile unsigned*)ctx->kernel_error)
        throw Error(PNT_EKERNEL, "a batch kernel aborted while tuning");
      if (ms < best_t) {
        best_t = ms;
        best_c = c;
      }
      save_or_restore(jobs, n, sv, false, s);
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto& S : sv) {
      dev_free(S.canvas);
      dev_free(S.pick);
      dev_free(S.snap);
      dev_free(S.dirty);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  for (auto& S : sv) {
    dev_free(S.canvas);
    dev_free(S.pick);
    dev_free(S.snap);
    dev_free(S.dirty);
  }
  PNT_HIP_CHECK(hipEventDestroy(e0));
  PNT_HIP_CHECK(hipEventDestroy(e1));
  B.cfg = best_c;
  if (best) *best = best_c;
  if (best_ms) *best_ms = best_t;
}

// ------------------------------------------------------- captured graph --
// What a captured plan baked in about a brush: buffers and parameters.
struct BrushKey {
  const float* fp;
  const float* pick;
  const float* snap;
  const uint8_t* dirty;
  int64_t snap_stride;
  int size_map;
  bool use_snap;
  double pickup_rate, dep_rate, cap;
  bool operator==(const BrushKey& o) const {
    return fp == o.fp && pick == o.pick && snap == o.snap && dirty == o.dirty &&
           snap_stride == o.snap_stride && size_map == o.size_map && use_snap == o.use_snap &&
           pickup_rate == o.pickup_rate && dep_rate == o.dep_rate && cap == o.cap;
  }
};

static BrushKey brush_key(const pnt_brush_s* b) {
  return BrushKey{b->fp,          b->pick,     b->snap,     b->dirty,        b->snap_stride,
                  b->size_map,    b->use_snap, b->pickup_rate, b->dep_rate, b->cap};
}

struct BatchGraph {
  std::vector<BrushKey> brushes;
  pnt_ctx ctx = nullptr;
  bool retained = false;
  std::vector<pnt_batch_job> jobs;
  std::vector<std::vector<pnt_stamp>> stamps;
  std::vector<uint64_t> epochs;  // canvas epochs a launch expects
  char* stage = nullptr;
  char* dev = nullptr;
  float* pool = nullptr;  // waves: pickup maps of the extra stroke chains
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Plans the batch once and captures plan copy + counter clear + k_batch into
// a graph; each launch replays the same jobs on the canvases' current state.
// The captured k_batch is a plain launch sized to the co-resident block
// count (every wait in it is bounded).
BatchGraph* capture_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n) {
  for (int j = 0; j < n; ++j)
    PNT_REQUIRE(jobs[j].n_stamps == 0 || (jobs[j].stamps[0].flags & PNT_STAMP_DIP),
                "a captured batch needs every stamp list to start with a dip "
                "(its paint is fixed at capture)");
  BatchState& B = *batch_state(ctx);
  auto* G = new BatchGraph;
  G->ctx = ctx;
  try {
    G->jobs.assign(jobs, jobs + n);
    G->stamps.resize(n);
    for (int j = 0; j < n; ++j) {  // own copies of the stamp lists
      G->stamps[j].assign(jobs[j].stamps, jobs[j].stamps + jobs[j].n_stamps);
      G->jobs[j].stamps = G->stamps[j].data();
    }
    hipStream_t s = ctx->stream;
    pnt_batch_config C = B.cfg;
    for (int j = 0; j < n; ++j)  // checkDry runs in the waves, not in the persistent kernel
      if (jobs[j].canvas->clock != PNT_CLOCK_OFF && !C.waves) C = pnt_batch_config{16, 1, 1, 4};
    BatchPlan P;
    BWPlan W;
    int64_t total = 0;
    if (C.waves) {  // the waves and their own pool of pickup maps
      G->pool = dev_alloc(std::max<int64_t>(plan_waves_pool(G->jobs.data(), n, C.chunk), 1));
      W = plan_waves(G->jobs.data(), n, C.chunk, G->pool, s);
      total = W.total;
    } else {
      P = plan_batch(ctx, G->jobs.data(), n, C);
      total = P.total;
    }
    PNT_HIP_CHECK(hipStreamSynchronize(s));  // snapshot preparation done before capture
    PNT_HIP_CHECK(hipHostMalloc((void**)&G->stage, (size_t)std::max<int64_t>(total, 1)));
    PNT_HIP_CHECK(hipMalloc((void**)&G->dev, (size_t)std::max<int64_t>(total, 1)));
    if (C.waves)
      std::memcpy(G->stage, W.blob.data(), W.blob.size());
    else
      std::memcpy(G->stage, P.blob.data(), P.blob.size());
    unsigned* host_err = nullptr;
    PNT_HIP_CHECK(hipHostGetDevicePointer((void**)&host_err, ctx->kernel_error, 0));
    PNT_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      if (C.waves)
        enqueue_waves(W, G->stage, G->dev, s);
      else
        enqueue_plan(P, G->stage, G->dev, s, host_err, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    PNT_HIP_CHECK(hipStreamEndCapture(s, &G->graph));
    PNT_HIP_CHECK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
    for (int j = 0; j < n; ++j) {
      G->epochs.push_back(G->jobs[j].canvas->epoch);
      G->brushes.push_back(brush_key(G->jobs[j].brush));
    }
  } catch (...) {
    release_batch_graph(G);
    throw;
  }
  ctx_retain(ctx);
  G->retained = true;
  return G;
}

void launch_batch_graph(BatchGraph* G) {
  pnt_ctx ctx = G->ctx;
  if (*(volatile unsigned*)ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  for (size_t j = 0; j < G->jobs.size(); ++j) {
    const pnt_batch_job& J = G->jobs[j];
    PNT_REQUIRE(J.canvas->epoch == G->epochs[j] && brush_key(J.brush) == G->brushes[j] &&
                    (!J.brush->use_snap || (J.brush->dirty_canvas == J.canvas &&
                                            J.brush->dirty_epoch == J.canvas->epoch)),
                "a canvas or brush of the captured batch changed outside it; capture again");
  }
  PNT_HIP_CHECK(hipGraphLaunch(G->exec, ctx->stream));
  order_job_streams(ctx, G->jobs.data(), (int)G->jobs.size());
  finish_jobs(G->jobs.data(), (int)G->jobs.size());
  for (size_t j = 0; j < G->jobs.size(); ++j) G->epochs[j] = G->jobs[j].canvas->epoch;
}

void release_batch_graph(BatchGraph*& G) {
  if (!G) return;
  if (G->ctx) (void)hipStreamSynchronize(G->ctx->stream);
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  if (G->stage) (void)hipHostFree(G->stage);
  if (G->dev) (void)hipFree(G->dev);
  if (G->pool) (void)hipFree(G->pool);
  pnt_ctx ctx = G->retained ? G->ctx : nullptr;
  delete G;
  G = nullptr;
  if (ctx) ctx_release(ctx);
}

// The context's BatchState (plan buffers of its batches), made on first use;
// it lives in the context handle and goes with it.
static BatchState* batch_state(pnt_ctx ctx) {
  if (ctx->batch) return ctx->batch;
  auto* B = new BatchState;
  try {
    PNT_HIP_CHECK(hipEventCreateWithFlags(&B->uploaded, hipEventDisableTiming));
    PNT_HIP_CHECK(hipEventCreateWithFlags(&B->finished, hipEventDisableTiming));
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    PNT_HIP_CHECK(hipGetDevice(&dev));
    PNT_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) throw Error(PNT_ESTATE, "device has no cooperative launch");
    PNT_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_batch<32>), 1024, 0));
    B->max_blocks32 = ncu * std::max(per_cu, 1);
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_batch<16>), 256, 0));
    B->max_blocks16 = ncu * std::max(per_cu, 1);
  } catch (...) {
    if (B->uploaded) (void)hipEventDestroy(B->uploaded);
    if (B->finished) (void)hipEventDestroy(B->finished);
    delete B;
    throw;
  }
  ctx->batch = B;
  return B;
}

void batch_release(pnt_ctx ctx) {
  BatchState* B = ctx->batch;
  if (!B) return;
  if (B->stage) (void)hipHostFree(B->stage);
  if (B->dev) (void)hipFree(B->dev);
  if (B->pool) (void)hipFree(B->pool);
  if (B->uploaded) (void)hipEventDestroy(B->uploaded);
  if (B->finished) (void)hipEventDestroy(B->finished);
  delete B;
  ctx->batch = nullptr;
}

}  // namespace pnt
