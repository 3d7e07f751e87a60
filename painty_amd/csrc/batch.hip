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
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "stampcore.hpp"

namespace pnt {

enum : uint32_t { kStDry = 8u };  // dryCanvas() after this stamp

namespace {

constexpr int kWT = 32;                  // window tile side
constexpr int kBT = kWT * kWT;           // threads per block
constexpr int64_t kComposeItem = 1 << 18;  // pixels per composite work item

struct BJob {
  Layer7 canvas, pick, snap;
  Rgb3 R0, out;
  float* h;
  const float* fp;
  uint8_t* dirty;
  int compose, use_snap;
  int rows, cols, hr, size_map;
  float pickup_rate, dep_rate, cap;
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

__global__ __launch_bounds__(kBT) void k_batch(BArgs A) {
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
        window_tile<kWT>(T, J.pick.p, J.fp, g, st, dup, pull, tile, tiles_x, sm, pk, ps);
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
            K[c] = J.canvas.p[c][i];
            S[c] = J.canvas.p[3 + c][i];
            R[c] = J.R0.p[c][i];
            nz |= (K[c] != 0.0f) | (S[c] != 0.0f);
          }
          const float v = J.canvas.p[6][i];
          nz |= v != 0.0f;
          if (!nz) continue;
          J.h[i] += v;
          km_pixel(K, S, R, v);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            J.R0.p[c][i] = R[c];
            J.canvas.p[c][i] = 0.0f;
            J.canvas.p[3 + c][i] = 0.0f;
          }
          J.canvas.p[6][i] = 0.0f;
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
        k[c] = reinterpret_cast<const float4*>(J.canvas.p[c])[q];
        s[c] = reinterpret_cast<const float4*>(J.canvas.p[3 + c])[q];
      }
      const float4 v = reinterpret_cast<const float4*>(J.canvas.p[6])[q];
      const float vv[4] = {v.x, v.y, v.z, v.w};
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
        K[c] = J.canvas.p[c][i];
        S[c] = J.canvas.p[3 + c][i];
        R[c] = J.R0.p[c][i];
      }
      km_pixel(K, S, R, J.canvas.p[6][i]);
#pragma unroll
      for (int c = 0; c < 3; ++c) J.out.p[c][i] = R[c];
    }
  }
}

}  // namespace

// ------------------------------------------------------------------- host --
struct BatchState {
  int max_blocks = 0;
  char* stage = nullptr;
  int64_t stage_cap = 0;
  char* dev = nullptr;
  int64_t dev_cap = 0;
  hipEvent_t uploaded = nullptr;
};

static BatchState* batch_state(pnt_ctx ctx);

// cost of one stamp for a group of g blocks, in tile-iteration units
static double stamp_cost(int hr, int g) {
  const int side = 2 * hr + 1;
  const int tiles = ((side + kWT - 1) / kWT) * ((side + kWT - 1) / kWT);
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
// job whose finish time is latest while that lowers it; jobs beyond the
// block count share blocks, longest first onto the least loaded block.
static Placement place_jobs(const std::vector<int>& hr, const std::vector<int64_t>& nst,
                            int max_blocks) {
  const int nj = (int)hr.size();
  Placement P;
  const int lanes = 8;
  int slots = std::max(lanes, (max_blocks / lanes) * lanes);
  std::vector<int> order(nj);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int j, int g) { return (double)nst[j] * stamp_cost(hr[j], g); };
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
      if (g >= 16) break;
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

void run_batch_one_launch(pnt_ctx ctx, const pnt_batch_job* jobs, int n) {
  hipStream_t s = ctx->stream;
  if (*(volatile unsigned*)ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  BatchState& B = *batch_state(ctx);
  // ---- plan
  std::vector<PStamp> h_st;
  std::vector<int2> h_dl;
  std::vector<int> h_off;
  std::vector<BJob> h_jobs(n);
  std::vector<int> hr(n);
  std::vector<int64_t> nst(n);
  std::vector<CItem> items;
  for (int j = 0; j < n; ++j) {
    pnt_brush b = jobs[j].brush;
    pnt_canvas cv = jobs[j].canvas;
    const int s0 = (int)h_st.size();
    plan_stamps(b, cv, jobs[j].stamps, jobs[j].n_stamps, h_st, h_dl, h_off);
    for (int64_t k = 0; k < jobs[j].n_stamps; ++k)
      if (jobs[j].stamps[k].flags & PNT_STAMP_DRY_AFTER) h_st[s0 + k].flags |= kStDry;
    brush_prepare_snapshot(b, cv, s);
    BJob& J = h_jobs[j];
    J = BJob{};
    J.canvas = cv->layer();
    J.pick = b->pick_planes();
    J.snap = b->use_snap ? b->snap_planes() : cv->layer();
    J.R0 = cv->R0();
    J.h = cv->h();
    J.fp = b->fp;
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
  const Placement PL = place_jobs(hr, nst, B.max_blocks);
  // composite items, jobs that finish stamping early first
  std::vector<int> by_cost(n);
  std::iota(by_cost.begin(), by_cost.end(), 0);
  std::vector<double> finish(n, 0.0);
  for (const BGroup& G : PL.groups) {
    double tacc = 0.0;
    for (int k = 0; k < G.njobs; ++k) {
      const int j = PL.job_list[G.job0 + k];
      tacc += (double)nst[j] * stamp_cost(hr[j], G.nblocks);
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
  // ---- one device buffer for the plan, through pinned staging
  auto align = [](int64_t x) { return (x + 255) / 256 * 256; };
  const int64_t o_jobs = 0;
  const int64_t o_groups = align(o_jobs + (int64_t)h_jobs.size() * sizeof(BJob));
  const int64_t o_bg = align(o_groups + (int64_t)PL.groups.size() * sizeof(BGroup));
  const int64_t o_jl = align(o_bg + (int64_t)PL.block_group.size() * sizeof(int));
  const int64_t o_st = align(o_jl + (int64_t)PL.job_list.size() * sizeof(int));
  const int64_t o_dl = align(o_st + (int64_t)h_st.size() * sizeof(PStamp));
  const int64_t o_off = align(o_dl + (int64_t)h_dl.size() * sizeof(int2));
  const int64_t o_items = align(o_off + (int64_t)h_off.size() * sizeof(int));
  const int64_t o_counters = align(o_items + (int64_t)items.size() * sizeof(CItem));
  const int64_t n_counters = 1 + n + (int64_t)PL.groups.size() * kBarWords;
  const int64_t total = align(o_counters + n_counters * (int64_t)sizeof(unsigned));
  PNT_HIP_CHECK(hipEventSynchronize(B.uploaded));
  if (total > B.stage_cap) {
    if (B.stage) PNT_HIP_CHECK(hipHostFree(B.stage));
    B.stage = nullptr;
    B.stage_cap = std::max<int64_t>(total, 2 * B.stage_cap);
    PNT_HIP_CHECK(hipHostMalloc((void**)&B.stage, (size_t)B.stage_cap));
  }
  if (total > B.dev_cap) {
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    if (B.dev) PNT_HIP_CHECK(hipFree(B.dev));
    B.dev = nullptr;
    B.dev_cap = std::max<int64_t>(total, 2 * B.dev_cap);
    PNT_HIP_CHECK(hipMalloc((void**)&B.dev, (size_t)B.dev_cap));
  }
  auto put = [&](int64_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(B.stage + off, src, bytes);
  };
  put(o_jobs, h_jobs.data(), h_jobs.size() * sizeof(BJob));
  put(o_groups, PL.groups.data(), PL.groups.size() * sizeof(BGroup));
  put(o_bg, PL.block_group.data(), PL.block_group.size() * sizeof(int));
  put(o_jl, PL.job_list.data(), PL.job_list.size() * sizeof(int));
  put(o_st, h_st.data(), h_st.size() * sizeof(PStamp));
  put(o_dl, h_dl.data(), h_dl.size() * sizeof(int2));
  put(o_off, h_off.data(), h_off.size() * sizeof(int));
  put(o_items, items.data(), items.size() * sizeof(CItem));
  PNT_HIP_CHECK(hipMemcpyAsync(B.dev, B.stage, (size_t)o_counters, hipMemcpyHostToDevice, s));
  PNT_HIP_CHECK(hipEventRecord(B.uploaded, s));
  unsigned* counters = reinterpret_cast<unsigned*>(B.dev + o_counters);
  launch_fill_u32(counters, n_counters, 0u, s);

  BArgs A;
  A.jobs = reinterpret_cast<const BJob*>(B.dev + o_jobs);
  A.groups = reinterpret_cast<const BGroup*>(B.dev + o_groups);
  A.block_group = reinterpret_cast<const int*>(B.dev + o_bg);
  A.job_list = reinterpret_cast<const int*>(B.dev + o_jl);
  A.st = reinterpret_cast<const PStamp*>(B.dev + o_st);
  A.dl = reinterpret_cast<const int2*>(B.dev + o_dl);
  A.dl_off = reinterpret_cast<const int*>(B.dev + o_off);
  A.items = reinterpret_cast<const CItem*>(B.dev + o_items);
  A.n_items = (int)items.size();
  A.item_next = counters;
  A.done = counters + 1;
  A.bars = counters + 1 + n;
  PNT_HIP_CHECK(hipHostGetDevicePointer((void**)&A.host_err, ctx->kernel_error, 0));
  void* kargs[] = {&A};
  PNT_HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_batch),
                                           dim3(PL.nblocks), dim3(kBT), kargs, 0, s));
  // host mirrors of the brush and canvas state
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

// One BatchState per context (the plan buffers of its batches).
static std::vector<std::pair<pnt_ctx, BatchState*>> g_batch_states;

static BatchState* batch_state(pnt_ctx ctx) {
  for (auto& e : g_batch_states)
    if (e.first == ctx) return e.second;
  auto* B = new BatchState;
  PNT_HIP_CHECK(hipEventCreateWithFlags(&B->uploaded, hipEventDisableTiming));
  int dev = 0, coop = 0, ncu = 0, per_cu = 0;
  PNT_HIP_CHECK(hipGetDevice(&dev));
  PNT_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) throw Error(PNT_ESTATE, "device has no cooperative launch");
  PNT_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(k_batch), kBT, 0));
  B->max_blocks = ncu * std::max(per_cu, 1);
  g_batch_states.push_back({ctx, B});
  return B;
}

void batch_release(pnt_ctx ctx) {
  for (size_t i = 0; i < g_batch_states.size(); ++i)
    if (g_batch_states[i].first == ctx) {
      BatchState* B = g_batch_states[i].second;
      if (B->stage) (void)hipHostFree(B->stage);
      if (B->dev) (void)hipFree(B->dev);
      if (B->uploaded) (void)hipEventDestroy(B->uploaded);
      delete B;
      g_batch_states.erase(g_batch_states.begin() + i);
      return;
    }
}

}  // namespace pnt
