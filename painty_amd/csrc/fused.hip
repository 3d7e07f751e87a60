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
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"

namespace pnt {

namespace {

struct FArgs {
  const PStamp* st;
  Layer7 canvas, snap, pick;  // snap == canvas planes when the snapshot is off
  const float* fp;
  uint8_t* dirty;
  const int2* dl;
  const int* dl_off;
  int use_snap, rows, cols, hr, size_map, tiles_x, ntiles;
  float pickup_rate, dep_rate, cap;
  float pk[3], ps[3];  // brush paint (the last dip)
};

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

// Stamp j: window tiles on blocks [0, win_blocks) (win_blocks is 0 or
// ntiles), ring scan on the next ring_blocks blocks.
template <int kT>
__global__ __launch_bounds__(kT* kT) void k_stamp_fused(FArgs A, int j, int win_blocks,
                                                        int ring_blocks) {
  __shared__ int sm[kT + 2][kT + 2];
  const PStamp& st = A.st[j];
  const StampGeom g = geom_of(A, st);
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  const bool pull = A.use_snap && !dup;
  if ((int)blockIdx.x < win_blocks) {
    const VisitTarget T{&A.canvas, &A.snap, A.dirty, A.use_snap, A.pickup_rate, A.dep_rate,
                        A.cap};
    window_tile<kT>(T, A.pick.p, A.fp, g, st, dup, pull, tile_of_block(blockIdx.x, A.ntiles),
                    A.tiles_x, sm, A.pk, A.ps);
  } else {
    ring_scan<2>(A.canvas, A.snap, A.dirty, A.fp, g, st.box, pull,
              ((int)blockIdx.x - win_blocks) * kT * kT + (int)threadIdx.x, ring_blocks * kT * kT);
  }
}

// Stamp j's visits on a duplicated canvas column/row 0 and their map-cell
// partners, in reference order by dependency level; one workgroup.
__global__ __launch_bounds__(256) void k_stamp_levels(FArgs A, int j) {
  const PStamp& st = A.st[j];
  const StampGeom g = geom_of(A, st);
  const VisitTarget T{&A.canvas, &A.snap, A.dirty, A.use_snap, A.pickup_rate, A.dep_rate, A.cap};
  for (int L = 0; L < st.dl_nlev; ++L) {
    const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) {
      const int2 v = A.dl[e];
      int64_t ic;
      const int m = stamp_visit(g, v.x, v.y, &ic);
      if (m >= 0 && A.fp[m] > 0.0f)
        cached_visit(T, GlobalPick{A.pick.p, m}, A.fp[m], ic, false, A.pk, A.ps);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this level's stores before the next level's loads
    __syncthreads();
  }
}

// 4-pixel groups of box b, and the blocks that scan them (each thread
// takes two groups per pass).
int ring_blocks_for(const RingBox& b, int threads) {
  if (b.a_x1 < b.a_x0 || b.a_y1 < b.a_y0) return 0;
  const int64_t span = b.a_x1 - b.a_x0 + 1;
  const int64_t groups = ((span + 3) / 4 + 1) * (int64_t)(b.a_y1 - b.a_y0 + 1);
  const int64_t blocks = (groups + 2 * threads - 1) / (2 * threads);
  return (int)std::min<int64_t>(std::max<int64_t>(blocks, 1), 1024);
}

template <int kT>
void run_fused_t(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  hipStream_t s = b->ctx->stream;
  const PlanDev D = upload_plan(b, cv, st, n);
  FArgs A;
  A.st = D.st;
  A.canvas = cv->layer();
  A.snap = b->use_snap ? b->snap_planes() : cv->layer();
  A.pick = b->pick_planes();
  A.fp = b->fp;
  A.dirty = b->dirty;
  A.dl = D.dl;
  A.dl_off = D.dl_off;
  A.use_snap = b->use_snap ? 1 : 0;
  A.rows = cv->rows;
  A.cols = cv->cols;
  A.hr = (b->size_map - 1) / 2;
  A.size_map = b->size_map;
  const int side = 2 * A.hr + 1;
  A.tiles_x = (side + kT - 1) / kT;
  A.ntiles = A.tiles_x * A.tiles_x;
  A.pickup_rate = b->pickup_rate;
  A.dep_rate = b->dep_rate;
  A.cap = b->cap;
  for (int64_t k = 0; k < n; ++k) {
    if (st[k].flags & PNT_STAMP_DIP) {  // dip(): clean() the pickup map, new paint
      launch_fill(b->pick, 7 * b->mstride, 0.0f, s);
      for (int c = 0; c < 3; ++c) {
        b->paintK[c] = st[k].K[c];
        b->paintS[c] = st[k].S[c];
      }
    }
    for (int c = 0; c < 3; ++c) {
      A.pk[c] = b->paintK[c];
      A.ps[c] = b->paintS[c];
    }
    const PStamp& p = D.host[k];
    const int rb = A.use_snap ? ring_blocks_for(p.box, kT * kT) : 0;
    const int jk = (int)k;
    if (!(p.flags & (kStDupX | kStDupY))) {
      k_stamp_fused<kT><<<A.ntiles + rb, kT * kT, 0, s>>>(A, jk, A.ntiles, rb);
    } else {  // snapshot refresh first: a duplicated pixel is not refreshed by its visits
      if (rb) k_stamp_fused<kT><<<rb, kT * kT, 0, s>>>(A, jk, 0, rb);
      k_stamp_fused<kT><<<A.ntiles, kT * kT, 0, s>>>(A, jk, A.ntiles, 0);
      if (p.dl_nlev) k_stamp_levels<<<1, 256, 0, s>>>(A, jk);
    }
    PNT_HIP_CHECK(hipGetLastError());
    if (st[k].flags & PNT_STAMP_DRY_AFTER) {
      if (b->use_snap)
        launch_dry_dirty(cv->n, cv->layer(), cv->R0(), cv->h(), b->dirty, s);
      else
        launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), s);
    }
  }
  cv->epoch++;
  if (b->use_snap) {
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  }
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
    ring_scan<2>(A.f.canvas, A.f.snap, A.f.dirty, A.f.fp, g, st.box, pull,
                 (off - r.x) * kT * kT + t, r.y * kT * kT);
  } else if (role == 2) {
    const VisitTarget T{&A.f.canvas, &A.f.snap, A.f.dirty, A.f.use_snap, A.f.pickup_rate,
                        A.f.dep_rate, A.f.cap};
    float pkr[3] = {pk[0], pk[1], pk[2]}, psr[3] = {ps[0], ps[1], ps[2]};
    window_tile<kT>(T, A.f.pick.p, A.f.fp, g, st, dup, pull, off - r.x - r.y, A.f.tiles_x, sm,
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

void run_fused(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  if (n == 0) return;
  // window tile side: 16 (256 threads), or 32 from PAINTY_FUSED_TILE
  int kT = 16;
  if (const char* e = std::getenv("PAINTY_FUSED_TILE")) kT = std::atoi(e) == 32 ? 32 : 16;
  if (kT == 32)
    run_fused_t<32>(b, cv, st, n);
  else
    run_fused_t<16>(b, cv, st, n);
}

}  // namespace pnt
