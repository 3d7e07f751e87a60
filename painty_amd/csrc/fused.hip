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

}  // namespace

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
