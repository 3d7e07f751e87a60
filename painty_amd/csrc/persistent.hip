// A whole stamp list (one brush geometry, one canvas) in one cooperative
// launch: FootprintBrush::imprint (renderer/FootprintBrush.hxx:73-143)
// repeated in the reference's order, with the state that one imprint hands
// to the next kept where it is cheapest.
//
//   * Pickup map: each workgroup owns one kT x kT tile of map cells (one per
//     thread) for the whole launch and keeps their K, S, V and footprint in
//     LDS.  Per stamp it finds the window pixels that round into its tile
//     (inverse rotation of the tile gives a bounding box; each candidate is
//     mapped forward with the reference's exact FP64 formula), at most two
//     per cell, and runs each cell's visits in row-major order.  Visits of
//     different cells touch different canvas pixels, except on canvas
//     column/row 0 (see k_stamp_main in kernels.hip): those visits and their
//     partners run afterwards by dependency level, one barrier per level.
//   * Snapshot ring (updateSnapshot(canvas, center), :278-319): a copy of
//     canvas into snapshot over the ring around the window.  A per-pixel
//     dirty byte is set by every write to canvas or snapshot of that pixel
//     and cleared by a copy, so a clean pixel already holds snapshot ==
//     canvas and its copy is skipped.  Visits on the window border copy their
//     own pixel (they read the canvas value the refresh would have copied),
//     so ring and visits run in one phase.
//   * Order: [ring(j) + visits(j)] | barrier | ... ; a stamp with shared
//     column/row-0 pixels runs ring(j) | barrier | visits(j) | barrier per
//     level.  Every barrier wait is bounded (stampcore.hpp).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"

namespace pnt {

namespace {

struct PArgs {
  const PStamp* st;
  int n;
  int hr, size_map, rows, cols, tiles_x, ntiles;
  const float* fp;
  Plane7 pick;
  Layer7 canvas, src;
  uint8_t* dirty;
  int use_snap;
  double pickup_rate, dep_rate, cap;
  float paintK[3], paintS[3];
  const int2* dl;     // shared-pixel visits (col,row), grouped by level
  const int* dl_off;  // level offsets into dl
  unsigned* bar;      // kBarWords barrier words
  unsigned* host_err; // pinned host word raised on abort
};

template <int kT>
struct Pick {  // LDS pickup state of this block's map tile
  float K[3][kT * kT];
  float S[3][kT * kT];
  float V[kT * kT];
  float F[kT * kT];
};

template <int kT>
struct LdsPick {  // pickup map cell t of the block's tile, in LDS
  Pick<kT>* P;
  int t;
  __device__ float pick(int c) const {
    return c < 3 ? P->K[c][t] : (c < 6 ? P->S[c - 3][t] : P->V[t]);
  }
  __device__ void set_pick(int c, float v) const {
    if (c < 3)
      P->K[c][t] = v;
    else if (c < 6)
      P->S[c - 3][t] = v;
    else
      P->V[t] = v;
  }
};

template <int kT>
__global__ __launch_bounds__(kT* kT) void k_persistent(PArgs A) {
  constexpr int kCells = kT * kT;  // one map cell per thread
  __shared__ Pick<kT> P;
  __shared__ int cnt[kCells];
  __shared__ int slot[kCells][2];
  const int t = threadIdx.x;
  const unsigned nb = gridDim.x;
  const int tile = tile_of_block(blockIdx.x, A.ntiles);
  const int mx0 = (tile % A.tiles_x) * kT, my0 = (tile / A.tiles_x) * kT;
  {
    const int mx = mx0 + t % kT, my = my0 + t / kT;
    const bool own = mx < A.size_map && my < A.size_map;
    const int m = my * A.size_map + mx;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      P.K[c][t] = own ? A.pick.p[c][m] : 0.0f;
      P.S[c][t] = own ? A.pick.p[3 + c][m] : 0.0f;
    }
    P.V[t] = own ? A.pick.p[6][m] : 0.0f;
    P.F[t] = own ? A.fp[m] : 0.0f;
  }
  float pk[3], ps[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pk[c] = A.paintK[c];
    ps[c] = A.paintS[c];
  }
  StampGeom g;
  g.hr = A.hr;
  g.size_map = A.size_map;
  g.rows = A.rows;
  g.cols = A.cols;
  VisitTarget T{&A.canvas, &A.src, A.dirty, A.use_snap, A.pickup_rate, A.dep_rate, A.cap};
  const int gtid = blockIdx.x * kCells + t;
  const int gstride = nb * kCells;

  for (int j = 0; j < A.n; ++j) {
    const PStamp& st = A.st[j];
    g.cx = st.cx;
    g.cy = st.cy;
    g.c = st.c;
    g.s = st.s;
    const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
    const bool pull = A.use_snap && !dup;
    if (st.flags & kStDip) {  // dip(): clean() the pickup map, new paint
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pk[c] = st.K[c];
        ps[c] = st.S[c];
        P.K[c][t] = 0.0f;
        P.S[c][t] = 0.0f;
      }
      P.V[t] = 0.0f;
    }
    // ---- snapshot ring refresh: skip clean pixels (snapshot == canvas) and,
    // with pull, the pixels this stamp's visits refresh themselves
    if (A.use_snap) {
      const RingBox& b = st.box;
      if (b.a_x1 >= b.a_x0 && b.a_y1 >= b.a_y0) {
        const int w = b.a_x1 - b.a_x0 + 1;
        const int total = w * (b.a_y1 - b.a_y0 + 1);
        for (int e = gtid; e < total; e += gstride) {
          const int row = b.a_y0 + e / w, col = b.a_x0 + e % w;
          if (row > b.tl_y && row < b.br_y && col > b.tl_x && col < b.br_x) continue;
          const int64_t i = (int64_t)row * A.cols + col;
          if (!A.dirty[i]) continue;
          if (pull && ring_pixel_visited(g, b, A.fp, col, row)) continue;
#pragma unroll
          for (int c = 0; c < 6; ++c) A.src.ks[c][i] = A.canvas.ks[c][i];
          A.src.v[i] = A.canvas.v[i];
          A.dirty[i] = 0;
        }
      }
      if (dup && !group_sync(A.bar, nb, blockIdx.x, A.host_err)) return;
    }
    // ---- visits of this block's map tile
    cnt[t] = 0;
    __syncthreads();
    {
      // window offsets whose rotated position can round into the tile
      const double u0 = mx0 - A.hr - 0.5 - 1e-6, u1 = mx0 + kT - 1 - A.hr + 0.5 + 1e-6;
      const double v0 = my0 - A.hr - 0.5 - 1e-6, v1 = my0 + kT - 1 - A.hr + 0.5 + 1e-6;
      // inverse rotation of (u, v): col = u c + v s, row = -u s + v c
      double cmin = 1e300, cmax = -1e300, rmin = 1e300, rmax = -1e300;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double u = (k & 1) ? u1 : u0, v = (k & 2) ? v1 : v0;
        const double cc = u * g.c + v * g.s, rr = -u * g.s + v * g.c;
        cmin = fmin(cmin, cc);
        cmax = fmax(cmax, cc);
        rmin = fmin(rmin, rr);
        rmax = fmax(rmax, rr);
      }
      const int c0 = max((int)floor(cmin) - 1, -A.hr), c1 = min((int)ceil(cmax) + 1, A.hr);
      const int r0 = max((int)floor(rmin) - 1, -A.hr), r1 = min((int)ceil(rmax) + 1, A.hr);
      const int bw = c1 - c0 + 1;
      const int area = (c1 >= c0 && r1 >= r0) ? bw * (r1 - r0 + 1) : 0;
      for (int e = t; e < area; e += kCells) {
        const int col = c0 + e % bw, row = r0 + e / bw;
        int64_t ic;
        const int mm = stamp_visit(g, col, row, &ic);
        if (mm < 0) continue;
        const int lx = mm % A.size_map - mx0, ly = mm / A.size_map - my0;
        if (lx < 0 || lx >= kT || ly < 0 || ly >= kT) continue;
        const int cell = ly * kT + lx;
        const int k = atomicAdd(&cnt[cell], 1);
        if (k < 2) slot[cell][k] = ((row + A.hr) << 16) | (col + A.hr);
      }
    }
    __syncthreads();
    {
      const int nv = min(cnt[t], 2);
      const float fh = P.F[t];
      if (nv > 0 && fh > 0.0f) {
        int a = slot[t][0], b = nv > 1 ? slot[t][1] : 0;
        if (nv > 1 && b < a) {  // row-major order: (row << 16 | col) compares that way
          const int x = a;
          a = b;
          b = x;
        }
        const int colA = (a & 0xffff) - A.hr, rowA = (a >> 16) - A.hr;
        const int colB = (b & 0xffff) - A.hr, rowB = (b >> 16) - A.hr;
        if (!(dup && (on_dup_line(st, colA, rowA) || (nv > 1 && on_dup_line(st, colB, rowB))))) {
          const LdsPick<kT> pick{&P, t};
          int64_t ic;
          stamp_visit(g, colA, rowA, &ic);
          cached_visit(T, pick, fh, ic,
                       pull && on_window_border(st.box, (int)(ic % A.cols), (int)(ic / A.cols)),
                       pk, ps);
          if (nv > 1) {
            stamp_visit(g, colB, rowB, &ic);
            cached_visit(
                T, pick, fh, ic,
                pull && on_window_border(st.box, (int)(ic % A.cols), (int)(ic / A.cols)), pk, ps);
          }
        }
      }
    }
    // ---- visits sharing canvas column/row 0, level by level
    for (int L = 0; L < st.dl_nlev; ++L) {
      if (!group_sync(A.bar, nb, blockIdx.x, A.host_err)) return;
      const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
      for (int e = e0 + t; e < e1; e += kCells) {
        const int2 v = A.dl[e];
        int64_t ic;
        const int mm = stamp_visit(g, v.x, v.y, &ic);
        if (mm < 0) continue;
        const int lx = mm % A.size_map - mx0, ly = mm / A.size_map - my0;
        if (lx < 0 || lx >= kT || ly < 0 || ly >= kT) continue;
        const int cell = ly * kT + lx;
        if (P.F[cell] > 0.0f)
          cached_visit(T, LdsPick<kT>{&P, cell}, P.F[cell], ic, false, pk, ps);
      }
    }
    if (!group_sync(A.bar, nb, blockIdx.x, A.host_err)) return;
  }
  {
    const int mx = mx0 + t % kT, my = my0 + t / kT;
    if (mx < A.size_map && my < A.size_map) {
      const int m = my * A.size_map + mx;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        A.pick.p[c][m] = P.K[c][t];
        A.pick.p[3 + c][m] = P.S[c][t];
      }
      A.pick.p[6][m] = P.V[t];
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* d, int64_t n, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* d, int64_t n, unsigned v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = v;
}

// dryCanvas (Canvas.hxx:105-121) that also raises the snapshot dirty byte of
// every pixel whose wet layer it changes.
__global__ __launch_bounds__(256) void k_dry_dirty(int64_t n, Layer7 L, Rgb3 R0, float* h,
                                                   uint8_t* dirty) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float K[3], S[3], R[3];
  bool nz = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = L.ks[c][i];
    S[c] = L.ks[3 + c][i];
    R[c] = R0.p[c][i];
    nz |= (K[c] != 0.0f) | (S[c] != 0.0f);
  }
  const double v = L.v[i];
  nz |= v != 0.0;
  if (!nz) return;  // drying an empty pixel changes nothing
  h[i] = h_plus(h[i], v);
  km_pixel(K, S, R, (float)v);
#pragma unroll
  for (int c = 0; c < 3; ++c) R0.p[c][i] = R[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) L.ks[c][i] = 0.0f;
  L.v[i] = 0.0;
  dirty[i] = 1;
}

inline unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_u8<<<blocks256(n), 256, 0, s>>>(p, n, v);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_u32<<<blocks256(n), 256, 0, s>>>(p, n, v);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_dry_dirty(int64_t n, const Layer7& L, const Rgb3& R0, float* h, uint8_t* dirty,
                      hipStream_t s) {
  if (n <= 0) return;
  k_dry_dirty<<<blocks256(n), 256, 0, s>>>(n, L, R0, h, dirty);
  PNT_HIP_CHECK(hipGetLastError());
}

void plan_stamps(pnt_brush b, pnt_canvas cv, const pnt_stamp* st, int64_t n,
                 std::vector<PStamp>& h_st, std::vector<int2>& h_dl, std::vector<int>& h_off) {
  plan_stamps_geom(b->size_map, b->radius, cv, st, n, h_st, h_dl, h_off);
}

void plan_stamps_geom(int size_map, double radius, pnt_canvas cv, const pnt_stamp* st, int64_t n,
                      std::vector<PStamp>& h_st, std::vector<int2>& h_dl, std::vector<int>& h_off) {
  const size_t first = h_st.size();
  plan_stamps_rc(size_map, radius, cv->rows, cv->cols, st, n, h_st, h_dl, h_off);
  if (cv->clock != PNT_CLOCK_OFF)  // each imprint's time point: system_clock::now() per imprint
    for (size_t k = first; k < h_st.size(); ++k) h_st[k].now_us = canvas_now_us(cv);
}

void plan_stamps_rc(int size_map, double radius, int rows, int cols, const pnt_stamp* st,
                    int64_t n, std::vector<PStamp>& h_st, std::vector<int2>& h_dl,
                    std::vector<int>& h_off) {
  std::vector<int2> list;
  std::vector<int> off;
  for (int64_t k = 0; k < n; ++k) {
    const double cx = st[k].cx, cy = st[k].cy, theta = st[k].theta;
    PNT_REQUIRE(std::isfinite(cx) && std::isfinite(cy) && std::isfinite(theta) &&
                    std::fabs(cx) < 1e9 && std::fabs(cy) < 1e9,
                "stamp centre/angle must be finite and |centre| < 1e9");
    StampArgs a{};
    a.g.cx = cx;
    a.g.cy = cy;
    a.g.c = std::cos(-theta);
    a.g.s = std::sin(-theta);
    a.g.hr = (size_map - 1) / 2;
    a.g.size_map = size_map;
    a.g.rows = rows;
    a.g.cols = cols;
    a.dupx = stamp_dup_axis(cx, a.g.hr) ? 1 : 0;
    a.dupy = stamp_dup_axis(cy, a.g.hr) ? 1 : 0;
    PStamp p{};
    p.cx = a.g.cx;
    p.cy = a.g.cy;
    p.c = a.g.c;
    p.s = a.g.s;
    for (int c = 0; c < 3; ++c) {
      p.K[c] = st[k].K[c];
      p.S[c] = st[k].S[c];
    }
    p.flags = (st[k].flags & PNT_STAMP_DIP) ? kStDip : 0u;
    if (a.dupx) p.flags |= kStDupX;
    if (a.dupy) p.flags |= kStDupY;
    p.box = ring_box(a.g, radius);
    if (a.dupx || a.dupy) {
      plan_shared_visits(a, list, off);
      if (!list.empty()) {
        const int base = (int)h_dl.size();
        p.dl_lev = (int)h_off.size();
        p.dl_nlev = (int)off.size() - 1;
        for (int o : off) h_off.push_back(base + o);
        h_dl.insert(h_dl.end(), list.begin(), list.end());
      }
    }
    h_st.push_back(p);
  }
}

void brush_prepare_snapshot(pnt_brush b, pnt_canvas cv, hipStream_t s) {
  if (!b->use_snap) return;
  if (b->dirty_n != cv->n || !b->dirty) {
    // padded: the ring scan reads the bytes in aligned groups
    const int64_t need = round_up(cv->n, 16) + 16;
    if (need > b->dirty_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->dirty);
      b->dirty = nullptr;
      PNT_HIP_CHECK(hipMalloc((void**)&b->dirty, (size_t)need));
      b->dirty_cap = need;
    }
    b->dirty_n = cv->n;
    b->dirty_canvas = nullptr;
  }
  if (b->snap_rows != cv->rows || b->snap_cols != cv->cols)
    brush_full_snapshot(b, cv);  // snapshot == canvas: dirty bytes all 0
  else if (b->dirty_canvas != cv || b->dirty_epoch != cv->epoch)
    launch_fill_u8(b->dirty, cv->n, 1, s);  // unknown state: copy each ring pixel once
}

// Host-side state of persistent runs (lives with the brush).
struct PersistentState {
  PStamp* d_st = nullptr;
  int64_t st_cap = 0;
  float* d_fp = nullptr;  // footprints of a stroke list's geometries
  int64_t fp_cap = 0;
  int2* d_dl = nullptr;
  int64_t dl_cap = 0;
  int* d_off = nullptr;
  int64_t off_cap = 0;
  unsigned* d_bar = nullptr;
  int max_blocks16 = 0, max_blocks32 = 0;
  // pinned staging of the plan; `uploaded` completes once the device copies
  // have read it, after which it may be rewritten
  char* stage = nullptr;
  int64_t stage_cap = 0;
  hipEvent_t uploaded = nullptr;
  std::vector<PStamp> h_st;
  std::vector<int2> h_dl;
  std::vector<int> h_off;
};

void persistent_release(PersistentState*& p) {
  if (!p) return;
  (void)hipFree(p->d_st);
  (void)hipFree(p->d_dl);
  (void)hipFree(p->d_off);
  (void)hipFree(p->d_fp);
  (void)hipFree(p->d_bar);
  if (p->stage) (void)hipHostFree(p->stage);
  if (p->uploaded) (void)hipEventDestroy(p->uploaded);
  delete p;
  p = nullptr;
}

template <class T>
static void ensure_dev(T*& p, int64_t& cap, int64_t need, hipStream_t s) {
  if (need <= cap) return;
  PNT_HIP_CHECK(hipStreamSynchronize(s));  // the old buffer may still be read
  if (p) PNT_HIP_CHECK(hipFree(p));
  p = nullptr;
  const int64_t c = std::max<int64_t>(need, 2 * cap);
  PNT_HIP_CHECK(hipMalloc((void**)&p, (size_t)c * sizeof(T)));
  cap = c;
}

static PersistentState& plan_state(pnt_brush_s* b) {
  if (!b->pers) {
    auto* P = new PersistentState;
    b->pers = P;
    PNT_HIP_CHECK(hipMalloc((void**)&P->d_bar, kBarWords * sizeof(unsigned)));
    PNT_HIP_CHECK(hipEventCreateWithFlags(&P->uploaded, hipEventDisableTiming));
  }
  return *b->pers;
}

// Plans stamps [0, n) on the host, readies the snapshot and dirty bytes,
// and uploads the plan through pinned staging.  The host copy of the plan
// stays in the brush's state until the next call.
PlanDev upload_plan(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  const PlanPart part{b->size_map, b->radius, st, n};
  return upload_plan_parts(b, cv, &part, 1, 0, nullptr, nullptr);
}

PlanDev upload_plan_parts(pnt_brush_s* b, pnt_canvas_s* cv, const PlanPart* parts, int np,
                          int64_t fp_floats, const std::function<void(float*)>& fill_fp,
                          const float** d_fp) {
  hipStream_t s = b->ctx->stream;
  PersistentState& P = plan_state(b);
  P.h_st.clear();
  P.h_dl.clear();
  P.h_off.clear();
  if (np == 1) {
    plan_stamps_geom(parts[0].size_map, parts[0].radius, cv, parts[0].st, parts[0].n, P.h_st,
                     P.h_dl, P.h_off);
  } else {  // parts planned on their own, then concatenated
    std::vector<PStamp> pst;
    std::vector<int2> pdl;
    std::vector<int> poff;
    for (int i = 0; i < np; ++i) {
      pst.clear();
      pdl.clear();
      poff.clear();
      plan_stamps_geom(parts[i].size_map, parts[i].radius, cv, parts[i].st, parts[i].n, pst, pdl,
                       poff);
      const int off0 = (int)P.h_off.size(), dl0 = (int)P.h_dl.size();
      for (PStamp p : pst) {
        if (p.dl_nlev > 0) p.dl_lev += off0;
        P.h_st.push_back(p);
      }
      for (int o : poff) P.h_off.push_back(o + dl0);
      P.h_dl.insert(P.h_dl.end(), pdl.begin(), pdl.end());
    }
  }
  brush_prepare_snapshot(b, cv, s);
  const size_t bst = P.h_st.size() * sizeof(PStamp);
  const size_t bdl = P.h_dl.size() * sizeof(int2);
  const size_t boff = P.h_off.size() * sizeof(int);
  const size_t bfp = (size_t)fp_floats * sizeof(float);
  const int64_t need = (int64_t)(bst + bdl + boff + bfp);
  PNT_HIP_CHECK(hipEventSynchronize(P.uploaded));  // previous plan has been read
  if (need > P.stage_cap) {
    if (P.stage) PNT_HIP_CHECK(hipHostFree(P.stage));
    P.stage = nullptr;
    const int64_t c = std::max<int64_t>(need, 2 * P.stage_cap);
    PNT_HIP_CHECK(hipHostMalloc((void**)&P.stage, (size_t)c));
    P.stage_cap = c;
  }
  std::memcpy(P.stage, P.h_st.data(), bst);
  if (bdl) std::memcpy(P.stage + bst, P.h_dl.data(), bdl);
  if (boff) std::memcpy(P.stage + bst + bdl, P.h_off.data(), boff);
  if (bfp) fill_fp(reinterpret_cast<float*>(P.stage + bst + bdl + boff));
  const int64_t n = (int64_t)P.h_st.size();
  ensure_dev(P.d_st, P.st_cap, std::max<int64_t>(n, 1), s);
  ensure_dev(P.d_dl, P.dl_cap, std::max<int64_t>((int64_t)P.h_dl.size(), 1), s);
  ensure_dev(P.d_off, P.off_cap, std::max<int64_t>((int64_t)P.h_off.size(), 1), s);
  PNT_HIP_CHECK(hipMemcpyAsync(P.d_st, P.stage, bst, hipMemcpyHostToDevice, s));
  if (bdl)
    PNT_HIP_CHECK(hipMemcpyAsync(P.d_dl, P.stage + bst, bdl, hipMemcpyHostToDevice, s));
  if (boff)
    PNT_HIP_CHECK(hipMemcpyAsync(P.d_off, P.stage + bst + bdl, boff, hipMemcpyHostToDevice, s));
  if (bfp) {
    ensure_dev(P.d_fp, P.fp_cap, fp_floats, s);
    PNT_HIP_CHECK(hipMemcpyAsync(P.d_fp, P.stage + bst + bdl + boff, bfp, hipMemcpyHostToDevice,
                                 s));
  }
  if (d_fp) *d_fp = P.d_fp;
  PNT_HIP_CHECK(hipEventRecord(P.uploaded, s));
  return PlanDev{P.d_st, P.d_dl, P.d_off, P.h_st.data(), P.h_off.data()};
}

void run_persistent(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  hipStream_t s = b->ctx->stream;
  if (*(volatile unsigned*)b->ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  if (n == 0) return;
  PersistentState& P = plan_state(b);
  if (!P.max_blocks16) {
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    PNT_HIP_CHECK(hipGetDevice(&dev));
    PNT_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) throw Error(PNT_ESTATE, "device has no cooperative launch");
    PNT_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_persistent<16>), 256, 0));
    P.max_blocks16 = ncu * per_cu;
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_persistent<32>), 1024, 0));
    P.max_blocks32 = ncu * per_cu;
  }
  // map tile per workgroup: 16x16 cells for small brushes, 32x32 (1024
  // threads) for large ones, where fewer workgroups meet at each barrier
  int kT = b->size_map >= 128 ? 32 : 16;
  if (const char* e = std::getenv("PAINTY_PERSIST_TILE")) kT = std::atoi(e) == 32 ? 32 : 16;
  const int tiles_x = (b->size_map + kT - 1) / kT;
  const int ntiles = tiles_x * tiles_x;
  if (ntiles > (kT == 32 ? P.max_blocks32 : P.max_blocks16))
    throw Error(PNT_ESTATE, "brush map of " + std::to_string(b->size_map) +
                                " cells per side needs more co-resident workgroups than the "
                                "device holds; use PNT_EXEC_PLAIN");
  upload_plan(b, cv, st, n);

  PArgs A;
  A.hr = (b->size_map - 1) / 2;
  A.size_map = b->size_map;
  A.rows = cv->rows;
  A.cols = cv->cols;
  A.tiles_x = tiles_x;
  A.ntiles = ntiles;
  A.fp = b->fp;
  A.pick = b->pick_planes();
  A.canvas = cv->layer();
  A.src = b->use_snap ? b->snap_planes() : cv->layer();
  A.dirty = b->dirty;
  A.use_snap = b->use_snap ? 1 : 0;
  A.pickup_rate = b->pickup_rate;
  A.dep_rate = b->dep_rate;
  A.cap = b->cap;
  A.bar = P.d_bar;
  A.dl = P.d_dl;
  A.dl_off = P.d_off;
  A.host_err = nullptr;
  PNT_HIP_CHECK(hipHostGetDevicePointer((void**)&A.host_err, b->ctx->kernel_error, 0));
  // ---- one launch per run of stamps up to a dryCanvas
  int64_t i = 0;
  while (i < n) {
    int64_t e = i;
    while (e < n - 1 && !(st[e].flags & PNT_STAMP_DRY_AFTER)) ++e;
    launch_fill_u32(P.d_bar, kBarWords, 0u, s);
    A.st = P.d_st + i;
    A.n = (int)(e - i + 1);
    for (int c = 0; c < 3; ++c) {
      A.paintK[c] = b->paintK[c];
      A.paintS[c] = b->paintS[c];
    }
    void* kargs[] = {&A};
    const void* fn = kT == 32 ? reinterpret_cast<const void*>(k_persistent<32>)
                              : reinterpret_cast<const void*>(k_persistent<16>);
    PNT_HIP_CHECK(hipLaunchCooperativeKernel(fn, dim3(ntiles), dim3(kT * kT), kargs, 0, s));
    for (int64_t k = i; k <= e; ++k)  // brush paint after the run: its last dip
      if (st[k].flags & PNT_STAMP_DIP)
        for (int c = 0; c < 3; ++c) {
          b->paintK[c] = st[k].K[c];
          b->paintS[c] = st[k].S[c];
        }
    if (st[e].flags & PNT_STAMP_DRY_AFTER) {
      if (b->use_snap)
        launch_dry_dirty(cv->n, cv->layer(), cv->R0(), cv->h(), b->dirty, s);
      else
        launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), s);
    }
    i = e + 1;
  }
  cv->epoch++;
  if (b->use_snap) {
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  }
}

}  // namespace pnt
