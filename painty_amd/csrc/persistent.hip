// A whole stamp list (one brush geometry, one canvas) in one cooperative
// launch: FootprintBrush::imprint (renderer/FootprintBrush.hxx:73-143)
// repeated in the reference's order, with the state that one imprint hands
// to the next kept where it is cheapest.
//
//   * Pickup map: each workgroup owns one 16x16 tile of map cells for the
//     whole launch and keeps their K, S, V (and the footprint) in LDS.  Per
//     stamp it finds the window pixels that round into its tile (inverse
//     rotation of the tile gives a bounding box; each candidate is mapped
//     forward with the reference's exact FP64 formula), at most two per cell,
//     and runs each cell's visits in row-major order.  Visits of different
//     cells touch different canvas pixels, except on canvas column/row 0
//     (see k_stamp_main in kernels.hip): those visits and their partners run
//     afterwards by dependency level, one grid barrier per level.
//   * Snapshot ring (updateSnapshot(canvas, center), :278-319): a copy of
//     canvas into snapshot over the ring around the window.  A per-pixel
//     dirty byte is set by every write to canvas or snapshot of that pixel
//     and cleared by a copy, so a clean pixel already holds snapshot ==
//     canvas and its copy is skipped.
//   * Order: ring(j) | barrier | visits(j) [| barrier | shared-pixel level k]
//     | barrier | ring(j+1) ...   Every barrier wait is bounded; on expiry
//     the kernel raises an abort flag that every workgroup observes at its
//     next barrier and exits.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "visit.hpp"
#include "persistent.hpp"

namespace pnt {

namespace {

constexpr unsigned kSpinLimit = 1u << 22;
// barrier words, one 128-byte line each: 8 group arrival counters, the
// group-leader counter, the generation, the abort flag
constexpr int kBarLine = 32;
constexpr int kBarGroups = 8;
constexpr int kBarTop = kBarGroups * kBarLine;
constexpr int kBarGen = kBarTop + kBarLine;
constexpr int kBarAbort = kBarGen + kBarLine;
constexpr int kBarWords = kBarAbort + kBarLine;

enum : uint32_t { kDip = 1u, kDupX = 2u, kDupY = 4u };

struct PStamp {
  double cx, cy, c, s;
  float K[3], S[3];
  uint32_t flags;
  int32_t dl_lev;   // first entry of this stamp's level offsets (dup stamps)
  int32_t dl_nlev;  // number of shared-pixel levels (0 if none)
  RingBox box;
  int32_t pad[6];
};
static_assert(sizeof(PStamp) % 16 == 0, "PStamp alignment");

struct PArgs {
  const PStamp* st;
  int n;
  int hr, size_map, rows, cols, tiles_x, ntiles;
  const float* fp;
  Layer7 pick, canvas, src;
  uint8_t* dirty;
  int use_snap;
  float pickup_rate, dep_rate, cap;
  float paintK[3], paintS[3];
  const int2* dl;     // shared-pixel visits (col,row), grouped by level
  const int* dl_off;  // level offsets into dl
  unsigned* bar;      // kBarWords barrier words
  unsigned* host_err; // pinned host word raised on abort
};

__device__ __forceinline__ int tile_of_block(int b, int ntiles) {
  // blocks go round-robin to the 8 XCDs; give each XCD a contiguous run of
  // tiles so one XCD works on one band of the map (and of the window)
  const int x = b & 7, q = b >> 3;
  const int base = ntiles >> 3, rem = ntiles & 7;
  return x * base + (x < rem ? x : rem) + q;
}

// Grid barrier over the co-resident blocks of the cooperative launch.
//  * Each wave drains its own loads and stores (s_waitcnt 0): on this part a
//    workgroup barrier alone does not wait for global stores.  Then one wave
//    per block writes its XCD's L2 back (agent-scope release): the L2s of
//    different XCDs are not coherent with each other.
//  * Arrival is two-level (8 groups by block index, then the group leaders)
//    on counters that only grow, so no counter is ever reset.
//  * Waiters spin on relaxed loads (an acquire load would invalidate the L2
//    on every poll) and acquire once the generation moves.
//  * The wait is bounded: on expiry the abort flag is raised, every block
//    leaves at its next barrier, and the pinned host word reports it.
__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned nblocks, unsigned* host_err) {
  __shared__ int s_ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned gen = __hip_atomic_load(&bar[kBarGen], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    const unsigned b = blockIdx.x;
    const unsigned grp = b & (kBarGroups - 1);
    const unsigned ngroups = nblocks < (unsigned)kBarGroups ? nblocks : (unsigned)kBarGroups;
    const unsigned in_grp = nblocks / kBarGroups + (grp < nblocks % kBarGroups ? 1u : 0u);
    const unsigned a = __hip_atomic_fetch_add(&bar[grp * kBarLine], 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    if ((a + 1) % in_grp == 0) {
      const unsigned t = __hip_atomic_fetch_add(&bar[kBarTop], 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
      last = (t + 1) % ngroups == 0;
    }
    if (last) {
      __hip_atomic_store(&bar[kBarGen], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&bar[kBarGen], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             gen) {
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 63) == 0) {
          if (__hip_atomic_load(&bar[kBarAbort], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
            break;
          }
          if (spins > kSpinLimit) {
            __hip_atomic_store(&bar[kBarAbort], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(host_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            ok = 0;
            break;
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&bar[kBarAbort], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

template <int kT>
struct Pick {  // LDS pickup state of this block's map tile
  float K[3][kT * kT];
  float S[3][kT * kT];
  float V[kT * kT];
  float F[kT * kT];
};

template <int kT>
struct LdsPick {  // pickup-map side of a visit accessor
  Pick<kT>& P;
  int t;
  __device__ float pick(int c) const {
    return c < 3 ? P.K[c][t] : (c < 6 ? P.S[c - 3][t] : P.V[t]);
  }
  __device__ void set_pick(int c, float v) const {
    if (c < 3)
      P.K[c][t] = v;
    else if (c < 6)
      P.S[c - 3][t] = v;
    else
      P.V[t] = v;
  }
};

// One visit with every value it may need loaded up front (one memory round
// trip): the canvas pixel, and the pickup source -- the snapshot pixel, the
// canvas itself when the snapshot is off, or, on the window border where
// this stamp's snapshot refresh copies canvas -> snapshot just before the
// visits (FootprintBrush.hxx:307-318), the canvas value that refresh copies.
template <int kT>
struct CachedVisit : LdsPick<kT> {
  float sv[7], cv[7];
  bool alias;      // pickup source is the canvas itself (snapshot off)
  bool src_moved;  // the pickup changed the source volume
  __device__ float src(int c) const { return sv[c]; }
  __device__ void set_src_v(float v) {
    sv[6] = v;
    if (alias) cv[6] = v;
    src_moved = true;
  }
  __device__ float canvas(int c) const { return cv[c]; }
  __device__ void set_canvas(int c, float v) { cv[c] = v; }
};

template <int kT>
__device__ __forceinline__ void visit_cell(const PArgs& A, Pick<kT>& P, int t, int64_t ic, bool pull,
                                           const float pk[3], const float ps[3]) {
  const float fh = P.F[t];
  if (!(fh > 0.0f)) return;
  CachedVisit<kT> acc{{P, t}, {}, {}, !A.use_snap, false};
#pragma unroll
  for (int c = 0; c < 7; ++c) acc.cv[c] = A.canvas.p[c][ic];
  if (A.use_snap && !pull) {
#pragma unroll
    for (int c = 0; c < 7; ++c) acc.sv[c] = A.src.p[c][ic];
  } else {
#pragma unroll
    for (int c = 0; c < 7; ++c) acc.sv[c] = acc.cv[c];
  }
  paint_visit_math(acc, fh, A.pickup_rate, A.dep_rate, A.cap, pk, ps);
#pragma unroll
  for (int c = 0; c < 7; ++c) A.canvas.p[c][ic] = acc.cv[c];
  if (A.use_snap) {
    if (pull) {
#pragma unroll
      for (int c = 0; c < 7; ++c) A.src.p[c][ic] = acc.sv[c];
    } else if (acc.src_moved) {
      A.src.p[6][ic] = acc.sv[6];
    }
    A.dirty[ic] = 1;
  }
}

__device__ __forceinline__ bool on_dup_line(const PStamp& st, int col, int row) {
  return ((st.flags & kDupX) && (int)(col + st.cx) == 0) ||
         ((st.flags & kDupY) && (int)(row + st.cy) == 0);
}

// The window offset k in [-hr, hr] with (int)(k + c) == x, when exactly one
// exists (no duplicate on this axis); false otherwise.
__device__ __forceinline__ bool window_offset(int x, double c, int hr, int* k) {
#pragma clang fp contract(off)
  const int k0 = x - (int)floor(c);
  for (int d = -1; d <= 1; ++d) {
    const int kk = k0 + d;
    if (kk >= -hr && kk <= hr && (int)(kk + c) == x) {
      *k = kk;
      return true;
    }
  }
  return false;
}

// Is canvas pixel (x, y), on the ring of this stamp, also visited with a
// footprint height > 0 by it (and therefore refreshed by that visit)?
__device__ __forceinline__ bool ring_pixel_visited(const PArgs& A, const StampGeom& g,
                                                   const RingBox& b, int x, int y) {
  if (x < b.tl_x || x > b.br_x || y < b.tl_y || y > b.br_y) return false;
  int col, row;
  if (!window_offset(x, g.cx, A.hr, &col) || !window_offset(y, g.cy, A.hr, &row)) return false;
  int64_t ic;
  const int m = stamp_visit(g, col, row, &ic);
  return m >= 0 && A.fp[m] > 0.0f;
}

template <int kT>
__global__ __launch_bounds__(kT* kT) void k_persistent(PArgs A) {
  constexpr int kCells = kT * kT;
  constexpr int kThreads = kCells;  // one map cell per thread
  constexpr int kCpt = 1;
  __shared__ Pick<kT> P;
  __shared__ int cnt[kCells];
  __shared__ int slot[kCells][2];
  const int t = threadIdx.x;
  const unsigned nb = gridDim.x;
  const int tile = tile_of_block(blockIdx.x, A.ntiles);
  const int mx0 = (tile % A.tiles_x) * kT, my0 = (tile / A.tiles_x) * kT;
#pragma unroll
  for (int q = 0; q < kCpt; ++q) {
    const int cell = t + q * kThreads;
    const int mx = mx0 + cell % kT, my = my0 + cell / kT;
    const bool own = mx < A.size_map && my < A.size_map;
    const int m = my * A.size_map + mx;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      P.K[c][cell] = own ? A.pick.p[c][m] : 0.0f;
      P.S[c][cell] = own ? A.pick.p[3 + c][m] : 0.0f;
    }
    P.V[cell] = own ? A.pick.p[6][m] : 0.0f;
    P.F[cell] = own ? A.fp[m] : 0.0f;
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
  const int gtid = blockIdx.x * kThreads + t;
  const int gstride = nb * kThreads;

  for (int j = 0; j < A.n; ++j) {
    const PStamp& st = A.st[j];
    g.cx = st.cx;
    g.cy = st.cy;
    g.c = st.c;
    g.s = st.s;
    const bool dup = (st.flags & (kDupX | kDupY)) != 0;
    const bool pull = A.use_snap && !dup;
    if (st.flags & kDip) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pk[c] = st.K[c];
        ps[c] = st.S[c];
      }
#pragma unroll
      for (int q = 0; q < kCpt; ++q) {
        const int cell = t + q * kThreads;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          P.K[c][cell] = 0.0f;
          P.S[c][cell] = 0.0f;
        }
        P.V[cell] = 0.0f;
      }
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
          if (pull && ring_pixel_visited(A, g, b, col, row)) continue;
#pragma unroll
          for (int c = 0; c < 7; ++c) A.src.p[c][i] = A.canvas.p[c][i];
          A.dirty[i] = 0;
        }
      }
      if (dup && !grid_sync(A.bar, nb, A.host_err)) return;
    }
    // ---- visits of this block's map tile
#pragma unroll
    for (int q = 0; q < kCpt; ++q) cnt[t + q * kThreads] = 0;
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
      for (int e = t; e < area; e += kThreads) {
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
#pragma unroll
    for (int q = 0; q < kCpt; ++q) {
      const int cell = t + q * kThreads;
      const int nv = min(cnt[cell], 2);
      if (nv == 0) continue;
      int a = slot[cell][0], b = nv > 1 ? slot[cell][1] : 0;
      if (nv > 1 && b < a) {  // row-major order: (row << 16 | col) compares that way
        const int x = a;
        a = b;
        b = x;
      }
      const int colA = (a & 0xffff) - A.hr, rowA = (a >> 16) - A.hr;
      const int colB = (b & 0xffff) - A.hr, rowB = (b >> 16) - A.hr;
      if (dup && (on_dup_line(st, colA, rowA) || (nv > 1 && on_dup_line(st, colB, rowB))))
        continue;
      const RingBox& bx = st.box;
      int64_t ic;
      stamp_visit(g, colA, rowA, &ic);
      {
        const int x = (int)(ic % A.cols), y = (int)(ic / A.cols);
        const bool border = x == bx.tl_x || x == bx.br_x || y == bx.tl_y || y == bx.br_y;
        visit_cell<kT>(A, P, cell, ic, pull && border, pk, ps);
      }
      if (nv > 1) {
        stamp_visit(g, colB, rowB, &ic);
        const int x = (int)(ic % A.cols), y = (int)(ic / A.cols);
        const bool border = x == bx.tl_x || x == bx.br_x || y == bx.tl_y || y == bx.br_y;
        visit_cell<kT>(A, P, cell, ic, pull && border, pk, ps);
      }
    }
    // ---- visits sharing canvas column/row 0, level by level
    for (int L = 0; L < st.dl_nlev; ++L) {
      if (!grid_sync(A.bar, nb, A.host_err)) return;
      const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
      for (int e = e0 + t; e < e1; e += kThreads) {
        const int2 v = A.dl[e];
        int64_t ic;
        const int mm = stamp_visit(g, v.x, v.y, &ic);
        if (mm < 0) continue;
        const int lx = mm % A.size_map - mx0, ly = mm / A.size_map - my0;
        if (lx < 0 || lx >= kT || ly < 0 || ly >= kT) continue;
        visit_cell<kT>(A, P, ly * kT + lx, ic, false, pk, ps);
      }
    }
    if (!grid_sync(A.bar, nb, A.host_err)) return;
  }
#pragma unroll
  for (int q = 0; q < kCpt; ++q) {
    const int cell = t + q * kThreads;
    const int mx = mx0 + cell % kT, my = my0 + cell / kT;
    if (mx < A.size_map && my < A.size_map) {
      const int m = my * A.size_map + mx;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        A.pick.p[c][m] = P.K[c][cell];
        A.pick.p[3 + c][m] = P.S[c][cell];
      }
      A.pick.p[6][m] = P.V[cell];
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* d, int64_t n, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = v;
}

__global__ __launch_bounds__(512) void k_reset_bar(unsigned* bar) {
  if (threadIdx.x < kBarWords) bar[threadIdx.x] = 0u;
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
    K[c] = L.p[c][i];
    S[c] = L.p[3 + c][i];
    R[c] = R0.p[c][i];
    nz |= (K[c] != 0.0f) | (S[c] != 0.0f);
  }
  const float v = L.p[6][i];
  nz |= v != 0.0f;
  if (!nz) return;  // drying an empty pixel changes nothing
  h[i] += v;
  km_pixel(K, S, R, v);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R0.p[c][i] = R[c];
    L.p[c][i] = 0.0f;
    L.p[3 + c][i] = 0.0f;
  }
  L.p[6][i] = 0.0f;
  dirty[i] = 1;
}

inline unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_u8<<<blocks256(n), 256, 0, s>>>(p, n, v);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_dry_dirty(int64_t n, const Layer7& L, const Rgb3& R0, float* h, uint8_t* dirty,
                      hipStream_t s) {
  if (n <= 0) return;
  k_dry_dirty<<<blocks256(n), 256, 0, s>>>(n, L, R0, h, dirty);
  PNT_HIP_CHECK(hipGetLastError());
}

// Host-side state of persistent runs (lives with the brush).
struct PersistentState {
  PStamp* d_st = nullptr;
  int64_t st_cap = 0;
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

static void ensure_dirty(pnt_brush_s* b, pnt_canvas_s* cv, hipStream_t s) {
  if (b->dirty_n == cv->n && b->dirty) return;
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  dev_free(b->dirty);
  b->dirty = nullptr;
  PNT_HIP_CHECK(hipMalloc((void**)&b->dirty, (size_t)std::max<int64_t>(cv->n, 1)));
  b->dirty_n = cv->n;
  b->dirty_canvas = nullptr;
}

void run_persistent(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n) {
  hipStream_t s = b->ctx->stream;
  if (*(volatile unsigned*)b->ctx->kernel_error)
    throw Error(PNT_EKERNEL, "an earlier persistent stamp kernel aborted");
  if (n == 0) return;
  if (!b->pers) {
    auto* P = new PersistentState;
    b->pers = P;
    PNT_HIP_CHECK(hipMalloc((void**)&P->d_bar, kBarWords * sizeof(unsigned)));
    PNT_HIP_CHECK(hipEventCreateWithFlags(&P->uploaded, hipEventDisableTiming));
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    PNT_HIP_CHECK(hipGetDevice(&dev));
    PNT_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) throw Error(PNT_ESTATE, "device has no cooperative launch");
    PNT_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_persistent<16>), 256, 0));
    P->max_blocks16 = ncu * per_cu;
    PNT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_persistent<32>), 1024, 0));
    P->max_blocks32 = ncu * per_cu;
  }
  PersistentState& P = *b->pers;
  // map tile per workgroup: 16x16 cells (one per thread) for small brushes,
  // 32x32 (four per thread) for large ones, where fewer workgroups meet at
  // each barrier
  int kT = b->size_map >= 128 ? 32 : 16;
  if (const char* e = std::getenv("PAINTY_PERSIST_TILE")) kT = std::atoi(e) == 32 ? 32 : 16;
  const int tiles_x = (b->size_map + kT - 1) / kT;
  const int ntiles = tiles_x * tiles_x;
  if (ntiles > (kT == 32 ? P.max_blocks32 : P.max_blocks16))
    throw Error(PNT_ESTATE, "brush map of " + std::to_string(b->size_map) +
                                " cells per side needs more co-resident workgroups than the "
                                "device holds; use PNT_EXEC_PLAIN");
  // ---- plan every stamp on the host (FP64 geometry, shared-pixel levels)
  P.h_st.resize((size_t)n);
  P.h_dl.clear();
  P.h_off.clear();
  std::vector<int2> list;
  std::vector<int> off;
  for (int64_t k = 0; k < n; ++k) {
    const StampArgs a = stamp_args(b, cv, st[k].cx, st[k].cy, st[k].theta);
    PStamp& p = P.h_st[(size_t)k];
    p = PStamp{};
    p.cx = a.g.cx;
    p.cy = a.g.cy;
    p.c = a.g.c;
    p.s = a.g.s;
    for (int c = 0; c < 3; ++c) {
      p.K[c] = st[k].K[c];
      p.S[c] = st[k].S[c];
    }
    p.flags = (st[k].flags & PNT_STAMP_DIP) ? kDip : 0u;
    if (a.dupx) p.flags |= kDupX;
    if (a.dupy) p.flags |= kDupY;
    p.box = ring_box(a.g, b->radius);
    if (a.dupx || a.dupy) {
      plan_shared_visits(a, list, off);
      if (!list.empty()) {
        const int base = (int)P.h_dl.size();
        p.dl_lev = (int)P.h_off.size();
        p.dl_nlev = (int)off.size() - 1;
        for (int o : off) P.h_off.push_back(base + o);
        P.h_dl.insert(P.h_dl.end(), list.begin(), list.end());
      }
    }
  }
  // ---- snapshot and its dirty bytes
  if (b->use_snap) {
    ensure_dirty(b, cv, s);
    if (b->snap_rows != cv->rows || b->snap_cols != cv->cols)
      brush_full_snapshot(b, cv);  // snapshot == canvas: dirty bytes all 0
    else if (b->dirty_canvas != cv || b->dirty_epoch != cv->epoch)
      launch_fill_u8(b->dirty, cv->n, 1, s);  // unknown state: copy each ring pixel once
  }
  // ---- upload the plan through pinned staging
  const size_t bst = P.h_st.size() * sizeof(PStamp);
  const size_t bdl = P.h_dl.size() * sizeof(int2);
  const size_t boff = P.h_off.size() * sizeof(int);
  const int64_t need = (int64_t)(bst + bdl + boff);
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
  ensure_dev(P.d_st, P.st_cap, n, s);
  ensure_dev(P.d_dl, P.dl_cap, std::max<int64_t>((int64_t)P.h_dl.size(), 1), s);
  ensure_dev(P.d_off, P.off_cap, std::max<int64_t>((int64_t)P.h_off.size(), 1), s);
  PNT_HIP_CHECK(hipMemcpyAsync(P.d_st, P.stage, bst, hipMemcpyHostToDevice, s));
  if (bdl)
    PNT_HIP_CHECK(hipMemcpyAsync(P.d_dl, P.stage + bst, bdl, hipMemcpyHostToDevice, s));
  if (boff)
    PNT_HIP_CHECK(hipMemcpyAsync(P.d_off, P.stage + bst + bdl, boff, hipMemcpyHostToDevice, s));
  PNT_HIP_CHECK(hipEventRecord(P.uploaded, s));

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
    k_reset_bar<<<1, 512, 0, s>>>(P.d_bar);
    PNT_HIP_CHECK(hipGetLastError());
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
