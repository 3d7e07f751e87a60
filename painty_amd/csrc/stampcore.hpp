// Device pieces shared by the launch-per-list stamp kernels (persistent.hip,
// batch.hip): the planned stamp descriptor, the bounded group barrier, the
// one-round-trip visit, and the snapshot-ring helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "dry.hpp"
#include "internal.hpp"
#include "visit.hpp"

namespace pnt {

// Bounding box (map cells, inclusive) of the footprint's nonzero values: a
// visit whose cell lies outside it has fh = 0 and changes nothing, so its
// loads are skipped (an ALU test, no memory round trip).
struct FpBox {
  int x0, y0, x1, y1;
};


enum : uint32_t { kStDip = 1u, kStDupX = 2u, kStDupY = 4u };

// One imprint, planned on the host (FP64 geometry as stamp_args computes it).
struct PStamp {
  double cx, cy, c, s;
  float K[3], S[3];  // paint of a dip before this stamp
  uint32_t flags;    // kStDip | kStDupX | kStDupY
  int32_t dl_lev;    // first entry of this stamp's level offsets (dup stamps)
  int32_t dl_nlev;   // number of shared-pixel levels (0 if none)
  RingBox box;
  int32_t pad0;
  int64_t now_us;  // the imprint's time point when the canvas has a drying clock
  int32_t pad[4];
};
static_assert(sizeof(PStamp) % 16 == 0, "PStamp alignment");

// ------------------------------------------------------------ barriers ---
// Barrier words of one group, one 128-byte line each: 8 arrival counters,
// the leader counter, the generation, the abort flag.
constexpr int kBarLine = 32;
constexpr int kBarGroups = 8;
constexpr int kBarTop = kBarGroups * kBarLine;
constexpr int kBarGen = kBarTop + kBarLine;
constexpr int kBarAbort = kBarGen + kBarLine;
constexpr int kBarWords = kBarAbort + kBarLine;
constexpr unsigned kSpinLimit = 1u << 22;

// Barrier over the `nblocks` co-resident blocks of a group (`rank` in
// [0, nblocks)).
//  * Each wave drains its own loads and stores (s_waitcnt 0): a workgroup
//    barrier alone does not wait for global stores.  Then one wave per block
//    writes its XCD's L2 back (agent-scope release): the L2s of different
//    XCDs are not coherent with each other.
//  * Arrival is two-level (8 sub-groups by rank, then their leaders) on
//    counters that only grow, so no counter is ever reset.
//  * Waiters spin on relaxed loads (an acquire load would invalidate the L2
//    on every poll) and acquire once the generation moves.
//  * The wait is bounded: on expiry the abort flag is raised, every block
//    leaves at its next barrier, and the pinned host word reports it.
// Returns false when the run is aborted.
__device__ __forceinline__ bool group_sync(unsigned* bar, unsigned nblocks, unsigned rank,
                                           unsigned* host_err) {
  __shared__ int s_ok;
  if (nblocks == 1) {  // one block: a workgroup barrier orders everything
    __syncthreads();
    return true;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned gen =
        __hip_atomic_load(&bar[kBarGen], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned grp = rank & (kBarGroups - 1);
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

// --------------------------------------------------------------- visits ---
// One visit with every value it may need loaded up front (one memory round
// trip): the canvas pixel, and the pickup source -- the snapshot pixel, the
// canvas itself when the snapshot is off, or, on the window border where
// this stamp's snapshot refresh copies canvas -> snapshot just before the
// visits (FootprintBrush.hxx:307-318), the canvas value that refresh copies.
// PickAcc supplies pick(c) / set_pick(c, v) for the visit's map cell.
template <class PickAcc>
struct CachedVisit : PickAcc {
  float sv[6], cv[6];  // K, S of the pickup source and of the canvas pixel
  double svv, cvv;     // their volumes
  bool alias;          // pickup source is the canvas itself (snapshot off)
  bool src_moved;      // the pickup changed the source volume
  __device__ float src(int c) const { return sv[c]; }
  __device__ double src_v() const { return svv; }
  __device__ void set_src_v(double v) {
    svv = v;
    if (alias) cvv = v;
    src_moved = true;
  }
  __device__ float canvas(int c) const { return cv[c]; }
  __device__ double canvas_v() const { return cvv; }
  __device__ void set_canvas(int c, float v) { cv[c] = v; }
  __device__ void set_canvas_v(double v) { cvv = v; }
  // the pickup source is the canvas pixel itself
  __device__ void src_from_canvas() {
#pragma unroll
    for (int c = 0; c < 6; ++c) sv[c] = cv[c];
    svv = cvv;
  }
};

// One pixel of a layer: K, S (FP32) and V (FP64).
__device__ __forceinline__ void ld_px(const Layer7& L, int64_t i, float ks[6], double& v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) ks[c] = L.ks[c][i];
  v = L.v[i];
}
__device__ __forceinline__ void st_px(const Layer7& L, int64_t i, const float ks[6], double v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) L.ks[c][i] = ks[c];
  L.v[i] = v;
}

struct GlobalPick {  // pickup map cell in global memory
  float* const* p;
  int im;
  __device__ float pick(int c) const { return p[c][im]; }
  __device__ void set_pick(int c, float v) const { p[c][im] = v; }
};

// Where one visit reads and writes.
struct VisitTarget {
  const Layer7* canvas;
  const Layer7* src;  // snapshot planes (unused when use_snap == 0)
  uint8_t* dirty;
  int use_snap;
  double pickup_rate, dep_rate, cap;
};

template <class PickAcc>
__device__ __forceinline__ void cached_visit(const VisitTarget& T, const PickAcc& pick, float fh,
                                             int64_t ic, bool pull, const float pk[3],
                                             const float ps[3]) {
  CachedVisit<PickAcc> acc{pick, {}, {}, 0.0, 0.0, !T.use_snap, false};
  ld_px(*T.canvas, ic, acc.cv, acc.cvv);
  if (T.use_snap && !pull)
    ld_px(*T.src, ic, acc.sv, acc.svv);
  else
    acc.src_from_canvas();
  paint_visit_math(acc, fh, T.pickup_rate, T.dep_rate, T.cap, pk, ps);
  st_px(*T.canvas, ic, acc.cv, acc.cvv);
  if (T.use_snap) {
    if (pull)
      st_px(*T.src, ic, acc.sv, acc.svv);
    else if (acc.src_moved)
      T.src->v[ic] = acc.svv;
    T.dirty[ic] = 1;
  }
}

// Tile of block b when blocks [0, ntiles) take one tile each: blocks go
// round-robin to the 8 XCDs, so each XCD gets a contiguous run of tiles and
// works on one band of the map (and of the window) in its own L2.
__device__ __forceinline__ int tile_of_block(int b, int ntiles) {
  const int x = b & 7, q = b >> 3;
  const int base = ntiles >> 3, rem = ntiles & 7;
  return x * base + (x < rem ? x : rem) + q;
}

// ----------------------------------------------------------- ring helpers -
// Does stamp p straddle canvas column/row 0 with shared-pixel visits (which
// then run by level on a block of their own)?
__host__ __device__ inline bool dup_levels(const PStamp& p) {
  return (p.flags & (kStDupX | kStDupY)) && p.dl_nlev > 0;
}

__device__ __forceinline__ bool on_dup_line(const PStamp& st, int col, int row) {
  return ((st.flags & kStDupX) && (int)(col + st.cx) == 0) ||
         ((st.flags & kStDupY) && (int)(row + st.cy) == 0);
}

// Every window offset k in [-hr, hr] with (int)(k + c) == x, ascending: one
// for most pixels, two on a duplicated canvas column/row 0, none off the
// window.  Returns how many.
__device__ __forceinline__ int window_offsets(int x, double c, int hr, int k[2]) {
#pragma clang fp contract(off)
  const int k0 = x - (int)floor(c);
  int n = 0;
  for (int d = -1; d <= 1; ++d) {
    const int kk = k0 + d;
    if (kk >= -hr && kk <= hr && (int)(kk + c) == x && n < 2) k[n++] = kk;
  }
  return n;
}

// Is canvas pixel (x, y), on the ring of this stamp, also visited with a
// footprint height > 0 by it (and therefore refreshed by that visit)?  With
// a drying clock every visited pixel counts: its window tile refreshes it
// before checkDry changes the canvas.  A pixel of a duplicated column/row 0
// counts when any of its visits does (the first of them refreshes it).
__device__ __forceinline__ bool ring_pixel_visited(const StampGeom& g, const RingBox& b,
                                                   const float* fp, int x, int y,
                                                   bool any_visit = false) {
  if (x < b.tl_x || x > b.br_x || y < b.tl_y || y > b.br_y) return false;
  int cs[2], rs[2];
  const int nc = window_offsets(x, g.cx, g.hr, cs), nr = window_offsets(y, g.cy, g.hr, rs);
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nc; ++c) {
      int64_t ic;
      const int m = stamp_visit(g, cs[c], rs[r], &ic);
      if (m >= 0 && (any_visit || fp[m] > 0.0f)) return true;
    }
  return false;
}

__device__ __forceinline__ bool on_window_border(const RingBox& b, int x, int y) {
  return x == b.tl_x || x == b.br_x || y == b.tl_y || y == b.br_y;
}

// Snapshot ring refresh of one stamp (updateSnapshot(canvas, center),
// FootprintBrush.hxx:307-318) over `nthreads` threads.  The allowed box is
// cut into 4-pixel groups at 4-aligned linear indices; a thread reads a
// group's four dirty bytes (a set byte: snapshot differs from canvas), and
// for the set pixels outside the window interior copies canvas -> snapshot
// and clears the byte.  With `pull`, window-border pixels that a visit of
// this stamp refreshes itself are left to that visit.  Each thread takes
// kG groups per pass (so that a stamp's ring needs few blocks) with all
// their dirty-word loads in flight together (the dirty
// strip behind a moving window is contiguous, so a pixel-serial copy would
// wait on one round trip per pixel).  Plane and dirty buffers are padded
// past the last pixel to a multiple of 16, so a group never leaves them.
template <int kG>
__device__ __forceinline__ void ring_scan(const Layer7& canvas, const Layer7& snap, uint8_t* dirty,
                                          const float* fp, const StampGeom& g, const RingBox& b,
                                          bool pull, int tid, int nthreads,
                                          bool clocked = false) {
  if (b.a_x1 < b.a_x0 || b.a_y1 < b.a_y0) return;
  const int rows = b.a_y1 - b.a_y0 + 1;
  const int span = b.a_x1 - b.a_x0 + 1;
  const int per_row = (span + 3) / 4 + 1;  // +1: misaligned start
  const int total = per_row * rows;
  for (int e0 = tid; e0 < total; e0 += kG * nthreads) {
    unsigned mask[kG];
    int64_t base[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      mask[u] = 0u;
      base[u] = 0;
      const int e = e0 + u * nthreads;
      if (e >= total) continue;
      // group e: row a_y0 + e / per_row, pixels [base, base + 4) with
      // base = (lo & ~3) + 4 * (e % per_row), lo = row*cols + a_x0
      const int row = b.a_y0 + e / per_row;
      const int64_t rowi = (int64_t)row * g.cols;
      const int64_t lo = rowi + b.a_x0;
      const int64_t bs = (lo & ~(int64_t)3) + 4 * (int64_t)(e % per_row);
      if (bs >= lo + span) continue;
      const bool interior_row = row > b.tl_y && row < b.br_y;
      const int x0 = (int)(bs - rowi);
      if (interior_row && x0 > b.tl_x && x0 + 3 < b.br_x) continue;  // inside the window interior
      const unsigned d = *reinterpret_cast<const unsigned*>(dirty + bs);
      if (!d) continue;
      unsigned m = 0u, border = 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!((d >> (8 * q)) & 0xffu)) continue;
        const int64_t i = bs + q;
        if (i < lo || i >= lo + span) continue;
        const int col = x0 + q;
        if (interior_row && col > b.tl_x && col < b.br_x) continue;
        m |= 1u << q;
        if (pull && on_window_border(b, col, row)) border |= 1u << q;
      }
      // window-border pixels a visit refreshes itself (few: one loop body)
#pragma unroll 1
      for (int q = 0; q < 4; ++q)
        if (((border >> q) & 1u) && ring_pixel_visited(g, b, fp, x0 + q, row, clocked))
          m &= ~(1u << q);
      mask[u] = m;
      base[u] = bs;
    }
    // the copies (few: the strip the last stamps left dirty), one group at a
    // time: two groups' planes in registers at once set the register peak
    // of every kernel that runs ring scans
    constexpr int kC = 1;
#pragma unroll
    for (int u0 = 0; u0 < kG; u0 += kC) {
      float4 v[kC][6];
      double2 w[kC][2];  // the four volumes
#pragma unroll
      for (int h = 0; h < kC; ++h) {
        const int u = u0 + h;
        if (!mask[u]) continue;
#pragma unroll
        for (int c = 0; c < 6; ++c) v[h][c] = reinterpret_cast<const float4*>(canvas.ks[c])[base[u] >> 2];
        w[h][0] = reinterpret_cast<const double2*>(canvas.v)[base[u] >> 1];
        w[h][1] = reinterpret_cast<const double2*>(canvas.v)[(base[u] >> 1) + 1];
      }
#pragma unroll
      for (int h = 0; h < kC; ++h) {
        const int u = u0 + h;
        if (!mask[u]) continue;
        if (mask[u] == 0xfu) {
#pragma unroll
          for (int c = 0; c < 6; ++c) reinterpret_cast<float4*>(snap.ks[c])[base[u] >> 2] = v[h][c];
          reinterpret_cast<double2*>(snap.v)[base[u] >> 1] = w[h][0];
          reinterpret_cast<double2*>(snap.v)[(base[u] >> 1) + 1] = w[h][1];
          *reinterpret_cast<unsigned*>(dirty + base[u]) = 0u;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (!((mask[u] >> q) & 1u)) continue;
#pragma unroll
            for (int c = 0; c < 6; ++c)
              snap.ks[c][base[u] + q] = reinterpret_cast<const float*>(&v[h][c])[q];
            snap.v[base[u] + q] = reinterpret_cast<const double*>(&w[h][0])[q];
            dirty[base[u] + q] = 0;
          }
        }
      }
    }
  }
}

// Pickup map cell held in registers between its load and its store.
struct RegPick {
  float v[7];
  __device__ float pick(int c) const { return v[c]; }
  __device__ void set_pick(int c, float x) { v[c] = x; }
};

__device__ __forceinline__ int64_t canvas_index(const StampGeom& g, int col, int row) {
#pragma clang fp contract(off)
  return (int64_t)(int)(row + g.cy) * g.cols + (int)(col + g.cx);
}

// Loads everything one visit reads: canvas and pickup-source pixel.
__device__ __forceinline__ void load_pixel(const VisitTarget& T, int64_t ic, bool pull,
                                           CachedVisit<RegPick>& acc) {
  ld_px(*T.canvas, ic, acc.cv, acc.cvv);
  if (T.use_snap && !pull)
    ld_px(*T.src, ic, acc.sv, acc.svv);
  else
    acc.src_from_canvas();
  acc.src_moved = false;
}

// Runs one visit on loaded values and stores the canvas / snapshot side.
__device__ __forceinline__ void run_pixel(const VisitTarget& T, int64_t ic, bool pull, float fh,
                                          const float pk[3], const float ps[3],
                                          CachedVisit<RegPick>& acc) {
  paint_visit_math(acc, fh, T.pickup_rate, T.dep_rate, T.cap, pk, ps);
  st_px(*T.canvas, ic, acc.cv, acc.cvv);
  if (T.use_snap) {
    if (pull)
      st_px(*T.src, ic, acc.sv, acc.svv);
    else if (acc.src_moved)
      T.src->v[ic] = acc.svv;
    T.dirty[ic] = 1;
  }
}

// Canvas::checkDry (Canvas.hxx:123-155) of a visited pixel whose canvas
// values are held in registers (ks[6], v); R0, h and the time map are in
// memory.  The same arithmetic as check_dry (dry.hpp).
__device__ __forceinline__ void check_dry_regs(const DryClock& c, float ks[6], double& v,
                                               int64_t ic) {
  if (c.dry_ms > 0 && v > 0.001) {
    const int64_t elapsed_ms = (c.now_us - c.tmap[ic]) / 1000;  // duration_cast truncation
    if (elapsed_ms >= c.dry_ms) {
      float K[3], S[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        K[ch] = ks[ch];
        S[ch] = ks[3 + ch];
      }
      dry_pixel_into(c, K, S, v, ic);
#pragma unroll
      for (int ch = 0; ch < 6; ++ch) ks[ch] = 0.0f;
      v = 0.0;
    }
  }
  c.tmap[ic] = c.now_us;
}

// checkDry of a visited pixel with footprint height <= 0 (no pickup or
// deposit follows).  On the window border of a stamp that refreshes its
// border itself (`pull`), the snapshot refresh comes first, as
// updateSnapshot runs before the visits (FootprintBrush.hxx:80-84).  A
// pixel that dries no longer equals its snapshot.
__device__ __forceinline__ void check_only(const VisitTarget& T, const DryClock& c, int64_t ic,
                                           bool pv) {
  if (pv) {
    float ks[6];
    double v;
    ld_px(*T.canvas, ic, ks, v);
    st_px(*T.src, ic, ks, v);
    T.dirty[ic] = 0;
  }
  const double v0 = T.canvas->v[ic];
  check_dry(c, *T.canvas, ic);
  if (T.use_snap && v0 > 0.001 && T.canvas->v[ic] != v0) T.dirty[ic] = 1;
}

// The visits of one kT x kT tile of a stamp's window, by one workgroup of
// kT*kT threads (all of which must call it).  The map cells of the tile and
// a one-pixel halo go to LDS (`sm`); a visit whose left or upper neighbour
// shares its cell is left to that neighbour, which runs both in row-major
// order (a cell has at most two visitors, and they are 4-neighbours).
// Visits on a duplicated canvas column/row 0 (and their partners) are left
// out; they run afterwards by dependency level.  Each visit's pickup cell,
// canvas pixel and source pixel are loaded in one batch.
//
// With a drying clock (clk.tmap != nullptr) every visited pixel -- in the canvas
// and in the map, whatever its footprint height -- gets Canvas::checkDry
// before its pickup and deposit (FootprintBrush.hxx:115), by the thread that
// runs its visit: map cells are then taken without the footprint-box cut,
// so the pairs are those of the plain kernel.

template <int kT>
__device__ __forceinline__ void window_tile(const VisitTarget& T, float* const* pick,
                                            const float* fp, const FpBox& fb, const StampGeom& g,
                                            const PStamp& st, bool dup, bool pull, int tile,
                                            int tiles_x, int (*sm)[kT + 2], const float pk[3],
                                            const float ps[3], const DryClock& clk = DryClock{}) {
  const bool clocked = clk.tmap != nullptr;
  const int t = threadIdx.x;
  const int lx = t % kT, ly = t / kT;
  const int col0 = -g.hr + (tile % tiles_x) * kT;
  const int row0 = -g.hr + (tile / tiles_x) * kT;
  // map cells of the tile and its halo; without a clock, cells outside the
  // footprint's nonzero box count as no visit (their visits have fh = 0)
  for (int e = t; e < (kT + 2) * (kT + 2); e += kT * kT) {
    const int hy = e / (kT + 2), hx = e % (kT + 2);
    int64_t ic;
    sm[hy][hx] = clocked ? stamp_visit(g, col0 + hx - 1, row0 + hy - 1, &ic)
                     : stamp_visit_in(g, col0 + hx - 1, row0 + hy - 1, fb.x0, fb.y0, fb.x1,
                                      fb.y1, &ic);
  }
  __syncthreads();
  const int mm = sm[ly + 1][lx + 1];
  // Tiles of up to 16x16 threads also load the partner visit's pixel in the
  // first batch (it is another canvas pixel, which no other visit of this
  // stamp touches), saving a dependent round trip; larger tiles keep the
  // registers and load it after the first visit.
  constexpr bool kPre = kT <= 16;
  bool act = false;
  int pdir = 0;
  int64_t ic = 0, pic = 0;
  bool pv = false, ppull = false;
  float fh = 0.0f;
  CachedVisit<RegPick> acc;
  float pcv[kPre ? 6 : 1], psv[kPre ? 6 : 1];
  double pcvv = 0.0, psvv = 0.0;
  if (mm >= 0 && sm[ly + 1][lx] != mm && sm[ly][lx + 1] != mm) {
    const int col = col0 + lx, row = row0 + ly;
    pdir = sm[ly + 1][lx + 2] == mm ? 1 : (sm[ly + 2][lx + 1] == mm ? 2 : 0);
    if (!(dup && (on_dup_line(st, col, row) ||
                  (pdir && on_dup_line(st, col + (pdir == 1), row + (pdir == 2)))))) {
      act = true;
      const int x = (int)(col + g.cx), y = (int)(row + g.cy);  // canvas_index's pixel
      ic = (int64_t)y * g.cols + x;
      pv = pull && on_window_border(st.box, x, y);
      // the footprint height and every value the visits read go out in one
      // batch (a visit with fh <= 0 then uses none of them)
      fh = fp[mm];
#pragma unroll
      for (int c = 0; c < 7; ++c) acc.v[c] = pick[c][mm];
      acc.alias = !T.use_snap;
      load_pixel(T, ic, pv, acc);
      if (pdir) {  // the second visitor of the cell, right or below
        const int px = (int)(col + (pdir == 1) + g.cx), py = (int)(row + (pdir == 2) + g.cy);
        pic = (int64_t)py * g.cols + px;
        ppull = pull && on_window_border(st.box, px, py);
        if constexpr (kPre) {
          ld_px(*T.canvas, pic, pcv, pcvv);
          if (T.use_snap && !ppull) {
            ld_px(*T.src, pic, psv, psvv);
          } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) psv[c] = pcv[c];
            psvv = pcvv;
          }
        }
      }
    }
  }
  if (act && clocked && !(fh > 0.0f)) {  // visits that only check
    check_only(T, clk, ic, pv);
    if (pdir) check_only(T, clk, pic, ppull);
  }
  act = act && fh > 0.0f;
  if (act) {
    if (clocked) {
      check_dry_regs(clk, acc.cv, acc.cvv, ic);
      if (!T.use_snap) acc.src_from_canvas();  // the pickup source is the canvas itself
    }
    run_pixel(T, ic, pv, fh, pk, ps, acc);
    if (pdir) {
      acc.alias = !T.use_snap;
      if constexpr (kPre) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          acc.cv[c] = pcv[c];
          acc.sv[c] = psv[c];
        }
        acc.cvv = pcvv;
        acc.svv = psvv;
        acc.src_moved = false;
      } else {
        load_pixel(T, pic, ppull, acc);  // the pickup cell stays in acc.v
      }
      if (clocked) {
        check_dry_regs(clk, acc.cv, acc.cvv, pic);
        if (!T.use_snap) acc.src_from_canvas();
      }
      run_pixel(T, pic, ppull, fh, pk, ps, acc);
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) pick[c][mm] = acc.v[c];
  }
  __syncthreads();  // sm is rewritten by the block's next tile
}

// Is visit (col, row), at canvas pixel ic on the window border of box b,
// the first visit of that pixel in reference order (rows, then columns)
// that acts: footprint height > 0, or any visit when `any_visit` (a drying
// clock)?  Only a pixel of a duplicated column/row 0 has more than one.
__device__ __forceinline__ bool first_acting_visit(const StampGeom& g, const RingBox& b,
                                                   const float* fp, int col, int row, int64_t ic,
                                                   bool any_visit) {
  const int x = (int)(ic % g.cols), y = (int)(ic / g.cols);
  if (!on_window_border(b, x, y)) return false;
  int cs[2], rs[2];
  const int nc = window_offsets(x, g.cx, g.hr, cs), nr = window_offsets(y, g.cy, g.hr, rs);
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nc; ++c) {
      int64_t i2;
      const int m = stamp_visit(g, cs[c], rs[r], &i2);
      if (m >= 0 && (any_visit || fp[m] > 0.0f)) return cs[c] == col && rs[r] == row;
    }
  return false;
}

// Visits sharing a canvas pixel on column/row 0 (and their map-cell
// partners), listed in reference order and grouped by dependency level; the
// threads of one workgroup, levels separated by barriers.  With a clock,
// each listed visit first checks its pixel (the second check of a pixel in
// one imprint finds its time already at the imprint's and changes nothing).
//
// With `pull` (the stamp's window tiles and ring scan run in the same
// launch), a listed visit on the window border refreshes its own snapshot
// pixel when it is the first visit of that pixel that acts (footprint
// height > 0, or any visit with a clock) in reference order; the ring scan
// leaves such pixels alone.
template <int kThreads>
__device__ __forceinline__ void level_visits(const VisitTarget& T, float* const* pick,
                                             const float* fp, const StampGeom& g, const PStamp& st,
                                             const int2* dl, const int* dl_off, const float pk[3],
                                             const float ps[3],
                                             const DryClock& clk = DryClock{},
                                             bool pull = false) {
  for (int L = 0; L < st.dl_nlev; ++L) {
    const int e0 = dl_off[st.dl_lev + L], e1 = dl_off[st.dl_lev + L + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kThreads) {
      const int2 v = dl[e];
      int64_t ic;
      const int m = stamp_visit(g, v.x, v.y, &ic);
      if (m < 0) continue;
      const bool pv = pull && first_acting_visit(g, st.box, fp, v.x, v.y, ic, clk.tmap != nullptr);
      if (clk.tmap) check_only(T, clk, ic, pv);  // with pv, the refresh comes first
      if (fp[m] > 0.0f)
        cached_visit(T, GlobalPick{pick, m}, fp[m], ic, pv && !clk.tmap, pk, ps);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this level's stores before the next level's loads
    __syncthreads();
  }
}

// ------------------------------------------------------------- host side --
// Plans stamps [0, n) of one list for brush b on canvas cv into h_st
// (appended), with the shared-pixel visit lists of dup stamps appended to
// h_dl / h_off (offsets are global indices into h_dl).
void plan_stamps(pnt_brush b, pnt_canvas cv, const pnt_stamp* st, int64_t n,
                 std::vector<PStamp>& h_st, std::vector<int2>& h_dl, std::vector<int>& h_off);
// The same for a brush geometry given explicitly (padded footprint side and
// radius), so that a list can change geometry between strokes.
void plan_stamps_geom(int size_map, double radius, pnt_canvas cv, const pnt_stamp* st, int64_t n,
                      std::vector<PStamp>& h_st, std::vector<int2>& h_dl, std::vector<int>& h_off);

}  // namespace pnt
