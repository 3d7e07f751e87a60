// Device pieces shared by the launch-per-list stamp kernels (persistent.hip,
// batch.hip): the planned stamp descriptor, the bounded group barrier, the
// one-round-trip visit, and the snapshot-ring helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "internal.hpp"
#include "visit.hpp"

namespace pnt {

enum : uint32_t { kStDip = 1u, kStDupX = 2u, kStDupY = 4u };

// One imprint, planned on the host (FP64 geometry as stamp_args computes it).
struct PStamp {
  double cx, cy, c, s;
  float K[3], S[3];  // paint of a dip before this stamp
  uint32_t flags;    // kStDip | kStDupX | kStDupY
  int32_t dl_lev;    // first entry of this stamp's level offsets (dup stamps)
  int32_t dl_nlev;   // number of shared-pixel levels (0 if none)
  RingBox box;
  int32_t pad[6];
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
  float pickup_rate, dep_rate, cap;
};

template <class PickAcc>
__device__ __forceinline__ void cached_visit(const VisitTarget& T, const PickAcc& pick, float fh,
                                             int64_t ic, bool pull, const float pk[3],
                                             const float ps[3]) {
  CachedVisit<PickAcc> acc{pick, {}, {}, !T.use_snap, false};
#pragma unroll
  for (int c = 0; c < 7; ++c) acc.cv[c] = T.canvas->p[c][ic];
  if (T.use_snap && !pull) {
#pragma unroll
    for (int c = 0; c < 7; ++c) acc.sv[c] = T.src->p[c][ic];
  } else {
#pragma unroll
    for (int c = 0; c < 7; ++c) acc.sv[c] = acc.cv[c];
  }
  paint_visit_math(acc, fh, T.pickup_rate, T.dep_rate, T.cap, pk, ps);
#pragma unroll
  for (int c = 0; c < 7; ++c) T.canvas->p[c][ic] = acc.cv[c];
  if (T.use_snap) {
    if (pull) {
#pragma unroll
      for (int c = 0; c < 7; ++c) T.src->p[c][ic] = acc.sv[c];
    } else if (acc.src_moved) {
      T.src->p[6][ic] = acc.sv[6];
    }
    T.dirty[ic] = 1;
  }
}

// ----------------------------------------------------------- ring helpers -
__device__ __forceinline__ bool on_dup_line(const PStamp& st, int col, int row) {
  return ((st.flags & kStDupX) && (int)(col + st.cx) == 0) ||
         ((st.flags & kStDupY) && (int)(row + st.cy) == 0);
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
__device__ __forceinline__ bool ring_pixel_visited(const StampGeom& g, const RingBox& b,
                                                   const float* fp, int x, int y) {
  if (x < b.tl_x || x > b.br_x || y < b.tl_y || y > b.br_y) return false;
  int col, row;
  if (!window_offset(x, g.cx, g.hr, &col) || !window_offset(y, g.cy, g.hr, &row)) return false;
  int64_t ic;
  const int m = stamp_visit(g, col, row, &ic);
  return m >= 0 && fp[m] > 0.0f;
}

__device__ __forceinline__ bool on_window_border(const RingBox& b, int x, int y) {
  return x == b.tl_x || x == b.br_x || y == b.tl_y || y == b.br_y;
}

// ------------------------------------------------------------- host side --
// Plans stamps [0, n) of one list for brush b on canvas cv into h_st
// (appended), with the shared-pixel visit lists of dup stamps appended to
// h_dl / h_off (offsets are global indices into h_dl).
void plan_stamps(pnt_brush b, pnt_canvas cv, const pnt_stamp* st, int64_t n,
                 std::vector<PStamp>& h_st, std::vector<int2>& h_dl, std::vector<int>& h_off);

}  // namespace pnt
