// One phase of a "wave": a single launch that advances many independent
// stamp chains at once, each on its own range of blocks (serve.hip runs a
// server's admitted requests this way, batch.hip a batch of canvases).
//
// A phase is one step of one chain: a canvas initialisation, a full
// snapshot copy (updateSnapshot(canvas), FootprintBrush.hxx:168-172), a
// pickup-map clear (dip()/clean(), :150-160), one imprint (:73-143) -- its
// window tiles, snapshot ring scan and, for an imprint that straddles canvas
// column/row 0, its shared-pixel visits by level, side by side -- a
// dryCanvas (Canvas.hxx:105-121), a composite (Renderer.hxx:26-41), or a
// pickup-map copy.  Phases of one wave touch disjoint memory; the stream
// orders the waves.  Every kind runs the same visit and KM arithmetic as the
// single-canvas kernels, so a chain gives the same bits as run alone.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"
#include "km.hpp"
#include "stampcore.hpp"

namespace pnt {

constexpr int kWT = 16;          // window tile side of wave phases
constexpr int kWT2 = kWT * kWT;  // threads per block
// Waves per SIMD the wave kernels keep (at most 128 VGPRs): four workgroups
// of a launch per CU.  They need 108-112; five (at most 102) spills.
constexpr int kWaveOccupancy = 4;

enum : int {
  kSInit = 0,  // canvas planes: no paint, R0 = bg, h = 0 (and the snapshot of it)
  kSSnap,      // snapshot = canvas, dirty bytes 0
  kSClear,     // pickup map = 0
  kSStamp,     // window tiles + ring scan (+ one block of shared-pixel levels) of one imprint
  kSDry,       // dryCanvas, marking the snapshot dirty bytes it changes
  kSCompose,   // composite of the wet layer over R0 into out
  kSCopyPick,  // pickup map -> pick_dst
};

// A canvas block: K0..2 S0..2 (FP32), V (FP64, two strides), R0_0..2, h;
// strides of cstride floats (kCanvasStrides of them).
struct SSlot {
  const PStamp* st;
  float* canvas;
  float* snap;      // a layer of cstride (the canvas itself when the snapshot is off)
  uint8_t* dirty;   // snapshot dirty bytes (dirty_bytes of them)
  float* pick;      // the chain's pickup map: 7 planes of mstride
  float* pick_dst;  // kSCopyPick target
  const float* fp;  // padded footprint, size_map^2
  const int2* dl;   // shared-pixel visit lists
  const int* dl_off;
  float* out[3];    // composite planes
  int cstride, mstride, dirty_bytes;
  int rows, cols;
  int kind, start, nblk, ntiles, rb, hr, size_map, tiles_x;
  int part, nparts;  // streaming kinds: part `part` of `nparts` of every range they cover
  int fb[4];
  float pk[3], ps[3];
  double pickup_rate, dep_rate, cap;
  int use_snap;
  float bg[3];
  int64_t* tmap;   // the canvas' drying-clock time map (nullptr: no clock)
  int64_t dry_ms;  // its drying time
};

// Part `part` of `nparts` of [0, n): [lo, hi).
__device__ __forceinline__ void part_range(int64_t n, int part, int nparts, int64_t& lo,
                                           int64_t& hi) {
  lo = n * part / nparts;
  hi = n * (part + 1) / nparts;
}

__device__ __forceinline__ void wave_fill4(float* p, int64_t lo, int64_t hi, float v, int tid,
                                           int nth) {
  const float4 x = make_float4(v, v, v, v);
  for (int64_t q = lo + tid; q < hi; q += nth) reinterpret_cast<float4*>(p)[q] = x;
}

// Canvas::dryCanvas (Canvas.hxx:105-121) of pixel i; marks the snapshot
// dirty byte of a pixel it changes (when dirty is given).
__device__ __forceinline__ void wave_dry_pixel(const Layer7& L, float* R0, float* h, int64_t cs,
                                               uint8_t* dirty, int64_t i) {
  float K[3], S[3], R[3];
  bool nz = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = L.ks[c][i];
    S[c] = L.ks[3 + c][i];
    R[c] = R0[c * cs + i];
    nz |= (K[c] != 0.0f) | (S[c] != 0.0f);
  }
  const double v = L.v[i];
  nz |= v != 0.0;
  if (!nz) return;  // drying an empty pixel changes nothing
  h[i] = h_plus(h[i], v);
  km_pixel(K, S, R, (float)v);
#pragma unroll
  for (int c = 0; c < 3; ++c) R0[c * cs + i] = R[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) L.ks[c][i] = 0.0f;
  L.v[i] = 0.0;
  if (dirty) dirty[i] = 1;
}

// Renderer::compose (Renderer.hxx:26-41) of pixel i.
__device__ __forceinline__ void wave_compose_pixel(const Layer7& L, const float* R0,
                                                   float* const out[3], int64_t cs, int64_t i) {
  float K[3], S[3], R[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = L.ks[c][i];
    S[c] = L.ks[3 + c][i];
    R[c] = R0[c * cs + i];
  }
  km_pixel(K, S, R, (float)L.v[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c][i] = R[c];
}

// Block b (of S.nblk) of phase S; all kWT2 threads of the block call it.
template <int kG = kRingGroups>
__device__ __forceinline__ void run_wave_slot(const SSlot& S, int b, int (*sm)[kWT + 2]) {
  const int tid = b * kWT2 + (int)threadIdx.x, nth = S.nblk * kWT2;
  const int64_t cs = S.cstride;
  const Layer7 canvas = layer_at(S.canvas, cs);
  const Layer7 snap = layer_at(S.snap, cs);
  float* R0 = S.canvas + 8 * cs;
  float* hp = S.canvas + 11 * cs;
  const int64_t n = (int64_t)S.rows * S.cols;
  // streaming kinds cover part S.part of S.nparts of each plane (float4s)
  // and of the dirty bytes (16-byte groups) or pixels they touch
  int64_t lo4, hi4, lod, hid;
  part_range(cs / 4, S.part, S.nparts, lo4, hi4);
  part_range(S.dirty_bytes / 16, S.part, S.nparts, lod, hid);
  switch (S.kind) {
    case kSInit: {
      // zero bits are 0.0f and 0.0: the wet layer (8 strides) and h
      for (int p = 0; p < kCanvasStrides; ++p)
        wave_fill4(S.canvas + p * cs, lo4, hi4, (p >= 8 && p < 11) ? S.bg[p - 8] : 0.0f, tid,
                   nth);
      if (S.use_snap) {  // a fresh brush's snapshot of the new canvas
        for (int p = 0; p < kLayerStrides; ++p) wave_fill4(S.snap + p * cs, lo4, hi4, 0.0f, tid, nth);
        for (int64_t q = lod + tid; q < hid; q += nth)
          reinterpret_cast<uint4*>(S.dirty)[q] = make_uint4(0u, 0u, 0u, 0u);
      }
      return;
    }
    case kSSnap: {
      for (int p = 0; p < kLayerStrides; ++p)  // raw copy: K, S and the FP64 V
        for (int64_t q = lo4 + tid; q < hi4; q += nth)
          reinterpret_cast<float4*>(S.snap + p * cs)[q] =
              reinterpret_cast<const float4*>(S.canvas + p * cs)[q];
      for (int64_t q = lod + tid; q < hid; q += nth)
        reinterpret_cast<uint4*>(S.dirty)[q] = make_uint4(0u, 0u, 0u, 0u);
      return;
    }
    case kSClear:
      wave_fill4(S.pick, 0, 7 * (int64_t)S.mstride / 4, 0.0f, tid, nth);
      return;
    case kSCopyPick:
      for (int64_t q = tid; q < 7 * (int64_t)S.mstride / 4; q += nth)
        reinterpret_cast<float4*>(S.pick_dst)[q] = reinterpret_cast<const float4*>(S.pick)[q];
      return;
    case kSDry: {  // with a clock, dryCanvas also stamps every pixel's time (its last imprint's)
      const int64_t now = S.tmap ? S.st->now_us : 0;
      int64_t lo, hi;
      part_range(n, S.part, S.nparts, lo, hi);
      for (int64_t i = lo + tid; i < hi; i += nth) {
        wave_dry_pixel(canvas, R0, hp, cs, S.use_snap ? S.dirty : nullptr, i);
        if (S.tmap) S.tmap[i] = now;
      }
      return;
    }
    case kSCompose: {
      int64_t lo, hi;
      part_range(n, S.part, S.nparts, lo, hi);
      for (int64_t i = lo + tid; i < hi; i += nth) wave_compose_pixel(canvas, R0, S.out, cs, i);
      return;
    }
    default:
      break;
  }
  const PStamp& st = *S.st;
  StampGeom g;
  g.cx = st.cx;
  g.cy = st.cy;
  g.c = st.c;
  g.s = st.s;
  g.hr = S.hr;
  g.size_map = S.size_map;
  g.rows = S.rows;
  g.cols = S.cols;
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  const bool pull = S.use_snap != 0;  // border visits refresh their own snapshot pixels
  const VisitTarget T{&canvas, &snap, S.dirty, S.use_snap, S.pickup_rate, S.dep_rate, S.cap};
  float* pick[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) pick[c] = S.pick + c * (int64_t)S.mstride;
  const FpBox fb{S.fb[0], S.fb[1], S.fb[2], S.fb[3]};
  // Canvas::checkDry on every visited pixel when the canvas has a clock
  const DryClock ck{S.tmap, st.now_us, S.dry_ms, Rgb3{{R0, R0 + cs, R0 + 2 * cs}}, hp};
  const int win = S.ntiles;
  if (b < win) {
    window_tile<kWT>(T, pick, S.fp, fb, g, st, dup, pull, tile_of_block(b, S.ntiles), S.tiles_x,
                     sm, S.pk, S.ps, ck);
  } else if (b < win + S.rb) {
    ring_scan<kG>(canvas, snap, S.dirty, S.fp, g, st.box, pull, (b - win) * kWT2 + (int)threadIdx.x,
                 S.rb * kWT2, S.tmap != nullptr);
  } else if (dup) {  // the shared-pixel visits, beside the tiles
    level_visits<kWT2>(T, pick, S.fp, g, st, S.dl, S.dl_off, S.pk, S.ps, ck, pull);
  }
}

// Blocks of a streaming phase over `items` work items (at most 2048).
inline int wave_stream_blocks(int64_t items) {
  const int64_t b = (items + kWT2 - 1) / kWT2;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

}  // namespace pnt
