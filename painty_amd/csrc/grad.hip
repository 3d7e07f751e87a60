// Differentiable composite + stamping: gradients of a loss on the composite
// R = Renderer::compose (Renderer.hxx:26-41) of a canvas after a stamp list
// (FootprintBrush::imprint, FootprintBrush.hxx:73-143, with each dip's paint
// from PaintMixer::mixSinglePaint, PaintMixer.cxx:327-355) with respect to
// the initial wet layer K, S, V, the substrate R0 and the mixing weights.
// The reference has no gradients; this is the extension it is measured on.
//
// Forward: the stamps run as PNT_EXEC_FUSED, split into segments of at most
// `seg_len` stamps that never cross a dip.  Before each segment the state it
// can touch is checkpointed: the canvas and snapshot inside the union of its
// stamps' snapshot boxes, and the pickup map.
//
// Backward, segment by segment from the last: restore the checkpoint and
// recompute the segment stamp by stamp (ring copy | tape | visits), where the
// tape of a stamp holds what its visits read -- the canvas and snapshot of
// its window box after the ring copy, and the pickup map.  Then, from the
// last stamp of the segment to the first:
//   visits   k_adj_visits: each cell's visits replayed from the tape (the
//            forward arithmetic of visit.hpp, so branches resolve the same
//            way) and their adjoints applied in reverse order;
//   shared   k_adj_shared: visits on a duplicated canvas column/row 0,
//            replayed on the tape and reversed level by level;
//   ring     k_adj_ring: the ring copy canvas -> snapshot moves the snapshot
//            adjoint onto the canvas;
//   dip      the pickup map was cleared: its adjoint is dropped.
// The state is recomputed, not stored: memory is the checkpoints (a few MB
// per segment) plus one segment of tapes.
#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "km.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"

namespace pnt {

namespace {

constexpr int kGT = 16;  // window tile of the adjoint kernel
constexpr int kBlk = 256;

// ------------------------------------------------------ visit adjoint ----
struct VState {
  float P[7];          // pickup cell (FP32, as stored)
  double C[7], X[7];   // canvas and source pixel (K, S from FP32; V is FP64)
};

// Adjoint of paint_visit_math (visit.hpp) at input state `in`.  On entry g
// holds the adjoints of the outputs (P, C and, with the snapshot, X); on
// return the adjoints of the inputs.  gB accumulates the adjoint of the
// brush paint.  The intermediates are recomputed with the forward's own
// operations (FP64 on the stored FP32 values), so every threshold branch is
// the one the forward took.
__device__ __forceinline__ void visit_adj(const VState& in, VState& g, float gB[6], float fh_f,
                                          double pr, double dr, double cap, const float pk[3],
                                          const float ps[3], bool alias) {
  const double fh = fh_f;
  const double B[6] = {pk[0], pk[1], pk[2], ps[0], ps[1], ps[2]};
  const float* Pf = in.P;
  const double* Cf = in.C;
  const double* Xf = alias ? in.C : in.X;
  double P[7], C[7], X[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    P[c] = Pf[c];
    C[c] = Cf[c];
    X[c] = Xf[c];
  }
  const double vP = P[6], vC = X[6];
  const double vLeave = (pr * vC) * fh;
  const bool a = vLeave > kMinVolume;
  double P1[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) P1[c] = P[c];
  double X6n = vC, iv1 = 0.0;
  bool bb = false;
  if (a) {
    X6n = vC - vLeave;  // the stored source volume (FP64)
    const double vt1 = vP + vLeave;
    bb = vt1 > kMinVolume;
    if (bb) {
      iv1 = 1.0 / vt1;
#pragma unroll
      for (int c = 0; c < 6; ++c) P1[c] = (float)blendd(vP, P[c], vLeave, X[c], iv1);
    }
    P1[6] = (float)(vP + vLeave);
  }
  const double u = P1[6];
  const bool f = cap - u > 0.0;
  const double vFree = fmax(0.0, cap - u);
  const double vt2 = u + vFree;
  const bool d = vt2 > kMinVolume;
  const double iv2 = d ? 1.0 / vt2 : 0.0;
  double src[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) src[c] = d ? blendd(u, P1[c], vFree, B[c], iv2) : P1[c];
  double Cd[7];  // the canvas pixel as depositPaint reads it
#pragma unroll
  for (int c = 0; c < 6; ++c) Cd[c] = C[c];
  Cd[6] = alias ? X6n : C[6];
  const double vB = cap * fh, vCc = Cd[6], vt3 = vB + vCc;
  const bool e = vt3 > kMinVolume;
  const double iv3 = e ? 1.0 / vt3 : 0.0;
  // ---- depositPaint
  double gCd[7], gsrc[6];
  double gvCc = g.C[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    if (e) {
      const double C1 = blendd(vB, src[c], vCc, Cd[c], iv3);
      gsrc[c] = g.C[c] * vB * iv3;
      gCd[c] = g.C[c] * vCc * iv3;
      gvCc += g.C[c] * (Cd[c] - C1) * iv3;
    } else {
      gsrc[c] = g.C[c];
      gCd[c] = 0.0;
    }
  }
  gCd[6] = gvCc;
  double gP1[7];
  double gu = g.P[6] * (1.0 - dr * fh);
  if (d) {
    double gvFree = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      gP1[c] = g.P[c] + gsrc[c] * u * iv2;
      gB[c] += (float)(gsrc[c] * vFree * iv2);
      gu += gsrc[c] * (P1[c] - src[c]) * iv2;
      gvFree += gsrc[c] * (B[c] - src[c]) * iv2;
    }
    if (f) gu -= gvFree;
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c) gP1[c] = g.P[c] + gsrc[c];
  }
  gP1[6] = gu;
  // ---- pickupPaint
  double gXo[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) gXo[c] = alias ? gCd[c] : g.X[c];
  double gPin[7], gXin[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) gXin[c] = gXo[c];
  if (a) {
    double gvP = gP1[6], gvLeave = gP1[6] - gXo[6], gvC = gXo[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (bb) {
        gPin[c] = gP1[c] * vP * iv1;
        gXin[c] += gP1[c] * vLeave * iv1;
        gvP += gP1[c] * (P[c] - P1[c]) * iv1;
        gvLeave += gP1[c] * (X[c] - P1[c]) * iv1;
      } else {
        gPin[c] = gP1[c];
      }
    }
    gvC += gvLeave * pr * fh;
    gPin[6] = gvP;
    gXin[6] = gvC;
  } else {
#pragma unroll
    for (int c = 0; c < 7; ++c) gPin[c] = gP1[c];
  }
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    g.P[c] = (float)gPin[c];
    g.C[c] = alias ? gXin[c] : gCd[c];
    g.X[c] = alias ? 0.0 : gXin[c];
  }
}

// The forward visit on a state held in registers (the same instructions as
// every stamp kernel).
__device__ __forceinline__ void visit_fwd(VState& s, float fh, double pr, double dr, double cap,
                                          const float pk[3], const float ps[3], bool alias) {
  CachedVisit<RegPick> acc;
#pragma unroll
  for (int c = 0; c < 7; ++c) acc.v[c] = s.P[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    acc.cv[c] = (float)s.C[c];
    acc.sv[c] = (float)(alias ? s.C[c] : s.X[c]);
  }
  acc.cvv = s.C[6];
  acc.svv = alias ? s.C[6] : s.X[6];
  acc.alias = alias;
  acc.src_moved = false;
  paint_visit_math(acc, fh, pr, dr, cap, pk, ps);
#pragma unroll
  for (int c = 0; c < 7; ++c) s.P[c] = acc.v[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    s.C[c] = acc.cv[c];
    s.X[c] = alias ? acc.cv[c] : acc.sv[c];
  }
  s.C[6] = acc.cvv;
  s.X[6] = alias ? acc.cvv : acc.svv;
}

// One stamp's tape: canvas and snapshot of the window box as the visits
// find them (7 FP64 planes of `area` each: K, S widened, V as stored), then
// the pickup map (7 FP32 planes of m).
struct Tape {
  const double* C;
  const double* X;
  const float* P;
  int64_t area, m;
  int bx0, by0, bw, bh;
};

struct AArgs {
  const PStamp* st;
  int j;  // stamp index into st
  int hr, size_map, rows, cols, tiles_x, ntiles, use_snap;
  const float* fp;
  Tape t;
  Plane7 gC, gX, gP;  // adjoints: canvas (caller's planes), snapshot, pickup map
  float* gB;          // 6 floats of this stamp's paint adjoint
  double pr, dr, cap;
  float pk[3], ps[3];
  // shared-pixel visits of this stamp (plan lists) and their scratch
  const int2* dl;
  const int* dl_off;
  float* scratch;  // per-visit records of the shared-pixel replay
};

__device__ __forceinline__ StampGeom geom_of(const AArgs& A, const PStamp& st) {
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

__device__ __forceinline__ int64_t tape_index(const Tape& t, int64_t ic, int cols) {
  const int x = (int)(ic % cols), y = (int)(ic / cols);
  return (int64_t)(y - t.by0) * t.bw + (x - t.bx0);
}

__device__ __forceinline__ void load_in(const AArgs& A, int64_t ic, int mm, VState& s) {
  const int64_t k = tape_index(A.t, ic, A.cols);
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    s.P[c] = A.t.P[c * A.t.m + mm];
    s.C[c] = A.t.C[c * A.t.area + k];
    s.X[c] = A.use_snap ? A.t.X[c * A.t.area + k] : 0.0;
  }
}

__device__ __forceinline__ void load_g_pixel(const AArgs& A, int64_t ic, VState& g) {
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    g.C[c] = A.gC.p[c][ic];
    g.X[c] = A.use_snap ? A.gX.p[c][ic] : 0.0f;
  }
}

__device__ __forceinline__ void store_g_pixel(const AArgs& A, int64_t ic, const VState& g) {
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    A.gC.p[c][ic] = (float)g.C[c];
    if (A.use_snap) A.gX.p[c][ic] = (float)g.X[c];
  }
}

__device__ __forceinline__ void flush_gB(const AArgs& A, const float gB[6]) {
  __shared__ float s_gB[6];
  if (threadIdx.x < 6) s_gB[threadIdx.x] = 0.0f;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 6; ++c)
    if (gB[c] != 0.0f) atomicAdd(&s_gB[c], gB[c]);
  __syncthreads();
  if (threadIdx.x < 6 && s_gB[threadIdx.x] != 0.0f) atomicAdd(&A.gB[threadIdx.x], s_gB[threadIdx.x]);
}

// Adjoint of the visits of one stamp that do not share a canvas pixel, one
// kGT x kGT window tile per block (cells and partners as window_tile).
__global__ __launch_bounds__(kGT* kGT) void k_adj_visits(AArgs A) {
  __shared__ int sm[kGT + 2][kGT + 2];
  const PStamp& st = A.st[A.j];
  const StampGeom g = geom_of(A, st);
  const bool dup = (st.flags & (kStDupX | kStDupY)) != 0;
  const int tile = blockIdx.x;
  const int t = threadIdx.x, lx = t % kGT, ly = t / kGT;
  const int col0 = -g.hr + (tile % A.tiles_x) * kGT;
  const int row0 = -g.hr + (tile / A.tiles_x) * kGT;
  for (int e = t; e < (kGT + 2) * (kGT + 2); e += kGT * kGT) {
    const int hy = e / (kGT + 2), hx = e % (kGT + 2);
    int64_t ic;
    sm[hy][hx] = stamp_visit(g, col0 + hx - 1, row0 + hy - 1, &ic);
  }
  __syncthreads();
  float gB[6] = {0, 0, 0, 0, 0, 0};
  const int mm = sm[ly + 1][lx + 1];
  if (mm >= 0 && sm[ly + 1][lx] != mm && sm[ly][lx + 1] != mm) {
    const int col = col0 + lx, row = row0 + ly;
    const int pdir = sm[ly + 1][lx + 2] == mm ? 1 : (sm[ly + 2][lx + 1] == mm ? 2 : 0);
    const bool skip = dup && (on_dup_line(st, col, row) ||
                              (pdir && on_dup_line(st, col + (pdir == 1), row + (pdir == 2))));
    const float fh = A.fp[mm];
    if (!skip && fh > 0.0f) {
      const bool alias = !A.use_snap;
      const int64_t ic1 = canvas_index(g, col, row);
      VState in1, gs;
      load_in(A, ic1, mm, in1);
#pragma unroll
      for (int c = 0; c < 7; ++c) gs.P[c] = A.gP.p[c][mm];
      if (pdir) {
        const int64_t ic2 = canvas_index(g, col + (pdir == 1), row + (pdir == 2));
        VState in2 = in1;
        visit_fwd(in2, fh, A.pr, A.dr, A.cap, A.pk, A.ps, alias);  // pickup cell after visit 1
        const int64_t k2 = tape_index(A.t, ic2, A.cols);
#pragma unroll
        for (int c = 0; c < 7; ++c) {
          in2.C[c] = A.t.C[c * A.t.area + k2];
          in2.X[c] = A.use_snap ? A.t.X[c * A.t.area + k2] : 0.0;
        }
        load_g_pixel(A, ic2, gs);
        visit_adj(in2, gs, gB, fh, A.pr, A.dr, A.cap, A.pk, A.ps, alias);
        store_g_pixel(A, ic2, gs);
      }
      load_g_pixel(A, ic1, gs);
      visit_adj(in1, gs, gB, fh, A.pr, A.dr, A.cap, A.pk, A.ps, alias);
      store_g_pixel(A, ic1, gs);
#pragma unroll
      for (int c = 0; c < 7; ++c) A.gP.p[c][mm] = gs.P[c];
    }
  }
  flush_gB(A, gB);
}

// Adjoint of a stamp's visits on a duplicated canvas column/row 0 and their
// partners.  Only these visits touch those pixels and cells, so the tape
// holds their true starting state; the tape of stamp j is not read after
// its adjoint, so the replay runs on it in place.  Visits are replayed level
// by level (a level's visits share no map cell and no canvas pixel, so each
// level runs in parallel), each visit's input recorded; then the adjoints
// run level by level in reverse.
__global__ __launch_bounds__(kBlk) void k_adj_shared(AArgs A) {
  const PStamp& st = A.st[A.j];
  if (st.dl_nlev <= 0) return;
  const int base = A.dl_off[st.dl_lev];
  VState* rec = reinterpret_cast<VState*>(A.scratch);
  const StampGeom g = geom_of(A, st);
  const bool alias = !A.use_snap;
  double* wC = const_cast<double*>(A.t.C);  // C, X (7 planes of area each), P (7 of m)
  double* wX = const_cast<double*>(A.t.X);
  float* wP = const_cast<float*>(A.t.P);
  for (int L = 0; L < st.dl_nlev; ++L) {  // replay in plan order, a level at a time
    const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlk) {
      const int2 v = A.dl[e];
      int64_t ic;
      const int mm = stamp_visit(g, v.x, v.y, &ic);
      if (mm < 0 || !(A.fp[mm] > 0.0f)) continue;
      const int64_t k = tape_index(A.t, ic, A.cols);
      VState sv;
#pragma unroll
      for (int c = 0; c < 7; ++c) {
        sv.P[c] = wP[c * A.t.m + mm];
        sv.C[c] = wC[c * A.t.area + k];
        sv.X[c] = A.use_snap ? wX[c * A.t.area + k] : 0.0;
      }
      rec[e - base] = sv;
      visit_fwd(sv, A.fp[mm], A.pr, A.dr, A.cap, A.pk, A.ps, alias);
#pragma unroll
      for (int c = 0; c < 7; ++c) {
        wP[c * A.t.m + mm] = sv.P[c];
        wC[c * A.t.area + k] = sv.C[c];
        if (A.use_snap) wX[c * A.t.area + k] = sv.X[c];
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // this level's stores before the next level's loads
    __syncthreads();
  }
  float gB[6] = {0, 0, 0, 0, 0, 0};
  for (int L = st.dl_nlev - 1; L >= 0; --L) {  // adjoints, levels in reverse
    const int e0 = A.dl_off[st.dl_lev + L], e1 = A.dl_off[st.dl_lev + L + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlk) {
      const int2 v = A.dl[e];
      int64_t ic;
      const int mm = stamp_visit(g, v.x, v.y, &ic);
      if (mm < 0 || !(A.fp[mm] > 0.0f)) continue;
      VState gs;
#pragma unroll
      for (int c = 0; c < 7; ++c) gs.P[c] = A.gP.p[c][mm];
      load_g_pixel(A, ic, gs);
      visit_adj(rec[e - base], gs, gB, A.fp[mm], A.pr, A.dr, A.cap, A.pk, A.ps, alias);
      store_g_pixel(A, ic, gs);
#pragma unroll
      for (int c = 0; c < 7; ++c) A.gP.p[c][mm] = gs.P[c];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  flush_gB(A, gB);
}

// Adjoint of the ring copy canvas -> snapshot (FootprintBrush.hxx:307-318).
__global__ __launch_bounds__(kBlk) void k_adj_ring(RingBox b, int cols, Plane7 gC, Plane7 gX) {
  const int w = b.a_x1 - b.a_x0 + 1;
  const int64_t total = (int64_t)w * (b.a_y1 - b.a_y0 + 1);
  for (int64_t e = (int64_t)blockIdx.x * kBlk + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kBlk) {
    const int row = b.a_y0 + (int)(e / w), col = b.a_x0 + (int)(e % w);
    if (row > b.tl_y && row < b.br_y && col > b.tl_x && col < b.br_x) continue;
    const int64_t i = (int64_t)row * cols + col;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      gC.p[c][i] += gX.p[c][i];
      gX.p[c][i] = 0.0f;
    }
  }
}

// Adjoint of the first imprint's full canvas -> snapshot copy.
__global__ __launch_bounds__(kBlk) void k_adj_full_copy(int64_t n, Plane7 gC, Plane7 gX) {
  for (int64_t i = (int64_t)blockIdx.x * kBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlk)
#pragma unroll
    for (int c = 0; c < 7; ++c) gC.p[c][i] += gX.p[c][i];
}

// Adjoint of ComputeReflectance (KubelkaMunk.hxx:25-83) in the stable form
// of km.hpp, with Q = S*b*coth(x) = B (1 + 2/E), N = S - R0 (A - Q),
// D = A - R0 S + Q, R = N/D, x = B d.  The numerators are expanded so that
// the terms in Q (~ 1/d for thin layers) cancel symbolically, not in FP32:
//   dR/dR0 = (N S - (A-Q) D)/D^2 = B^2 csch^2 x / D^2
//   dR/dA  = -(R0 D + N)/D^2     = -(S (1 - R0^2) + 2 R0 Q)/D^2
//   dR/dQ  = (R0 D - N)/D^2      = (2 R0 A - R0^2 S - S)/D^2
//   dR/dS|direct = (D + R0 N)/D^2 = (A (1 - R0^2) + Q (1 + R0^2))/D^2
//   dQ/dB = coth x - x csch^2 x = d/dx (x coth x), by its Taylor series
//           2x/3 - 4x^3/45 + 12x^5/945 - 8x^7/4725 + 20x^9/93555 below x = 0.5
//   dQ/dd = -B^2 csch^2 x,  coth x = 1 + 2/E,  csch^2 x = 4(E+1)/E^2,
//   dB/dK = A/B,  dB/dS = K/B.
// Where x > 20 the reference's coth is the constant 1 (csch^2 -> 0).
__device__ __forceinline__ void km_adj_channel(float K, float Sin, float R0, float d, float gR,
                                               float* gK, float* gS, float* gd, float* gR0) {
  const bool clamp = !(__builtin_fabsf(Sin) > kKmThr);
  const float S = clamp ? kSClamp : Sin;
  const float A = S + K;
  const float q = K * (K + 2.0f * S);
  const float B = __builtin_amdgcn_sqrtf(q > 0.0f ? q : 0.0f);
  const float x = B * d;
  float y = 2.0f * x;
  const bool flat = y >= 40.0f;
  y = flat ? 40.0f : y;
  const float E = km_expm1(y);
  const float iE = rcpf(E);
  const float ct = 1.0f + 2.0f * iE;
  const float csch2 = flat ? 0.0f : 4.0f * (E + 1.0f) * iE * iE;
  const float Q = B * ct;
  const float D = A - R0 * S + Q;
  const float iD2 = rcpf(D * D);
  const float R02 = R0 * R0;
  const float dR0 = (B * B * csch2) * iD2;
  const float dA = -(S * (1.0f - R02) + 2.0f * R0 * Q) * iD2;
  const float dQ = (2.0f * R0 * A - R02 * S - S) * iD2;
  const float dSd = (A * (1.0f - R02) + Q * (1.0f + R02)) * iD2;
  const float x2 = x * x;
  const float series =
      x * (2.0f / 3.0f +
           x2 * (-4.0f / 45.0f + x2 * (12.0f / 945.0f + x2 * (-8.0f / 4725.0f + x2 * (20.0f / 93555.0f)))));
  const float dQdB = x < 0.5f ? series : ct - x * csch2;
  const float dQdd = -B * B * csch2;
  const float iB = B > 0.0f ? rcpf(B) : 0.0f;
  *gK = gR * (dA + dQ * dQdB * A * iB);
  *gS = clamp ? 0.0f : gR * (dSd + dA + dQ * dQdB * K * iB);
  *gd += gR * dQ * dQdd;
  *gR0 = gR * dR0;
}

__global__ __launch_bounds__(kBlk) void k_km_adj(int64_t n, Layer7 L, Rgb3 R0, Rgb3 gR, Plane7 gL,
                                                 Rgb3 gR0) {
  for (int64_t i = (int64_t)blockIdx.x * kBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlk) {
    const float d = (float)L.v[i];
    float gd = 0.0f;
    if (__builtin_fabsf(d) < kKmThr) {  // R = R0
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gL.p[c][i] = 0.0f;
        gL.p[3 + c][i] = 0.0f;
        gR0.p[c][i] = gR.p[c][i];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float gk, gs, gr0;
        km_adj_channel(L.ks[c][i], L.ks[3 + c][i], R0.p[c][i], d, gR.p[c][i], &gk, &gs, &gd, &gr0);
        gL.p[c][i] = gk;
        gL.p[3 + c][i] = gs;
        gR0.p[c][i] = gr0;
      }
    }
    gL.p[6][i] = gd;
  }
}

// A w x h box at (x0, y0) of a layer's planes (row stride `cols`) <-> a
// dense buffer of 7 planes (plane stride `stride`): FP64 planes for a
// canvas or snapshot layer (K, S widened exactly, V as stored), FP32 planes
// for the pickup map.  Up to three boxes per launch.
struct BoxRegion {
  Layer7 L;   // when layer
  Plane7 P;   // otherwise
  int layer;
  int cols, x0, y0, w, h;
  void* buf;  // double* (layer) or float*
  int64_t stride;
};

inline BoxRegion layer_box(const Layer7& L, int cols, int x0, int y0, int w, int h, double* buf,
                           int64_t stride) {
  BoxRegion r{};
  r.L = L;
  r.layer = 1;
  r.cols = cols;
  r.x0 = x0;
  r.y0 = y0;
  r.w = w;
  r.h = h;
  r.buf = buf;
  r.stride = stride;
  return r;
}

inline BoxRegion plane_box(const Plane7& P, int cols, int x0, int y0, int w, int h, float* buf,
                           int64_t stride) {
  BoxRegion r{};
  r.P = P;
  r.layer = 0;
  r.cols = cols;
  r.x0 = x0;
  r.y0 = y0;
  r.w = w;
  r.h = h;
  r.buf = buf;
  r.stride = stride;
  return r;
}

__global__ __launch_bounds__(kBlk) void k_box_copy(BoxRegion r0, BoxRegion r1, BoxRegion r2,
                                                   int64_t n0, int64_t n1, int64_t n2,
                                                   int to_buf) {
  const int64_t total = n0 + n1 + n2;
  for (int64_t t = (int64_t)blockIdx.x * kBlk + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kBlk) {
    const BoxRegion& r = t < n0 ? r0 : (t < n0 + n1 ? r1 : r2);
    const int64_t e = t < n0 ? t : (t < n0 + n1 ? t - n0 : t - n0 - n1);
    const int64_t i = (int64_t)(r.y0 + e / r.w) * r.cols + (r.x0 + e % r.w);
    if (r.layer) {
      double* b = static_cast<double*>(r.buf);
      if (to_buf) {
#pragma unroll
        for (int c = 0; c < 6; ++c) b[c * r.stride + e] = r.L.ks[c][i];
        b[6 * r.stride + e] = r.L.v[i];
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) r.L.ks[c][i] = (float)b[c * r.stride + e];
        r.L.v[i] = b[6 * r.stride + e];
      }
    } else {
      float* b = static_cast<float*>(r.buf);
#pragma unroll
      for (int c = 0; c < 7; ++c) {
        if (to_buf)
          b[c * r.stride + e] = r.P.p[c][i];
        else
          r.P.p[c][i] = b[c * r.stride + e];
      }
    }
  }
}

inline unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlk - 1) / kBlk, 4096));
}

// Copies up to three boxes in one launch (boxes with w or h <= 0 are skipped).
void box_copy(std::initializer_list<BoxRegion> regions, bool to_buf, hipStream_t s) {
  BoxRegion r[3] = {};
  int64_t n[3] = {0, 0, 0};
  int k = 0;
  for (const BoxRegion& b : regions) {
    if (b.w <= 0 || b.h <= 0) continue;
    r[k] = b;
    n[k] = (int64_t)b.w * b.h;
    ++k;
  }
  const int64_t total = n[0] + n[1] + n[2];
  if (total == 0) return;
  for (int i = k; i < 3; ++i) r[i] = r[0];
  k_box_copy<<<grid_for(total), kBlk, 0, s>>>(r[0], r[1], r[2], n[0], n[1], n[2], to_buf ? 1 : 0);
  PNT_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace pnt

// ===================================================================== host
using namespace pnt;

struct pnt_grad_s {
  pnt_ctx ctx = nullptr;
  pnt_brush b = nullptr;
  pnt_canvas cv = nullptr;  // owned working canvas
  int rows = 0, cols = 0;
  std::vector<pnt_stamp> st;
  std::vector<int> paint_of;
  int n_pig = 0, n_paint = 0;
  std::vector<double> palK, palS, W;
  int seg_len = 32;
  struct Seg {
    int s0, s1;
    int x0, y0, w, h;  // checkpoint box
    int64_t off;       // checkpoint offset (floats)
  };
  std::vector<Seg> segs;
  float* ckpt = nullptr;
  int64_t ckpt_cap = 0;
  float* tape = nullptr;
  int64_t tape_cap = 0;
  float* gX = nullptr;  // 7 planes of round_up(n, 64) (snapshot adjoint)
  float* gP = nullptr;  // 7 planes of m (pickup adjoint)
  float* gB = nullptr;  // n_stamps x 6
  int64_t gX_cap = 0, gP_cap = 0, gB_cap = 0;
  float* scratch = nullptr;
  int64_t scratch_cap = 0;
  bool have_forward = false;
  int64_t n_stride = 0, m = 0;
};

namespace pnt {
int guard_call(const std::function<void()>& f);  // api.hip: status + pnt_last_error
}

namespace {

template <class T>
void ensure(T*& p, int64_t& cap, int64_t need, hipStream_t s) {
  if (need <= cap) return;
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  if (p) PNT_HIP_CHECK(hipFree(p));
  p = nullptr;
  PNT_HIP_CHECK(hipMalloc((void**)&p, (size_t)need * sizeof(T)));
  cap = need;
}

Plane7 dense7(float* base, int64_t stride) {
  Plane7 L;
  for (int c = 0; c < 7; ++c) L.p[c] = base + c * stride;
  return L;
}

// Window box of a stamp (canvas pixels its visits can touch), clipped.
void window_box(const PStamp& p, int rows, int cols, int* x0, int* y0, int* w, int* h) {
  const int a = std::max(p.box.tl_x, 0), c = std::min(p.box.br_x, cols - 1);
  const int e = std::max(p.box.tl_y, 0), f = std::min(p.box.br_y, rows - 1);
  *x0 = a;
  *y0 = e;
  *w = std::max(c - a + 1, 0);
  *h = std::max(f - e + 1, 0);
}

}  // namespace

static void check_status(int st) {
  if (st != PNT_OK) throw Error(st, pnt_last_error());
}

extern "C" {

int pnt_grad_create(pnt_ctx ctx, int rows, int cols, pnt_brush brush, const pnt_stamp* stamps,
                    int64_t n, const int32_t* paint_of, int n_pig, const double* palK,
                    const double* palS, int n_paint, int seg_len, pnt_grad* out) {
  return guard_call([&] {
    PNT_REQUIRE(ctx && brush && out, "NULL argument");
    PNT_REQUIRE(brush->ctx->device == ctx->device, "brush on another device");
    PNT_REQUIRE(rows > 0 && cols > 0 && (int64_t)rows * cols < ((int64_t)1 << 31),
                "rows/cols out of range");
    PNT_REQUIRE(n > 0 && stamps && paint_of, "empty stamp list");
    PNT_REQUIRE(n_pig > 0 && n_paint > 0 && palK && palS, "empty palette");
    PNT_REQUIRE(seg_len > 0, "segment length must be positive");
    PNT_REQUIRE(stamps[0].flags & PNT_STAMP_DIP, "the first stamp must dip (its paint is mixed)");
    for (int64_t i = 0; i < n; ++i) {
      PNT_REQUIRE(!(stamps[i].flags & PNT_STAMP_DRY_AFTER),
                  "dryCanvas inside a differentiable stamp list is not supported");
      PNT_REQUIRE(std::isfinite(stamps[i].cx) && std::isfinite(stamps[i].cy) &&
                      std::isfinite(stamps[i].theta),
                  "non-finite stamp");
      if (stamps[i].flags & PNT_STAMP_DIP)
        PNT_REQUIRE(paint_of[i] >= 0 && paint_of[i] < n_paint, "paint index out of range");
    }
    auto* g = new pnt_grad_s;
    g->ctx = ctx;
    g->b = brush;
    g->rows = rows;
    g->cols = cols;
    g->st.assign(stamps, stamps + n);
    g->paint_of.assign(paint_of, paint_of + n);
    g->n_pig = n_pig;
    g->n_paint = n_paint;
    g->palK.assign(palK, palK + 3 * n_pig);
    g->palS.assign(palS, palS + 3 * n_pig);
    g->seg_len = seg_len;
    const int st_create = pnt_canvas_create(ctx, rows, cols, &g->cv);
    if (st_create != PNT_OK) {
      const std::string msg = pnt_last_error();
      delete g;
      throw Error(st_create, msg);
    }
    ctx_retain(ctx);
    *out = g;
  });
}

int pnt_grad_destroy(pnt_grad g) {
  return guard_call([&] {
    if (!g) return;
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->cv) pnt_canvas_destroy(g->cv);
    (void)hipFree(g->ckpt);
    (void)hipFree(g->tape);
    (void)hipFree(g->gX);
    (void)hipFree(g->gP);
    (void)hipFree(g->gB);
    (void)hipFree(g->scratch);
    pnt_ctx ctx = g->ctx;
    delete g;
    ctx_release(ctx);
  });
}

int pnt_grad_forward(pnt_grad g, pnt_dev_layer wet0, pnt_dev_rgb R0, const double* W,
                     pnt_dev_rgb R) {
  return guard_call([&] {
    PNT_REQUIRE(g && W, "NULL argument");
    for (int c = 0; c < 7; ++c) PNT_REQUIRE(wet0.p[c], "NULL wet-layer plane");
    for (int c = 0; c < 3; ++c) PNT_REQUIRE(R0.p[c] && R.p[c], "NULL plane");
    g->ctx->use();
    pnt_brush b = g->b;
    pnt_canvas cv = g->cv;
    hipStream_t s = g->ctx->stream;
    const int64_t n = cv->n;
    PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    const Layer7 L = cv->layer();
    const Rgb3 R0c = cv->R0();
    Plane7 W0;
    for (int c = 0; c < 7; ++c) W0.p[c] = wet0.p[c];
    launch_layer_from_planes(n, W0, L, s);  // V widened to FP64
    for (int c = 0; c < 3; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(R0c.p[c], R0.p[c], n * sizeof(float), hipMemcpyDeviceToDevice, s));
    cv->epoch++;
    g->W.assign(W, W + (int64_t)g->n_paint * g->n_pig);
    // each dip's paint: mixSinglePaint(W[paint_of[i]]) in FP64, stored as FP32
    for (size_t i = 0; i < g->st.size(); ++i) {
      if (!(g->st[i].flags & PNT_STAMP_DIP)) continue;
      double K[3], S[3];
      check_status(pnt_mix_single(g->n_pig, g->palK.data(), g->palS.data(), g->n_pig,
                                  g->W.data() + (int64_t)g->paint_of[i] * g->n_pig, K, S));
      for (int c = 0; c < 3; ++c) {
        g->st[i].K[c] = (float)K[c];
        g->st[i].S[c] = (float)S[c];
      }
    }
    // the first imprint on a canvas copies it whole into the snapshot
    if (b->use_snap) brush_full_snapshot(b, cv);
    // segments (at most seg_len stamps, a dip starts a new one) and their
    // checkpoint boxes: the union of their snapshot boxes (window boxes
    // without the snapshot)
    g->segs.clear();
    const int ns = (int)g->st.size();
    int64_t total = 0;
    const int64_t m = b->m;
    for (int s0 = 0; s0 < ns;) {
      int s1 = s0 + 1;
      while (s1 < ns && s1 - s0 < g->seg_len && !(g->st[s1].flags & PNT_STAMP_DIP)) ++s1;
      int x0 = cv->cols, y0 = cv->rows, x1 = -1, y1 = -1;
      for (int i = s0; i < s1; ++i) {
        const StampArgs a = stamp_args(b, cv, g->st[i].cx, g->st[i].cy, g->st[i].theta);
        const RingBox rb = ring_box(a.g, b->radius);
        int bx0 = rb.a_x0, by0 = rb.a_y0, bx1 = rb.a_x1, by1 = rb.a_y1;
        if (!b->use_snap) {
          bx0 = std::max(rb.tl_x, 0);
          by0 = std::max(rb.tl_y, 0);
          bx1 = std::min(rb.br_x, cv->cols - 1);
          by1 = std::min(rb.br_y, cv->rows - 1);
        }
        if (bx1 < bx0 || by1 < by0) continue;
        x0 = std::min(x0, bx0);
        y0 = std::min(y0, by0);
        x1 = std::max(x1, bx1);
        y1 = std::max(y1, by1);
      }
      pnt_grad_s::Seg sg{s0, s1, x0, y0, std::max(x1 - x0 + 1, 0), std::max(y1 - y0 + 1, 0), total};
      // canvas and snapshot boxes as 7 FP64 planes (14 floats per pixel
      // each), the pickup map as 7 FP32 planes; offsets kept 8-byte aligned
      total += round_up((int64_t)sg.w * sg.h * (b->use_snap ? 28 : 14) + 7 * m, 2);
      g->segs.push_back(sg);
      s0 = s1;
    }
    ensure(g->ckpt, g->ckpt_cap, std::max<int64_t>(total, 1), s);
    const Plane7 pick = b->pick_planes();
    for (const auto& sg : g->segs) {
      const int64_t area = (int64_t)sg.w * sg.h;
      float* c0 = g->ckpt + sg.off;
      box_copy({layer_box(L, cv->cols, sg.x0, sg.y0, sg.w, sg.h, reinterpret_cast<double*>(c0), area),
                layer_box(b->snap_planes(), cv->cols, sg.x0, sg.y0, b->use_snap ? sg.w : 0, sg.h,
                          reinterpret_cast<double*>(c0 + 14 * area), area),
                plane_box(pick, b->size_map, 0, 0, b->size_map, b->size_map,
                          c0 + (b->use_snap ? 28 : 14) * area, m)},
               true, s);
      run_fused(b, cv, g->st.data() + sg.s0, sg.s1 - sg.s0);
    }
    launch_compose(n, L, R0c, Rgb3{{R.p[0], R.p[1], R.p[2]}}, s);
    g->have_forward = true;
  });
}

int pnt_grad_backward(pnt_grad g, pnt_dev_rgb gR, pnt_dev_layer gwet, pnt_dev_rgb gR0,
                      double* gW) {
  return guard_call([&] {
    PNT_REQUIRE(g && gW, "NULL argument");
    PNT_REQUIRE(g->have_forward, "backward before forward");
    for (int c = 0; c < 7; ++c) PNT_REQUIRE(gwet.p[c], "NULL gradient plane");
    for (int c = 0; c < 3; ++c) PNT_REQUIRE(gR.p[c] && gR0.p[c], "NULL plane");
    g->ctx->use();
    pnt_brush b = g->b;
    pnt_canvas cv = g->cv;
    hipStream_t s = g->ctx->stream;
    const int64_t n = cv->n, m = b->m;
    const int sm = b->size_map, hr = (sm - 1) / 2;
    const int ns = (int)g->st.size();
    const Layer7 L = cv->layer();
    const Plane7 gC{{gwet.p[0], gwet.p[1], gwet.p[2], gwet.p[3], gwet.p[4], gwet.p[5], gwet.p[6]}};
    // ---- composite adjoint at the final state
    k_km_adj<<<grid_for(n), kBlk, 0, s>>>(n, L, cv->R0(), Rgb3{{gR.p[0], gR.p[1], gR.p[2]}}, gC,
                                          Rgb3{{gR0.p[0], gR0.p[1], gR0.p[2]}});
    PNT_HIP_CHECK(hipGetLastError());
    const int64_t nst = round_up(n, 64);
    if (b->use_snap) {
      ensure(g->gX, g->gX_cap, 7 * nst, s);
      launch_fill(g->gX, 7 * nst, 0.0f, s);
    }
    ensure(g->gP, g->gP_cap, 7 * m, s);
    ensure(g->gB, g->gB_cap, 6 * (int64_t)ns, s);
    launch_fill(g->gP, 7 * m, 0.0f, s);
    launch_fill(g->gB, 6 * (int64_t)ns, 0.0f, s);
    const Plane7 gXL = b->use_snap ? dense7(g->gX, nst) : gC;
    const Plane7 gPL = dense7(g->gP, m);
    // tapes of one segment
    const int64_t bwmax = 2 * (int64_t)hr + 2;
    const int64_t area = bwmax * bwmax;
    // per stamp: C and X as 7 FP64 planes of area (28 floats per pixel),
    // then P as 7 FP32 planes of m
    const int64_t slot = round_up(28 * area + 7 * m, 2);
    ensure(g->tape, g->tape_cap, slot * g->seg_len, s);
    const int side = 2 * hr + 1;
    const int tiles_x = (side + kGT - 1) / kGT;
    const Plane7 pick = b->pick_planes();
    std::vector<float> paints(6 * (size_t)g->seg_len);
    for (int k = (int)g->segs.size() - 1; k >= 0; --k) {
      const auto& sg = g->segs[k];
      const int cnt = sg.s1 - sg.s0;
      const PlanDev D = upload_plan(b, cv, g->st.data() + sg.s0, cnt);
      // scratch of the shared-pixel replay: one record per visit
      int64_t max_list = 0;
      for (int q = 0; q < cnt; ++q)
        if (D.host[q].dl_nlev > 0)
          max_list = std::max<int64_t>(
              max_list, D.host_off[D.host[q].dl_lev + D.host[q].dl_nlev] - D.host_off[D.host[q].dl_lev]);
      if (max_list)
        ensure(g->scratch, g->scratch_cap, max_list * (int64_t)(sizeof(VState) / sizeof(float)), s);
      // ---- restore the segment's checkpoint
      const int64_t carea = (int64_t)sg.w * sg.h;
      float* c0 = g->ckpt + sg.off;
      box_copy({layer_box(L, cv->cols, sg.x0, sg.y0, sg.w, sg.h, reinterpret_cast<double*>(c0), carea),
                layer_box(b->snap_planes(), cv->cols, sg.x0, sg.y0, b->use_snap ? sg.w : 0, sg.h,
                          reinterpret_cast<double*>(c0 + 14 * carea), carea),
                plane_box(pick, sm, 0, 0, sm, sm, c0 + (b->use_snap ? 28 : 14) * carea, m)},
               false, s);
      {  // brush paint at the segment start: its last dip (stamp 0 dips)
        int i = sg.s0;
        while (!(g->st[i].flags & PNT_STAMP_DIP)) --i;
        for (int c = 0; c < 3; ++c) {
          b->paintK[c] = g->st[i].K[c];
          b->paintS[c] = g->st[i].S[c];
        }
      }
      // ---- recompute the segment, taping each stamp
      std::vector<Tape> tapes(cnt);
      for (int q = 0; q < cnt; ++q) {
        const pnt_stamp& sq = g->st[sg.s0 + q];
        if (sq.flags & PNT_STAMP_DIP) {
          launch_fill(b->pick, 7 * b->mstride, 0.0f, s);
          for (int c = 0; c < 3; ++c) {
            b->paintK[c] = sq.K[c];
            b->paintS[c] = sq.S[c];
          }
        }
        for (int c = 0; c < 3; ++c) {
          paints[6 * q + c] = b->paintK[c];
          paints[6 * q + 3 + c] = b->paintS[c];
        }
        const StampArgs a = stamp_args(b, cv, sq.cx, sq.cy, sq.theta);
        const PStamp& p = D.host[q];
        if (b->use_snap) launch_ring_refresh(p.box, cv->cols, L, b->snap_planes(), s);
        Tape& t = tapes[q];
        float* base = g->tape + slot * q;
        double* tC = reinterpret_cast<double*>(base);
        double* tX = reinterpret_cast<double*>(base + 14 * area);
        float* tP = base + 28 * area;
        t.C = tC;
        t.X = tX;
        t.P = tP;
        t.area = area;
        t.m = m;
        window_box(p, cv->rows, cv->cols, &t.bx0, &t.by0, &t.bw, &t.bh);
        box_copy({layer_box(L, cv->cols, t.bx0, t.by0, t.bw, t.bh, tC, area),
                  layer_box(b->snap_planes(), cv->cols, t.bx0, t.by0, b->use_snap ? t.bw : 0, t.bh,
                            tX, area),
                  plane_box(pick, sm, 0, 0, sm, sm, tP, m)},
                 true, s);
        launch_stamp_main(a, s);
        if (p.dl_nlev > 0) launch_stamp_serial(a, D.dl, D.dl_off + p.dl_lev, p.dl_nlev, s);
      }
      // ---- adjoints, last stamp first
      for (int q = cnt - 1; q >= 0; --q) {
        const int i = sg.s0 + q;
        const PStamp& p = D.host[q];
        AArgs A;
        A.st = D.st;
        A.j = q;
        A.hr = hr;
        A.size_map = sm;
        A.rows = cv->rows;
        A.cols = cv->cols;
        A.tiles_x = tiles_x;
        A.ntiles = tiles_x * tiles_x;
        A.use_snap = b->use_snap ? 1 : 0;
        A.fp = b->fp;
        A.t = tapes[q];
        A.gC = gC;
        A.gX = gXL;
        A.gP = gPL;
        A.gB = g->gB + 6 * (int64_t)i;
        A.pr = b->pickup_rate;
        A.dr = b->dep_rate;
        A.cap = b->cap;
        for (int c = 0; c < 3; ++c) {
          A.pk[c] = paints[6 * q + c];
          A.ps[c] = paints[6 * q + 3 + c];
        }
        A.dl = D.dl;
        A.dl_off = D.dl_off;
        A.scratch = g->scratch;
        if (A.t.bw > 0 && A.t.bh > 0) {
          k_adj_visits<<<(unsigned)A.ntiles, kGT * kGT, 0, s>>>(A);
          PNT_HIP_CHECK(hipGetLastError());
          if (p.dl_nlev > 0) {
            k_adj_shared<<<1, kBlk, 0, s>>>(A);
            PNT_HIP_CHECK(hipGetLastError());
          }
        }
        if (b->use_snap && p.box.a_x1 >= p.box.a_x0 && p.box.a_y1 >= p.box.a_y0) {
          const int64_t tot = (int64_t)(p.box.a_x1 - p.box.a_x0 + 1) * (p.box.a_y1 - p.box.a_y0 + 1);
          k_adj_ring<<<grid_for(tot), kBlk, 0, s>>>(p.box, cv->cols, gC, gXL);
          PNT_HIP_CHECK(hipGetLastError());
        }
        if (g->st[i].flags & PNT_STAMP_DIP) launch_fill(g->gP, 7 * m, 0.0f, s);
      }
      PNT_HIP_CHECK(hipStreamSynchronize(s));  // tapes and the plan are reused next segment
    }
    if (b->use_snap) {
      k_adj_full_copy<<<grid_for(n), kBlk, 0, s>>>(n, gC, gXL);
      PNT_HIP_CHECK(hipGetLastError());
    }
    // the recompute ran kernels that do not keep the snapshot dirty bytes
    cv->epoch++;
    b->dirty_canvas = nullptr;
    // ---- mixing weights: dB/dW of mixSinglePaint (PaintMixer.cxx:327-355)
    std::vector<float> hB(6 * (size_t)ns);
    PNT_HIP_CHECK(hipMemcpyAsync(hB.data(), g->gB, hB.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    std::fill(gW, gW + (int64_t)g->n_paint * g->n_pig, 0.0);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = ns - 1; i >= 0; --i) {
      for (int c = 0; c < 6; ++c) acc[c] += hB[6 * (size_t)i + c];
      if (!(g->st[i].flags & PNT_STAMP_DIP)) continue;
      const double* w = g->W.data() + (int64_t)g->paint_of[i] * g->n_pig;
      double wsum = 0.0;
      for (int l = 0; l < g->n_pig; ++l) wsum += w[l];
      const bool unit = std::fabs(wsum - 1.0) < 10e-6;
      const double norm = unit ? 1.0 : 1.0 / wsum;
      double Bm[6] = {0, 0, 0, 0, 0, 0};
      for (int l = 0; l < g->n_pig; ++l)
        for (int c = 0; c < 3; ++c) {
          Bm[c] += norm * w[l] * g->palK[3 * l + c];
          Bm[3 + c] += norm * w[l] * g->palS[3 * l + c];
        }
      for (int l = 0; l < g->n_pig; ++l) {
        double v = 0.0;
        for (int c = 0; c < 3; ++c) {
          const double dK = unit ? g->palK[3 * l + c] : norm * (g->palK[3 * l + c] - Bm[c]);
          const double dS = unit ? g->palS[3 * l + c] : norm * (g->palS[3 * l + c] - Bm[3 + c]);
          v += acc[c] * dK + acc[3 + c] * dS;
        }
        gW[(int64_t)g->paint_of[i] * g->n_pig + l] += v;
      }
      for (double& a : acc) a = 0.0;
    }
  });
}

}  // extern "C"
