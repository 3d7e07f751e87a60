// The arithmetic of one imprint visit, shared by every stamp kernel.
//
// pickupPaint (renderer/FootprintBrush.hxx:349-384), then depositPaint
// (:393-431), with blend() (:331-340) as blendf below.  Storage is reached
// through an accessor so the plain kernel (pickup map in global memory) and
// the persistent kernel (pickup map in LDS) execute the same instructions and
// produce the same bits.  Products that the reference rounds separately are
// combined only through explicit fmaf, never by compiler contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace pnt {

// MinVolume = 0.001 (FootprintBrush.hxx:26) compared in FP64; for a float
// x, x > 0.001 holds exactly when x >= 0.001f.
constexpr float kMinVolumeF = 0.001f;

// (va*a + vb*b) / vt, with inv_vt = 1/vt shared by the channels of one blend
// (v_rcp_f32, 1 ulp; the product adds at most 1.5 ulp over a divide).
__device__ __forceinline__ float blendf(float va, float a, float vb, float b, float inv_vt) {
  return __builtin_fmaf(va, a, vb * b) * inv_vt;
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// Accessor interface:
//   float pick(int c) / void set_pick(int c, float)   c = 0..2 K, 3..5 S, 6 V
//   float src(int c)  / void set_src_v(float)          pickup source (snapshot or canvas)
//   float canvas(int c) / void set_canvas(int c, float)
template <class Acc>
__device__ __forceinline__ void paint_visit_math(Acc& a, float fh, float pickup_rate,
                                                 float dep_rate, float cap, const float pK[3],
                                                 const float pS[3]) {
  {  // pickupPaint
    const float vP = a.pick(6);
    const float vC = a.src(6);
    const float vLeave = (pickup_rate * vC) * fh;
    if (vLeave >= kMinVolumeF) {
      a.set_src_v(vC - vLeave);
      const float vt = vP + vLeave;
      if (vt >= kMinVolumeF) {
        const float iv = rcpf(vt);
#pragma unroll
        for (int c = 0; c < 6; ++c) a.set_pick(c, blendf(vP, a.pick(c), vLeave, a.src(c), iv));
      }
      a.set_pick(6, vP + vLeave);
    }
  }
  {  // depositPaint
    const float vP = a.pick(6);
    const float vFree = fmaxf(0.0f, cap - vP);
    float src[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) src[c] = a.pick(c);
    const float vt = vP + vFree;
    if (vt >= kMinVolumeF) {
      const float iv = rcpf(vt);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        src[c] = blendf(vP, src[c], vFree, pK[c], iv);
        src[3 + c] = blendf(vP, src[3 + c], vFree, pS[c], iv);
      }
    }
    a.set_pick(6, __builtin_fmaf(-(dep_rate * vP), fh, vP));
    const float vC = a.canvas(6);
    const float vB = cap * fh;
    const float vt2 = vB + vC;
    if (vt2 >= kMinVolumeF) {
      const float iv = rcpf(vt2);
#pragma unroll
      for (int c = 0; c < 6; ++c) a.set_canvas(c, blendf(vB, src[c], vC, a.canvas(c), iv));
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.set_canvas(c, src[c]);
    }
    a.set_canvas(6, vB + vC);
  }
}

}  // namespace pnt
