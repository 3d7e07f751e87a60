// The arithmetic of one imprint visit, shared by every stamp kernel.
//
// pickupPaint (renderer/FootprintBrush.hxx:349-384), then depositPaint
// (:393-431), with blend() (:331-340) as blendd below.  Storage is reached
// through an accessor so the plain kernel (pickup map in global memory) and
// the persistent kernel (pickup map in LDS) execute the same instructions and
// produce the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace pnt {

// MinVolume = 0.001 (FootprintBrush.hxx:26).
constexpr double kMinVolume = 0.001;

// The arithmetic runs in FP64, as the reference's does (vec3 is Eigen's
// double vector): loads widen the stored FP32 values, every product, sum,
// quotient and threshold test is the reference's own in double precision,
// and each K/S result is rounded to FP32 once, when it is stored (canvas
// and snapshot volumes are stored in FP64).  Over the
// hundreds of dependent visits a pixel and a pickup cell take in a long
// stamp list, only that storage rounding accumulates (an FP32 chain would
// add the rounding of every intermediate step as well).
//
// blend(va, a, vb, b) = vt > MinVolume ? (va*a + vb*b) / vt : a with
// vt = va + vb (FootprintBrush.hxx:331-340); inv_vt = 1/vt is shared by the
// channels of one blend.
__device__ __forceinline__ double blendd(double va, double a, double vb, double b, double inv_vt) {
  return (va * a + vb * b) * inv_vt;
}

// Kept for the FP32 adjoint (grad.hip): the FP32 form of blend.
__device__ __forceinline__ float blendf(float va, float a, float vb, float b, float inv_vt) {
  return __builtin_fmaf(va, a, vb * b) * inv_vt;
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// Accessor interface:
//   float pick(int c) / void set_pick(int c, float)   c = 0..2 K, 3..5 S, 6 V
//   float src(int c) (c < 6) / double src_v() / void set_src_v(double)
//                                                      pickup source (snapshot or canvas)
//   float canvas(int c) (c < 6) / double canvas_v()
//   void set_canvas(int c, float) (c < 6) / void set_canvas_v(double)
// Canvas and source volumes are FP64 in memory, so every volume that a
// MinVolume test sees is the reference's own number; contraction is off so
// that each product and sum rounds as the reference's does.
template <class Acc>
__device__ __forceinline__ void paint_visit_math(Acc& a, float fh_f, double pickup_rate,
                                                 double dep_rate, double cap, const float pK[3],
                                                 const float pS[3]) {
#pragma clang fp contract(off)
  const double fh = fh_f;
  {  // pickupPaint (FootprintBrush.hxx:349-384)
    const double vP = a.pick(6);
    const double vC = a.src_v();
    const double vLeave = (pickup_rate * vC) * fh;
    if (vLeave > kMinVolume) {
      a.set_src_v(vC - vLeave);
      const double vt = vP + vLeave;
      if (vt > kMinVolume) {
        const double iv = 1.0 / vt;
        double nv[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) nv[c] = blendd(vP, a.pick(c), vLeave, a.src(c), iv);
#pragma unroll
        for (int c = 0; c < 6; ++c) a.set_pick(c, (float)nv[c]);
      }
      a.set_pick(6, (float)(vP + vLeave));
    }
  }
  {  // depositPaint (FootprintBrush.hxx:393-431)
    const double vP = a.pick(6);
    const double vFree = fmax(0.0, cap - vP);
    double src[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) src[c] = a.pick(c);
    const double vt = vP + vFree;
    if (vt > kMinVolume) {
      const double iv = 1.0 / vt;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        src[c] = blendd(vP, src[c], vFree, (double)pK[c], iv);
        src[3 + c] = blendd(vP, src[3 + c], vFree, (double)pS[c], iv);
      }
    }
    a.set_pick(6, (float)(vP - (dep_rate * vP) * fh));
    const double vC = a.canvas_v();
    const double vB = cap * fh;
    const double vt2 = vB + vC;
    if (vt2 > kMinVolume) {
      const double iv = 1.0 / vt2;
      double nv[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) nv[c] = blendd(vB, src[c], vC, a.canvas(c), iv);
#pragma unroll
      for (int c = 0; c < 6; ++c) a.set_canvas(c, (float)nv[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.set_canvas(c, (float)src[c]);
    }
    a.set_canvas_v(vB + vC);
  }
}

}  // namespace pnt
