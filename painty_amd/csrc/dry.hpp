// Canvas::checkDry (Canvas.hxx:123-155) on the device, shared by the stamp
// and texture-brush kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "km.hpp"

namespace pnt {

// Canvas::checkDry (Canvas.hxx:123-155) of a visited pixel, before its
// pickup and deposit (FootprintBrush.hxx:115, TextureBrush.hxx:164): paint
// older than the drying time dries into the substrate as dryCanvas does; the
// pixel's time becomes the imprint's.  V is FP64, so V > 0.001 is the
// reference's own test.  The reference's partial branch needs the integer
// quotient elapsed / drying time above 0.01, which is 0 whenever elapsed <
// drying time, so it never runs.  h accumulates in FP64 and is stored
// rounded.
// The height after `v` of paint dries onto h (FP64 sum, stored rounded).
__device__ __forceinline__ float h_plus(float h, double v) { return (float)((double)h + v); }

__device__ __forceinline__ void dry_pixel_into(const DryClock& c, float K[3], float S[3],
                                               double v, int64_t ic) {
  float R[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) R[ch] = c.R0.p[ch][ic];
  c.h[ic] = h_plus(c.h[ic], v);
  km_pixel(K, S, R, (float)v);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) c.R0.p[ch][ic] = R[ch];
}

__device__ __forceinline__ void check_dry(const DryClock& c, const Layer7& canvas, int64_t ic) {
  const double v = canvas.v[ic];
  if (c.dry_ms > 0 && v > 0.001) {
    const int64_t elapsed_ms = (c.now_us - c.tmap[ic]) / 1000;  // duration_cast truncation
    if (elapsed_ms >= c.dry_ms) {
      float K[3], S[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        K[ch] = canvas.ks[ch][ic];
        S[ch] = canvas.ks[3 + ch][ic];
      }
      dry_pixel_into(c, K, S, v, ic);
#pragma unroll
      for (int ch = 0; ch < 6; ++ch) canvas.ks[ch][ic] = 0.0f;
      canvas.v[ic] = 0.0;
    }
  }
  c.tmap[ic] = c.now_us;
}

}  // namespace pnt
