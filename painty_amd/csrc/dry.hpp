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
// pixel's time becomes the imprint's.  V > 0.001 in FP64 is V >= 0.001f for a float V (0.001f is
// the first float above 0.001).  The reference's partial branch needs the
// integer quotient elapsed / drying time above 0.01, which is 0 whenever
// elapsed < drying time, so it never runs.
__device__ __forceinline__ void check_dry(const DryClock& c, const Layer7& canvas, int64_t ic) {
  const float v = canvas.p[6][ic];
  if (c.dry_ms > 0 && v >= 0.001f) {
    const int64_t elapsed_ms = (c.now_us - c.tmap[ic]) / 1000;  // duration_cast truncation
    if (elapsed_ms >= c.dry_ms) {
      float K[3], S[3], R[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        K[ch] = canvas.p[ch][ic];
        S[ch] = canvas.p[3 + ch][ic];
        R[ch] = c.R0.p[ch][ic];
      }
      c.h[ic] += v;
      km_pixel(K, S, R, v);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        c.R0.p[ch][ic] = R[ch];
        canvas.p[ch][ic] = 0.0f;
        canvas.p[3 + ch][ic] = 0.0f;
      }
      canvas.p[6][ic] = 0.0f;
    }
  }
  c.tmap[ic] = c.now_us;
}

}  // namespace pnt
