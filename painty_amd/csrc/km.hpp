// Kubelka-Munk reflectance in FP32 for gfx950.
//
// Restates core/KubelkaMunk.hxx:25-83 (ComputeReflectance) and
// core/Math.hxx:157-169 (coth) in a form that is stable in FP32 and costs
// two reciprocals, one square root and one exp2 per channel.
//
// Reference, per channel (S clamped to 1e-11 when |S| <= 1e4*DBL_EPSILON):
//   a = 1 + K/S,  b = sqrt(max(a^2-1, 0)),  x = b*S*d
//   R = (1 - R0*(a - b*coth x)) / (a - R0 + b*coth x)
// Multiplying numerator and denominator by S, with A = S*a = S+K and
// B = S*b = sqrt(K*(K+2S)), x = B*d, E = expm1(2x), coth x = 1 + 2/E, and
// a - b = 1/(a+b), so that S*(a - b*coth x) = S - S*(K+B)/(A+B) - 2B/E:
//   numerator   S*(1 - R0*(a - b coth x)) = S*(1-R0) + R0*(S*(K+B)/(A+B) + 2B/E)
//   denominator S*(a - R0 + b coth x)     = S*(1-R0) + K + B + 2B/E
// For 0 <= R0 <= 1 and K, S >= 0 every term is non-negative: no subtraction
// of nearly equal terms anywhere (1 - R0 is exact for R0 >= 0.5), so the
// result stays within a few ulp even where the reference's own form cancels
// (R0 near 1 under a weakly absorbing, thick layer).  The reference returns
// coth = 1 for x > 20; here 2x is clamped to 40, where 2/E = 8.5e-18
// relative to 1, i.e. the same number in FP32.
//
// Thresholds: 1e4*DBL_EPSILON = 10000*2^-52 = 0x1.388p-39 is exact in FP32,
// so |d| < thr and |S| > thr compare exactly as the FP64 reference does.
#pragma once

#include <hip/hip_runtime.h>

namespace pnt {

constexpr float kKmThr = 0x1.388p-39f;   // numeric_limits<double>::epsilon()*1e4
constexpr float kSClamp = 1e-11f;        // KubelkaMunk.hxx:41

// expm1(y) for y in [0, 40].  Below 0.5: Taylor series of e^y - 1 to y^10
// (truncation < 5e-10 relative).  Above: 2^(y*log2 e) with the product split
// into hi + lo (fma) so the exponent carries ~48 bits, then -1 (y >= 0.5 keeps
// the subtraction within 1.3 bits).
__device__ __forceinline__ float km_expm1(float y) {
#pragma clang fp contract(off)
  // Taylor: y*(1 + y/2*(1 + y/3*(1 + ... )))
  float p = 1.0f + y * (1.0f / 10.0f);
  p = 1.0f + y * (1.0f / 9.0f) * p;
  p = 1.0f + y * (1.0f / 8.0f) * p;
  p = 1.0f + y * (1.0f / 7.0f) * p;
  p = 1.0f + y * (1.0f / 6.0f) * p;
  p = 1.0f + y * (1.0f / 5.0f) * p;
  p = 1.0f + y * (1.0f / 4.0f) * p;
  p = 1.0f + y * (1.0f / 3.0f) * p;
  p = 1.0f + y * 0.5f * p;
  const float small = y * p;
  constexpr float kL2eHi = 1.44269502162933349609375f;   // float(log2 e)
  constexpr float kL2eLo = 1.925963033500011079013347e-8f;  // log2 e - hi
  constexpr float kLn2 = 0.693147180559945309f;
  const float t = y * kL2eHi;
  const float err = __builtin_fmaf(y, kL2eHi, -t) + y * kL2eLo;
  const float e = __builtin_amdgcn_exp2f(t);
  const float big = __builtin_fmaf(e, err * kLn2, e) - 1.0f;
  return (y < 0.5f) ? small : big;
}

// One channel, d already known to satisfy |d| >= kKmThr.  Contraction is
// off and every fused multiply-add is written out, so each kernel that
// inlines this computes the same bits (composite, dry, checkDry, waves).
__device__ __forceinline__ float km_channel(float K, float S, float R0, float d) {
#pragma clang fp contract(off)
  S = (__builtin_fabsf(S) > kKmThr) ? S : kSClamp;
  const float q = K * __builtin_fmaf(2.0f, S, K);
  const float B = __builtin_amdgcn_sqrtf(q > 0.0f ? q : 0.0f);
  float y = 2.0f * B * d;
  y = y < 40.0f ? y : 40.0f;
  const float E = km_expm1(y);
  const float KB = K + B;
  const float AB = S + KB;
  const float r = __builtin_amdgcn_rcpf(E * AB);  // 1/(E (A+B))
  const float t = 2.0f * B * (AB * r);            // 2B/E
  const float w = S * (1.0f - R0);
  const float num = __builtin_fmaf(R0, __builtin_fmaf(S * KB, E * r, t), w);
  const float den = w + KB + t;
  return num * __builtin_amdgcn_rcpf(den);
}

// ComputeReflectance for one pixel (3 channels, shared thickness d).
__device__ __forceinline__ void km_pixel(const float K[3], const float S[3], float R0[3],
                                         float d) {
  if (__builtin_fabsf(d) < kKmThr) return;  // KubelkaMunk.hxx:31-34
#pragma unroll
  for (int c = 0; c < 3; ++c) R0[c] = km_channel(K[c], S[c], R0[c], d);
}

// The GLSL shaders' variant (renderer/shaders/DryingStep.compute.glsl:31-49):
// no S clamp; r0 returned when any channel has |b S d| < 1e-5.
__device__ __forceinline__ void km_pixel_glsl(const float K[3], const float S[3], float R0[3],
                                              float d) {
#pragma clang fp contract(off)
  bool small = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float q = K[c] * (K[c] + 2.0f * S[c]);
    const float B = __builtin_amdgcn_sqrtf(q > 0.0f ? q : 0.0f);
    small |= __builtin_fabsf(B * d) < 10e-6f;
  }
  if (small) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float s = S[c];
    const float q = K[c] * __builtin_fmaf(2.0f, s, K[c]);
    const float B = __builtin_amdgcn_sqrtf(q > 0.0f ? q : 0.0f);
    float y = 2.0f * B * d;
    y = y < 40.0f ? y : 40.0f;
    const float E = km_expm1(y);
    const float KB = K[c] + B;
    const float AB = s + KB;
    const float r = __builtin_amdgcn_rcpf(E * AB);
    const float t = 2.0f * B * (AB * r);
    const float w = s * (1.0f - R0[c]);
    const float num = __builtin_fmaf(R0[c], __builtin_fmaf(s * KB, E * r, t), w);
    R0[c] = num * __builtin_amdgcn_rcpf(w + KB + t);
  }
}

}  // namespace pnt
