// ComputeReflectance -- drop-in for core/KubelkaMunk.hxx:25-83, evaluated by
// the device kernel through the C ABI (for whole images use
// Renderer::compose; this per-pixel form exists for callers of the free
// function).
#pragma once

#include "painty/core/Vec.hxx"
#include "painty/gpu/Device.hxx"

namespace painty {

template <class Float, std::size_t N>
vec<Float, N> ComputeReflectance(const vec<Float, N>& K, const vec<Float, N>& S,
                                 const vec<Float, N>& R0, const Float d) {
  static_assert(N == 3, "the device path is RGB (3 channels)");
  const float k[3] = {float(K[0]), float(K[1]), float(K[2])};
  const float s[3] = {float(S[0]), float(S[1]), float(S[2])};
  const float r0[3] = {float(R0[0]), float(R0[1]), float(R0[2])};
  const float df = static_cast<float>(d);
  float r[3];
  gpu::check(pnt_compute_reflectance(gpu::Device::get(), 1, k, s, r0, &df, r));
  return vec<Float, N>(r[0], r[1], r[2]);
}

}  // namespace painty
