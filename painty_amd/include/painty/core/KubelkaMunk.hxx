// ComputeReflectance -- drop-in for core/KubelkaMunk.hxx:25-83, evaluated by
// the device kernel through the C ABI (for whole images use
// Renderer::compose; this per-pixel form exists for callers of the free
// function).
#pragma once

#include <stdexcept>
#include <vector>

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

// The same over many pixels in one device call: R[i] = ComputeReflectance(
// K[i], S[i], R0[i], d[i]), equal to the single-pixel form bit for bit.  The
// single-pixel form is a device round trip per call; a caller that walks an
// image pixel by pixel gathers its pixels and calls this once.
template <class Float, std::size_t N>
std::vector<vec<Float, N>> ComputeReflectance(const std::vector<vec<Float, N>>& K,
                                              const std::vector<vec<Float, N>>& S,
                                              const std::vector<vec<Float, N>>& R0,
                                              const std::vector<Float>& d) {
  static_assert(N == 3, "the device path is RGB (3 channels)");
  const std::size_t n = K.size();
  if (S.size() != n || R0.size() != n || d.size() != n)
    throw std::invalid_argument("ComputeReflectance: K, S, R0 and d differ in length");
  std::vector<float> k(3 * n), s(3 * n), r0(3 * n), df(n), r(3 * n);
  for (std::size_t i = 0; i < n; ++i) {
    for (std::size_t c = 0; c < 3; ++c) {
      k[3 * i + c] = float(K[i][c]);
      s[3 * i + c] = float(S[i][c]);
      r0[3 * i + c] = float(R0[i][c]);
    }
    df[i] = static_cast<float>(d[i]);
  }
  std::vector<vec<Float, N>> R(n);
  if (n == 0) return R;
  gpu::check(pnt_compute_reflectance(gpu::Device::get(), (int64_t)n, k.data(), s.data(),
                                     r0.data(), df.data(), r.data()));
  for (std::size_t i = 0; i < n; ++i) R[i] = vec<Float, N>(r[3 * i], r[3 * i + 1], r[3 * i + 2]);
  return R;
}

}  // namespace painty
