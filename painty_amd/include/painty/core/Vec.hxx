// Fixed-size vectors of the drop-in interface.
//
// The reference's `vec<T, N>` is an Eigen column vector (core/Vec.hxx:19-40);
// the only operations its hot-path interface needs from callers are
// construction from a brace list and element access, which this provides.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>

namespace painty {

template <class T, std::size_t N>
struct vec {
  std::array<T, N> v{};
  vec() = default;
  template <class... A>
  vec(A... a) : v{{static_cast<T>(a)...}} {
    static_assert(sizeof...(A) == N, "vec: wrong number of components");
  }
  T& operator[](std::size_t i) { return v[i]; }
  const T& operator[](std::size_t i) const { return v[i]; }
  T* data() { return v.data(); }
  const T* data() const { return v.data(); }
  static constexpr std::size_t size() { return N; }
};

using vec2 = vec<double, 2>;
using vec3 = vec<double, 3>;
using vec3f = vec<float, 3>;

// Channel type and dimension of a pixel type (core/Types.hxx DataType).
template <class V>
struct DataType;
template <class T, std::size_t N>
struct DataType<vec<T, N>> {
  using channel_type = T;
  static constexpr std::size_t dim = N;
};

}  // namespace painty
