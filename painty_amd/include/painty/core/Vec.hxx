// Fixed-size vectors of the drop-in interface.
//
// The reference's `vec<T, N>` is an Eigen column vector (core/Vec.hxx:19-40).
// Its callers on this path construct vectors from brace lists, index them,
// and do the component-wise arithmetic of the spline helpers and the GUI's
// stroke driver (sums, differences, scaling, norm(), normalized()), which
// this provides.
#pragma once

#include <array>
#include <cmath>
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

  vec& operator+=(const vec& o) {
    for (std::size_t i = 0; i < N; ++i) v[i] += o.v[i];
    return *this;
  }
  vec& operator-=(const vec& o) {
    for (std::size_t i = 0; i < N; ++i) v[i] -= o.v[i];
    return *this;
  }
  vec& operator*=(T s) {
    for (std::size_t i = 0; i < N; ++i) v[i] *= s;
    return *this;
  }
  vec& operator/=(T s) {
    for (std::size_t i = 0; i < N; ++i) v[i] /= s;
    return *this;
  }
  friend vec operator+(vec a, const vec& b) { return a += b; }
  friend vec operator-(vec a, const vec& b) { return a -= b; }
  friend vec operator-(vec a) {
    for (std::size_t i = 0; i < N; ++i) a.v[i] = -a.v[i];
    return a;
  }
  friend vec operator*(vec a, T s) { return a *= s; }
  friend vec operator*(T s, vec a) {
    for (std::size_t i = 0; i < N; ++i) a.v[i] = s * a.v[i];
    return a;
  }
  friend vec operator/(vec a, T s) { return a /= s; }
  friend bool operator==(const vec& a, const vec& b) { return a.v == b.v; }
  friend bool operator!=(const vec& a, const vec& b) { return a.v != b.v; }
  T dot(const vec& o) const {
    T s = 0;
    for (std::size_t i = 0; i < N; ++i) s += v[i] * o.v[i];
    return s;
  }
  T squaredNorm() const { return dot(*this); }
  T norm() const { return std::sqrt(squaredNorm()); }
  // Eigen's normalized(): unchanged when the norm is zero
  vec normalized() const {
    const T n2 = squaredNorm();
    if (!(n2 > T(0))) return *this;
    const T n = std::sqrt(n2);
    vec r = *this;
    for (std::size_t i = 0; i < N; ++i) r.v[i] /= n;
    return r;
  }
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
template <>
struct DataType<double> {
  using channel_type = double;
  static constexpr std::size_t dim = 1;
};
template <>
struct DataType<float> {
  using channel_type = float;
  static constexpr std::size_t dim = 1;
};

}  // namespace painty
