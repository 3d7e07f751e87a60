// Host image of the drop-in interface.
//
// The reference's `Mat<T>` is `cv::Mat_<T>` (image/Mat.hxx:41): row-major,
// continuous, with `rows`, `cols`, `(i, j)` and linear `(k)` access.  This
// class provides exactly those members over a std::vector.
#pragma once

#include <cstdint>
#include <vector>

namespace painty {

template <class T>
class Mat {
 public:
  int32_t rows = 0, cols = 0;

  Mat() = default;
  Mat(int32_t r, int32_t c) : rows(r), cols(c), _d(static_cast<std::size_t>(r) * c) {}
  Mat(int32_t r, int32_t c, const T& v) : rows(r), cols(c), _d(static_cast<std::size_t>(r) * c, v) {}

  T& operator()(int32_t i, int32_t j) { return _d[static_cast<std::size_t>(i) * cols + j]; }
  const T& operator()(int32_t i, int32_t j) const {
    return _d[static_cast<std::size_t>(i) * cols + j];
  }
  T& operator()(int64_t k) { return _d[static_cast<std::size_t>(k)]; }
  const T& operator()(int64_t k) const { return _d[static_cast<std::size_t>(k)]; }

  std::size_t total() const { return _d.size(); }
  bool empty() const { return _d.empty(); }
  T* data() { return _d.data(); }
  const T* data() const { return _d.data(); }
  typename std::vector<T>::iterator begin() { return _d.begin(); }
  typename std::vector<T>::iterator end() { return _d.end(); }
  typename std::vector<T>::const_iterator begin() const { return _d.begin(); }
  typename std::vector<T>::const_iterator end() const { return _d.end(); }
  Mat clone() const { return *this; }

 private:
  std::vector<T> _d;
};

}  // namespace painty
