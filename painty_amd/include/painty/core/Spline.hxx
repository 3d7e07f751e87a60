// Spline helpers -- drop-in for core/Spline.hxx:15-221: linear, Catmull-Rom
// (tau = 1/2) and cubic interpolation between control points, and SplineEval
// over a point sequence, where u in [0, 1] spans the n - 1 segments and the
// control indices are clamped to the ends.  Host code: the reference's
// stroke drivers (FootprintBrush::paintStroke, the GUI's mouse handler,
// apps/painty_gui/DigitalCanvas.cxx:98-125, 128-152) place stamps and
// texture strokes with them.  The polynomial forms are the reference's, so
// positions and headings agree to the bit.
//
// Those forms follow painty's core/Spline.hxx, which carries this notice:
//
//   MIT License
//
//   Copyright (c) 2020 Thomas Lindemeier
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
#pragma once

#include <cmath>
#include <cstdint>
#include <iterator>

#include "painty/core/Vec.hxx"

namespace painty {

template <class T>
inline T LinearDerivative(const T& p0, const T& p1) {
  return p1 - p0;
}

template <class T>
inline T Linear(const T& p0, const T& p1, const typename DataType<T>::channel_type t) {
  using F = typename DataType<T>::channel_type;
  return p0 * (F(1) - t) + p1 * t;
}

// Spline.hxx:27-47: the Catmull-Rom basis with tension tau = 1/2.
template <class T>
inline T CatmullRom(const T& p_1, const T& p0, const T& p1, const T& p2,
                    typename DataType<T>::channel_type t) {
  using F = typename DataType<T>::channel_type;
  const F tau = F(0.5), t2 = t * t, t3 = t2 * t;
  const F b1 = -tau * t + F(2) * tau * t2 - tau * t3;
  const F b2 = F(1) + (tau - F(3)) * t2 + (F(2) - tau) * t3;
  const F b3 = tau * t + (F(3) - F(2) * tau) * t2 + (tau - F(2)) * t3;
  const F b4 = -tau * t2 + tau * t3;
  return p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4;
}

// Spline.hxx:49-73: its derivative in t.
template <class T>
inline T CatmullRomDerivativeFirst(const T& p_1, const T& p0, const T& p1, const T& p2,
                                   typename DataType<T>::channel_type t) {
  using F = typename DataType<T>::channel_type;
  const F tau = F(0.5), t2 = t * t;
  const F b1 = tau * (F(-3) * t2 + F(4) * t - F(1));
  const F b2 = -t * (F(-2) * tau + F(3) * (tau - F(2)) * t + F(6));
  const F b3 = (t - F(1)) * (F(3) * (tau - F(2)) * t - tau);
  const F b4 = tau * t * (F(3) * t - F(2));
  return p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4;
}

// Spline.hxx:75-93: the cubic through p0 (t = 0) and p1 (t = 1).
template <class T>
inline T Cubic(const T& p_1, const T& p0, const T& p1, const T& p2,
               typename DataType<T>::channel_type t) {
  const auto t2 = t * t;
  const T a0 = p2 - p1 - p_1 + p0;
  const T a1 = p_1 - p0 - a0;
  const T a2 = p1 - p_1;
  const T a3 = p0;
  return a0 * t * t2 + a1 * t2 + a2 * t + a3;
}

template <class T>
inline T CubicDerivativeFirst(const T& p_1, const T& p0, const T& p1, const T& p2,
                              typename DataType<T>::channel_type t) {
  using F = typename DataType<T>::channel_type;
  const F t2 = t * t;
  const T a0 = p2 - p1 - p_1 + p0;
  const T a1 = p_1 - p0 - a0;
  const T a2 = p1 - p_1;
  return F(3) * a0 * t2 + F(2) * a1 * t + a2;
}

template <class T>
inline T CubicDerivativeSecond(const T& p_1, const T& p0, const T& p1, const T& p2,
                               typename DataType<T>::channel_type t) {
  using F = typename DataType<T>::channel_type;
  const T a0 = p2 - p1 - p_1 + p0;
  const T a1 = p_1 - p0 - a0;
  return F(2) * (F(3) * a0 * t + a1);
}

// Spline.hxx:127-221
template <class ContainerIt>
class SplineEval {
  using value_type = typename std::iterator_traits<ContainerIt>::value_type;
  using value = typename DataType<value_type>::channel_type;

 public:
  SplineEval(ContainerIt begin, ContainerIt end) : _begin(begin), _end(end) {}

  value_type linear(value u) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    return Linear(at(i), at(i + 1), t);
  }
  value_type linearDerivative(value u) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    return LinearDerivative(at(i), at(i + 1));
  }
  value_type cubic(value u) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    return Cubic(at(i - 1), at(i), at(i + 1), at(i + 2), t);
  }
  value_type cubicDerivative(value u, int32_t order) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    if (order == 2) return CubicDerivativeSecond(at(i - 1), at(i), at(i + 1), at(i + 2), t);
    return CubicDerivativeFirst(at(i - 1), at(i), at(i + 1), at(i + 2), t);
  }
  value_type catmullRom(value u) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    return CatmullRom(at(i - 1), at(i), at(i + 1), at(i + 2), t);
  }
  value_type catmullRomDerivativeFirst(value u) const {
    int32_t i = 0;
    value t;
    control(u, i, t);
    return CatmullRomDerivativeFirst(at(i - 1), at(i), at(i + 1), at(i + 2), t);
  }

 private:
  void control(value u, int32_t& index, value& t) const {
    const auto n = std::distance(_begin, _end);
    const value x = static_cast<value>(n - 1) * u;
    index = static_cast<int32_t>(x);
    t = x - std::floor(x);
  }
  const value_type& at(int32_t index) const {
    const auto n = std::distance(_begin, _end);
    if (index < 0) return *_begin;
    if (index >= static_cast<int32_t>(n)) return *(_end - 1);
    return *(_begin + index);
  }

  ContainerIt _begin, _end;
};

}  // namespace painty
