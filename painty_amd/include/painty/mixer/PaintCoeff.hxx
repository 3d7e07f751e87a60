// PaintCoeff -- drop-in for mixer/PaintCoeff.hxx:18-36.
#pragma once

#include "painty/core/Vec.hxx"

namespace painty {

constexpr auto CoeffSamplesCount = 3;
using CoeffPrecision = double;

struct PaintCoeff {
  using VecType = vec<CoeffPrecision, CoeffSamplesCount>;
  VecType K;  // absorption
  VecType S;  // scattering
};

}  // namespace painty
