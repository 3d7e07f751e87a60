// Palette -- drop-in for mixer/Palette.hxx:18-20.
#pragma once

#include <vector>

#include "painty_hip.h"
#include "painty/mixer/PaintCoeff.hxx"

namespace painty {

using Palette = std::vector<PaintCoeff>;

// mixer/src/Palette.cxx:453-463: the entry of data/paint_palettes/
// thinning_medium.json.
inline auto getThinningMedium() -> PaintCoeff {
  double K[3], S[3];
  pnt_thinning_medium(K, S);
  PaintCoeff p;
  for (int c = 0; c < 3; ++c) {
    p.K[c] = K[c];
    p.S[c] = S[c];
  }
  return p;
}

}  // namespace painty
