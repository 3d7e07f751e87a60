// Palette -- drop-in for mixer/Palette.hxx:18.
#pragma once

#include <vector>

#include "painty/mixer/PaintCoeff.hxx"

namespace painty {

using Palette = std::vector<PaintCoeff>;

}  // namespace painty
