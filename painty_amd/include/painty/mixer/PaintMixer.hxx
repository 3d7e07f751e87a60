// PaintMixer -- drop-in for the mixing members of mixer/PaintMixer.hxx:22-57
// (mixSinglePaint, mixed, get/setUnderlyingPalette) through the C ABI.  The
// Ceres-fitted members (mixFromInputPicture, getWeightsForMixingTargetPaint,
// getMixtureWeightsForReflectance, mixClosestFit) are not part of this
// library.
#pragma once

#include <vector>

#include "painty/gpu/Device.hxx"
#include "painty/mixer/Palette.hxx"

namespace painty {

class PaintMixer {
 public:
  explicit PaintMixer(const Palette& basePalette) : _basePalette(basePalette) {}

  // PaintMixer.cxx:327-355; std::invalid_argument when the weight count
  // differs from the palette size.
  auto mixSinglePaint(const std::vector<CoeffPrecision>& weights) const -> PaintCoeff {
    std::vector<double> k, s;
    for (const auto& p : _basePalette)
      for (int c = 0; c < 3; ++c) {
        k.push_back(p.K[c]);
        s.push_back(p.S[c]);
      }
    PaintCoeff out;
    gpu::check(pnt_mix_single((int)_basePalette.size(), k.data(), s.data(), (int)weights.size(),
                              weights.data(), out.K.data(), out.S.data()));
    return out;
  }

  auto getUnderlyingPalette() const -> const Palette& { return _basePalette; }
  void setUnderlyingPalette(const Palette& palette) { _basePalette = palette; }

  // PaintMixer.cxx:536-548
  auto mixed(const PaintCoeff& paint, const double paintVolume, const PaintCoeff& other,
             const double otherVolume) -> PaintCoeff {
    PaintCoeff out;
    gpu::check(pnt_mixed(paint.K.data(), paint.S.data(), paintVolume, other.K.data(),
                         other.S.data(), otherVolume, out.K.data(), out.S.data()));
    return out;
  }

 private:
  Palette _basePalette;
};

}  // namespace painty
