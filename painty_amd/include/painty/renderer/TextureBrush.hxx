// TextureBrush<vector_type> -- drop-in for renderer/TextureBrush.hxx:19-238:
// a brush stroke texture warped along the stroke, smudging, volume-weighted
// mixing, on the device through the C ABI.
//
// The reference builds its BrushStrokeSample from a directory and reads the
// thickness map from sampleDir/thickness_map.png
// (renderer/src/BrushStrokeSample.cxx:20-22, 164).  Image IO is the caller's
// here: the constructor takes the thickness map itself, or a directory and a
// sample loader (setSampleLoader); without a loader the directory form
// throws std::ios_base::failure, as the reference does for an unreadable
// file.
#pragma once

#include <array>
#include <functional>
#include <ios>
#include <string>
#include <vector>

#include "painty/renderer/BrushBase.hxx"

namespace painty {

template <class vector_type>
class TextureBrush final : public BrushBase<vector_type> {
 public:
  using SampleLoader = std::function<Mat<double>(const std::string& sampleDir)>;

  // TextureBrush.hxx:25-31: radius 0, paint 0.1, smudging on.
  explicit TextureBrush(const Mat<double>& thicknessMap) {
    gpu::check(pnt_tbrush_create(gpu::Device::get(), &_h));
    setThicknessMap(thicknessMap);
  }
  explicit TextureBrush(const std::string& sampleDir) : TextureBrush(load(sampleDir)) {}
  TextureBrush(const TextureBrush&) = delete;
  TextureBrush& operator=(const TextureBrush&) = delete;
  ~TextureBrush() override {
    if (_h) pnt_tbrush_destroy(_h);
  }

  static void setSampleLoader(SampleLoader loader) { loader_slot() = std::move(loader); }

  // TextureBrush.hxx:33-41
  void setRadius(const double radius) override { gpu::check(pnt_tbrush_set_radius(_h, radius)); }

  // TextureBrush.hxx:48-50
  void dip(const std::array<vector_type, 2UL>& paint) override {
    const double k[3] = {double(paint[0][0]), double(paint[0][1]), double(paint[0][2])};
    const double s[3] = {double(paint[1][0]), double(paint[1][1]), double(paint[1][2])};
    gpu::check(pnt_tbrush_dip(_h, k, s));
  }

  // TextureBrush.hxx:52-205
  void paintStroke(const std::vector<vec2>& verticesArg, Canvas<vector_type>& canvas) override {
    gpu::check(pnt_tbrush_set_thickness_scale(_h, double(this->getThicknessScale())));
    std::vector<double> xy(2 * verticesArg.size());
    for (std::size_t i = 0; i < verticesArg.size(); ++i) {
      xy[2 * i] = verticesArg[i][0];
      xy[2 * i + 1] = verticesArg[i][1];
    }
    canvas.flush();
    gpu::check(pnt_tbrush_paint_stroke(_h, canvas.handle(), xy.data(), (int64_t)verticesArg.size()));
    canvas.invalidate();
  }

  // TextureBrush.hxx:207-209
  void enableSmudge(const bool enable) { gpu::check(pnt_tbrush_set_use_smudge(_h, enable ? 1 : 0)); }

  // BrushStrokeSample::setThicknessMap (renderer/src/BrushStrokeSample.cxx:28-31)
  void setThicknessMap(const Mat<double>& thicknessMap) {
    gpu::check(pnt_tbrush_set_thickness_map(_h, thicknessMap.data(), thicknessMap.rows,
                                            thicknessMap.cols));
  }

  // The smudge's pickup map: side x side cells of K, S (AoS) and V.
  int getSmudgeMap(std::vector<double>* K, std::vector<double>* S, std::vector<double>* V) const {
    int side = 0;
    gpu::check(pnt_tbrush_get_smudge_map(_h, &side, nullptr, nullptr, nullptr));
    const std::size_t m = static_cast<std::size_t>(side) * side;
    if (K) K->resize(3 * m);
    if (S) S->resize(3 * m);
    if (V) V->resize(m);
    gpu::check(pnt_tbrush_get_smudge_map(_h, &side, K ? K->data() : nullptr,
                                         S ? S->data() : nullptr, V ? V->data() : nullptr));
    return side;
  }

  pnt_tbrush handle() const { return _h; }

 private:
  static SampleLoader& loader_slot() {
    static SampleLoader l;
    return l;
  }
  static Mat<double> load(const std::string& sampleDir) {
    if (!loader_slot()) throw std::ios_base::failure(sampleDir + "/thickness_map.png");
    return loader_slot()(sampleDir);
  }

  pnt_tbrush _h = nullptr;
};

}  // namespace painty
