// FootprintBrush<vector_type> -- drop-in for renderer/FootprintBrush.hxx:
// 22-502: the imprint brush with pickup map and snapshot buffer, run on the
// device through the C ABI.
//
// The reference reads its footprint from ./data/footprint/footprint.png and
// resizes it to width x width (width = 2*ceil(radius)+1) with Lanczos-4
// (FootprintBrush.hxx:48-55).  Here the footprint is a width x width image
// passed to the constructor / setRadius, or, when none is passed, comes from
// the brush's footprint source.  The default source does what the reference
// does when ./data/footprint/footprint.png can be read (pnt_image_read_gray,
// then pnt_resize_lanczos4 on the device; the decoded image is kept per
// path), and gives a Gaussian (sigma = width/4, peak 1) when it cannot,
// where the reference would throw.  setFootprintSource replaces it, e.g.
// with imageSource(path).
#pragma once

#include <array>
#include <cmath>
#include <fstream>
#include <functional>
#include <ios>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "painty/renderer/BrushBase.hxx"

namespace painty {

template <class vector_type>
class FootprintBrush final : public BrushBase<vector_type> {
  using T = typename DataType<vector_type>::channel_type;

 public:
  using FootprintSource = std::function<Mat<double>(int32_t width)>;

  // FootprintBrush.hxx:29-44
  explicit FootprintBrush(const double radius) : FootprintBrush(radius, source()(width_of(radius))) {}
  FootprintBrush(const double radius, const Mat<double>& footprint) {
    const auto f = gpu::to_f32(footprint.data(), footprint.total());
    check_square(footprint);
    gpu::check(pnt_brush_create(gpu::Device::get(), radius, f.data(), footprint.rows, &_h));
  }
  FootprintBrush(const FootprintBrush&) = delete;
  FootprintBrush& operator=(const FootprintBrush&) = delete;
  ~FootprintBrush() override {
    if (_h) pnt_brush_destroy(_h);
  }

  // FootprintBrush.hxx:46-63 (nothing happens when |radius - current| < 0.5)
  void setRadius(const double radius) override { setRadius(radius, source()(width_of(radius))); }
  void setRadius(const double radius, const Mat<double>& footprint) {
    check_square(footprint);
    const auto f = gpu::to_f32(footprint.data(), footprint.total());
    gpu::check(pnt_brush_set_radius(_h, radius, f.data(), footprint.rows));
    _fp_valid = false;
  }

  // FootprintBrush.hxx:73-143
  void imprint(const vec2& center, const double theta, Canvas<vector_type>& canvas) {
    canvas.flush();
    gpu::check(pnt_imprint(_h, canvas.handle(), center[0], center[1], theta));
    canvas.invalidate();
  }

  // FootprintBrush.hxx:150-154: clean(), then the brush paint is {K, S}.
  void dip(const std::array<vector_type, 2UL>& paint) override {
    const float k[3] = {float(paint[0][0]), float(paint[0][1]), float(paint[0][2])};
    const float s[3] = {float(paint[1][0]), float(paint[1][1]), float(paint[1][2])};
    gpu::check(pnt_brush_dip(_h, k, s));
  }

  // FootprintBrush.hxx:160-166
  void clean() { gpu::check(pnt_brush_clean(_h)); }

  // FootprintBrush.hxx:168-172
  void updateSnapshot(const Canvas<vector_type>& canvas) {
    canvas.flush();
    gpu::check(pnt_brush_update_snapshot(_h, canvas.handle()));
  }

  // FootprintBrush.hxx:174-176: a copy of the pickup map (sizeMap x sizeMap).
  const PaintLayer<vector_type>& getPickupMap() const {
    int m = 0;
    gpu::check(pnt_brush_size_map(_h, &m));
    std::vector<float> k(3 * (size_t)m * m), s(3 * (size_t)m * m), v((size_t)m * m);
    gpu::check(pnt_brush_get_pickup_map(_h, k.data(), s.data(), v.data(), PNT_AOS));
    if (!_pickup || _pickup->getRows() != m) _pickup.reset(new PaintLayer<vector_type>(m, m));
    gpu::check(pnt_layer_upload(_pickup->handle(), k.data(), s.data(), v.data(), PNT_AOS));
    _pickup->invalidate();
    return *_pickup;
  }

  // FootprintBrush.hxx:178-180 (padded to sizeMap x sizeMap)
  const Mat<double>& getFootprint() const {
    if (!_fp_valid) {
      int m = 0;
      gpu::check(pnt_brush_size_map(_h, &m));
      std::vector<float> f((size_t)m * m);
      gpu::check(pnt_brush_get_footprint(_h, f.data()));
      _fp = Mat<double>(m, m);
      gpu::from_f32(f, _fp.data());
      _fp_valid = true;
    }
    return _fp;
  }

  // FootprintBrush.hxx:182-204
  void setPickupRate(const T rate) { gpu::check(pnt_brush_set_pickup_rate(_h, double(rate))); }
  void setDepositionRate(const T rate) {
    gpu::check(pnt_brush_set_deposition_rate(_h, double(rate)));
  }
  T getPickupRate() const {
    double r = 0;
    gpu::check(pnt_brush_get_pickup_rate(_h, &r));
    return static_cast<T>(r);
  }
  T getDepositionRate() const {
    double r = 0;
    gpu::check(pnt_brush_get_deposition_rate(_h, &r));
    return static_cast<T>(r);
  }
  bool getUseSnapshotBuffer() const {
    int u = 0;
    gpu::check(pnt_brush_get_use_snapshot(_h, &u));
    return u != 0;
  }
  void setUseSnapshotBuffer(const bool use) { gpu::check(pnt_brush_set_use_snapshot(_h, use)); }

  // FootprintBrush.hxx:206-268: Catmull-Rom placement at 1 px chord steps,
  // one imprint per step (the first segment uses p_pre = path[0]).
  void paintStroke(const std::vector<vec2>& path, Canvas<vector_type>& canvas) override {
    std::vector<double> xy;
    xy.reserve(2 * path.size());
    for (const auto& p : path) {
      xy.push_back(p[0]);
      xy.push_back(p[1]);
    }
    canvas.flush();
    gpu::check(pnt_paint_stroke(_h, canvas.handle(), xy.data(), (int64_t)path.size()));
    canvas.invalidate();
  }

  // ---- beyond the reference
  // An ordered dip / imprint / dryCanvas list in one call (PNT_EXEC_*).
  void imprintStamps(const std::vector<pnt_stamp>& stamps, Canvas<vector_type>& canvas,
                     int mode = PNT_EXEC_FUSED) {
    canvas.flush();
    gpu::check(pnt_imprint_stamps(_h, canvas.handle(), stamps.data(), (int64_t)stamps.size(),
                                  mode));
    canvas.invalidate();
  }
  // A painting session in one call: per stroke setRadius(radius, footprint)
  // then its stamps (pnt_imprint_strokes); under PNT_EXEC_FUSED strokes whose
  // snapshot boxes do not meet run concurrently, with the same result.
  struct Stroke {
    double radius;
    Mat<double> footprint;
    std::vector<pnt_stamp> stamps;
  };
  void imprintStrokes(const std::vector<Stroke>& strokes, Canvas<vector_type>& canvas,
                      int mode = PNT_EXEC_FUSED) {
    std::vector<std::vector<float>> fps;
    std::vector<pnt_stroke> list;
    fps.reserve(strokes.size());
    for (const Stroke& s : strokes) {
      check_square(s.footprint);
      fps.push_back(gpu::to_f32(s.footprint.data(), s.footprint.total()));
      list.push_back(pnt_stroke{s.radius, fps.back().data(), s.footprint.rows, s.stamps.data(),
                                (int64_t)s.stamps.size()});
    }
    canvas.flush();
    gpu::check(pnt_imprint_strokes(_h, canvas.handle(), list.data(), (int64_t)list.size(), mode));
    canvas.invalidate();
    _fp_valid = false;
  }
  pnt_brush handle() const { return _h; }
  static void setFootprintSource(FootprintSource s) { source() = std::move(s); }
  // ScaledMat(imRead(path, gray, true), width, width) (FootprintBrush.hxx:
  // 48-55); throws std::ios_base::failure when the file cannot be read, as
  // imRead does.
  static Mat<double> fromImage(const std::string& path, int32_t width) {
    const Mat<double>& full = decoded(path);
    Mat<double> out(width, width);
    gpu::check(pnt_resize_lanczos4(gpu::Device::get(), full.data(), full.rows, full.cols, width,
                                   width, out.data()));
    return out;
  }
  static FootprintSource imageSource(std::string path) {
    return [path](int32_t width) { return fromImage(path, width); };
  }
  // The reference's footprint width for a radius (FootprintBrush.hxx:50).
  static int32_t width_of(double radius) {
    return 2 * static_cast<int32_t>(std::ceil(radius)) + 1;
  }

 private:
  static void check_square(const Mat<double>& f) {
    if (f.rows != f.cols || f.rows < 1 || f.rows % 2 == 0)
      throw std::invalid_argument("footprint must be an odd-sized square image");
  }
  static Mat<double> gaussian(int32_t width) {
    Mat<double> f(width, width);
    const double c = (width - 1) / 2.0, sigma = width / 4.0;
    for (int32_t i = 0; i < width; ++i)
      for (int32_t j = 0; j < width; ++j)
        f(i, j) = std::exp(-((i - c) * (i - c) + (j - c) * (j - c)) / (2.0 * sigma * sigma));
    return f;
  }
  static Mat<double> reference_or_gaussian(int32_t width) {
    static const std::string path = "./data/footprint/footprint.png";
    if (std::ifstream(path, std::ios::binary).good()) return fromImage(path, width);
    return gaussian(width);
  }
  // the linear grey image of each footprint file read so far
  static const Mat<double>& decoded(const std::string& path) {
    static std::mutex mu;
    static std::map<std::string, Mat<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(path);
    if (it != cache.end()) return it->second;
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::ios_base::failure(path);
    const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)),
                                     std::istreambuf_iterator<char>());
    int rows = 0, cols = 0;
    gpu::check(pnt_image_read_gray(bytes.data(), (int64_t)bytes.size(), 1, &rows, &cols, nullptr, 0));
    Mat<double> img(rows, cols);
    gpu::check(pnt_image_read_gray(bytes.data(), (int64_t)bytes.size(), 1, &rows, &cols, img.data(),
                                   (int64_t)img.total()));
    return cache.emplace(path, std::move(img)).first->second;
  }
  static FootprintSource& source() {
    static FootprintSource s = &FootprintBrush::reference_or_gaussian;
    return s;
  }

  pnt_brush _h = nullptr;
  mutable std::unique_ptr<PaintLayer<vector_type>> _pickup;
  mutable Mat<double> _fp;
  mutable bool _fp_valid = false;
};

}  // namespace painty
