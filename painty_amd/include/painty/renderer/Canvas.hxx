// Canvas<vector_type> -- drop-in for renderer/Canvas.hxx:20-198, backed by a
// device canvas (wet layer, R0, h) through the C ABI.
//
// Host mirrors stand in for the reference's buffers, as in PaintLayer.hxx.
// As in the reference, every imprint runs checkDry against the system clock
// and the canvas keeps a time map (Canvas.hxx:97-103, 123-155); useClock()
// switches to a manual clock or turns the check off (deterministic drying).
#pragma once

#include <chrono>
#include <cstdint>
#include <vector>

#include "painty/renderer/PaintLayer.hxx"

namespace painty {

template <class vector_type>
class Canvas final {
  using T = typename DataType<vector_type>::channel_type;

 public:
  // Canvas.hxx:26-35: background (1,1,1), empty wet layer, h = 0.
  Canvas(const int32_t rows, const int32_t cols) : _rows(rows), _cols(cols) {
    gpu::check(pnt_canvas_create(gpu::Device::get(), rows, cols, &_h));
    gpu::check(pnt_canvas_set_clock(_h, PNT_CLOCK_SYSTEM));
    pnt_layer view = nullptr;
    gpu::check(pnt_canvas_layer(_h, &view));
    _layer = new PaintLayer<vector_type>(view);
  }
  Canvas(const Canvas&) = delete;
  Canvas& operator=(const Canvas&) = delete;
  ~Canvas() {
    delete _layer;
    if (_h) pnt_canvas_destroy(_h);
  }

  // Canvas.hxx:37-58
  void clear() {
    gpu::check(pnt_canvas_clear(_h));
    invalidate();
  }

  const Mat<vector_type>& getR0() const { return pull(), _R0; }
  const Mat<T>& get_h() const { return pull(), _hb; }
  Mat<vector_type>& getR0() { return pull(), _m = gpu::Mirror::HostDirty, _R0; }
  Mat<T>& get_h() { return pull(), _m = gpu::Mirror::HostDirty, _hb; }

  // Canvas.hxx:76-78
  Mat<vector_type> getReflectanceLayerDry() const { return getR0().clone(); }

  // Canvas.hxx:80-87: clear(), then R0 = background.
  void setBackground(const Mat<vector_type>& background) {
    if (background.rows != _rows || background.cols != _cols)
      throw std::invalid_argument("background size differs from the canvas");
    const auto r = gpu::to_f32(background.data()->data(), background.total() * 3);
    gpu::check(pnt_canvas_set_background(_h, r.data(), PNT_AOS));
    invalidate();
  }

  const PaintLayer<vector_type>& getPaintLayer() const { return *_layer; }
  PaintLayer<vector_type>& getPaintLayer() { return *_layer; }

  // Canvas.hxx:101-103: one time point per pixel (a host mirror; writes
  // through this reference reach the device at the next flush)
  std::vector<std::chrono::system_clock::time_point>& getTimeMap() {
    pull_time();
    _tm = gpu::Mirror::HostDirty;
    return _time;
  }

  // Canvas.hxx:105-121
  void dryCanvas() {
    flush();
    gpu::check(pnt_canvas_dry(_h));
    invalidate();
  }

  // Canvas.hxx:123-155 for one pixel at time point `timePoint`.
  void checkDry(int32_t x, int32_t y, const std::chrono::system_clock::time_point& timePoint) {
    if (x < 0 || y < 0 || x >= _cols || y >= _rows)
      throw std::out_of_range("checkDry outside the canvas");
    flush();
    gpu::check(pnt_canvas_check_dry(_h, x, y, to_us(timePoint)));
    invalidate();
  }

  // Canvas.hxx:157-163
  std::chrono::milliseconds getDryingTime() {
    int64_t ms = 0;
    gpu::check(pnt_canvas_get_drying_time_ms(_h, &ms));
    return std::chrono::milliseconds(ms);
  }
  void setDryingTime(std::chrono::milliseconds msecs) {
    gpu::check(pnt_canvas_set_drying_time_ms(_h, static_cast<int64_t>(msecs.count())));
  }

  // ---- beyond the reference: the device handle and mirror control
  // PNT_CLOCK_SYSTEM (the default), PNT_CLOCK_MANUAL (time set by setTime) or
  // PNT_CLOCK_OFF (no checkDry; stamp lists may then take the fast paths).
  void useClock(int clock) {
    flush();
    gpu::check(pnt_canvas_set_clock(_h, clock));
    _tm = gpu::Mirror::Stale;
  }
  void setTime(const std::chrono::system_clock::time_point& t) {
    gpu::check(pnt_canvas_set_time_us(_h, to_us(t)));
  }
  pnt_canvas handle() const { return _h; }
  int32_t getRows() const { return _rows; }
  int32_t getCols() const { return _cols; }
  // Uploads every mirror written through a mutable getter.
  void flush() const {
    _layer->flush();
    if (_tm == gpu::Mirror::HostDirty) {
      std::vector<int64_t> t(_time.size());
      for (std::size_t i = 0; i < t.size(); ++i) t[i] = to_us(_time[i]);
      gpu::check(pnt_canvas_upload_time_map(_h, t.data()));
      _tm = gpu::Mirror::Clean;
    }
    if (_m != gpu::Mirror::HostDirty) return;
    const auto r = gpu::to_f32(_R0.data()->data(), _R0.total() * 3);
    const auto h = gpu::to_f32(_hb.data(), _hb.total());
    gpu::check(pnt_canvas_upload_R0(_h, r.data(), PNT_AOS));
    gpu::check(pnt_canvas_upload_h(_h, h.data()));
    _m = gpu::Mirror::Clean;
  }
  // The device canvas changed behind the mirrors.
  void invalidate() const {
    _layer->invalidate();
    _m = gpu::Mirror::Stale;
    _tm = gpu::Mirror::Stale;
  }

 private:
  static int64_t to_us(const std::chrono::system_clock::time_point& t) {
    return std::chrono::duration_cast<std::chrono::microseconds>(t.time_since_epoch()).count();
  }
  void pull_time() const {
    if (_tm != gpu::Mirror::Stale) return;
    const std::size_t n = static_cast<std::size_t>(_rows) * _cols;
    std::vector<int64_t> t(n);
    gpu::check(pnt_canvas_download_time_map(_h, t.data()));
    _time.resize(n);
    for (std::size_t i = 0; i < n; ++i)
      _time[i] = std::chrono::system_clock::time_point(
          std::chrono::duration_cast<std::chrono::system_clock::duration>(
              std::chrono::microseconds(t[i])));
    _tm = gpu::Mirror::Clean;
  }
  void pull() const {
    if (_m != gpu::Mirror::Stale) return;
    const std::size_t n = static_cast<std::size_t>(_rows) * _cols;
    std::vector<float> r(3 * n), h(n);
    gpu::check(pnt_canvas_download_R0(_h, r.data(), PNT_AOS));
    gpu::check(pnt_canvas_download_h(_h, h.data()));
    _R0 = Mat<vector_type>(_rows, _cols);
    _hb = Mat<T>(_rows, _cols);
    gpu::from_f32(r, _R0.data()->data());
    gpu::from_f32(h, _hb.data());
    _m = gpu::Mirror::Clean;
  }

  int32_t _rows, _cols;
  pnt_canvas _h = nullptr;
  PaintLayer<vector_type>* _layer = nullptr;
  mutable Mat<vector_type> _R0;
  mutable Mat<T> _hb;
  mutable gpu::Mirror _m = gpu::Mirror::Stale;
  mutable std::vector<std::chrono::system_clock::time_point> _time;
  mutable gpu::Mirror _tm = gpu::Mirror::Stale;
};

}  // namespace painty
