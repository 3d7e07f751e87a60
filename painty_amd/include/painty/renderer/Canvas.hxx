// Canvas<vector_type> -- drop-in for renderer/Canvas.hxx:20-198, backed by a
// device canvas (wet layer, R0, h) through the C ABI.
//
// Host mirrors stand in for the reference's buffers, as in PaintLayer.hxx.
// Drying is deterministic: only dryCanvas() (and stamp lists that ask for
// it) dries paint; the wall-clock checkDry / time map of the reference
// (Canvas.hxx:97-103, 123-155) are not kept.
#pragma once

#include <chrono>
#include <cstdint>

#include "painty/renderer/PaintLayer.hxx"

namespace painty {

template <class vector_type>
class Canvas final {
  using T = typename DataType<vector_type>::channel_type;

 public:
  // Canvas.hxx:26-35: background (1,1,1), empty wet layer, h = 0.
  Canvas(const int32_t rows, const int32_t cols) : _rows(rows), _cols(cols) {
    gpu::check(pnt_canvas_create(gpu::Device::get(), rows, cols, &_h));
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

  // Canvas.hxx:105-121
  void dryCanvas() {
    flush();
    gpu::check(pnt_canvas_dry(_h));
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
  pnt_canvas handle() const { return _h; }
  int32_t getRows() const { return _rows; }
  int32_t getCols() const { return _cols; }
  // Uploads every mirror written through a mutable getter.
  void flush() const {
    _layer->flush();
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
  }

 private:
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
};

}  // namespace painty
