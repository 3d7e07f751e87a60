// PaintLayer<vector_type> -- drop-in for renderer/PaintLayer.hxx:21-145,
// backed by a device layer (7 FP32 planes) through the C ABI.
//
// The reference hands out its cv::Mat buffers by reference.  Here the
// getters return host mirrors: a const getter downloads when the device is
// newer; a mutable getter also marks the mirror as written, and it is
// uploaded before the next device operation on the layer.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "painty/core/Vec.hxx"
#include "painty/gpu/Device.hxx"
#include "painty/image/Mat.hxx"

namespace painty {

template <class vector_type>
class Canvas;

template <class vector_type>
class PaintLayer final {
  using T = typename DataType<vector_type>::channel_type;
  static constexpr auto N = DataType<vector_type>::dim;
  static_assert(N == 3, "the device path is RGB (3 channels)");

 public:
  // PaintLayer.hxx:27-30: all buffers zero.
  PaintLayer(int32_t rows, int32_t cols) : _rows(rows), _cols(cols), _owner(true) {
    gpu::check(pnt_layer_create(gpu::Device::get(), rows, cols, &_h));
  }
  PaintLayer(const PaintLayer& o) : PaintLayer(o._rows, o._cols) { o.copyTo(*this); }
  PaintLayer& operator=(const PaintLayer& o) {
    if (this != &o) o.copyTo(*this);
    return *this;
  }
  ~PaintLayer() {
    if (_h) pnt_layer_destroy(_h);
  }

  const Mat<vector_type>& getK_buffer() const { return pull(), _K; }
  const Mat<vector_type>& getS_buffer() const { return pull(), _S; }
  const Mat<T>& getV_buffer() const { return pull(), _V; }
  Mat<vector_type>& getK_buffer() { return touch(), _K; }
  Mat<vector_type>& getS_buffer() { return touch(), _S; }
  Mat<T>& getV_buffer() { return touch(), _V; }

  int32_t getRows() const { return _rows; }
  int32_t getCols() const { return _cols; }

  // PaintLayer.hxx:68-74
  void clear() {
    gpu::check(pnt_layer_clear(_h));
    _m = gpu::Mirror::Stale;
  }

  // PaintLayer.hxx:81-96: R0 = KM(layer over R0) per pixel; R0 is
  // reallocated and filled with 1.0 when its size differs.
  void composeOnto(Mat<vector_type>& R0) const {
    flush();
    if (R0.rows != _rows || R0.cols != _cols) R0 = Mat<vector_type>(_rows, _cols, ones());
    std::vector<float> r = gpu::to_f32(R0.data()->data(), R0.total() * 3);
    gpu::check(pnt_layer_compose_onto_host(_h, r.data(), _rows, _cols, PNT_AOS));
    gpu::from_f32(r, R0.data()->data());
  }

  // PaintLayer.hxx:103-111 (deep copy; reallocates on a size mismatch).
  void copyTo(PaintLayer& other) const {
    flush();
    gpu::check(pnt_layer_copy_to(_h, other._h));
    gpu::check(pnt_layer_shape(other._h, &other._rows, &other._cols));
    other._m = gpu::Mirror::Stale;
  }

  // PaintLayer.hxx:122-127
  void set(int32_t i, int32_t j, const vector_type& k, const vector_type& s, const T v) {
    flush();
    const float kf[3] = {float(k[0]), float(k[1]), float(k[2])};
    const float sf[3] = {float(s[0]), float(s[1]), float(s[2])};
    gpu::check(pnt_layer_set(_h, i, j, kf, sf, static_cast<float>(v)));
    if (_m == gpu::Mirror::Clean) {  // keep the mirror in step
      _K(i, j) = vector_type(kf[0], kf[1], kf[2]);
      _S(i, j) = vector_type(sf[0], sf[1], sf[2]);
      _V(i, j) = static_cast<T>(static_cast<float>(v));
    }
  }

  // ---- beyond the reference: the device handle and mirror control
  pnt_layer handle() const { return _h; }
  // Uploads mirrors written through a mutable getter.
  void flush() const {
    if (_m != gpu::Mirror::HostDirty) return;
    const std::size_t n = static_cast<std::size_t>(_rows) * _cols;
    const auto k = gpu::to_f32(_K.data()->data(), 3 * n);
    const auto s = gpu::to_f32(_S.data()->data(), 3 * n);
    gpu::check(pnt_layer_upload(_h, k.data(), s.data(), nullptr, PNT_AOS));
    // V travels in FP64, as the device holds it
    std::vector<double> v(_V.data(), _V.data() + n);
    gpu::check(pnt_layer_upload_volume(_h, v.data()));
    _m = gpu::Mirror::Clean;
  }
  // The device data changed behind the mirrors.
  void invalidate() const { _m = gpu::Mirror::Stale; }

 private:
  friend class Canvas<vector_type>;
  // A view on a canvas' wet layer (not owned).
  explicit PaintLayer(pnt_layer view) : _owner(false), _h(view) {
    gpu::check(pnt_layer_shape(_h, &_rows, &_cols));
  }

  static vector_type ones() { return vector_type(1.0, 1.0, 1.0); }

  void pull() const {
    if (_m != gpu::Mirror::Stale) return;
    const std::size_t n = static_cast<std::size_t>(_rows) * _cols;
    std::vector<float> k(3 * n), s(3 * n);
    std::vector<double> v(n);
    gpu::check(pnt_layer_download(_h, k.data(), s.data(), nullptr, PNT_AOS));
    gpu::check(pnt_layer_download_volume(_h, v.data()));
    _K = Mat<vector_type>(_rows, _cols);
    _S = Mat<vector_type>(_rows, _cols);
    _V = Mat<T>(_rows, _cols);
    gpu::from_f32(k, _K.data()->data());
    gpu::from_f32(s, _S.data()->data());
    for (std::size_t i = 0; i < n; ++i) _V.data()[i] = static_cast<T>(v[i]);
    _m = gpu::Mirror::Clean;
  }
  void touch() {
    pull();
    _m = gpu::Mirror::HostDirty;
  }

  int32_t _rows = 0, _cols = 0;
  bool _owner = true;
  pnt_layer _h = nullptr;
  mutable Mat<vector_type> _K, _S;
  mutable Mat<T> _V;
  mutable gpu::Mirror _m = gpu::Mirror::Stale;
};

}  // namespace painty
