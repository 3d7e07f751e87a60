// Renderer<vector_type> -- drop-in for renderer/Renderer.hxx:26-156: the KM
// composite of a wet layer over R0 (compose) and the lit display image
// (render), on the device.
#pragma once

#include "painty/renderer/Canvas.hxx"

namespace painty {

template <class vector_type>
class Renderer final {
 public:
  // Renderer.hxx:26-41: R = ComputeReflectance(K, S, R0, V) per pixel.
  Mat<vector_type> compose(const PaintLayer<vector_type>& paintLayer,
                           const Mat<vector_type>& R0_buffer) const {
    if (R0_buffer.rows != paintLayer.getRows() || R0_buffer.cols != paintLayer.getCols())
      throw std::invalid_argument("R0 size differs from the paint layer");
    paintLayer.flush();
    const auto r0 = gpu::to_f32(R0_buffer.data()->data(), R0_buffer.total() * 3);
    std::vector<float> r(r0.size());
    gpu::check(pnt_compose_layer_host(paintLayer.handle(), r0.data(), r.data(), PNT_AOS));
    Mat<vector_type> R1(R0_buffer.rows, R0_buffer.cols);
    gpu::from_f32(r, R1.data()->data());
    return R1;
  }

  // Renderer.hxx:48-53: compose(canvas.getPaintLayer(), canvas.getR0()).
  Mat<vector_type> compose(const Canvas<vector_type>& canvas) const {
    canvas.flush();
    std::vector<float> r(static_cast<std::size_t>(canvas.getRows()) * canvas.getCols() * 3);
    gpu::check(pnt_compose(canvas.handle(), r.data(), PNT_AOS));
    Mat<vector_type> R1(canvas.getRows(), canvas.getCols());
    gpu::from_f32(r, R1.data()->data());
    return R1;
  }

  // Renderer.hxx:60-156: the composite lit by a point light with a
  // Beckmann / Cook-Torrance specular term over the wet layer's thickness,
  // clamped to [0, 1].
  Mat<vector_type> render(const Canvas<vector_type>& canvas) const {
    canvas.flush();
    std::vector<float> r(static_cast<std::size_t>(canvas.getRows()) * canvas.getCols() * 3);
    gpu::check(pnt_render(canvas.handle(), r.data(), PNT_AOS));
    Mat<vector_type> rgb(canvas.getRows(), canvas.getCols());
    gpu::from_f32(r, rgb.data()->data());
    return rgb;
  }
};

}  // namespace painty
