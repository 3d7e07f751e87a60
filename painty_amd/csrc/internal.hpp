// Internal interfaces between the C ABI (api.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "painty_hip.h"
#include "stamp_geom.hpp"

namespace pnt {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PNT_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::pnt::Error(PNT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A paint layer on the device: FP32 planes K0 K1 K2 S0 S1 S2 and the
// volume plane V in FP64.  Volume feeds the reference's MinVolume and
// drying thresholds (FootprintBrush.hxx:26, Canvas.hxx:123-155); held in
// FP64, as the reference holds it, every such test takes the reference's
// branch.  V lies at float offset 6*stride of the layer's block (two float
// strides wide).
struct Layer7 {
  float* ks[6];
  double* v;
};
// Seven FP32 planes K0 K1 K2 S0 S1 S2 V: the pickup map, and caller-owned
// device layers (pnt_dev_layer).
struct Plane7 {
  float* p[7];
};
// The layer whose planes start at `base`, `stride` floats apart.
__host__ __device__ inline Layer7 layer_at(float* base, int64_t stride) {
  Layer7 L;
  for (int i = 0; i < 6; ++i) L.ks[i] = base + i * stride;
  L.v = reinterpret_cast<double*>(base + 6 * stride);
  return L;
}
// Strides a layer takes (6 FP32 planes + the FP64 plane), and a canvas
// block: its wet layer, R0 (3 planes, strides 8..10) and h (stride 11).
constexpr int kLayerStrides = 8;
constexpr int kCanvasStrides = 12;
struct Rgb3 {
  float* p[3];
};

// A pointer in the global address space.  Kernels that read device-memory
// records holding pointers view them through a mirror record with these
// fields: a pointer loaded as global keeps that address space through the
// code that uses it, so its loads and stores compile to global_* memory
// operations instead of flat ones (which also wait on LDS traffic).
template <class T>
using gptr = __attribute__((address_space(1))) T*;

// Brush state and parameters one imprint needs.
struct BrushArgs {
  const float* fp;   // footprint, size_map^2
  Plane7 pick;       // pickup map, size_map^2 planes
  double pickup_rate, dep_rate, cap;
  float paintK[3], paintS[3];
};

// Canvas::checkDry state of one imprint (Canvas.hxx:123-155); a null tmap
// switches the check off.
struct DryClock {
  int64_t* tmap;    // time map of the canvas, microseconds per pixel
  int64_t now_us;   // the imprint's time point
  int64_t dry_ms;   // drying time
  Rgb3 R0;          // substrate and height the drying writes
  float* h;
};

// One imprint: geometry + where its visits read and write.
struct StampArgs {
  StampGeom g;
  BrushArgs b;
  Layer7 canvas;     // wet layer of the canvas
  Layer7 src;        // pickup source: the snapshot, or the canvas itself
  int dupx, dupy;    // canvas column / row 0 visited twice (see stamp_dup_axis)
  DryClock clk;
};

// 4-pixel groups a ring-scan thread takes per pass (ring_blocks_for).  More
// groups per thread means fewer ring blocks per stamp but longer ones, and
// a stamp launch lasts as long as its slowest block: two measured fastest.
constexpr int kRingGroups = 2;

// Snapshot ring of updateSnapshot(canvas, center), FootprintBrush.hxx:286-305.
struct RingBox {
  int a_x0, a_y0, a_x1, a_y1;  // allowed box (inclusive, clipped to the canvas)
  int tl_x, tl_y, br_x, br_y;  // window corners; strict interior is skipped
};

// ---- composite / dry (kernels.hip) ----
void launch_compose(int64_t n, const Layer7& L, const Rgb3& R0, const Rgb3& R, hipStream_t s);
void launch_compose2(int64_t n, const Layer7& dry, const Layer7& wet, const Rgb3& R0,
                     const Rgb3& R, hipStream_t s);
// The same over caller-owned FP32 planes (pnt_dev_layer).
void launch_compose(int64_t n, const Plane7& L, const Rgb3& R0, const Rgb3& R, hipStream_t s);
void launch_compose2(int64_t n, const Plane7& dry, const Plane7& wet, const Rgb3& R0,
                     const Rgb3& R, hipStream_t s);
void launch_dry(int64_t n, const Layer7& L, const Rgb3& R0, float* h, hipStream_t s);
void launch_dry_step(int64_t n, const Layer7& L, const Rgb3& R0, float portion,
                     hipStream_t s);
void launch_reflectance_aos(int64_t n, const float* K, const float* S, const float* R0,
                            const float* d, float* R, hipStream_t s);
void launch_fill(float* p, int64_t n, float v, hipStream_t s);
// `steps` dry steps of `portion` (k_dry_step repeated per pixel); *wet is
// raised when a pixel still holds paint after them (zero it first).
void launch_dry_steps(int64_t n, const Layer7& L, const Rgb3& R0, float portion, int steps,
                      unsigned* wet, hipStream_t s);
void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);
void launch_fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t s);
void launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s);

// ---- ingest (ingest.hip) ----
void launch_scatter_absorb(int64_t n3, const double* Rb, const double* Rw, double* K, double* S,
                           unsigned* bad, hipStream_t s);

// ---- display (display.hip) ----
void launch_render(int rows, int cols, const Layer7& L, const Rgb3& R0, const Rgb3& out,
                   hipStream_t s);
void launch_encode(int64_t n, const Rgb3& in, int to_srgb, int bits, void* out, hipStream_t s);
void launch_check_dry_pixel(const DryClock& c, const Layer7& canvas, int64_t ic, hipStream_t s);
void launch_copy_layer(int64_t n, const Layer7& src, const Layer7& dst, hipStream_t s);
// A layer from seven FP32 planes (V widened exactly), and back (V rounded).
void launch_layer_from_planes(int64_t n, const Plane7& src, const Layer7& dst, hipStream_t s);
void launch_planes_from_layer(int64_t n, const Layer7& src, const Plane7& dst, hipStream_t s);

// ---- stamping (kernels.hip) ----
void launch_ring_refresh(const RingBox& box, int cols, const Layer7& canvas, const Layer7& snap,
                         hipStream_t s);
void launch_stamp_main(const StampArgs& a, hipStream_t s);
// Visits that share a canvas pixel at column/row 0 (and their map-cell
// partners), listed in reference order and grouped by dependency level.
void launch_stamp_serial(const StampArgs& a, const int2* list, const int* level_off, int nlevels,
                         hipStream_t s);

RingBox ring_box(const StampGeom& g, double radius);

}  // namespace pnt
