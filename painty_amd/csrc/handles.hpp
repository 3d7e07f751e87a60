// Definitions of the opaque handles of painty_hip.h (shared by api.hip and
// persistent.hip).
#pragma once

#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "internal.hpp"

#define PNT_REQUIRE(cond, msg)                        \
  do {                                                \
    if (!(cond)) throw ::pnt::Error(PNT_EINVAL, msg); \
  } while (0)

namespace pnt {
inline int64_t round_up(int64_t n, int64_t a) { return (n + a - 1) / a * a; }
float* dev_alloc(int64_t nfloats);
void dev_free(void* p);
struct PersistentState;
struct ChainState;
struct FusedState;
struct BatchState;
struct KernelTimer;
}  // namespace pnt

struct pnt_ctx_s {
  // One reference for the caller's handle plus one per object made on the
  // context (layers, canvases, brushes, gradient runners, captured batches);
  // the stream and buffers go with the last one, so objects may be destroyed
  // after pnt_ctx_destroy in any order.
  std::atomic<int> refs{1};
  int device = 0;
  hipStream_t stream = nullptr;
  // Pinned host word that persistent kernels raise when a bounded wait
  // expires; checked by pnt_ctx_synchronize and at the next persistent run.
  unsigned* kernel_error = nullptr;
  // plan buffers and launch configuration of the context's batches
  // (batch.hip); made by the first batch call, freed with the context
  pnt::BatchState* batch = nullptr;
  void use() const { PNT_HIP_CHECK(hipSetDevice(device)); }
};

namespace pnt {
void ctx_retain(pnt_ctx ctx);
void ctx_release(pnt_ctx ctx);
pnt_ctx ctx_create_own_queue(int device);  // api.hip
}  // namespace pnt

struct pnt_canvas_s;

struct pnt_layer_s {
  pnt_ctx ctx = nullptr;
  int rows = 0, cols = 0;
  int64_t n = 0, stride = 0;
  float* base = nullptr;
  bool owner = true;
  pnt_canvas_s* parent = nullptr;  // the canvas whose wet layer this views
  pnt::Layer7 planes() const { return pnt::layer_at(base, stride); }
  void alloc(int r, int c) {
    rows = r;
    cols = c;
    n = (int64_t)r * c;
    stride = pnt::round_up(n > 0 ? n : 1, 64);
    base = pnt::dev_alloc(pnt::kLayerStrides * stride);
  }
  void touched();  // the wet layer changed outside a brush's tracking
};

struct pnt_canvas_s {
  pnt_ctx ctx = nullptr;
  int rows = 0, cols = 0;
  int64_t n = 0, stride = 0;
  float* base = nullptr;     // wet layer (strides 0..7: K, S, then V in FP64), 8..10 R0, 11 h
  bool owns_base = true;     // false: base belongs to the creator (canvas_over)
  float* scratch = nullptr;  // 3 planes for host-side compose results
  pnt_layer_s view;
  int64_t drying_ms = 15000000;  // Canvas.hxx:32
  // checkDry clock (PNT_CLOCK_*); tmap holds one time point (us) per pixel
  // while a clock is on
  int clock = PNT_CLOCK_OFF;
  int64_t manual_us = 0;
  // drying timer (pnt_canvas_set_dry_timer): dryStep(dry_portion) every
  // dry_period_us of canvas time, last applied at dry_last_us; 0: off
  int64_t dry_period_us = 0, dry_last_us = 0;
  float dry_portion = 0.01f;
  unsigned* dry_wet = nullptr;  // device word: a pixel still holds paint
  // the GL renderer's lighting thickness (its S.a plane), made zero by the
  // first GL texture-brush stroke; nullptr until then
  float* light = nullptr;
  int64_t* tmap = nullptr;
  float bg[3] = {1.f, 1.f, 1.f};
  // Bumped by every change of the wet layer; a brush's snapshot dirty bytes
  // are valid only for the (canvas, epoch) they were last brought up to.
  uint64_t epoch = 1;
  pnt::Layer7 layer() const { return view.planes(); }
  pnt::Rgb3 R0() const {
    return pnt::Rgb3{{base + 8 * stride, base + 9 * stride, base + 10 * stride}};
  }
  float* h() const { return base + 11 * stride; }
};

inline void pnt_layer_s::touched() {
  if (parent) parent->epoch++;
}

struct pnt_brush_s {
  pnt_ctx ctx = nullptr;
  double radius = 0.0;
  int width = 0, size_map = 0;
  int64_t m = 0, mstride = 0;
  float* fp = nullptr;         // footprint, size_map^2
  int64_t fp_cap = 0, pick_cap = 0;  // floats allocated at fp / pick
  std::vector<float> fp_host;  // same footprint on the host (traffic replay)
  int fbox[4] = {0, 0, 1 << 30, 1 << 30};  // nonzero footprint cells: x0 y0 x1 y1
  float* pick = nullptr;       // pickup map: 7 FP32 planes of mstride
  float* snap = nullptr;       // snapshot: a layer (8 strides of snap_stride)
  int snap_rows = 0, snap_cols = 0;
  int64_t snap_stride = 0, snap_cap = 0;  // floats allocated at snap
  bool use_snap = true;
  double pickup_rate = 0.9, dep_rate = 0.05, cap = 1.0;  // FootprintBrush.hxx:483-495
  float paintK[3] = {0, 0, 0}, paintS[3] = {0, 0, 0};
  // dependency lists of the current stamp's shared-pixel visits (plain path)
  int2* d_list = nullptr;
  int* d_off = nullptr;
  int64_t d_cap = 0;
  // snapshot dirty bytes (persistent path): dirty[i] == 0 means snapshot
  // pixel i equals canvas pixel i of (dirty_canvas, dirty_epoch)
  uint8_t* dirty = nullptr;
  int64_t dirty_n = 0, dirty_cap = 0;  // pixels covered / bytes allocated
  const pnt_canvas_s* dirty_canvas = nullptr;
  uint64_t dirty_epoch = 0;
  pnt::PersistentState* pers = nullptr;
  pnt::ChainState* chain = nullptr;  // buffers of PNT_EXEC_CHAINED runs
  pnt::FusedState* fused = nullptr;  // streams and maps of concurrent fused strokes
  int64_t last_launches = 0;  // stamp-kernel launches of the last fused stamp-list call
  pnt::KernelTimer* ktimer = nullptr;  // per-launch event pairs (pnt_brush_set_kernel_timing)

  pnt::Plane7 pick_planes() const {
    pnt::Plane7 L;
    for (int i = 0; i < 7; ++i) L.p[i] = pick + i * mstride;
    return L;
  }
  pnt::Layer7 snap_planes() const { return pnt::layer_at(snap, snap_stride); }
  pnt::BrushArgs args() const {
    pnt::BrushArgs b;
    b.fp = fp;
    b.pick = pick_planes();
    b.pickup_rate = pickup_rate;
    b.dep_rate = dep_rate;
    b.cap = cap;
    for (int c = 0; c < 3; ++c) {
      b.paintK[c] = paintK[c];
      b.paintS[c] = paintS[c];
    }
    return b;
  }
};

namespace pnt {
void brush_full_snapshot(pnt_brush b, pnt_canvas cv);
StampArgs stamp_args(pnt_brush b, pnt_canvas cv, double cx, double cy, double theta);
void plan_shared_visits(const StampArgs& a, std::vector<int2>& list, std::vector<int>& off);
void launch_dry_dirty(int64_t n, const Layer7& L, const Rgb3& R0, float* h, uint8_t* dirty,
                      hipStream_t s);
// Snapshot of brush b ready for stamping canvas cv (allocated, copied when
// the size changed) and its dirty bytes valid for cv.
void brush_prepare_snapshot(pnt_brush b, pnt_canvas cv, hipStream_t s);
// FootprintBrush::setRadius geometry with an injected width x width footprint
// (api.hip): new footprint and a cleared pickup map.
void brush_set_geometry(pnt_brush b, double radius, const float* core, int width);
// The footprint padded with zeros to size_map^2 (FootprintBrush.hxx:52-58,
// padded to exactly sizeMap), after the checks brush_set_geometry makes.
std::vector<float> padded_footprint(double radius, const float* core, int width, int* size_map);
// The two steps of padded_footprint: the map side (validating radius and
// width), and one pass that validates the core, writes it centred in a
// size_map^2 zero map at dst and gives the box of its nonzero cells.
int footprint_size_map(double radius, const float* core, int width);
void pad_footprint_into(const float* core, int width, int size_map, float* dst, int box[4]);
// The bounding box (x0, y0, x1, y1) of the nonzero cells of a padded
// size_map^2 footprint; empty (x1 < x0) when all are zero.
void footprint_box(const std::vector<float>& fp, int size_map, int box[4]);
// Palette JSON (mixer/src/Serialization.cxx:26-53), ingest.hip.
void palette_from_json(const char* text, int64_t len, std::vector<double>& K,
                       std::vector<double>& S);
std::string palette_to_json(const double* K, const double* S, int64_t n);
// A canvas on a caller buffer of at least kCanvasStrides*round_up(rows*cols, 64) floats
// that pnt_canvas_destroy does not free: cleared (background 1) when `clear`,
// else a view of the planes as they are.
pnt_canvas canvas_over(pnt_ctx ctx, int rows, int cols, float* base, bool clear = true);
// The canvas clock's current time point (microseconds).
int64_t canvas_now_us(const pnt_canvas_s* cv);
// Runs the drying-timer steps whose ticks have passed (nothing when the
// timer or the clock is off); every call that reads or changes a canvas
// starts with it.
void canvas_tick(pnt_canvas_s* cv);
// Every pixel's time point set to now (Canvas::clear, dryCanvas), if a clock is on.
void canvas_stamp_time(pnt_canvas cv);
}  // namespace pnt
