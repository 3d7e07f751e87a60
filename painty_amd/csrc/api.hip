// C ABI of painty_hip.h: handles, host<->device transfer, stamp planning.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <climits>
#include <utility>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"

using namespace pnt;

namespace {

// Message of the last failed call on this thread.  A plain character array:
// it has no destructor, so calls made while the process exits (a caller's
// static objects releasing their handles run after thread-local objects are
// destroyed) still find valid storage.
thread_local char g_err[1024];

void set_err(const char* m) {
  std::strncpy(g_err, m, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = '\0';
}

template <class F>
int guard(F&& f) {
  try {
    f();
    g_err[0] = '\0';
    return PNT_OK;
  } catch (const Error& e) {
    set_err(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return PNT_ENOMEM;
  } catch (const std::exception& e) {
    set_err(e.what());
    return PNT_EINVAL;
  }
}

// Host AoS/planar <-> device planes.
void upload_planes(float* const* dst, int nplanes, const float* src, int64_t n, int layout,
                   hipStream_t s) {
  if (nplanes == 1 || layout == PNT_PLANAR) {
    for (int c = 0; c < nplanes; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(dst[c], src + c * n, n * sizeof(float), hipMemcpyHostToDevice, s));
  } else {
    PNT_REQUIRE(layout == PNT_AOS, "layout must be PNT_AOS or PNT_PLANAR");
    std::vector<float> tmp((size_t)(n * nplanes));
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < nplanes; ++c) tmp[(size_t)(c * n + i)] = src[i * nplanes + c];
    for (int c = 0; c < nplanes; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(dst[c], tmp.data() + c * n, n * sizeof(float),
                                   hipMemcpyHostToDevice, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  PNT_HIP_CHECK(hipStreamSynchronize(s));
}

// Host FP32 volumes <-> a layer's FP64 V plane (widened exactly, rounded back).
void upload_volume(double* dst, const float* src, int64_t n, hipStream_t s) {
  std::vector<double> tmp(src, src + n);
  PNT_HIP_CHECK(hipMemcpyAsync(dst, tmp.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  PNT_HIP_CHECK(hipStreamSynchronize(s));
}

void download_volume(float* dst, const double* src, int64_t n, hipStream_t s) {
  std::vector<double> tmp((size_t)n);
  PNT_HIP_CHECK(hipMemcpyAsync(tmp.data(), src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  for (int64_t i = 0; i < n; ++i) dst[i] = (float)tmp[(size_t)i];
}

void download_planes(float* dst, int nplanes, float* const* src, int64_t n, int layout,
                     hipStream_t s) {
  if (nplanes == 1 || layout == PNT_PLANAR) {
    for (int c = 0; c < nplanes; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(dst + c * n, src[c], n * sizeof(float), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    PNT_REQUIRE(layout == PNT_AOS, "layout must be PNT_AOS or PNT_PLANAR");
    std::vector<float> tmp((size_t)(n * nplanes));
    for (int c = 0; c < nplanes; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(tmp.data() + c * n, src[c], n * sizeof(float),
                                   hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < nplanes; ++c) dst[i * nplanes + c] = tmp[(size_t)(c * n + i)];
  }
}

}  // namespace

namespace pnt {

int guard_call(const std::function<void()>& f) { return guard(f); }

float* dev_alloc(int64_t nfloats) {
  void* p = nullptr;
  if (nfloats <= 0) nfloats = 1;
  const hipError_t e = hipMalloc(&p, (size_t)nfloats * sizeof(float));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(PNT_ENOMEM, std::string("hipMalloc of ") + std::to_string(nfloats * 4) +
                                " bytes failed: " + hipGetErrorString(e));
  }
  return static_cast<float*>(p);
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}


// FootprintBrush::setRadius geometry (FootprintBrush.hxx:52-61) with an
// injected width x width footprint, padded with zeros to size_map^2.
int footprint_size_map(double radius, const float* core, int width) {
  PNT_REQUIRE(core != nullptr, "footprint is NULL");
  PNT_REQUIRE(std::isfinite(radius) && radius > 0.0 && radius < 4096.0,
              "radius must be in (0, 4096)");
  const int w = (int)(2.0 * std::ceil(radius) + 1.0);
  PNT_REQUIRE(w == width, "footprint width must be 2*ceil(radius)+1");
  return (int)std::ceil(std::sqrt(2.0) * w);
}

void pad_footprint_into(const float* core, int w, int size_map, float* dst, int box[4]) {
  const int pad = (size_map - w) / 2;
  box[0] = box[1] = size_map;
  box[2] = box[3] = -1;
  std::memset(dst, 0, sizeof(float) * (size_t)pad * size_map);
  for (int i = 0; i < w; ++i) {
    float* row = dst + (size_t)(i + pad) * size_map;
    const float* src = core + (size_t)i * w;
    std::memset(row, 0, sizeof(float) * (size_t)pad);
    for (int j = 0; j < w; ++j) {
      const float v = src[j];
      PNT_REQUIRE(std::isfinite(v), "footprint values must be finite");
      row[pad + j] = v;
      if (v > 0.0f) {
        box[0] = std::min(box[0], pad + j);
        box[2] = std::max(box[2], pad + j);
        box[1] = std::min(box[1], pad + i);
        box[3] = std::max(box[3], pad + i);
      }
    }
    std::memset(row + pad + w, 0, sizeof(float) * (size_t)(size_map - pad - w));
  }
  std::memset(dst + (size_t)(pad + w) * size_map, 0,
              sizeof(float) * (size_t)(size_map - pad - w) * size_map);
}

std::vector<float> padded_footprint(double radius, const float* core, int width, int* size_map_out) {
  const int size_map = footprint_size_map(radius, core, width);
  std::vector<float> full((size_t)size_map * size_map);
  int box[4];
  pad_footprint_into(core, width, size_map, full.data(), box);
  *size_map_out = size_map;
  return full;
}

void footprint_box(const std::vector<float>& fp, int size_map, int box[4]) {
  box[0] = box[1] = size_map;
  box[2] = box[3] = -1;
  for (int y = 0; y < size_map; ++y)
    for (int x = 0; x < size_map; ++x)
      if (fp[(size_t)y * size_map + x] > 0.0f) {
        box[0] = std::min(box[0], x);
        box[1] = std::min(box[1], y);
        box[2] = std::max(box[2], x);
        box[3] = std::max(box[3], y);
      }
}

void brush_set_geometry(pnt_brush b, double radius, const float* core, int width) {
  int size_map = 0;
  std::vector<float> full = padded_footprint(radius, core, width, &size_map);
  const int w = width;
  const int64_t m = (int64_t)size_map * size_map;
  const int64_t mstride = round_up(m, 64);
  hipStream_t s = b->ctx->stream;
  // buffers keep their capacity: a smaller brush reuses them (no free, which
  // would wait for the whole device)
  float* fp = b->fp;
  float* pick = b->pick;
  if (m > b->fp_cap || 7 * mstride > b->pick_cap) {
    fp = dev_alloc(m);
    try {
      pick = dev_alloc(7 * mstride);
    } catch (...) {
      dev_free(fp);
      throw;
    }
  }
  try {
    PNT_HIP_CHECK(hipMemcpyAsync(fp, full.data(), m * sizeof(float), hipMemcpyHostToDevice, s));
    launch_fill(pick, 7 * mstride, 0.0f, s);
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    if (fp != b->fp) {  // the new buffers are not the brush's yet
      dev_free(fp);
      dev_free(pick);
    }
    throw;
  }
  if (fp != b->fp) {
    dev_free(b->fp);
    dev_free(b->pick);
    b->fp_cap = m;
    b->pick_cap = 7 * mstride;
  }
  b->fp = fp;
  footprint_box(full, size_map, b->fbox);
  b->fp_host.swap(full);
  b->pick = pick;
  b->radius = radius;
  b->width = w;
  b->size_map = size_map;
  b->m = m;
  b->mstride = mstride;
}

// PaintLayer::copyTo into the snapshot (FootprintBrush.hxx:281-284, 168-172).
void brush_full_snapshot(pnt_brush b, pnt_canvas cv) {
  hipStream_t s = b->ctx->stream;
  if (b->snap_rows != cv->rows || b->snap_cols != cv->cols) {
    const int64_t stride = round_up(std::max<int64_t>(cv->n, 1), 64);
    if (kLayerStrides * stride > b->snap_cap) {  // grow; a smaller canvas reuses the buffer
      float* snap = dev_alloc(kLayerStrides * stride);
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->snap);
      b->snap = snap;
      b->snap_cap = kLayerStrides * stride;
    }
    b->snap_stride = stride;
    b->snap_rows = cv->rows;
    b->snap_cols = cv->cols;
  }
  launch_copy_layer(cv->n, cv->layer(), b->snap_planes(), s);
  if (b->dirty && round_up(cv->n, 16) + 16 <= b->dirty_cap) {
    b->dirty_n = cv->n;
    launch_fill_u8(b->dirty, cv->n, 0, s);
    b->dirty_canvas = cv;
    b->dirty_epoch = cv->epoch;
  } else {
    b->dirty_canvas = nullptr;
  }
}

StampArgs stamp_args(pnt_brush b, pnt_canvas cv, double cx, double cy, double theta) {
  PNT_REQUIRE(std::isfinite(cx) && std::isfinite(cy) && std::isfinite(theta) &&
                  std::fabs(cx) < 1e9 && std::fabs(cy) < 1e9,
              "stamp centre/angle must be finite and |centre| < 1e9");
  StampArgs a;
  a.g.cx = cx;
  a.g.cy = cy;
  a.g.c = std::cos(-theta);
  a.g.s = std::sin(-theta);
  a.g.hr = (b->size_map - 1) / 2;
  a.g.size_map = b->size_map;
  a.g.rows = cv->rows;
  a.g.cols = cv->cols;
  a.b = b->args();
  a.canvas = cv->layer();
  a.src = b->use_snap ? b->snap_planes() : cv->layer();
  a.dupx = stamp_dup_axis(cx, a.g.hr) ? 1 : 0;
  a.dupy = stamp_dup_axis(cy, a.g.hr) ? 1 : 0;
  a.clk = DryClock{};
  return a;
}

int64_t canvas_now_us(const pnt_canvas_s* cv) {
  if (cv->clock == PNT_CLOCK_MANUAL) return cv->manual_us;
  return std::chrono::duration_cast<std::chrono::microseconds>(
             std::chrono::system_clock::now().time_since_epoch())
      .count();
}

// The drying timer of the reference's GPU render thread
// (SbrRenderThread.cxx:37-41): every 250 ms, CanvasGpu::dryStep(0.01)
// (CanvasGpu.hxx:29).  Here it ticks on the canvas clock and is applied
// lazily: the steps whose ticks have passed since the last call run before
// the next call that reads or changes the canvas.  The step is per pixel, so
// k steps run as k iterations per pixel in one launch, cut short once no
// pixel holds paint (steps then change nothing).
void canvas_tick(pnt_canvas cv) {
  if (cv->dry_period_us <= 0 || cv->clock == PNT_CLOCK_OFF) return;
  const int64_t now = canvas_now_us(cv);
  if (now <= cv->dry_last_us) return;
  int64_t k = (now - cv->dry_last_us) / cv->dry_period_us;
  if (k <= 0) return;
  cv->dry_last_us += k * cv->dry_period_us;
  cv->ctx->use();
  hipStream_t s = cv->ctx->stream;
  cv->epoch++;
  if (!cv->dry_wet) cv->dry_wet = reinterpret_cast<unsigned*>(dev_alloc(1));
  constexpr int64_t kChunk = 4096;  // iterations per launch
  while (k > 0) {
    const int steps = (int)std::min<int64_t>(k, kChunk);
    unsigned wet = 0;
    PNT_HIP_CHECK(hipMemcpyAsync(cv->dry_wet, &wet, sizeof(unsigned), hipMemcpyHostToDevice, s));
    launch_dry_steps(cv->n, cv->layer(), cv->R0(), cv->dry_portion, steps, cv->dry_wet, s);
    PNT_HIP_CHECK(hipMemcpyAsync(&wet, cv->dry_wet, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    k -= steps;
    if (!wet) break;
  }
}

void canvas_stamp_time(pnt_canvas cv) {
  if (cv->tmap) launch_fill_i64(cv->tmap, cv->n, canvas_now_us(cv), cv->ctx->stream);
}

// Visits of a stamp that share a canvas pixel on column/row 0, plus their
// map-cell partners, in reference (row-major) order, with dependency levels:
// a visit's level is one more than the highest level of the earlier visits
// with its map cell or its canvas pixel.
//
// Canvas column 0 receives the two window columns k0, k0 + 1 with
// (int)(k + cx) == 0.  A visit's map-cell partner is a 4-neighbour; for a
// visit on those columns the partners off them can only be (k0 - 1, row) and
// (k0 + 2, row), so the visits come from the band of columns k0-1 .. k0+2
// (rows likewise).  The levels of cells and pixels are kept in flat
// per-thread arrays (pixels relative to the window's canvas box), cleared
// entry by entry after use.
void plan_shared_visits(const StampArgs& a, std::vector<int2>& list, std::vector<int>& off) {
  list.clear();
  off.clear();
  if (!a.dupx && !a.dupy) return;
  const StampGeom& g = a.g;
  const int hr = g.hr;
  const int side = 2 * hr + 1;
  struct Visit {
    int row, col, m;
    int64_t ic;
  };
  thread_local std::vector<Visit> vis, vis2;
  thread_local std::vector<int> lvl_cell, lvl_px, lvl, band;
  thread_local std::vector<int64_t> band_ic;
  vis.clear();
  // first window offset k in [-hr, hr - 1] with (int)(k + c) == (int)(k + 1 + c) == 0
  auto first_dup = [&](double c) {
    const double f = std::floor(-c);
    for (int d = -1; d <= 1; ++d) {
      const double kd = f + (double)d;
      if (kd < (double)-hr || kd > (double)(hr - 1)) continue;
      const int k = (int)kd;
      if ((int)(k + c) == 0 && (int)(k + 1 + c) == 0) return k;
    }
    return INT32_MIN;
  };
  // map cells of the band: 4 lines of `side` entries (line j = k0 - 1 + j)
  band.resize((size_t)4 * side);
  band_ic.resize((size_t)4 * side);
  auto fill_band = [&](int k0, bool columns) {
    for (int j = 0; j < 4; ++j)
      for (int t = 0; t < side; ++t) {
        const int line = k0 - 1 + j, along = t - hr;
        int64_t ic = 0;
        band[(size_t)j * side + t] = columns ? stamp_visit(g, line, along, &ic)
                                             : stamp_visit(g, along, line, &ic);
        band_ic[(size_t)j * side + t] = ic;
      }
  };
  auto keep = [&](int j, int t) {  // band entry j, t is on the dup lines or a partner
    const int m = band[(size_t)j * side + t];
    if (m < 0) return false;
    if (j == 1 || j == 2) return true;
    return m == band[(size_t)(j == 0 ? 1 : 2) * side + t];
  };
  if (a.dupx) {
    const int k0 = first_dup(g.cx);
    fill_band(k0, true);
    for (int t = 0; t < side; ++t)  // row-major: rows outer, band columns inner
      for (int j = 0; j < 4; ++j)
        if (keep(j, t))
          vis.push_back(Visit{t - hr, k0 - 1 + j, band[(size_t)j * side + t],
                              band_ic[(size_t)j * side + t]});
  }
  if (a.dupy) {
    const int k0 = first_dup(g.cy);
    fill_band(k0, false);
    vis2.clear();
    for (int j = 0; j < 4; ++j)
      for (int t = 0; t < side; ++t)
        if (keep(j, t))
          vis2.push_back(Visit{k0 - 1 + j, t - hr, band[(size_t)j * side + t],
                               band_ic[(size_t)j * side + t]});
    if (vis.empty()) {
      vis.swap(vis2);
    } else {  // both axes: merge the two row-major lists, dropping repeats
      const size_t n1 = vis.size();
      vis.insert(vis.end(), vis2.begin(), vis2.end());
      auto lt = [](const Visit& x, const Visit& y) {
        return x.row < y.row || (x.row == y.row && x.col < y.col);
      };
      std::inplace_merge(vis.begin(), vis.begin() + (std::ptrdiff_t)n1, vis.end(), lt);
      vis.erase(std::unique(vis.begin(), vis.end(),
                            [](const Visit& x, const Visit& y) {
                              return x.row == y.row && x.col == y.col;
                            }),
                vis.end());
    }
  }
  // levels; canvas pixels of window offsets lie in [x_lo, x_lo + span)^2
  const int x_lo = (int)(g.cx - hr) - 1, y_lo = (int)(g.cy - hr) - 1;
  const int span = 2 * hr + 4;
  const size_t ncell = (size_t)g.size_map * g.size_map;
  if (lvl_cell.size() < ncell) lvl_cell.assign(ncell, -1);
  if (lvl_px.size() < (size_t)span * span) lvl_px.assign((size_t)span * span, -1);
  auto px_slot = [&](int64_t ic) {
    return (size_t)((int)(ic / g.cols) - y_lo) * span + (size_t)((int)(ic % g.cols) - x_lo);
  };
  lvl.resize(vis.size());
  int maxl = -1;
  for (size_t i = 0; i < vis.size(); ++i) {
    const size_t p = px_slot(vis[i].ic);
    const int L = std::max(lvl_cell[(size_t)vis[i].m], lvl_px[p]) + 1;
    lvl_cell[(size_t)vis[i].m] = L;
    lvl_px[p] = L;
    lvl[i] = L;
    maxl = std::max(maxl, L);
  }
  for (const Visit& v : vis) {  // clear what this stamp set
    lvl_cell[(size_t)v.m] = -1;
    lvl_px[px_slot(v.ic)] = -1;
  }
  // counting sort by level, keeping reference order within a level
  off.assign((size_t)maxl + 2, 0);
  for (int L : lvl) off[(size_t)L + 1]++;
  for (size_t i = 1; i < off.size(); ++i) off[i] += off[i - 1];
  list.resize(vis.size());
  std::vector<int> pos(off.begin(), off.end() - 1);
  for (size_t i = 0; i < vis.size(); ++i)
    list[(size_t)pos[(size_t)lvl[i]]++] = make_int2(vis[i].col, vis[i].row);
}

}  // namespace pnt

namespace {

void brush_dip(pnt_brush b, const float K[3], const float S[3]) {
  launch_fill(b->pick, 7 * b->mstride, 0.0f, b->ctx->stream);
  for (int c = 0; c < 3; ++c) {
    b->paintK[c] = K[c];
    b->paintS[c] = S[c];
  }
}

// FootprintBrush::imprint (FootprintBrush.hxx:73-143) as kernel launches.
void imprint_plain(pnt_brush b, pnt_canvas cv, double cx, double cy, double theta) {
  hipStream_t s = b->ctx->stream;
  cv->epoch++;
  b->dirty_canvas = nullptr;
  if (b->use_snap) {
    if (b->snap_rows != cv->rows || b->snap_cols != cv->cols) brush_full_snapshot(b, cv);
  }
  StampArgs a = stamp_args(b, cv, cx, cy, theta);
  if (cv->tmap)  // Canvas::checkDry on every visit (FootprintBrush.hxx:86, 115)
    a.clk = DryClock{cv->tmap, canvas_now_us(cv), cv->drying_ms, cv->R0(), cv->h()};
  if (b->use_snap) launch_ring_refresh(ring_box(a.g, b->radius), cv->cols, cv->layer(), b->snap_planes(), s);
  launch_stamp_main(a, s);
  if (a.dupx || a.dupy) {
    std::vector<int2> list;
    std::vector<int> off;
    plan_shared_visits(a, list, off);
    if (!list.empty()) {
      const int64_t need = (int64_t)list.size();
      if (need > b->d_cap || (int64_t)off.size() > b->d_cap + 1) {
        PNT_HIP_CHECK(hipStreamSynchronize(s));
        dev_free(b->d_list);
        dev_free(b->d_off);
        b->d_list = nullptr;
        b->d_off = nullptr;
        const int64_t cap = std::max<int64_t>(need, (int64_t)off.size()) * 2;
        PNT_HIP_CHECK(hipMalloc(&b->d_list, cap * sizeof(int2)));
        PNT_HIP_CHECK(hipMalloc(&b->d_off, (cap + 1) * sizeof(int)));
        b->d_cap = cap;
      }
      PNT_HIP_CHECK(hipMemcpyAsync(b->d_list, list.data(), need * sizeof(int2),
                                   hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(b->d_off, off.data(), off.size() * sizeof(int),
                                   hipMemcpyHostToDevice, s));
      launch_stamp_serial(a, b->d_list, b->d_off, (int)off.size() - 1, s);
      PNT_HIP_CHECK(hipStreamSynchronize(s));  // host lists die here
    }
  }
}

void check_layout(int layout) {
  PNT_REQUIRE(layout == PNT_AOS || layout == PNT_PLANAR, "layout must be PNT_AOS or PNT_PLANAR");
}

void check_same_ctx(pnt_brush b, pnt_canvas cv) {
  PNT_REQUIRE(b && cv, "NULL handle");
  PNT_REQUIRE(b->ctx->device == cv->ctx->device, "brush and canvas live on different devices");
}

// painty/core/Spline.hxx:27-45 and :47-73 (tau = 0.5), FP64 without contraction.
double catmull_rom(double p_1, double p0, double p1, double p2, double t) {
#pragma clang fp contract(off)
  const double tau = 0.5;
  const double t2 = t * t;
  const double t3 = t2 * t;
  const double b1 = -tau * t + 2.0 * tau * t2 - tau * t3;
  const double b2 = 1.0 + (tau - 3.0) * t2 + (2.0 - tau) * t3;
  const double b3 = tau * t + (3.0 - 2.0 * tau) * t2 + (tau - 2.0) * t3;
  const double b4 = -tau * t2 + tau * t3;
  return (p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4);
}

double catmull_rom_d1(double p_1, double p0, double p1, double p2, double t) {
#pragma clang fp contract(off)
  const double tau = 0.5;
  const double t2 = t * t;
  const double b1 = tau * (-3.0 * t2 + 4.0 * t - 1.0);
  const double b2 = -t * (-2.0 * tau + 3.0 * (tau - 2.0) * t + 6.0);
  const double b3 = (t - 1.0) * (3.0 * (tau - 2.0) * t - tau);
  const double b4 = tau * t * (3.0 * t - 2.0);
  return (p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4);
}

// PaintMixer::mixSinglePaint, mixer/src/PaintMixer.cxx:327-355 (Eps = 10e-6, :278).
void mix_single(int n, const double* bK, const double* bS, const double* w, double K[3],
                double S[3]) {
#pragma clang fp contract(off)
  double norm = 1.0, wsum = 0.0;
  for (int i = 0; i < n; ++i) wsum += w[i];
  if (!(std::fabs(wsum - 1.0) < 10e-6)) norm = 1. / wsum;
  double k[3] = {0, 0, 0}, s[3] = {0, 0, 0};
  for (int l = 0; l < n; ++l)
    for (int c = 0; c < 3; ++c) {
      k[c] += norm * w[l] * bK[3 * l + c];
      s[c] += norm * w[l] * bS[3 * l + c];
    }
  for (int c = 0; c < 3; ++c) {
    K[c] = k[c];
    S[c] = s[c];
  }
}

// PaintMixer::mixed, PaintMixer.cxx:536-548.
void mixed(const double K1[3], const double S1[3], double v1, const double K2[3],
           const double S2[3], double v2, double K[3], double S[3]) {
#pragma clang fp contract(off)
  const double inv = 1.0 / (v1 + v2);
  for (int c = 0; c < 3; ++c) {
    const double k = ((v1 * K1[c]) + (v2 * K2[c])) * inv;
    const double s = ((v1 * S1[c]) + (v2 * S2[c])) * inv;
    K[c] = k;
    S[c] = s;
  }
}

}  // namespace

// ================================================================== C ABI ==
extern "C" {

const char* pnt_last_error(void) { return g_err; }

int pnt_version(void) { return 100; }

}  // extern "C"

namespace pnt {
void ctx_retain(pnt_ctx ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }

void ctx_release(pnt_ctx ctx) {
  if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  batch_release(ctx);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->kernel_error) (void)hipHostFree(ctx->kernel_error);
  delete ctx;
}

// A context whose stream has a hardware queue of its own: a stream made with
// a CU mask (every CU of the device) gets its own queue instead of sharing
// one of the process's pool (GPU_MAX_HW_QUEUES, 4 by default) with the
// streams made before it.  For streams whose kernels must not queue behind
// another stream's (the server's wave lanes).
pnt_ctx ctx_create_own_queue(int device) {
  auto* c = new pnt_ctx_s;
  c->device = device;
  try {
    c->use();
    int cus = 0;
    PNT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
    for (int i = 0; i < cus; ++i) mask[(size_t)i / 32] |= 1u << (i % 32);
    PNT_HIP_CHECK(hipExtStreamCreateWithCUMask(&c->stream, (uint32_t)mask.size(), mask.data()));
    PNT_HIP_CHECK(hipHostMalloc((void**)&c->kernel_error, sizeof(unsigned), hipHostMallocMapped));
    *c->kernel_error = 0;
  } catch (...) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    throw;
  }
  return c;
}
}  // namespace pnt

extern "C" {

int pnt_ctx_create(int device, pnt_ctx* out) {
  return guard([&] {
    PNT_REQUIRE(out != nullptr, "out is NULL");
    int ndev = 0;
    PNT_HIP_CHECK(hipGetDeviceCount(&ndev));
    PNT_REQUIRE(device >= 0 && device < ndev, "no such HIP device");
    auto* c = new pnt_ctx_s;
    c->device = device;
    try {
      c->use();
      PNT_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      PNT_HIP_CHECK(hipHostMalloc((void**)&c->kernel_error, sizeof(unsigned), hipHostMallocMapped));
      *c->kernel_error = 0;
    } catch (...) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
  });
}

int pnt_ctx_destroy(pnt_ctx ctx) {
  return guard([&] {
    if (ctx) ctx_release(ctx);
  });
}

int pnt_ctx_synchronize(pnt_ctx ctx) {
  return guard([&] {
    PNT_REQUIRE(ctx, "NULL context");
    ctx->use();
    PNT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (*(volatile unsigned*)ctx->kernel_error) {
      *ctx->kernel_error = 0;
      throw Error(PNT_EKERNEL, "a persistent stamp kernel aborted (grid barrier wait expired); "
                               "canvas and brush state are undefined");
    }
  });
}

void* pnt_ctx_stream(pnt_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ------------------------------------------------------------- PaintLayer --
int pnt_layer_create(pnt_ctx ctx, int rows, int cols, pnt_layer* out) {
  return guard([&] {
    PNT_REQUIRE(ctx && out, "NULL argument");
    PNT_REQUIRE(rows >= 0 && cols >= 0 && (int64_t)rows * cols < ((int64_t)1 << 31),
                "rows/cols out of range");
    ctx->use();
    auto* L = new pnt_layer_s;
    L->ctx = ctx;
    try {
      L->alloc(rows, cols);
      launch_fill(L->base, kLayerStrides * L->stride, 0.0f, ctx->stream);  // zero bits: 0.0 too
      PNT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      dev_free(L->base);
      delete L;
      throw;
    }
    ctx_retain(ctx);
    *out = L;
  });
}

int pnt_layer_destroy(pnt_layer L) {
  return guard([&] {
    if (!L) return;
    if (L->owner) {
      pnt_ctx ctx = L->ctx;
      (void)hipSetDevice(ctx->device);
      (void)hipStreamSynchronize(ctx->stream);
      dev_free(L->base);
      delete L;
      ctx_release(ctx);
    }
  });
}

int pnt_layer_shape(pnt_layer L, int* rows, int* cols) {
  return guard([&] {
    PNT_REQUIRE(L && rows && cols, "NULL argument");
    *rows = L->rows;
    *cols = L->cols;
  });
}

int pnt_layer_clear(pnt_layer L) {
  return guard([&] {
    PNT_REQUIRE(L, "NULL layer");
    L->ctx->use();
    L->touched();
    launch_fill(L->base, kLayerStrides * L->stride, 0.0f, L->ctx->stream);
  });
}

int pnt_layer_upload(pnt_layer L, const float* K, const float* S, const float* V, int layout) {
  return guard([&] {
    PNT_REQUIRE(L, "NULL layer");
    check_layout(layout);
    L->ctx->use();
    if (L->parent) canvas_tick(L->parent);
    L->touched();
    const Layer7 P = L->planes();
    if (K) upload_planes(P.ks, 3, K, L->n, layout, L->ctx->stream);
    if (S) upload_planes(P.ks + 3, 3, S, L->n, layout, L->ctx->stream);
    if (V) upload_volume(P.v, V, L->n, L->ctx->stream);
  });
}

int pnt_layer_download(pnt_layer L, float* K, float* S, float* V, int layout) {
  return guard([&] {
    PNT_REQUIRE(L, "NULL layer");
    check_layout(layout);
    L->ctx->use();
    if (L->parent) canvas_tick(L->parent);
    const Layer7 P = L->planes();
    if (K) download_planes(K, 3, P.ks, L->n, layout, L->ctx->stream);
    if (S) download_planes(S, 3, P.ks + 3, L->n, layout, L->ctx->stream);
    if (V) download_volume(V, P.v, L->n, L->ctx->stream);
  });
}

int pnt_layer_upload_volume(pnt_layer L, const double* V) {
  return guard([&] {
    PNT_REQUIRE(L && V, "NULL argument");
    L->ctx->use();
    if (L->parent) canvas_tick(L->parent);
    L->touched();
    PNT_HIP_CHECK(hipMemcpyAsync(L->planes().v, V, (size_t)L->n * sizeof(double),
                                 hipMemcpyHostToDevice, L->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(L->ctx->stream));
  });
}

int pnt_layer_download_volume(pnt_layer L, double* V) {
  return guard([&] {
    PNT_REQUIRE(L && V, "NULL argument");
    L->ctx->use();
    if (L->parent) canvas_tick(L->parent);
    PNT_HIP_CHECK(hipMemcpyAsync(V, L->planes().v, (size_t)L->n * sizeof(double),
                                 hipMemcpyDeviceToHost, L->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(L->ctx->stream));
  });
}

int pnt_layer_copy_to(pnt_layer src, pnt_layer dst) {
  return guard([&] {
    PNT_REQUIRE(src && dst, "NULL layer");
    PNT_REQUIRE(src->ctx->device == dst->ctx->device, "layers on different devices");
    src->ctx->use();
    dst->touched();
    hipStream_t s = src->ctx->stream;
    if (dst->rows != src->rows || dst->cols != src->cols) {
      PNT_REQUIRE(dst->owner, "cannot resize the paint layer of a canvas");
      PNT_HIP_CHECK(hipStreamSynchronize(dst->ctx->stream));
      float* nb = dev_alloc(kLayerStrides * round_up(std::max<int64_t>(src->n, 1), 64));
      dev_free(dst->base);
      dst->base = nb;
      dst->rows = src->rows;
      dst->cols = src->cols;
      dst->n = src->n;
      dst->stride = round_up(std::max<int64_t>(src->n, 1), 64);
    }
    if (dst->ctx->stream != s) PNT_HIP_CHECK(hipStreamSynchronize(dst->ctx->stream));
    launch_copy_layer(src->n, src->planes(), dst->planes(), s);
    if (dst->ctx->stream != s) PNT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int pnt_layer_set(pnt_layer L, int i, int j, const float k[3], const float s[3], float v) {
  return guard([&] {
    PNT_REQUIRE(L && k && s, "NULL argument");
    PNT_REQUIRE(i >= 0 && i < L->rows && j >= 0 && j < L->cols, "index out of range");
    L->ctx->use();
    L->touched();
    const int64_t idx = (int64_t)i * L->cols + j;
    const float vals[6] = {k[0], k[1], k[2], s[0], s[1], s[2]};
    const double vd = v;
    const Layer7 P = L->planes();
    for (int c = 0; c < 6; ++c)
      PNT_HIP_CHECK(hipMemcpyAsync(P.ks[c] + idx, &vals[c], sizeof(float), hipMemcpyHostToDevice,
                                   L->ctx->stream));
    PNT_HIP_CHECK(hipMemcpyAsync(P.v + idx, &vd, sizeof(double), hipMemcpyHostToDevice,
                                 L->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(L->ctx->stream));
  });
}

int pnt_layer_compose_onto_host(pnt_layer L, float* R0, int r0_rows, int r0_cols, int layout) {
  return guard([&] {
    PNT_REQUIRE(L && R0, "NULL argument");
    check_layout(layout);
    L->ctx->use();
    hipStream_t s = L->ctx->stream;
    const int64_t n = L->n;
    if (r0_rows != L->rows || r0_cols != L->cols)  // PaintLayer.hxx:82-87
      for (int64_t i = 0; i < 3 * n; ++i) R0[i] = 1.0f;
    const int64_t stride = round_up(std::max<int64_t>(n, 1), 64);
    float* buf = dev_alloc(3 * stride);
    try {
      Rgb3 r{{buf, buf + stride, buf + 2 * stride}};
      upload_planes(r.p, 3, R0, n, layout, s);
      launch_compose(n, L->planes(), r, r, s);
      download_planes(R0, 3, r.p, n, layout, s);
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dev_free(buf);
      throw;
    }
    dev_free(buf);
  });
}

int pnt_layer_compose_onto_canvas(pnt_layer L, pnt_canvas cv) {
  return guard([&] {
    PNT_REQUIRE(L && cv, "NULL argument");
    PNT_REQUIRE(L->rows == cv->rows && L->cols == cv->cols, "layer and canvas sizes differ");
    PNT_REQUIRE(L->ctx->device == cv->ctx->device, "layer and canvas on different devices");
    cv->ctx->use();
    canvas_tick(cv);
    if (L->ctx->stream != cv->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(L->ctx->stream));
    launch_compose(L->n, L->planes(), cv->R0(), cv->R0(), cv->ctx->stream);
  });
}

// ----------------------------------------------------------------- Canvas --
static void canvas_clear(pnt_canvas cv) {
  hipStream_t s = cv->ctx->stream;
  cv->epoch++;
  launch_fill(cv->base, kLayerStrides * cv->stride, 0.0f, s);  // zero bits: V = 0.0 too
  for (int c = 0; c < 3; ++c) launch_fill(cv->R0().p[c], cv->n, cv->bg[c], s);
  launch_fill(cv->h(), cv->n, 0.0f, s);
  canvas_stamp_time(cv);  // Canvas.hxx:47-50
}

}  // extern "C"

namespace pnt {
// A canvas on a caller-provided buffer of at least
// kCanvasStrides*round_up(rows*cols, 64) floats, which pnt_canvas_destroy
// leaves alone.
pnt_canvas canvas_over(pnt_ctx ctx, int rows, int cols, float* base, bool clear) {
  PNT_REQUIRE(ctx && base, "NULL argument");
  PNT_REQUIRE(rows > 0 && cols > 0 && (int64_t)rows * cols < ((int64_t)1 << 31),
              "rows/cols out of range");
  ctx->use();
  auto* cv = new pnt_canvas_s;
  cv->ctx = ctx;
  cv->rows = rows;
  cv->cols = cols;
  cv->n = (int64_t)rows * cols;
  cv->stride = round_up(cv->n, 64);
  cv->base = base;
  cv->owns_base = false;
  cv->view.ctx = ctx;
  cv->view.rows = rows;
  cv->view.cols = cols;
  cv->view.n = cv->n;
  cv->view.stride = cv->stride;
  cv->view.base = cv->base;
  cv->view.owner = false;
  cv->view.parent = cv;
  try {
    if (clear) canvas_clear(cv);
  } catch (...) {
    delete cv;
    throw;
  }
  ctx_retain(ctx);
  return cv;
}
}  // namespace pnt

extern "C" {

int pnt_canvas_create(pnt_ctx ctx, int rows, int cols, pnt_canvas* out) {
  return guard([&] {
    PNT_REQUIRE(ctx && out, "NULL argument");
    PNT_REQUIRE(rows > 0 && cols > 0 && (int64_t)rows * cols < ((int64_t)1 << 31),
                "rows/cols out of range");
    ctx->use();
    auto* cv = new pnt_canvas_s;
    cv->ctx = ctx;
    cv->rows = rows;
    cv->cols = cols;
    cv->n = (int64_t)rows * cols;
    cv->stride = round_up(cv->n, 64);
    try {
      cv->base = dev_alloc(kCanvasStrides * cv->stride);
      canvas_clear(cv);
      PNT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      dev_free(cv->base);
      delete cv;
      throw;
    }
    cv->view.ctx = ctx;
    cv->view.rows = rows;
    cv->view.cols = cols;
    cv->view.n = cv->n;
    cv->view.stride = cv->stride;
    cv->view.base = cv->base;
    cv->view.owner = false;
    cv->view.parent = cv;
    ctx_retain(ctx);
    *out = cv;
  });
}

int pnt_canvas_destroy(pnt_canvas cv) {
  return guard([&] {
    if (!cv) return;
    pnt_ctx ctx = cv->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (cv->owns_base) dev_free(cv->base);
    dev_free(cv->scratch);
    dev_free(cv->tmap);
    dev_free(cv->dry_wet);
    dev_free(cv->light);
    delete cv;
    ctx_release(ctx);
  });
}

int pnt_canvas_shape(pnt_canvas cv, int* rows, int* cols) {
  return guard([&] {
    PNT_REQUIRE(cv && rows && cols, "NULL argument");
    *rows = cv->rows;
    *cols = cv->cols;
  });
}

int pnt_canvas_layer(pnt_canvas cv, pnt_layer* out) {
  return guard([&] {
    PNT_REQUIRE(cv && out, "NULL argument");
    *out = &cv->view;
  });
}

int pnt_canvas_clear(pnt_canvas cv) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    cv->ctx->use();
    canvas_tick(cv);
    canvas_clear(cv);
  });
}

int pnt_canvas_set_background(pnt_canvas cv, const float* R0, int layout) {
  return guard([&] {
    PNT_REQUIRE(cv && R0, "NULL argument");
    check_layout(layout);
    cv->ctx->use();
    canvas_tick(cv);
    canvas_clear(cv);
    Rgb3 r = cv->R0();
    upload_planes(r.p, 3, R0, cv->n, layout, cv->ctx->stream);
  });
}

int pnt_canvas_upload_R0(pnt_canvas cv, const float* R0, int layout) {
  return guard([&] {
    PNT_REQUIRE(cv && R0, "NULL argument");
    check_layout(layout);
    cv->ctx->use();
    canvas_tick(cv);
    Rgb3 r = cv->R0();
    upload_planes(r.p, 3, R0, cv->n, layout, cv->ctx->stream);
  });
}

int pnt_canvas_download_R0(pnt_canvas cv, float* R0, int layout) {
  return guard([&] {
    PNT_REQUIRE(cv && R0, "NULL argument");
    check_layout(layout);
    cv->ctx->use();
    canvas_tick(cv);
    Rgb3 r = cv->R0();
    download_planes(R0, 3, r.p, cv->n, layout, cv->ctx->stream);
  });
}

int pnt_canvas_upload_h(pnt_canvas cv, const float* h) {
  return guard([&] {
    PNT_REQUIRE(cv && h, "NULL argument");
    cv->ctx->use();
    canvas_tick(cv);
    float* p = cv->h();
    upload_planes(&p, 1, h, cv->n, PNT_PLANAR, cv->ctx->stream);
  });
}

int pnt_canvas_download_h(pnt_canvas cv, float* h) {
  return guard([&] {
    PNT_REQUIRE(cv && h, "NULL argument");
    cv->ctx->use();
    canvas_tick(cv);
    float* p = cv->h();
    download_planes(h, 1, &p, cv->n, PNT_PLANAR, cv->ctx->stream);
  });
}

int pnt_canvas_dry(pnt_canvas cv) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    cv->ctx->use();
    canvas_tick(cv);
    cv->epoch++;
    launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), cv->ctx->stream);
    canvas_stamp_time(cv);  // Canvas.hxx:106, 118
  });
}

int pnt_canvas_dry_step(pnt_canvas cv, float portion) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    PNT_REQUIRE(std::isfinite(portion), "portion must be finite");
    cv->ctx->use();
    canvas_tick(cv);
    cv->epoch++;
    launch_dry_step(cv->n, cv->layer(), cv->R0(), portion, cv->ctx->stream);
  });
}

int pnt_canvas_set_dry_timer(pnt_canvas cv, int64_t period_ms, float portion) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    PNT_REQUIRE(period_ms >= 0, "timer period must be >= 0");
    PNT_REQUIRE(period_ms == 0 || (std::isfinite(portion) && portion > 0.0f),
                "dry portion must be positive");
    canvas_tick(cv);  // steps of the previous timer up to now
    cv->dry_period_us = period_ms * 1000;
    cv->dry_portion = portion;
    cv->dry_last_us = canvas_now_us(cv);
  });
}

int pnt_canvas_set_drying_time_ms(pnt_canvas cv, int64_t ms) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    PNT_REQUIRE(ms >= 0, "drying time must be >= 0");
    cv->drying_ms = ms;
  });
}

int pnt_canvas_set_clock(pnt_canvas cv, int clock) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    PNT_REQUIRE(clock == PNT_CLOCK_OFF || clock == PNT_CLOCK_SYSTEM || clock == PNT_CLOCK_MANUAL,
                "unknown clock");
    cv->ctx->use();
    if (clock == PNT_CLOCK_OFF) {
      if (cv->tmap) {
        PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
        dev_free(cv->tmap);
        cv->tmap = nullptr;
      }
    } else if (!cv->tmap) {
      cv->tmap = reinterpret_cast<int64_t*>(dev_alloc(2 * cv->n));
      cv->clock = clock;
      canvas_stamp_time(cv);  // every pixel starts at now, as at construction
    }
    cv->clock = clock;
  });
}

int pnt_canvas_check_dry(pnt_canvas cv, int x, int y, int64_t now_us) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    PNT_REQUIRE(x >= 0 && y >= 0 && x < cv->cols && y < cv->rows, "pixel outside the canvas");
    PNT_REQUIRE(cv->tmap, "the canvas has no clock (pnt_canvas_set_clock)");
    cv->ctx->use();
    canvas_tick(cv);
    cv->epoch++;
    const DryClock c{cv->tmap, now_us, cv->drying_ms, cv->R0(), cv->h()};
    launch_check_dry_pixel(c, cv->layer(), (int64_t)y * cv->cols + x, cv->ctx->stream);
  });
}

int pnt_canvas_get_clock(pnt_canvas cv, int* clock) {
  return guard([&] {
    PNT_REQUIRE(cv && clock, "NULL argument");
    *clock = cv->clock;
  });
}

int pnt_canvas_set_time_us(pnt_canvas cv, int64_t now_us) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    cv->manual_us = now_us;
  });
}

int pnt_canvas_upload_time_map(pnt_canvas cv, const int64_t* tmap_us) {
  return guard([&] {
    PNT_REQUIRE(cv && tmap_us, "NULL argument");
    PNT_REQUIRE(cv->tmap, "the canvas has no clock (pnt_canvas_set_clock)");
    cv->ctx->use();
    PNT_HIP_CHECK(hipMemcpyAsync(cv->tmap, tmap_us, cv->n * sizeof(int64_t),
                                 hipMemcpyHostToDevice, cv->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
  });
}

int pnt_canvas_download_time_map(pnt_canvas cv, int64_t* tmap_us) {
  return guard([&] {
    PNT_REQUIRE(cv && tmap_us, "NULL argument");
    PNT_REQUIRE(cv->tmap, "the canvas has no clock (pnt_canvas_set_clock)");
    cv->ctx->use();
    PNT_HIP_CHECK(hipMemcpyAsync(tmap_us, cv->tmap, cv->n * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, cv->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
  });
}

int pnt_canvas_get_drying_time_ms(pnt_canvas cv, int64_t* ms) {
  return guard([&] {
    PNT_REQUIRE(cv && ms, "NULL argument");
    *ms = cv->drying_ms;
  });
}

int pnt_canvas_device_planes(pnt_canvas cv, pnt_dev_wet* layer, pnt_dev_rgb* R0, float** h) {
  return guard([&] {
    PNT_REQUIRE(cv, "NULL canvas");
    canvas_tick(cv);
    cv->epoch++;  // the caller may write the planes
    if (layer) {
      const Layer7 L = cv->layer();
      for (int i = 0; i < 6; ++i) layer->ks[i] = L.ks[i];
      layer->v = L.v;
    }
    if (R0) {
      const Rgb3 r = cv->R0();
      for (int i = 0; i < 3; ++i) R0->p[i] = r.p[i];
    }
    if (h) *h = cv->h();
  });
}

// --------------------------------------------------------------- Renderer --
int pnt_compose(pnt_canvas cv, float* R, int layout) {
  return guard([&] {
    PNT_REQUIRE(cv && R, "NULL argument");
    check_layout(layout);
    cv->ctx->use();
    canvas_tick(cv);
    if (!cv->scratch) cv->scratch = dev_alloc(3 * cv->stride);
    Rgb3 out{{cv->scratch, cv->scratch + cv->stride, cv->scratch + 2 * cv->stride}};
    launch_compose(cv->n, cv->layer(), cv->R0(), out, cv->ctx->stream);
    download_planes(R, 3, out.p, cv->n, layout, cv->ctx->stream);
  });
}

int pnt_render(pnt_canvas cv, float* R, int layout) {
  return guard([&] {
    PNT_REQUIRE(cv && R, "NULL argument");
    check_layout(layout);
    cv->ctx->use();
    canvas_tick(cv);
    if (!cv->scratch) cv->scratch = dev_alloc(3 * cv->stride);
    Rgb3 out{{cv->scratch, cv->scratch + cv->stride, cv->scratch + 2 * cv->stride}};
    launch_render(cv->rows, cv->cols, cv->layer(), cv->R0(), out, cv->ctx->stream);
    download_planes(R, 3, out.p, cv->n, layout, cv->ctx->stream);
  });
}

int pnt_dev_render(pnt_canvas cv, pnt_dev_rgb R) {
  return guard([&] {
    PNT_REQUIRE(cv && R.p[0] && R.p[1] && R.p[2], "NULL argument");
    cv->ctx->use();
    canvas_tick(cv);
    Rgb3 out{{R.p[0], R.p[1], R.p[2]}};
    launch_render(cv->rows, cv->cols, cv->layer(), cv->R0(), out, cv->ctx->stream);
  });
}

int pnt_dev_encode(pnt_ctx ctx, int64_t n, pnt_dev_rgb rgb, int to_srgb, int bits, void* out,
                   void* stream) {
  return guard([&] {
    PNT_REQUIRE(ctx && out && rgb.p[0] && rgb.p[1] && rgb.p[2], "NULL argument");
    PNT_REQUIRE(n >= 0, "n must be >= 0");
    PNT_REQUIRE(bits == 8 || bits == 16, "bits must be 8 or 16");
    ctx->use();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    launch_encode(n, Rgb3{{rgb.p[0], rgb.p[1], rgb.p[2]}}, to_srgb ? 1 : 0, bits, out, s);
  });
}

int pnt_encode(pnt_ctx ctx, int64_t n, const float* rgb, int layout, int to_srgb, int bits,
               void* out) {
  return guard([&] {
    PNT_REQUIRE(ctx && rgb && out, "NULL argument");
    PNT_REQUIRE(n >= 0, "n must be >= 0");
    PNT_REQUIRE(bits == 8 || bits == 16, "bits must be 8 or 16");
    check_layout(layout);
    ctx->use();
    hipStream_t s = ctx->stream;
    const int64_t stride = round_up(std::max<int64_t>(n, 1), 64);
    const size_t out_bytes = (size_t)(3 * n) * (bits == 16 ? 2 : 1);
    float* buf = dev_alloc(3 * stride + (int64_t)(out_bytes + 3) / 4);
    try {
      Rgb3 in{{buf, buf + stride, buf + 2 * stride}};
      upload_planes(in.p, 3, rgb, n, layout, s);
      void* dout = buf + 3 * stride;
      launch_encode(n, in, to_srgb ? 1 : 0, bits, dout, s);
      PNT_HIP_CHECK(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dev_free(buf);
      throw;
    }
    dev_free(buf);
  });
}

int pnt_compose_layer_host(pnt_layer L, const float* R0, float* R, int layout) {
  return guard([&] {
    PNT_REQUIRE(L && R0 && R, "NULL argument");
    check_layout(layout);
    L->ctx->use();
    hipStream_t s = L->ctx->stream;
    const int64_t n = L->n;
    const int64_t stride = round_up(std::max<int64_t>(n, 1), 64);
    float* buf = dev_alloc(6 * stride);
    try {
      Rgb3 r0{{buf, buf + stride, buf + 2 * stride}};
      Rgb3 r1{{buf + 3 * stride, buf + 4 * stride, buf + 5 * stride}};
      upload_planes(r0.p, 3, R0, n, layout, s);
      launch_compose(n, L->planes(), r0, r1, s);
      download_planes(R, 3, r1.p, n, layout, s);
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dev_free(buf);
      throw;
    }
    dev_free(buf);
  });
}

int pnt_compute_reflectance(pnt_ctx ctx, int64_t n, const float* K, const float* S,
                            const float* R0, const float* d, float* R) {
  return guard([&] {
    PNT_REQUIRE(ctx && K && S && R0 && d && R, "NULL argument");
    PNT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return;
    ctx->use();
    hipStream_t s = ctx->stream;
    float* buf = dev_alloc(13 * n);
    try {
      PNT_HIP_CHECK(hipMemcpyAsync(buf, K, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(buf + 3 * n, S, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(buf + 6 * n, R0, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(buf + 9 * n, d, n * sizeof(float), hipMemcpyHostToDevice, s));
      launch_reflectance_aos(n, buf, buf + 3 * n, buf + 6 * n, buf + 9 * n, buf + 10 * n, s);
      PNT_HIP_CHECK(hipMemcpyAsync(R, buf + 10 * n, 3 * n * sizeof(float), hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dev_free(buf);
      throw;
    }
    dev_free(buf);
  });
}

static Plane7 to7(const pnt_dev_layer& l) {
  Plane7 L;
  for (int i = 0; i < 7; ++i) {
    PNT_REQUIRE(l.p[i] != nullptr, "NULL device plane");
    L.p[i] = l.p[i];
  }
  return L;
}
static Layer7 to_wet(const pnt_dev_wet& l) {
  Layer7 L;
  for (int i = 0; i < 6; ++i) {
    PNT_REQUIRE(l.ks[i] != nullptr, "NULL device plane");
    L.ks[i] = l.ks[i];
  }
  PNT_REQUIRE(l.v != nullptr, "NULL device plane");
  L.v = l.v;
  return L;
}
static Rgb3 to3(const pnt_dev_rgb& r) {
  Rgb3 R;
  for (int i = 0; i < 3; ++i) {
    PNT_REQUIRE(r.p[i] != nullptr, "NULL device plane");
    R.p[i] = r.p[i];
  }
  return R;
}

int pnt_dev_compose(pnt_ctx ctx, int64_t n, pnt_dev_layer layer, pnt_dev_rgb R0, pnt_dev_rgb R,
                    void* stream) {
  return guard([&] {
    PNT_REQUIRE(ctx && n >= 0, "invalid argument");
    ctx->use();
    launch_compose(n, to7(layer), to3(R0), to3(R),
                   stream ? (hipStream_t)stream : ctx->stream);
  });
}

int pnt_dev_compose2(pnt_ctx ctx, int64_t n, pnt_dev_layer dry, pnt_dev_layer wet,
                     pnt_dev_rgb R0, pnt_dev_rgb R, void* stream) {
  return guard([&] {
    PNT_REQUIRE(ctx && n >= 0, "invalid argument");
    ctx->use();
    launch_compose2(n, to7(dry), to7(wet), to3(R0), to3(R),
                    stream ? (hipStream_t)stream : ctx->stream);
  });
}

int pnt_dev_compose_wet(pnt_ctx ctx, int64_t n, pnt_dev_wet layer, pnt_dev_rgb R0, pnt_dev_rgb R,
                        void* stream) {
  return guard([&] {
    PNT_REQUIRE(ctx && n >= 0, "invalid argument");
    ctx->use();
    launch_compose(n, to_wet(layer), to3(R0), to3(R), stream ? (hipStream_t)stream : ctx->stream);
  });
}

int pnt_dev_compose2_wet(pnt_ctx ctx, int64_t n, pnt_dev_wet dry, pnt_dev_wet wet,
                         pnt_dev_rgb R0, pnt_dev_rgb R, void* stream) {
  return guard([&] {
    PNT_REQUIRE(ctx && n >= 0, "invalid argument");
    ctx->use();
    launch_compose2(n, to_wet(dry), to_wet(wet), to3(R0), to3(R),
                    stream ? (hipStream_t)stream : ctx->stream);
  });
}

// ---------------------------------------------------------- FootprintBrush --
int pnt_brush_create(pnt_ctx ctx, double radius, const float* footprint, int width,
                     pnt_brush* out) {
  return guard([&] {
    PNT_REQUIRE(ctx && out, "NULL argument");
    ctx->use();
    auto* b = new pnt_brush_s;
    b->ctx = ctx;
    try {
      brush_set_geometry(b, radius, footprint, width);
    } catch (...) {
      delete b;
      throw;
    }
    ctx_retain(ctx);
    *out = b;
  });
}

int pnt_brush_destroy(pnt_brush b) {
  return guard([&] {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    persistent_release(b->pers);
    chain_release(b->chain);
    fused_release(b->fused);
    ktimer_release(b->ktimer);
    dev_free(b->fp);
    dev_free(b->pick);
    dev_free(b->snap);
    dev_free(b->d_list);
    dev_free(b->d_off);
    dev_free(b->dirty);
    pnt_ctx ctx = b->ctx;
    delete b;
    ctx_release(ctx);
  });
}

int pnt_brush_set_radius(pnt_brush b, double radius, const float* footprint, int width) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    if (std::fabs(b->radius - radius) < 0.5) return;  // fuzzyCompare(_radius, radius, 0.5)
    b->ctx->use();
    PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    brush_set_geometry(b, radius, footprint, width);
  });
}

int pnt_brush_size_map(pnt_brush b, int* size_map) {
  return guard([&] {
    PNT_REQUIRE(b && size_map, "NULL argument");
    *size_map = b->size_map;
  });
}

int pnt_brush_get_footprint(pnt_brush b, float* out) {
  return guard([&] {
    PNT_REQUIRE(b && out, "NULL argument");
    b->ctx->use();
    PNT_HIP_CHECK(hipMemcpyAsync(out, b->fp, b->m * sizeof(float), hipMemcpyDeviceToHost,
                                 b->ctx->stream));
    PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pnt_brush_get_pickup_map(pnt_brush b, float* K, float* S, float* V, int layout) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    check_layout(layout);
    b->ctx->use();
    const Plane7 P = b->pick_planes();
    if (K) download_planes(K, 3, P.p, b->m, layout, b->ctx->stream);
    if (S) download_planes(S, 3, P.p + 3, b->m, layout, b->ctx->stream);
    if (V) download_planes(V, 1, P.p + 6, b->m, layout, b->ctx->stream);
  });
}

int pnt_brush_set_pickup_rate(pnt_brush b, double rate) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->pickup_rate = rate;
  });
}
int pnt_brush_get_pickup_rate(pnt_brush b, double* rate) {
  return guard([&] {
    PNT_REQUIRE(b && rate, "NULL argument");
    *rate = b->pickup_rate;
  });
}
int pnt_brush_set_deposition_rate(pnt_brush b, double rate) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->dep_rate = rate;
  });
}
int pnt_brush_get_deposition_rate(pnt_brush b, double* rate) {
  return guard([&] {
    PNT_REQUIRE(b && rate, "NULL argument");
    *rate = b->dep_rate;
  });
}
int pnt_brush_set_use_snapshot(pnt_brush b, int use) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->use_snap = use != 0;
  });
}
int pnt_brush_get_use_snapshot(pnt_brush b, int* use) {
  return guard([&] {
    PNT_REQUIRE(b && use, "NULL argument");
    *use = b->use_snap ? 1 : 0;
  });
}

int pnt_brush_dip(pnt_brush b, const float K[3], const float S[3]) {
  return guard([&] {
    PNT_REQUIRE(b && K && S, "NULL argument");
    b->ctx->use();
    brush_dip(b, K, S);
  });
}

int pnt_brush_clean(pnt_brush b) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->ctx->use();
    launch_fill(b->pick, 7 * b->mstride, 0.0f, b->ctx->stream);
  });
}

int pnt_brush_update_snapshot(pnt_brush b, pnt_canvas cv) {
  return guard([&] {
    check_same_ctx(b, cv);
    b->ctx->use();
    canvas_tick(cv);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    // the dirty bytes that go with the snapshot (all clean after a full copy)
    const int64_t need = round_up(cv->n, 16) + 16;
    if (need > b->dirty_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
      dev_free(b->dirty);
      b->dirty = nullptr;
      PNT_HIP_CHECK(hipMalloc((void**)&b->dirty, (size_t)need));
      b->dirty_cap = need;
    }
    b->dirty_n = cv->n;
    brush_full_snapshot(b, cv);
  });
}

int pnt_imprint(pnt_brush b, pnt_canvas cv, double cx, double cy, double theta) {
  return guard([&] {
    check_same_ctx(b, cv);
    b->ctx->use();
    canvas_tick(cv);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    imprint_plain(b, cv, cx, cy, theta);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pnt_imprint_stamps(pnt_brush b, pnt_canvas cv, const pnt_stamp* st, int64_t n, int mode) {
  return guard([&] {
    check_same_ctx(b, cv);
    PNT_REQUIRE(n >= 0 && (st || n == 0), "invalid stamp list");
    PNT_REQUIRE(mode == PNT_EXEC_PLAIN || mode == PNT_EXEC_PERSISTENT || mode == PNT_EXEC_FUSED ||
                    mode == PNT_EXEC_CHAINED,
                "unknown mode");
    b->ctx->use();
    canvas_tick(cv);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    // checkDry runs on the plain and fused paths; the one-launch variants
    // (persistent, chained) of a canvas with a drying clock run fused
    if (cv->clock != PNT_CLOCK_OFF && (mode == PNT_EXEC_PERSISTENT || mode == PNT_EXEC_CHAINED))
      mode = PNT_EXEC_FUSED;
    if (mode == PNT_EXEC_PERSISTENT) {
      run_persistent(b, cv, st, n);
    } else if (mode == PNT_EXEC_FUSED) {
      run_fused(b, cv, st, n);
    } else if (mode == PNT_EXEC_CHAINED) {
      run_chained(b, cv, st, n);
    } else {
      for (int64_t i = 0; i < n; ++i) {
        if (st[i].flags & PNT_STAMP_DIP) brush_dip(b, st[i].K, st[i].S);
        imprint_plain(b, cv, st[i].cx, st[i].cy, st[i].theta);
        if (st[i].flags & PNT_STAMP_DRY_AFTER) {
          cv->epoch++;
          launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), cv->ctx->stream);
          canvas_stamp_time(cv);
        }
      }
    }
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pnt_imprint_stamps_noalloc(pnt_brush b, pnt_canvas cv, const pnt_stamp* st, int64_t n) {
  return guard([&] {
    check_same_ctx(b, cv);
    PNT_REQUIRE(n >= 0 && (st || n == 0), "invalid stamp list");
    PNT_REQUIRE(cv->clock == PNT_CLOCK_OFF,
                "the no-allocation call takes canvases without a drying clock");
    b->ctx->use();
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    run_fused_noalloc(b, cv, st, n);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pnt_imprint_strokes(pnt_brush b, pnt_canvas cv, const pnt_stroke* S, int64_t ns, int mode) {
  return guard([&] {
    check_same_ctx(b, cv);
    PNT_REQUIRE(ns >= 0 && (S || ns == 0), "invalid stroke list");
    PNT_REQUIRE(mode == PNT_EXEC_PLAIN || mode == PNT_EXEC_PERSISTENT || mode == PNT_EXEC_FUSED ||
                    mode == PNT_EXEC_CHAINED,
                "unknown mode");
    for (int64_t i = 0; i < ns; ++i)
      PNT_REQUIRE(S[i].n_stamps >= 0 && (S[i].stamps || S[i].n_stamps == 0) && S[i].footprint,
                  "invalid stroke");
    b->ctx->use();
    canvas_tick(cv);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    if (mode == PNT_EXEC_FUSED) {
      run_fused_strokes(b, cv, S, ns);
    } else {
      for (int64_t i = 0; i < ns; ++i) {
        if (!(std::fabs(b->radius - S[i].radius) < 0.5)) {  // FootprintBrush::setRadius
          PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
          brush_set_geometry(b, S[i].radius, S[i].footprint, S[i].width);
        }
        const int rc = pnt_imprint_stamps(b, cv, S[i].stamps, S[i].n_stamps, mode);
        if (rc != PNT_OK) throw Error(rc, pnt_last_error());
      }
    }
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pnt_brush_last_launches(pnt_brush b, int64_t* launches) {
  return guard([&] {
    PNT_REQUIRE(b && launches, "NULL argument");
    *launches = b->last_launches;
  });
}

int pnt_brush_set_kernel_timing(pnt_brush b, int on) {
  return guard([&] {
    PNT_REQUIRE(b, "NULL brush");
    if (!b->ktimer) b->ktimer = new KernelTimer;
    b->ktimer->on = on != 0;
    b->ktimer->used = 0;
  });
}

int pnt_brush_kernel_time(pnt_brush b, double* ms, int64_t* launches) {
  return guard([&] {
    PNT_REQUIRE(b && ms && launches, "NULL argument");
    b->ctx->use();
    *ms = b->ktimer ? ktimer_total_ms(*b->ktimer) : 0.0;
    *launches = b->ktimer ? (int64_t)(b->ktimer->used / 2) : 0;
  });
}

int pnt_brush_stamp_traffic(pnt_brush b, int rows, int cols, const pnt_stamp* st, int64_t n,
                            int64_t* visits, int64_t* ring) {
  return guard([&] {
    PNT_REQUIRE(b && visits && ring && (st || n == 0), "NULL argument");
    PNT_REQUIRE(rows > 0 && cols > 0, "rows/cols must be positive");
    int64_t nv = 0, nr = 0;
    StampGeom g;
    g.hr = (b->size_map - 1) / 2;
    g.size_map = b->size_map;
    g.rows = rows;
    g.cols = cols;
    for (int64_t k = 0; k < n; ++k) {
      g.cx = st[k].cx;
      g.cy = st[k].cy;
      g.c = std::cos(-st[k].theta);
      g.s = std::sin(-st[k].theta);
      for (int row = -g.hr; row <= g.hr; ++row)
        for (int col = -g.hr; col <= g.hr; ++col) {
          int64_t ic;
          const int m = stamp_visit(g, col, row, &ic);
          if (m >= 0 && b->fp_host[(size_t)m] > 0.0f) ++nv;
        }
      if (b->use_snap) {
        const RingBox box = ring_box(g, b->radius);
        if (box.a_x1 >= box.a_x0 && box.a_y1 >= box.a_y0) {
          const int64_t all = (int64_t)(box.a_x1 - box.a_x0 + 1) * (box.a_y1 - box.a_y0 + 1);
          const int ix0 = std::max(box.tl_x + 1, box.a_x0), ix1 = std::min(box.br_x - 1, box.a_x1);
          const int iy0 = std::max(box.tl_y + 1, box.a_y0), iy1 = std::min(box.br_y - 1, box.a_y1);
          const int64_t inner =
              (ix1 >= ix0 && iy1 >= iy0) ? (int64_t)(ix1 - ix0 + 1) * (iy1 - iy0 + 1) : 0;
          nr += all - inner;
        }
      }
    }
    *visits = nv;
    *ring = nr;
  });
}

// The checks every batch entry point makes on its jobs.
static void validate_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs) {
  PNT_REQUIRE(ctx && (jobs || n_jobs == 0) && n_jobs >= 0, "invalid batch");
  ctx->use();
  std::vector<const void*> seen;
  for (int j = 0; j < n_jobs; ++j) {
    const pnt_batch_job& J = jobs[j];
    PNT_REQUIRE(J.canvas && J.brush, "batch job without canvas or brush");
    PNT_REQUIRE(J.canvas->ctx->device == ctx->device && J.brush->ctx->device == ctx->device,
                "batch job on another device");
    PNT_REQUIRE(J.n_stamps >= 0 && (J.stamps || J.n_stamps == 0), "invalid stamp list");
    if (J.out[0])
      for (int c = 0; c < 3; ++c)
        PNT_REQUIRE(J.out[c] && ((uintptr_t)J.out[c] & 15u) == 0,
                    "composite planes must be 16-byte aligned device pointers");
    seen.push_back(J.canvas);
    seen.push_back(J.brush);
  }
  std::sort(seen.begin(), seen.end());
  PNT_REQUIRE(std::adjacent_find(seen.begin(), seen.end()) == seen.end(),
              "canvases and brushes must be distinct across the jobs of a batch");
  for (int j = 0; j < n_jobs; ++j) canvas_tick(jobs[j].canvas);  // due drying-timer steps
  for (int j = 0; j < n_jobs; ++j)  // other streams' work on these objects comes first
    if (jobs[j].canvas->ctx->stream != ctx->stream || jobs[j].brush->ctx->stream != ctx->stream) {
      PNT_HIP_CHECK(hipStreamSynchronize(jobs[j].canvas->ctx->stream));
      PNT_HIP_CHECK(hipStreamSynchronize(jobs[j].brush->ctx->stream));
    }
}

static void check_config(const pnt_batch_config* c) {
  PNT_REQUIRE(c, "NULL configuration");
  PNT_REQUIRE(c->tile == 16 || c->tile == 32, "tile must be 16 or 32");
  PNT_REQUIRE(c->max_group >= 1 && c->max_group <= 256, "max_group must be in [1, 256]");
  PNT_REQUIRE(c->waves == 0 || c->waves == 1, "waves must be 0 or 1");
  PNT_REQUIRE(c->chunk == 1 || c->chunk == 2 || c->chunk == 4, "chunk must be 1, 2 or 4");
}

int pnt_batch_submit(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs) {
  return pnt_batch_run(ctx, jobs, n_jobs, PNT_BATCH_ONE_LAUNCH);
}

int pnt_batch_wait(pnt_ctx ctx) { return pnt_ctx_synchronize(ctx); }

int pnt_batch_run(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, int mode) {
  return guard([&] {
    PNT_REQUIRE(mode == PNT_BATCH_ONE_LAUNCH || mode == PNT_BATCH_PLAIN, "unknown batch mode");
    validate_batch(ctx, jobs, n_jobs);
    if (n_jobs == 0) return;
    if (mode == PNT_BATCH_ONE_LAUNCH) {
      run_batch_one_launch(ctx, jobs, n_jobs, nullptr);
      return;
    }
    for (int j = 0; j < n_jobs; ++j) {  // plain: the job's launches on its brush's stream
      const pnt_batch_job& J = jobs[j];
      pnt_brush b = J.brush;
      pnt_canvas cv = J.canvas;
      hipStream_t s = b->ctx->stream;
      for (int64_t i = 0; i < J.n_stamps; ++i) {
        const pnt_stamp& st = J.stamps[i];
        if (st.flags & PNT_STAMP_DIP) brush_dip(b, st.K, st.S);
        imprint_plain(b, cv, st.cx, st.cy, st.theta);
        if (st.flags & PNT_STAMP_DRY_AFTER) {
          cv->epoch++;
          launch_dry(cv->n, cv->layer(), cv->R0(), cv->h(), s);
        }
      }
      if (J.out[0])
        launch_compose(cv->n, cv->layer(), cv->R0(), Rgb3{{J.out[0], J.out[1], J.out[2]}}, s);
      if (s != ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(s));
    }
  });
}

int pnt_batch_tune(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, pnt_batch_config* best,
                   double* best_ms) {
  return guard([&] {
    validate_batch(ctx, jobs, n_jobs);
    PNT_REQUIRE(n_jobs > 0, "empty batch");
    tune_batch(ctx, jobs, n_jobs, best, best_ms);
  });
}

int pnt_batch_profile(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, int cap,
                      int* job_first, int* job_last, int* wave_blocks, float* wave_end_ms,
                      int* n_waves) {
  return guard([&] {
    validate_batch(ctx, jobs, n_jobs);
    PNT_REQUIRE(n_jobs > 0, "empty batch");
    PNT_REQUIRE(cap >= 0, "negative cap");
    BatchWaveProfile P;
    profile_batch_waves(ctx, jobs, n_jobs, &P);
    const int nw = (int)P.wave_blocks.size();
    for (int k = 0; k < n_jobs; ++k) {
      if (job_first) job_first[k] = P.job_first[(size_t)k];
      if (job_last) job_last[k] = P.job_last[(size_t)k];
    }
    for (int w = 0; w < std::min(nw, cap); ++w) {
      if (wave_blocks) wave_blocks[w] = P.wave_blocks[(size_t)w];
      if (wave_end_ms) wave_end_ms[w] = P.wave_end_ms[(size_t)w];
    }
    if (n_waves) *n_waves = nw;
  });
}

int pnt_batch_set_config(pnt_ctx ctx, const pnt_batch_config* cfg) {
  return guard([&] {
    PNT_REQUIRE(ctx, "NULL context");
    check_config(cfg);
    ctx->use();
    pnt_batch_config c = *cfg;
    set_batch_config(ctx, c);
  });
}

int pnt_batch_get_config(pnt_ctx ctx, pnt_batch_config* cfg) {
  return guard([&] {
    PNT_REQUIRE(ctx && cfg, "NULL argument");
    ctx->use();
    *cfg = get_batch_config(ctx);
  });
}

int pnt_batch_capture(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, pnt_batch_graph* out) {
  return guard([&] {
    PNT_REQUIRE(out, "NULL argument");
    validate_batch(ctx, jobs, n_jobs);
    PNT_REQUIRE(n_jobs > 0, "empty batch");
    *out = reinterpret_cast<pnt_batch_graph>(capture_batch(ctx, jobs, n_jobs));
  });
}

int pnt_batch_graph_launch(pnt_batch_graph g) {
  return guard([&] {
    PNT_REQUIRE(g, "NULL graph");
    launch_batch_graph(reinterpret_cast<BatchGraph*>(g));
  });
}

int pnt_batch_graph_destroy(pnt_batch_graph g) {
  return guard([&] {
    BatchGraph* G = reinterpret_cast<BatchGraph*>(g);
    release_batch_graph(G);
  });
}

int pnt_paint_stroke(pnt_brush b, pnt_canvas cv, const double* p, int64_t n) {
  return guard([&] {
    check_same_ctx(b, cv);
    PNT_REQUIRE(n >= 0 && (p || n == 0), "invalid path");
    b->ctx->use();
    canvas_tick(cv);
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    // FootprintBrush.hxx:251-267
    for (int64_t i = 0; i + 1 < n; ++i) {
      const int64_t ipre = i > 0 ? i - 1 : 0;
      const int64_t inext = (i + 2 < n) ? i + 2 : n - 1;
      const double *pp = p + 2 * ipre, *p0 = p + 2 * i, *p1 = p + 2 * (i + 1), *pn = p + 2 * inext;
      double dist;
      {
#pragma clang fp contract(off)
        const double dx = p1[0] - p0[0], dy = p1[1] - p0[1];
        dist = std::sqrt(dx * dx + dy * dy);
      }
      for (int pd = 1; pd <= (int)dist; ++pd) {
        const double t = (double)pd / dist;
        const double ddx = catmull_rom_d1(pp[0], p0[0], p1[0], pn[0], t);
        const double ddy = catmull_rom_d1(pp[1], p0[1], p1[1], pn[1], t);
        imprint_plain(b, cv, catmull_rom(pp[0], p0[0], p1[0], pn[0], t),
                      catmull_rom(pp[1], p0[1], p1[1], pn[1], t), std::atan2(ddy, ddx));
      }
    }
    if (cv->ctx->stream != b->ctx->stream) PNT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

// -------------------------------------------------------------- PaintMixer --
int pnt_compute_scattering_absorption(pnt_ctx ctx, int64_t n, const double* Rb, const double* Rw,
                                      double* K, double* S) {
  return guard([&] {
    PNT_REQUIRE(ctx && (n == 0 || (Rb && Rw && K && S)), "NULL argument");
    PNT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return;
    ctx->use();
    hipStream_t s = ctx->stream;
    const int64_t n3 = 3 * n;
    double* buf = nullptr;
    const size_t bytes = (size_t)(4 * n3) * sizeof(double) + 256;
    PNT_HIP_CHECK(hipMalloc((void**)&buf, bytes));
    try {
      unsigned* bad = reinterpret_cast<unsigned*>(buf + 4 * n3);
      launch_fill_u32(bad, 1, 0u, s);
      PNT_HIP_CHECK(hipMemcpyAsync(buf, Rb, n3 * sizeof(double), hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(buf + n3, Rw, n3 * sizeof(double), hipMemcpyHostToDevice, s));
      launch_scatter_absorb(n3, buf, buf + n3, buf + 2 * n3, buf + 3 * n3, bad, s);
      unsigned h_bad = 0;
      PNT_HIP_CHECK(hipMemcpyAsync(&h_bad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipMemcpyAsync(K, buf + 2 * n3, n3 * sizeof(double), hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipMemcpyAsync(S, buf + 3 * n3, n3 * sizeof(double), hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      PNT_REQUIRE(!h_bad, "on black or white inputs violate one of the following conditions: "
                          "0 < black < white < 1");
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(buf);
      throw;
    }
    PNT_HIP_CHECK(hipFree(buf));
  });
}

int pnt_palette_from_json(const char* text, int64_t len, double* K, double* S, int64_t capacity,
                          int64_t* n) {
  return guard([&] {
    PNT_REQUIRE(text && n && len >= 0 && capacity >= 0, "invalid argument");
    std::vector<double> k, s;
    palette_from_json(text, len, k, s);
    *n = (int64_t)(k.size() / 3);
    PNT_REQUIRE(*n <= capacity, "palette has more entries than the capacity given");
    PNT_REQUIRE((K && S) || *n == 0, "NULL output");
    if (*n) {
      std::memcpy(K, k.data(), k.size() * sizeof(double));
      std::memcpy(S, s.data(), s.size() * sizeof(double));
    }
  });
}

int pnt_palette_to_json(const double* K, const double* S, int64_t n, char* text, int64_t capacity,
                        int64_t* len) {
  return guard([&] {
    PNT_REQUIRE(len && n >= 0 && (n == 0 || (K && S)), "invalid argument");
    const std::string js = palette_to_json(K, S, n);
    *len = (int64_t)js.size();
    PNT_REQUIRE(text && capacity > *len, "text buffer too small (needs *len + 1 bytes)");
    std::memcpy(text, js.c_str(), js.size() + 1);
  });
}

void pnt_thinning_medium(double K[3], double S[3]) {
  for (int c = 0; c < 3; ++c) K[c] = S[c] = 0.01;
}

int pnt_mix_single(int n_palette, const double* base_K, const double* base_S, int n_weights,
                   const double* w, double K[3], double S[3]) {
  return guard([&] {
    if (n_weights != n_palette)
      throw Error(PNT_EINVAL, "Palette size does not match underlying size.");
    PNT_REQUIRE((base_K && base_S && w) || n_palette == 0, "NULL argument");
    PNT_REQUIRE(K && S, "NULL argument");
    mix_single(n_palette, base_K, base_S, w, K, S);
  });
}

int pnt_mixed(const double K1[3], const double S1[3], double v1, const double K2[3],
              const double S2[3], double v2, double K[3], double S[3]) {
  return guard([&] {
    PNT_REQUIRE(K1 && S1 && K2 && S2 && K && S, "NULL argument");
    mixed(K1, S1, v1, K2, S2, v2, K, S);
  });
}

}  // extern "C"
