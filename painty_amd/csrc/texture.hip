// TextureBrush (renderer/TextureBrush.hxx:19-238) with its Smudge
// (renderer/Smudge.hxx:22-229): a brush stroke texture -- a thickness map --
// warped along the stroke by generalized barycentric coordinates
// (core/Math.hxx:73-133, the mean value coordinates of Hormann and Floater
// over the stroke's frame polygon), an optional smudge pass that trades
// paint between the canvas and a rotating pickup map along the stroke, then
// volume-weighted mixing with V = max(Vtex, Vcanvas).
//
// Kernels, one thread per pixel (or map cell):
//   k_tex_sample   texture coordinates by the barycentric interpolation, in
//                  FP64 as the reference; the texture sampled bilinearly
//                  (Interpolate, image/Mat.hxx:69-100, reflected borders);
//                  Canvas::checkDry on textured pixels when a clock is on;
//                  the largest Vtex folded in per workgroup (the smudge's
//                  maximum thickness);
//   k_smudge_step  one spine position of Smudge::smudge: the pickup map
//                  turned (updateOrientation, resampled at the rotated
//                  positions into the other of two FP64 buffers), then the
//                  cell's exchange with the canvas pixel under it;
//   k_tex_deposit  the volume-weighted mix.
// The smudge runs one step per unit of stroke length, in stream order.  A
// step is independent per cell: a cell reads the previous map and writes
// only itself, and a map cell meets one canvas pixel at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "dry.hpp"
#include "handles.hpp"
#include "internal.hpp"

// FP64 expressions evaluate as written (no fused multiply-adds), as the
// reference's do: the texture coordinates decide which pixels a stroke
// covers, and the smudge's map offsets are truncations.
#pragma clang fp contract(off)

namespace pnt {
namespace {

constexpr int kBlock = 256;
// k_smudge_step covers one small map per launch (6400 cells at radius 40):
// 64-thread workgroups spread a step over four times the CUs (6.7 against
// 7.3 us per launch at radius 40).
constexpr int kStepBlock = 64;
constexpr double kPi = 3.141592653589793238462643383279502884;  // Math.hxx:21-22

__host__ __device__ __forceinline__ int reflect_index(int p, int len) {
  if (len == 1) return 0;
  do {
    if (p < 0)
      p = -p - 1;
    else if (p >= len)
      p = 2 * len - p - 1;
  } while (p < 0 || p >= len);
  return p;
}

// Interpolate (image/Mat.hxx:69-100): bilinear with BORDER_REFLECT; element
// (y, x) of the image is img[(y * cols + x) * step].
template <class T>
__device__ __forceinline__ double interp(const T* img, int rows, int cols, int step, double x,
                                         double y) {
  const int xi = (int)floor(x), yi = (int)floor(y);
  const int x0 = reflect_index(xi, cols), x1 = reflect_index(xi + 1, cols);
  const int y0 = reflect_index(yi, rows), y1 = reflect_index(yi + 1, rows);
  const double a = x - (double)xi, c = y - (double)yi;
  const auto at = [&](int yy, int xx) { return (double)img[((int64_t)yy * cols + xx) * step]; };
  return (at(y0, x0) * (1.0 - a) + at(y0, x1) * a) * (1.0 - c) +
         (at(y1, x0) * (1.0 - a) + at(y1, x1) * a) * c;
}

// generalizedBarycentricCoordinatesInterpolate (core/Math.hxx:73-133) of
// the 2-d values `val` at the np vertices of polygon `poly`.
__device__ void gbc(const double2* poly, const double2* val, int np, double px, double py,
                    double* ou, double* ov) {
  const double Eps = DBL_EPSILON * 100.0;
  if (np == 1) {
    *ou = val[0].x;
    *ov = val[0].y;
    return;
  }
  // r, A, D of each vertex on the fly: vertex i needs those of i - 1 and i
  auto sv = [&](int i) { return make_double2(poly[i].x - px, poly[i].y - py); };
  // the early exits of the reference scan the vertices in order first
  for (int i = 0; i < np; ++i) {
    const int i1 = i == np - 1 ? 0 : i + 1;
    const double2 si = sv(i), si1 = sv(i1);
    const double ri = sqrt(si.x * si.x + si.y * si.y);
    if (fabs(ri - 0.0) < Eps) {
      *ou = val[i].x;
      *ov = val[i].y;
      return;
    }
    const double A = (si.x * si1.y - si1.x * si.y) / 2.0;
    const double D = si.x * si1.x + si.y * si1.y;
    if (fabs(A - 0.0) < Eps && D < 0.0) {
      const double ri1 = sqrt(si1.x * si1.x + si1.y * si1.y);
      const double w = 1.0 / (ri + ri1);
      *ou = (ri1 * val[i].x + ri * val[i1].x) * w;
      *ov = (ri1 * val[i].y + ri * val[i1].y) * w;
      return;
    }
  }
  double fu = 0.0, fv = 0.0, W = 0.0;
  // A, D, r of vertex np - 1 (the predecessor of vertex 0)
  double2 sp = sv(np - 1), s0 = sv(0);
  double rp = sqrt(sp.x * sp.x + sp.y * sp.y);
  double Ap = (sp.x * s0.y - s0.x * sp.y) / 2.0, Dp = sp.x * s0.x + sp.y * s0.y;
  double2 si = s0;
  double ri = sqrt(si.x * si.x + si.y * si.y);
  for (int i = 0; i < np; ++i) {
    const int i1 = i == np - 1 ? 0 : i + 1;
    const double2 si1 = sv(i1);
    const double ri1 = sqrt(si1.x * si1.x + si1.y * si1.y);
    const double Ai = (si.x * si1.y - si1.x * si.y) / 2.0;
    const double Di = si.x * si1.x + si.y * si1.y;
    double w = 0.0;
    if (Ap != 0.0) w = w + (rp - Dp / ri) / Ap;
    if (Ai != 0.0) w = w + (ri1 - Di / ri) / Ai;
    fu = fu + w * val[i].x;
    fv = fv + w * val[i].y;
    W = W + w;
    rp = ri;
    Ap = Ai;
    Dp = Di;
    si = si1;
    ri = ri1;
  }
  if (!(fabs(W - 0.0) < Eps)) {
    *ou = fu * (1.0 / W);
    *ov = fv * (1.0 / W);
  } else {
    *ou = val[0].x;
    *ov = val[0].y;
  }
}

struct TexArgs {
  Layer7 canvas;
  int rows, cols;
  const double2* poly;  // frame polygon, np vertices
  const double2* uv;    // texture coordinates at the vertices
  int np;
  const double* thick;  // brush stroke sample thickness map (FP64, as the reference's Mat<double>)
  int t_rows, t_cols;
  double scale;         // thickness scale
  int bx0, by0, bx1, by1;  // pixel loop: (int) of the bounding box corners
  int mx0, my0;         // (int) boundMin: origin of the stroke's thickness map
  int trows, tcols;     // its size
  double* tmap;         // the stroke's thickness map (trows x tcols), zeroed
  DryClock clk;         // checkDry; clk.tmap == nullptr: off
  double paint[6];      // the brush's K, S
  unsigned long long* maxd;  // largest Vtex, as the bits of a non-negative double (zeroed)
};

__device__ double tex_sample_pixel(const TexArgs& A, int64_t e) {
  const int w = A.bx1 - A.bx0 + 1;
  const int64_t total = (int64_t)w * (A.by1 - A.by0 + 1);
  if (e >= total) return 0.0;
  const int x = A.bx0 + (int)(e % w), y = A.by0 + (int)(e / w);
  if (x < 0 || x >= A.cols || y < 0 || y >= A.rows) return 0.0;
  double u, v;
  gbc(A.poly, A.uv, A.np, (double)x, (double)y, &u, &v);
  if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) return 0.0;
  u *= A.t_cols;
  v *= A.t_rows;
  const double vt = A.scale * interp(A.thick, A.t_rows, A.t_cols, 1, u, v);
  if (!(vt > 0.0)) return 0.0;
  const int s = x - A.mx0, t = y - A.my0;
  if (s < 0 || t < 0 || s >= A.tcols || t >= A.trows) return 0.0;
  if (A.clk.tmap) check_dry(A.clk, A.canvas, (int64_t)y * A.cols + x);
  A.tmap[(int64_t)t * A.tcols + s] = vt;
  return vt;
}

// One thread per bounding-box pixel; each workgroup also folds the largest
// Vtex it stored into *A.maxd (ordered as integers: Vtex > 0).
__global__ __launch_bounds__(kBlock) void k_tex_sample(TexArgs A) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = tex_sample_pixel(A, (int64_t)blockIdx.x * kBlock + threadIdx.x);
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] > 0.0)
    atomicMax(A.maxd, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ __launch_bounds__(kBlock) void k_tex_deposit(TexArgs A) {
  const int w = A.bx1 - A.bx0 + 1;
  const int64_t total = (int64_t)w * (A.by1 - A.by0 + 1);
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= total) return;
  const int x = A.bx0 + (int)(e % w), y = A.by0 + (int)(e / w);
  if (x < 0 || x >= A.cols || y < 0 || y >= A.rows) return;
  const int s = x - A.mx0, t = y - A.my0;
  if (s < 0 || t < 0 || s >= A.tcols || t >= A.trows) return;
  const double vt = A.tmap[(int64_t)t * A.tcols + s];
  if (!(vt > 0.0)) return;  // not textured: not in the reference's pixel list
  const int64_t ic = (int64_t)y * A.cols + x;
  const double vc = A.canvas.v[ic];
  const double vs = vc + vt;
  if (vs > 0.0) {
    const double sc = 1.0 / vs;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A.canvas.ks[c][ic] = (float)((vc * A.canvas.ks[c][ic] + vt * A.paint[c]) * sc);
      A.canvas.ks[3 + c][ic] = (float)((vc * A.canvas.ks[3 + c][ic] + vt * A.paint[3 + c]) * sc);
    }
    A.canvas.v[ic] = fmax(vt, vc);
  }
}

// A smudge pickup map: K, S (3 channels, interleaved per cell) and V.
struct SMap {
  double* K;
  double* S;
  double* V;
};

// Where a smudge step meets the canvas: the spine position and the map's
// top-left canvas pixel, the stroke's thickness map and its largest value.
struct TradeArgs {
  Layer7 canvas;
  int rows, cols;
  int roi_x, roi_y;
  double bmin_x, bmin_y, ctr_x, ctr_y, radius;
  const double* tmap;
  int trows, tcols;
  double maxD;
};

// One spine position of Smudge::smudge (Smudge.hxx:49-157), for map cell
// ip = y * n + x:
//  - updateOrientation (Smudge.hxx:164-198): dst(x, y) = src at (x, y)
//    rotated by dtheta about (max/2, max/2), bilinear; src(x, y) where the
//    rotated position leaves the map;
//  - the exchange of dst(x, y) with the canvas pixel under it.
__global__ __launch_bounds__(kStepBlock) void k_smudge_step(SMap src, SMap dst, int n, double cx,
                                                            double cy, double sn, double cs,
                                                            TradeArgs T) {
  const int64_t ip = (int64_t)blockIdx.x * kStepBlock + threadIdx.x;
  if (ip >= (int64_t)n * n) return;
  const int x = (int)(ip % n), y = (int)(ip / n);
  double pK[3], pS[3], pV;
  {
    const double px = x - cx, py = y - cy;
    const double rx = px * cs - py * sn + cx, ry = px * sn + py * cs + cy;
    if (rx < 0 || ry < 0 || rx >= n || ry >= n) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pK[c] = src.K[3 * ip + c];
        pS[c] = src.S[3 * ip + c];
      }
      pV = src.V[ip];
    } else {
      pV = interp(src.V, n, n, 1, rx, ry);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pK[c] = interp(src.K + c, n, n, 3, rx, ry);
        pS[c] = interp(src.S + c, n, n, 3, rx, ry);
      }
    }
  }
  const int cpx = x + T.roi_x, cpy = y + T.roi_y;
  const int tpx = (int)(cpx - T.bmin_x), tpy = (int)(cpy - T.bmin_y);
  bool trade = !(cpx < 0 || cpy < 0 || cpx >= T.cols || cpy >= T.rows) &&
               !(tpx < 0 || tpy < 0 || tpx >= T.tcols || tpy >= T.trows);
  double Dv = 0.0;
  if (trade) {
    const double dx = T.ctr_x - (double)cpx, dy = T.ctr_y - (double)cpy;
    trade = !(sqrt(dx * dx + dy * dy) > T.radius);
  }
  if (trade) {
    Dv = T.tmap[(int64_t)tpy * T.tcols + tpx];
    trade = !(Dv <= 0.0);
  }
  if (trade) {
    constexpr double pickupRate = 0.1, depositionRate = 0.1, MinVolume = 0.001;
    const double md = T.maxD;
    const int64_t ic = (int64_t)cpy * T.cols + cpx;
    const double cV = T.canvas.v[ic];
    const double cVl = cV * depositionRate * Dv / md, cVr = cV - cVl;
    const double pVl = pV * pickupRate * Dv / md, pVr = pV - pVl;
    double cK[3], cS[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      cK[c] = T.canvas.ks[c][ic];
      cS[c] = T.canvas.ks[3 + c][ic];
    }
    const double cVnew = cVr + pVl;
    if (cVnew > MinVolume) {
      const double inv = 1. / cVnew;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T.canvas.ks[c][ic] = (float)(inv * (cVr * cK[c] + pVl * pK[c]));
        T.canvas.ks[3 + c][ic] = (float)(inv * (cVr * cS[c] + pVl * pS[c]));
      }
      T.canvas.v[ic] = fmax(cVnew, 0.0);
    }
    const double pVnew = pVr + cVl;
    if (pVnew > MinVolume) {
      const double inv = 1.0 / pVnew;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double k = inv * (pVr * pK[c] + cVl * cK[c]);
        const double q = inv * (pVr * pS[c] + cVl * cS[c]);
        pK[c] = k;
        pS[c] = q;
      }
      pV = fmax(pVnew, 0.0);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dst.K[3 * ip + c] = pK[c];
    dst.S[3 * ip + c] = pS[c];
  }
  dst.V[ip] = pV;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned step_blocks_for(int64_t n) {
  return (unsigned)((n + kStepBlock - 1) / kStepBlock);
}

// SplineEval::catmullRom / catmullRomDerivativeFirst (core/Spline.hxx:150-
// 221) over points v: u in [0, 1] spans the segments, indices clamped.
void spline_cr(const std::vector<double2>& v, double u, bool deriv, double out[2]) {
  const int n = (int)v.size();
  const double x = (double)(n - 1) * u;
  const int index = (int)x;
  const double t = x - std::floor(x);
  const double tau = 0.5, t2 = t * t, t3 = t2 * t;
  double b[4];
  if (!deriv) {
    b[0] = -tau * t + 2.0 * tau * t2 - tau * t3;
    b[1] = 1.0 + (tau - 3.0) * t2 + (2.0 - tau) * t3;
    b[2] = tau * t + (3.0 - 2.0 * tau) * t2 + (tau - 2.0) * t3;
    b[3] = -tau * t2 + tau * t3;
  } else {
    b[0] = tau * (-3.0 * t2 + 4.0 * t - 1.0);
    b[1] = -t * (-2.0 * tau + 3.0 * (tau - 2.0) * t + 6.0);
    b[2] = (t - 1.0) * (3.0 * (tau - 2.0) * t - tau);
    b[3] = tau * t * (3.0 * t - 2.0);
  }
  double2 p[4];
  for (int q = 0; q < 4; ++q) p[q] = v[(size_t)std::min(std::max(index - 1 + q, 0), n - 1)];
  out[0] = p[0].x * b[0] + p[1].x * b[1] + p[2].x * b[2] + p[3].x * b[3];
  out[1] = p[0].y * b[0] + p[1].y * b[1] + p[2].y * b[2] + p[3].y * b[3];
}

void normalize2(double t[2]) {  // Eigen's normalized(): unchanged when zero
  const double z = t[0] * t[0] + t[1] * t[1];
  if (z > 0.0) {
    const double n = std::sqrt(z);
    t[0] /= n;
    t[1] /= n;
  }
}

double normalize_angle(double a) {  // Smudge.hxx:200-210
  const double half = 0.5 * kPi;
  while (a <= -half) a += kPi;
  while (a > half) a -= kPi;
  return a;
}

}  // namespace

int guard_call(const std::function<void()>& f);  // api.hip: status + pnt_last_error
}  // namespace pnt

using namespace pnt;

struct pnt_tbrush_s {
  pnt_ctx ctx = nullptr;
  double radius = 0.0;
  double paint[6] = {0.1, 0.1, 0.1, 0.1, 0.1, 0.1};  // TextureBrush.hxx:27-29
  double* thick = nullptr;  // thickness map of the brush stroke sample
  int t_rows = 0, t_cols = 0;
  double scale = 1.0;      // BrushBase thickness scale
  bool use_smudge = true;  // TextureBrush.hxx:237
  // Smudge(size): maps of size x size cells, _maxSize = size made odd
  int map = 0, max_size = 1;
  double* smap[2] = {nullptr, nullptr};  // ping-pong: 7 doubles per cell
  int cur = 0;             // the map in use
  double cur_rot = 0.0;
  // scratch of a stroke
  double* geo = nullptr;   // polygon + uv + maxD
  int64_t geo_cap = 0;
  double* tmap = nullptr;
  int64_t tmap_cap = 0;
  SMap maps(int k) const {
    const int64_t m = (int64_t)map * map;
    return SMap{smap[k], smap[k] + 3 * m, smap[k] + 6 * m};
  }
};

namespace pnt {
namespace {
void tb_smudge_reset(pnt_tbrush_s* b, int size) {
  hipStream_t s = b->ctx->stream;
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  for (double*& p : b->smap) {
    dev_free(p);
    p = nullptr;
  }
  b->map = std::max(size, 0);
  b->max_size = (size % 2 == 0) ? size + 1 : size;
  const int64_t m = (int64_t)b->map * b->map;
  for (double*& p : b->smap) {
    p = reinterpret_cast<double*>(dev_alloc(2 * std::max<int64_t>(7 * m, 1)));
    launch_fill(reinterpret_cast<float*>(p), 2 * std::max<int64_t>(7 * m, 1), 0.0f, s);
  }
  b->cur = 0;
  b->cur_rot = 0.0;
}
}  // namespace
}  // namespace pnt

extern "C" {

int pnt_tbrush_create(pnt_ctx ctx, pnt_tbrush* out) {
  return guard_call([&] {
    PNT_REQUIRE(ctx && out, "NULL argument");
    ctx->use();
    auto* b = new pnt_tbrush_s;
    b->ctx = ctx;
    try {
      tb_smudge_reset(b, 0);  // Smudge(2.0 * _radius) with _radius = 0
    } catch (...) {
      for (double* p : b->smap) dev_free(p);
      delete b;
      throw;
    }
    ctx_retain(ctx);
    *out = b;
  });
}

int pnt_tbrush_destroy(pnt_tbrush b) {
  return guard_call([&] {
    if (!b) return;
    pnt_ctx ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (double* p : b->smap) dev_free(p);
    dev_free(b->thick);
    dev_free(b->geo);
    dev_free(b->tmap);
    delete b;
    ctx_release(ctx);
  });
}

int pnt_tbrush_set_radius(pnt_tbrush b, double radius) {
  return guard_call([&] {
    PNT_REQUIRE(b, "NULL brush");
    PNT_REQUIRE(std::isfinite(radius) && radius >= 0.0 && radius < 4096.0,
                "radius must be in [0, 4096)");
    b->ctx->use();
    if (!(std::fabs(b->radius - radius) < 0.5)) {  // TextureBrush.hxx:31-39
      b->radius = radius;
      if (b->use_smudge) tb_smudge_reset(b, (int)(2.0 * radius));
    }
  });
}

int pnt_tbrush_dip(pnt_tbrush b, const double K[3], const double S[3]) {
  return guard_call([&] {
    PNT_REQUIRE(b && K && S, "NULL argument");
    for (int c = 0; c < 3; ++c) {
      b->paint[c] = K[c];
      b->paint[3 + c] = S[c];
    }
  });
}

int pnt_tbrush_set_thickness_map(pnt_tbrush b, const double* map, int rows, int cols) {
  return guard_call([&] {
    PNT_REQUIRE(b && map, "NULL argument");
    PNT_REQUIRE(rows > 0 && cols > 0 && (int64_t)rows * cols < ((int64_t)1 << 28),
                "thickness map size out of range");
    for (int64_t i = 0; i < (int64_t)rows * cols; ++i)
      PNT_REQUIRE(std::isfinite(map[i]), "thickness map values must be finite");
    b->ctx->use();
    hipStream_t s = b->ctx->stream;
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    dev_free(b->thick);
    b->thick = nullptr;
    b->thick = reinterpret_cast<double*>(dev_alloc(2 * (int64_t)rows * cols));
    PNT_HIP_CHECK(hipMemcpyAsync(b->thick, map, (size_t)rows * cols * sizeof(double),
                                 hipMemcpyHostToDevice, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    b->t_rows = rows;
    b->t_cols = cols;
  });
}

int pnt_tbrush_set_thickness_scale(pnt_tbrush b, double scale) {
  return guard_call([&] {
    PNT_REQUIRE(b && std::isfinite(scale), "invalid argument");
    b->scale = scale;
  });
}

int pnt_tbrush_set_use_smudge(pnt_tbrush b, int on) {
  return guard_call([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->use_smudge = on != 0;
  });
}

int pnt_tbrush_get_smudge_map(pnt_tbrush b, int* side, double* K, double* S, double* V) {
  return guard_call([&] {
    PNT_REQUIRE(b && side, "NULL argument");
    b->ctx->use();
    *side = b->map;
    const int64_t m = (int64_t)b->map * b->map;
    hipStream_t s = b->ctx->stream;
    const SMap mp = b->maps(b->cur);
    if (K) PNT_HIP_CHECK(hipMemcpyAsync(K, mp.K, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
    if (S) PNT_HIP_CHECK(hipMemcpyAsync(S, mp.S, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
    if (V) PNT_HIP_CHECK(hipMemcpyAsync(V, mp.V, m * sizeof(double), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int pnt_tbrush_paint_stroke(pnt_tbrush b, pnt_canvas cv, const double* xy, int64_t n) {
  return guard_call([&] {
    PNT_REQUIRE(b && cv && (xy || n == 0) && n >= 0, "invalid argument");
    PNT_REQUIRE(b->ctx->device == cv->ctx->device, "brush and canvas on different devices");
    for (int64_t i = 0; i < 2 * n; ++i) PNT_REQUIRE(std::isfinite(xy[i]), "non-finite vertex");
    if (n < 2) return;  // TextureBrush.hxx:54-56
    PNT_REQUIRE(b->thick, "no thickness map (pnt_tbrush_set_thickness_map)");
    b->ctx->use();
    canvas_tick(cv);
    hipStream_t s = b->ctx->stream;
    if (cv->ctx->stream != s) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    cv->epoch++;
    const double r = b->radius;
    // the stroke extended by the radius at both ends (TextureBrush.hxx:58-66)
    std::vector<double2> v;
    v.reserve((size_t)n + 2);
    double d0[2] = {xy[2] - xy[0], xy[3] - xy[1]};
    double d1[2] = {xy[2 * n - 2] - xy[2 * n - 4], xy[2 * n - 1] - xy[2 * n - 3]};
    normalize2(d0);
    normalize2(d1);
    v.push_back(make_double2(xy[0] - d0[0] * r, xy[1] - d0[1] * r));
    for (int64_t i = 0; i < n; ++i) v.push_back(make_double2(xy[2 * i], xy[2 * i + 1]));
    v.push_back(make_double2(xy[2 * n - 2] + d1[0] * r, xy[2 * n - 1] + d1[1] * r));
    const int nv = (int)v.size();
    double bmin[2] = {v[0].x, v[0].y}, bmax[2] = {v[0].x, v[0].y};
    for (const double2& p : v) {
      bmin[0] = std::min(bmin[0], p.x);
      bmin[1] = std::min(bmin[1], p.y);
      bmax[0] = std::max(bmax[0], p.x);
      bmax[1] = std::max(bmax[1], p.y);
    }
    bmin[0] = std::max(bmin[0] - r, 0.0);
    bmax[0] = std::min(bmax[0] + r, (double)(cv->cols - 1));
    bmin[1] = std::max(bmin[1] - r, 0.0);
    bmax[1] = std::min(bmax[1] + r, (double)(cv->rows - 1));
    double length = 0.0;
    for (int i = 1; i < nv; ++i)
      length += std::sqrt((v[i].x - v[i - 1].x) * (v[i].x - v[i - 1].x) +
                          (v[i].y - v[i - 1].y) * (v[i].y - v[i - 1].y));
    // frame polygon: the right side backwards, then the left side (:95-121)
    const int np = 2 * nv;
    std::vector<double2> poly((size_t)np), uv((size_t)np);
    for (int i = 0; i < nv; ++i) {
      const double u = (double)i / (double)(nv - 1);
      double c[2], t[2];
      spline_cr(v, u, false, c);
      spline_cr(v, u, true, t);
      normalize2(t);
      const double d[2] = {-t[1], t[0]};
      poly[(size_t)(nv + i)] = make_double2(c[0] - r * d[0], c[1] - r * d[1]);
      uv[(size_t)(nv + i)] = make_double2(u, 0.0);
      poly[(size_t)(nv - 1 - i)] = make_double2(c[0] + r * d[0], c[1] + r * d[1]);
      uv[(size_t)(nv - 1 - i)] = make_double2(u, 1.0);
    }
    const int trows = (int)(bmax[1] - bmin[1] + 1), tcols = (int)(bmax[0] - bmin[0] + 1);
    // an empty thickness map covers no pixel; a negative size is an error
    // (the reference's Mat would refuse it)
    PNT_REQUIRE(trows >= 0 && tcols >= 0,
                "stroke entirely off the canvas: its thickness map would have a negative size");
    if (trows == 0 || tcols == 0) return;
    const int64_t need = 4 * (int64_t)np + 1;
    if (need > b->geo_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->geo);
      b->geo = nullptr;
      b->geo = reinterpret_cast<double*>(dev_alloc(2 * need));
      b->geo_cap = need;
    }
    const int64_t tn = (int64_t)trows * tcols;
    if (tn > b->tmap_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->tmap);
      b->tmap = nullptr;
      b->tmap = reinterpret_cast<double*>(dev_alloc(2 * tn));
      b->tmap_cap = tn;
    }
    // polygon, texture coordinates, and a zero for the largest Vtex
    std::vector<double> hg((size_t)need, 0.0);
    std::memcpy(hg.data(), poly.data(), np * sizeof(double2));
    std::memcpy(hg.data() + 2 * np, uv.data(), np * sizeof(double2));
    PNT_HIP_CHECK(hipMemcpyAsync(b->geo, hg.data(), need * sizeof(double), hipMemcpyHostToDevice, s));
    launch_fill(reinterpret_cast<float*>(b->tmap), 2 * tn, 0.0f, s);
    TexArgs A{};
    A.canvas = cv->layer();
    A.rows = cv->rows;
    A.cols = cv->cols;
    A.poly = reinterpret_cast<const double2*>(b->geo);
    A.uv = reinterpret_cast<const double2*>(b->geo + 2 * np);
    A.np = np;
    A.thick = b->thick;
    A.t_rows = b->t_rows;
    A.t_cols = b->t_cols;
    A.scale = b->scale;
    A.bx0 = (int)bmin[0];
    A.by0 = (int)bmin[1];
    A.bx1 = (int)bmax[0];
    A.by1 = (int)bmax[1];
    A.mx0 = (int)bmin[0];
    A.my0 = (int)bmin[1];
    A.trows = trows;
    A.tcols = tcols;
    A.tmap = b->tmap;
    A.maxd = reinterpret_cast<unsigned long long*>(b->geo + 4 * np);
    for (int c = 0; c < 6; ++c) A.paint[c] = b->paint[c];
    A.clk = DryClock{};
    if (cv->tmap)  // Canvas::checkDry on every textured pixel (TextureBrush.hxx:133, 164)
      A.clk = DryClock{cv->tmap, canvas_now_us(cv), cv->drying_ms, cv->R0(), cv->h()};
    const int64_t npx = (int64_t)(A.bx1 - A.bx0 + 1) * (A.by1 - A.by0 + 1);
    k_tex_sample<<<blocks_for(npx), kBlock, 0, s>>>(A);
    PNT_HIP_CHECK(hipGetLastError());
    if (b->use_smudge && b->map > 0) {  // Smudge::smudge (Smudge.hxx:40-158)
      double maxD_h = 0.0;
      PNT_HIP_CHECK(hipMemcpyAsync(&maxD_h, b->geo + 4 * np, sizeof(double), hipMemcpyDeviceToHost,
                                   s));
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      // one turn and one trade per unit of length: bound the launches
      PNT_REQUIRE(!(length > 0.0) || 1.0 / length >= 1.0 / (double)(1 << 24),
                  "stroke too long for the smudge (more than 2^24 spine steps)");
      PNT_REQUIRE(!(length > 0.0) || length >= 1e-3,
                  "stroke shorter than 1e-3 pixels but not empty: too many smudge steps");
      const int64_t m = (int64_t)b->map * b->map;
      TradeArgs T{};
      T.canvas = cv->layer();
      T.rows = cv->rows;
      T.cols = cv->cols;
      T.bmin_x = bmin[0];
      T.bmin_y = bmin[1];
      T.radius = b->max_size * 0.5;
      T.tmap = b->tmap;
      T.trows = trows;
      T.tcols = tcols;
      T.maxD = maxD_h;
      for (double u = 0.0; maxD_h > 0.0 && u <= 1.0; u += 1. / length) {
        double center[2], t[2];
        spline_cr(v, u, false, center);
        spline_cr(v, u, true, t);
        normalize2(t);
        const double theta = std::atan2(t[1], t[0]);
        const double dtheta = normalize_angle(theta - b->cur_rot);
        b->cur_rot = theta;
        T.roi_x = (int)(center[0] - b->map / 2.0);
        T.roi_y = (int)(center[1] - b->map / 2.0);
        T.ctr_x = center[0];
        T.ctr_y = center[1];
        k_smudge_step<<<step_blocks_for(m), kStepBlock, 0, s>>>(
            b->maps(b->cur), b->maps(1 - b->cur), b->map, b->max_size / 2.0, b->max_size / 2.0,
            std::sin(dtheta), std::cos(dtheta), T);
        PNT_HIP_CHECK(hipGetLastError());
        b->cur = 1 - b->cur;
      }
    }
    k_tex_deposit<<<blocks_for(npx), kBlock, 0, s>>>(A);
    PNT_HIP_CHECK(hipGetLastError());
    PNT_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors above end here
  });
}

}  // extern "C"

// ===================================================================== GL --
// TextureBrushGpu (renderer/src/TextureBrushGpu.cxx:17-305): the reference's
// GL texture brush.  Its images are y-flipped (gpu/GpuMat.hxx:43-56,
// 107-132), so GL row g is canvas row H-1-g; the warped brush texture is
// kept in GL row order here too.  Where GL leaves behaviour to the
// implementation this follows one fixed choice:
//  - rasterisation: vertices snapped to 1/256 pixel; a pixel is covered when
//    its centre lies inside a triangle, edges by the top-left rule in window
//    coordinates (y up); texture coordinates interpolated linearly at the
//    centre; later triangles of the strip overwrite earlier ones;
//  - the brush texture is sampled with GL_LINEAR and GL_CLAMP_TO_EDGE;
//  - the smudge shader's writes to one pickup cell: every invocation reads
//    the pickup state from before the dispatch and the highest invocation
//    index wins.
// Shader arithmetic is FP32 (highp float), host geometry FP64.
namespace pnt {
namespace {

struct GlTri {  // a strip vertex: snapped window position, texture coordinates
  long long X, Y;
  float s, t;
};

__device__ __forceinline__ long long gl_cross(long long ax, long long ay, long long bx, long long by,
                                              long long px, long long py) {
  return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}
__device__ __forceinline__ bool gl_top_left(long long ax, long long ay, long long bx, long long by) {
  const long long dx = bx - ax, dy = by - ay;
  return (dy == 0 && dx < 0) || dy < 0;
}

struct GlWarpArgs {
  const GlTri* v;  // strip vertices
  int nv;
  int px0, py0, pw, ph;  // pixel box (GL coordinates) the strip can cover
  int cols;
  const float* tex;  // brush texture as uploaded (t_rows x t_cols)
  int t_rows, t_cols;
  float* warp;       // canvas-sized, GL rows
};

__device__ double gl_sample(const GlWarpArgs& A, double s, double t) {
  const double x = s * A.t_cols - 0.5, y = t * A.t_rows - 0.5;
  const double fx0 = floor(x), fy0 = floor(y);
  const double ax = x - fx0, ay = y - fy0;
  const int x0 = (int)fx0, y0 = (int)fy0;
  const int xa = min(max(x0, 0), A.t_cols - 1), xb = min(max(x0 + 1, 0), A.t_cols - 1);
  const int ya = min(max(y0, 0), A.t_rows - 1), yb = min(max(y0 + 1, 0), A.t_rows - 1);
  const auto at = [&](int yy, int xx) { return (double)A.tex[(int64_t)yy * A.t_cols + xx]; };
  return (at(ya, xa) * (1.0 - ax) + at(ya, xb) * ax) * (1.0 - ay) +
         (at(yb, xa) * (1.0 - ax) + at(yb, xb) * ax) * ay;
}

// A strip triangle's pixel box and the offset of its pixels in the flattened
// claim launch.
struct GlBox {
  int tx0, ty0, tw, th;
  long long off;
};

// The triangle j of the strip (vertices j, j+1, j+2), counter-clockwise in
// window coordinates (y up); false when it has no area.
__device__ __forceinline__ bool gl_tri(const GlTri* v, int j, long long X[3], long long Y[3],
                                       int id[3], long long* area) {
  id[0] = j;
  id[1] = j + 1;
  id[2] = j + 2;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    X[q] = v[id[q]].X;
    Y[q] = v[id[q]].Y;
  }
  long long a = gl_cross(X[0], Y[0], X[1], Y[1], X[2], Y[2]);
  if (a == 0) return false;
  if (a < 0) {
    long long tx = X[1], ty = Y[1];
    X[1] = X[2];
    Y[1] = Y[2];
    X[2] = tx;
    Y[2] = ty;
    const int ti = id[1];
    id[1] = id[2];
    id[2] = ti;
    a = -a;
  }
  *area = a;
  return true;
}

// Edge functions of a pixel centre against triangle (X, Y); covered by the
// top-left rule.
__device__ __forceinline__ bool gl_covers(const long long X[3], const long long Y[3], long long cx,
                                          long long cy, long long e[3]) {
  e[0] = gl_cross(X[1], Y[1], X[2], Y[2], cx, cy);
  e[1] = gl_cross(X[2], Y[2], X[0], Y[0], cx, cy);
  e[2] = gl_cross(X[0], Y[0], X[1], Y[1], cx, cy);
  const bool in0 = e[0] > 0 || (e[0] == 0 && gl_top_left(X[1], Y[1], X[2], Y[2]));
  const bool in1 = e[1] > 0 || (e[1] == 0 && gl_top_left(X[2], Y[2], X[0], Y[0]));
  const bool in2 = e[2] > 0 || (e[2] == 0 && gl_top_left(X[0], Y[0], X[1], Y[1]));
  return in0 && in1 && in2;
}

// First pass: one thread per pixel of every triangle's box; a covered pixel
// is claimed by the latest triangle of the strip that covers it (draw order:
// later triangles overwrite earlier ones).
__global__ __launch_bounds__(kBlock) void k_gl_claim(const GlTri* v, const GlBox* bx, int ntri,
                                                     long long total, int cols, int* claim) {
  const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = ntri - 1;  // the triangle whose pixel range holds e
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (bx[mid].off <= e) lo = mid;
    else hi = mid - 1;
  }
  const GlBox B = bx[lo];
  const long long k = e - B.off;
  if (k >= (long long)B.tw * B.th) return;
  const int px = B.tx0 + (int)(k % B.tw), py = B.ty0 + (int)(k / B.tw);
  long long X[3], Y[3], area, ed[3];
  int id[3];
  if (!gl_tri(v, lo, X, Y, id, &area)) return;
  if (gl_covers(X, Y, (long long)px * 256 + 128, (long long)py * 256 + 128, ed))
    atomicMax(claim + (int64_t)py * cols + px, lo + 1);
}

// Second pass: one thread per pixel of the strip's box; the claiming
// triangle's texture coordinates at the pixel centre, sampled
// (BrushTextureWarp.frag: texture().r).  Clears the claim for the next
// stroke.
__global__ __launch_bounds__(kBlock) void k_gl_shade(GlWarpArgs A, int* claim) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (int64_t)A.pw * A.ph) return;
  const int px = A.px0 + (int)(e % A.pw), py = A.py0 + (int)(e / A.pw);
  const int64_t ip = (int64_t)py * A.cols + px;
  const int c = claim[ip];
  if (c <= 0) return;
  claim[ip] = 0;
  long long X[3], Y[3], area, ed[3];
  int id[3];
  gl_tri(A.v, c - 1, X, Y, id, &area);
  gl_covers(X, Y, (long long)px * 256 + 128, (long long)py * 256 + 128, ed);
  const double w0 = (double)ed[0] / (double)area, w1 = (double)ed[1] / (double)area,
               w2 = (double)ed[2] / (double)area;
  const double s = w0 * A.v[id[0]].s + w1 * A.v[id[1]].s + w2 * A.v[id[2]].s;
  const double t = w0 * A.v[id[0]].t + w1 * A.v[id[1]].t + w2 * A.v[id[2]].t;
  A.warp[ip] = (float)gl_sample(A, (double)(float)s, (double)(float)t);
}

struct GlStep {
  int tlx, tly;  // dispatch offset (GL coordinates)
  float cs, sn;  // cos / sin of the heading
};

struct GlSmudgeArgs {
  Layer7 canvas;
  float* light;
  int rows, cols;
  const float* warp;
  float scale;
  float* smK;  // m*m RGBA
  float* smS;
  int m, dw;   // pickup side; dispatch width (m rounded up to 16)
  float rc;    // rotation centre m * 0.5
  // Claims and candidates, two of each, by step parity.  Per cell:
  // (step << 32) | (invocation + 1); per invocation: 8 candidate values
  // (K rgb V, S rgb V).
  unsigned long long* owner[2];
  float* cand[2];
};

// The pickup cell ip as the step before left it, where that step claimed
// the cell (step > 1): the winning candidate.  nullptr: the map holds it.
__device__ __forceinline__ const float* gl_cell(const GlSmudgeArgs& A, unsigned long long step,
                                                int64_t ip) {
  if (step > 1) {
    const unsigned long long w = A.owner[(step - 1) & 1][ip];
    if ((w >> 32) == step - 1)
      return A.cand[(step - 1) & 1] + 8 * (int64_t)((w & 0xffffffffull) - 1);
  }
  return nullptr;
}

// Smudge.compute for one spine step.  Each invocation updates its canvas
// pixel and leaves a candidate for its pickup cell with a claim on the cell;
// the highest claiming invocation wins the cell.  The winners are applied to
// the map by the next step's launch (or k_gl_smudge_flush after the last):
// there, threads ip < m*m copy the previous step's winners into the map while
// every reader of a cell the previous step claimed takes the winner's
// candidate directly, so no thread reads a map cell that another writes.
__global__ __launch_bounds__(kBlock) void k_gl_smudge(GlSmudgeArgs A, GlStep St,
                                                      unsigned long long step) {
  const int64_t inv = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (inv >= (int64_t)A.dw * A.dw) return;
  if (inv < (int64_t)A.m * A.m) {  // apply the step before's winner of cell inv
    const float* o = gl_cell(A, step, inv);
    if (o) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        A.smK[4 * inv + c] = o[c];
        A.smS[4 * inv + c] = o[4 + c];
      }
    }
  }
  const int gx = (int)(inv % A.dw), gy = (int)(inv / A.dw);
  const int px = St.tlx + gx, pyg = St.tly + gy;
  if (px < 0 || pyg < 0 || px >= A.cols || pyg >= A.rows) return;
  const float vTex = A.scale * A.warp[(int64_t)pyg * A.cols + px];
  if (vTex < 10e-6f) return;
  const float ptx = (float)gx - A.rc, pty = (float)gy - A.rc;
  const float rx = (ptx * St.cs - pty * St.sn) + A.rc, ry = (ptx * St.sn + pty * St.cs) + A.rc;
  const int qx = (int)rx, qy = (int)ry;
  if (qx < 0 || qy < 0 || qx >= A.cols || qy >= A.rows) return;
  const int64_t ic = (int64_t)(A.rows - 1 - pyg) * A.cols + px;
  const bool inmap = qx < A.m && qy < A.m;
  const int64_t ip = inmap ? (int64_t)qy * A.m + qx : 0;
  float pK[3] = {0.0f, 0.0f, 0.0f}, pS[3] = {0.0f, 0.0f, 0.0f}, pV = 0.0f;
  if (inmap) {
    const float* o = gl_cell(A, step, ip);
    const float* k = o ? o : A.smK + 4 * ip;
    const float* sv = o ? o + 4 : A.smS + 4 * ip;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pK[c] = k[c];
      pS[c] = sv[c];
    }
    pV = k[3];
  }
  float cK[3], cS[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cK[c] = A.canvas.ks[c][ic];
    cS[c] = A.canvas.ks[3 + c][ic];
  }
  const float cV = (float)A.canvas.v[ic];
  const float cVl = cV * 0.7f * vTex, cVr = cV - cVl;
  const float pVl = pV * 0.7f * vTex, pVr = pV - pVl;
  const float pVnew = pVr + cVl;
  if (pVnew > 0.0f && inmap) {
    float* o = A.cand[step & 1] + 8 * inv;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o[c] = (pVr * pK[c] + cVl * cK[c]) / pVnew;
      o[4 + c] = (pVr * pS[c] + cVl * cS[c]) / pVnew;
    }
    o[3] = o[7] = fmaxf(pVnew, 0.0f);
    atomicMax(A.owner[step & 1] + ip, (step << 32) | (unsigned long long)(inv + 1));
  }
  const float cVnew = cVr + pVl;
  if (cVnew > 0.0f) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A.canvas.ks[c][ic] = (cVr * pK[c] + pVl * cK[c]) / cVnew;
      A.canvas.ks[3 + c][ic] = (cVr * pS[c] + pVl * cS[c]) / cVnew;
    }
    A.canvas.v[ic] = fmaxf(cVnew, 0.0f);
    A.light[ic] = vTex;
  }
}

// After the last step (number step - 1): its winners into the map.
__global__ __launch_bounds__(kBlock) void k_gl_smudge_flush(GlSmudgeArgs A,
                                                            unsigned long long step) {
  const int64_t ip = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (ip >= (int64_t)A.m * A.m) return;
  const float* o = gl_cell(A, step, ip);
  if (!o) return;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    A.smK[4 * ip + c] = o[c];
    A.smS[4 * ip + c] = o[4 + c];
  }
}

// PaintTextureFootprintOnCanvas.compute then ClearTextureFootprint.compute
// over the stroke's box rounded up to 16 x 16 workgroups.
__global__ __launch_bounds__(kBlock) void k_gl_imprint(Layer7 canvas, float* light, int rows,
                                                       int cols, float* warp, float scale,
                                                       float3 Kb, float3 Sb, int ox, int oy,
                                                       int gw, int gh) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (int64_t)gw * gh) return;
  const int gx = (int)(e % gw), gy = (int)(e / gw);
  const int px = ox + gx, pyg = rows - (oy + gy) - 1;
  if (px < 0 || pyg < 0 || px >= cols || pyg >= rows) return;
  const int64_t iw = (int64_t)pyg * cols + px;
  const int64_t ic = (int64_t)(rows - 1 - pyg) * cols + px;
  const float vCan = (float)canvas.v[ic], vTex = scale * warp[iw];
  const float vSum = vCan + vTex;
  if (!(vSum < 10e-6f)) {
    const float kb[3] = {Kb.x, Kb.y, Kb.z}, sb[3] = {Sb.x, Sb.y, Sb.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      canvas.ks[c][ic] = (vCan * canvas.ks[c][ic] + vTex * kb[c]) / vSum;
      canvas.ks[3 + c][ic] = (vCan * canvas.ks[3 + c][ic] + vTex * sb[c]) / vSum;
    }
    canvas.v[ic] = vSum;
    light[ic] = (vTex > 0.0f) ? vTex : vCan;
  }
  warp[iw] = 0.0f;
}

int round16(int v) { return (v + 15) / 16 * 16; }

}  // namespace
}  // namespace pnt

struct pnt_tbgl_s {
  pnt_ctx ctx = nullptr;
  double radius = 0.0;
  float paint[6] = {0, 0, 0, 0, 0, 0};
  float* tex = nullptr;
  int t_rows = 0, t_cols = 0;
  double scale = 1.0;
  bool smudge = false;  // TextureBrushGpu.hxx:63
  int m = 0;            // smudge window side
  float* smK = nullptr;
  float* smS = nullptr;
  float* warp = nullptr;  // canvas-sized, GL rows
  int w_rows = 0, w_cols = 0;
  int* claim = nullptr;    // canvas-sized, zero between strokes
  // per-stroke scratch
  GlTri* verts = nullptr;
  int64_t verts_cap = 0;
  GlBox* boxes = nullptr;
  int64_t boxes_cap = 0;
  unsigned long long* owner = nullptr;
  float* cand = nullptr;
  int64_t sm_cap = 0;  // cells for owner / invocations for cand
};

namespace pnt {
namespace {
void gl_reset_smudge(pnt_tbgl_s* b, int m) {
  hipStream_t s = b->ctx->stream;
  PNT_HIP_CHECK(hipStreamSynchronize(s));
  dev_free(b->smK);
  dev_free(b->smS);
  dev_free(b->owner);
  dev_free(b->cand);
  b->smK = b->smS = b->cand = nullptr;
  b->owner = nullptr;
  b->sm_cap = 0;
  b->m = m;
  const int64_t cells = (int64_t)m * m;
  b->smK = dev_alloc(4 * cells + 4);
  b->smS = dev_alloc(4 * cells + 4);
  launch_fill(b->smK, 4 * cells + 4, 0.0f, s);
  launch_fill(b->smS, 4 * cells + 4, 0.0f, s);
  const int dw = round16(m);
  b->owner = reinterpret_cast<unsigned long long*>(dev_alloc(4 * cells + 4));  // two by parity
  b->cand = dev_alloc(16 * (int64_t)dw * dw + 8);
  launch_fill(reinterpret_cast<float*>(b->owner), 4 * cells + 4, 0.0f, s);
  b->sm_cap = cells;
}
void gl_set_radius(pnt_tbgl_s* b, double radius) {  // TextureBrushGpu.cxx:48-60
  if (!(std::fabs(b->radius - radius) < 0.5)) {
    b->radius = radius;
    gl_reset_smudge(b, (int)(uint32_t)(2.0 * radius) + 1);
  }
}
}  // namespace
}  // namespace pnt

extern "C" {

int pnt_tbgl_create(pnt_ctx ctx, pnt_tbgl* out) {
  return guard_call([&] {
    PNT_REQUIRE(ctx && out, "NULL argument");
    ctx->use();
    auto* b = new pnt_tbgl_s;
    b->ctx = ctx;
    try {
      gl_set_radius(b, 5.0);  // TextureBrushGpu.cxx:45
    } catch (...) {
      dev_free(b->smK);
      dev_free(b->smS);
      dev_free(b->owner);
      dev_free(b->cand);
      delete b;
      throw;
    }
    ctx_retain(ctx);
    *out = b;
  });
}

int pnt_tbgl_destroy(pnt_tbgl b) {
  return guard_call([&] {
    if (!b) return;
    pnt_ctx ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : {(void*)b->tex, (void*)b->smK, (void*)b->smS, (void*)b->warp,
                    (void*)b->claim, (void*)b->verts, (void*)b->boxes, (void*)b->owner,
                    (void*)b->cand})
      dev_free(p);
    delete b;
    ctx_release(ctx);
  });
}

int pnt_tbgl_set_radius(pnt_tbgl b, double radius) {
  return guard_call([&] {
    PNT_REQUIRE(b, "NULL brush");
    PNT_REQUIRE(std::isfinite(radius) && radius >= 0.0 && radius < 2048.0,
                "radius must be in [0, 2048)");
    b->ctx->use();
    gl_set_radius(b, radius);
  });
}

int pnt_tbgl_dip(pnt_tbgl b, const double K[3], const double S[3]) {
  return guard_call([&] {
    PNT_REQUIRE(b && K && S, "NULL argument");
    for (int c = 0; c < 3; ++c) {
      b->paint[c] = (float)K[c];
      b->paint[3 + c] = (float)S[c];
    }
  });
}

int pnt_tbgl_set_texture(pnt_tbgl b, const double* img, int rows, int cols) {
  return guard_call([&] {
    PNT_REQUIRE(b && img, "NULL argument");
    PNT_REQUIRE(rows > 0 && cols > 0 && (int64_t)rows * cols < ((int64_t)1 << 28),
                "texture size out of range");
    const int64_t n = (int64_t)rows * cols;
    double lo = img[0], hi = img[0];
    for (int64_t i = 0; i < n; ++i) {
      PNT_REQUIRE(std::isfinite(img[i]), "texture values must be finite");
      lo = std::min(lo, img[i]);
      hi = std::max(hi, img[i]);
    }
    // TextureBrushDictionary.cxx:106-123: NORM_MINMAX, flipped, FP32
    const double sc = hi > lo ? 1.0 / (hi - lo) : 0.0;
    std::vector<float> t((size_t)n);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c)
        t[(size_t)(rows - 1 - r) * cols + c] = (float)((img[(int64_t)r * cols + c] - lo) * sc);
    b->ctx->use();
    hipStream_t s = b->ctx->stream;
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    dev_free(b->tex);
    b->tex = nullptr;
    b->tex = dev_alloc(n);
    PNT_HIP_CHECK(hipMemcpyAsync(b->tex, t.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    b->t_rows = rows;
    b->t_cols = cols;
  });
}

int pnt_tbgl_set_thickness_scale(pnt_tbgl b, double scale) {
  return guard_call([&] {
    PNT_REQUIRE(b && std::isfinite(scale), "invalid argument");
    b->scale = scale;
  });
}

int pnt_tbgl_enable_smudge(pnt_tbgl b, int enable) {
  return guard_call([&] {
    PNT_REQUIRE(b, "NULL brush");
    b->smudge = enable != 0;
  });
}

int pnt_tbgl_get_smudge(pnt_tbgl b, int* side, float* K4, float* S4) {
  return guard_call([&] {
    PNT_REQUIRE(b && side, "NULL argument");
    b->ctx->use();
    *side = b->m;
    const int64_t c = (int64_t)b->m * b->m;
    hipStream_t s = b->ctx->stream;
    if (K4) PNT_HIP_CHECK(hipMemcpyAsync(K4, b->smK, 4 * c * sizeof(float), hipMemcpyDeviceToHost, s));
    if (S4) PNT_HIP_CHECK(hipMemcpyAsync(S4, b->smS, 4 * c * sizeof(float), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int pnt_canvas_download_lighting(pnt_canvas cv, float* out) {
  return guard_call([&] {
    PNT_REQUIRE(cv && out, "NULL argument");
    cv->ctx->use();
    canvas_tick(cv);
    hipStream_t s = cv->ctx->stream;
    if (cv->light)
      PNT_HIP_CHECK(hipMemcpyAsync(out, cv->light, cv->n * sizeof(float), hipMemcpyDeviceToHost, s));
    PNT_HIP_CHECK(hipStreamSynchronize(s));
    if (!cv->light) std::fill(out, out + cv->n, 0.0f);
  });
}

int pnt_tbgl_paint_stroke(pnt_tbgl b, pnt_canvas cv, const double* xy, int64_t n) {
  return guard_call([&] {
    PNT_REQUIRE(b && cv && (xy || n == 0) && n >= 0, "invalid argument");
    PNT_REQUIRE(b->ctx->device == cv->ctx->device, "brush and canvas on different devices");
    PNT_REQUIRE(n <= (1 << 20), "too many vertices");
    for (int64_t i = 0; i < 2 * n; ++i) PNT_REQUIRE(std::isfinite(xy[i]), "non-finite vertex");
    if (n < 2) return;
    PNT_REQUIRE(b->tex, "no brush texture (pnt_tbgl_set_texture)");
    b->ctx->use();
    canvas_tick(cv);
    hipStream_t s = b->ctx->stream;
    if (cv->ctx->stream != s) PNT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
    const int rows = cv->rows, cols = cv->cols;
    cv->epoch++;
    if (!cv->light) {
      cv->light = dev_alloc(cv->n);
      launch_fill(cv->light, cv->n, 0.0f, s);
    }
    if (!b->warp || b->w_rows != rows || b->w_cols != cols) {  // TextureBrushGpu.cxx:170-178
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->warp);
      dev_free(b->claim);
      b->warp = nullptr;
      b->claim = nullptr;
      b->warp = dev_alloc((int64_t)rows * cols);
      b->claim = reinterpret_cast<int*>(dev_alloc((int64_t)rows * cols));
      launch_fill(b->warp, (int64_t)rows * cols, 0.0f, s);
      launch_fill(reinterpret_cast<float*>(b->claim), (int64_t)rows * cols, 0.0f, s);
      b->w_rows = rows;
      b->w_cols = cols;
    }
    const double r = b->radius;
    std::vector<double2> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[(size_t)i] = make_double2(xy[2 * i], xy[2 * i + 1]);
    double bmin[2] = {xy[0], xy[1]}, bmax[2] = {xy[0], xy[1]};
    for (int64_t i = 0; i < n; ++i) {
      bmin[0] = std::min(bmin[0], xy[2 * i]);
      bmin[1] = std::min(bmin[1], xy[2 * i + 1]);
      bmax[0] = std::max(bmax[0], xy[2 * i]);
      bmax[1] = std::max(bmax[1], xy[2 * i + 1]);
    }
    bmin[0] = std::max(bmin[0] - r, 0.0);
    bmax[0] = std::min(bmax[0] + r, (double)(cols - 1));
    bmin[1] = std::max(bmin[1] - r, 0.0);
    bmax[1] = std::min(bmax[1] + r, (double)(rows - 1));
    // the strip r0 l0 r1 l1 ... (generateWarpedTexture, TextureBrushGpu.cxx:186-212)
    const int nv = (int)(2 * n);
    std::vector<GlTri> tri((size_t)nv);
    long long xmin = LLONG_MAX, xmax = LLONG_MIN, ymin = LLONG_MAX, ymax = LLONG_MIN;
    for (int i = 0; i < (int)n; ++i) {
      const double u = (double)i / (double)(n - 1);
      double c[2], t[2];
      spline_cr(v, u, false, c);
      spline_cr(v, u, true, t);
      normalize2(t);
      const double d[2] = {-t[1], t[0]};
      const float rx = (float)(c[0] + r * d[0]), ry = (float)(c[1] + r * d[1]);
      const float lx = (float)(c[0] - r * d[0]), ly = (float)(c[1] - r * d[1]);
      const float pos[2][2] = {{rx, ry}, {lx, ly}};
      for (int k = 0; k < 2; ++k) {
        GlTri& g = tri[(size_t)(2 * i + k)];
        g.X = (long long)std::llrint((double)pos[k][0] * 256.0);
        g.Y = (long long)std::llrint((double)((float)rows - pos[k][1]) * 256.0);
        g.s = (float)u;
        g.t = k == 0 ? 0.0f : 1.0f;
        xmin = std::min(xmin, g.X);
        xmax = std::max(xmax, g.X);
        ymin = std::min(ymin, g.Y);
        ymax = std::max(ymax, g.Y);
      }
    }
    const int px0 = (int)std::max<long long>(0, (xmin - 128) >> 8);
    const int px1 = (int)std::min<long long>(cols - 1, ((xmax - 128) >> 8) + 1);
    const int py0 = (int)std::max<long long>(0, (ymin - 128) >> 8);
    const int py1 = (int)std::min<long long>(rows - 1, ((ymax - 128) >> 8) + 1);
    if (nv > b->verts_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->verts);
      b->verts = nullptr;
      b->verts = reinterpret_cast<GlTri*>(dev_alloc((int64_t)nv * (int64_t)(sizeof(GlTri) / 4)));
      b->verts_cap = nv;
    }
    PNT_HIP_CHECK(hipMemcpyAsync(b->verts, tri.data(), nv * sizeof(GlTri), hipMemcpyHostToDevice, s));
    // each triangle's pixel box (the centres it may cover), flattened
    const int ntri = nv - 2;
    std::vector<GlBox> bx((size_t)std::max(ntri, 1));
    long long total = 0;
    for (int j = 0; j < ntri; ++j) {
      long long x0 = LLONG_MAX, x1 = LLONG_MIN, y0 = LLONG_MAX, y1 = LLONG_MIN;
      for (int q = 0; q < 3; ++q) {
        x0 = std::min(x0, tri[(size_t)(j + q)].X);
        x1 = std::max(x1, tri[(size_t)(j + q)].X);
        y0 = std::min(y0, tri[(size_t)(j + q)].Y);
        y1 = std::max(y1, tri[(size_t)(j + q)].Y);
      }
      const long long tx0 = std::max<long long>(0, (x0 - 128) >> 8);
      const long long tx1 = std::min<long long>(cols - 1, ((x1 - 128) >> 8) + 1);
      const long long ty0 = std::max<long long>(0, (y0 - 128) >> 8);
      const long long ty1 = std::min<long long>(rows - 1, ((y1 - 128) >> 8) + 1);
      GlBox& B = bx[(size_t)j];
      B.tx0 = (int)std::min<long long>(tx0, cols - 1);
      B.ty0 = (int)std::min<long long>(ty0, rows - 1);
      B.tw = tx1 >= tx0 ? (int)(tx1 - tx0 + 1) : 0;
      B.th = ty1 >= ty0 ? (int)(ty1 - ty0 + 1) : 0;
      B.off = total;
      total += (long long)B.tw * B.th;
    }
    PNT_REQUIRE(total < ((long long)1 << 40), "stroke too large");
    if (ntri > b->boxes_cap) {
      PNT_HIP_CHECK(hipStreamSynchronize(s));
      dev_free(b->boxes);
      b->boxes = nullptr;
      b->boxes = reinterpret_cast<GlBox*>(dev_alloc((int64_t)ntri * (int64_t)(sizeof(GlBox) / 4)));
      b->boxes_cap = ntri;
    }
    PNT_HIP_CHECK(hipMemcpyAsync(b->boxes, bx.data(), ntri * sizeof(GlBox), hipMemcpyHostToDevice,
                                 s));
    if (total > 0 && px1 >= px0 && py1 >= py0) {
      k_gl_claim<<<(unsigned)((total + kBlock - 1) / kBlock), kBlock, 0, s>>>(b->verts, b->boxes,
                                                                              ntri, total, cols,
                                                                              b->claim);
      PNT_HIP_CHECK(hipGetLastError());
      GlWarpArgs W{b->verts, nv, px0, py0, px1 - px0 + 1, py1 - py0 + 1, cols, b->tex,
                   b->t_rows, b->t_cols, b->warp};
      k_gl_shade<<<blocks_for((int64_t)W.pw * W.ph), kBlock, 0, s>>>(W, b->claim);
      PNT_HIP_CHECK(hipGetLastError());
    }
    if (b->smudge) {  // TextureBrushGpu::smudge (TextureBrushGpu.cxx:253-305)
      double length = 0.0;
      for (int64_t i = 1; i < n; ++i)
        length += std::sqrt((xy[2 * i] - xy[2 * i - 2]) * (xy[2 * i] - xy[2 * i - 2]) +
                            (xy[2 * i + 1] - xy[2 * i - 1]) * (xy[2 * i + 1] - xy[2 * i - 1]));
      PNT_REQUIRE(!(length > 0.0) || length <= (double)(1 << 22),
                  "stroke too long for the smudge");
      PNT_REQUIRE(!(length > 0.0) || length >= 1e-3,
                  "stroke shorter than 1e-3 pixels but not empty: too many smudge steps");
      GlSmudgeArgs A{};
      A.canvas = cv->layer();
      A.light = cv->light;
      A.rows = rows;
      A.cols = cols;
      A.warp = b->warp;
      A.scale = (float)b->scale;
      A.smK = b->smK;
      A.smS = b->smS;
      A.m = b->m;
      A.dw = round16(b->m);
      A.rc = (float)(b->m * 0.5);
      A.owner[0] = b->owner;
      A.owner[1] = b->owner + (int64_t)A.m * A.m;
      A.cand[0] = b->cand;
      A.cand[1] = b->cand + 8 * (int64_t)A.dw * A.dw;
      unsigned long long step = 0;
      int lastx = -1, lasty = -1;
      for (double u = 0.0; u <= 1.0; u += (0.25 / length)) {
        double c[2], t[2];
        spline_cr(v, u, false, c);
        if ((int)c[0] == lastx && (int)c[1] == lasty) continue;
        lastx = (int)c[0];
        lasty = (int)c[1];
        spline_cr(v, u, true, t);
        normalize2(t);
        const float th = (float)std::atan2(t[1], t[0]);
        GlStep St{(int)((double)(int)c[0] - (double)b->m * 0.5),
                  (int)((double)(int)c[1] - (double)b->m * 0.5), std::cos(th), std::sin(th)};
        ++step;
        PNT_REQUIRE(step < (1ull << 31), "too many smudge steps");
        k_gl_smudge<<<blocks_for((int64_t)A.dw * A.dw), kBlock, 0, s>>>(A, St, step);
        PNT_HIP_CHECK(hipGetLastError());
      }
      if (step > 0) {
        k_gl_smudge_flush<<<blocks_for((int64_t)A.m * A.m), kBlock, 0, s>>>(A, step + 1);
        PNT_HIP_CHECK(hipGetLastError());
      }
      // the claims carry the step number; start the next stroke from zero
      launch_fill(reinterpret_cast<float*>(b->owner), 4 * (int64_t)b->m * b->m + 4, 0.0f, s);
    }
    const int aw = (int)(bmax[0] - bmin[0]) + 1, ah = (int)(bmax[1] - bmin[1]) + 1;
    if (aw > 0 && ah > 0) {
      const int gw = round16(aw), gh = round16(ah);
      k_gl_imprint<<<blocks_for((int64_t)gw * gh), kBlock, 0, s>>>(
          cv->layer(), cv->light, rows, cols, b->warp, (float)b->scale,
          make_float3(b->paint[0], b->paint[1], b->paint[2]),
          make_float3(b->paint[3], b->paint[4], b->paint[5]), (int)bmin[0], (int)bmin[1], gw, gh);
      PNT_HIP_CHECK(hipGetLastError());
    }
    PNT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
