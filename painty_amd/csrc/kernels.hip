// gfx950 kernels: KM composite, drying, imprint stamping.
//
// Composite and dry are streaming kernels over the layer planes (FP32 K/S,
// FP64 V; HBM-bound): one thread handles four consecutive pixels with
// 16-byte loads from every FP32 plane and 32 bytes of volumes.
// Stamping follows renderer/FootprintBrush.hxx:73-143 visit by visit; see the
// comment above k_stamp_main for how the reference's serial order is kept.
#include <algorithm>

#include "internal.hpp"
#include "km.hpp"
#include "visit.hpp"
#include "dry.hpp"

namespace pnt {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float4 ld4(const float* p, int64_t i) {
  return *reinterpret_cast<const float4*>(p + 4 * i);
}
__device__ __forceinline__ void st4(float* p, int64_t i, const float4& v) {
  *reinterpret_cast<float4*>(p + 4 * i) = v;
}
__device__ __forceinline__ float& lane(float4& v, int j) {
  return reinterpret_cast<float*>(&v)[j];
}

// Four FP64 volumes of pixels 4i..4i+3 (32-byte aligned).
__device__ __forceinline__ void ldd4(const double* p, int64_t i, double v[4]) {
  const double2 a = reinterpret_cast<const double2*>(p)[2 * i];
  const double2 b = reinterpret_cast<const double2*>(p)[2 * i + 1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void std4(double* p, int64_t i, const double v[4]) {
  reinterpret_cast<double2*>(p)[2 * i] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(p)[2 * i + 1] = make_double2(v[2], v[3]);
}

// ------------------------------------------------------------ composite ---
// The K/S planes and the thickness (FP32 for the KM evaluation) of a
// canvas layer (FP64 volume) or of a caller's seven FP32 planes.
__device__ __forceinline__ const float* ksp(const Layer7& L, int c) { return L.ks[c]; }
__device__ __forceinline__ const float* ksp(const Plane7& L, int c) { return L.p[c]; }
__device__ __forceinline__ float4 thick4(const Layer7& L, int64_t i) {
  double v[4];
  ldd4(L.v, i, v);
  return make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}
__device__ __forceinline__ float4 thick4(const Plane7& L, int64_t i) { return ld4(L.p[6], i); }
__device__ __forceinline__ float thick1(const Layer7& L, int64_t i) { return (float)L.v[i]; }
__device__ __forceinline__ float thick1(const Plane7& L, int64_t i) { return L.p[6][i]; }

// Applies one layer (K, S, V planes) to the reflectance r of pixels 4i..4i+3.
template <class LT>
__device__ __forceinline__ void apply_layer4(const LT& L, int64_t i, float4 r[3]) {
  float4 k[3], s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    k[c] = ld4(ksp(L, c), i);
    s[c] = ld4(ksp(L, 3 + c), i);
  }
  float4 v = thick4(L, i);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float K[3], S[3], R[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      K[c] = lane(k[c], j);
      S[c] = lane(s[c], j);
      R[c] = lane(r[c], j);
    }
    km_pixel(K, S, R, lane(v, j));
#pragma unroll
    for (int c = 0; c < 3; ++c) lane(r[c], j) = R[c];
  }
}

template <class LT>
__device__ __forceinline__ void apply_layer1(const LT& L, int64_t i, float R[3]) {
  float K[3], S[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = ksp(L, c)[i];
    S[c] = ksp(L, 3 + c)[i];
  }
  km_pixel(K, S, R, thick1(L, i));
}

// Renderer::compose (Renderer.hxx:26-41) for kLayers == 1; for kLayers == 2
// the dry layer A is composed first (PaintLayer::composeOnto, :81-96) and the
// wet layer B over the result, without writing the intermediate.
template <int kLayers, class LT>
__global__ __launch_bounds__(kBlock) void k_compose_vec(int64_t n4, LT A, LT B, Rgb3 R0, Rgb3 R) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n4) return;
  float4 r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = ld4(R0.p[c], i);
  apply_layer4(A, i, r);
  if (kLayers == 2) apply_layer4(B, i, r);
#pragma unroll
  for (int c = 0; c < 3; ++c) st4(R.p[c], i, r[c]);
}

template <int kLayers, class LT>
__global__ __launch_bounds__(kBlock) void k_compose_scalar(int64_t begin, int64_t n, LT A, LT B,
                                                           Rgb3 R0, Rgb3 R) {
  const int64_t i = begin + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = R0.p[c][i];
  apply_layer1(A, i, r);
  if (kLayers == 2) apply_layer1(B, i, r);
#pragma unroll
  for (int c = 0; c < 3; ++c) R.p[c][i] = r[c];
}

// ------------------------------------------------------------------ dry ---
// Canvas::dryCanvas (Canvas.hxx:105-121): h += V; R0 = KM(K, S, R0, V);
// V = 0; K = S = 0.
__device__ __forceinline__ void dry1(const Layer7& L, const Rgb3& R0, float* h, int64_t i) {
  float K[3], S[3], R[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = L.ks[c][i];
    S[c] = L.ks[3 + c][i];
    R[c] = R0.p[c][i];
  }
  const double v = L.v[i];
  h[i] = h_plus(h[i], v);
  km_pixel(K, S, R, (float)v);
#pragma unroll
  for (int c = 0; c < 3; ++c) R0.p[c][i] = R[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) L.ks[c][i] = 0.0f;
  L.v[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_dry_vec(int64_t n4, Layer7 L, Rgb3 R0, float* h) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n4) return;
  float4 r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = ld4(R0.p[c], i);
  float4 hv = ld4(h, i);
  double v[4];
  ldd4(L.v, i, v);
#pragma unroll
  for (int j = 0; j < 4; ++j) lane(hv, j) = h_plus(lane(hv, j), v[j]);
  apply_layer4(L, i, r);
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int c = 0; c < 3; ++c) st4(R0.p[c], i, r[c]);
#pragma unroll
  for (int c = 0; c < 6; ++c) st4(L.ks[c], i, z);
  const double zd[4] = {0.0, 0.0, 0.0, 0.0};
  std4(L.v, i, zd);
  st4(h, i, hv);
}

__global__ __launch_bounds__(kBlock) void k_dry_scalar(int64_t begin, int64_t n, Layer7 L,
                                                       Rgb3 R0, float* h) {
  const int64_t i = begin + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dry1(L, R0, h, i);
}

// DryingStep.compute.glsl:66-76: V leaves up to `portion`; R0 takes it.
// The volume arithmetic is FP64 (V is); the shader's KM runs in FP32.
__device__ __forceinline__ double dry_step_px(double d, float portion, float K[3], float S[3],
                                              float R[3]) {
#pragma clang fp contract(off)
  const double rem = fmax(0.0, d - (double)portion);
  const float leave = (float)(d - rem);
  if (__builtin_fabsf(leave) > 10e-6f) km_pixel_glsl(K, S, R, leave);
  return rem;
}

__global__ __launch_bounds__(kBlock) void k_dry_step(int64_t n, Layer7 L, Rgb3 R0,
                                                     float portion) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float K[3], S[3], R[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = L.ks[c][i];
    S[c] = L.ks[3 + c][i];
    R[c] = R0.p[c][i];
  }
  L.v[i] = dry_step_px(L.v[i], portion, K, S, R);
#pragma unroll
  for (int c = 0; c < 3; ++c) R0.p[c][i] = R[c];
}

// `steps` applications of k_dry_step's per-pixel update, stopping early once
// the pixel holds no paint (further steps change nothing).
__global__ __launch_bounds__(kBlock) void k_dry_steps(int64_t n, Layer7 L, Rgb3 R0, float portion,
                                                      int steps, unsigned* wet) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool still = false;
  if (i < n) {
    double d = L.v[i];
    if (d != 0.0) {  // V = 0 stays as it is under any number of steps
      float K[3], S[3], R[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        K[c] = L.ks[c][i];
        S[c] = L.ks[3 + c][i];
        R[c] = R0.p[c][i];
      }
      for (int k = 0; k < steps; ++k) {
        d = dry_step_px(d, portion, K, S, R);
        if (!(d > 0.0)) break;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) R0.p[c][i] = R[c];
      L.v[i] = d;
      still = d > 0.0;
    }
  }
  // one atomic per wave that still holds paint
  if (__ballot(still) != 0 && (threadIdx.x & 63) == 0) atomicOr(wet, 1u);
}

// ComputeReflectance over AoS arrays (KubelkaMunk.hxx:25-83).
__global__ __launch_bounds__(kBlock) void k_reflectance_aos(int64_t n, const float* K,
                                                            const float* S, const float* R0,
                                                            const float* d, float* R) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float k[3], s[3], r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    k[c] = K[3 * i + c];
    s[c] = S[3 * i + c];
    r[c] = R0[3 * i + c];
  }
  km_pixel(k, s, r, d[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) R[3 * i + c] = r[c];
}

__global__ __launch_bounds__(kBlock) void k_fill(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// PaintLayer::copyTo: pixels 4i..4i+3 of every plane per thread (16-byte
// FP32 and 32-byte FP64 accesses), and the ragged tail one pixel a thread.
__global__ __launch_bounds__(kBlock) void k_copy_layer(int64_t n4, Layer7 src, Layer7 dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n4) return;
  float4 k[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) k[c] = ld4(src.ks[c], i);
  double v[4];
  ldd4(src.v, i, v);
#pragma unroll
  for (int c = 0; c < 6; ++c) st4(dst.ks[c], i, k[c]);
  std4(dst.v, i, v);
}

__global__ __launch_bounds__(kBlock) void k_copy_layer_tail(int64_t i0, int64_t n, Layer7 src,
                                                            Layer7 dst) {
  const int64_t i = i0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.ks[c][i] = src.ks[c][i];
  dst.v[i] = src.v[i];
}

// A layer from seven FP32 planes (V widened exactly) and back (V rounded).
__global__ __launch_bounds__(kBlock) void k_layer_from_planes(int64_t n, Plane7 src, Layer7 dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.ks[c][i] = src.p[c][i];
  dst.v[i] = (double)src.p[6][i];
}

__global__ __launch_bounds__(kBlock) void k_planes_from_layer(int64_t n, Layer7 src, Plane7 dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst.p[c][i] = src.ks[c][i];
  dst.p[6][i] = (float)src.v[i];
}

// ------------------------------------------------------------ stamping ---
// One visit of an imprint for canvas pixel ic and map cell im: the shared
// visit arithmetic (visit.hpp) on the pickup map in global memory.
struct GlobalVisit {
  const StampArgs& a;
  int64_t ic;
  int im;
  __device__ float pick(int c) const { return a.b.pick.p[c][im]; }
  __device__ void set_pick(int c, float v) const { a.b.pick.p[c][im] = v; }
  __device__ float src(int c) const { return a.src.ks[c][ic]; }
  __device__ double src_v() const { return a.src.v[ic]; }
  __device__ void set_src_v(double v) const { a.src.v[ic] = v; }
  __device__ float canvas(int c) const { return a.canvas.ks[c][ic]; }
  __device__ double canvas_v() const { return a.canvas.v[ic]; }
  __device__ void set_canvas(int c, float v) const { a.canvas.ks[c][ic] = v; }
  __device__ void set_canvas_v(double v) const { a.canvas.v[ic] = v; }
};

// Canvas::checkDry of one pixel, called directly (Canvas.hxx:123-155).
__global__ __launch_bounds__(64) void k_check_dry_pixel(DryClock c, Layer7 canvas, int64_t ic) {
  if (threadIdx.x == 0) check_dry(c, canvas, ic);
}

__device__ __forceinline__ void paint_visit(const StampArgs& a, int64_t ic, int im) {
  const float fh = a.b.fp[im];
  if (!(fh > 0.0f)) return;
  GlobalVisit acc{a, ic, im};
  paint_visit_math(acc, fh, a.b.pickup_rate, a.b.dep_rate, a.b.cap, a.b.paintK, a.b.paintS);
}

// updateSnapshot(canvas, center), FootprintBrush.hxx:307-318: copy the wet
// layer into the snapshot over the allowed box except the window interior.
__global__ __launch_bounds__(kBlock) void k_ring_refresh(RingBox b, int cols, Layer7 canvas,
                                                         Layer7 snap) {
  const int w = b.a_x1 - b.a_x0 + 1;
  const int hgt = b.a_y1 - b.a_y0 + 1;
  const int64_t total = (int64_t)w * hgt;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kBlock) {
    const int row = b.a_y0 + (int)(e / w);
    const int col = b.a_x0 + (int)(e % w);
    if (row > b.tl_y && row < b.br_y && col > b.tl_x && col < b.br_x) continue;
    const int64_t i = (int64_t)row * cols + col;
#pragma unroll
    for (int c = 0; c < 6; ++c) snap.ks[c][i] = canvas.ks[c][i];
    snap.v[i] = canvas.v[i];
  }
}

// One imprint's visits, spread over the GPU.
//
// The reference visits window pixels in row-major order.  Two visits can
// touch the same state only when
//   (a) they round to the same pickup-map cell: rotation preserves distances
//       and a half-open unit cell holds no two points >= 1 apart except at
//       distance exactly 1, so such pairs are 4-neighbours in the window and a
//       cell has at most two visitors; or
//   (b) they truncate to the same canvas pixel: only column 0 / row 0 when the
//       stamp straddles the left / top edge with a fractional centre.
// Each workgroup takes a 16x16 tile of the window and computes the map cells
// of the tile plus a one-pixel halo in LDS.  A visit whose earlier neighbour
// (left or up) shares its cell is skipped; the earlier one runs both, in
// order.  Under (b), visits on canvas column/row 0 and their partners are
// left to k_stamp_serial, which runs them in reference order by dependency
// level.  No other visit shares state with them, so the two kernels commute.
constexpr int kTile = 16;

__device__ __forceinline__ bool on_dup_line(const StampArgs& a, int col, int row) {
  return (a.dupx && (int)(col + a.g.cx) == 0) || (a.dupy && (int)(row + a.g.cy) == 0);
}

__global__ __launch_bounds__(kTile* kTile) void k_stamp_main(StampArgs a, int tiles_x) {
  __shared__ int sm[kTile + 2][kTile + 2];
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int col0 = -a.g.hr + tx * kTile, row0 = -a.g.hr + ty * kTile;
  for (int e = threadIdx.x; e < (kTile + 2) * (kTile + 2); e += kTile * kTile) {
    const int ly = e / (kTile + 2), lx = e % (kTile + 2);
    int64_t ic;
    sm[ly][lx] = stamp_visit(a.g, col0 + lx - 1, row0 + ly - 1, &ic);
  }
  __syncthreads();
  const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
  const int m = sm[ly + 1][lx + 1];
  if (m < 0) return;
  if (sm[ly + 1][lx] == m || sm[ly][lx + 1] == m) return;  // earlier visitor owns the cell
  const int col = col0 + lx, row = row0 + ly;
  int pcol = col, prow = row;
  bool partner = false;
  if (sm[ly + 1][lx + 2] == m) {
    pcol = col + 1;
    partner = true;
  } else if (sm[ly + 2][lx + 1] == m) {
    prow = row + 1;
    partner = true;
  }
  if ((a.dupx | a.dupy) && (on_dup_line(a, col, row) || (partner && on_dup_line(a, pcol, prow))))
    return;
  int64_t ic, pic = 0;
  stamp_visit(a.g, col, row, &ic);
  if (partner) stamp_visit(a.g, pcol, prow, &pic);
  if (a.clk.tmap) check_dry(a.clk, a.canvas, ic);
  paint_visit(a, ic, m);
  if (partner) {
    if (a.clk.tmap) check_dry(a.clk, a.canvas, pic);
    paint_visit(a, pic, m);
  }
}

// Visits listed in reference order, grouped by level: a visit's level is one
// more than the highest level of the earlier visits sharing its map cell or
// canvas pixel.  One workgroup; levels are separated by barriers.
__global__ __launch_bounds__(kBlock) void k_stamp_serial(StampArgs a, const int2* list,
                                                         const int* level_off, int nlevels) {
  for (int L = 0; L < nlevels; ++L) {
    for (int k = level_off[L] + threadIdx.x; k < level_off[L + 1]; k += kBlock) {
      const int2 v = list[k];
      int64_t ic;
      const int m = stamp_visit(a.g, v.x, v.y, &ic);
      if (m >= 0) {
        if (a.clk.tmap) check_dry(a.clk, a.canvas, ic);
        paint_visit(a, ic, m);
      }
    }
    __syncthreads();
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }
inline bool aligned16(const Layer7& L) {
  for (auto* p : L.ks)
    if (!aligned16(p)) return false;
  return ((uintptr_t)L.v & 31u) == 0;
}
inline bool aligned16(const Plane7& L) {
  for (auto* p : L.p)
    if (!aligned16(p)) return false;
  return true;
}
inline bool aligned16(const Rgb3& L) {
  for (auto* p : L.p)
    if (!aligned16(p)) return false;
  return true;
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <int kL, class LT>
static void launch_compose_t(int64_t n, const LT& A, const LT& B, const Rgb3& R0, const Rgb3& R,
                             hipStream_t s) {
  if (n <= 0) return;
  int64_t done = 0;
  if (aligned16(A) && aligned16(B) && aligned16(R0) && aligned16(R)) {
    const int64_t n4 = n / 4;
    if (n4 > 0) k_compose_vec<kL, LT><<<blocks_for(n4), kBlock, 0, s>>>(n4, A, B, R0, R);
    done = n4 * 4;
  }
  if (done < n)
    k_compose_scalar<kL, LT><<<blocks_for(n - done), kBlock, 0, s>>>(done, n, A, B, R0, R);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_compose(int64_t n, const Layer7& L, const Rgb3& R0, const Rgb3& R, hipStream_t s) {
  launch_compose_t<1>(n, L, L, R0, R, s);
}
void launch_compose(int64_t n, const Plane7& L, const Rgb3& R0, const Rgb3& R, hipStream_t s) {
  launch_compose_t<1>(n, L, L, R0, R, s);
}

void launch_compose2(int64_t n, const Layer7& dry, const Layer7& wet, const Rgb3& R0,
                     const Rgb3& R, hipStream_t s) {
  launch_compose_t<2>(n, dry, wet, R0, R, s);
}
void launch_compose2(int64_t n, const Plane7& dry, const Plane7& wet, const Rgb3& R0,
                     const Rgb3& R, hipStream_t s) {
  launch_compose_t<2>(n, dry, wet, R0, R, s);
}

void launch_dry(int64_t n, const Layer7& L, const Rgb3& R0, float* h, hipStream_t s) {
  if (n <= 0) return;
  int64_t done = 0;
  if (aligned16(L) && aligned16(R0) && aligned16(h)) {
    const int64_t n4 = n / 4;
    if (n4 > 0) k_dry_vec<<<blocks_for(n4), kBlock, 0, s>>>(n4, L, R0, h);
    done = n4 * 4;
  }
  if (done < n) k_dry_scalar<<<blocks_for(n - done), kBlock, 0, s>>>(done, n, L, R0, h);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_dry_step(int64_t n, const Layer7& L, const Rgb3& R0, float portion, hipStream_t s) {
  if (n <= 0) return;
  k_dry_step<<<blocks_for(n), kBlock, 0, s>>>(n, L, R0, portion);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_reflectance_aos(int64_t n, const float* K, const float* S, const float* R0,
                            const float* d, float* R, hipStream_t s) {
  if (n <= 0) return;
  k_reflectance_aos<<<blocks_for(n), kBlock, 0, s>>>(n, K, S, R0, d, R);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_dry_steps(int64_t n, const Layer7& L, const Rgb3& R0, float portion, int steps,
                      unsigned* wet, hipStream_t s) {
  if (n <= 0 || steps <= 0) return;
  k_dry_steps<<<blocks_for(n), kBlock, 0, s>>>(n, L, R0, portion, steps, wet);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_fill(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return;
  k_fill<<<blocks_for(n), kBlock, 0, s>>>(p, n, v);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_check_dry_pixel(const DryClock& c, const Layer7& canvas, int64_t ic, hipStream_t s) {
  k_check_dry_pixel<<<1, 64, 0, s>>>(c, canvas, ic);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_i64<<<blocks_for(n), kBlock, 0, s>>>(p, n, v);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_layer_from_planes(int64_t n, const Plane7& src, const Layer7& dst, hipStream_t s) {
  if (n <= 0) return;
  k_layer_from_planes<<<blocks_for(n), kBlock, 0, s>>>(n, src, dst);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_planes_from_layer(int64_t n, const Layer7& src, const Plane7& dst, hipStream_t s) {
  if (n <= 0) return;
  k_planes_from_layer<<<blocks_for(n), kBlock, 0, s>>>(n, src, dst);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_copy_layer(int64_t n, const Layer7& src, const Layer7& dst, hipStream_t s) {
  if (n <= 0) return;
  int64_t done = 0;
  if (aligned16(src) && aligned16(dst)) {
    const int64_t n4 = n / 4;
    if (n4 > 0) k_copy_layer<<<blocks_for(n4), kBlock, 0, s>>>(n4, src, dst);
    done = n4 * 4;
  }
  if (done < n) k_copy_layer_tail<<<blocks_for(n - done), kBlock, 0, s>>>(done, n, src, dst);
  PNT_HIP_CHECK(hipGetLastError());
}

RingBox ring_box(const StampGeom& g, double radius) {
#pragma clang fp contract(off)
  RingBox b;
  b.tl_x = (int)(g.cx - g.hr);
  b.tl_y = (int)(g.cy - g.hr);
  b.br_x = (int)(g.cx + g.hr);
  b.br_y = (int)(g.cy + g.hr);
  b.a_x0 = std::max((int)(g.cx - g.hr - radius), 0);
  b.a_y0 = std::max((int)(g.cy - g.hr - radius), 0);
  b.a_x1 = std::min((int)(g.cx + g.hr + radius), g.cols - 1);
  b.a_y1 = std::min((int)(g.cy + g.hr + radius), g.rows - 1);
  return b;
}

void launch_ring_refresh(const RingBox& box, int cols, const Layer7& canvas, const Layer7& snap,
                         hipStream_t s) {
  if (box.a_x1 < box.a_x0 || box.a_y1 < box.a_y0) return;
  const int64_t total = (int64_t)(box.a_x1 - box.a_x0 + 1) * (box.a_y1 - box.a_y0 + 1);
  const unsigned grid = (unsigned)std::min<int64_t>(blocks_for(total), 4096);
  k_ring_refresh<<<grid, kBlock, 0, s>>>(box, cols, canvas, snap);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_stamp_main(const StampArgs& a, hipStream_t s) {
  const int side = 2 * a.g.hr + 1;
  const int tiles = (side + kTile - 1) / kTile;
  k_stamp_main<<<(unsigned)(tiles * tiles), kTile * kTile, 0, s>>>(a, tiles);
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_stamp_serial(const StampArgs& a, const int2* list, const int* level_off, int nlevels,
                         hipStream_t s) {
  if (nlevels <= 0) return;
  k_stamp_serial<<<1, kBlock, 0, s>>>(a, list, level_off, nlevels);
  PNT_HIP_CHECK(hipGetLastError());
}

}  // namespace pnt
