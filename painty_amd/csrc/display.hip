// Display output: Renderer::render (renderer/Renderer.hxx:60-156) -- the
// composite lit by a point light with a Beckmann / Cook-Torrance specular
// term over the height field of the wet layer's thickness V -- and the pixel
// transform of io::imSave (io/src/ImageIO.cxx:158-183): linear -> sRGB and
// quantisation to 8 or 16 bits.
//
// Both are per-pixel HBM-bound kernels.  render reads K, S, V, R0 (40 B) and
// writes R (12 B): 52 B/px, the same as the composite it contains; the four
// V neighbours come from L2.  encode reads 12 B and writes 3 or 6 B per
// pixel.
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"
#include "km.hpp"

namespace pnt {
namespace {

constexpr int kBlock = 256;
constexpr float kPiF = 3.14159265358979323846f;  // Math.hxx:21-22

// cv::borderInterpolate(p, len, BORDER_REFLECT) for p in [-1, len] (the
// default border of Interpolate, image/Mat.hxx:69-71); at integer positions
// Interpolate returns the bordered pixel itself.
__device__ __forceinline__ int reflect_index(int p, int len) {
  if (len == 1) return 0;
  if (p < 0) return -p - 1;
  if (p >= len) return 2 * len - p - 1;
  return p;
}

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Eigen's normalized(): a / sqrt(squaredNorm), unchanged when the norm is 0.
__device__ __forceinline__ F3 normalized(F3 a) {
  const float z = dot3(a, a);
  if (!(z > 0.0f)) return a;
  const float n = sqrtf(z);
  return F3{a.x / n, a.y / n, a.z / n};
}

struct RenderArgs {
  Layer7 L;
  Rgb3 R0, out;
  int rows, cols;
};

// One thread per pixel.  Renderer.hxx:97-152, evaluated in FP32.  The
// Beckmann exponent tan(acos(x))^2 is evaluated as (1 - x^2) / x^2 (the same
// value; FP32 acos loses the digits near 1 that this form keeps; x = 0 gives
// +inf and a zero term, as tan(acos(0))^2 ~ 2.7e32 does in FP64).
__global__ __launch_bounds__(kBlock) void k_render(RenderArgs a) {
  const int64_t n = (int64_t)a.rows * a.cols;
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p / a.cols), j = (int)(p % a.cols);
  float K[3], S[3], Kd[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    K[c] = a.L.ks[c][p];
    S[c] = a.L.ks[3 + c][p];
    Kd[c] = a.R0.p[c][p];
  }
  const double* V = a.L.v;  // thickness, evaluated in FP32 below
  const float s11 = (float)V[p];
  km_pixel(K, S, Kd, s11);  // compR = compose(canvas)
  const float s01 = (float)V[(int64_t)i * a.cols + reflect_index(j - 1, a.cols)];
  const float s21 = (float)V[(int64_t)i * a.cols + reflect_index(j + 1, a.cols)];
  const float s10 = (float)V[(int64_t)reflect_index(i - 1, a.rows) * a.cols + j];
  const float s12 = (float)V[(int64_t)reflect_index(i + 1, a.rows) * a.cols + j];
  // normal: normalized(va x vb) with z negated (Renderer.hxx:113-121)
  const F3 va = normalized(F3{2.0f, 0.0f, s21 - s01});
  const F3 vb = normalized(F3{0.0f, 2.0f, s12 - s10});
  F3 nrm = normalized(F3{va.y * vb.z - va.z * vb.y, va.z * vb.x - va.x * vb.z,
                         va.x * vb.y - va.y * vb.x});
  nrm.z = -nrm.z;
  // light at (-200, -1500, -2000), eye at (cols/2, rows/2, -100)
  const float px = (float)j, py = (float)i;
  const F3 ld = normalized(F3{-200.0f - px, -1500.0f - py, -2000.0f - s11});
  const F3 l = normalized(ld);
  const F3 v = normalized(F3{0.5f * (float)a.cols - px, 0.5f * (float)a.rows - py, -100.0f - s11});
  const F3 h = normalized(F3{v.x + l.x, v.y + l.y, v.z + l.z});
  const float NdotH = fmaxf(0.0f, dot3(nrm, h));
  const float VdotH = fmaxf(0.0f, dot3(v, h));
  const float NdotV = fmaxf(0.0f, dot3(nrm, v));
  const float NdotL = fmaxf(0.0f, dot3(nrm, l));
  constexpr float m = 0.5f, s = 0.2f, Ks = 1.0f, lightPower = 15.0f;
  float spec = 0.0f;
  if (NdotL > 0.0f && NdotV > 0.0f) {
    const float x2 = NdotH * NdotH;
    const float A = 1.0f / (m * m + x2 * x2 * kPiF);
    const float B = expf(-((1.0f - x2) / x2) / (m * m));
    const float G1 = 2.0f * NdotH * NdotV / VdotH;
    const float G2 = 2.0f * NdotH * NdotL / VdotH;
    const float g = G2 < G1 ? G2 : G1;  // std::min(G1, G2)
    const float G = g < 1.0f ? g : 1.0f;  // std::min(1.0, g)
    const float q = 1.0f - VdotH;
    const float RF = Ks + (1.0f - Ks) * (q * q * q * q * q);
    spec = (A * B * G * RF) / (NdotL * NdotV);
  }
  const float beta = lightPower * (1.0f / (4.0f * kPiF * dot3(ld, ld)));
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float r = (beta * NdotL) * ((1.0f - s) * Kd[c] + s * spec) + (Kd[c] * 0.2f) * Kd[c];
    a.out.p[c][p] = fminf(fmaxf(r, 0.0f), 1.0f);
  }
}

// ColorConverter::rgb2srgb (core/Color.hxx:197-205) and Mat::convertTo's
// saturate_cast (round half to even), in FP64 so that the integer result
// matches the FP64 reference on the same float input.
template <class T>
__global__ __launch_bounds__(kBlock) void k_encode(int64_t n, Rgb3 in, int to_srgb, double scale,
                                                   T* out) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double x = (double)in.p[c][p];
    if (to_srgb) x = x <= 0.00313066844250063 ? x * 12.92 : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
    double r = rint(x * scale);
    r = r < 0.0 ? 0.0 : (r > scale ? scale : r);
    out[3 * p + c] = (T)r;
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_render(int rows, int cols, const Layer7& L, const Rgb3& R0, const Rgb3& out,
                   hipStream_t s) {
  const int64_t n = (int64_t)rows * cols;
  if (n <= 0) return;
  k_render<<<blocks_for(n), kBlock, 0, s>>>(RenderArgs{L, R0, out, rows, cols});
  PNT_HIP_CHECK(hipGetLastError());
}

void launch_encode(int64_t n, const Rgb3& in, int to_srgb, int bits, void* out, hipStream_t s) {
  if (n <= 0) return;
  if (bits == 16)
    k_encode<uint16_t><<<blocks_for(n), kBlock, 0, s>>>(n, in, to_srgb, 65535.0,
                                                        static_cast<uint16_t*>(out));
  else
    k_encode<uint8_t><<<blocks_for(n), kBlock, 0, s>>>(n, in, to_srgb, 255.0,
                                                       static_cast<uint8_t*>(out));
  PNT_HIP_CHECK(hipGetLastError());
}

}  // namespace pnt
