// Footprint ingest (FootprintBrush.hxx:48-58): the reference reads
// ./data/footprint/footprint.png with io::imRead(path, Mat<double>&, true)
// (io/src/ImageIO.cxx:78-113; the in-memory form :115-150), then resizes it
// with ScaledMat (image/Mat.hxx:141-148), i.e. cv::resize with
// INTER_LANCZOS4, before padding it to the pickup map's size.
//
//  - PNG decoding is host work: the file is parsed and inflated (zlib), rows
//    unfiltered, and grey samples (8 or 16 bit, alpha dropped as
//    IMREAD_GRAYSCALE does) scaled by 1/maxval in FP64 and, on request,
//    converted from sRGB to linear (core/Color.hxx:189-195).
//  - The Lanczos-4 resize runs on the device in FP64: a horizontal pass over
//    every source row, then a vertical pass, 8 taps each, with OpenCV's
//    float coefficients and replicated borders, summed in tap order (FP
//    contraction off, so the sums round as OpenCV's do).
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace pnt {
int guard_call(const std::function<void()>& f);  // api.hip: status + pnt_last_error

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------ PNG --
uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

int paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  if (pb <= pc) return b;
  return c;
}

struct Gray {
  int rows = 0, cols = 0, maxval = 255;
  std::vector<uint16_t> v;  // grey samples, row-major
};

// A PNG's grey samples.  Supported: colour types 0 (grey) and 4 (grey +
// alpha) at 8 or 16 bits, not interlaced -- the reference's footprint is
// 8-bit grey + alpha.  Anything else fails with a message rather than being
// converted differently from OpenCV.
Gray decode_png(const uint8_t* d, int64_t len) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (len < 8 || std::memcmp(d, sig, 8) != 0) throw Error(PNT_EINVAL, "not a PNG file");
  int64_t p = 8;
  bool have_hdr = false, done = false;
  int width = 0, height = 0, depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> idat;
  while (!done) {
    if (p + 12 > len) throw Error(PNT_EINVAL, "PNG: truncated chunk");
    const uint32_t n = be32(d + p);
    if ((int64_t)n > len - p - 12) throw Error(PNT_EINVAL, "PNG: truncated chunk");
    const uint8_t* type = d + p + 4;
    const uint8_t* data = d + p + 8;
    const uint32_t crc = be32(data + n);
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, n + 4) != crc)
      throw Error(PNT_EINVAL, "PNG: chunk CRC mismatch");
    if (std::memcmp(type, "IHDR", 4) == 0) {
      if (n != 13) throw Error(PNT_EINVAL, "PNG: bad IHDR");
      width = (int)be32(data);
      height = (int)be32(data + 4);
      depth = data[8];
      ctype = data[9];
      interlace = data[12];
      if (data[10] != 0 || data[11] != 0) throw Error(PNT_EINVAL, "PNG: unknown method");
      have_hdr = true;
    } else if (std::memcmp(type, "IDAT", 4) == 0) {
      idat.insert(idat.end(), data, data + n);
    } else if (std::memcmp(type, "IEND", 4) == 0) {
      done = true;
    } else if (!(type[0] & 0x20)) {  // an unknown critical chunk (PLTE: palette images)
      throw Error(PNT_EINVAL, "PNG: colour or palette image (grey or grey+alpha expected)");
    }
    p += 12 + (int64_t)n;
  }
  if (!have_hdr) throw Error(PNT_EINVAL, "PNG: no IHDR");
  if (ctype != 0 && ctype != 4)
    throw Error(PNT_EINVAL, "PNG: colour image (grey or grey+alpha expected)");
  if (depth != 8 && depth != 16) throw Error(PNT_EINVAL, "PNG: 8- or 16-bit samples expected");
  if (interlace != 0) throw Error(PNT_EINVAL, "PNG: interlaced images are not supported");
  if (width <= 0 || height <= 0 || (int64_t)width * height > ((int64_t)1 << 28))
    throw Error(PNT_EINVAL, "PNG: image size out of range");
  const int chans = ctype == 4 ? 2 : 1, bps = depth / 8, bpp = chans * bps;
  const int64_t stride = (int64_t)width * bpp;
  std::vector<uint8_t> raw((size_t)((stride + 1) * height));
  z_stream zs{};
  if (inflateInit(&zs) != Z_OK) throw Error(PNT_EINVAL, "PNG: zlib init failed");
  zs.next_in = idat.data();
  zs.avail_in = (uInt)idat.size();
  zs.next_out = raw.data();
  zs.avail_out = (uInt)raw.size();
  const int zr = inflate(&zs, Z_FINISH);
  const uLong got = zs.total_out;
  inflateEnd(&zs);
  if (zr != Z_STREAM_END || got != raw.size()) throw Error(PNT_EINVAL, "PNG: bad image data");
  // undo the row filters (PNG spec 9.2) in place
  std::vector<uint8_t> prev((size_t)stride, 0);
  Gray g;
  g.rows = height;
  g.cols = width;
  g.maxval = depth == 16 ? 65535 : 255;
  g.v.resize((size_t)width * height);
  for (int y = 0; y < height; ++y) {
    uint8_t* row = raw.data() + (size_t)y * (stride + 1);
    const int f = row[0];
    uint8_t* x = row + 1;
    for (int64_t i = 0; i < stride; ++i) {
      const int a = i >= bpp ? x[i - bpp] : 0, b = prev[(size_t)i],
                c = i >= bpp ? prev[(size_t)(i - bpp)] : 0;
      int r;
      switch (f) {
        case 0: r = x[i]; break;
        case 1: r = x[i] + a; break;
        case 2: r = x[i] + b; break;
        case 3: r = x[i] + ((a + b) >> 1); break;
        case 4: r = x[i] + paeth(a, b, c); break;
        default: throw Error(PNT_EINVAL, "PNG: unknown row filter");
      }
      x[i] = (uint8_t)r;
    }
    std::memcpy(prev.data(), x, (size_t)stride);
    for (int i = 0; i < width; ++i) {
      const uint8_t* s = x + (size_t)i * bpp;  // the grey sample comes first
      g.v[(size_t)y * width + i] = bps == 2 ? (uint16_t)((s[0] << 8) | s[1]) : s[0];
    }
  }
  return g;
}

// ColorConverter::srgb2rgb (core/Color.hxx:189-195)
double srgb2rgb(double s) {
  return s <= 0.0404482362771082 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4);
}

// ------------------------------------------------------- Lanczos resize --
// OpenCV's interpolateLanczos4: the 8 weights of the Lanczos-4 kernel at
// t = x + 3 - i through one sin/cos pair and the angle-sum table, rounded
// to float, normalised by their float sum; x = 0 gives the unit tap.
void lanczos4_coeffs(float x, float c[8]) {
  static const double s45 = 0.70710678118654752440084436210485;
  static const double cs[8][2] = {{1, 0},  {-s45, -s45}, {0, 1},  {s45, -s45},
                                  {-1, 0}, {s45, s45},   {0, -1}, {-s45, s45}};
  if (x < FLT_EPSILON) {
    for (int i = 0; i < 8; ++i) c[i] = 0.0f;
    c[3] = 1.0f;
    return;
  }
  const double pi = 3.14159265358979323846;
  const float xp3 = x + 3.0f;
  const double y0 = -(double)xp3 * pi * 0.25, s0 = std::sin(y0), c0 = std::cos(y0);
  float sum = 0.0f;
  for (int i = 0; i < 8; ++i) {
    const float xi = x + 3.0f - (float)i;
    const double y = -(double)xi * pi * 0.25;
    c[i] = (float)((cs[i][0] * s0 + cs[i][1] * c0) / (y * y));
    sum += c[i];
  }
  sum = 1.0f / sum;
  for (int i = 0; i < 8; ++i) c[i] *= sum;
}

// One axis: the first tap's source index and the 8 weights per output.
void lanczos_axis(int ssize, int dsize, std::vector<int>& ofs, std::vector<double>& w) {
  const double scale = 1.0 / ((double)dsize / ssize);
  ofs.resize((size_t)dsize);
  w.resize(8 * (size_t)dsize);
  for (int d = 0; d < dsize; ++d) {
    float f = (float)((d + 0.5) * scale - 0.5);
    const int s = (int)std::floor(f);
    f -= (float)s;
    ofs[(size_t)d] = s - 3;
    float c[8];
    lanczos4_coeffs(f, c);
    for (int j = 0; j < 8; ++j) w[8 * (size_t)d + j] = (double)c[j];
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// tmp[r][dx] = sum_j src[r][clamp(xofs[dx] + j)] * xw[dx][j]
__global__ __launch_bounds__(kBlock) void k_resize_h(const double* src, int rows, int cols,
                                                     int ocols, const int* xofs, const double* xw,
                                                     double* tmp) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (int64_t)rows * ocols) return;
  const int r = (int)(e / ocols), dx = (int)(e % ocols);
  const double* S = src + (int64_t)r * cols;
  const int s0 = xofs[dx];
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) v += S[clampi(s0 + j, 0, cols - 1)] * xw[8 * dx + j];
  tmp[e] = v;
}

// dst[dy][dx] = sum_k tmp[clamp(yofs[dy] + k)][dx] * yw[dy][k]
__global__ __launch_bounds__(kBlock) void k_resize_v(const double* tmp, int rows, int ocols,
                                                     int orows, const int* yofs, const double* yw,
                                                     double* dst) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (int64_t)orows * ocols) return;
  const int dy = (int)(e / ocols), dx = (int)(e % ocols);
  const int s0 = yofs[dy];
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    v += tmp[(int64_t)clampi(s0 + k, 0, rows - 1) * ocols + dx] * yw[8 * dy + k];
  dst[e] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace
}  // namespace pnt

using namespace pnt;

extern "C" {

int pnt_image_read_gray(const uint8_t* data, int64_t len, int srgb, int* rows, int* cols,
                        double* out, int64_t capacity) {
  return guard_call([&] {
    PNT_REQUIRE(data && len > 0 && rows && cols, "invalid argument");
    const Gray g = decode_png(data, len);
    *rows = g.rows;
    *cols = g.cols;
    if (!out) return;
    PNT_REQUIRE(capacity >= (int64_t)g.rows * g.cols, "output capacity below rows * cols");
    const double scale = g.maxval == 65535 ? 1.0 / 0xffff : 1.0 / 0xff;  // Mat::convertTo
    for (size_t i = 0; i < g.v.size(); ++i) {
      const double x = (double)g.v[i] * scale;
      out[i] = srgb ? srgb2rgb(x) : x;
    }
  });
}

int pnt_resize_lanczos4(pnt_ctx ctx, const double* src, int rows, int cols, int out_rows,
                        int out_cols, double* out) {
  return guard_call([&] {
    PNT_REQUIRE(ctx && src && out, "NULL argument");
    PNT_REQUIRE(rows > 0 && cols > 0 && out_rows > 0 && out_cols > 0 &&
                    (int64_t)rows * cols <= ((int64_t)1 << 28) &&
                    (int64_t)out_rows * out_cols <= ((int64_t)1 << 28) &&
                    (int64_t)rows * out_cols <= ((int64_t)1 << 28),
                "image sizes out of range");
    ctx->use();
    hipStream_t s = ctx->stream;
    std::vector<int> xofs, yofs;
    std::vector<double> xw, yw;
    lanczos_axis(cols, out_cols, xofs, xw);
    lanczos_axis(rows, out_rows, yofs, yw);
    const int64_t n_src = (int64_t)rows * cols, n_tmp = (int64_t)rows * out_cols,
                  n_dst = (int64_t)out_rows * out_cols;
    // one device block: src, tmp, dst, weights (doubles), then offsets (ints)
    const int64_t nd = n_src + n_tmp + n_dst + 8 * (int64_t)(out_cols + out_rows);
    const int64_t ni = out_cols + out_rows;
    double* d = reinterpret_cast<double*>(dev_alloc(2 * nd + ni + 2));
    try {
      double* d_src = d;
      double* d_tmp = d_src + n_src;
      double* d_dst = d_tmp + n_tmp;
      double* d_xw = d_dst + n_dst;
      double* d_yw = d_xw + 8 * (int64_t)out_cols;
      int* d_xofs = reinterpret_cast<int*>(d_yw + 8 * (int64_t)out_rows);
      int* d_yofs = d_xofs + out_cols;
      PNT_HIP_CHECK(hipMemcpyAsync(d_src, src, n_src * sizeof(double), hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(d_xw, xw.data(), xw.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(d_yw, yw.data(), yw.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(d_xofs, xofs.data(), xofs.size() * sizeof(int),
                                   hipMemcpyHostToDevice, s));
      PNT_HIP_CHECK(hipMemcpyAsync(d_yofs, yofs.data(), yofs.size() * sizeof(int),
                                   hipMemcpyHostToDevice, s));
      k_resize_h<<<blocks_for(n_tmp), kBlock, 0, s>>>(d_src, rows, cols, out_cols, d_xofs, d_xw,
                                                      d_tmp);
      PNT_HIP_CHECK(hipGetLastError());
      k_resize_v<<<blocks_for(n_dst), kBlock, 0, s>>>(d_tmp, rows, out_cols, out_rows, d_yofs,
                                                      d_yw, d_dst);
      PNT_HIP_CHECK(hipGetLastError());
      PNT_HIP_CHECK(hipMemcpyAsync(out, d_dst, n_dst * sizeof(double), hipMemcpyDeviceToHost, s));
      PNT_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dev_free(d);
      throw;
    }
    dev_free(d);
  });
}

}  // extern "C"
