// Ingest of paint data: ComputeScatteringAndAbsorption (the Curtis et al.
// inversion of core/KubelkaMunk.hxx:84-124, turning a colour over black and
// over white into K and S) as a device kernel, and the palette JSON format
// of mixer/src/Serialization.cxx:26-53 ([{"K": [3], "S": [3]}, ...]).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"

namespace pnt {
namespace {

constexpr int kBlock = 256;

// acoth (core/Math.hxx:182-192): +inf when fuzzyCompare(x, 1, 100 eps)
// (|x - 1| < 100 eps, Math.hxx:33-37).
__device__ __forceinline__ double acoth_d(double x) {
  if (fabs(x - 1.0) < 100.0 * DBL_EPSILON) return INFINITY;
  return log((x + 1.0) / (x - 1.0)) / 2.0;
}

// One channel of one colour per thread, in FP64 as the reference computes
// it; an input outside 0 < Rb < Rw < 1 raises *bad (the reference throws).
__global__ __launch_bounds__(kBlock) void k_scatter_absorb(int64_t n3, const double* Rb,
                                                           const double* Rw, double* K, double* S,
                                                           unsigned* bad) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n3) return;
  const double b0 = Rb[i], w0 = Rw[i];
  if (!(b0 < w0 && b0 > 0 && b0 < 1.0 && w0 > 0 && w0 < 1.0)) {
    *bad = 1u;
    K[i] = S[i] = 0.0;
    return;
  }
  const double a = 0.5 * (w0 + (b0 - w0 + 1.) / b0);
  const double b = sqrt(a * a - 1.);
  const double arg = (b * b - (a - w0) * (a - 1.)) / (b * (1. - w0));
  const double s = (1. / b) * acoth_d(arg);
  S[i] = s;
  K[i] = s * (a - 1.);
}

}  // namespace

void launch_scatter_absorb(int64_t n3, const double* Rb, const double* Rw, double* K, double* S,
                           unsigned* bad, hipStream_t s) {
  if (n3 <= 0) return;
  k_scatter_absorb<<<(unsigned)((n3 + kBlock - 1) / kBlock), kBlock, 0, s>>>(n3, Rb, Rw, K, S,
                                                                             bad);
  PNT_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ palette JSON
namespace {

struct JsonIn {
  const char* p;
  const char* end;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool eat(char c) {
    ws();
    if (p < end && *p == c) {
      ++p;
      return true;
    }
    return false;
  }
  void need(char c, const char* what) {
    if (!eat(c)) throw Error(PNT_EINVAL, std::string("palette JSON: expected ") + what);
  }
  std::string str() {
    need('"', "a string");
    std::string s;
    while (p < end && *p != '"') {
      if (*p == '\\') {  // keys of this format need no escapes beyond the simple ones
        ++p;
        if (p >= end) break;
      }
      s.push_back(*p++);
    }
    need('"', "the end of a string");
    return s;
  }
  double num() {
    ws();
    char* e = nullptr;
    errno = 0;
    const std::string tok(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
    const double v = std::strtod(tok.c_str(), &e);
    if (e == tok.c_str()) throw Error(PNT_EINVAL, "palette JSON: expected a number");
    p += e - tok.c_str();
    return v;
  }
  void skip_value() {  // any JSON value (keys other than K and S)
    ws();
    if (p >= end) throw Error(PNT_EINVAL, "palette JSON: unexpected end");
    if (*p == '"') {
      str();
    } else if (*p == '[' || *p == '{') {
      const char open = *p, close = open == '[' ? ']' : '}';
      ++p;
      if (eat(close)) return;
      do {
        if (open == '{') {
          str();
          need(':', "':'");
        }
        skip_value();
      } while (eat(','));
      need(close, open == '[' ? "']'" : "'}'");
    } else if (std::strncmp(p, "true", 4) == 0 || std::strncmp(p, "null", 4) == 0) {
      p += 4;
    } else if (std::strncmp(p, "false", 5) == 0) {
      p += 5;
    } else {
      num();
    }
  }
  void triple(double out[3]) {
    need('[', "'[' of a 3-vector");
    for (int c = 0; c < 3; ++c) {
      if (c) need(',', "',' in a 3-vector");
      out[c] = num();
    }
    need(']', "']' after 3 values");
  }
};

// The shortest decimal that reads back as v (Python's repr convention:
// positional for 1e-4 <= |v| < 1e16, else d.ddde+XX / e-XX).
void put_double(std::string& out, double v) {
  if (!std::isfinite(v)) throw Error(PNT_EINVAL, "palette JSON: non-finite value");
  char buf[40];
  int prec = 1;
  for (; prec <= 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  // buf = [-]d.ddddde[+-]XX with `prec` significant digits
  const char* q = buf;
  std::string sign;
  if (*q == '-') {
    sign = "-";
    ++q;
  }
  std::string digits;
  for (; *q && *q != 'e'; ++q)
    if (*q != '.') digits.push_back(*q);
  const int exp10 = std::atoi(q + 1);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string s;
  if (exp10 >= -4 && exp10 < 16) {
    if (exp10 >= 0) {
      std::string ip = digits.substr(0, std::min<size_t>(digits.size(), (size_t)exp10 + 1));
      while ((int)ip.size() < exp10 + 1) ip.push_back('0');
      std::string fp = digits.size() > (size_t)exp10 + 1 ? digits.substr((size_t)exp10 + 1) : "";
      s = ip + "." + (fp.empty() ? "0" : fp);
    } else {
      s = "0." + std::string((size_t)(-exp10 - 1), '0') + digits;
    }
  } else {
    s = digits.substr(0, 1);
    if (digits.size() > 1) s += "." + digits.substr(1);
    char e[8];
    std::snprintf(e, sizeof e, "e%c%02d", exp10 < 0 ? '-' : '+', std::abs(exp10));
    s += e;
  }
  out += sign + s;
}

}  // namespace

void palette_from_json(const char* text, int64_t len, std::vector<double>& K,
                       std::vector<double>& S) {
  JsonIn in{text, text + len};
  in.need('[', "'[' (a palette is a JSON array)");
  K.clear();
  S.clear();
  if (!in.eat(']')) {
    do {
      in.need('{', "'{' (a paint coefficient object)");
      double k[3], s[3];
      bool hk = false, hs = false;
      if (!in.eat('}')) {
        do {
          const std::string key = in.str();
          in.need(':', "':'");
          if (key == "K") {
            in.triple(k);
            hk = true;
          } else if (key == "S") {
            in.triple(s);
            hs = true;
          } else {
            in.skip_value();
          }
        } while (in.eat(','));
        in.need('}', "'}'");
      }
      if (!hk || !hs) throw Error(PNT_EINVAL, "palette JSON: an entry lacks \"K\" or \"S\"");
      K.insert(K.end(), k, k + 3);
      S.insert(S.end(), s, s + 3);
    } while (in.eat(','));
    in.need(']', "']' at the end of the palette");
  }
  in.ws();
  if (in.p != in.end) throw Error(PNT_EINVAL, "palette JSON: trailing characters");
}

std::string palette_to_json(const double* K, const double* S, int64_t n) {
  std::string out = "[";
  for (int64_t l = 0; l < n; ++l) {
    if (l) out += ",";
    out += "{\"K\":[";
    for (int c = 0; c < 3; ++c) {
      if (c) out += ",";
      put_double(out, K[3 * l + c]);
    }
    out += "],\"S\":[";
    for (int c = 0; c < 3; ++c) {
      if (c) out += ",";
      put_double(out, S[3 * l + c]);
    }
    out += "]}";
  }
  return out + "]";
}

}  // namespace pnt
