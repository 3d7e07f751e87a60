// LoadPalette / SavePalette -- drop-in for mixer/Serialization.hxx:22-25
// (mixer/src/Serialization.cxx:26-53) over the library's palette JSON.
// VisualizePalette (a display helper of the GUI) is not provided.
#pragma once

#include <iostream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "painty/gpu/Device.hxx"
#include "painty/mixer/Palette.hxx"

namespace painty {

// Appends the stream's palette entries, as the reference does.
inline void LoadPalette(std::istream& stream, Palette& palette) {
  const std::string text((std::istreambuf_iterator<char>(stream)), std::istreambuf_iterator<char>());
  int64_t n = 0;
  const int64_t cap = 1 + (int64_t)text.size() / 16;  // every entry takes more than 16 chars
  std::vector<double> K(3 * cap), S(3 * cap);
  gpu::check(pnt_palette_from_json(text.data(), (int64_t)text.size(), K.data(), S.data(), cap, &n));
  for (int64_t l = 0; l < n; ++l) {
    PaintCoeff p;
    for (int c = 0; c < 3; ++c) {
      p.K[c] = K[3 * l + c];
      p.S[c] = S[3 * l + c];
    }
    palette.push_back(p);
  }
}

inline void SavePalette(std::ostream& stream, const Palette& palette) {
  std::vector<double> K, S;
  for (const PaintCoeff& p : palette)
    for (int c = 0; c < 3; ++c) {
      K.push_back(p.K[c]);
      S.push_back(p.S[c]);
    }
  std::vector<char> buf(64 + 320 * palette.size());
  int64_t len = 0;
  gpu::check(pnt_palette_to_json(K.data(), S.data(), (int64_t)palette.size(), buf.data(),
                                 (int64_t)buf.size(), &len));
  stream << std::string(buf.data(), (size_t)len);
}

}  // namespace painty
