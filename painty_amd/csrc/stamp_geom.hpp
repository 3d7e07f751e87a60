// Index math of one imprint, shared by host (stamp planning) and device.
//
// Restates renderer/FootprintBrush.hxx:90-113: for window offsets
// row, col in [-hr, hr] (rows outer, cols inner)
//   canvas pixel  (x, y) = ((int)(col + cx), (int)(row + cy))    truncation
//   map cell    (mx, my) = (round(col*c - row*s + hr), round(col*s + row*c + hr))
// with c = cos(-theta), s = sin(-theta) computed once on the host in FP64.
// The products and sums are evaluated in FP64 without contraction so that
// round() sees exactly the value the reference computes (x86-64 without FMA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pnt {

struct StampGeom {
  double cx, cy;  // brush centre (canvas coordinates)
  double c, s;    // cos(-theta), sin(-theta)
  int hr;         // window half size = (sizeMap-1)/2 (wr == hr)
  int size_map;   // pickup map / padded footprint side
  int rows, cols; // canvas
};

// Map cell index (my*size_map + mx) of window offset (col, row), or -1 when
// the reference skips it (outside the window, the canvas or the map); the
// canvas pixel index goes to *ic.
__host__ __device__ __forceinline__ int stamp_visit(const StampGeom& g, int col, int row,
                                                    int64_t* ic) {
#pragma clang fp contract(off)
  if (col < -g.hr || col > g.hr || row < -g.hr || row > g.hr) return -1;
  const int x = (int)(col + g.cx);
  const int y = (int)(row + g.cy);
  if (y < 0 || x < 0 || x >= g.cols || y >= g.rows) return -1;
  const double dc = (double)col, dr = (double)row;
  const double rc = dc * g.c - dr * g.s;
  const double rr = dc * g.s + dr * g.c;
  const int mx = (int)round(rc + (double)g.hr);
  const int my = (int)round(rr + (double)g.hr);
  if (my < 0 || mx < 0 || mx >= g.size_map || my >= g.size_map) return -1;
  *ic = (int64_t)y * g.cols + x;
  return my * g.size_map + mx;
}

// stamp_visit, and -1 also for a map cell outside the box [fx0, fx1] x
// [fy0, fy1] (the footprint's nonzero cells: such a visit has fh = 0 and
// changes nothing).
__host__ __device__ __forceinline__ int stamp_visit_in(const StampGeom& g, int col, int row,
                                                       int fx0, int fy0, int fx1, int fy1,
                                                       int64_t* ic) {
#pragma clang fp contract(off)
  if (col < -g.hr || col > g.hr || row < -g.hr || row > g.hr) return -1;
  const int x = (int)(col + g.cx);
  const int y = (int)(row + g.cy);
  if (y < 0 || x < 0 || x >= g.cols || y >= g.rows) return -1;
  const double dc = (double)col, dr = (double)row;
  const double rc = dc * g.c - dr * g.s;
  const double rr = dc * g.s + dr * g.c;
  const int mx = (int)round(rc + (double)g.hr);
  const int my = (int)round(rr + (double)g.hr);
  if (my < 0 || mx < 0 || mx >= g.size_map || my >= g.size_map) return -1;
  if (mx < fx0 || mx > fx1 || my < fy0 || my > fy1) return -1;
  *ic = (int64_t)y * g.cols + x;
  return my * g.size_map + mx;
}

// Canvas column 0 (row 0) receives two window columns (rows) when a stamp
// straddles the left (top) edge with a fractional centre: (int)(z) == 0 for
// z in (-1, 1).  Returns true when some col in [-hr, hr-1] has
// (int)(col + cx) == 0 == (int)(col + 1 + cx).  Both truncate to 0 only for
// -1 < col + cx < 0, so col lies within 1 of -cx: the three integers around
// floor(-cx) are the only candidates (|cx| < 1e9 keeps them exact).
__host__ __device__ __forceinline__ bool stamp_dup_axis(double c, int hr) {
#pragma clang fp contract(off)
  const double f = floor(-c);
  for (int d = -1; d <= 1; ++d) {
    const double kd = f + (double)d;
    if (kd < (double)-hr || kd > (double)(hr - 1)) continue;
    const int k = (int)kd;
    if ((int)(k + c) == 0 && (int)(k + 1 + c) == 0) return true;
  }
  return false;
}

}  // namespace pnt
