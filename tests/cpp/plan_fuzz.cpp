// Host-side planners of the stamp paths, driven on the CPU with random stamp
// lists and checked against their invariants -- built, with the library's
// host code, under AddressSanitizer and UndefinedBehaviorSanitizer by
// sanitize_planners.sh.  No device is used.
//
//   plan_stamps_rc      geometry, ring boxes, the shared-pixel visit levels of
//                       stamps that straddle canvas column/row 0
//   plan_units          stamps cut into units and assigned to chains
//   plan_stamp_needs    per stamp, the leading stamps of earlier units it waits on
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <utility>
#include <vector>

#include "handles.hpp"
#include "internal.hpp"
#include "persistent.hpp"
#include "stampcore.hpp"

using namespace pnt;

namespace {

int g_fail = 0;
#define CHECK(c, ...)                                    \
  do {                                                   \
    if (!(c)) {                                          \
      if (g_fail++ < 20) {                               \
        std::fprintf(stderr, "plan_fuzz: %s: ", #c);     \
        std::fprintf(stderr, __VA_ARGS__);               \
        std::fprintf(stderr, "\n");                      \
      }                                                  \
    }                                                    \
  } while (0)

struct Case {
  int rows, cols, size_map;
  double radius;
  std::vector<pnt_stamp> st;
};

// A visit on a duplicated canvas column / row 0 (the device's on_dup_line).
bool dup_line(const PStamp& p, int col, int row) {
  return ((p.flags & kStDupX) && (int)(col + p.cx) == 0) ||
         ((p.flags & kStDupY) && (int)(row + p.cy) == 0);
}

Case make_case(std::mt19937_64& rng, int it) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  Case c;
  c.rows = 1 + (int)(U(rng) * 200);
  c.cols = 1 + (int)(U(rng) * 200);
  c.size_map = 1 + 2 * (int)(U(rng) * 24);  // odd, 1..49
  const int hr = (c.size_map - 1) / 2;
  c.radius = hr + U(rng) * 3.0;
  const int n = 1 + (int)(U(rng) * 60);
  double x = -hr - 2 + U(rng) * (c.cols + 2 * hr + 4);
  double y = -hr - 2 + U(rng) * (c.rows + 2 * hr + 4);
  for (int k = 0; k < n; ++k) {
    pnt_stamp s{};
    if (it % 3 == 0) {  // strokes along the top / left edge: fractional centres near 0
      s.cx = (it % 2 ? -0.99 + 1.98 * U(rng) : x);
      s.cy = (it % 2 ? y : -0.99 + 1.98 * U(rng));
    } else {
      s.cx = x;
      s.cy = y;
    }
    s.theta = U(rng) * 6.283185307179586;
    if (it % 5 == 0) s.theta = (int)(U(rng) * 4) * 1.5707963267948966;
    for (int ch = 0; ch < 3; ++ch) {
      s.K[ch] = (float)U(rng);
      s.S[ch] = (float)U(rng);
    }
    s.flags = (k == 0 || U(rng) < 0.1) ? PNT_STAMP_DIP : 0u;
    if (U(rng) < 0.05) s.flags |= PNT_STAMP_DRY_AFTER;
    c.st.push_back(s);
    x += -2.0 + 4.0 * U(rng);
    y += -2.0 + 4.0 * U(rng);
    if (U(rng) < 0.1) {  // a jump: a new stroke elsewhere
      x = -hr - 2 + U(rng) * (c.cols + 2 * hr + 4);
      y = -hr - 2 + U(rng) * (c.rows + 2 * hr + 4);
    }
  }
  return c;
}

void check_stamps(const Case& c, const std::vector<PStamp>& h, const std::vector<int2>& dl,
                  const std::vector<int>& off, int it) {
  const int hr = (c.size_map - 1) / 2;
  CHECK(h.size() == c.st.size(), "case %d: %zu plans for %zu stamps", it, h.size(), c.st.size());
  for (size_t k = 0; k < h.size(); ++k) {
    const PStamp& p = h[k];
    const RingBox& b = p.box;
    CHECK(b.a_x0 >= 0 && b.a_y0 >= 0 && b.a_x1 <= c.cols - 1 && b.a_y1 <= c.rows - 1,
          "case %d stamp %zu: ring box outside the canvas", it, k);
    const bool dup = (p.flags & (kStDupX | kStDupY)) != 0;
    if (!dup) {
      CHECK(p.dl_nlev == 0, "case %d stamp %zu: levels without a duplicated line", it, k);
      continue;
    }
    if (p.dl_nlev == 0) continue;
    CHECK(p.dl_lev >= 0 && p.dl_lev + p.dl_nlev < (int)off.size(),
          "case %d stamp %zu: level table out of range", it, k);
    if (!(p.dl_lev >= 0 && p.dl_lev + p.dl_nlev < (int)off.size())) continue;
    StampGeom g{p.cx, p.cy, p.c, p.s, hr, c.size_map, c.rows, c.cols};
    std::set<std::pair<int, int>> listed;
    for (int l = 0; l < p.dl_nlev; ++l) {
      const int a = off[(size_t)(p.dl_lev + l)], e = off[(size_t)(p.dl_lev + l + 1)];
      CHECK(0 <= a && a <= e && e <= (int)dl.size(), "case %d stamp %zu: level %d [%d, %d)", it,
            k, l, a, e);
      if (!(0 <= a && a <= e && e <= (int)dl.size())) break;
      for (int q = a; q < e; ++q) {
        const int2 v = dl[(size_t)q];
        int64_t ic = 0;
        CHECK(stamp_visit(g, v.x, v.y, &ic) >= 0, "case %d stamp %zu: listed (%d, %d) is no visit",
              it, k, v.x, v.y);
        CHECK(listed.insert({v.x, v.y}).second, "case %d stamp %zu: (%d, %d) listed twice", it, k,
              v.x, v.y);
      }
    }
    // every visit on a duplicated canvas column / row 0 is listed
    for (int row = -hr; row <= hr; ++row)
      for (int col = -hr; col <= hr; ++col) {
        int64_t ic = 0;
        if (stamp_visit(g, col, row, &ic) < 0 || !dup_line(p, col, row)) continue;
        CHECK(listed.count({col, row}) == 1, "case %d stamp %zu: shared visit (%d, %d) unlisted",
              it, k, col, row);
      }
  }
}

void check_units(const Case& c, const std::vector<PStamp>& h, int nchains, int it) {
  const int64_t n = (int64_t)c.st.size();
  std::vector<uint8_t> clear((size_t)n);
  for (int64_t k = 0; k < n; ++k) clear[(size_t)k] = (c.st[(size_t)k].flags & PNT_STAMP_DIP) != 0;
  const std::vector<Unit> U = plan_units(c.st.data(), h.data(), clear.data(), n, nchains);
  int64_t at = 0;
  std::vector<int> unit_of((size_t)n, -1);
  for (size_t i = 0; i < U.size(); ++i) {
    const Unit& u = U[i];
    CHECK(u.s0 == at && u.s1 > u.s0 && u.s1 <= n, "case %d unit %zu: [%lld, %lld) after %lld", it,
          i, (long long)u.s0, (long long)u.s1, (long long)at);
    CHECK(u.chain >= 0 && u.chain < nchains, "case %d unit %zu: chain %d of %d", it, i, u.chain,
          nchains);
    for (int d : u.deps) CHECK(d >= 0 && d < (int)i, "case %d unit %zu: dependency %d", it, i, d);
    for (int64_t k = u.s0; k < u.s1 && k < n && k >= 0; ++k) unit_of[(size_t)k] = (int)i;
    at = u.s1;
  }
  CHECK(at == n, "case %d: units cover %lld of %lld stamps", it, (long long)at, (long long)n);
  if (at != n) return;
  const StampNeeds N = plan_stamp_needs(U, h.data());
  CHECK((int64_t)N.off.size() == n + 1 && N.off[0] == 0 && N.off.back() == (int)N.need.size(),
        "case %d: need table of %zu offsets", it, N.off.size());
  if (!((int64_t)N.off.size() == n + 1)) return;
  for (int64_t k = 0; k < n; ++k) {
    CHECK(N.off[(size_t)k] <= N.off[(size_t)k + 1], "case %d stamp %lld: offsets fall", it,
          (long long)k);
    for (int e = N.off[(size_t)k]; e < N.off[(size_t)k + 1] && e < (int)N.need.size(); ++e) {
      const int2 nd = N.need[(size_t)e];
      const int u = unit_of[(size_t)k];
      CHECK(nd.x >= 0 && nd.x < u, "case %d stamp %lld (unit %d): waits on unit %d", it,
            (long long)k, u, nd.x);
      if (nd.x >= 0 && nd.x < (int)U.size())
        CHECK(nd.y >= 1 && nd.y <= U[(size_t)nd.x].s1 - U[(size_t)nd.x].s0,
              "case %d stamp %lld: waits on %d stamps of unit %d", it, (long long)k, nd.y, nd.x);
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  const int cases = argc > 1 ? std::atoi(argv[1]) : 3000;
  std::mt19937_64 rng(20261203);
  int dup_cases = 0;
  for (int it = 0; it < cases; ++it) {
    const Case c = make_case(rng, it);
    std::vector<PStamp> h;
    std::vector<int2> dl;
    std::vector<int> off;
    plan_stamps_rc(c.size_map, c.radius, c.rows, c.cols, c.st.data(), (int64_t)c.st.size(), h, dl,
                   off);
    check_stamps(c, h, dl, off, it);
    for (const PStamp& p : h) dup_cases += p.dl_nlev > 0;
    for (int nchains : {1, 3, 8}) check_units(c, h, nchains, it);
  }
  std::printf("plan_fuzz: %d cases, %d stamps with shared-pixel levels, %d failures\n", cases,
              dup_cases, g_fail);
  return g_fail ? 1 : 0;
}
