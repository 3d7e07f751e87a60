/*
 * painty_oracle.c -- FP64 CPU restatement of lindemeier/painty's paint
 * rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP library
 * and the timed CPU baseline; the product never links or calls it.
 *
 * Built with -ffp-contract=off so that a*b+c is never fused: the reference is
 * compiled for generic x86-64 (no FMA), so every product and sum rounds on its
 * own, and the rounding of the stamp index math (FootprintBrush.hxx:95-100)
 * depends on that.
 *
 * Parity pins (tests/test_oracle.py):
 *   coth / ComputeReflectance   reference goldens KubelkaMunkTest.cxx:14-31 and
 *                               MathTest.cxx:83-91; closed forms (d->inf gives
 *                               R_inf = a-b, d=0 gives R0, R0=R_inf fixed
 *                               point, Beer-Lambert at S=0, the K=0 limit,
 *                               additivity of two coats); 50-digit mpmath
 *                               evaluation of Kubelka's solution.
 *   compose / composeOnto / dry closed forms against ComputeReflectance and
 *                               the dry invariants (h += V, layer zeroed).
 *   imprint / pickup / deposit  brute force (an independent pure-python
 *                               restatement on tiny canvases), the pickup-off
 *                               closed form (volume-weighted mean), bijective
 *                               map at theta=0, volume conservation of pickup,
 *                               thin-paint cases around the MinVolume threshold.
 *   mixSinglePaint / mixed      closed form (matrix product) on the reference's
 *                               palette data files.
 *   CatmullRom                  reference golden SplineTest.cxx:26.
 */
#include "painty_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* painty/core/Math.hxx:157-169 */
double orc_coth(double x) {
  if (x > 20.0) return 1.0;
  if (fabs(x) > 0.0) {
    const double res = cosh(x) / sinh(x);
    return isnan(res) ? 1.0 : res;
  }
  return INFINITY;
}

/* painty/core/Math.hxx:181-191 (fuzzyCompare, Math.hxx:34-38) */
double orc_acoth(double x) {
  if (fabs(x - 1.0) < 100.0 * DBL_EPSILON) return INFINITY;
  return log((x + 1.0) / (x - 1.0)) / 2.0;
}

/* painty/core/KubelkaMunk.hxx:25-83 */
void orc_compute_reflectance(const double K[3], const double S_in[3],
                             const double R0[3], double d, double R[3]) {
  const double thr = DBL_EPSILON * 10000.0;
  if (fabs(d) < thr) {
    R[0] = R0[0];
    R[1] = R0[1];
    R[2] = R0[2];
    return;
  }
  double out[3];
  for (int i = 0; i < 3; i++) {
    const double S = (fabs(S_in[i]) > thr) ? S_in[i] : 0.00000000001;
    const double K_S = K[i] / S;
    const double a = 1.0 + K_S;
    const double asq = a * a;
    const double v = asq - 1.0;
    const double b = (v < 0.0) ? 0.0 : sqrt(v);
    const double bSh = b * S * d;
    const double bcothbSh = b * orc_coth(bSh);
    out[i] = (1.0 - R0[i] * (a - bcothbSh)) / (a - R0[i] + bcothbSh);
  }
  R[0] = out[0];
  R[1] = out[1];
  R[2] = out[2];
}

/* painty/core/KubelkaMunk.hxx:89-124 */
int orc_compute_scattering_absorption(const double Rb[3], const double Rw[3],
                                      double K[3], double S[3]) {
  for (int i = 0; i < 3; i++) {
    if (!(Rb[i] < Rw[i] && Rb[i] > 0 && Rb[i] < 1.0 && Rw[i] > 0 &&
          Rw[i] < 1.0))
      return -1;
  }
  double a[3], b[3], arg[3];
  for (int i = 0; i < 3; i++) a[i] = 0.5 * (Rw[i] + (Rb[i] - Rw[i] + 1.) / Rb[i]);
  for (int i = 0; i < 3; i++) b[i] = sqrt(a[i] * a[i] - 1.);
  for (int i = 0; i < 3; i++)
    arg[i] = (b[i] * b[i] - (a[i] - Rw[i]) * (a[i] - 1.)) / (b[i] * (1. - Rw[i]));
  for (int i = 0; i < 3; i++) S[i] = (1. / b[i]) * orc_acoth(arg[i]);
  for (int i = 0; i < 3; i++) K[i] = S[i] * (a[i] - 1.);
  return 0;
}

/* painty/renderer/shaders/DryingStep.compute.glsl:31-49 (and the identical
 * computeReflectance of ComposeLayerOnSubstrate.compute.glsl:29-47) */
void orc_compute_reflectance_glsl(const double K[3], const double S[3],
                                  const double R0[3], double d, double R[3]) {
  double a[3], b[3], bSh[3];
  int any_small = 0;
  for (int i = 0; i < 3; i++) {
    a[i] = 1.0 + K[i] / S[i];
    const double v = a[i] * a[i] - 1.0;
    b[i] = sqrt(v > 0.0 ? v : 0.0);
    bSh[i] = b[i] * S[i] * d;
    if (fabs(bSh[i]) < 10e-6) any_small = 1;
  }
  if (any_small) {
    R[0] = R0[0];
    R[1] = R0[1];
    R[2] = R0[2];
    return;
  }
  double out[3];
  for (int i = 0; i < 3; i++) {
    /* GLSL coth: 1 for x > 20, cosh/sinh otherwise (shader :13-22) */
    const double c = (bSh[i] > 20.0) ? 1.0 : cosh(bSh[i]) / sinh(bSh[i]);
    const double bc = b[i] * c;
    out[i] = (1.0 - R0[i] * (a[i] - bc)) / (a[i] - R0[i] + bc);
  }
  R[0] = out[0];
  R[1] = out[1];
  R[2] = out[2];
}

/* painty/renderer/Renderer.hxx:26-41 */
void orc_compose(int64_t n, const double* K, const double* S, const double* V,
                 const double* R0, double* R) {
  for (int64_t i = 0; i < n; i++)
    orc_compute_reflectance(K + 3 * i, S + 3 * i, R0 + 3 * i, V[i], R + 3 * i);
}

/* painty/renderer/PaintLayer.hxx:81-96 (the resize-to-ones branch :82-87 is
 * the caller's business: sizes are explicit here) */
void orc_compose_onto(int64_t n, const double* K, const double* S,
                      const double* V, double* R0) {
  for (int64_t i = 0; i < n; i++)
    orc_compute_reflectance(K + 3 * i, S + 3 * i, R0 + 3 * i, V[i], R0 + 3 * i);
}

/* painty/renderer/Canvas.hxx:105-121 */
void orc_dry_canvas(int64_t n, double* K, double* S, double* V, double* R0,
                    double* h) {
  for (int64_t i = 0; i < n; i++) {
    const double v = V[i];
    h[i] += v;
    orc_compute_reflectance(K + 3 * i, S + 3 * i, R0 + 3 * i, v, R0 + 3 * i);
    V[i] = 0.0;
    K[3 * i + 0] = K[3 * i + 1] = K[3 * i + 2] = 0.0;
    S[3 * i + 0] = S[3 * i + 1] = S[3 * i + 2] = 0.0;
  }
}

/* painty/renderer/shaders/DryingStep.compute.glsl:66-76 */
void orc_dry_step(int64_t n, const double* K, const double* S, double* V,
                  double* R0, double portion) {
  for (int64_t i = 0; i < n; i++) {
    const double d = V[i];
    const double rem = (d - portion > 0.0) ? d - portion : 0.0;
    const double leave = d - rem;
    if (fabs(leave) > 10e-6)
      orc_compute_reflectance_glsl(K + 3 * i, S + 3 * i, R0 + 3 * i, leave,
                                   R0 + 3 * i);
    V[i] = rem;
  }
}

/* ------------------------------------------------------------------------ */
/* FootprintBrush (painty/renderer/FootprintBrush.hxx) */
struct orc_brush {
  double radius;
  int size_map;
  double* footprint; /* size_map^2 */
  double *pK, *pS, *pV; /* pickup map (PaintLayer sizeMap x sizeMap) */
  int snap_rows, snap_cols;
  double *sK, *sS, *sV; /* snapshot buffer (PaintLayer rows x cols) */
  int use_snapshot;
  double cap, pickup_rate, deposition_rate;
  double paintK[3], paintS[3];
};

static const double kMinVolume = 0.001; /* FootprintBrush.hxx:26 */

/* FootprintBrush.hxx:29-37, 46-63, 477-495 */
orc_brush* orc_brush_create(double radius, const double* core, int width) {
  const int w = (int)(2.0 * ceil(radius) + 1.0);
  if (w != width || core == NULL) return NULL;
  orc_brush* b = (orc_brush*)calloc(1, sizeof(orc_brush));
  b->radius = radius;
  b->size_map = (int)ceil(sqrt(2.0) * w);
  const int pad = (b->size_map - w) / 2;
  const int64_t m = (int64_t)b->size_map * b->size_map;
  b->footprint = (double*)calloc((size_t)m, sizeof(double));
  for (int i = 0; i < w; i++)
    for (int j = 0; j < w; j++)
      b->footprint[(int64_t)(i + pad) * b->size_map + (j + pad)] = core[i * w + j];
  b->pK = (double*)calloc((size_t)m * 3, sizeof(double));
  b->pS = (double*)calloc((size_t)m * 3, sizeof(double));
  b->pV = (double*)calloc((size_t)m, sizeof(double));
  b->use_snapshot = 1;
  b->cap = 1.0;
  b->pickup_rate = 0.9;
  b->deposition_rate = 0.05;
  return b;
}

/* FootprintBrush.hxx:46-63 */
int orc_brush_set_radius(orc_brush* b, double radius, const double* core,
                         int width) {
  if (fabs(b->radius - radius) < 0.5) return 0; /* fuzzyCompare(.., 0.5) */
  orc_brush* nb = orc_brush_create(radius, core, width);
  if (!nb) return -1;
  free(b->footprint);
  free(b->pK);
  free(b->pS);
  free(b->pV);
  b->radius = nb->radius;
  b->size_map = nb->size_map;
  b->footprint = nb->footprint;
  b->pK = nb->pK;
  b->pS = nb->pS;
  b->pV = nb->pV;
  free(nb);
  return 0;
}

void orc_brush_destroy(orc_brush* b) {
  if (!b) return;
  free(b->footprint);
  free(b->pK);
  free(b->pS);
  free(b->pV);
  free(b->sK);
  free(b->sS);
  free(b->sV);
  free(b);
}

int orc_brush_size_map(const orc_brush* b) { return b->size_map; }
const double* orc_brush_footprint(const orc_brush* b) { return b->footprint; }

void orc_brush_set_rates(orc_brush* b, double pickup, double deposition) {
  b->pickup_rate = pickup;
  b->deposition_rate = deposition;
}

void orc_brush_set_use_snapshot(orc_brush* b, int use) { b->use_snapshot = use; }

/* FootprintBrush.hxx:160-166 */
void orc_brush_clean(orc_brush* b) {
  const int64_t m = (int64_t)b->size_map * b->size_map;
  memset(b->pK, 0, sizeof(double) * 3 * (size_t)m);
  memset(b->pS, 0, sizeof(double) * 3 * (size_t)m);
  memset(b->pV, 0, sizeof(double) * (size_t)m);
}

/* FootprintBrush.hxx:150-154 */
void orc_brush_dip(orc_brush* b, const double K[3], const double S[3]) {
  orc_brush_clean(b);
  for (int c = 0; c < 3; c++) {
    b->paintK[c] = K[c];
    b->paintS[c] = S[c];
  }
}

/* PaintLayer::copyTo (PaintLayer.hxx:103-111), used by
 * FootprintBrush.hxx:168-172 and :281-284 */
void orc_brush_update_snapshot(orc_brush* b, int rows, int cols,
                               const double* K, const double* S,
                               const double* V) {
  const int64_t n = (int64_t)rows * cols;
  if (b->snap_rows != rows || b->snap_cols != cols) {
    free(b->sK);
    free(b->sS);
    free(b->sV);
    b->sK = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    b->sS = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    b->sV = (double*)malloc(sizeof(double) * (size_t)n);
    b->snap_rows = rows;
    b->snap_cols = cols;
  }
  memcpy(b->sK, K, sizeof(double) * 3 * (size_t)n);
  memcpy(b->sS, S, sizeof(double) * 3 * (size_t)n);
  memcpy(b->sV, V, sizeof(double) * (size_t)n);
}

void orc_brush_get_pickup(const orc_brush* b, double* K, double* S, double* V) {
  const int64_t m = (int64_t)b->size_map * b->size_map;
  memcpy(K, b->pK, sizeof(double) * 3 * (size_t)m);
  memcpy(S, b->pS, sizeof(double) * 3 * (size_t)m);
  memcpy(V, b->pV, sizeof(double) * (size_t)m);
}

/* FootprintBrush.hxx:278-319 */
static void update_snapshot_ring(orc_brush* b, int rows, int cols,
                                 const double* K, const double* S,
                                 const double* V, double cx, double cy) {
  if (b->snap_cols != cols || b->snap_rows != rows)
    orc_brush_update_snapshot(b, rows, cols, K, S, V);
  const int h = b->size_map;
  const int hr = (h - 1) / 2;
  const int wr = hr;
  /* Eigen int vectors built from doubles: truncation toward zero */
  const int tl_x = (int)(cx - wr), tl_y = (int)(cy - hr);
  const int br_x = (int)(cx + wr), br_y = (int)(cy + hr);
  int a_tl_x = (int)(cx - wr - b->radius);
  int a_tl_y = (int)(cy - hr - b->radius);
  int a_br_x = (int)(cx + wr + b->radius);
  int a_br_y = (int)(cy + hr + b->radius);
  if (a_tl_x < 0) a_tl_x = 0;
  if (a_tl_y < 0) a_tl_y = 0;
  if (a_br_x > cols - 1) a_br_x = cols - 1;
  if (a_br_y > rows - 1) a_br_y = rows - 1;
  for (int row = a_tl_y; row <= a_br_y; row++) {
    for (int col = a_tl_x; col <= a_br_x; col++) {
      if (row > tl_y && row < br_y && col > tl_x && col < br_x) continue;
      const int64_t i = (int64_t)row * cols + col;
      for (int c = 0; c < 3; c++) {
        b->sK[3 * i + c] = K[3 * i + c];
        b->sS[3 * i + c] = S[3 * i + c];
      }
      b->sV[i] = V[i];
    }
  }
}

/* FootprintBrush.hxx:331-340 */
static void blend3(double va, const double* a, double vb, const double* b,
                   double* out) {
  const double vt = va + vb;
  double r0 = a[0], r1 = a[1], r2 = a[2];
  if (vt > kMinVolume) {
    r0 = (va * a[0] + vb * b[0]) / vt;
    r1 = (va * a[1] + vb * b[1]) / vt;
    r2 = (va * a[2] + vb * b[2]) / vt;
  }
  out[0] = r0;
  out[1] = r1;
  out[2] = r2;
}

/* FootprintBrush.hxx:349-384 */
static void pickup_paint(orc_brush* b, int64_t ic, int64_t im, double* srcK,
                         double* srcS, double* srcV) {
  const double fh = b->footprint[im];
  if (fh > 0.0) {
    const double vP = b->pV[im];
    const double vC = srcV[ic];
    const double vLeave = b->pickup_rate * vC * fh;
    if (vLeave > kMinVolume) {
      srcV[ic] = vC - vLeave;
      double k[3], s[3];
      blend3(vP, b->pK + 3 * im, vLeave, srcK + 3 * ic, k);
      blend3(vP, b->pS + 3 * im, vLeave, srcS + 3 * ic, s);
      for (int c = 0; c < 3; c++) {
        b->pK[3 * im + c] = k[c];
        b->pS[3 * im + c] = s[c];
      }
      b->pV[im] = vP + vLeave;
    }
  }
}

/* FootprintBrush.hxx:393-431 */
static void deposit_paint(orc_brush* b, int64_t ic, int64_t im, double* K,
                          double* S, double* V) {
  const double fh = b->footprint[im];
  if (fh > 0.0) {
    const double vP = b->pV[im];
    const double vFree = (b->cap - vP > 0.0) ? b->cap - vP : 0.0; /* max(0,.) */
    double ks[3], ss[3];
    blend3(vP, b->pK + 3 * im, vFree, b->paintK, ks);
    blend3(vP, b->pS + 3 * im, vFree, b->paintS, ss);
    const double vLeave = b->deposition_rate * vP * fh;
    b->pV[im] = vP - vLeave;
    const double vC = V[ic];
    const double vB = b->cap * fh;
    double k[3], s[3];
    blend3(vB, ks, vC, K + 3 * ic, k);
    blend3(vB, ss, vC, S + 3 * ic, s);
    for (int c = 0; c < 3; c++) {
      K[3 * ic + c] = k[c];
      S[3 * ic + c] = s[c];
    }
    V[ic] = vB + vC;
  }
}

/* FootprintBrush.hxx:73-143 */
/* Canvas::checkDry (Canvas.hxx:123-155) for pixel (x, y) at clock time
 * now_us; the time map holds microseconds, and the elapsed time is truncated
 * to whole milliseconds as duration_cast<milliseconds> does. */
static void check_dry(const orc_clock* c, int cols, int x, int y, double* K,
                      double* S, double* V) {
  const int64_t ic = (int64_t)y * cols + x;
  const double v = V[ic];
  if (c->dry_ms > 0 && v > 0.001) {
    const int64_t dur = (c->now_us - c->tmap[ic]) / 1000;
    if (dur >= c->dry_ms) {
      c->h[ic] += v;
      orc_compute_reflectance(K + 3 * ic, S + 3 * ic, c->R0 + 3 * ic, v,
                              c->R0 + 3 * ic);
      V[ic] = 0.0;
      K[3 * ic + 0] = K[3 * ic + 1] = K[3 * ic + 2] = 0.0;
      S[3 * ic + 0] = S[3 * ic + 1] = S[3 * ic + 2] = 0.0;
    } else {
      /* integer quotient, as in the reference (Canvas.hxx:141) */
      const double rate = (double)(dur / c->dry_ms);
      if (rate > 0.01) {
        const double vl = rate * v;
        c->h[ic] += vl;
        orc_compute_reflectance(K + 3 * ic, S + 3 * ic, c->R0 + 3 * ic, vl,
                                c->R0 + 3 * ic);
        V[ic] = v - vl;
      }
    }
  }
  c->tmap[ic] = c->now_us;
}

static void imprint_impl(orc_brush* b, int rows, int cols, double* K, double* S,
                         double* V, double cx, double cy, double theta,
                         const orc_clock* clk) {
  const int h = b->size_map; /* footprint rows (padded to sizeMap) */
  const int hr = (h - 1) / 2;
  const int wr = hr;
  double *srcK = K, *srcS = S, *srcV = V;
  if (b->use_snapshot) {
    update_snapshot_ring(b, rows, cols, K, S, V, cx, cy);
    srcK = b->sK;
    srcS = b->sS;
    srcV = b->sV;
  }
  const double cosT = cos(-theta);
  const double sinT = sin(-theta);
  for (int row = -hr; row <= hr; row++) {
    for (int col = -wr; col <= wr; col++) {
      const int x = (int)(col + cx);
      const int y = (int)(row + cy);
      const double rc = col * cosT - row * sinT;
      const double rr = col * sinT + row * cosT;
      const int mx = (int)round(rc + wr);
      const int my = (int)round(rr + hr);
      if (y < 0 || x < 0 || x >= cols || y >= rows) continue;
      if (my < 0 || mx < 0 || mx >= b->size_map || my >= b->size_map) continue;
      if (clk) check_dry(clk, cols, x, y, K, S, V); /* FootprintBrush.hxx:115 */
      const int64_t ic = (int64_t)y * cols + x;
      const int64_t im = (int64_t)my * b->size_map + mx;
      pickup_paint(b, ic, im, srcK, srcS, srcV);
      deposit_paint(b, ic, im, K, S, V);
    }
  }
}

void orc_brush_imprint(orc_brush* b, int rows, int cols, double* K, double* S,
                       double* V, double cx, double cy, double theta) {
  imprint_impl(b, rows, cols, K, S, V, cx, cy, theta, NULL);
}

void orc_brush_imprint_timed(orc_brush* b, int rows, int cols, double* K,
                             double* S, double* V, double cx, double cy,
                             double theta, const orc_clock* clk) {
  imprint_impl(b, rows, cols, K, S, V, cx, cy, theta, clk);
}

void orc_run_stamps(orc_brush* b, int rows, int cols, double* K, double* S,
                    double* V, double* R0, double* h, const orc_stamp* stamps,
                    int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    const orc_stamp* st = stamps + i;
    if (st->flags & ORC_DIP) orc_brush_dip(b, st->K, st->S);
    orc_brush_imprint(b, rows, cols, K, S, V, st->cx, st->cy, st->theta);
    if (st->flags & ORC_DRY_AFTER)
      orc_dry_canvas((int64_t)rows * cols, K, S, V, R0, h);
  }
}

/* ------------------------------------------------------------------------ */
/* painty/mixer/src/PaintMixer.cxx:327-355 (Eps = 10e-6 at :278) */
int orc_mix_single(int n_palette, const double* baseK, const double* baseS,
                   int n_weights, const double* w, double K[3], double S[3]) {
  if (n_weights != n_palette) return -1;
  double norm = 1.0, wsum = 0.0;
  for (int i = 0; i < n_weights; i++) wsum += w[i];
  if (!(fabs(wsum - 1.0) < 10e-6)) norm = 1. / wsum;
  for (int c = 0; c < 3; c++) {
    K[c] = 0.0;
    S[c] = 0.0;
  }
  for (int l = 0; l < n_palette; l++) {
    for (int c = 0; c < 3; c++) {
      K[c] += norm * w[l] * baseK[3 * l + c];
      S[c] += norm * w[l] * baseS[3 * l + c];
    }
  }
  return 0;
}

/* painty/mixer/src/PaintMixer.cxx:536-548 */
void orc_mixed(const double K1[3], const double S1[3], double v1,
               const double K2[3], const double S2[3], double v2, double K[3],
               double S[3]) {
  const double totalV = v1 + v2;
  const double inv = 1.0 / totalV;
  for (int c = 0; c < 3; c++) {
    K[c] = ((v1 * K1[c]) + (v2 * K2[c])) * inv;
    S[c] = ((v1 * S1[c]) + (v2 * S2[c])) * inv;
  }
}

/* ------------------------------------------------------------------------ */
/* painty/core/Spline.hxx:27-45 */
double orc_catmull_rom(double p_1, double p0, double p1, double p2, double t) {
  const double tau = 0.5;
  const double t2 = t * t;
  const double t3 = t2 * t;
  const double b1 = -tau * t + 2.0 * tau * t2 - tau * t3;
  const double b2 = 1.0 + (tau - 3.0) * t2 + (2.0 - tau) * t3;
  const double b3 = tau * t + (3.0 - 2.0 * tau) * t2 + (tau - 2.0) * t3;
  const double b4 = -tau * t2 + tau * t3;
  return (p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4);
}

/* painty/core/Spline.hxx:47-73 */
double orc_catmull_rom_d1(double p_1, double p0, double p1, double p2,
                          double t) {
  const double tau = 0.5;
  const double t2 = t * t;
  const double b1 = tau * (-3.0 * t2 + 4.0 * t - 1.0);
  const double b2 = -t * (-2.0 * tau + 3.0 * (tau - 2.0) * t + 6.0);
  const double b3 = (t - 1.0) * (3.0 * (tau - 2.0) * t - tau);
  const double b4 = tau * t * (3.0 * t - 2.0);
  return (p_1 * b1 + p0 * b2 + p1 * b3 + p2 * b4);
}

/* painty/renderer/FootprintBrush.hxx:251-267 */
int64_t orc_stroke_stamps(const double* p, int64_t n, double* out,
                          int64_t max_out) {
  int64_t k = 0;
  for (int64_t i = 0; i + 1 < n; i++) {
    const int64_t ipre = (i > 0) ? i - 1 : 0;
    const int64_t inext = (i + 2 < n) ? i + 2 : n - 1;
    const double* pp = p + 2 * ipre;
    const double* p0 = p + 2 * i;
    const double* p1 = p + 2 * (i + 1);
    const double* pn = p + 2 * inext;
    const double dx = p1[0] - p0[0], dy = p1[1] - p0[1];
    const double dist = sqrt(dx * dx + dy * dy);
    for (int pd = 1; pd <= (int)dist; pd++) {
      const double t = (double)pd / dist;
      const double ddx = orc_catmull_rom_d1(pp[0], p0[0], p1[0], pn[0], t);
      const double ddy = orc_catmull_rom_d1(pp[1], p0[1], p1[1], pn[1], t);
      if (k < max_out) {
        out[3 * k + 0] = orc_catmull_rom(pp[0], p0[0], p1[0], pn[0], t);
        out[3 * k + 1] = orc_catmull_rom(pp[1], p0[1], p1[1], pn[1], t);
        out[3 * k + 2] = atan2(ddy, ddx);
      }
      k++;
    }
  }
  return k;
}

/* ------------------------------------------------------------------------ */
/* Gradient of composite + stamp chain, FP64 reverse mode.
 *
 * Forward: the stamps of orc_run_stamps (dips take their paint from
 * orc_mix_single of a weight row), then Renderer::compose.  Backward: the
 * adjoint of each operation in reverse order, from per-stamp tapes (the
 * window of canvas and snapshot as the visits see them, and the pickup map
 * before the stamp).  Visits inside a stamp are replayed in the reference's
 * row-major order from the tape, so each visit's adjoint sees its exact
 * inputs.  Branches (volume thresholds, max(0, .), coth's x > 20 clamp, the
 * d and S thresholds of ComputeReflectance) are differentiated as the
 * branch taken. */

/* One visit (pickupPaint then depositPaint, FootprintBrush.hxx:349-431) on
 * explicit state: P = pickup cell, C = canvas pixel, X = pickup source pixel
 * (the snapshot; the canvas pixel itself when `alias`).  Same operations,
 * in the same order, as pickup_paint / deposit_paint above. */
typedef struct {
  double P[7], C[7], X[7]; /* K0 K1 K2 S0 S1 S2 V */
} orc_vstate;

static void visit_fwd(orc_vstate* s, double fh, double pr, double dr,
                      double cap, const double B[6], int alias) {
  double* P = s->P;
  double* C = s->C;
  double* X = alias ? s->C : s->X;
  const double vP = P[6], vC = X[6];
  const double vLeave = pr * vC * fh;
  if (vLeave > kMinVolume) {
    X[6] = vC - vLeave;
    double k[3], ss[3];
    blend3(vP, P, vLeave, X, k);
    blend3(vP, P + 3, vLeave, X + 3, ss);
    for (int c = 0; c < 3; c++) {
      P[c] = k[c];
      P[3 + c] = ss[c];
    }
    P[6] = vP + vLeave;
  }
  const double u = P[6];
  const double vFree = (cap - u > 0.0) ? cap - u : 0.0;
  double ks[3], sk[3];
  blend3(u, P, vFree, B, ks);
  blend3(u, P + 3, vFree, B + 3, sk);
  P[6] = u - dr * u * fh;
  const double vCc = C[6], vB = cap * fh;
  double k2[3], s2[3];
  blend3(vB, ks, vCc, C, k2);
  blend3(vB, sk, vCc, C + 3, s2);
  for (int c = 0; c < 3; c++) {
    C[c] = k2[c];
    C[3 + c] = s2[c];
  }
  C[6] = vB + vCc;
}

/* Adjoint of visit_fwd at input state `pre`.  On entry g holds the adjoints
 * of the outputs (P, C and, unless alias, X); on return the adjoints of the
 * inputs.  gB accumulates the adjoint of the brush paint (K0..S2). */
static void visit_adj(const orc_vstate* pre, orc_vstate* g, double gB[6],
                      double fh, double pr, double dr, double cap,
                      const double B[6], int alias) {
  const double* P = pre->P;
  const double* C = pre->C;
  const double* X = alias ? pre->C : pre->X;
  /* forward intermediates */
  const double vP = P[6], vC = X[6];
  const double vLeave = pr * vC * fh;
  const int a = vLeave > kMinVolume;
  double P1[7];
  for (int c = 0; c < 7; c++) P1[c] = P[c];
  double X6n = vC, vt1 = 0.0;
  int bb = 0;
  if (a) {
    X6n = vC - vLeave;
    vt1 = vP + vLeave;
    bb = vt1 > kMinVolume;
    if (bb)
      for (int c = 0; c < 6; c++) P1[c] = (vP * P[c] + vLeave * X[c]) / vt1;
    P1[6] = vP + vLeave;
  }
  const double u = P1[6];
  const int f = cap - u > 0.0;
  const double vFree = f ? cap - u : 0.0;
  const double vt2 = u + vFree;
  const int d = vt2 > kMinVolume;
  double src[6];
  for (int c = 0; c < 6; c++) src[c] = d ? (u * P1[c] + vFree * B[c]) / vt2 : P1[c];
  double Cd[7]; /* the canvas pixel as depositPaint reads it */
  for (int c = 0; c < 6; c++) Cd[c] = C[c];
  Cd[6] = alias ? X6n : C[6];
  const double vB = cap * fh, vCc = Cd[6], vt3 = vB + vCc;
  const int e = vt3 > kMinVolume;
  double C1[6];
  for (int c = 0; c < 6; c++) C1[c] = e ? (vB * src[c] + vCc * Cd[c]) / vt3 : src[c];

  /* depositPaint */
  double gCd[7], gsrc[6];
  double gvCc = g->C[6]; /* C.V = vB + vCc */
  for (int c = 0; c < 6; c++) {
    if (e) {
      gsrc[c] = g->C[c] * vB / vt3;
      gCd[c] = g->C[c] * vCc / vt3;
      gvCc += g->C[c] * (Cd[c] - C1[c]) / vt3;
    } else {
      gsrc[c] = g->C[c];
      gCd[c] = 0.0;
    }
  }
  gCd[6] = gvCc;
  double gP1[7];
  double gu = g->P[6] * (1.0 - dr * fh); /* P.V = u - dr*u*fh */
  if (d) {
    double gvFree = 0.0;
    for (int c = 0; c < 6; c++) {
      gP1[c] = g->P[c] + gsrc[c] * u / vt2;
      gB[c] += gsrc[c] * vFree / vt2;
      gu += gsrc[c] * (P1[c] - src[c]) / vt2;
      gvFree += gsrc[c] * (B[c] - src[c]) / vt2;
    }
    if (f) gu -= gvFree;
  } else {
    for (int c = 0; c < 6; c++) gP1[c] = g->P[c] + gsrc[c];
  }
  gP1[6] = gu;
  /* the source pixel after pickupPaint */
  double gXo[7];
  for (int c = 0; c < 7; c++) gXo[c] = alias ? gCd[c] : g->X[c];
  /* pickupPaint */
  double gPin[7], gXin[7];
  for (int c = 0; c < 7; c++) gXin[c] = gXo[c];
  if (a) {
    double gvP = gP1[6], gvLeave = gP1[6] - gXo[6], gvC = gXo[6];
    for (int c = 0; c < 6; c++) {
      if (bb) {
        gPin[c] = gP1[c] * vP / vt1;
        gXin[c] += gP1[c] * vLeave / vt1;
        gvP += gP1[c] * (P[c] - P1[c]) / vt1;
        gvLeave += gP1[c] * (X[c] - P1[c]) / vt1;
      } else {
        gPin[c] = gP1[c];
      }
    }
    gvC += gvLeave * pr * fh;
    gPin[6] = gvP;
    gXin[6] = gvC;
  } else {
    for (int c = 0; c < 7; c++) gPin[c] = gP1[c];
  }
  for (int c = 0; c < 7; c++) g->P[c] = gPin[c];
  if (alias) {
    for (int c = 0; c < 7; c++) g->C[c] = gXin[c];
  } else {
    for (int c = 0; c < 7; c++) {
      g->C[c] = gCd[c];
      g->X[c] = gXin[c];
    }
  }
}

/* Adjoint of ComputeReflectance (KubelkaMunk.hxx:25-83) for one pixel:
 * gK, gS, gR0 set, *gd set. */
void orc_compute_reflectance_adj(const double K[3], const double S_in[3],
                                 const double R0[3], double d,
                                 const double gR[3], double gK[3],
                                 double gS[3], double* gd, double gR0[3]) {
  const double thr = DBL_EPSILON * 10000.0;
  *gd = 0.0;
  if (fabs(d) < thr) {
    for (int i = 0; i < 3; i++) {
      gR0[i] = gR[i];
      gK[i] = gS[i] = 0.0;
    }
    return;
  }
  for (int i = 0; i < 3; i++) {
    const int clamp = !(fabs(S_in[i]) > thr);
    const double S = clamp ? 0.00000000001 : S_in[i];
    const double a = 1.0 + K[i] / S;
    const double v = a * a - 1.0;
    const double b = (v < 0.0) ? 0.0 : sqrt(v);
    const double x = b * S * d;
    const double ct = orc_coth(x);
    double csch2 = 0.0; /* -d coth / dx; 0 where coth is clamped to 1 */
    if (x <= 20.0 && fabs(x) > 0.0) {
      const double sh = sinh(x);
      csch2 = 1.0 / (sh * sh);
    }
    const double q = b * ct;
    const double N = 1.0 - R0[i] * (a - q);
    const double D = a - R0[i] + q;
    const double dR_dR0 = (N - (a - q) * D) / (D * D);
    const double dR_da = (-R0[i] * D - N) / (D * D);
    const double dR_dq = (R0[i] * D - N) / (D * D);
    const double dq_db = ct - x * csch2;
    const double dq_dS = -b * b * d * csch2;
    const double dq_dd = -b * b * S * csch2;
    const double db_da = (v > 0.0) ? a / b : 0.0;
    const double dRa = dR_da + dR_dq * dq_db * db_da;
    gK[i] = gR[i] * dRa / S;
    gS[i] = clamp ? 0.0 : gR[i] * (dRa * (-K[i] / (S * S)) + dR_dq * dq_dS);
    *gd += gR[i] * dR_dq * dq_dd;
    gR0[i] = gR[i] * dR_dR0;
  }
}

/* Window box of a stamp: canvas pixels its visits can touch. */
static void window_box(const orc_brush* b, int rows, int cols, double cx,
                       double cy, int* x0, int* y0, int* x1, int* y1) {
  const int hr = (b->size_map - 1) / 2;
  int a = (int)(-hr + cx), c = (int)(hr + cx), e = (int)(-hr + cy), f = (int)(hr + cy);
  if (a < 0) a = 0;
  if (e < 0) e = 0;
  if (c > cols - 1) c = cols - 1;
  if (f > rows - 1) f = rows - 1;
  *x0 = a;
  *x1 = c;
  *y0 = e;
  *y1 = f;
}

typedef struct {
  int x0, y0, x1, y1; /* window box (empty when x1 < x0 or y1 < y0) */
  double* C;          /* box pixels x 7, canvas as the visits find it */
  double* X;          /* box pixels x 7, snapshot as the visits find it */
  double* P;          /* size_map^2 x 7, pickup map before the stamp */
  double B[6];        /* brush paint */
} orc_tape;

static void get7(const double* K, const double* S, const double* V, int64_t i,
                 double* o) {
  for (int c = 0; c < 3; c++) {
    o[c] = K[3 * i + c];
    o[3 + c] = S[3 * i + c];
  }
  o[6] = V[i];
}

int orc_grad_run(orc_brush* b, int rows, int cols, double* K, double* S,
                 double* V, const double* R0, const orc_stamp* st, int64_t n,
                 const int* paint_of, int n_pig, const double* palK,
                 const double* palS, int n_paint, const double* W,
                 const double* gR, double* R, double* gK, double* gS,
                 double* gV, double* gR0, double* gW) {
  const int64_t npx = (int64_t)rows * cols;
  const int alias = !b->use_snapshot;
  if (n > 0 && !(st[0].flags & ORC_DIP)) return -1; /* paint must come from a dip */
  for (int64_t i = 0; i < n; i++) {
    if (st[i].flags & ORC_DRY_AFTER) return -1;
    if ((st[i].flags & ORC_DIP) && (paint_of[i] < 0 || paint_of[i] >= n_paint)) return -1;
  }
  if (b->snap_rows == rows && b->snap_cols == cols) return -1; /* needs a fresh snapshot */
  const int sm = b->size_map;
  const int64_t m = (int64_t)sm * sm;
  orc_tape* tape = (orc_tape*)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_tape));
  double paint[6] = {0, 0, 0, 0, 0, 0};
  /* ---- forward, taping each stamp */
  for (int64_t i = 0; i < n; i++) {
    if (st[i].flags & ORC_DIP) {
      double k[3], s[3];
      orc_mix_single(n_pig, palK, palS, n_pig, W + (int64_t)paint_of[i] * n_pig, k, s);
      orc_brush_dip(b, k, s);
      for (int c = 0; c < 3; c++) {
        paint[c] = k[c];
        paint[3 + c] = s[c];
      }
    }
    orc_tape* t = tape + i;
    for (int c = 0; c < 6; c++) t->B[c] = paint[c];
    t->P = (double*)malloc(sizeof(double) * 7 * (size_t)m);
    for (int64_t q = 0; q < m; q++) get7(b->pK, b->pS, b->pV, q, t->P + 7 * q);
    if (!alias) update_snapshot_ring(b, rows, cols, K, S, V, st[i].cx, st[i].cy);
    window_box(b, rows, cols, st[i].cx, st[i].cy, &t->x0, &t->y0, &t->x1, &t->y1);
    const int bw = t->x1 - t->x0 + 1, bh = t->y1 - t->y0 + 1;
    if (bw > 0 && bh > 0) {
      t->C = (double*)malloc(sizeof(double) * 7 * (size_t)bw * bh);
      t->X = (double*)malloc(sizeof(double) * 7 * (size_t)bw * bh);
      for (int y = t->y0; y <= t->y1; y++)
        for (int x = t->x0; x <= t->x1; x++) {
          const int64_t ic = (int64_t)y * cols + x;
          const int64_t k = (int64_t)(y - t->y0) * bw + (x - t->x0);
          get7(K, S, V, ic, t->C + 7 * k);
          if (!alias) get7(b->sK, b->sS, b->sV, ic, t->X + 7 * k);
        }
    }
    /* the visits (the snapshot ring was refreshed above) */
    const int hr = (sm - 1) / 2;
    const double cosT = cos(-st[i].theta), sinT = sin(-st[i].theta);
    for (int row = -hr; row <= hr; row++)
      for (int col = -hr; col <= hr; col++) {
        const int x = (int)(col + st[i].cx), y = (int)(row + st[i].cy);
        const double rc = col * cosT - row * sinT, rr = col * sinT + row * cosT;
        const int mx = (int)round(rc + hr), my = (int)round(rr + hr);
        if (y < 0 || x < 0 || x >= cols || y >= rows) continue;
        if (my < 0 || mx < 0 || mx >= sm || my >= sm) continue;
        const int64_t ic = (int64_t)y * cols + x, im = (int64_t)my * sm + mx;
        if (alias) {
          pickup_paint(b, ic, im, K, S, V);
        } else {
          pickup_paint(b, ic, im, b->sK, b->sS, b->sV);
        }
        deposit_paint(b, ic, im, K, S, V);
      }
  }
  /* ---- composite and its adjoint */
  for (int64_t i = 0; i < npx; i++) {
    orc_compute_reflectance(K + 3 * i, S + 3 * i, R0 + 3 * i, V[i], R + 3 * i);
    double gd;
    orc_compute_reflectance_adj(K + 3 * i, S + 3 * i, R0 + 3 * i, V[i], gR + 3 * i,
                                gK + 3 * i, gS + 3 * i, &gd, gR0 + 3 * i);
    gV[i] = gd;
  }
  double* gsK = (double*)calloc((size_t)npx * 3, sizeof(double));
  double* gsS = (double*)calloc((size_t)npx * 3, sizeof(double));
  double* gsV = (double*)calloc((size_t)npx, sizeof(double));
  double* gP = (double*)calloc((size_t)m * 7, sizeof(double));
  for (int64_t q = 0; q < (int64_t)n_paint * n_pig; q++) gW[q] = 0.0;
  double gpaint[6] = {0, 0, 0, 0, 0, 0};
  /* per-visit replay buffers */
  const int side = sm;
  orc_vstate* pre = (orc_vstate*)malloc(sizeof(orc_vstate) * (size_t)side * side);
  int64_t* vic = (int64_t*)malloc(sizeof(int64_t) * (size_t)side * side);
  int64_t* vim = (int64_t*)malloc(sizeof(int64_t) * (size_t)side * side);
  double* Pm = (double*)malloc(sizeof(double) * 7 * (size_t)m);
  /* ---- backward over the stamps */
  for (int64_t i = n - 1; i >= 0; i--) {
    const orc_tape* t = tape + i;
    const int bw = t->x1 - t->x0 + 1, bh = t->y1 - t->y0 + 1;
    const int64_t nb = (bw > 0 && bh > 0) ? (int64_t)bw * bh : 0;
    double* Cw = (double*)malloc(sizeof(double) * 7 * (size_t)(nb ? nb : 1));
    double* Xw = (double*)malloc(sizeof(double) * 7 * (size_t)(nb ? nb : 1));
    if (nb) {
      memcpy(Cw, t->C, sizeof(double) * 7 * (size_t)nb);
      if (!alias) memcpy(Xw, t->X, sizeof(double) * 7 * (size_t)nb);
    }
    memcpy(Pm, t->P, sizeof(double) * 7 * (size_t)m);
    /* replay the visits in reference order, recording each one's input */
    int64_t nv = 0;
    const int hr = (sm - 1) / 2;
    const double cosT = cos(-st[i].theta), sinT = sin(-st[i].theta);
    for (int row = -hr; row <= hr; row++)
      for (int col = -hr; col <= hr; col++) {
        const int x = (int)(col + st[i].cx), y = (int)(row + st[i].cy);
        const double rc = col * cosT - row * sinT, rr = col * sinT + row * cosT;
        const int mx = (int)round(rc + hr), my = (int)round(rr + hr);
        if (y < 0 || x < 0 || x >= cols || y >= rows) continue;
        if (my < 0 || mx < 0 || mx >= sm || my >= sm) continue;
        const int64_t im = (int64_t)my * sm + mx;
        const double fh = b->footprint[im];
        if (!(fh > 0.0)) continue;
        const int64_t k = (int64_t)(y - t->y0) * bw + (x - t->x0);
        orc_vstate s;
        memset(&s, 0, sizeof(s));
        memcpy(s.P, Pm + 7 * im, sizeof(s.P));
        memcpy(s.C, Cw + 7 * k, sizeof(s.C));
        if (!alias) memcpy(s.X, Xw + 7 * k, sizeof(s.X));
        pre[nv] = s;
        vic[nv] = (int64_t)y * cols + x;
        vim[nv] = im;
        visit_fwd(&s, fh, b->pickup_rate, b->deposition_rate, b->cap, t->B, alias);
        memcpy(Pm + 7 * im, s.P, sizeof(s.P));
        memcpy(Cw + 7 * k, s.C, sizeof(s.C));
        if (!alias) memcpy(Xw + 7 * k, s.X, sizeof(s.X));
        nv++;
      }
    /* adjoints in reverse order */
    for (int64_t v = nv - 1; v >= 0; v--) {
      const int64_t ic = vic[v], im = vim[v];
      orc_vstate g;
      memset(&g, 0, sizeof(g));
      memcpy(g.P, gP + 7 * im, sizeof(g.P));
      get7(gK, gS, gV, ic, g.C);
      if (!alias) get7(gsK, gsS, gsV, ic, g.X);
      visit_adj(pre + v, &g, gpaint, b->footprint[im], b->pickup_rate,
                b->deposition_rate, b->cap, t->B, alias);
      memcpy(gP + 7 * im, g.P, sizeof(g.P));
      for (int c = 0; c < 3; c++) {
        gK[3 * ic + c] = g.C[c];
        gS[3 * ic + c] = g.C[3 + c];
        if (!alias) {
          gsK[3 * ic + c] = g.X[c];
          gsS[3 * ic + c] = g.X[3 + c];
        }
      }
      gV[ic] = g.C[6];
      if (!alias) gsV[ic] = g.X[6];
    }
    free(Cw);
    free(Xw);
    if (!alias) {
      /* the ring refresh copied canvas -> snapshot (FootprintBrush.hxx:307-318) */
      const int tl_x = (int)(st[i].cx - hr), tl_y = (int)(st[i].cy - hr);
      const int br_x = (int)(st[i].cx + hr), br_y = (int)(st[i].cy + hr);
      int ax0 = (int)(st[i].cx - hr - b->radius), ay0 = (int)(st[i].cy - hr - b->radius);
      int ax1 = (int)(st[i].cx + hr + b->radius), ay1 = (int)(st[i].cy + hr + b->radius);
      if (ax0 < 0) ax0 = 0;
      if (ay0 < 0) ay0 = 0;
      if (ax1 > cols - 1) ax1 = cols - 1;
      if (ay1 > rows - 1) ay1 = rows - 1;
      for (int y = ay0; y <= ay1; y++)
        for (int x = ax0; x <= ax1; x++) {
          if (y > tl_y && y < br_y && x > tl_x && x < br_x) continue;
          const int64_t p = (int64_t)y * cols + x;
          for (int c = 0; c < 3; c++) {
            gK[3 * p + c] += gsK[3 * p + c];
            gS[3 * p + c] += gsS[3 * p + c];
            gsK[3 * p + c] = gsS[3 * p + c] = 0.0;
          }
          gV[p] += gsV[p];
          gsV[p] = 0.0;
        }
      if (i == 0) { /* the first imprint copied the whole canvas (:281-284) */
        for (int64_t p = 0; p < npx; p++) {
          for (int c = 0; c < 3; c++) {
            gK[3 * p + c] += gsK[3 * p + c];
            gS[3 * p + c] += gsS[3 * p + c];
          }
          gV[p] += gsV[p];
        }
      }
    }
    if (st[i].flags & ORC_DIP) {
      /* clean() zeroed the pickup map: nothing before flows through it */
      memset(gP, 0, sizeof(double) * 7 * (size_t)m);
      /* paint = mixSinglePaint(W[p]) (PaintMixer.cxx:327-355) */
      const double* w = W + (int64_t)paint_of[i] * n_pig;
      double wsum = 0.0;
      for (int l = 0; l < n_pig; l++) wsum += w[l];
      const int unit = fabs(wsum - 1.0) < 10e-6;
      const double norm = unit ? 1.0 : 1.0 / wsum;
      for (int l = 0; l < n_pig; l++) {
        double acc = 0.0;
        for (int c = 0; c < 3; c++) {
          const double dK = unit ? palK[3 * l + c] : norm * (palK[3 * l + c] - t->B[c]);
          const double dS = unit ? palS[3 * l + c] : norm * (palS[3 * l + c] - t->B[3 + c]);
          acc += gpaint[c] * dK + gpaint[3 + c] * dS;
        }
        gW[(int64_t)paint_of[i] * n_pig + l] += acc;
      }
      for (int c = 0; c < 6; c++) gpaint[c] = 0.0;
    }
  }
  for (int64_t i = 0; i < n; i++) {
    free(tape[i].C);
    free(tape[i].X);
    free(tape[i].P);
  }
  free(tape);
  free(pre);
  free(vic);
  free(vim);
  free(Pm);
  free(gsK);
  free(gsS);
  free(gsV);
  free(gP);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Renderer::render (renderer/Renderer.hxx:60-156): the composite lit by a
 * point light with a Beckmann / Cook-Torrance specular term over the height
 * field of the wet layer's thickness V. */

static const double kPi = 3.141592653589793238462643383279502884; /* Math.hxx:21-22 */

/* cv::borderInterpolate with BORDER_REFLECT (the default of Interpolate,
 * image/Mat.hxx:69-71) for p in [-1, len]: fedcba|abcdefgh|hgfedcb */
static int reflect_index(int p, int len) {
  if (len == 1) return 0;
  if (p < 0) return -p - 1;
  if (p >= len) return 2 * len - p - 1;
  return p;
}

static void v3_normalized(const double a[3], double out[3]) {
  /* Eigen's normalized(): a / sqrt(squaredNorm), unchanged when zero */
  const double z = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  const double n = z > 0.0 ? sqrt(z) : 1.0;
  for (int i = 0; i < 3; i++) out[i] = z > 0.0 ? a[i] / n : a[i];
}

static double v3_dot(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static double render_G(double NdotH, double NdotV, double VdotH, double NdotL) {
  const double G1 = 2.0 * NdotH * NdotV / VdotH;
  const double G2 = 2.0 * NdotH * NdotL / VdotH;
  const double g = (G2 < G1) ? G2 : G1; /* std::min(G1, G2) */
  return (g < 1.0) ? g : 1.0;           /* std::min(1.0, g) */
}

static double render_beckmann(double NdotH, double m) {
  const double A = 1.0 / (pow(m, 2.0) + pow(NdotH, 4.0) * kPi);
  const double B = exp(-pow(tan(acos(NdotH)), 2.0) / pow(m, 2.0));
  return A * B;
}

void orc_render_pixel_normal(double s01, double s21, double s10, double s12,
                             double n[3]) {
  const double a[3] = {2.0, 0.0, s21 - s01}, b[3] = {0.0, 2.0, s12 - s10};
  double va[3], vb[3], c[3];
  v3_normalized(a, va);
  v3_normalized(b, vb);
  c[0] = va[1] * vb[2] - va[2] * vb[1];
  c[1] = va[2] * vb[0] - va[0] * vb[2];
  c[2] = va[0] * vb[1] - va[1] * vb[0];
  v3_normalized(c, n);
  n[2] *= -1.0;
}

void orc_render(int rows, int cols, const double* K, const double* S,
                const double* V, const double* R0, double* out) {
  const double lightPos[3] = {-200.0, -1500.0, -2000.0};
  const double eyePos[3] = {cols / 2.0, rows / 2.0, -100.0};
  const double lightPower = 15.0, Ks = 1.0, m = 0.5, s = 0.2;
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) {
      const int64_t p = (int64_t)i * cols + j;
      double Kd[3];
      orc_compute_reflectance(K + 3 * p, S + 3 * p, R0 + 3 * p, V[p], Kd);
      const double s11 = V[p];
      const double s01 = V[(int64_t)i * cols + reflect_index(j - 1, cols)];
      const double s21 = V[(int64_t)i * cols + reflect_index(j + 1, cols)];
      const double s10 = V[(int64_t)reflect_index(i - 1, rows) * cols + j];
      const double s12 = V[(int64_t)reflect_index(i + 1, rows) * cols + j];
      double n[3];
      orc_render_pixel_normal(s01, s21, s10, s12, n);
      const double pix[3] = {(double)j, (double)i, s11};
      double ld0[3], ld[3], l[3], v0[3], v[3], h0[3], h[3];
      for (int k = 0; k < 3; k++) ld0[k] = lightPos[k] - pix[k];
      v3_normalized(ld0, ld);
      v3_normalized(ld, l);
      for (int k = 0; k < 3; k++) v0[k] = eyePos[k] - pix[k];
      v3_normalized(v0, v);
      for (int k = 0; k < 3; k++) h0[k] = v[k] + l[k];
      v3_normalized(h0, h);
      const double NdotH = fmax(0.0, v3_dot(n, h));
      const double VdotH = fmax(0.0, v3_dot(v, h));
      const double NdotV = fmax(0.0, v3_dot(n, v));
      const double NdotL = fmax(0.0, v3_dot(n, l));
      double spec[3] = {0.0, 0.0, 0.0};
      if (NdotL > 0.0 && NdotV > 0.0) {
        const double bg = render_beckmann(NdotH, m) * render_G(NdotH, NdotV, VdotH, NdotL);
        const double rf = Ks + (1.0 - Ks) * pow(1.0 - VdotH, 5.0);
        for (int k = 0; k < 3; k++) spec[k] = (bg * rf) / (NdotL * NdotV);
      }
      const double ldn = sqrt(v3_dot(ld, ld));
      const double beta = lightPower * (1.0 / (4.0 * kPi * pow(ldn, 2.0)));
      for (int k = 0; k < 3; k++) {
        const double r = (beta * NdotL) * ((1.0 - s) * Kd[k] + s * spec[k]) +
                         (Kd[k] * 0.2) * Kd[k];
        out[3 * p + k] = fmin(fmax(r, 0.0), 1.0);
      }
    }
  }
}

/* ColorConverter::rgb2srgb (core/Color.hxx:197-205) */
double orc_rgb2srgb(double l) {
  if (l <= 0.00313066844250063) return l * 12.92;
  return 1.055 * pow(l, 1.0 / 2.4) - 0.055;
}

/* io::imSave's pixel transform (io/src/ImageIO.cxx:158-183): optional
 * linear -> sRGB, then Mat::convertTo to 8 or 16 bits (scale 2^bits - 1,
 * cvRound = round half to even, saturate).  Output interleaved RGB (the
 * reference swaps to BGR only for OpenCV's writer). */
void orc_encode(int64_t n, const double* rgb, int to_srgb, int bits,
                uint16_t* out) {
  const double scale = bits == 16 ? 65535.0 : 255.0;
  for (int64_t i = 0; i < 3 * n; i++) {
    const double x = to_srgb ? orc_rgb2srgb(rgb[i]) : rgb[i];
    double r = nearbyint(x * scale);
    if (r < 0.0) r = 0.0;
    if (r > scale) r = scale;
    out[i] = (uint16_t)r;
  }
}
