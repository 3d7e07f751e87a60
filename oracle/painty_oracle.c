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
void orc_brush_imprint(orc_brush* b, int rows, int cols, double* K, double* S,
                       double* V, double cx, double cy, double theta) {
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
      const int64_t ic = (int64_t)y * cols + x;
      const int64_t im = (int64_t)my * b->size_map + mx;
      pickup_paint(b, ic, im, srcK, srcS, srcV);
      deposit_paint(b, ic, im, K, S, V);
    }
  }
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
