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
  do {
    if (p < 0)
      p = -p - 1;
    else if (p >= len)
      p = 2 * len - p - 1;
  } while (p < 0 || p >= len);
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

/* ------------------------------------------------------------------------ */
/* TextureBrush (renderer/TextureBrush.hxx:52-205) with Smudge
 * (renderer/Smudge.hxx:22-229): a brush-stroke texture (thickness map)
 * warped along the stroke by generalized barycentric coordinates
 * (core/Math.hxx:73-133, Hormann & Floater mean value coordinates), an
 * optional smudge pass, then volume-weighted mixing with V = max. */

/* SplineEval::catmullRom / catmullRomDerivativeFirst (core/Spline.hxx:150-
 * 221) over n 2-d points: u in [0, 1] spans the n - 1 segments, indices
 * clamped to the ends. */
static void spline_cr(const double* v, int n, double u, int deriv, double out[2]) {
  const double x = (double)(n - 1) * u;
  const int index = (int)x;
  const double t = x - floor(x);
  const int id[4] = {index - 1, index, index + 1, index + 2};
  double p[4][2];
  for (int q = 0; q < 4; q++) {
    const int i = id[q] < 0 ? 0 : (id[q] >= n ? n - 1 : id[q]);
    p[q][0] = v[2 * i];
    p[q][1] = v[2 * i + 1];
  }
  for (int c = 0; c < 2; c++)
    out[c] = deriv ? orc_catmull_rom_d1(p[0][c], p[1][c], p[2][c], p[3][c], t)
                   : orc_catmull_rom(p[0][c], p[1][c], p[2][c], p[3][c], t);
}

/* generalizedBarycentricCoordinatesInterpolate (core/Math.hxx:73-133) of
 * 2-d values at the np vertices of a polygon. */
void orc_gbc(const double* poly, int np, double px, double py, const double* vals,
             double out[2]) {
  const double Eps = DBL_EPSILON * 100.0;
  if (np == 1) {
    out[0] = vals[0];
    out[1] = vals[1];
    return;
  }
  double* s = (double*)malloc(sizeof(double) * 2 * np);
  double* r = (double*)malloc(sizeof(double) * np);
  double* A = (double*)malloc(sizeof(double) * np);
  double* D = (double*)malloc(sizeof(double) * np);
  for (int i = 0; i < np; i++) {
    s[2 * i] = poly[2 * i] - px;
    s[2 * i + 1] = poly[2 * i + 1] - py;
  }
  int done = 0;
  for (int i = 0; i < np && !done; i++) {
    const int i1 = (i == np - 1) ? 0 : i + 1;
    const double* si = s + 2 * i;
    const double* si1 = s + 2 * i1;
    r[i] = sqrt(si[0] * si[0] + si[1] * si[1]);
    if (fabs(r[i] - 0.0) < Eps) {
      out[0] = vals[2 * i];
      out[1] = vals[2 * i + 1];
      done = 1;
      break;
    }
    const double det = si[0] * si1[1] - si1[0] * si[1];
    A[i] = det / 2.0;
    D[i] = si[0] * si1[0] + si[1] * si1[1];
    if (fabs(A[i] - 0.0) < Eps && D[i] < 0.0) {
      const double ri1 = sqrt(si1[0] * si1[0] + si1[1] * si1[1]);
      if (i == np - 1) r[0] = ri1; else r[i + 1] = ri1;
      for (int c = 0; c < 2; c++)
        out[c] = (ri1 * vals[2 * i + c] + r[i] * vals[2 * i1 + c]) * (1.0 / (r[i] + ri1));
      done = 1;
    }
  }
  if (!done) {
    double f[2] = {0.0, 0.0}, W = 0.0;
    for (int i = 0; i < np; i++) {
      double w = 0.0;
      const int im = (i == 0) ? np - 1 : i - 1;
      const int ip = (i == np - 1) ? 0 : i + 1;
      if (A[im] != 0.0) w = w + (r[im] - D[im] / r[i]) / A[im];
      if (A[i] != 0.0) w = w + (r[ip] - D[i] / r[i]) / A[i];
      f[0] = f[0] + w * vals[2 * i];
      f[1] = f[1] + w * vals[2 * i + 1];
      W = W + w;
    }
    if (!(fabs(W - 0.0) < Eps)) {
      out[0] = f[0] * (1.0 / W);
      out[1] = f[1] * (1.0 / W);
    } else {
      out[0] = vals[0];
      out[1] = vals[1];
    }
  }
  free(s);
  free(r);
  free(A);
  free(D);
}

/* Interpolate (image/Mat.hxx:69-100) of a single-channel image with
 * BORDER_REFLECT. */
static double interp_reflect_n(const double* img, int rows, int cols, int nch, double x,
                               double y) {
  const int xi = (int)floor(x), yi = (int)floor(y);
  const int x0 = reflect_index(xi, cols), x1 = reflect_index(xi + 1, cols);
  const int y0 = reflect_index(yi, rows), y1 = reflect_index(yi + 1, rows);
  const double a = x - (double)xi, c = y - (double)yi;
#define PX(yy, xx) img[((int64_t)(yy) * cols + (xx)) * nch]
  return (PX(y0, x0) * (1.0 - a) + PX(y0, x1) * a) * (1.0 - c) +
         (PX(y1, x0) * (1.0 - a) + PX(y1, x1) * a) * c;
#undef PX
}

static double interp_reflect(const double* img, int rows, int cols, double x, double y) {
  return interp_reflect_n(img, rows, cols, 1, x, y);
}

/* The smudge pickup maps: AoS K, S (3 per cell) and V of map x map cells. */
struct orc_tbrush {
  double radius;
  double paint[6]; /* K, S */
  double* thick;   /* the brush stroke sample's thickness map */
  int t_rows, t_cols;
  double thickness_scale;
  int use_smudge;
  int map;     /* smudge map side (Smudge(size): size x size) */
  int max_size; /* size, made odd */
  double *dK, *dS, *dV, *sK, *sS, *sV; /* _pickupMapDst / _pickupMapSrc */
  double cur_rot;
};

static void tb_smudge_reset(orc_tbrush* b, int size) {
  free(b->dK); free(b->dS); free(b->dV); free(b->sK); free(b->sS); free(b->sV);
  b->map = size < 0 ? 0 : size;
  b->max_size = (size % 2 == 0) ? size + 1 : size;
  const size_t m = (size_t)b->map * b->map;
  b->dK = (double*)calloc(3 * m + 1, sizeof(double));
  b->dS = (double*)calloc(3 * m + 1, sizeof(double));
  b->dV = (double*)calloc(m + 1, sizeof(double));
  b->sK = (double*)calloc(3 * m + 1, sizeof(double));
  b->sS = (double*)calloc(3 * m + 1, sizeof(double));
  b->sV = (double*)calloc(m + 1, sizeof(double));
  b->cur_rot = 0.0;
}

orc_tbrush* orc_tbrush_create(void) {
  orc_tbrush* b = (orc_tbrush*)calloc(1, sizeof(orc_tbrush));
  for (int c = 0; c < 6; c++) b->paint[c] = 0.1; /* TextureBrush.hxx:27-29 */
  b->thickness_scale = 1.0;
  b->use_smudge = 1;
  tb_smudge_reset(b, 0); /* Smudge(2.0 * _radius) with _radius = 0 */
  return b;
}

void orc_tbrush_destroy(orc_tbrush* b) {
  if (!b) return;
  free(b->dK); free(b->dS); free(b->dV); free(b->sK); free(b->sS); free(b->sV);
  free(b->thick);
  free(b);
}

/* TextureBrush.hxx:31-39 */
void orc_tbrush_set_radius(orc_tbrush* b, double radius) {
  if (!(fabs(b->radius - radius) < 0.5)) {
    b->radius = radius;
    if (b->use_smudge) tb_smudge_reset(b, (int)(2.0 * radius));
  }
}

void orc_tbrush_dip(orc_tbrush* b, const double K[3], const double S[3]) {
  for (int c = 0; c < 3; c++) {
    b->paint[c] = K[c];
    b->paint[3 + c] = S[c];
  }
}

void orc_tbrush_set_thickness_map(orc_tbrush* b, const double* map, int rows, int cols) {
  free(b->thick);
  b->thick = (double*)malloc(sizeof(double) * (size_t)rows * cols);
  memcpy(b->thick, map, sizeof(double) * (size_t)rows * cols);
  b->t_rows = rows;
  b->t_cols = cols;
}

void orc_tbrush_set_thickness_scale(orc_tbrush* b, double s) { b->thickness_scale = s; }
void orc_tbrush_set_use_smudge(orc_tbrush* b, int on) { b->use_smudge = on; }

void orc_tbrush_get_smudge_map(const orc_tbrush* b, int* side, double* K, double* S, double* V) {
  *side = b->map;
  const size_t m = (size_t)b->map * b->map;
  if (K) memcpy(K, b->dK, sizeof(double) * 3 * m);
  if (S) memcpy(S, b->dS, sizeof(double) * 3 * m);
  if (V) memcpy(V, b->dV, sizeof(double) * m);
}

static double normalize_angle(double a) { /* Smudge.hxx:200-210 */
  const double HALF_PI = 0.5 * kPi;
  while (a <= -HALF_PI) a += kPi;
  while (a > HALF_PI) a -= kPi;
  return a;
}

/* Smudge::updateOrientation (Smudge.hxx:164-198) */
static void tb_update_orientation(orc_tbrush* b, const double heading[2]) {
  const double theta = atan2(heading[1], heading[0]);
  const double dtheta = normalize_angle(theta - b->cur_rot);
  b->cur_rot = theta;
  const int n = b->map;
  const size_t m = (size_t)n * n;
  memcpy(b->sK, b->dK, sizeof(double) * 3 * m);
  memcpy(b->sS, b->dS, sizeof(double) * 3 * m);
  memcpy(b->sV, b->dV, sizeof(double) * m);
  const double cx = b->max_size / 2.0, cy = b->max_size / 2.0;
  const double sn = sin(dtheta), cs = cos(dtheta);
  for (int x = 0; x < n; x++) {
    for (int y = 0; y < n; y++) {
      const size_t i = (size_t)y * n + x;
      const double px = x - cx, py = y - cy;
      const double rx = px * cs - py * sn + cx, ry = px * sn + py * cs + cy;
      if (rx < 0 || ry < 0 || rx >= n || ry >= n) {
        for (int c = 0; c < 3; c++) {
          b->dK[3 * i + c] = b->sK[3 * i + c];
          b->dS[3 * i + c] = b->sS[3 * i + c];
        }
        b->dV[i] = b->sV[i];
      } else {
        b->dV[i] = interp_reflect(b->sV, n, n, rx, ry);
        for (int c = 0; c < 3; c++) {
          b->dK[3 * i + c] = interp_reflect_n(b->sK + c, n, n, 3, rx, ry);
          b->dS[3 * i + c] = interp_reflect_n(b->sS + c, n, n, 3, rx, ry);
        }
      }
    }
  }
}

/* Test hooks: load the pickup map, and turn it to a heading. */
void orc_tbrush_set_smudge_map(orc_tbrush* b, const double* K, const double* S,
                               const double* V) {
  const size_t m = (size_t)b->map * b->map;
  memcpy(b->dK, K, sizeof(double) * 3 * m);
  memcpy(b->dS, S, sizeof(double) * 3 * m);
  memcpy(b->dV, V, sizeof(double) * m);
}

void orc_tbrush_turn(orc_tbrush* b, double hx, double hy) {
  const double h[2] = {hx, hy};
  tb_update_orientation(b, h);
}

/* Smudge::smudge (Smudge.hxx:40-158) */
static void tb_smudge(orc_tbrush* b, int rows, int cols, double* K, double* S, double* V,
                      const double bmin[2], const double* verts, int nv, double length,
                      const double* tmap, int trows, int tcols) {
  double maxD = 0.0;
  for (int64_t i = 0; i < (int64_t)trows * tcols; i++) maxD = fmax(tmap[i], maxD);
  if (maxD <= 0.0) return;
  const double pickupRate = 0.1, depositionRate = 0.1, MinVolume = 0.001;
  for (double u = 0.0; u <= 1.0; u += 1. / length) {
    double center[2], t[2];
    spline_cr(verts, nv, u, 0, center);
    spline_cr(verts, nv, u, 1, t);
    const double tn = sqrt(t[0] * t[0] + t[1] * t[1]);
    if (tn > 0.0) { t[0] /= tn; t[1] /= tn; }
    tb_update_orientation(b, t);
    const double radius = b->max_size * 0.5;
    const int roi_x = (int)(center[0] - b->map / 2.0);
    const int roi_y = (int)(center[1] - b->map / 2.0);
    for (int x = 0; x < b->map; x++) {
      for (int y = 0; y < b->map; y++) {
        const int cpx = x + roi_x, cpy = y + roi_y;
        const int tpx = (int)(cpx - bmin[0]), tpy = (int)(cpy - bmin[1]);
        if (cpx < 0 || cpy < 0 || cpx >= cols || cpy >= rows) continue;
        if (tpx < 0 || tpy < 0 || tpx >= tcols || tpy >= trows) continue;
        const double dx = center[0] - (double)cpx, dy = center[1] - (double)cpy;
        if (sqrt(dx * dx + dy * dy) > radius) continue;
        const double Dv = tmap[(int64_t)tpy * tcols + tpx];
        if (Dv <= 0.0) continue;
        const int64_t ic = (int64_t)cpy * cols + cpx;
        const size_t ip = (size_t)y * b->map + x;
        const double cV = V[ic], pV = b->dV[ip];
        const double cVl = cV * depositionRate * Dv / maxD, cVr = cV - cVl;
        const double pVl = pV * pickupRate * Dv / maxD, pVr = pV - pVl;
        double cK[3], cS[3], pK[3], pS[3];
        for (int c = 0; c < 3; c++) {
          cK[c] = K[3 * ic + c];
          cS[c] = S[3 * ic + c];
          pK[c] = b->dK[3 * ip + c];
          pS[c] = b->dS[3 * ip + c];
        }
        const double pVnew = pVr + cVl;
        if (pVnew > MinVolume) {
          const double inv = 1.0 / pVnew;
          for (int c = 0; c < 3; c++) {
            b->dK[3 * ip + c] = inv * (pVr * pK[c] + cVl * cK[c]);
            b->dS[3 * ip + c] = inv * (pVr * pS[c] + cVl * cS[c]);
          }
          b->dV[ip] = fmax(pVnew, 0.0);
        }
        const double cVnew = cVr + pVl;
        if (cVnew > MinVolume) {
          const double inv = 1. / cVnew;
          for (int c = 0; c < 3; c++) {
            K[3 * ic + c] = inv * (cVr * cK[c] + pVl * pK[c]);
            S[3 * ic + c] = inv * (cVr * cS[c] + pVl * pS[c]);
          }
          V[ic] = fmax(cVnew, 0.0);
        }
      }
    }
  }
}

/* TextureBrush::paintStroke (TextureBrush.hxx:52-205).  clk: Canvas::
 * checkDry on every textured pixel (NULL: off). */
void orc_tbrush_paint_stroke(orc_tbrush* b, int rows, int cols, double* K, double* S, double* V,
                             const double* xy, int n, const orc_clock* clk) {
  if (n < 2 || !b->thick) return;
  const double r = b->radius;
  const int nv = n + 2;
  double* v = (double*)malloc(sizeof(double) * 2 * nv);
  {
    double d0[2] = {xy[2] - xy[0], xy[3] - xy[1]};
    double l0 = sqrt(d0[0] * d0[0] + d0[1] * d0[1]);
    double d1[2] = {xy[2 * n - 2] - xy[2 * n - 4], xy[2 * n - 1] - xy[2 * n - 3]};
    double l1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1]);
    if (l0 > 0.0) { d0[0] /= l0; d0[1] /= l0; }
    if (l1 > 0.0) { d1[0] /= l1; d1[1] /= l1; }
    v[0] = xy[0] - d0[0] * r;
    v[1] = xy[1] - d0[1] * r;
    memcpy(v + 2, xy, sizeof(double) * 2 * n);
    v[2 * nv - 2] = xy[2 * n - 2] + d1[0] * r;
    v[2 * nv - 1] = xy[2 * n - 1] + d1[1] * r;
  }
  double bmin[2] = {v[0], v[1]}, bmax[2] = {v[0], v[1]};
  for (int i = 0; i < nv; i++) {
    bmin[0] = fmin(bmin[0], v[2 * i]);
    bmin[1] = fmin(bmin[1], v[2 * i + 1]);
    bmax[0] = fmax(bmax[0], v[2 * i]);
    bmax[1] = fmax(bmax[1], v[2 * i + 1]);
  }
  bmin[0] = fmax(bmin[0] - r, 0.0);
  bmax[0] = fmin(bmax[0] + r, (double)(cols - 1));
  bmin[1] = fmax(bmin[1] - r, 0.0);
  bmax[1] = fmin(bmax[1] + r, (double)(rows - 1));
  double length = 0.0;
  for (int i = 1; i < nv; i++) {
    const double dx = v[2 * i] - v[2 * i - 2], dy = v[2 * i + 1] - v[2 * i - 1];
    length += sqrt(dx * dx + dy * dy);
  }
  /* the frame polygon: the right side backwards, then the left side */
  const int np = 2 * nv;
  double* poly = (double*)malloc(sizeof(double) * 2 * np);
  double* uv = (double*)malloc(sizeof(double) * 2 * np);
  for (int i = 0; i < nv; i++) {
    const double u = (double)i / (double)(nv - 1);
    double c[2], t[2];
    spline_cr(v, nv, u, 0, c);
    spline_cr(v, nv, u, 1, t);
    const double tn = sqrt(t[0] * t[0] + t[1] * t[1]);
    if (tn > 0.0) { t[0] /= tn; t[1] /= tn; }
    const double d[2] = {-t[1], t[0]};
    /* up (left) at nv + i; down (right) reversed at nv - 1 - i */
    poly[2 * (nv + i)] = c[0] - r * d[0];
    poly[2 * (nv + i) + 1] = c[1] - r * d[1];
    uv[2 * (nv + i)] = u;
    uv[2 * (nv + i) + 1] = 0.0;
    poly[2 * (nv - 1 - i)] = c[0] + r * d[0];
    poly[2 * (nv - 1 - i) + 1] = c[1] + r * d[1];
    uv[2 * (nv - 1 - i)] = u;
    uv[2 * (nv - 1 - i) + 1] = 1.0;
  }
  const int trows = (int)(bmax[1] - bmin[1] + 1), tcols = (int)(bmax[0] - bmin[0] + 1);
  double* tmap = (double*)calloc((size_t)(trows > 0 ? trows : 1) * (tcols > 0 ? tcols : 1), sizeof(double));
  for (int x = (int)bmin[0]; x <= (int)bmax[0]; x++) {
    for (int y = (int)bmin[1]; y <= (int)bmax[1]; y++) {
      if (x < 0 || x >= cols || y < 0 || y >= rows) continue;
      double tp[2];
      orc_gbc(poly, np, (double)x, (double)y, uv, tp);
      if (tp[0] < 0.0 || tp[0] > 1.0 || tp[1] < 0.0 || tp[1] > 1.0) continue;
      tp[0] *= b->t_cols;
      tp[1] *= b->t_rows;
      const double vt = b->thickness_scale * interp_reflect(b->thick, b->t_rows, b->t_cols, tp[0], tp[1]);
      if (vt > 0.0) {
        const int s2 = x - (int)bmin[0], t2 = y - (int)bmin[1];
        if (s2 >= 0 && t2 >= 0 && s2 < tcols && t2 < trows) {
          if (clk) check_dry(clk, cols, x, y, K, S, V);
          tmap[(int64_t)t2 * tcols + s2] = vt;
        }
      }
    }
  }
  if (b->use_smudge)
    tb_smudge(b, rows, cols, K, S, V, bmin, v, nv, length, tmap, trows, tcols);
  for (int x = (int)bmin[0]; x <= (int)bmax[0]; x++) {
    for (int y = (int)bmin[1]; y <= (int)bmax[1]; y++) {
      if (x < 0 || x >= cols || y < 0 || y >= rows) continue;
      const int s2 = x - (int)bmin[0], t2 = y - (int)bmin[1];
      if (!(s2 >= 0 && t2 >= 0 && s2 < tcols && t2 < trows)) continue;
      const double vt = tmap[(int64_t)t2 * tcols + s2];
      if (!(vt > 0.0)) continue; /* not in the reference's pixel list */
      const int64_t ic = (int64_t)y * cols + x;
      const double vc = V[ic], vs = vc + vt;
      if (vs > 0.0) {
        const double sc = 1.0 / vs;
        for (int c = 0; c < 3; c++) {
          K[3 * ic + c] = (vc * K[3 * ic + c] + vt * b->paint[c]) * sc;
          S[3 * ic + c] = (vc * S[3 * ic + c] + vt * b->paint[3 + c]) * sc;
        }
        V[ic] = fmax(vt, vc);
      }
    }
  }
  free(tmap);
  free(poly);
  free(uv);
  free(v);
}

/* ------------------------------------------------------------------------ */
/* Footprint ingest (FootprintBrush.hxx:48-58): the PNG's grey values through
 * io::imRead, then ScaledMat.  Decoding the PNG itself is checked against an
 * independent decoder in the tests, not restated here. */

/* ColorConverter::srgb2rgb (core/Color.hxx:189-195) */
double orc_srgb2rgb(double s) {
  if (s <= 0.0404482362771082) return s / 12.92;
  return pow((s + 0.055) / 1.055, 2.4);
}

/* imRead's scale (io/src/ImageIO.cxx:91-104): Mat::convertTo(CV_64FC1, scale)
 * multiplies by scale = 1.0 / 0xff or 1.0 / 0xffff */
void orc_gray_from_u16(int64_t n, const uint16_t* v, int maxval, int srgb, double* out) {
  const double scale = maxval == 65535 ? 1.0 / 0xffff : 1.0 / 0xff;
  for (int64_t i = 0; i < n; i++) {
    const double x = (double)v[i] * scale;
    out[i] = srgb ? orc_srgb2rgb(x) : x;
  }
}

/* OpenCV's interpolateLanczos4: L(t) = sin(pi t) sin(pi t / 4) / (pi^2 t^2 / 4)
 * at t = x + 3 - i, i = 0..7, through one sin/cos pair and the angle-sum
 * table, each rounded to float, then normalised by the float sum. */
void orc_lanczos4_coeffs(float x, float c[8]) {
  static const double s45 = 0.70710678118654752440084436210485;
  static const double cs[8][2] = {{1, 0},   {-s45, -s45}, {0, 1},  {s45, -s45},
                                  {-1, 0},  {s45, s45},   {0, -1}, {-s45, s45}};
  if (x < FLT_EPSILON) {
    for (int i = 0; i < 8; i++) c[i] = 0.0f;
    c[3] = 1.0f;
    return;
  }
  float sum = 0.0f;
  const double pi = 3.14159265358979323846;
  const float xp3 = x + 3.0f;
  const double y0 = -(double)xp3 * pi * 0.25, s0 = sin(y0), c0 = cos(y0);
  for (int i = 0; i < 8; i++) {
    const float xi = x + 3.0f - (float)i;
    const double y = -(double)xi * pi * 0.25;
    c[i] = (float)((cs[i][0] * s0 + cs[i][1] * c0) / (y * y));
    sum += c[i];
  }
  sum = 1.0f / sum;
  for (int i = 0; i < 8; i++) c[i] *= sum;
}

static void lanczos_axis(int ssize, int dsize, int* ofs, float* alpha) {
  const double inv_scale = (double)dsize / ssize;
  const double scale = 1.0 / inv_scale;
  for (int d = 0; d < dsize; d++) {
    float f = (float)((d + 0.5) * scale - 0.5);
    const int s = (int)floorf(f);
    f -= (float)s;
    ofs[d] = s;
    orc_lanczos4_coeffs(f, alpha + 8 * d);
  }
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_resize_lanczos4(const double* src, int rows, int cols, int out_rows, int out_cols,
                         double* dst) {
  int* xofs = (int*)malloc(sizeof(int) * (size_t)out_cols);
  int* yofs = (int*)malloc(sizeof(int) * (size_t)out_rows);
  float* xa = (float*)malloc(sizeof(float) * 8 * (size_t)out_cols);
  float* ya = (float*)malloc(sizeof(float) * 8 * (size_t)out_rows);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)rows * out_cols);
  lanczos_axis(cols, out_cols, xofs, xa);
  lanczos_axis(rows, out_rows, yofs, ya);
  for (int r = 0; r < rows; r++)
    for (int dx = 0; dx < out_cols; dx++) {
      double v = 0.0;
      for (int j = 0; j < 8; j++)
        v += src[(int64_t)r * cols + clampi(xofs[dx] - 3 + j, 0, cols - 1)] * (double)xa[8 * dx + j];
      tmp[(int64_t)r * out_cols + dx] = v;
    }
  for (int dy = 0; dy < out_rows; dy++)
    for (int dx = 0; dx < out_cols; dx++) {
      double v = 0.0;
      for (int k = 0; k < 8; k++)
        v += tmp[(int64_t)clampi(yofs[dy] - 3 + k, 0, rows - 1) * out_cols + dx] *
             (double)ya[8 * dy + k];
      dst[(int64_t)dy * out_cols + dx] = v;
    }
  free(xofs);
  free(yofs);
  free(xa);
  free(ya);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* TextureBrushGpu (renderer/src/TextureBrushGpu.cxx:17-305) and its shaders,
 * in FP32 where the shaders compute (highp float) and FP64 where the host
 * code does.  The canvas planes are FP32 in canvas row order; the warped
 * brush texture is kept in GL row order (row g = canvas row H-1-g). */
struct orc_tbgl {
  double radius;
  float paint[6];
  float* tex; /* brush texture as uploaded: min-max normalised, flipped */
  int t_rows, t_cols;
  double thickness_scale;
  int smudge;
  int n;          /* smudge window side: (int)(2 r) + 1 */
  float* smK;     /* n*n RGBA: K rgb, V */
  float* smS;     /* n*n RGBA: S rgb, V */
  float* warp;    /* rows*cols, GL rows */
  int w_rows, w_cols;
};

static void tbgl_reset_smudge(orc_tbgl* b) {
  free(b->smK);
  free(b->smS);
  b->smK = (float*)calloc(4 * (size_t)b->n * b->n + 4, sizeof(float));
  b->smS = (float*)calloc(4 * (size_t)b->n * b->n + 4, sizeof(float));
}

/* TextureBrushGpu.cxx:48-60 */
void orc_tbgl_set_radius(orc_tbgl* b, double radius) {
  if (!(fabs(b->radius - radius) < 0.5)) {
    b->radius = radius;
    b->n = (int)(unsigned)(2.0 * radius) + 1;
    tbgl_reset_smudge(b);
  }
}

orc_tbgl* orc_tbgl_create(void) {
  orc_tbgl* b = (orc_tbgl*)calloc(1, sizeof(orc_tbgl));
  b->thickness_scale = 1.0;
  b->smudge = 0;          /* TextureBrushGpu.hxx:63 */
  b->n = 11;              /* _smudgeK(Size{10, 10}) then setRadius(5.0) */
  tbgl_reset_smudge(b);
  orc_tbgl_set_radius(b, 5.0);
  return b;
}

void orc_tbgl_destroy(orc_tbgl* b) {
  if (!b) return;
  free(b->smK); free(b->smS); free(b->tex); free(b->warp);
  free(b);
}

void orc_tbgl_dip(orc_tbgl* b, const double K[3], const double S[3]) {
  for (int c = 0; c < 3; c++) {
    b->paint[c] = (float)K[c];
    b->paint[3 + c] = (float)S[c];
  }
}

/* TextureBrushDictionary.cxx:106-123: NORM_MINMAX to [0, 1], flipped, FP32 */
void orc_tbgl_set_texture(orc_tbgl* b, const double* img, int rows, int cols) {
  free(b->tex);
  b->tex = (float*)malloc(sizeof(float) * (size_t)rows * cols);
  double lo = img[0], hi = img[0];
  for (int64_t i = 0; i < (int64_t)rows * cols; i++) {
    lo = fmin(lo, img[i]);
    hi = fmax(hi, img[i]);
  }
  const double sc = hi > lo ? 1.0 / (hi - lo) : 0.0;
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++)
      b->tex[(int64_t)(rows - 1 - r) * cols + c] = (float)((img[(int64_t)r * cols + c] - lo) * sc);
  b->t_rows = rows;
  b->t_cols = cols;
}

void orc_tbgl_set_thickness_scale(orc_tbgl* b, double s) { b->thickness_scale = s; }
void orc_tbgl_enable_smudge(orc_tbgl* b, int on) { b->smudge = on; }

void orc_tbgl_get_smudge(const orc_tbgl* b, int* side, float* K4, float* S4) {
  *side = b->n;
  if (K4) memcpy(K4, b->smK, sizeof(float) * 4 * (size_t)b->n * b->n);
  if (S4) memcpy(S4, b->smS, sizeof(float) * 4 * (size_t)b->n * b->n);
}

/* GL_LINEAR + GL_CLAMP_TO_EDGE at texture coordinates (s, t) */
static double tbgl_sample(const orc_tbgl* b, double s, double t) {
  const double x = s * b->t_cols - 0.5, y = t * b->t_rows - 0.5;
  const double fx0 = floor(x), fy0 = floor(y);
  const double ax = x - fx0, ay = y - fy0;
  const int x0 = (int)fx0, y0 = (int)fy0;
  const int xa = clampi(x0, 0, b->t_cols - 1), xb = clampi(x0 + 1, 0, b->t_cols - 1);
  const int ya = clampi(y0, 0, b->t_rows - 1), yb = clampi(y0 + 1, 0, b->t_rows - 1);
#define TX(yy, xx) ((double)b->tex[(int64_t)(yy) * b->t_cols + (xx)])
  return (TX(ya, xa) * (1.0 - ax) + TX(ya, xb) * ax) * (1.0 - ay) +
         (TX(yb, xa) * (1.0 - ax) + TX(yb, xb) * ax) * ay;
#undef TX
}

static int64_t tbgl_cross(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t px, int64_t py) {
  return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}

static int tbgl_top_left(int64_t ax, int64_t ay, int64_t bx, int64_t by) {
  const int64_t dx = bx - ax, dy = by - ay;
  return (dy == 0 && dx < 0) || dy < 0;
}

/* generateWarpedTexture (TextureBrushGpu.cxx:168-247) + BrushTextureWarp:
 * the strip r0 l0 r1 l1 ... drawn into the canvas-sized warp image */
static void tbgl_warp(orc_tbgl* b, int rows, int cols, const double* xy, int n) {
  const int nvb = 2 * n;
  float* vx = (float*)malloc(sizeof(float) * nvb);
  float* vy = (float*)malloc(sizeof(float) * nvb);
  float* tu = (float*)malloc(sizeof(float) * nvb);
  float* tv = (float*)malloc(sizeof(float) * nvb);
  for (int i = 0; i < n; i++) {
    const double u = (double)i / (double)(n - 1);
    double c[2], t[2];
    spline_cr(xy, n, u, 0, c);
    spline_cr(xy, n, u, 1, t);
    const double tn = sqrt(t[0] * t[0] + t[1] * t[1]);
    if (tn > 0.0) { t[0] /= tn; t[1] /= tn; }
    const double d[2] = {-t[1], t[0]};
    vx[2 * i] = (float)(c[0] + b->radius * d[0]);  /* r: texture v = 0 */
    vy[2 * i] = (float)(c[1] + b->radius * d[1]);
    vx[2 * i + 1] = (float)(c[0] - b->radius * d[0]);  /* l: v = 1 */
    vy[2 * i + 1] = (float)(c[1] - b->radius * d[1]);
    tu[2 * i] = tu[2 * i + 1] = (float)u;
    tv[2 * i] = 0.0f;
    tv[2 * i + 1] = 1.0f;
  }
  for (int j = 0; j + 2 < nvb; j++) {
    int id[3] = {j, j + 1, j + 2};
    int64_t X[3], Y[3];
    for (int q = 0; q < 3; q++) {  /* window coordinates (y up), 1/256 px */
      X[q] = (int64_t)llrint((double)vx[id[q]] * 256.0);
      Y[q] = (int64_t)llrint((double)((float)rows - vy[id[q]]) * 256.0);
    }
    int64_t area = tbgl_cross(X[0], Y[0], X[1], Y[1], X[2], Y[2]);
    if (area == 0) continue;
    if (area < 0) {  /* counter-clockwise in y-up coordinates */
      int64_t tx = X[1], ty = Y[1];
      X[1] = X[2]; Y[1] = Y[2]; X[2] = tx; Y[2] = ty;
      int ti = id[1]; id[1] = id[2]; id[2] = ti;
      area = -area;
    }
    int64_t xmin = X[0], xmax = X[0], ymin = Y[0], ymax = Y[0];
    for (int q = 1; q < 3; q++) {
      xmin = X[q] < xmin ? X[q] : xmin; xmax = X[q] > xmax ? X[q] : xmax;
      ymin = Y[q] < ymin ? Y[q] : ymin; ymax = Y[q] > ymax ? Y[q] : ymax;
    }
    int px0 = (int)((xmin - 128) >> 8), px1 = (int)((xmax - 128) >> 8) + 1;
    int py0 = (int)((ymin - 128) >> 8), py1 = (int)((ymax - 128) >> 8) + 1;
    px0 = px0 < 0 ? 0 : px0; py0 = py0 < 0 ? 0 : py0;
    px1 = px1 > cols - 1 ? cols - 1 : px1; py1 = py1 > rows - 1 ? rows - 1 : py1;
    for (int py = py0; py <= py1; py++)
      for (int px = px0; px <= px1; px++) {
        const int64_t cx = (int64_t)px * 256 + 128, cy = (int64_t)py * 256 + 128;
        const int64_t e0 = tbgl_cross(X[1], Y[1], X[2], Y[2], cx, cy);
        const int64_t e1 = tbgl_cross(X[2], Y[2], X[0], Y[0], cx, cy);
        const int64_t e2 = tbgl_cross(X[0], Y[0], X[1], Y[1], cx, cy);
        const int in0 = e0 > 0 || (e0 == 0 && tbgl_top_left(X[1], Y[1], X[2], Y[2]));
        const int in1 = e1 > 0 || (e1 == 0 && tbgl_top_left(X[2], Y[2], X[0], Y[0]));
        const int in2 = e2 > 0 || (e2 == 0 && tbgl_top_left(X[0], Y[0], X[1], Y[1]));
        if (!(in0 && in1 && in2)) continue;
        const double w0 = (double)e0 / (double)area, w1 = (double)e1 / (double)area,
                     w2 = (double)e2 / (double)area;
        const double s = w0 * tu[id[0]] + w1 * tu[id[1]] + w2 * tu[id[2]];
        const double t = w0 * tv[id[0]] + w1 * tv[id[1]] + w2 * tv[id[2]];
        b->warp[(int64_t)py * cols + px] = (float)tbgl_sample(b, (double)(float)s, (double)(float)t);
      }
  }
  free(vx); free(vy); free(tu); free(tv);
}

static int tbgl_round16(int v) { return (v + 15) / 16 * 16; }

/* TextureBrushGpu::smudge (TextureBrushGpu.cxx:253-305) + Smudge.compute */
static void tbgl_smudge(orc_tbgl* b, int rows, int cols, float* K, float* S, float* V,
                        float* light, const double* xy, int n) {
  double length = 0.0;
  for (int i = 1; i < n; i++) {
    const double dx = xy[2 * i] - xy[2 * i - 2], dy = xy[2 * i + 1] - xy[2 * i - 1];
    length += sqrt(dx * dx + dy * dy);
  }
  const int m = b->n;
  const int dw = tbgl_round16(m), dh = tbgl_round16(m);
  const float rcx = (float)(m * 0.5), rcy = (float)(m * 0.5);
  const float scale = (float)b->thickness_scale;
  const size_t cells = (size_t)m * m;
  float* oK = (float*)malloc(sizeof(float) * 4 * cells);  /* state before the dispatch */
  float* oS = (float*)malloc(sizeof(float) * 4 * cells);
  int* owner = (int*)malloc(sizeof(int) * cells);
  float* nK = (float*)malloc(sizeof(float) * 4 * cells);
  float* nS = (float*)malloc(sizeof(float) * 4 * cells);
  int lastx = -1, lasty = -1;
  for (double u = 0.0; u <= 1.0; u += (0.25 / length)) {
    double c[2], t[2];
    spline_cr(xy, n, u, 0, c);
    if ((int)c[0] == lastx && (int)c[1] == lasty) continue;
    lastx = (int)c[0];
    lasty = (int)c[1];
    spline_cr(xy, n, u, 1, t);
    const double tn = sqrt(t[0] * t[0] + t[1] * t[1]);
    if (tn > 0.0) { t[0] /= tn; t[1] /= tn; }
    const double theta = atan2(t[1], t[0]);
    const int tlx = (int)((double)(int)c[0] - (double)m * 0.5);
    const int tly = (int)((double)(int)c[1] - (double)m * 0.5);
    const float th = (float)theta, cs = cosf(th), sn = sinf(th);
    memcpy(oK, b->smK, sizeof(float) * 4 * cells);
    memcpy(oS, b->smS, sizeof(float) * 4 * cells);
    for (size_t i = 0; i < cells; i++) owner[i] = -1;
    for (int gy = 0; gy < dh; gy++)
      for (int gx = 0; gx < dw; gx++) {
        const int px = tlx + gx, pyg = tly + gy; /* GL coordinates */
        if (px < 0 || pyg < 0 || px >= cols || pyg >= rows) continue;
        const float vTex = scale * b->warp[(int64_t)pyg * cols + px];
        if (vTex < 10e-6f) continue;
        const float ptx = (float)gx - rcx, pty = (float)gy - rcy;
        const float rx = (ptx * cs - pty * sn) + rcx, ry = (ptx * sn + pty * cs) + rcy;
        const int qx = (int)rx, qy = (int)ry;
        if (qx < 0 || qy < 0 || qx >= cols || qy >= rows) continue;
        const int64_t ic = (int64_t)(rows - 1 - pyg) * cols + px;  /* canvas row */
        const int inmap = qx < m && qy < m;
        const size_t ip = inmap ? (size_t)qy * m + qx : 0;
        float pK[3] = {0, 0, 0}, pS[3] = {0, 0, 0}, pV = 0.0f;
        if (inmap) {
          for (int ch = 0; ch < 3; ch++) { pK[ch] = oK[4 * ip + ch]; pS[ch] = oS[4 * ip + ch]; }
          pV = oK[4 * ip + 3];
        }
        float cK[3], cS[3];
        for (int ch = 0; ch < 3; ch++) { cK[ch] = K[3 * ic + ch]; cS[ch] = S[3 * ic + ch]; }
        const float cV = V[ic];
        const float cVl = cV * 0.7f * vTex, cVr = cV - cVl;
        const float pVl = pV * 0.7f * vTex, pVr = pV - pVl;
        const float pVnew = pVr + cVl;
        if (pVnew > 0.0f && inmap) {
          const int inv = gx + gy * dw;
          if (inv > owner[ip]) {
            owner[ip] = inv;
            for (int ch = 0; ch < 3; ch++) {
              nK[4 * ip + ch] = (pVr * pK[ch] + cVl * cK[ch]) / pVnew;
              nS[4 * ip + ch] = (pVr * pS[ch] + cVl * cS[ch]) / pVnew;
            }
            nK[4 * ip + 3] = fmaxf(pVnew, 0.0f);
            nS[4 * ip + 3] = fmaxf(pVnew, 0.0f);
          }
        }
        const float cVnew = cVr + pVl;
        if (cVnew > 0.0f) {
          for (int ch = 0; ch < 3; ch++) {
            K[3 * ic + ch] = (cVr * pK[ch] + pVl * cK[ch]) / cVnew;
            S[3 * ic + ch] = (cVr * pS[ch] + pVl * cS[ch]) / cVnew;
          }
          V[ic] = fmaxf(cVnew, 0.0f);
          light[ic] = vTex;
        }
      }
    for (size_t i = 0; i < cells; i++)
      if (owner[i] >= 0) {
        memcpy(b->smK + 4 * i, nK + 4 * i, sizeof(float) * 4);
        memcpy(b->smS + 4 * i, nS + 4 * i, sizeof(float) * 4);
      }
  }
  free(oK); free(oS); free(owner); free(nK); free(nS);
}

/* TextureBrushGpu::paintStroke(vertices, CanvasGpu&) (TextureBrushGpu.cxx:73-166) */
void orc_tbgl_paint_stroke(orc_tbgl* b, int rows, int cols, float* K, float* S, float* V,
                           float* light, const double* xy, int n) {
  if (n < 2 || !b->tex) return;
  if (!b->warp || b->w_rows != rows || b->w_cols != cols) {
    free(b->warp);
    b->warp = (float*)calloc((size_t)rows * cols, sizeof(float));
    b->w_rows = rows;
    b->w_cols = cols;
  }
  double bmin[2] = {xy[0], xy[1]}, bmax[2] = {xy[0], xy[1]};
  for (int i = 0; i < n; i++) {
    bmin[0] = fmin(bmin[0], xy[2 * i]); bmin[1] = fmin(bmin[1], xy[2 * i + 1]);
    bmax[0] = fmax(bmax[0], xy[2 * i]); bmax[1] = fmax(bmax[1], xy[2 * i + 1]);
  }
  bmin[0] = fmax(bmin[0] - b->radius, 0.0);
  bmax[0] = fmin(bmax[0] + b->radius, (double)(cols - 1));
  bmin[1] = fmax(bmin[1] - b->radius, 0.0);
  bmax[1] = fmin(bmax[1] + b->radius, (double)(rows - 1));
  tbgl_warp(b, rows, cols, xy, n);
  const int aw = (int)(bmax[0] - bmin[0]) + 1, ah = (int)(bmax[1] - bmin[1]) + 1;
  const int ox = (int)bmin[0], oy = (int)bmin[1];
  if (b->smudge) tbgl_smudge(b, rows, cols, K, S, V, light, xy, n);
  const float Kb[3] = {b->paint[0], b->paint[1], b->paint[2]};
  const float Sb[3] = {b->paint[3], b->paint[4], b->paint[5]};
  const float scale = (float)b->thickness_scale;
  const int gw = tbgl_round16(aw), gh = tbgl_round16(ah);
  for (int gy = 0; gy < gh; gy++)  /* PaintTextureFootprintOnCanvas */
    for (int gx = 0; gx < gw; gx++) {
      const int px = ox + gx, pyg = rows - (oy + gy) - 1;
      if (px < 0 || pyg < 0 || px >= cols || pyg >= rows) continue;
      const int64_t ic = (int64_t)(rows - 1 - pyg) * cols + px;
      const float vCan = V[ic], vTex = scale * b->warp[(int64_t)pyg * cols + px];
      const float vSum = vCan + vTex;
      if (vSum < 10e-6f) continue;
      for (int ch = 0; ch < 3; ch++) {
        K[3 * ic + ch] = (vCan * K[3 * ic + ch] + vTex * Kb[ch]) / vSum;
        S[3 * ic + ch] = (vCan * S[3 * ic + ch] + vTex * Sb[ch]) / vSum;
      }
      V[ic] = vSum;
      light[ic] = (vTex > 0.0f) ? vTex : vCan;
    }
  for (int gy = 0; gy < gh; gy++)  /* ClearTextureFootprint */
    for (int gx = 0; gx < gw; gx++) {
      const int px = ox + gx, pyg = rows - (oy + gy) - 1;
      if (px < 0 || pyg < 0 || px >= cols || pyg >= rows) continue;
      b->warp[(int64_t)pyg * cols + px] = 0.0f;
    }
}
