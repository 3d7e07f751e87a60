/*
 * painty_oracle.h -- FP64 CPU restatement of lindemeier/painty's paint
 * rendering hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * that the HIP implementation is compared against (and the timed CPU
 * baseline).  Nothing in the product (painty_amd/, include/) links or calls it.
 *
 * Layout conventions follow the reference's cv::Mat_ storage: row-major,
 * continuous; K, S and R0 are AoS triples (3 doubles per pixel, like
 * Mat<vec3>), V and h are one double per pixel.
 *
 * Every function cites the reference file:line it restates.  Where the
 * reference has undefined behaviour the documented intent is implemented
 * instead (noted at the function).
 */
#ifndef PAINTY_ORACLE_H
#define PAINTY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* painty/core/Math.hxx:157-169 */
double orc_coth(double x);
/* painty/core/Math.hxx:181-191 */
double orc_acoth(double x);

/* painty/core/KubelkaMunk.hxx:25-83 (N = 3, Float = double) */
void orc_compute_reflectance(const double K[3], const double S_in[3],
                             const double R0[3], double d, double R[3]);

/* painty/core/KubelkaMunk.hxx:89-124; returns -1 where the reference throws
 * std::invalid_argument, 0 otherwise. */
int orc_compute_scattering_absorption(const double Rb[3], const double Rw[3],
                                      double K[3], double S[3]);

/* Same formula as the GLSL compute shaders use
 * (painty/renderer/shaders/DryingStep.compute.glsl:31-49): no S clamp,
 * returns r0 when any |b S d| < 1e-5. */
void orc_compute_reflectance_glsl(const double K[3], const double S[3],
                                  const double R0[3], double d, double R[3]);

/* painty/renderer/Renderer.hxx:26-41 (out of place) */
void orc_compose(int64_t n, const double* K, const double* S, const double* V,
                 const double* R0, double* R);

/* painty/renderer/PaintLayer.hxx:81-96 (in place onto R0) */
void orc_compose_onto(int64_t n, const double* K, const double* S,
                      const double* V, double* R0);

/* painty/renderer/Canvas.hxx:105-121 */
void orc_dry_canvas(int64_t n, double* K, double* S, double* V, double* R0,
                    double* h);

/* painty/renderer/shaders/DryingStep.compute.glsl:56-77 (partial dry) */
void orc_dry_step(int64_t n, const double* K, const double* S, double* V,
                  double* R0, double portion);

/* ---- imprint ("footprint") brush: painty/renderer/FootprintBrush.hxx ---- */
typedef struct orc_brush orc_brush;

/* FootprintBrush(radius) + setRadius (FootprintBrush.hxx:29-63) with the
 * footprint supplied directly as a width x width core (width =
 * 2*ceil(radius)+1) instead of loading ./data/footprint/footprint.png.
 * The core is padded with zeros to exactly sizeMap x sizeMap (the reference
 * pads to width+2*pad, one short when sizeMap-width is odd, and then reads
 * row/col sizeMap-1 out of bounds; here that row/col reads 0). */
orc_brush* orc_brush_create(double radius, const double* core, int width);
void orc_brush_destroy(orc_brush* b);
/* setRadius (FootprintBrush.hxx:46-63): when |radius - current| >= 0.5, a new
 * footprint and a new, cleared pickup map; the snapshot buffer is kept.
 * Returns -1 when width != 2*ceil(radius)+1. */
int orc_brush_set_radius(orc_brush* b, double radius, const double* core,
                         int width);
int orc_brush_size_map(const orc_brush* b);
const double* orc_brush_footprint(const orc_brush* b); /* sizeMap^2 */
void orc_brush_set_rates(orc_brush* b, double pickup, double deposition);
void orc_brush_set_use_snapshot(orc_brush* b, int use);
/* FootprintBrush.hxx:150-154 */
void orc_brush_dip(orc_brush* b, const double K[3], const double S[3]);
/* FootprintBrush.hxx:160-166 */
void orc_brush_clean(orc_brush* b);
/* FootprintBrush.hxx:168-172 (full canvas -> snapshot copy) */
void orc_brush_update_snapshot(orc_brush* b, int rows, int cols,
                               const double* K, const double* S,
                               const double* V);
/* FootprintBrush.hxx:174-176 (pickup map, sizeMap^2 AoS / scalar) */
void orc_brush_get_pickup(const orc_brush* b, double* K, double* S, double* V);
/* FootprintBrush.hxx:73-143 with updateSnapshot(canvas, center) :278-319,
 * pickupPaint :349-384, depositPaint :393-431, blend :331-340.
 * Canvas::checkDry (:115) is a no-op on K/S/V for drying time 0 and for any
 * run shorter than the default drying time (Canvas.hxx:32,126). */
void orc_brush_imprint(orc_brush* b, int rows, int cols, double* K, double* S,
                       double* V, double cx, double cy, double theta);

/* The same imprint with Canvas::checkDry (Canvas.hxx:123-155) run on every
 * visited pixel before its pickup and deposit, as FootprintBrush.hxx:115
 * does.  The clock: tmap (rows*cols time points in microseconds), the
 * imprint's time point now_us, the drying time in milliseconds, and the
 * canvas R0 (AoS) and h the drying writes. */
typedef struct {
  int64_t* tmap;
  int64_t now_us;
  int64_t dry_ms;
  double* R0;
  double* h;
} orc_clock;
void orc_brush_imprint_timed(orc_brush* b, int rows, int cols, double* K,
                             double* S, double* V, double cx, double cy,
                             double theta, const orc_clock* clk);

/* A stamp list driver (the GUI calls dip / imprint / dryCanvas in sequence,
 * apps/painty_gui/DigitalCanvas.cxx:49-123). */
typedef struct {
  double cx, cy, theta;
  double K[3], S[3]; /* paint for ORC_DIP */
  uint32_t flags;
} orc_stamp;
enum { ORC_DIP = 1u, ORC_DRY_AFTER = 2u };
void orc_run_stamps(orc_brush* b, int rows, int cols, double* K, double* S,
                    double* V, double* R0, double* h, const orc_stamp* stamps,
                    int64_t n);

/* ---- gradient of composite + stamping (FP64 reverse mode) ----
 * Forward: stamps as orc_run_stamps (no ORC_DRY_AFTER; stamp 0 must dip),
 * each dip's paint = orc_mix_single(palette, W[paint_of[i]]), then
 * R = Renderer::compose of the final wet layer over R0 (Renderer.hxx:26-41).
 * Backward: gK, gS, gV (initial wet layer), gR0 and gW (n_paint x n_pig)
 * of the loss whose gradient w.r.t. R is gR.  K, S, V are updated in place
 * to the final state.  The brush must be fresh (no snapshot yet).  Returns
 * -1 for an unsupported stamp list. */
int orc_grad_run(orc_brush* b, int rows, int cols, double* K, double* S,
                 double* V, const double* R0, const orc_stamp* st, int64_t n,
                 const int* paint_of, int n_pig, const double* palK,
                 const double* palS, int n_paint, const double* W,
                 const double* gR, double* R, double* gK, double* gS,
                 double* gV, double* gR0, double* gW);
/* Adjoint of ComputeReflectance (KubelkaMunk.hxx:25-83) for one pixel. */
void orc_compute_reflectance_adj(const double K[3], const double S_in[3],
                                 const double R0[3], double d,
                                 const double gR[3], double gK[3],
                                 double gS[3], double* gd, double gR0[3]);

/* ---- mixer: painty/mixer/src/PaintMixer.cxx ---- */
/* :327-355; returns -1 (reference: std::invalid_argument) on size mismatch */
int orc_mix_single(int n_palette, const double* baseK, const double* baseS,
                   int n_weights, const double* w, double K[3], double S[3]);
/* :536-548 */
void orc_mixed(const double K1[3], const double S1[3], double v1,
               const double K2[3], const double S2[3], double v2, double K[3],
               double S[3]);

/* ---- stroke placement: painty/core/Spline.hxx:27-73 and
 * FootprintBrush::paintStroke (FootprintBrush.hxx:251-267) ---- */
double orc_catmull_rom(double p_1, double p0, double p1, double p2, double t);
double orc_catmull_rom_d1(double p_1, double p0, double p1, double p2,
                          double t);
/* Writes up to max_out (cx, cy, theta) triples; returns the count the stroke
 * produces (may exceed max_out).  The reference reads path[ULONG_MAX] at
 * i = 0 (:252, max(i-1UL, 0UL) wraps); the intent, p_pre = path[0], is used. */
int64_t orc_stroke_stamps(const double* path_xy, int64_t n_points,
                          double* out, int64_t max_out);

/* ---- display (Renderer::render, Renderer.hxx:60-156; imSave's pixel
 * transform, io/src/ImageIO.cxx:158-183) ----
 * orc_render: K, S, R0 AoS (rows*cols*3), V (rows*cols) -> out AoS, clamped
 * to [0, 1].  orc_render_pixel_normal: the surface normal from the four
 * reflect-bordered neighbours of a pixel's V (Renderer.hxx:104-121). */
void orc_render(int rows, int cols, const double* K, const double* S,
                const double* V, const double* R0, double* out);
void orc_render_pixel_normal(double s01, double s21, double s10, double s12,
                             double n[3]);
/* ColorConverter::rgb2srgb (core/Color.hxx:197-205) */
double orc_rgb2srgb(double l);
/* optional rgb2srgb, scale by 2^bits - 1 (bits 8 or 16), round half to even,
 * saturate; interleaved RGB out */
void orc_encode(int64_t n, const double* rgb, int to_srgb, int bits,
                uint16_t* out);

/* ---- TextureBrush (renderer/TextureBrush.hxx:19-238) with Smudge
 * (renderer/Smudge.hxx:22-229) ----
 * A brush stroke texture (thickness map) warped along a stroke polyline by
 * generalized barycentric coordinates (core/Math.hxx:73-133), an optional
 * smudge pass, then volume-weighted mixing with V = max(Vtex, Vcanvas).
 * K, S of the canvas are AoS (rows*cols*3), V rows*cols; xy holds n points.
 * orc_gbc: the interpolation of 2-d values at a polygon's np vertices. */
typedef struct orc_tbrush orc_tbrush;
orc_tbrush* orc_tbrush_create(void);
void orc_tbrush_destroy(orc_tbrush* b);
void orc_tbrush_set_radius(orc_tbrush* b, double radius);
void orc_tbrush_dip(orc_tbrush* b, const double K[3], const double S[3]);
void orc_tbrush_set_thickness_map(orc_tbrush* b, const double* map, int rows, int cols);
void orc_tbrush_set_thickness_scale(orc_tbrush* b, double s);
void orc_tbrush_set_use_smudge(orc_tbrush* b, int on);
void orc_tbrush_get_smudge_map(const orc_tbrush* b, int* side, double* K, double* S,
                               double* V);
void orc_tbrush_paint_stroke(orc_tbrush* b, int rows, int cols, double* K, double* S,
                             double* V, const double* xy, int n, const orc_clock* clk);
void orc_gbc(const double* poly, int np, double px, double py, const double* vals,
             double out[2]);
/* Test hooks: overwrite the smudge's pickup map (side x side cells), and
 * Smudge::updateOrientation to heading (hx, hy) (Smudge.hxx:164-198). */
void orc_tbrush_set_smudge_map(orc_tbrush* b, const double* K, const double* S,
                               const double* V);
void orc_tbrush_turn(orc_tbrush* b, double hx, double hy);

/* ---- Footprint ingest: io::imRead's pixel transform and ScaledMat ----
 * orc_srgb2rgb: ColorConverter::srgb2rgb (core/Color.hxx:189-195).
 * orc_gray_from_u16: imRead(.., Mat<double>&, convertFrom_sRGB)
 * (io/src/ImageIO.cxx:78-113): v * (1/maxval) in FP64 (Mat::convertTo),
 * then srgb2rgb when asked; maxval 255 or 65535.
 * orc_lanczos4_coeffs / orc_resize_lanczos4: ScaledMat (image/Mat.hxx:
 * 141-148) = cv::resize(INTER_LANCZOS4) of an FP64 single-channel image,
 * restating OpenCV 4.x's published resize (imgproc/src/resize.cpp:
 * interpolateLanczos4, resizeGeneric_ with HResizeLanczos4 /
 * VResizeLanczos4<double, double, float>): source position
 * fx = (float)((dx + 0.5) * scale - 0.5) with scale = 1 / (dst / src),
 * 8 taps from floor(fx) - 3, float coefficients, replicated borders, FP64
 * sums in tap order. */
double orc_srgb2rgb(double s);
void orc_gray_from_u16(int64_t n, const uint16_t* v, int maxval, int srgb, double* out);
void orc_lanczos4_coeffs(float x, float c[8]);
void orc_resize_lanczos4(const double* src, int rows, int cols, int out_rows, int out_cols,
                         double* dst);

/* ---- TextureBrushGpu (renderer/src/TextureBrushGpu.cxx:17-305) with its
 * shaders BrushTextureWarp.vert/.frag, Smudge.compute,
 * PaintTextureFootprintOnCanvas.compute and ClearTextureFootprint.compute.
 * The GL renderer keeps its images y-flipped (gpu/GpuMat.hxx:43-56,
 * 107-132): GL row g is canvas row H-1-g.  Where GL leaves behaviour to the
 * implementation this restatement fixes it (and the product follows):
 * vertices snapped to 1/256 pixel, a pixel covered when its centre is inside
 * a triangle with the top-left rule on edges (window y up), texture
 * coordinates interpolated linearly at the centre, the brush texture
 * sampled with GL_LINEAR and GL_CLAMP_TO_EDGE, later triangles of the strip
 * overwriting earlier ones, and the smudge shader's writes to one pickup
 * cell resolved as: every invocation reads the state before the dispatch,
 * the highest invocation index (x + y * width of the dispatch) wins.
 * light: the canvas S.a plane (the GL renderer's lighting thickness). */
typedef struct orc_tbgl orc_tbgl;
orc_tbgl* orc_tbgl_create(void);
void orc_tbgl_destroy(orc_tbgl* b);
void orc_tbgl_set_radius(orc_tbgl* b, double radius);
void orc_tbgl_dip(orc_tbgl* b, const double K[3], const double S[3]);
void orc_tbgl_set_texture(orc_tbgl* b, const double* img, int rows, int cols);
void orc_tbgl_set_thickness_scale(orc_tbgl* b, double s);
void orc_tbgl_enable_smudge(orc_tbgl* b, int on);
void orc_tbgl_get_smudge(const orc_tbgl* b, int* side, float* K4, float* S4);
void orc_tbgl_paint_stroke(orc_tbgl* b, int rows, int cols, float* K, float* S, float* V,
                           float* light, const double* xy, int n);

#ifdef __cplusplus
}
#endif
#endif
