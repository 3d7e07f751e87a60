/*
 * painty_hip.h -- C ABI of the gfx950 (MI355X) implementation of painty's
 * paint-rendering hot path: Kubelka-Munk compositing of wet and dry paint
 * layers over the canvas substrate, and imprint-brush stamping with
 * volume-weighted K/S mixing, pickup and drying.
 *
 * The reference (lindemeier/painty) exposes this path as header-only C++
 * templates instantiated with vec3 = Eigen::Matrix<double,3,1>; it has no
 * C/FFI layer.  Each entry point below replaces one member or free function
 * of those templates; the citation gives the reference file:line.  The C++
 * classes in painty_amd/include/painty/ (Canvas, PaintLayer, FootprintBrush,
 * TextureBrush, Renderer, PaintMixer) are the drop-in binding over this ABI.
 *
 * Conventions
 *   - Every function returns PNT_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is pnt_last_error().
 *     PNT_EINVAL is where the reference throws std::invalid_argument.
 *   - Handles are opaque, own their device memory, and are not thread-safe
 *     (as the reference classes are not).  Work on one context is ordered on
 *     that context's HIP stream.
 *   - Host buffers belong to the caller and are only read/written during the
 *     call.  Uploads are synchronous, downloads synchronise the stream.
 *   - Host images are row-major.  layout PNT_AOS: K, S, R0 interleaved as
 *     3 floats per pixel (the memory layout of the reference's Mat<vec3>);
 *     PNT_PLANAR: three full planes one after another.  V and h are one float
 *     per pixel.
 *   - Device state is FP32 structure-of-arrays: a paint layer is 7 planes
 *     (K0 K1 K2 S0 S1 S2 V), a canvas adds R0 (3 planes) and h (1 plane).
 *   - Numeric domain: K >= 0, S >= 0, V >= 0 (the reference's own domain);
 *     S within 1e4*DBL_EPSILON of 0 is clamped to 1e-11 as in the reference.
 */
#ifndef PAINTY_HIP_H
#define PAINTY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNT_OK 0
#define PNT_EINVAL (-1)  /* invalid argument (reference: std::invalid_argument) */
#define PNT_EHIP (-2)    /* HIP runtime failure */
#define PNT_ENOMEM (-3)  /* device allocation failed / over the memory budget */
#define PNT_ESTATE (-4)  /* call not valid in the object's current state */
#define PNT_EKERNEL (-5) /* a kernel reported an error (bounded wait expired) */

#define PNT_AOS 0
#define PNT_PLANAR 1

typedef struct pnt_ctx_s* pnt_ctx;
typedef struct pnt_layer_s* pnt_layer;
typedef struct pnt_canvas_s* pnt_canvas;
typedef struct pnt_brush_s* pnt_brush;

/* ------------------------------------------------------------ library --- */
/* Message of the last failed call on this thread ("" if none). */
const char* pnt_last_error(void);
/* ABI version: major*10000 + minor*100 + patch. */
int pnt_version(void);

/* ------------------------------------------------------------ context --- */
/* A device and one HIP stream (default priority) on it. */
int pnt_ctx_create(int device, pnt_ctx* out);
int pnt_ctx_destroy(pnt_ctx ctx);
int pnt_ctx_synchronize(pnt_ctx ctx);
/* The context's hipStream_t, for callers that order their own work with it. */
void* pnt_ctx_stream(pnt_ctx ctx);

/* ------------------------------------------- PaintLayer<vec3> ----------- */
/* PaintLayer(rows, cols) -- renderer/PaintLayer.hxx:27-30 (zero-filled). */
int pnt_layer_create(pnt_ctx ctx, int rows, int cols, pnt_layer* out);
int pnt_layer_destroy(pnt_layer layer);
/* getRows / getCols -- PaintLayer.hxx:56-62 */
int pnt_layer_shape(pnt_layer layer, int* rows, int* cols);
/* clear() -- PaintLayer.hxx:68-74 */
int pnt_layer_clear(pnt_layer layer);
/* Write / read the K, S, V buffers (getK_buffer/getS_buffer/getV_buffer,
 * PaintLayer.hxx:32-54).  Any pointer may be NULL to skip that buffer.  The
 * device holds K and S in FP32 and V in FP64 (as the reference's Mat<double>
 * V): FP32 V is widened exactly on upload and rounded on download. */
int pnt_layer_upload(pnt_layer layer, const float* K, const float* S,
                     const float* V, int layout);
int pnt_layer_download(pnt_layer layer, float* K, float* S, float* V,
                       int layout);
/* The V buffer in FP64, exactly as the device holds it (rows*cols doubles). */
int pnt_layer_upload_volume(pnt_layer layer, const double* V);
int pnt_layer_download_volume(pnt_layer layer, double* V);
/* copyTo(other) -- PaintLayer.hxx:103-111 (reallocates on size mismatch) */
int pnt_layer_copy_to(pnt_layer src, pnt_layer dst);
/* set(i, j, k, s, v) -- PaintLayer.hxx:122-127 */
int pnt_layer_set(pnt_layer layer, int i, int j, const float k[3],
                  const float s[3], float v);
/* composeOnto(R0) -- PaintLayer.hxx:81-96, with R0 a host image (rows x cols
 * x 3 in `layout`), updated in place.  When rows/cols differ from the
 * layer's, R0 must have room for the layer's size and is first filled with
 * 1.0 (PaintLayer.hxx:82-87); pass r0_rows/r0_cols = the current size. */
int pnt_layer_compose_onto_host(pnt_layer layer, float* R0, int r0_rows,
                                int r0_cols, int layout);
/* composeOnto(canvas.getR0()) -- the dry layer baked over the substrate. */
int pnt_layer_compose_onto_canvas(pnt_layer layer, pnt_canvas canvas);

/* --------------------------------------------------- Canvas<vec3> ------- */
/* Canvas(rows, cols) -- renderer/Canvas.hxx:26-35: background (1,1,1),
 * wet layer zero, h zero. */
int pnt_canvas_create(pnt_ctx ctx, int rows, int cols, pnt_canvas* out);
int pnt_canvas_destroy(pnt_canvas canvas);
int pnt_canvas_shape(pnt_canvas canvas, int* rows, int* cols);
/* getPaintLayer() -- Canvas.hxx:89-95: a view on the canvas' wet layer,
 * valid while the canvas lives; destroying the view does not free it. */
int pnt_canvas_layer(pnt_canvas canvas, pnt_layer* out);
/* clear() -- Canvas.hxx:37-58 */
int pnt_canvas_clear(pnt_canvas canvas);
/* setBackground(R0) -- Canvas.hxx:80-87 (clear(), then R0 = background) */
int pnt_canvas_set_background(pnt_canvas canvas, const float* R0, int layout);
/* getR0() / get_h() writes and reads -- Canvas.hxx:60-74 */
int pnt_canvas_upload_R0(pnt_canvas canvas, const float* R0, int layout);
int pnt_canvas_download_R0(pnt_canvas canvas, float* R0, int layout);
int pnt_canvas_upload_h(pnt_canvas canvas, const float* h);
int pnt_canvas_download_h(pnt_canvas canvas, float* h);
/* dryCanvas() -- Canvas.hxx:105-121 */
int pnt_canvas_dry(pnt_canvas canvas);
/* Partial drying, the GPU renderer's DryingStep
 * (renderer/shaders/DryingStep.compute.glsl:56-77): V leaves up to
 * `portion`, R0 takes the leaving thickness, K/S stay. */
int pnt_canvas_dry_step(pnt_canvas canvas, float portion);
/* get/setDryingTime -- Canvas.hxx:157-163.  Any value >= 0 is stored; the
 * default is the reference's 15000000 ms (Canvas.hxx:32).  The drying time
 * acts through the canvas clock below. */
int pnt_canvas_set_drying_time_ms(pnt_canvas canvas, int64_t ms);
/* The drying timer of the reference's GPU render thread
 * (SbrRenderThread.cxx:37-41: every 250 ms, CanvasGpu::dryStep(0.01),
 * CanvasGpu.hxx:29), on the canvas clock: every period_ms of clock time
 * one pnt_canvas_dry_step(portion).  The steps whose ticks have passed run
 * at the start of the next call that reads or changes the canvas (so a
 * manual clock makes them reproducible); the timer starts now.
 * period_ms = 0 stops it; it does nothing while the clock is off. */
int pnt_canvas_set_dry_timer(pnt_canvas canvas, int64_t period_ms, float portion);
int pnt_canvas_get_drying_time_ms(pnt_canvas canvas, int64_t* ms);
/* The clock of Canvas::checkDry (Canvas.hxx:123-155), which the reference's
 * imprint runs on every visited pixel before its pickup
 * (FootprintBrush.hxx:86, 115): paint older than the drying time dries into
 * the substrate as dryCanvas does, and the pixel's time point becomes the
 * imprint's.
 *   PNT_CLOCK_OFF (the default): no check -- stamping is deterministic, and
 *     only pnt_canvas_dry / PNT_STAMP_DRY_AFTER / pnt_canvas_dry_step dry.
 *   PNT_CLOCK_SYSTEM: each imprint takes system_clock::now(), as the
 *     reference does.
 *   PNT_CLOCK_MANUAL: each imprint takes the time set by
 *     pnt_canvas_set_time_us (reproducible runs and tests).
 * Switching a clock on sets every pixel's time point to now, as the
 * reference's construction does (Canvas.hxx:45-50); pnt_canvas_clear and
 * pnt_canvas_dry do the same (Canvas.hxx:47, 118).  Time points are
 * microseconds; elapsed time is truncated to whole milliseconds.  With a
 * clock on, pnt_imprint_stamps runs the plain path whatever mode it is
 * given, and pnt_batch_* refuse the canvas. */
#define PNT_CLOCK_OFF 0
#define PNT_CLOCK_SYSTEM 1
#define PNT_CLOCK_MANUAL 2
int pnt_canvas_set_clock(pnt_canvas canvas, int clock);
int pnt_canvas_get_clock(pnt_canvas canvas, int* clock);
int pnt_canvas_set_time_us(pnt_canvas canvas, int64_t now_us);
/* Canvas::checkDry (Canvas.hxx:123-155) of pixel (x, y) at time point now_us,
 * called directly; needs a clock (the time map). */
int pnt_canvas_check_dry(pnt_canvas canvas, int x, int y, int64_t now_us);
/* Canvas::getTimeMap (Canvas.hxx:101-103): rows*cols time points (us). */
int pnt_canvas_upload_time_map(pnt_canvas canvas, const int64_t* tmap_us);
int pnt_canvas_download_time_map(pnt_canvas canvas, int64_t* tmap_us);

/* ------------------------------------------------- Renderer<vec3> ------- */
/* compose(canvas) -- renderer/Renderer.hxx:48-53: R = KM(wet layer over R0),
 * out of place, into a host image. */
int pnt_compose(pnt_canvas canvas, float* R, int layout);
/* Renderer::render -- Renderer.hxx:60-156: the composite lit by a point
 * light at (-200, -1500, -2000) with a Beckmann / Cook-Torrance specular
 * term over the height field of the wet layer's V (reflect-bordered
 * neighbours), clamped to [0, 1].  pnt_render downloads it (layout as for
 * downloads); pnt_dev_render (below) writes caller device planes. */
int pnt_render(pnt_canvas canvas, float* rgb, int layout);
/* io::imSave's pixel transform (io/src/ImageIO.cxx:158-183):
 * ColorConverter::rgb2srgb when to_srgb (core/Color.hxx:197-205), then
 * scale by 2^bits - 1 (bits 8: uint8_t, 16: uint16_t), round half to even
 * and saturate, as Mat::convertTo does.  Output: n interleaved RGB triples
 * (the reference swaps to BGR only for OpenCV's file writer; writing files
 * is left to the caller).  pnt_encode goes host to host; pnt_dev_encode
 * (below, with the device-plane entry points) stays on the device. */
int pnt_encode(pnt_ctx ctx, int64_t n, const float* rgb, int layout, int to_srgb,
               int bits, void* out);
/* compose(paintLayer, R0_buffer) -- Renderer.hxx:26-41 with host images. */
int pnt_compose_layer_host(pnt_layer layer, const float* R0, float* R,
                           int layout);

/* ------------------------------------- FootprintBrush<vec3> ------------- */
/* FootprintBrush(radius) -- renderer/FootprintBrush.hxx:29-63.  The
 * reference loads ./data/footprint/footprint.png and resizes it to
 * width = 2*ceil(radius)+1 (pnt_image_read_gray + pnt_resize_lanczos4 do
 * the same); here the width x width footprint (finite values; the Lanczos
 * resize rings a little below 0 and above 1, and cells <= 0 take no part
 * in pickup or deposit, FootprintBrush.hxx:355, 398) is passed in.  It is
 * padded with zeros to sizeMap =
 * ceil(sqrt(2)*width) on each side (the reference pads one short when
 * sizeMap-width is odd and reads past its buffer; that row/col reads 0). */
int pnt_brush_create(pnt_ctx ctx, double radius, const float* footprint,
                     int width, pnt_brush* out);
int pnt_brush_destroy(pnt_brush brush);
/* setRadius(radius) -- FootprintBrush.hxx:46-63: new footprint, pickup map
 * re-created and cleared.  Like the reference, nothing happens when
 * |radius - current| < 0.5 (fuzzyCompare with Eps 0.5). */
int pnt_brush_set_radius(pnt_brush brush, double radius, const float* footprint,
                         int width);
int pnt_brush_size_map(pnt_brush brush, int* size_map);
/* getFootprint() -- FootprintBrush.hxx:178-180 (sizeMap x sizeMap) */
int pnt_brush_get_footprint(pnt_brush brush, float* out);
/* getPickupMap() -- FootprintBrush.hxx:174-176 (sizeMap x sizeMap layer) */
int pnt_brush_get_pickup_map(pnt_brush brush, float* K, float* S, float* V,
                             int layout);
/* set/getPickupRate, set/getDepositionRate, set/getUseSnapshotBuffer --
 * FootprintBrush.hxx:182-204 (defaults 0.9, 0.05, true: :477-495).  The
 * rates are T = double in the reference and here: the visit arithmetic runs
 * in FP64 on the stored FP32 values. */
int pnt_brush_set_pickup_rate(pnt_brush brush, double rate);
int pnt_brush_get_pickup_rate(pnt_brush brush, double* rate);
int pnt_brush_set_deposition_rate(pnt_brush brush, double rate);
int pnt_brush_get_deposition_rate(pnt_brush brush, double* rate);
int pnt_brush_set_use_snapshot(pnt_brush brush, int use);
int pnt_brush_get_use_snapshot(pnt_brush brush, int* use);
/* dip(paint) -- FootprintBrush.hxx:150-154: clean(), then paint = {K, S} */
int pnt_brush_dip(pnt_brush brush, const float K[3], const float S[3]);
/* clean() -- FootprintBrush.hxx:160-166 */
int pnt_brush_clean(pnt_brush brush);
/* updateSnapshot(canvas) -- FootprintBrush.hxx:168-172 */
int pnt_brush_update_snapshot(pnt_brush brush, pnt_canvas canvas);
/* imprint(center, theta, canvas) -- FootprintBrush.hxx:73-143 */
int pnt_imprint(pnt_brush brush, pnt_canvas canvas, double cx, double cy,
                double theta);

/* An ordered stamp list: the GUI's dip / imprint / dryCanvas sequence
 * (apps/painty_gui/DigitalCanvas.cxx:49-123) as one call.  Each stamp is
 * imprint(center=(cx,cy), theta); PNT_STAMP_DIP dips the brush in (K, S)
 * first, PNT_STAMP_DRY_AFTER runs dryCanvas() after it.  The result is the
 * same as issuing the calls one by one. */
typedef struct {
  double cx, cy, theta;
  float K[3], S[3];
  uint32_t flags;
  int32_t reserved;
} pnt_stamp;
#define PNT_STAMP_DIP 1u
#define PNT_STAMP_DRY_AFTER 2u
/* mode: PNT_EXEC_PLAIN issues each stamp as its own kernel launches
 * (snapshot ring copy, then the visits); PNT_EXEC_PERSISTENT runs the whole
 * list in one persistent launch; PNT_EXEC_FUSED plans the list once and
 * issues one launch per stamp in which the snapshot refresh (dirty pixels
 * only) and the visits run side by side; PNT_EXEC_CHAINED runs the same
 * work for the whole list in one launch, each stamp's blocks starting when
 * the previous stamp's blocks signal completion through device counters. */
#define PNT_EXEC_PLAIN 0
#define PNT_EXEC_PERSISTENT 1
#define PNT_EXEC_FUSED 2
#define PNT_EXEC_CHAINED 3
int pnt_imprint_stamps(pnt_brush brush, pnt_canvas canvas,
                       const pnt_stamp* stamps, int64_t n, int mode);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the same stamps and results as
 * pnt_imprint_stamps(PNT_EXEC_FUSED) (stamps in stream order on the brush's
 * own pickup map), with the stamp plan in a fixed 4 MiB device array of the
 * library instead of library-allocated buffers; the call allocates no device
 * or pinned memory.  Limits: at most 24576 stamps per call (and 1 MiB of
 * shared-pixel visit lists), else PNT_EINVAL.  The brush's snapshot buffers
 * must already fit the canvas (a previous stamp call on a canvas at least as
 * large, or pnt_brush_update_snapshot), else PNT_EINVAL; so is a canvas
 * with a drying clock (pnt_canvas_set_clock).  Returns once the
 * plan is on the device and the stamps are enqueued.  Every call on one GPU
 * uses that GPU's one array: calls that overlap on one GPU (other threads or
 * streams) race and must be serialised by the caller. */
int pnt_imprint_stamps_noalloc(pnt_brush brush, pnt_canvas canvas, const pnt_stamp* stamps,
                               int64_t n);

/* A stroke list: for each stroke, FootprintBrush::setRadius(radius) with
 * its width x width footprint (FootprintBrush.hxx:46-63: nothing changes
 * within 0.5 of the current radius, otherwise a new footprint and a clean
 * pickup map), then its stamps as pnt_imprint_stamps would run them.  The
 * brush ends with the last stroke's geometry.  Under PNT_EXEC_FUSED the
 * strokes whose snapshot boxes do not meet run concurrently (each with its
 * own pickup map; the result is the same to the bit); other modes run the
 * strokes one after another. */
typedef struct {
  double radius;
  const float* footprint;
  int width;
  const pnt_stamp* stamps;
  int64_t n_stamps;
} pnt_stroke;
int pnt_imprint_strokes(pnt_brush brush, pnt_canvas canvas, const pnt_stroke* strokes,
                        int64_t n_strokes, int mode);

/* Many independent (canvas, brush, stamp list) jobs at once, each then
 * composited (Renderer::compose, Renderer.hxx:48-53) into `out` -- three
 * 16-byte-aligned device planes of rows*cols floats, or out[0] == NULL for
 * no composite.  Canvases and brushes must be distinct across the jobs.
 * PNT_BATCH_ONE_LAUNCH runs every job and composite in one cooperative
 * launch (jobs packed onto groups of workgroups by cost, composites taken
 * by whichever workgroups are free); PNT_BATCH_PLAIN issues each job's
 * stamps and composite as separate launches, one job after another.  Each
 * job's result equals pnt_imprint_stamps + pnt_compose on its own. */
typedef struct {
  pnt_canvas canvas;
  pnt_brush brush;
  const pnt_stamp* stamps;
  int64_t n_stamps;
  float* out[3];
} pnt_batch_job;
#define PNT_BATCH_ONE_LAUNCH 0
#define PNT_BATCH_PLAIN 1
int pnt_batch_run(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, int mode);
/* The same one-launch batch as a submit / wait pair (no reference
 * counterpart; for callers that overlap host work with the batch): submit
 * enqueues it on the context's stream and returns once the plan is staged
 * (the stamp lists may then be reused); wait blocks until it is done and
 * reports an aborted kernel as PNT_EKERNEL. */
int pnt_batch_submit(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs);
int pnt_batch_wait(pnt_ctx ctx);

/* How a batch runs (no reference counterpart).  waves == 0: one persistent
 * cooperative launch, window tiles of tile x tile pixels (16 or 32; one
 * thread each) and at most max_group workgroups on one job's stamps.
 * waves == 1: the batch planned on the host into a sequence of launches
 * ("waves"), each carrying the next step of every ready stroke of every
 * canvas, with at most `chunk` strokes (1, 2 or 4) of one canvas in flight
 * at once (each with its own pickup map); 16 x 16 tiles.  A context starts
 * with {32, 16, 0, 1}. */
typedef struct {
  int tile;
  int max_group;
  int waves;
  int chunk;
} pnt_batch_config;
int pnt_batch_set_config(pnt_ctx ctx, const pnt_batch_config* cfg);
int pnt_batch_get_config(pnt_ctx ctx, pnt_batch_config* cfg);
/* Autotuner: runs the batch under each candidate configuration, timing it
 * on the context's stream, restores every canvas and brush of the batch
 * after each run, and keeps the fastest configuration for the context.
 * The jobs' state is unchanged on return. */
int pnt_batch_tune(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs,
                   pnt_batch_config* best, double* best_ms);
/* Diagnostic, no counterpart in the reference: runs the batch once as waves
 * (the context's chunk), as pnt_batch_run would, with a timing event after
 * every wave launch, and reports the schedule.  job_first[k] / job_last[k]
 * (n_jobs entries each): the first and last wave holding a phase of job k
 * (-1: none).  wave_blocks[w], wave_end_ms[w] (cap entries each): the
 * workgroups of wave w and the time from the batch's start (after its image
 * upload) to the end of wave w.  *n_waves: the wave count; waves beyond cap
 * run but are not reported.  Any output pointer may be NULL. */
int pnt_batch_profile(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs, int cap,
                      int* job_first, int* job_last, int* wave_blocks, float* wave_end_ms,
                      int* n_waves);
/* The batch planned once and captured as a HIP graph (plan copy, counter
 * clear, k_batch -- or, in the waves configuration, the plan copy and every
 * wave); each launch replays it on the jobs' current state -- the
 * same result as pnt_batch_run(PNT_BATCH_ONE_LAUNCH) with the same jobs.
 * Every stamp list must start with PNT_STAMP_DIP; a canvas or brush changed
 * outside the graph since its last launch makes the next launch fail
 * (PNT_EINVAL).  The captured kernel is a plain launch of at most the
 * co-resident workgroup count: run it on an otherwise idle device (its
 * waits are bounded, so a violation reports PNT_EKERNEL, it cannot hang). */
typedef struct pnt_batch_graph_s* pnt_batch_graph;
int pnt_batch_capture(pnt_ctx ctx, const pnt_batch_job* jobs, int n_jobs,
                      pnt_batch_graph* out);
int pnt_batch_graph_launch(pnt_batch_graph graph);
int pnt_batch_graph_destroy(pnt_batch_graph graph);

/* paintStroke(path, canvas) -- FootprintBrush.hxx:206-268: Catmull-Rom
 * placement at 1 px chord steps (core/Spline.hxx:27-73), one imprint per
 * step.  path_xy holds n_points (x, y) pairs.  The reference reads
 * path[ULONG_MAX] for the first segment (:252); its intent, p_pre = path[0],
 * is used. */
int pnt_paint_stroke(pnt_brush brush, pnt_canvas canvas, const double* path_xy,
                     int64_t n_points);

/* Measurement helper (no reference counterpart): replays the index math of
 * `n` stamps on a rows x cols canvas with this brush's footprint and counts
 * the active visits (in canvas, in map, footprint > 0) and the in-canvas
 * snapshot-ring pixels the stamps refresh (0 when the snapshot is off).
 * Host only; the brush and canvas state are not touched. */
int pnt_brush_stamp_traffic(pnt_brush brush, int rows, int cols,
                            const pnt_stamp* stamps, int64_t n,
                            int64_t* active_visits, int64_t* ring_pixels);

/* Measurement helper (no reference counterpart): the number of stamp-kernel
 * launches (k_stamp_fused, k_stamp_levels, k_stamp_wave) the brush's last
 * PNT_EXEC_FUSED stamp-list call issued; 0 after other modes. */
int pnt_brush_last_launches(pnt_brush brush, int64_t* launches);
/* Per-launch kernel timing of a brush's stamp-list calls (measurement; no
 * counterpart in the reference, which has no device kernels): while on, each
 * stamp-kernel launch of the next pnt_imprint_stamps / pnt_imprint_strokes
 * call on the brush's own stream is bracketed by a pair of HIP events.
 * pnt_brush_kernel_time waits for that call and gives the sum of the
 * bracketed kernels' durations (ms) and how many were timed. */
int pnt_brush_set_kernel_timing(pnt_brush brush, int on);
int pnt_brush_kernel_time(pnt_brush brush, double* ms, int64_t* launches);

/* --------------------------------------------------------- PaintMixer --- */
/* mixSinglePaint(weights) -- mixer/src/PaintMixer.cxx:327-355: K = sum_l
 * norm*w_l*K_l (same for S), norm = 1/sum(w) unless |sum(w)-1| < 1e-5.
 * base_K / base_S: n_palette x 3; PNT_EINVAL when n_weights != n_palette. */
/* ComputeScatteringAndAbsorption -- KubelkaMunk.hxx:84-124 (Curtis et al.
 * 1997): K and S of n colours from their reflectance over black (Rb) and
 * over white (Rw), AoS n*3 doubles each, computed on the device in FP64.
 * Fails with PNT_EINVAL unless 0 < Rb < Rw < 1 on every channel (the
 * reference throws std::invalid_argument). */
int pnt_compute_scattering_absorption(pnt_ctx ctx, int64_t n, const double* Rb,
                                      const double* Rw, double* K, double* S);
/* LoadPalette / SavePalette -- mixer/src/Serialization.cxx:26-53: the JSON
 * array [{"K": [k0, k1, k2], "S": [s0, s1, s2]}, ...].  from_json reads up
 * to `capacity` entries into K, S (AoS n*3) and sets *n (other keys are
 * ignored, "K" and "S" are required); to_json writes the compact form with
 * the shortest round-trip decimals, NUL-terminated, and sets *len (fails
 * when capacity <= *len). */
int pnt_palette_from_json(const char* text, int64_t len, double* K, double* S,
                          int64_t capacity, int64_t* n);
int pnt_palette_to_json(const double* K, const double* S, int64_t n, char* text,
                        int64_t capacity, int64_t* len);
/* getThinningMedium -- mixer/src/Palette.cxx:453-463: the single entry of the
 * reference's data/paint_palettes/thinning_medium.json, K = S = 0.01. */
void pnt_thinning_medium(double K[3], double S[3]);
int pnt_mix_single(int n_palette, const double* base_K, const double* base_S,
                   int n_weights, const double* weights, double K[3],
                   double S[3]);
/* mixed(paint, vol, other, vol) -- PaintMixer.cxx:536-548 */
int pnt_mixed(const double K1[3], const double S1[3], double v1,
              const double K2[3], const double S2[3], double v2, double K[3],
              double S[3]);

/* ------------------------------------------ KubelkaMunk free functions -- */
/* ComputeReflectance(K, S, R0, d) over n pixels of host arrays (each K, S,
 * R0, R n x 3 AoS, d n floats) -- core/KubelkaMunk.hxx:25-83 on the device. */
int pnt_compute_reflectance(pnt_ctx ctx, int64_t n, const float* K,
                            const float* S, const float* R0, const float* d,
                            float* R);

/* ---------------------------------------------- device-pointer kernels -- */
/* The same kernels on caller-owned device planes (e.g. torch tensors), on
 * `stream` (a hipStream_t; NULL = the context's stream).  A layer is 7
 * planes of n floats: K0 K1 K2 S0 S1 S2 V; rgb is 3 planes.  Planes are
 * float pointers; 16-byte aligned planes use the vectorised path. */
typedef struct {
  float* p[7];
} pnt_dev_layer;
typedef struct {
  float* p[3];
} pnt_dev_rgb;
/* A canvas' wet layer on the device: K0 K1 K2 S0 S1 S2 as FP32 planes and
 * the volume V as an FP64 plane (the reference holds V in double, and its
 * MinVolume and drying tests see it). */
typedef struct {
  float* ks[6];
  double* v;
} pnt_dev_wet;
/* Renderer::render into caller device planes (see pnt_render). */
int pnt_dev_render(pnt_canvas canvas, pnt_dev_rgb rgb);
/* The imSave pixel transform (see pnt_encode) from device planes to device
 * memory, on `stream` (NULL: the context's stream). */
int pnt_dev_encode(pnt_ctx ctx, int64_t n, pnt_dev_rgb rgb, int to_srgb, int bits,
                   void* out, void* stream);
/* R = KM(layer over R0); R may equal R0 (composeOnto). */
int pnt_dev_compose(pnt_ctx ctx, int64_t n, pnt_dev_layer layer,
                    pnt_dev_rgb R0, pnt_dev_rgb R, void* stream);
/* R = KM(wet over KM(dry over R0)); R0 is not modified. */
int pnt_dev_compose2(pnt_ctx ctx, int64_t n, pnt_dev_layer dry,
                     pnt_dev_layer wet, pnt_dev_rgb R0, pnt_dev_rgb R,
                     void* stream);
/* The same two on canvas wet layers (pnt_canvas_device_planes). */
int pnt_dev_compose_wet(pnt_ctx ctx, int64_t n, pnt_dev_wet layer,
                        pnt_dev_rgb R0, pnt_dev_rgb R, void* stream);
int pnt_dev_compose2_wet(pnt_ctx ctx, int64_t n, pnt_dev_wet dry,
                         pnt_dev_wet wet, pnt_dev_rgb R0, pnt_dev_rgb R,
                         void* stream);

/* ------------------------------------- differentiable composite+stamp -- */
/* No reference counterpart: the reference has no gradients.  A stamp list
 * (PNT_STAMP_DIP on stamp 0; no PNT_STAMP_DRY_AFTER) run with `brush` on a
 * rows x cols canvas, each dip's paint K/S = mixSinglePaint(W[paint_of[i]])
 * (PaintMixer.cxx:327-355) over an n_pig palette (palK/palS n_pig x 3), then
 * R = Renderer::compose (Renderer.hxx:26-41).  pnt_grad_backward gives the
 * gradient of <gR, R> w.r.t. the initial wet layer (K0..S2, V planes), the
 * substrate R0 and W (n_paint x n_pig, row-major), by recomputing the stamp
 * state from checkpoints taken every seg_len stamps (and at every dip).
 * Device planes are caller-owned, n = rows*cols floats each; the brush's
 * pickup map and snapshot are used as scratch by both calls. */
typedef struct pnt_grad_s* pnt_grad;
int pnt_grad_create(pnt_ctx ctx, int rows, int cols, pnt_brush brush,
                    const pnt_stamp* stamps, int64_t n, const int32_t* paint_of,
                    int n_pig, const double* palK, const double* palS,
                    int n_paint, int seg_len, pnt_grad* out);
int pnt_grad_destroy(pnt_grad g);
/* R = compose(stamps(wet0 over R0)); W: n_paint x n_pig host doubles. */
int pnt_grad_forward(pnt_grad g, pnt_dev_layer wet0, pnt_dev_rgb R0,
                     const double* W, pnt_dev_rgb R);
/* After pnt_grad_forward: gradients of <gR, R>.  gW: host n_paint x n_pig. */
int pnt_grad_backward(pnt_grad g, pnt_dev_rgb gR, pnt_dev_layer gwet,
                      pnt_dev_rgb gR0, double* gW);

/* ------------------------------------------------------------- server -- */
/* No reference counterpart (the reference paints on one canvas in one GUI
 * thread; its GPU renderer queues tasks FIFO, gpu/GpuTaskQueue.hxx:32-46).
 * A request server for stroke + composite work on canvases kept resident on
 * one device, with continuous batching: up to n_slots admitted requests run
 * at once, and their stamps share device launches ("waves": one launch
 * carries the next phase of every ready stroke of every admitted request);
 * requests that arrive meanwhile join the next wave.  Requests on one canvas
 * run in submission order, one at a time; a request whose canvas is not
 * resident is admitted when it fits hbm_budget_bytes (canvas planes +
 * composite planes, 56 B/px), spilling idle canvases to host memory, least
 * recently used first; a later request on a spilled canvas restores it.
 * Each request runs as a fresh FootprintBrush (FootprintBrush.hxx:29-63):
 * its stamps (with any PNT_STAMP_DRY_AFTER), then Renderer::compose
 * (Renderer.hxx:48-53) into the canvas' resident composite planes, copied
 * to `out` when it is not NULL.  Each request's result is bit for bit the
 * same as running it alone through pnt_imprint_stamps(PNT_EXEC_FUSED) with
 * a fresh brush, then pnt_compose. */
typedef struct pnt_server_s* pnt_server;
typedef struct {
  int64_t canvas_id;       /* resident canvas key; a new id creates a canvas */
  int rows, cols;          /* its size (must match an existing canvas) */
  float background[3];     /* substrate R0 of a new canvas */
  double radius;           /* brush radius and its width x width footprint */
  const float* footprint;  /* (copied at submit) */
  int width;
  double pickup_rate, deposition_rate;
  int use_snapshot;
  const pnt_stamp* stamps; /* stamp list (copied at submit); PNT_STAMP_DIP sets paint */
  int64_t n_stamps;
  float* out;              /* host composite, 3 planes of rows*cols, or NULL */
} pnt_request;
int pnt_server_create(int device, int n_slots, int64_t hbm_budget_bytes,
                      pnt_server* out);
/* Queues a request; returns at once with a ticket. */
int pnt_server_submit(pnt_server server, const pnt_request* request,
                      int64_t* ticket);
/* Blocks until the ticket's request has run; its latency (submit to done,
 * ms) and its status. */
int pnt_server_wait(pnt_server server, int64_t ticket, double* latency_ms);
/* A resident canvas (no request running on it) and its composite planes. */
int pnt_server_canvas(pnt_server server, int64_t canvas_id, pnt_canvas* canvas,
                      pnt_dev_rgb* composite);
/* Ends a canvas: once no request on it is queued or running, its device
 * block returns to the server's pool and any host copy is freed; a later
 * request with the same id starts a new canvas.  A request submitted on the
 * id before that happens cancels the release. */
int pnt_server_release(pnt_server server, int64_t canvas_id);
int pnt_server_stats(pnt_server server, int64_t* resident_bytes,
                     int64_t* resident_canvases, int64_t* completed);
/* Finishes the queued work, then releases everything. */
int pnt_server_destroy(pnt_server server);

/* ------------------------------------------------------- TextureBrush --- */
/* TextureBrush<vector_type> -- renderer/TextureBrush.hxx:19-238 with its
 * Smudge (renderer/Smudge.hxx:22-229): a brush stroke texture (thickness
 * map) warped along a stroke by mean value coordinates over the stroke's
 * frame polygon (core/Math.hxx:73-133), an optional smudge pass, then
 * volume-weighted mixing with V = max(Vtex, Vcanvas).  Paint, thickness map
 * and smudge maps are FP64 as the reference's; the canvas is FP32. */
typedef struct pnt_tbrush_s* pnt_tbrush;
/* TextureBrush(sampleDir) -- TextureBrush.hxx:25-31: radius 0, paint 0.1,
 * smudge on, thickness scale 1; the sample's thickness map is set apart
 * (pnt_tbrush_set_thickness_map) since image IO is the caller's. */
int pnt_tbrush_create(pnt_ctx ctx, pnt_tbrush* out);
int pnt_tbrush_destroy(pnt_tbrush brush);
/* setRadius -- TextureBrush.hxx:33-41: a change of 0.5 or more resets the
 * smudge to a (int)(2 radius) square map when smudging is on. */
int pnt_tbrush_set_radius(pnt_tbrush brush, double radius);
/* dip -- TextureBrush.hxx:48-50 */
int pnt_tbrush_dip(pnt_tbrush brush, const double K[3], const double S[3]);
/* BrushStrokeSample::getThicknessMap (renderer/src/BrushStrokeSample.cxx:
 * 24-31, 164): rows x cols row-major values, finite. */
int pnt_tbrush_set_thickness_map(pnt_tbrush brush, const double* map, int rows,
                                 int cols);
/* BrushBase::setThicknessScale -- renderer/BrushBase.hxx:31-33 */
int pnt_tbrush_set_thickness_scale(pnt_tbrush brush, double scale);
/* enableSmudge -- TextureBrush.hxx:207-209 */
int pnt_tbrush_set_use_smudge(pnt_tbrush brush, int enable);
/* The smudge's pickup map (Smudge.hxx _pickupMapDst): side x side cells,
 * K and S AoS (3 per cell) and V; any output may be NULL. */
int pnt_tbrush_get_smudge_map(pnt_tbrush brush, int* side, double* K,
                              double* S, double* V);
/* paintStroke -- TextureBrush.hxx:52-205: xy holds n vertices (x, y);
 * fewer than 2 do nothing.  Canvas::checkDry runs on textured pixels when
 * the canvas clock is on.  Fails with PNT_EINVAL without a thickness map,
 * or when the smudge would take more than 2^24 spine steps. */
int pnt_tbrush_paint_stroke(pnt_tbrush brush, pnt_canvas canvas,
                            const double* xy, int64_t n);

/* ---------------------------------------------------- TextureBrushGpu --- */
/* TextureBrushGpu -- renderer/TextureBrushGpu.hxx:21-64,
 * renderer/src/TextureBrushGpu.cxx:17-305: the reference's GL texture brush
 * (used by its render thread, SbrRenderThread).  A stroke rasterises the
 * brush texture along a triangle strip around the spline through the
 * vertices (BrushTextureWarp.vert/.frag), optionally smudges
 * (Smudge.compute: pickup and deposition rates 0.7, the pickup map read at
 * the rotated position), then mixes by volume with V = Vcan + Vtex and
 * stores the lighting thickness (PaintTextureFootprintOnCanvas.compute), and
 * clears the warped texture (ClearTextureFootprint.compute).  The GL images
 * are y-flipped (gpu/GpuMat.hxx:43-56), and the smudge's dispatch offset is
 * taken from the canvas position unflipped, as in the reference.  Where GL
 * leaves behaviour to the implementation this fixes one choice: vertices
 * snapped to 1/256 pixel, pixel centres inside a triangle with the top-left
 * rule (window y up), linear interpolation of texture coordinates, later
 * triangles overwriting earlier ones, GL_LINEAR + GL_CLAMP_TO_EDGE sampling,
 * and for the smudge's writes to one pickup cell: every invocation reads the
 * state before the dispatch and the highest invocation index wins.  The
 * reference picks the brush texture at random from its texture dictionary
 * (TextureBrushDictionary.cxx:24-68); here the caller sets it. */
typedef struct pnt_tbgl_s* pnt_tbgl;
/* TextureBrushGpu() -- TextureBrushGpu.cxx:17-46: radius 5 (an 11 x 11
 * smudge window), smudging off, paint 0. */
int pnt_tbgl_create(pnt_ctx ctx, pnt_tbgl* out);
int pnt_tbgl_destroy(pnt_tbgl brush);
/* setRadius -- TextureBrushGpu.cxx:48-60: a change of 0.5 or more makes new,
 * cleared smudge maps of (int)(2 radius) + 1 square. */
int pnt_tbgl_set_radius(pnt_tbgl brush, double radius);
/* dip -- TextureBrushGpu.cxx:62-64 */
int pnt_tbgl_dip(pnt_tbgl brush, const double K[3], const double S[3]);
/* The brush texture as a texture dictionary entry makes it
 * (TextureBrushDictionary.cxx:106-123): the height map (rows x cols,
 * row-major, finite) normalised to [0, 1] by its minimum and maximum,
 * flipped, FP32. */
int pnt_tbgl_set_texture(pnt_tbgl brush, const double* img, int rows, int cols);
/* BrushBase::setThicknessScale -- renderer/BrushBase.hxx:31-33 */
int pnt_tbgl_set_thickness_scale(pnt_tbgl brush, double scale);
/* enableSmudge -- TextureBrushGpu.cxx:249-251 */
int pnt_tbgl_enable_smudge(pnt_tbgl brush, int enable);
/* The smudge maps (side x side RGBA: K or S, then V); either may be NULL. */
int pnt_tbgl_get_smudge(pnt_tbgl brush, int* side, float* K4, float* S4);
/* paintStroke(vertices, CanvasGpu&) -- TextureBrushGpu.cxx:73-166; fewer
 * than 2 vertices do nothing; PNT_EINVAL without a brush texture. */
int pnt_tbgl_paint_stroke(pnt_tbgl brush, pnt_canvas canvas, const double* xy,
                          int64_t n);
/* The canvas' lighting thickness (the GL renderer's S.a plane), rows x cols;
 * zero until a pnt_tbgl stroke writes it. */
int pnt_canvas_download_lighting(pnt_canvas canvas, float* out);

/* ---------------------------------------------------- footprint ingest - */
/* io::imRead(buffer, Mat<double>& gray, convertFrom_sRGB) -- io/src/
 * ImageIO.cxx:115-150 (the file form, :78-113, reads the same bytes): a PNG
 * file's grey samples (8 or 16 bit; colour type grey or grey + alpha, the
 * alpha dropped as IMREAD_GRAYSCALE does; not interlaced), times 1/255 or
 * 1/65535, then sRGB -> linear (core/Color.hxx:189-195) when srgb != 0.
 * Sets *rows, *cols; fills out (row-major, capacity >= rows*cols) unless it
 * is NULL.  Other PNG kinds fail with PNT_EINVAL. */
int pnt_image_read_gray(const uint8_t* data, int64_t len, int srgb, int* rows,
                        int* cols, double* out, int64_t capacity);
/* ScaledMat(input, rows, cols) -- image/Mat.hxx:141-148: cv::resize with
 * INTER_LANCZOS4 (OpenCV's 8-tap kernel, float weights, replicated borders,
 * FP64 sums) of a single-channel FP64 image, on the device.  The footprint
 * of FootprintBrush::setRadius (FootprintBrush.hxx:48-58) is
 * ScaledMat(imRead(footprint.png), width, width). */
int pnt_resize_lanczos4(pnt_ctx ctx, const double* src, int rows, int cols,
                        int out_rows, int out_cols, double* out);

/* ---------------------------------------------------- canvas device view - */
/* Device planes of a canvas (valid while it lives): wet layer (K, S FP32,
 * V FP64), R0 (3), h. */
int pnt_canvas_device_planes(pnt_canvas canvas, pnt_dev_wet* layer,
                             pnt_dev_rgb* R0, float** h);

#ifdef __cplusplus
}
#endif
#endif /* PAINTY_HIP_H */
