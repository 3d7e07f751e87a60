// Stamp lists planned on the host and run from a device-side plan:
// PNT_EXEC_PERSISTENT (persistent.hip), PNT_EXEC_FUSED (fused.hip) and
// batches (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "painty_hip.h"

struct pnt_brush_s;
struct pnt_canvas_s;

namespace pnt {

struct PersistentState;
struct PStamp;

// A stamp list's plan on the device, and its host copy (valid until the
// brush plans again).
struct PlanDev {
  const PStamp* st;
  const int2* dl;
  const int* dl_off;
  const PStamp* host;
  const int* host_off;  // host copy of dl_off
};
PlanDev upload_plan(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
// A list cut into parts that each have their own brush geometry (padded
// footprint side, radius).  The footprints of the list's geometries
// (fp_floats floats) travel with the plan, to *d_fp: fill_fp writes them
// into the pinned staging buffer once the previous plan's upload is done.
struct PlanPart {
  int size_map;
  double radius;
  const pnt_stamp* st;
  int64_t n;
};
PlanDev upload_plan_parts(pnt_brush_s* b, pnt_canvas_s* cv, const PlanPart* parts, int np,
                          int64_t fp_floats, const std::function<void(float*)>& fill_fp,
                          const float** d_fp);

// chains: concurrent stroke chains at most (0: PAINTY_FUSED_CHAINS or 4)
void run_fused(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n, int chains = 0);
void run_fused_strokes(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stroke* strokes, int64_t ns);
void run_chained(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
struct ChainState;
void chain_release(ChainState*& c);
struct FusedState;
void fused_release(FusedState*& f);
void run_persistent(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
void persistent_release(PersistentState*& p);

// pnt_batch_run(..., PNT_BATCH_ONE_LAUNCH), pnt_batch_tune and captured
// batches (batch.hip).  cfg NULL: the context's configuration.
void run_batch_one_launch(pnt_ctx ctx, const pnt_batch_job* jobs, int n,
                          const pnt_batch_config* cfg);
void tune_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n, pnt_batch_config* best,
                double* best_ms);
void set_batch_config(pnt_ctx ctx, const pnt_batch_config& c);
pnt_batch_config get_batch_config(pnt_ctx ctx);
struct BatchGraph;
BatchGraph* capture_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n);
void launch_batch_graph(BatchGraph* g);
void release_batch_graph(BatchGraph*& g);
void batch_release(pnt_ctx ctx);

}  // namespace pnt
