// Whole stamp lists in one persistent launch (PNT_EXEC_PERSISTENT).
#pragma once

#include <stdint.h>

#include "painty_hip.h"

struct pnt_brush_s;
struct pnt_canvas_s;

namespace pnt {

struct PersistentState;

void run_persistent(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
void persistent_release(PersistentState*& p);

// pnt_batch_run(..., PNT_BATCH_ONE_LAUNCH) (batch.hip)
void run_batch_one_launch(pnt_ctx ctx, const pnt_batch_job* jobs, int n);
void batch_release(pnt_ctx ctx);

}  // namespace pnt
