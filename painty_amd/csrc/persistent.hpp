// Stamp lists planned on the host and run from a device-side plan:
// PNT_EXEC_PERSISTENT (persistent.hip), PNT_EXEC_FUSED (fused.hip) and
// batches (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "painty_hip.h"
#include "internal.hpp"

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

// A run of stamps of one stroke (from a dip to the next dip) within one
// drying phase, and the earlier units of its phase whose snapshot boxes meet
// its own (fused.hip, plan_units).
struct Unit {
  int64_t s0, s1;             // stamps [s0, s1)
  int x0, y0, x1, y1;         // union of the allowed boxes (x1 < x0: touches nothing)
  int chain;
  bool dry_after;             // the unit's last stamp dries the canvas
  std::vector<int> deps;      // earlier units of the phase it must follow
};
// Cuts stamps [0, n) (planned in h) into units and assigns them to at most
// nchains pickup-map chains by list scheduling; clear[k]: the map is clean
// before stamp k.
std::vector<Unit> plan_units(const pnt_stamp* st, const PStamp* h, const uint8_t* clear,
                             int64_t n, int nchains);
// Per stamp, the units it depends on and how many of their leading stamps
// must have run first (need[off[k] .. off[k+1]) = (unit, count)).
struct StampNeeds {
  std::vector<int> off;
  std::vector<int2> need;
};
StampNeeds plan_stamp_needs(const std::vector<Unit>& U, const PStamp* h);
// Blocks of `threads` threads that scan the snapshot ring of box b.
// groups: 4-pixel groups per ring-scan thread (the ring_scan<groups> the
// launch runs).
int ring_blocks_for(const RingBox& b, int threads, int groups = kRingGroups);
// Plans stamps [0, n) for a brush of footprint side size_map and radius on a
// rows x cols canvas (FP64 geometry, ring boxes, shared-pixel levels).
void plan_stamps_rc(int size_map, double radius, int rows, int cols, const pnt_stamp* st,
                    int64_t n, std::vector<PStamp>& h_st, std::vector<int2>& h_dl,
                    std::vector<int>& h_off);

// chains: concurrent stroke chains at most (0: PAINTY_FUSED_CHAINS or 4)
void run_fused(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n, int chains = 0);
void run_fused_strokes(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stroke* strokes, int64_t ns);
// The fused stamps of a list in stream order with the plan in a file-scope
// device array (pnt_imprint_stamps_noalloc).
void run_fused_noalloc(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
void run_chained(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
struct ChainState;
void chain_release(ChainState*& c);
struct FusedState;
void fused_release(FusedState*& f);
// Event pairs around each stamp-kernel launch of a brush's stamp-list calls,
// for measuring the kernels' own durations on their stream.
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;  // start, end of launch k at 2k, 2k+1
  size_t used = 0;
};
void ktimer_release(KernelTimer*& t);
// Sum of the timed launches' durations in ms (waits for the last one).
double ktimer_total_ms(KernelTimer& t);
void run_persistent(pnt_brush_s* b, pnt_canvas_s* cv, const pnt_stamp* st, int64_t n);
void persistent_release(PersistentState*& p);

// pnt_batch_run(..., PNT_BATCH_ONE_LAUNCH), pnt_batch_tune and captured
// batches (batch.hip).  cfg NULL: the context's configuration.
void run_batch_one_launch(pnt_ctx ctx, const pnt_batch_job* jobs, int n,
                          const pnt_batch_config* cfg);
// pnt_batch_profile: the batch run once as waves, timed wave by wave.
struct BatchWaveProfile {
  std::vector<float> wave_end_ms;  // from the image upload to the end of each wave
  std::vector<int> wave_blocks, job_first, job_last;
};
void profile_batch_waves(pnt_ctx ctx, const pnt_batch_job* jobs, int n, BatchWaveProfile* prof);
void tune_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n, pnt_batch_config* best,
                double* best_ms);
void set_batch_config(pnt_ctx ctx, const pnt_batch_config& c);
pnt_batch_config get_batch_config(pnt_ctx ctx);
struct BatchGraph;
BatchGraph* capture_batch(pnt_ctx ctx, const pnt_batch_job* jobs, int n);
void launch_batch_graph(BatchGraph* g);
void release_batch_graph(BatchGraph*& g);
void batch_release(pnt_ctx ctx);
// A batch as waves (batchwave.hip): the device image (jobs, slots, group
// tables, stamps, visit lists, paints) and where each wave's slots and group
// table start; pickup maps beyond the brushes' own come from a pool given
// when the image is built (plan_waves_pool floats).
struct BWPlan {
  std::vector<char> blob;
  int64_t total = 0;
  int64_t pool_floats = 0;
  int64_t o_jobs = 0;
  std::vector<int64_t> wave_slots, wave_grp;
  std::vector<int> wave_blocks;
  std::vector<int> job_first, job_last;  // each job's first and last wave (-1: none)
  int njobs = 0;
};
int64_t plan_waves_pool(const pnt_batch_job* jobs, int n, int chunk);
BWPlan plan_waves(const pnt_batch_job* jobs, int n, int chunk, float* pool, hipStream_t s);
// ev (optional, wave_blocks.size() + 1 events): ev[0] is recorded after the
// image copy, ev[w + 1] after wave w.
void enqueue_waves(const BWPlan& W, const char* stage, char* dev, hipStream_t s,
                   const hipEvent_t* ev = nullptr);

}  // namespace pnt
