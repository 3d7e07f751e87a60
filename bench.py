"""Benchmark: KM composite + imprint stamping on one MI355X.

The default run (`python bench.py`) measures every leg of the metric and
prints them in ONE JSON line:

* value: stamps/s of the large single-canvas job -- a 4096x4096 FP32 canvas
  (K log-uniform [0.01,10], S log-uniform [0.1,5], thickness U(0,0.5) with 30%
  zeros, substrate U(0.7,0.95)), 2000 imprint stamps along seeded Bezier
  strokes at 2 px spacing with per-stroke footprints of 129-257 px (Gaussian x
  value noise), pickup on, dryCanvas() after every 500th stamp, then a full
  composite.  All inputs resident in HBM before the timed region.
* roofline: the stamp kernel of that job (k_stamp_wave), its average launch
  duration measured with HIP events around each launch on its stream.
* composite: the KM composite on paint-bearing input (the same canvas before
  stamping: 70% of pixels carry paint in each layer), single layer
  (Renderer::compose, 56 B/px) and two layers (dry composed onto the
  substrate, then wet, 88 B/px), with rooflines.
* batched: 256 canvases (sides 256..2048, 50..500 stamps, footprints
  32..128 px) stamped and composited in one launch vs the same jobs as plain
  per-canvas launches; `tail`: one more run timed wave by wave
  (pnt_batch_profile), when the jobs end and how much of the batch runs in
  waves too small to fill the device.
* served: Poisson arrivals of stroke + composite requests (canvases
  512..4096, 10..1000 stamps, 30% follow-ups), p50/p99 latency and req/s per
  offered rate, against the plain per-request path; and the same on 2048x2048
  canvases.
* diff: forward + backward of the differentiable 2048x2048, 500-stamp job.
* cpu_baseline: the FP64 oracle on the host (the reference's loops): the
  whole single-canvas job on one core, and the composite on all cores.

`--config single|batched|served|diff|texture|copy` runs one leg alone and
prints its own line (for profiling runs).  Synthetic data throughout; see
painty_amd/workload.py.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gc
import json
import math
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KM composite Gpixel/s + stamps/s; p50/p99 served latency vs plain GPU path"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes (FP32 SoA minimum traffic)
# K, S, R0, R, h and the pickup map are FP32; the volume V is FP64 (8 B).
BYTES_COMPOSE1 = 56   # K3 S3 V R0_3 in, R_3 out
BYTES_COMPOSE2 = 88   # dry K3 S3 V + wet K3 S3 V + R0_3 in, R_3 out
BYTES_DRY = 96        # K3 S3 V R0_3 h read, the same written
BYTES_VISIT = 164     # canvas K,S,V rw 64 + snapshot K,S,V r + V w 40 + pickup rw 56 + footprint 4
BYTES_RING = 64       # snapshot refresh: K,S,V read (32) + written (32)


class HipEvents:
    """hipEvent timing on a given stream (the stream the kernels run on)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b) -> float:
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def groups(stamps):
    ids = stamps["brush"]
    cut = np.flatnonzero(np.diff(ids)) + 1
    starts = np.concatenate([[0], cut])
    ends = np.concatenate([cut, [len(stamps)]])
    return [(int(ids[s]), stamps[s:e]) for s, e in zip(starts, ends)]


def host_description():
    """Host cores this process may use, the machine's count, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "model": model}


def profiled_traffic(kernel, per="launch"):
    """HBM bytes per launch (or per batch) of `kernel` from the committed
    counter summaries (profiles/*.md lines `traffic_bytes_per_<per>[<kernel>]:
    N`: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when no summary covers it."""
    import glob
    import re
    pat = re.compile(r"traffic_bytes_per_%s\[(.+?)\]:\s*(\d+)" % per)
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.search(line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    return None


def free_device():
    gc.collect()


# ------------------------------------------------------------ CPU baseline --
def cpu_baseline(wl, n_stamps: int, comp_rows: int, threads: int):
    """FP64 oracle, the reference's loops (test infrastructure used here only
    as the CPU baseline):
      * stamps: single thread, as the reference; n_stamps = 0 runs the whole
        job as the GPU does (every stamp and its dries, then the composite);
        otherwise the first n_stamps on the full canvas without dries;
      * composite on paint-bearing input (the canvas before stamping): one
        thread over the first comp_rows rows, and all `threads` host threads
        over the whole canvas (row bands, single and two-layer)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    R0s = wl.R0.astype(np.float64)  # raw substrate
    R0 = R0s.copy()
    O.compose_onto(wl.dry.K, wl.dry.S, wl.dry.V, R0)
    wK, wS, wV = (x.astype(np.float64) for x in (wl.wet.K, wl.wet.S, wl.wet.V))
    dK, dS, dV = (x.astype(np.float64) for x in (wl.dry.K, wl.dry.S, wl.dry.V))
    rows, cols = wl.rows, wl.cols
    n = rows * cols
    # composite, one thread, a band of rows
    comp_rows = rows if comp_rows <= 0 else min(comp_rows, rows)
    t0 = time.perf_counter()
    O.compose(wK[:comp_rows], wS[:comp_rows], wV[:comp_rows], R0[:comp_rows])
    t_c1 = time.perf_counter() - t0
    # composite, all threads, row bands (ctypes releases the GIL)
    bands = [(r, min(rows, r + max(1, rows // (4 * threads)))) for r in
             range(0, rows, max(1, rows // (4 * threads)))]

    def one_layer(b):
        O.compose(wK[b[0]:b[1]], wS[b[0]:b[1]], wV[b[0]:b[1]], R0[b[0]:b[1]])

    def two_layer(b):
        R = R0s[b[0]:b[1]].copy()
        O.compose_onto(dK[b[0]:b[1]], dS[b[0]:b[1]], dV[b[0]:b[1]], R)
        O.compose(wK[b[0]:b[1]], wS[b[0]:b[1]], wV[b[0]:b[1]], R)

    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(one_layer, bands))
        t_cn = time.perf_counter() - t0
        t0 = time.perf_counter()
        list(ex.map(two_layer, bands))
        t_cn2 = time.perf_counter() - t0
    # the stamp job, one thread
    cv = O.Canvas.from_arrays(wK.copy(), wS.copy(), wV.copy(), R0)
    whole = n_stamps <= 0 or n_stamps >= len(wl.stamps)
    sample = wl.stamps.copy() if whole else wl.stamps[:n_stamps].copy()
    if not whole:
        sample["flags"] &= ~np.uint32(2)  # no dry inside a partial sample
    brush = None
    t0 = time.perf_counter()
    for bid, grp in groups(sample):
        if brush is None:
            brush = O.Brush(wl.radii[bid], wl.cores[bid].astype(np.float64))
            brush.set_rates(wl.pickup_rate, wl.deposition_rate)
        else:
            brush.set_radius(wl.radii[bid], wl.cores[bid].astype(np.float64))
        brush.run(cv, grp)
    t_st = time.perf_counter() - t0
    t0 = time.perf_counter()
    cv.compose()
    t_fc = time.perf_counter() - t0
    return {"stamps_s": len(sample) / (t_st + t_fc) if whole else len(sample) / t_st,
            "t_stamps_s": t_st, "t_final_composite_s": t_fc, "whole": whole,
            "n_stamps": len(sample),
            "composite_gpixel_s_1core": comp_rows * cols / t_c1 / 1e9, "comp_rows": comp_rows,
            "composite_gpixel_s_all_cores": n / t_cn / 1e9,
            "composite2_gpixel_s_all_cores": n / t_cn2 / 1e9, "threads": threads}


# ------------------------------------------------------ large single canvas --
def leg_composite(wl, reps: int = 20):
    """The KM composite on paint-bearing input: the canvas's wet layer over
    the raw substrate (k_compose_vec<1>, Renderer::compose, 56 B/px), and the
    dry layer then the wet layer over the substrate in one pass
    (k_compose_vec<2>, 88 B/px).  Live HIP events on the library's stream."""
    from painty_amd import _lib
    from painty_amd import hip as H
    ctx = H.default_context()
    ev = HipEvents()
    n = wl.rows * wl.cols
    wet = H.Canvas(wl.rows, wl.cols)
    wet.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    wet.setR0(wl.R0)
    dry = H.Canvas(wl.rows, wl.cols)  # its wet-layer planes hold the dry layer
    dry.getPaintLayer().setBuffers(wl.dry.K, wl.dry.S, wl.dry.V)
    out = H.Canvas(wl.rows, wl.cols)
    Lw, R0, _ = wet.device_planes()
    Ld, _, _ = dry.device_planes()
    _, Ro, _ = out.device_planes()
    lw, ld = Lw, Ld
    r0 = _lib.DevRgb((C.c_void_p * 3)(*R0))
    ro = _lib.DevRgb((C.c_void_p * 3)(*Ro))
    ctx.synchronize()
    res = {"pixels": n,
           "wet_fraction": round(float((wl.wet.V != 0).mean()), 4),
           "dry_fraction": round(float((wl.dry.V != 0).mean()), 4)}
    for name, call, nbytes, kernel in (
            ("single_layer", lambda: _lib.call("pnt_dev_compose_wet", ctx.h, n, lw, r0, ro, None),
             BYTES_COMPOSE1, "k_compose_vec<1>"),
            ("two_layer", lambda: _lib.call("pnt_dev_compose2_wet", ctx.h, n, ld, lw, r0, ro, None),
             BYTES_COMPOSE2, "k_compose_vec<2>")):
        call()
        a, b = ev.event(), ev.event()
        ev.record(a, ctx.stream)
        for _ in range(reps):
            call()
        ev.record(b, ctx.stream)
        ms = ev.elapsed_ms(a, b) / reps
        gbs = nbytes * n / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "gpixel_s": round(n / (ms * 1e-3) / 1e9, 2),
                     "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                                  "traffic": profiled_traffic(kernel),
                                  "algorithmic_bytes_per_launch": nbytes * n}}
        log("bench: composite %s %.3f ms, %.1f Gpixel/s, %.0f GB/s" %
            (name, ms, n / (ms * 1e-3) / 1e9, gbs))
    del wet, dry, out
    free_device()
    return res


def leg_single(args, wl):
    """The large single canvas: one step = the whole stamp job (one
    pnt_imprint_strokes call: per-stroke footprints, dries inside, concurrent
    strokes as waves) and the composite of the canvas."""
    from painty_amd import _lib
    from painty_amd import hip as H
    from painty_amd import workload as W
    ctx = H.default_context()
    stream = ctx.stream
    ev = HipEvents()
    cv = H.Canvas(wl.rows, wl.cols)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    dry = H.PaintLayer(wl.rows, wl.cols)
    dry.setBuffers(wl.dry.K, wl.dry.S, wl.dry.V)
    dry.composeOntoCanvas(cv)
    del dry
    out = H.Canvas(wl.rows, wl.cols)  # its R0 planes receive the composite
    Lc, Rc, _ = cv.device_planes()
    _, Ro, _ = out.device_planes()
    dev_layer = Lc
    dev_r0 = _lib.DevRgb((C.c_void_p * 3)(*Rc))
    dev_out = _lib.DevRgb((C.c_void_p * 3)(*Ro))
    n_px = wl.rows * wl.cols
    mode = {"plain": _lib.EXEC_PLAIN, "persistent": _lib.EXEC_PERSISTENT,
            "fused": _lib.EXEC_FUSED, "chained": _lib.EXEC_CHAINED}[args.mode]

    # segments: one brush geometry, no dryCanvas inside (it runs between them)
    segs = []
    for bid, g in groups(wl.stamps):
        cut = np.flatnonzero(g["flags"] & W.DRY_AFTER) + 1
        for part in np.split(g, cut):
            if len(part):
                p2 = part.copy()
                p2["flags"] &= ~np.uint32(W.DRY_AFTER)
                segs.append((bid, H.stamps_to_c(p2), bool(part["flags"][-1] & W.DRY_AFTER)))
    brush = H.FootprintBrush(wl.radii[segs[0][0]], wl.cores[segs[0][0]])
    brush.setPickupRate(wl.pickup_rate)
    brush.setDepositionRate(wl.deposition_rate)

    # algorithmic traffic of the stamps (exact host replay of the index math)
    visits = ring = 0
    for bid, arr, _ in segs:
        brush.setRadius(wl.radii[bid], wl.cores[bid])
        a, b = C.c_int64(), C.c_int64()
        _lib.call("pnt_brush_stamp_traffic", brush.h, wl.rows, wl.cols, arr, len(arr),
                  C.byref(a), C.byref(b))
        visits += a.value
        ring += b.value
    n_dry = sum(1 for s_ in segs if s_[2])
    stamp_bytes = BYTES_VISIT * visits + BYTES_RING * ring
    log("bench: %d segments, %d active visits, %d ring pixels, %d dries per step" %
        (len(segs), visits, ring, n_dry))
    seg_events = [(ev.event(), ev.event()) for _ in segs]
    one_call = args.mode == "fused" and not args.per_stroke
    strokes = (_lib.PntStroke * len(segs))()
    keep = []
    for k, (bid, arr, dry_after) in enumerate(segs):
        st2 = (_lib.PntStamp * len(arr))()
        C.memmove(st2, arr, C.sizeof(arr))
        if dry_after:
            st2[len(arr) - 1].flags |= W.DRY_AFTER
        core = np.ascontiguousarray(wl.cores[bid], dtype=np.float32)
        keep.append((st2, core))
        strokes[k].radius = float(wl.radii[bid])
        strokes[k].footprint = core.ctypes.data_as(C.POINTER(C.c_float))
        strokes[k].width = int(core.shape[0])
        strokes[k].stamps = C.cast(st2, C.POINTER(_lib.PntStamp))
        strokes[k].n_stamps = len(arr)
    call_events = (ev.event(), ev.event())

    def step():
        if one_call:
            ev.record(call_events[0], stream)
            _lib.call("pnt_imprint_strokes", brush.h, cv.h, strokes, len(segs), mode)
            ev.record(call_events[1], stream)
            _lib.call("pnt_dev_compose_wet", ctx.h, n_px, dev_layer, dev_r0, dev_out, None)
            return
        for k, (bid, arr, dry_after) in enumerate(segs):
            brush.setRadius(wl.radii[bid], wl.cores[bid])
            ev.record(seg_events[k][0], stream)
            _lib.call("pnt_imprint_stamps", brush.h, cv.h, arr, len(arr), mode)
            ev.record(seg_events[k][1], stream)
            if dry_after:
                _lib.call("pnt_canvas_dry", cv.h)
        _lib.call("pnt_dev_compose_wet", ctx.h, n_px, dev_layer, dev_r0, dev_out, None)

    def stamp_ms():
        if one_call:  # the dries run inside the call
            return ev.elapsed_ms(*call_events)
        return sum(ev.elapsed_ms(a, b) for a, b in seg_events)

    for i in range(args.warmup):
        step()
        ctx.synchronize()
        log("bench: warmup step %d done" % i)
    e0, e1 = ev.event(), ev.event()
    ctx.synchronize()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    st_ms = 0.0
    for i in range(args.steps):
        step()
        if i + 1 < args.steps:
            ctx.synchronize()  # read this step's stamp events before they are reused
        st_ms += stamp_ms()
    ev.record(e1, stream)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev.elapsed_ms(e0, e1)
    ms_per_step = wall * 1e3 / args.steps
    st_ms /= args.steps
    log("bench: %d steps, %.2f ms/step wall, %.2f ms/step stream, stamp calls %.2f ms/step" %
        (args.steps, ms_per_step, gpu_ms / args.steps, st_ms))
    launches = None
    if one_call:  # stamp-kernel launches of one step's stroke-list call
        nl = C.c_int64(0)
        _lib.call("pnt_brush_last_launches", brush.h, C.byref(nl))
        launches = int(nl.value)

    # the stamp kernel's own launch durations: the same steps again with a
    # HIP event pair around every stamp-kernel launch on the brush's stream
    kern_ms, kern_n = None, None
    if one_call:
        _lib.call("pnt_brush_set_kernel_timing", brush.h, 1)
        tot, cnt = 0.0, 0
        for _ in range(args.steps):
            step()
            ms, nl2 = C.c_double(), C.c_int64()
            _lib.call("pnt_brush_kernel_time", brush.h, C.byref(ms), C.byref(nl2))
            tot += ms.value
            cnt += nl2.value
        _lib.call("pnt_brush_set_kernel_timing", brush.h, 0)
        ctx.synchronize()
        kern_ms, kern_n = tot, cnt
        log("bench: stamp kernel %d launches, %.3f us average launch duration" %
            (kern_n, kern_ms * 1e3 / max(kern_n, 1)))

    stamps_s = len(wl.stamps) / (ms_per_step * 1e-3)
    waves = one_call and os.environ.get("PAINTY_FUSED_SCHED", "waves") != "streams"
    tile = 32 if os.environ.get("PAINTY_FUSED_TILE", "16") == "32" else 16
    kernel = {"plain": "k_stamp_main+k_ring_refresh", "persistent": "k_persistent",
              "fused": ("k_stamp_wave<%d>" if waves else "k_stamp_fused<%d>") % tile,
              "chained": "k_stamp_chain<16>"}[args.mode]
    if kern_n:
        avg_us = kern_ms * 1e3 / kern_n
        per_launch = stamp_bytes / launches
        achieved = per_launch / (avg_us * 1e-6) / 1e9
        how = ("algorithmic bytes per launch / average launch duration (HIP event pair "
               "around each of %d launches on the kernel's stream)" % kern_n)
    else:
        avg_us = st_ms * 1e3 / launches if launches else None
        per_launch = stamp_bytes / launches if launches else None
        achieved = stamp_bytes / (st_ms * 1e-3) / 1e9
        how = "algorithmic stamp bytes / stream time of the stamp calls"
    roofline = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(kernel), "launches_per_step": launches,
                "algorithmic_bytes_per_launch": int(per_launch) if per_launch else None,
                "avg_launch_us": round(avg_us, 3) if avg_us else None,
                "algorithmic_bytes_per_step": int(stamp_bytes), "measured": how,
                "stamp_call_stream_ms_per_step": round(st_ms, 3),
                "stamp_call_gbs": round(stamp_bytes / (st_ms * 1e-3) / 1e9, 1)}
    res = {
        "value": round(stamps_s, 1), "unit": "stamps/s", "ms_per_step": round(ms_per_step, 3),
        "config": {"workload": "large single canvas %dx%d, %d stamps (footprint 129-257 px, "
                               "pickup on, dry every 500) + composite" % (wl.rows, wl.cols,
                                                                          len(wl.stamps)),
                   "mode": args.mode,
                   "call": "pnt_imprint_strokes (one call, dries inside)" if one_call
                           else "pnt_imprint_stamps per stroke segment",
                   "scheduler": ("waves" if waves else "streams") if one_call else None},
        "stamps_per_s_stamp_calls": round(len(wl.stamps) / (st_ms * 1e-3), 1),
        "roofline": roofline,
    }
    del cv, out, brush
    free_device()
    return res


# ------------------------------------------------------------------ batched --
def batch_tail(H, jobs, lens):
    """One more run of the batch as waves with a timing event after every
    wave (pnt_batch_profile): when each job's last wave ends, and how much of
    the batch runs in waves too small to fill the device."""
    import torch
    jf, jl, wb, we = H.batch_profile(jobs)
    end = we[jl]
    dt = np.diff(np.concatenate([[0.0], we]))
    resident = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    small = wb < resident
    k = int(np.argmax(end))
    out = {
        "waves": int(len(wb)), "profiled_ms": round(float(we[-1]), 3),
        "job_end_ms": {q: round(float(np.percentile(end, p)), 3)
                       for q, p in (("p50", 50), ("p90", 90), ("p99", 99), ("max", 100))},
        "after_p90_ms": round(float(we[-1] - np.percentile(end, 90)), 3),
        "workgroups_per_wave": {"p50": int(np.median(wb)), "min": int(wb.min()),
                                "max": int(wb.max())},
        "waves_below_one_round": int(small.sum()), "ms_in_waves_below_one_round":
            round(float(dt[small].sum()), 3), "one_round_workgroups": int(resident),
        "last_job": {"stamps": lens[k][0], "rows": lens[k][1], "cols": lens[k][2],
                     "footprint_px": lens[k][3], "waves": int(jl[k] - jf[k] + 1)},
        "measured": "HIP event after every wave launch (pnt_batch_profile), one eager run",
    }
    log("bench: batch tail %s" % json.dumps(out))
    return out


def leg_batched(args):
    """Batched configuration: 256 canvases (sides U{256..2048}, 50..500 stamps,
    footprints 32..128 px, pickup on) stamped and composited in one launch
    (autotuned, eager and as a captured-graph replay), against the same jobs
    issued one after another as plain per-canvas launches."""
    from painty_amd import _lib
    from painty_amd import hip as H
    from painty_amd import workload as W
    t0 = time.perf_counter()
    bw = W.batch256(n_canvas=args.canvases)
    log("bench: %s generated in %.1fs" % (bw.name, time.perf_counter() - t0))
    ctx = H.default_context()
    stream = ctx.stream
    ev = HipEvents()
    jobs, lens, visits, ring, n_px, n_st = [], [], 0, 0, 0, 0
    t0 = time.perf_counter()
    for wl in bw.jobs:
        cv = H.Canvas(wl.rows, wl.cols)
        cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
        cv.setR0(wl.R0)
        br = H.FootprintBrush(wl.radii[0], wl.cores[0])
        out = H.Canvas(wl.rows, wl.cols)
        arr = H.stamps_to_c(wl.stamps)
        a, b = C.c_int64(), C.c_int64()
        _lib.call("pnt_brush_stamp_traffic", br.h, wl.rows, wl.cols, arr, len(arr),
                  C.byref(a), C.byref(b))
        visits += a.value
        ring += b.value
        n_px += wl.rows * wl.cols
        n_st += len(arr)
        jobs.append((cv, br, arr, out))
        lens.append((len(arr), wl.rows, wl.cols, wl.cores[0].shape[0]))
        wl.wet = wl.dry = wl.R0 = None  # host copies no longer needed
    del bw
    log("bench: uploaded %d canvases, %.1f Mpx, %d stamps in %.1fs" %
        (len(jobs), n_px / 1e6, n_st, time.perf_counter() - t0))
    alg_bytes = BYTES_VISIT * visits + BYTES_RING * ring + BYTES_COMPOSE1 * n_px

    def run(mode):
        H.batch_run(jobs, mode=mode)

    tuned = None
    if args.batch_tile:
        H.batch_set_config(args.batch_tile, args.batch_group, args.batch_waves, args.batch_chunk)
        tuned = {"config": list(H.batch_get_config()), "ms": None, "fixed": True}
    for _ in range(args.warmup):
        run(_lib.BATCH_ONE_LAUNCH)
        ctx.synchronize()
    if not args.batch_tile and not args.no_tune:
        t0 = time.perf_counter()
        cfg, tms = H.batch_tune(jobs)
        tuned = {"config": list(cfg), "ms": round(tms, 3)}
        log("bench: tuned %s (%.2f ms) in %.1fs" % (cfg, tms, time.perf_counter() - t0))
    steps = max(1, args.batch_steps)

    def timed(step):
        a, b = ev.event(), ev.event()
        t0 = time.perf_counter()
        ev.record(a, stream)
        for _ in range(steps):
            step()
        ev.record(b, stream)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, ev.elapsed_ms(a, b) / steps

    eager_wall, eager_gpu = timed(lambda: run(_lib.BATCH_ONE_LAUNCH))
    log("bench: one launch %.2f ms/batch wall, %.2f ms/batch stream" % (eager_wall, eager_gpu))
    wall, gpu, path = eager_wall, eager_gpu, "eager one-launch"
    graph_ms = None
    if not args.no_graph:
        g = H.BatchGraph(jobs)
        g.launch()
        ctx.synchronize()
        gw, gg = timed(g.launch)
        g.close()
        graph_ms = round(gw, 3)
        log("bench: graph replay %.2f ms/batch wall, %.2f ms/batch stream" % (gw, gg))
        if gw < wall:
            wall, gpu, path = gw, gg, "captured graph replay"
    tail = batch_tail(H, jobs, lens) if H.batch_get_config()[2] else None
    plain = None
    if args.plain_steps:
        a2, b2 = ev.event(), ev.event()
        t0 = time.perf_counter()
        ev.record(a2, stream)
        for _ in range(args.plain_steps):
            run(_lib.BATCH_PLAIN)
        ev.record(b2, stream)
        ctx.synchronize()
        plain = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        log("bench: plain %.2f ms/batch" % plain)
    gbs = alg_bytes / (gpu * 1e-3) / 1e9
    tile, _, waves, chunk = H.batch_get_config()
    batch_kernel = "k_batch_wave" if waves else "k_batch<%d>" % tile
    del jobs
    free_device()
    return {
        "workload": "%d canvases (sides 256..2048), %d stamps (footprint 32-128 px, pickup "
                    "on) + composites, one launch" % (args.canvases, n_st),
        "ms_per_batch": round(wall, 3), "stamps_s": round(n_st / (wall * 1e-3), 1),
        "composite_gpixel_s_in_batch": round(n_px / (wall * 1e-3) / 1e9, 2),
        "path": path, "eager_ms": round(eager_wall, 3), "graph_ms": graph_ms, "steps": steps,
        "tuned": tuned,
        "plain_sequential_ms": round(plain, 2) if plain else None,
        "speedup_vs_plain": round(plain / wall, 2) if plain else None,
        "tail": tail,
        "roofline": {"bound": "hbm", "kernel": batch_kernel, "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(batch_kernel, "batch"),
                     "traffic_per": "batch",
                     "algorithmic_bytes_per_batch": int(alg_bytes),
                     "measured": "algorithmic bytes of the batch / stream time of one batch"
                                 + (" (every wave launch)" if waves else " (one launch)")},
    }


# ------------------------------------------------------------------- served --
def _serve_rate(srv, reqs, rate, seed, offset):
    """Open-loop Poisson arrivals of `reqs` at `rate` req/s; returns the
    latencies (ms) and the wall time."""
    import threading

    from painty_amd import _lib
    n = len(reqs)
    arr = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n))
    tickets = [None] * n
    lat = [None] * n
    err = [0]

    def waiter():
        for i in range(n):
            while tickets[i] is None:
                time.sleep(0.0002)
            try:
                lat[i] = srv.wait(tickets[i])
            except _lib.PaintyError:
                err[0] += 1
    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i, q in enumerate(reqs):
        while time.perf_counter() - t0 < arr[i]:
            time.sleep(0.0001)
        tickets[i] = srv.submit(q.canvas_id + offset, q.side, q.side, q.stamps, q.radius,
                                q.footprint, background=q.background)
        if q.last:  # the client ends the canvas after this request
            srv.release_canvas(q.canvas_id + offset)
    th.join()
    el = time.perf_counter() - t0
    ok = np.array([x for x in lat if x is not None])
    return {"requests": n, "errors": err[0], "req_s": round(len(ok) / el, 1),
            "p50_ms": round(float(np.percentile(ok, 50)), 2),
            "p99_ms": round(float(np.percentile(ok, 99)), 2)}


def _plain_requests(reqs):
    """The plain per-request path: each request alone -- canvas created on
    the background, PNT_EXEC_PLAIN stamps, composite, synchronised."""
    from painty_amd import _lib
    from painty_amd import hip as H
    ctx = H.default_context()
    lat = []
    t0 = time.perf_counter()
    for q in reqs:
        a = time.perf_counter()
        cv = H.Canvas(q.side, q.side)
        cv.setR0(np.full((q.side, q.side, 3), q.background[0], np.float32))
        br = H.FootprintBrush(q.radius, q.footprint)
        br.imprintStamps(q.stamps, cv, _lib.EXEC_PLAIN)
        H.Renderer.compose(cv)
        ctx.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
        del cv, br
    rps = len(reqs) / (time.perf_counter() - t0)
    return {"req_s": round(rps, 2), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "sample": len(reqs)}


def leg_served(args):
    """Served configuration: open-loop Poisson arrivals at each offered rate;
    a request = stamps + composite on a resident (or new) canvas.  Then the
    same with every canvas 2048x2048.  Each against the plain per-request
    path; `ratio_at_p99_20ms` = the req/s of the highest offered rate whose
    p99 is at most 20 ms, over the plain path's req/s."""
    from painty_amd import serve as SV
    from painty_amd import workload as W
    rates = [int(x) for x in args.rates.split(",")]
    rates2048 = [int(x) for x in args.rates2048.split(",")] if args.rates2048 else []
    srv = SV.Server(n_slots=args.slots, hbm_budget_bytes=int(args.hbm_gb * (1 << 30)))
    res = {"slots": args.slots, "hbm_budget_gb": args.hbm_gb, "per_rate": {}, "canvas2048": {}}
    seed = 5
    for rate in rates:
        n = max(50, int(rate * args.seconds))
        r = _serve_rate(srv, W.served_requests(n, seed), rate, seed, 1_000_000 * seed)
        seed += 1
        res["per_rate"][str(rate)] = r
        log("bench: served %d req/s offered: %s" % (rate, r))
    for rate in rates2048:
        n = max(50, int(rate * args.seconds))
        r = _serve_rate(srv, W.served_requests(n, seed, side=(2048, 2048)), rate, seed,
                        1_000_000 * seed)
        seed += 1
        res["canvas2048"][str(rate)] = r
        log("bench: served 2048x2048, %d req/s offered: %s" % (rate, r))
    res["resident_at_end"] = srv.stats()
    srv.close()
    free_device()
    res["plain"] = _plain_requests(W.served_requests(args.plain_requests, 99))
    log("bench: plain per-request path: %s" % res["plain"])
    if rates2048:
        res["plain2048"] = _plain_requests(W.served_requests(args.plain_requests, 98,
                                                             side=(2048, 2048)))
        log("bench: plain per-request path, 2048x2048: %s" % res["plain2048"])

    def best(per, plain):
        ok = [r for r in per.values() if r["p99_ms"] <= 20.0 and not r["errors"]]
        if not ok or not plain:
            return None
        top = max(ok, key=lambda r: r["req_s"])
        return round(top["req_s"] / plain["req_s"], 1)
    res["ratio_at_p99_20ms"] = best(res["per_rate"], res["plain"])
    if rates2048:
        res["ratio_at_p99_20ms_2048"] = best(res["canvas2048"], res["plain2048"])
    top = res["per_rate"][str(rates[-1])]
    res["speedup_vs_plain_at_top_rate"] = round(top["req_s"] / res["plain"]["req_s"], 1)
    return res


# --------------------------------------------------------------------- diff --
def leg_diff(args):
    """Differentiable configuration: 2048^2 canvas, 500 stamps (65 px
    Gaussian footprint, pickup on), 64 paints x 14 pigments; one step =
    forward (stamps + composite, checkpoints) + L2 loss gradient + backward
    (recompute + adjoints) for K, S, V, R0 and W."""
    import torch

    from painty_amd import diff as D
    from painty_amd import hip as H
    from painty_amd import workload as W
    t0 = time.perf_counter()
    dw = W.diff2048_500stamps()
    wl = dw.wl
    log("bench: %s generated in %.1fs" % (wl.name, time.perf_counter() - t0))
    dev = "cuda:0"
    pl = lambda a: torch.from_numpy(np.ascontiguousarray(a.transpose(2, 0, 1) if a.ndim == 3 else a)).to(dev)
    K, S, V, R0 = pl(wl.wet.K), pl(wl.wet.S), pl(wl.wet.V), pl(wl.R0)
    T = pl(dw.target)
    br = H.FootprintBrush(wl.radii[0], wl.cores[0])
    br.setPickupRate(wl.pickup_rate)
    run = D.GradRunner(wl.rows, wl.cols, br, wl.stamps, dw.paint_of, dw.palK, dw.palS,
                       dw.W.shape[0], seg_len=args.seg_len)
    for _ in range(max(1, args.warmup)):
        R = run.forward(K, S, V, R0, dw.W)
        torch.cuda.synchronize()
        run.backward(2.0 * (R - T))
    torch.cuda.synchronize()
    steps = max(1, args.diff_steps)
    fw = bw = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        R = run.forward(K, S, V, R0, dw.W)
        gR = 2.0 * (R - T)
        torch.cuda.synchronize()
        b = time.perf_counter()
        run.backward(gR)
        torch.cuda.synchronize()
        fw += b - a
        bw += time.perf_counter() - b
    wall = (time.perf_counter() - t0) * 1e3 / steps
    run.close()
    del br, K, S, V, R0, T
    torch.cuda.empty_cache()
    free_device()
    return {"workload": "%dx%d canvas, %d stamps (65 px footprint, pickup on), 64 paints x 14 "
                        "pigments, L2 loss; gradients of K, S, V, R0, W"
                        % (wl.rows, wl.cols, len(wl.stamps)),
            "segment": args.seg_len, "steps": steps, "step_ms": round(wall, 3),
            "forward_ms": round(fw * 1e3 / steps, 3), "backward_ms": round(bw * 1e3 / steps, 3),
            "stamps_s_fwd_bwd": round(len(wl.stamps) / (wall * 1e-3), 1)}


def bench_copy(args):
    """Calibration: PaintLayer::copyTo of a size x size layer (k_copy_layer,
    six FP32 planes and the FP64 V plane read and written: 32 B/px each
    way).  Gives the achievable copy bandwidth and, under a counter
    profiler, the FETCH_SIZE / WRITE_SIZE that a kernel of known traffic
    reports."""
    from painty_amd import _lib
    from painty_amd import hip as H
    n = args.size * args.size
    a = H.PaintLayer(args.size, args.size)
    b = H.PaintLayer(args.size, args.size)
    ctx = H.default_context()
    ev = HipEvents()
    reps = max(args.steps, 1)
    a.copyTo(b)
    ctx.synchronize()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, ctx.stream)
    for _ in range(reps):
        _lib.call("pnt_layer_copy_to", a.h, b.h)
    ev.record(e1, ctx.stream)
    ms = ev.elapsed_ms(e0, e1) / reps
    nbytes = 32 * n
    print(json.dumps({"calibration": "k_copy_layer", "pixels": n, "bytes_read": nbytes,
                      "bytes_written": nbytes, "ms": round(ms, 4),
                      "gb_s": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)


def bench_texture(args):
    """TextureBrush::paintStroke with smudging on the reference test's stroke
    (renderer/test/src/TextureBrushTest.cxx:16-36): a 768 x 1024 canvas,
    radius 40, the path (50, 250) -> (400, 250) -> (650, 250), over a canvas
    already holding paint (one stroke before the timed ones), with a
    synthetic 43 x 200 brush stroke sample.  One step is one stroke."""
    from painty_amd import hip as H
    rows, cols, radius = 768, 1024, 40.0
    rng = np.random.default_rng(7)
    ty, tx = np.mgrid[0:43, 0:200]
    tex = np.clip(np.sin(np.pi * ty / 42.0) * (0.6 + 0.4 * rng.random((43, 200))), 0, 1)
    path = np.array([[50.0, 250.0], [400.0, 250.0], [650.0, 250.0]])
    ctx = H.default_context()
    cv = H.Canvas(rows, cols)
    cv.clear()
    br = H.TextureBrush(tex)
    br.setRadius(radius)
    br.dip(((0.2, 0.3, 0.4), (0.1, 0.23, 0.14)))
    br.paintStroke(path + [0.0, 3.5], cv)
    ev = HipEvents()
    for _ in range(args.warmup):
        br.paintStroke(path, cv)
    ctx.synchronize()
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, ctx.stream)
    for _ in range(args.steps):
        br.paintStroke(path, cv)
    ev.record(e1, ctx.stream)
    ctx.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gpu_ms = ev.elapsed_ms(e0, e1) / args.steps
    # the GL brush (TextureBrushGpu) on the same stroke, smudging on as in the
    # reference's GpuTest.cxx:29-45
    gl = H.TextureBrushGpu()
    gl.setRadius(radius)
    gl.setTexture(tex)
    gl.enableSmudge(True)
    gl.dip(((0.2, 0.3, 0.4), (0.1, 0.23, 0.14)))
    for _ in range(max(args.warmup, 1)):
        gl.paintStroke(path, cv)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        gl.paintStroke(path, cv)
    ctx.synchronize()
    gl_ms = (time.perf_counter() - t1) * 1e3 / args.steps
    result = {"metric": "texture-brush strokes/s (smudge on)", "value": round(1e3 / wall_ms, 2),
              "unit": "strokes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "ms_per_step": round(wall_ms, 3), "stream_ms_per_step": round(gpu_ms, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic",
              "config": {"workload": "TextureBrushTest stroke: 768x1024 canvas, r=40, "
                                     "(50,250)->(400,250)->(650,250), smudge on"},
              "gl_brush_ms_per_stroke": round(gl_ms, 3)}
    if not args.no_cpu:
        from oracle import oracle as O
        oc = O.Canvas(rows, cols)
        ob = O.TextureBrush()
        ob.set_radius(radius)
        ob.set_thickness_map(tex)
        ob.dip([0.2, 0.3, 0.4], [0.1, 0.23, 0.14])
        ob.paint_stroke(path + [0.0, 3.5], oc)
        t0 = time.perf_counter()
        ob.paint_stroke(path, oc)
        t = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": round(1.0 / t, 3), "unit": "strokes/s", "cores": 1,
                                  "kind": "port", "sample": "one stroke of the same job, FP64 "
                                  "oracle (%.2fs)" % t}
    print(json.dumps(result), flush=True)




# --------------------------------------------------------------------- main --
def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["all", "single", "batched", "copy", "diff", "served",
                                         "texture"], default="all")
    ap.add_argument("--mode", choices=["plain", "persistent", "fused", "chained"], default="fused")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--stamps", type=int, default=2000)
    ap.add_argument("--per-stroke", action="store_true",
                    help="single, fused: one pnt_imprint_stamps call per stroke segment")
    ap.add_argument("--cpu-stamps", type=int, default=0, help="0: the whole job")
    ap.add_argument("--cpu-rows", type=int, default=256, help="one-thread composite sample rows")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seg-len", type=int, default=32)
    ap.add_argument("--diff-steps", type=int, default=3)
    ap.add_argument("--rates", default="200,500,1000,1250,1500,2000")
    ap.add_argument("--rates2048", default="500,1000,1250,1500,2000")
    ap.add_argument("--seconds", type=float, default=3.0)
    # 16 request slots: p99 <= 20 ms up to 1500 req/s on 2048x2048 canvases;
    # 32 saturate 4% higher but pass 20 ms from 1250 req/s on
    ap.add_argument("--slots", type=int, default=16,
                    help="served: request slots of the server (admitted requests at once)")
    ap.add_argument("--hbm-gb", type=float, default=96.0)
    ap.add_argument("--plain-requests", type=int, default=40)
    ap.add_argument("--canvases", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-tune", action="store_true", help="batched: keep the default launch shape")
    ap.add_argument("--no-graph", action="store_true", help="batched: skip the captured graph")
    ap.add_argument("--batch-tile", type=int, default=0, help="batched: fixed tile (16 or 32)")
    ap.add_argument("--batch-group", type=int, default=16,
                    help="batched: workgroups per job with --batch-tile")
    ap.add_argument("--batch-waves", type=int, default=0,
                    help="batched with --batch-tile: 1 runs the batch as waves")
    ap.add_argument("--batch-chunk", type=int, default=4,
                    help="batched waves: strokes of one canvas in flight")
    ap.add_argument("--skip", default="", help="comma list of legs to leave out of --config all "
                                               "(batched, served, diff, composite)")
    return ap.parse_args()


def header(args):
    return {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic"}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(canvases would shard one per GPU with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.config == "copy":
        return bench_copy(args)
    if args.config == "texture":
        return bench_texture(args)
    if args.config in ("batched", "served", "diff"):
        leg = {"batched": leg_batched, "served": leg_served, "diff": leg_diff}[args.config]
        print(json.dumps(dict(header(args), leg=args.config, **leg(args))), flush=True)
        return

    from painty_amd import workload as W
    t0 = time.perf_counter()
    wl = W.canvas4096_2000stamps(size=args.size, n_stamps=args.stamps)
    log("bench: workload %s generated in %.1fs (%d strokes)" % (wl.name, time.perf_counter() - t0,
                                                               len(wl.cores)))
    skip = set(x for x in args.skip.split(",") if x)
    result = header(args)
    result.update(leg_single(args, wl))
    if "composite" not in skip:
        result["composite"] = leg_composite(wl)
    if not args.no_cpu:
        threads = min(16, host_description()["nproc"] or 1)
        log("bench: cpu baseline (FP64 oracle): %s on one thread, composite on %d threads" %
            ("the whole job" if args.cpu_stamps <= 0 else "%d stamps" % args.cpu_stamps,
             threads))
        cb = cpu_baseline(wl, args.cpu_stamps, args.cpu_rows, threads)
        result["cpu_baseline"] = {
            "value": round(cb["stamps_s"], 3), "unit": "stamps/s", "cores": 1, "kind": "port",
            "sample": ("the whole job on one thread: %d stamps with their dries (%.1fs) + "
                       "composite of %dx%d (%.2fs), FP64 oracle" %
                       (cb["n_stamps"], cb["t_stamps_s"], wl.rows, wl.cols,
                        cb["t_final_composite_s"]))
                      if cb["whole"] else
                      ("first %d stamps of the job on the full canvas, no dries (%.1fs), one "
                       "thread, FP64 oracle" % (cb["n_stamps"], cb["t_stamps_s"])),
            "host": host_description(),
            "composite": {
                "input": "the canvas before stamping (paint in 70% of pixels per layer)",
                "gpixel_s_1_core": round(cb["composite_gpixel_s_1core"], 5),
                "sample_1_core": "%d rows x %d" % (cb["comp_rows"], wl.cols),
                "gpixel_s_all_cores": round(cb["composite_gpixel_s_all_cores"], 4),
                "two_layer_gpixel_s_all_cores": round(cb["composite2_gpixel_s_all_cores"], 4),
                "cores_all": cb["threads"]}}
    del wl
    free_device()
    for name, leg in (("batched", leg_batched), ("served", leg_served), ("diff", leg_diff)):
        if args.config != "all" or name in skip:
            continue
        try:
            result[name] = leg(args)
        except Exception as e:  # the headline line is still printed
            traceback.print_exc()
            result[name] = {"error": "%s: %s" % (type(e).__name__, e)}
        free_device()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
