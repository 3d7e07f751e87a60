"""Benchmark: KM composite + imprint stamping on one MI355X.

Workload (one step): the large single-canvas configuration -- a 4096x4096
FP32 canvas (K log-uniform [0.01,10], S log-uniform [0.1,5], thickness
U(0,0.5) with 30% zeros, substrate U(0.7,0.95)), 2000 imprint stamps along
seeded Bezier strokes at 2 px spacing with a per-stroke footprint of 129-257
px (Gaussian x value noise), pickup on, dryCanvas() after every 500th stamp,
then a full composite of the canvas.  All inputs are resident in HBM before
the timed region.  Synthetic data; see painty_amd/workload.py.

Prints one JSON line.  `value` is stamps/s of the whole job (stamps + drying
+ composite); the composite alone is reported as Gpixel/s beside it.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KM composite Gpixel/s + stamps/s; p50/p99 served latency vs plain GPU path"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes (FP32 SoA minimum traffic)
BYTES_COMPOSE1 = 52   # K3 S3 V R0_3 in, R_3 out
BYTES_DRY = 88        # 11 planes read, 11 written
BYTES_VISIT = 148     # canvas K,S,V rw 56 + snapshot K,S,V r + V w 32 + pickup rw 56 + footprint 4
BYTES_RING = 56       # snapshot refresh: 7 floats read + 7 written


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


def cpu_baseline(wl, n_stamps: int, comp_rows: int):
    """FP64 oracle (single thread, the reference's loops).  n_stamps = 0 runs
    the whole job as the GPU does (every stamp, its dries, then the composite
    of the whole canvas); otherwise a bounded sample: the first n_stamps on
    the full canvas without dries, and a composite of the first comp_rows
    rows (0: all)."""
    from oracle import oracle as O
    R0 = wl.R0.astype(np.float64)
    O.compose_onto(wl.dry.K, wl.dry.S, wl.dry.V, R0)
    cv = O.Canvas.from_arrays(wl.wet.K.astype(np.float64), wl.wet.S.astype(np.float64),
                              wl.wet.V.astype(np.float64), R0)
    whole = n_stamps <= 0 or n_stamps >= len(wl.stamps)
    sample = wl.stamps.copy() if whole else wl.stamps[:n_stamps].copy()
    if not whole:
        sample["flags"] &= ~np.uint32(2)  # no dry inside a partial sample
    n_stamps = len(sample)
    comp_rows = wl.rows if comp_rows <= 0 else min(comp_rows, wl.rows)
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
    n = comp_rows * wl.cols
    t0 = time.perf_counter()
    O.compose(cv.K[:comp_rows], cv.S[:comp_rows], cv.V[:comp_rows], cv.R0[:comp_rows])
    t_c = time.perf_counter() - t0
    # the whole job is timed like the GPU's (stamps, dries and composite);
    # a partial sample's rate is its stamps alone
    return {"stamps_s": n_stamps / (t_st + t_c) if whole else n_stamps / t_st,
            "composite_gpixel_s": n / t_c / 1e9, "t_stamps_s": t_st, "t_composite_s": t_c,
            "whole": whole, "n_stamps": n_stamps, "rows": comp_rows}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summaries
    (profiles/*.md lines `traffic_bytes_per_launch[<kernel>]: N`: FETCH_SIZE
    and WRITE_SIZE per dispatch, scaled by the copy-kernel calibration), or
    None when no summary covers it."""
    import glob
    import re
    pat = re.compile(r"traffic_bytes_per_launch\[(.+?)\]:\s*(\d+)")
    for path in sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                              "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.search(line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    return None


def bench_batched(args):
    """Batched configuration: N canvases (sides U{256..2048}, 50..500 stamps,
    footprints 32..128 px, pickup on) stamped and composited in one launch,
    against the same jobs issued one after another as plain launches."""
    from painty_amd import _lib
    from painty_amd import hip as H
    from painty_amd import workload as W
    t0 = time.perf_counter()
    bw = W.batch256(n_canvas=args.canvases)
    log("bench: %s generated in %.1fs" % (bw.name, time.perf_counter() - t0))
    ctx = H.default_context()
    stream = ctx.stream
    ev = HipEvents()
    jobs, visits, ring, n_px, n_st = [], 0, 0, 0, 0
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
        wl.wet = wl.dry = wl.R0 = None  # host copies no longer needed
    log("bench: uploaded %d canvases, %.1f Mpx, %d stamps in %.1fs" %
        (len(jobs), n_px / 1e6, n_st, time.perf_counter() - t0))
    alg_bytes = BYTES_VISIT * visits + BYTES_RING * ring + BYTES_COMPOSE1 * n_px

    def run(mode):
        H.batch_run(jobs, mode=mode)

    tuned = None
    if args.batch_tile:
        H.batch_set_config(args.batch_tile, args.batch_group)
        tuned = {"tile": args.batch_tile, "max_group": args.batch_group, "ms": None}
    for _ in range(args.warmup):
        run(_lib.BATCH_ONE_LAUNCH)
        ctx.synchronize()
    if not args.batch_tile and not args.no_tune:
        t0 = time.perf_counter()
        (tile, group), tms = H.batch_tune(jobs)
        tuned = {"tile": tile, "max_group": group, "ms": round(tms, 3)}
        log("bench: tuned tile %d, max_group %d (%.2f ms) in %.1fs" %
            (tile, group, tms, time.perf_counter() - t0))

    def timed(step):
        a, b = ev.event(), ev.event()
        t0 = time.perf_counter()
        ev.record(a, stream)
        for _ in range(args.steps):
            step()
        ev.record(b, stream)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, ev.elapsed_ms(a, b) / args.steps

    eager_wall, eager_gpu = timed(lambda: run(_lib.BATCH_ONE_LAUNCH))
    log("bench: one launch %.2f ms/step wall, %.2f ms/step stream" % (eager_wall, eager_gpu))
    wall, gpu, path = eager_wall, eager_gpu, "eager one-launch"
    graph_ms = None
    if not args.no_graph:
        g = H.BatchGraph(jobs)
        g.launch()
        ctx.synchronize()
        gw, gg = timed(g.launch)
        g.close()
        graph_ms = round(gw, 3)
        log("bench: graph replay %.2f ms/step wall, %.2f ms/step stream" % (gw, gg))
        if gw < wall:
            wall, gpu, path = gw, gg, "captured graph replay"
    a2, b2 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(a2, stream)
    for _ in range(args.plain_steps):
        run(_lib.BATCH_PLAIN)
    ev.record(b2, stream)
    ctx.synchronize()
    plain = (time.perf_counter() - t0) * 1e3 / max(1, args.plain_steps)
    log("bench: plain %.2f ms/step" % plain)
    gbs = alg_bytes / (gpu * 1e-3) / 1e9
    batch_kernel = "k_batch<%d>" % H.batch_get_config()[0]
    result = {
        "metric": METRIC, "value": round(n_st / (wall * 1e-3), 1), "unit": "stamps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: %d canvases (sides 256..2048), %d stamps "
                               "(footprint 32-128 px, pickup on) + composites, one launch"
                               % (len(jobs), n_st)},
        "composite_gpixel_s_in_batch": round(n_px / (wall * 1e-3) / 1e9, 2),
        "path": path, "eager_ms_per_step": round(eager_wall, 3), "graph_ms_per_step": graph_ms,
        "tuned": tuned,
        "plain_sequential_ms_per_step": round(plain, 2),
        "speedup_vs_plain": round(plain / wall, 2),
        "roofline": {"bound": "hbm", "kernel": batch_kernel, "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(batch_kernel),
                     "algorithmic_bytes_per_step": int(alg_bytes)},
    }
    print(json.dumps(result), flush=True)


def bench_diff(args):
    """Differentiable configuration: 2048^2 canvas, 500 stamps (65 px
    Gaussian footprint, pickup on), 64 paints x 14 pigments; one step =
    forward (stamps + composite, checkpoints) + L2 loss gradient + backward
    (recompute + adjoints) for K, S, V, R0 and W."""
    import torch

    from painty_amd import diff as D
    from painty_amd import hip as H
    from painty_amd import workload as W
    t0 = time.perf_counter()
    dw = W.diff2048_500stamps(size=args.size if args.size != 4096 else 2048, n_stamps=500)
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

    def step():
        R = run.forward(K, S, V, R0, dw.W)
        gR = 2.0 * (R - T)
        torch.cuda.synchronize()
        return run.backward(gR)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    fw = bw = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        R = run.forward(K, S, V, R0, dw.W)
        gR = 2.0 * (R - T)
        torch.cuda.synchronize()
        b = time.perf_counter()
        run.backward(gR)
        torch.cuda.synchronize()
        fw += b - a
        bw += time.perf_counter() - b
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    n_st = len(wl.stamps)
    result = {
        "metric": METRIC, "value": round(n_st / (wall * 1e-3), 1), "unit": "stamps/s (fwd+bwd)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiable: %dx%d canvas, %d stamps (65 px footprint, pickup "
                               "on), 64 paints x 14 pigments, L2 loss; gradients of K, S, V, R0, W"
                               % (wl.rows, wl.cols, n_st), "segment": args.seg_len},
        "forward_ms": round(fw * 1e3 / args.steps, 3),
        "backward_ms": round(bw * 1e3 / args.steps, 3),
    }
    print(json.dumps(result), flush=True)


def served_requests(n, seed, side=(512, 4096), stamps=(10, 1000), foot=(32, 128),
                    follow=0.3, recent=32):
    """The served configuration's request stream: square canvases with side
    log-uniform in `side`, stamps U{stamps}, footprint diameter U{foot},
    `follow` of the requests going to one of the `recent` last new canvases."""
    from painty_amd import workload as W
    rs = W.Stream(seed, 95)
    pool = [W.strokes(1000, 1000, stamps[1], seed * 1000 + k, stamps_per_stroke=(20, 80),
                      reach=0.15) for k in range(16)]
    cores = {}
    reqs, created = [], []
    for i in range(n):
        if created and rs.uniform(1)[0] < follow:
            cid, s_ = created[int(rs.integers(1, max(0, len(created) - recent), len(created) - 1)[0])]
        else:
            s_ = int(round(math.exp(rs.uniform(1, math.log(side[0]), math.log(side[1]))[0])))
            cid = len(created)
            created.append((cid, s_))
        n_st = int(rs.integers(1, stamps[0], stamps[1])[0])
        st = pool[i % len(pool)][:n_st].copy()
        st["cx"] *= s_ / 1000.0
        st["cy"] *= s_ / 1000.0
        st["flags"][0] |= W.DIP
        d = int(rs.integers(1, foot[0], foot[1])[0])
        r = max(1, d // 2)
        if r not in cores:
            cores[r] = W.gaussian_core(2 * r + 1, (2 * r + 1) / 4.0)
        reqs.append(dict(canvas_id=cid, rows=s_, cols=s_, stamps=st, radius=float(r),
                         footprint=cores[r], background=(0.85, 0.85, 0.85)))
    return reqs


def bench_served(args):
    """Served configuration: open-loop Poisson arrivals at each rate; request =
    stamps + composite on a resident (or new) canvas; p50/p99 latency and
    requests/s, against the plain per-request path (each request alone:
    canvas created, PNT_EXEC_PLAIN stamps, composite, synchronised)."""
    import threading

    from painty_amd import _lib
    from painty_amd import hip as H
    from painty_amd import serve as SV
    rates = [int(x) for x in args.rates.split(",")]
    srv = SV.Server(n_slots=args.slots, hbm_budget_bytes=int(args.hbm_gb * (1 << 30)))
    per_rate = {}
    seed = 5
    for rate in rates:
        n = max(50, int(rate * args.seconds))
        reqs = served_requests(n, seed)
        seed += 1
        offset = 1_000_000 * seed  # fresh canvas ids per rate
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
            q = dict(q, canvas_id=q["canvas_id"] + offset)
            tickets[i] = srv.submit(**q)
        th.join()
        el = time.perf_counter() - t0
        ok = np.array([x for x in lat if x is not None])
        per_rate[rate] = {"requests": n, "errors": err[0], "req_s": round(len(ok) / el, 1),
                          "p50_ms": round(float(np.percentile(ok, 50)), 2),
                          "p99_ms": round(float(np.percentile(ok, 99)), 2)}
        log("bench: served %d req/s offered: %s" % (rate, per_rate[rate]))
    st = srv.stats()
    srv.close()
    # plain per-request path on a sample of the same stream
    sample = served_requests(args.plain_requests, 99)
    ctx = H.default_context()
    t0 = time.perf_counter()
    lat_plain = []
    for q in sample:
        a = time.perf_counter()
        cv = H.Canvas(q["rows"], q["cols"])
        cv.setR0(np.full((q["rows"], q["cols"], 3), 0.85, np.float32))
        br = H.FootprintBrush(q["radius"], q["footprint"])
        br.imprintStamps(q["stamps"], cv, _lib.EXEC_PLAIN)
        H.Renderer.compose(cv)
        ctx.synchronize()
        lat_plain.append((time.perf_counter() - a) * 1e3)
        del cv, br
    plain_rps = len(sample) / (time.perf_counter() - t0)
    top = per_rate[rates[-1]]
    result = {
        "metric": METRIC, "value": top["req_s"], "unit": "req/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "served: Poisson arrivals, square canvases 512..4096 "
                               "(log-uniform side), 10..1000 stamps (footprint 32-128), "
                               "composite per request, 30% follow-ups to resident canvases",
                   "slots": args.slots, "hbm_budget_gb": args.hbm_gb},
        "per_rate": per_rate,
        "plain_per_request": {"req_s": round(plain_rps, 2),
                              "p50_ms": round(float(np.percentile(lat_plain, 50)), 2),
                              "p99_ms": round(float(np.percentile(lat_plain, 99)), 2),
                              "sample": len(sample)},
        "speedup_vs_plain_at_top_rate": round(top["req_s"] / plain_rps, 1),
        "resident_at_end": st,
    }
    print(json.dumps(result), flush=True)


def bench_copy(args):
    """Calibration: PaintLayer::copyTo of a size x size layer (k_copy_layer,
    7 FP32 planes read and written).  Gives the achievable copy bandwidth and,
    under a counter profiler, the FETCH_SIZE / WRITE_SIZE that a kernel of
    known traffic reports."""
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
    nbytes = 28 * n
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["plain", "persistent", "fused", "chained"], default="fused")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--stamps", type=int, default=2000)
    ap.add_argument("--cpu-stamps", type=int, default=0, help="0: the whole job")
    ap.add_argument("--cpu-rows", type=int, default=0, help="0: the whole canvas")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", choices=["single", "batched", "copy", "diff", "served",
                                         "texture"],
                    default="single")
    ap.add_argument("--seg-len", type=int, default=32)
    ap.add_argument("--rates", default="200,500,1000,2000")
    ap.add_argument("--seconds", type=float, default=3.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--hbm-gb", type=float, default=128.0)
    ap.add_argument("--plain-requests", type=int, default=40)
    ap.add_argument("--canvases", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-tune", action="store_true", help="batched: keep the default launch shape")
    ap.add_argument("--per-stroke", action="store_true",
                    help="single, fused: one pnt_imprint_stamps call per stroke segment")
    ap.add_argument("--no-graph", action="store_true", help="batched: skip the captured graph")
    ap.add_argument("--batch-tile", type=int, default=0, help="batched: fixed tile (16 or 32)")
    ap.add_argument("--batch-group", type=int, default=16,
                    help="batched: workgroups per job with --batch-tile")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(canvases would shard one per GPU with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.config == "batched":
        return bench_batched(args)
    if args.config == "copy":
        return bench_copy(args)
    if args.config == "diff":
        return bench_diff(args)
    if args.config == "served":
        return bench_served(args)
    if args.config == "texture":
        return bench_texture(args)

    from painty_amd import _lib
    from painty_amd import hip as H
    from painty_amd import workload as W

    t0 = time.perf_counter()
    wl = W.canvas4096_2000stamps(size=args.size, n_stamps=args.stamps)
    log("bench: workload %s generated in %.1fs (%d strokes)" % (wl.name, time.perf_counter() - t0,
                                                               len(wl.cores)))
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
    dev_layer = _lib.DevLayer((C.c_void_p * 7)(*Lc))
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
    # fused: the whole job as one stroke list (per-stroke footprints; strokes
    # whose snapshot boxes do not meet run concurrently), dries included
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
            _lib.call("pnt_dev_compose", ctx.h, n_px, dev_layer, dev_r0, dev_out, None)
            return
        for k, (bid, arr, dry_after) in enumerate(segs):
            brush.setRadius(wl.radii[bid], wl.cores[bid])
            ev.record(seg_events[k][0], stream)
            _lib.call("pnt_imprint_stamps", brush.h, cv.h, arr, len(arr), mode)
            ev.record(seg_events[k][1], stream)
            if dry_after:
                _lib.call("pnt_canvas_dry", cv.h)
        _lib.call("pnt_dev_compose", ctx.h, n_px, dev_layer, dev_r0, dev_out, None)

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
    log("bench: %d steps, %.2f ms/step wall, %.2f ms/step stream, stamp launches %.2f ms/step" %
        (args.steps, ms_per_step, gpu_ms / args.steps, st_ms))

    # composite kernel alone (the KM composite metric), live HIP events
    reps = 20
    a, b = ev.event(), ev.event()
    ev.record(a, stream)
    for _ in range(reps):
        _lib.call("pnt_dev_compose", ctx.h, n_px, dev_layer, dev_r0, dev_out, None)
    ev.record(b, stream)
    comp_ms = ev.elapsed_ms(a, b) / reps
    comp_gpx = n_px / (comp_ms * 1e-3) / 1e9
    comp_gbs = BYTES_COMPOSE1 * n_px / (comp_ms * 1e-3) / 1e9
    log("bench: composite %.3f ms, %.1f Gpixel/s, %.0f GB/s" % (comp_ms, comp_gpx, comp_gbs))

    stamps_s = args.stamps / (ms_per_step * 1e-3)
    stamp_gbs = stamp_bytes / (st_ms * 1e-3) / 1e9
    waves = one_call and os.environ.get("PAINTY_FUSED_SCHED", "waves") != "streams"
    tile = 32 if os.environ.get("PAINTY_FUSED_TILE", "16") == "32" else 16
    kernel = {"plain": "k_stamp_main+k_ring_refresh", "persistent": "k_persistent",
              "fused": ("k_stamp_wave<%d>" if waves else "k_stamp_fused<%d>") % tile,
              "chained": "k_stamp_chain<16>"}[args.mode]
    launches = None
    if one_call:  # stamp-kernel launches of one step's stroke-list call
        nl = C.c_int64(0)
        _lib.call("pnt_brush_last_launches", brush.h, C.byref(nl))
        launches = int(nl.value)
    result = {
        "metric": METRIC, "value": round(stamps_s, 1), "unit": "stamps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large single canvas %dx%d, %d stamps (footprint 129-257 px, "
                               "pickup on, dry every 500) + composite" % (wl.rows, wl.cols,
                                                                          args.stamps),
                   "mode": args.mode,
                   "call": "pnt_imprint_strokes (one call, dries inside)" if one_call
                           else "pnt_imprint_stamps per stroke segment"},
        "stamps_per_s_kernel": round(args.stamps / (st_ms * 1e-3), 1),
        "composite_gpixel_s": round(comp_gpx, 2),
        "composite_ms": round(comp_ms, 4),
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": round(stamp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(stamp_gbs / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(kernel),
                     "launches_per_step": launches,
                     "algorithmic_bytes_per_launch": int(stamp_bytes / launches)
                     if launches else None,
                     "avg_launch_us": round(st_ms * 1e3 / launches, 3) if launches else None,
                     "algorithmic_bytes_per_step": int(stamp_bytes),
                     "stamp_time": "stream time of the stamp calls"
                                   + (" (incl. %d dries and the gaps between launches; "
                                      "avg_launch_us = that time / launches)" % n_dry
                                      if one_call else ""),
                     "concurrent_chains": int(os.environ.get("PAINTY_FUSED_CHAINS",
                                                             "8" if waves else "4"))
                     if one_call else 1,
                     "scheduler": ("waves" if waves else "streams") if one_call else None},
        "roofline_composite": {"bound": "hbm", "kernel": "k_compose_vec<1>",
                               "achieved": round(comp_gbs, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(comp_gbs / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu:
        log("bench: cpu baseline (oracle, 1 thread): %s" %
            ("the whole job" if args.cpu_stamps <= 0 else
             "%d stamps + %d-row composite" % (args.cpu_stamps, args.cpu_rows)))
        cb = cpu_baseline(wl, args.cpu_stamps, args.cpu_rows)
        result["cpu_baseline"] = {
            "value": round(cb["stamps_s"], 3), "unit": "stamps/s", "cores": 1, "kind": "port",
            "composite_gpixel_s": round(cb["composite_gpixel_s"], 5),
            "sample": ("the whole job: %d stamps with their %d dries (%.1fs) + composite of "
                       "%dx%d (%.1fs), FP64 oracle" % (cb["n_stamps"], n_dry, cb["t_stamps_s"],
                                                       cb["rows"], wl.cols, cb["t_composite_s"]))
                      if cb["whole"] else
                      ("first %d stamps of the job on the full canvas, no dries (%.1fs) + "
                       "composite of %dx%d (%.1fs), FP64 oracle" %
                       (cb["n_stamps"], cb["t_stamps_s"], cb["rows"], wl.cols,
                        cb["t_composite_s"]))}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
