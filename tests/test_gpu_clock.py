"""Canvas::checkDry inside imprint (Canvas.hxx:123-155, FootprintBrush.hxx:
86, 115) on the GPU -- the plain per-stamp path, the fused stamp lists
(stream order and concurrent strokes as waves), stroke lists and batches --
against the FP64 oracle's timed imprint: wet layer, substrate, height and
composite within 1e-5 relative per channel, time maps exactly.  The canvas
clock is PNT_CLOCK_MANUAL, so every imprint's time point is set by the
test."""
import numpy as np
import pytest

from painty_amd import _lib
from painty_amd import workload as W
from tests.parity import assert_close, compare_canvas
from tests.test_gpu_stamp import PA, PB

pytestmark = pytest.mark.gpu
NOW = 50_000_000  # microseconds

# (cx, cy, theta, paint or None): interior stamps, a re-dip, and stamps that
# straddle the top-left corner with fractional centres (canvas column/row 0
# visited twice)
STAMPS = [(11.3, 9.6, 0.4, PA), (13.1, 11.4, 0.45, None), (15.2, 12.9, 0.5, None),
          (0.6, 0.3, 1.2, PB), (1.7, 2.2, 0.9, None), (20.5, 18.25, 2.0, None),
          (22.0, 19.0, 0.0, None)]


def _pair(hip, oracle, rows, cols, radius, seed, snapshot, age_us, dry_ms):
    wl = W.small(rows, cols, 0, radius, seed, pickup=True, noisy=True)
    tmap = (NOW - np.broadcast_to(np.asarray(age_us, np.int64), (rows, cols))).copy()
    oc = oracle.Canvas.from_arrays(wl.wet.K.astype(np.float64), wl.wet.S.astype(np.float64),
                                   wl.wet.V.astype(np.float64), wl.R0.astype(np.float64))
    oc.tmap[...] = tmap
    ob = oracle.Brush(wl.radii[0], wl.cores[0].astype(np.float64))
    ob.set_rates(wl.pickup_rate, wl.deposition_rate)
    ob.set_use_snapshot(snapshot)
    gc = hip.Canvas(rows, cols)
    gc.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    gc.setR0(wl.R0)
    gc.setDryingTime(dry_ms)
    gc.setClock(_lib.CLOCK_MANUAL)
    gc.setTimeMap(tmap)
    gb = hip.FootprintBrush(wl.radii[0], wl.cores[0])
    gb.setPickupRate(wl.pickup_rate)
    gb.setDepositionRate(wl.deposition_rate)
    gb.setUseSnapshotBuffer(snapshot)
    return oc, ob, gc, gb


def _one(cx, cy, th, paint):
    """One stamp as a stamp-list entry (with its dip)."""
    st = np.zeros(1, dtype=W.STAMP_DTYPE)
    st["cx"], st["cy"], st["theta"] = cx, cy, th
    if paint is not None:
        st["K"][0], st["S"][0] = paint[0], paint[1]
        st["flags"][0] = W.DIP
    return st


def _run(oc, ob, gc, gb, stamps, dry_ms, step_us=1000, mode=_lib.EXEC_PLAIN):
    """Each stamp at its own time point (step_us apart): the oracle's timed
    imprint, and on the GPU one call per stamp -- pnt_imprint on the plain
    path, or a one-stamp list in `mode`."""
    for k, (cx, cy, th, paint) in enumerate(stamps):
        t = NOW + step_us * k
        if paint is not None:
            ob.dip(*paint)
        ob.imprint_timed(cx, cy, th, oc, t, dry_ms)
        gc.setTime(t)
        if mode == _lib.EXEC_PLAIN:
            if paint is not None:
                gb.dip(paint)
            gb.imprint((cx, cy), th, gc)
        else:
            gb.imprintStamps(_one(cx, cy, th, paint), gc, mode)


MODES = [_lib.EXEC_PLAIN, _lib.EXEC_FUSED]


@pytest.mark.parametrize("mode", MODES, ids=["plain", "fused"])
@pytest.mark.parametrize("snapshot", [True, False])
@pytest.mark.parametrize("age_ms,dry_ms", [(0, 0), (10, 20), (25, 20), (90, 20), (5000, 1),
                                           (3, 2)])
def test_timed_imprint_matches_oracle(hip, oracle, snapshot, age_ms, dry_ms, mode):
    oc, ob, gc, gb = _pair(hip, oracle, 30, 34, 5, 21, snapshot, age_ms * 1000, dry_ms)
    _run(oc, ob, gc, gb, STAMPS, dry_ms, mode=mode)
    tag = "clock%d/snap%d/age%d/dry%d" % (mode, snapshot, age_ms, dry_ms)
    compare_canvas(tag, gc, oc, pickup=(gb, ob))
    assert np.array_equal(gc.getTimeMap(), oc.tmap), tag + ".tmap"


def _wl_stamps(wl):
    return [(float(s["cx"]), float(s["cy"]), float(s["theta"]),
             (s["K"].astype(np.float64), s["S"].astype(np.float64))
             if s["flags"] & W.DIP else None) for s in wl.stamps]


@pytest.mark.parametrize("mode", MODES, ids=["plain", "fused"])
def test_timed_strokes_with_redrying(hip, oracle, mode):
    """Strokes 1 ms apart with a 2 ms drying time: pixels dry again and
    again under later stamps of the same stroke; half the canvas starts old."""
    rows, cols = 96, 112
    age = np.zeros((rows, cols), np.int64)
    age[:, : cols // 2] = 40_000
    oc, ob, gc, gb = _pair(hip, oracle, rows, cols, 8, 5, True, age, 2)
    wl = W.small(rows, cols, 40, 8, 5, pickup=True, noisy=True)
    _run(oc, ob, gc, gb, _wl_stamps(wl), 2, mode=mode)
    compare_canvas("clock%d/strokes" % mode, gc, oc, pickup=(gb, ob))
    assert np.array_equal(gc.getTimeMap(), oc.tmap)


@pytest.mark.parametrize("snapshot", [True, False])
@pytest.mark.parametrize("how", ["fused", "strokes", "batch"])
def test_timed_stamp_list_matches_oracle(hip, oracle, how, snapshot):
    """A whole stamp list in one call at one time point -- fused (concurrent
    strokes as waves), a stroke list with a geometry change and a dryCanvas
    inside, and a batch of two canvases -- against the oracle's timed imprint
    of every stamp at that time: some of the canvas is older than the drying
    time, so the first stamps over it dry it, and corners are straddled."""
    rows, cols = 90, 100
    age = np.zeros((rows, cols), np.int64)
    age[: rows // 2, :] = 60_000
    dry_ms = 20
    oc, ob, gc, gb = _pair(hip, oracle, rows, cols, 6, 17, snapshot, age, dry_ms)
    wl = W.small(rows, cols, 48, 6, 17, pickup=True, noisy=True, reach=0.35)
    st = wl.stamps.copy()
    st["cx"][5], st["cy"][5] = 0.4, 0.7  # straddles the top-left corner
    st["flags"][23] |= W.DRY_AFTER
    core2 = W.noisy_core(2 * 9 + 1, W.Stream(17, 41))
    for k, (cx, cy, th, paint) in enumerate(_wl_stamps(W.Workload(
            wl.name, rows, cols, wl.dry, wl.wet, wl.R0, st, wl.cores, wl.radii))):
        if how == "strokes" and k == 30:
            ob.set_radius(9.0, core2.astype(np.float64))
        if paint is not None:
            ob.dip(*paint)
        ob.imprint_timed(cx, cy, th, oc, NOW, dry_ms)
        if st["flags"][k] & W.DRY_AFTER:
            oc.dry()
            oc.tmap[...] = NOW
    gc.setTime(NOW)
    if how == "fused":
        gb.imprintStamps(st, gc, _lib.EXEC_FUSED)
    elif how == "strokes":
        gb.imprintStrokes([(wl.radii[0], wl.cores[0], st[:30]), (9.0, core2, st[30:])], gc)
    else:
        hip.batch_run([(gc, gb, st, None)])
    tag = "clocklist/%s/snap%d" % (how, snapshot)
    compare_canvas(tag, gc, oc, pickup=(gb, ob))
    assert np.array_equal(gc.getTimeMap(), oc.tmap), tag + ".tmap"


@pytest.mark.parametrize("mode", [_lib.EXEC_PERSISTENT, _lib.EXEC_FUSED, _lib.EXEC_CHAINED])
def test_clocked_stamp_lists_equal_plain_path(hip, oracle, mode):
    """With a clock on, a stamp list in any mode gives the plain path's result
    to the bit (every stamp of the list takes the canvas's current time
    point; the one-launch modes run fused)."""
    results = []
    for m in (_lib.EXEC_PLAIN, mode):
        _, _, gc, gb = _pair(hip, oracle, 40, 44, 5, 31, True, 30_000, 20)
        gc.setTime(NOW)
        wl = W.small(40, 44, 25, 5, 31, pickup=True, noisy=True)
        gb.imprintStamps(wl.stamps, gc, m)
        lay = gc.getPaintLayer()
        results.append([lay.getK_buffer(), lay.getV_buffer(), gc.getR0(), gc.get_h(),
                        gc.getTimeMap()])
    for a, b in zip(*results):
        assert np.array_equal(a, b)


def test_clock_clear_dry_and_system_time(hip, oracle):
    """Switching the clock on, clear and dryCanvas stamp every pixel with the
    current time; the system clock reads microseconds since the epoch."""
    import time
    gc = hip.Canvas(16, 20)
    t0 = int(time.time() * 1e6)
    gc.setClock(_lib.CLOCK_SYSTEM)
    t1 = int(time.time() * 1e6)
    tm = gc.getTimeMap()
    assert tm.min() >= t0 - 2000 and tm.max() <= t1 + 2000 and (tm == tm[0, 0]).all()
    gc.setClock(_lib.CLOCK_MANUAL)
    gc.setTime(123_456)
    gc.dryCanvas()
    assert (gc.getTimeMap() == 123_456).all()
    gc.setTime(777)
    gc.clear()
    assert (gc.getTimeMap() == 777).all()
    assert gc.getClock() == _lib.CLOCK_MANUAL
    gc.setClock(_lib.CLOCK_OFF)
    with pytest.raises(hip.InvalidArgument):
        gc.getTimeMap()


def test_default_drying_time_changes_no_paint(hip, oracle):
    """The reference's default drying time (15,000,000 ms) under the system
    clock: the same paint as with no clock."""
    results = []
    for clock in (_lib.CLOCK_OFF, _lib.CLOCK_SYSTEM):
        wl = W.small(40, 44, 20, 5, 41, pickup=True, noisy=True)
        gc = hip.Canvas(40, 44)
        gc.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
        gc.setR0(wl.R0)
        gc.setClock(clock)
        gb = hip.FootprintBrush(wl.radii[0], wl.cores[0])
        gb.imprintStamps(wl.stamps, gc, _lib.EXEC_PLAIN)
        lay = gc.getPaintLayer()
        results.append([lay.getK_buffer(), lay.getS_buffer(), lay.getV_buffer(), gc.getR0()])
    for a, b in zip(*results):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("waves", [0, 1])
def test_batch_with_clocked_canvas_equals_fused(hip, oracle, waves):
    """A batch with one clocked and one unclocked canvas (the batch runs as
    waves whatever its configuration): each canvas equals the same stamps as
    one fused call, to the bit."""
    saved = hip.batch_get_config()
    hip.batch_set_config(16, 4, waves, 2)
    try:
        res = []
        for batch in (True, False):
            jobs = []
            for clocked in (True, False):
                _, _, gc, gb = _pair(hip, oracle, 50, 60, 5, 33 + clocked, True, 30_000, 20)
                if not clocked:
                    gc.setClock(_lib.CLOCK_OFF)
                else:
                    gc.setTime(NOW)
                wl = W.small(50, 60, 30, 5, 33 + clocked, pickup=True, noisy=True)
                jobs.append((gc, gb, wl.stamps, None))
            if batch:
                hip.batch_run(jobs)
            else:
                for gc, gb, st, _ in jobs:
                    gb.imprintStamps(st, gc, _lib.EXEC_FUSED)
            out = []
            for gc, gb, _, _ in jobs:
                lay = gc.getPaintLayer()
                out += [lay.getK_buffer(), lay.getV_buffer(), gc.getR0(), gc.get_h()]
            out.append(jobs[0][0].getTimeMap())
            res.append(out)
        for a, b in zip(*res):
            assert np.array_equal(a, b)
    finally:
        hip.batch_set_config(*saved)


def test_dry_timer(hip, oracle):
    """The GPU render thread's drying timer (SbrRenderThread.cxx:37-41):
    dryStep(0.01) every 250 ms of canvas clock, applied before the next call
    that reads the canvas; against the oracle's DryingStep repeated."""
    rows, cols = 24, 32
    rng = np.random.default_rng(8)
    K = rng.uniform(0.05, 1.0, (rows, cols, 3)).astype(np.float32)
    S = rng.uniform(0.05, 1.0, (rows, cols, 3)).astype(np.float32)
    V = rng.uniform(0.0, 0.08, (rows, cols)).astype(np.float32)
    V[:, :4] = 0.0
    R0 = rng.uniform(0.2, 1.0, (rows, cols, 3)).astype(np.float32)
    oK, oS, oV, oR0 = (a.astype(np.float64) for a in (K, S, V, R0))
    gc = hip.Canvas(rows, cols)
    gc.getPaintLayer().setBuffers(K, S, V)
    gc.setR0(R0)
    gc.setClock(_lib.CLOCK_MANUAL)
    gc.setTime(NOW)
    gc.setDryTimer(250, 0.01)
    done = 0
    for t_ms, steps in [(1100, 4), (1300, 5), (1300, 5), (3_600_000, 14400)]:
        gc.setTime(NOW + t_ms * 1000)
        R = hip.Renderer.compose(gc)  # reads the canvas: the timer's steps run first
        for _ in range(min(steps, 20) - done):  # V <= 0.08 is gone after 8 steps
            oracle.dry_step(oK, oS, oV, oR0, 0.01)
        done = min(steps, 20)
        tag = "timer/%dms" % t_ms
        # V - k * 0.01 in FP32 (as DryingStep.compute.glsl computes it) loses
        # the low bits of what remains: an absolute bound of a few FP32 ulps
        np.testing.assert_allclose(gc.getPaintLayer().getV_buffer(), oV, rtol=1e-5, atol=5e-8,
                                   err_msg=tag + ".V")
        assert_close(tag + ".R0", gc.getR0(), oR0)
        assert_close(tag + ".R", R, oracle.compose(oK, oS, oV, oR0))
    assert (gc.getPaintLayer().getV_buffer() == 0).all()
    gc.getPaintLayer().setBuffers(V=V)  # paint again, then stop the timer: nothing dries
    before = gc.getR0()
    gc.setDryTimer(0)
    gc.setTime(NOW + 7_200_000_000)
    np.testing.assert_array_equal(gc.getR0(), before)
    np.testing.assert_array_equal(gc.getPaintLayer().getV_buffer(), V)
