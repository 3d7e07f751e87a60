"""The planned stamp paths -- one persistent launch per list
(PNT_EXEC_PERSISTENT), one fused launch per stamp (PNT_EXEC_FUSED) and one
counter-chained launch per list (PNT_EXEC_CHAINED) --
against the FP64 oracle, and the bookkeeping that lets them skip no-op
snapshot copies: every way the canvas or the snapshot can change behind
their back must make them copy again.  Tolerance 1e-5 relative per
channel."""
import math

import numpy as np
import pytest

from painty_amd import workload as W
from tests.parity import compare_canvas, run_gpu, run_oracle
from tests.test_gpu_stamp import HAND, PA, PB, _stamp_array

pytestmark = pytest.mark.gpu
PERSIST = 1
PLAIN = 0
FUSED = 2
CHAINED = 3


@pytest.fixture(params=[PERSIST, FUSED, CHAINED], ids=["persistent", "fused", "chained"])
def lmode(request):
    return request.param


@pytest.mark.parametrize("case", sorted(HAND))
@pytest.mark.parametrize("snapshot", [True, False])
@pytest.mark.parametrize("radius", [4, 9, 20])
def test_hand_cases(hip, oracle, case, snapshot, radius, lmode):
    wl = W.small(28, 24, 0, radius, 7)
    st = _stamp_array(HAND[case])
    ocv, ob = run_oracle(wl, st, use_snapshot=snapshot)
    gcv, gb = run_gpu(wl, st, use_snapshot=snapshot, mode=lmode)
    compare_canvas("%s/" % lmode + "%s/snap%d/r%d" % (case, snapshot, radius), gcv, ocv, pickup=(gb, ob))


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("radius,noisy", [(3, False), (8, True), (16, False), (40, True)])
def test_seeded_strokes(hip, oracle, seed, radius, noisy, lmode):
    wl = W.small(96, 112, 40, radius, seed, pickup=True, noisy=noisy)
    ocv, ob = run_oracle(wl)
    gcv, gb = run_gpu(wl, mode=lmode)
    compare_canvas("%s/" % lmode + "strokes/s%d/r%d" % (seed, radius), gcv, ocv, pickup=(gb, ob))


def test_canvas1024_pickup(hip, oracle, lmode):
    wl = W.canvas1024_200stamps(pickup=True)
    ocv, ob = run_oracle(wl)
    gcv, gb = run_gpu(wl, mode=lmode)
    compare_canvas("%s/" % lmode + "canvas1024", gcv, ocv, pickup=(gb, ob))


def test_multi_brush_with_drying(hip, oracle, lmode):
    wl = W.canvas4096_2000stamps(seed=4, size=512, n_stamps=150, dry_every=50, foot=(41, 129))
    ocv, ob = run_oracle(wl)
    gcv, gb = run_gpu(wl, mode=lmode)
    compare_canvas("%s/" % lmode + "canvas-small", gcv, ocv, pickup=(gb, ob))


def _pair(rows, cols, radius, seed):
    wl = W.small(rows, cols, 0, radius, seed)
    core = wl.cores[0]
    ob = oracle_brush(wl)
    gb = hip_brush(wl)
    from oracle import oracle as O
    ocv = O.Canvas.from_arrays(wl.wet.K.astype(np.float64), wl.wet.S.astype(np.float64),
                               wl.wet.V.astype(np.float64), wl.R0.astype(np.float64))
    from painty_amd import hip as H
    gcv = H.Canvas(rows, cols)
    gcv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    gcv.setR0(wl.R0)
    return wl, core, ob, gb, ocv, gcv


def oracle_brush(wl):
    from oracle import oracle as O
    b = O.Brush(wl.radii[0], wl.cores[0].astype(np.float64))
    b.set_rates(wl.pickup_rate, wl.deposition_rate)
    return b


def hip_brush(wl):
    from painty_amd import hip as H
    return H.FootprintBrush(wl.radii[0], wl.cores[0])


def _stroke(x0, y0, n, dx=1.7, dy=0.9, th=0.4, paint=PA):
    return _stamp_array([(x0 + k * dx, y0 + k * dy, th + 0.05 * k, paint if k == 0 else None, 0)
                         for k in range(n)])


def _run(ob, gb, ocv, gcv, st, mode):
    ob.run(ocv, st)
    gb.imprintStamps(st, gcv, mode)


def test_mode_mixing_and_external_writes(hip, oracle, lmode):
    wl, core, ob, gb, ocv, gcv = _pair(60, 70, 6, 3)
    _run(ob, gb, ocv, gcv, _stroke(10, 12, 8), lmode)
    _run(ob, gb, ocv, gcv, _stroke(24, 20, 6, paint=PB), PLAIN)      # plain in between
    _run(ob, gb, ocv, gcv, _stroke(30, 25, 8, paint=None), lmode)
    # external write of the wet layer (host upload) between persistent runs
    K = gcv.getPaintLayer().getK_buffer()
    K[20:40, 20:40] *= 0.5
    gcv.getPaintLayer().setBuffers(K=K)
    ocv.K[...] = K.astype(np.float64)
    _run(ob, gb, ocv, gcv, _stroke(25, 22, 10, dx=-1.1, paint=None), lmode)
    # dryCanvas through the API, then more stamps
    ocv.dry()
    gcv.dryCanvas()
    _run(ob, gb, ocv, gcv, _stroke(15, 30, 10, paint=PB), lmode)
    compare_canvas("%s/" % lmode + "mixing", gcv, ocv, pickup=(gb, ob))


def test_two_brushes_one_canvas(hip, oracle, lmode):
    wl, core, ob1, gb1, ocv, gcv = _pair(64, 64, 5, 4)
    ob2 = oracle.Brush(9.0, W.gaussian_core(19, 4.0).astype(np.float64))
    from painty_amd import hip as H
    gb2 = H.FootprintBrush(9.0, W.gaussian_core(19, 4.0))
    for k in range(3):
        _run(ob1, gb1, ocv, gcv, _stroke(10 + 5 * k, 10 + 3 * k, 7, paint=PA if k == 0 else None),
             lmode)
        _run(ob2, gb2, ocv, gcv, _stroke(30 - 4 * k, 12 + 5 * k, 7, dx=-1.3,
                                         paint=PB if k == 0 else None), lmode)
    compare_canvas("%s/" % lmode + "two-brushes", gcv, ocv, pickup=(gb1, ob1))


def test_brush_on_two_canvases_same_size(hip, oracle, lmode):
    # the snapshot is not re-copied for a canvas of the same size (the
    # reference keeps the stale snapshot, FootprintBrush.hxx:281-284)
    wl, core, ob, gb, ocv, gcv = _pair(50, 50, 5, 5)
    from oracle import oracle as O
    from painty_amd import hip as H
    wl2 = W.small(50, 50, 0, 5, 6)
    ocv2 = O.Canvas.from_arrays(wl2.wet.K.astype(np.float64), wl2.wet.S.astype(np.float64),
                                wl2.wet.V.astype(np.float64), wl2.R0.astype(np.float64))
    gcv2 = H.Canvas(50, 50)
    gcv2.getPaintLayer().setBuffers(wl2.wet.K, wl2.wet.S, wl2.wet.V)
    gcv2.setR0(wl2.R0)
    _run(ob, gb, ocv, gcv, _stroke(10, 10, 9), lmode)
    _run(ob, gb, ocv2, gcv2, _stroke(12, 11, 9, paint=None), lmode)
    _run(ob, gb, ocv, gcv, _stroke(20, 15, 9, paint=None), lmode)
    compare_canvas("%s/" % lmode + "canvas1", gcv, ocv, pickup=(gb, ob))
    compare_canvas("%s/" % lmode + "canvas2", gcv2, ocv2)


def test_planned_paths_equal_plain_bitwise_on_one_stroke(hip, lmode):
    """All GPU paths run the same FP32 visit code; with one stroke of stamps
    (no shared pixels) they agree to the last bit."""
    wl = W.small(80, 80, 30, 7, 9, pickup=True, noisy=True, reach=0.2)
    a, _ = run_gpu(wl, mode=PLAIN)
    b, _ = run_gpu(wl, mode=lmode)
    for get in ("getK_buffer", "getS_buffer", "getV_buffer"):
        x = getattr(a.getPaintLayer(), get)()
        y = getattr(b.getPaintLayer(), get)()
        assert np.array_equal(x, y), get


@pytest.mark.parametrize("sched", ["waves", "streams"])
@pytest.mark.parametrize("chains", [2, 4, 8])
def test_fused_concurrent_strokes_equal_one_chain(hip, oracle, chains, sched, monkeypatch):
    """Strokes with disjoint snapshot boxes run side by side (as waves of one
    stream, or on concurrent streams), each with its own pickup map: the
    result equals the one-chain run to the bit (and the oracle), with drying
    between phases and the last stroke's pickup map left in the brush."""
    wl = W.canvas4096_2000stamps(seed=7, size=1024, n_stamps=400, dry_every=130, foot=(17, 41))
    monkeypatch.setenv("PAINTY_FUSED_SCHED", sched)
    monkeypatch.setenv("PAINTY_FUSED_CHAINS", "1")
    a, ab = run_gpu(wl, mode=FUSED)
    monkeypatch.setenv("PAINTY_FUSED_CHAINS", str(chains))
    b, bb = run_gpu(wl, mode=FUSED)
    for get in ("getK_buffer", "getS_buffer", "getV_buffer"):
        assert np.array_equal(getattr(a.getPaintLayer(), get)(), getattr(b.getPaintLayer(), get)()), get
    assert np.array_equal(a.getR0(), b.getR0()) and np.array_equal(a.get_h(), b.get_h())
    for x, y in zip(ab.getPickupMap(), bb.getPickupMap()):
        assert np.array_equal(x, y)
    ocv, ob = run_oracle(wl)
    compare_canvas("fused%d" % chains, b, ocv, pickup=(bb, ob))


def _strokes(wl):
    from tests.parity import _groups
    return [(wl.radii[bid], wl.cores[bid], g) for bid, g in _groups(wl.stamps)]


def _fresh(hip, wl):
    from painty_amd import hip as H
    cv = H.Canvas(wl.rows, wl.cols)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    dry = H.PaintLayer(wl.rows, wl.cols)
    dry.setBuffers(wl.dry.K, wl.dry.S, wl.dry.V)
    dry.composeOntoCanvas(cv)
    br = H.FootprintBrush(wl.radii[0], wl.cores[0])
    br.setPickupRate(wl.pickup_rate)
    br.setDepositionRate(wl.deposition_rate)
    return cv, br


def _same(a, ab, b, bb):
    for get in ("getK_buffer", "getS_buffer", "getV_buffer"):
        assert np.array_equal(getattr(a.getPaintLayer(), get)(), getattr(b.getPaintLayer(), get)()), get
    assert np.array_equal(a.getR0(), b.getR0()) and np.array_equal(a.get_h(), b.get_h())
    assert ab.getSizeMap() == bb.getSizeMap()
    for x, y in zip(ab.getPickupMap(), bb.getPickupMap()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("sched", ["waves", "streams"])
@pytest.mark.parametrize("chains", [1, 4])
def test_imprint_strokes_equals_stroke_by_stroke(hip, oracle, chains, sched, monkeypatch):
    """pnt_imprint_strokes (one call, per-stroke footprints, concurrent
    strokes) equals setRadius + imprintStamps stroke by stroke, to the bit,
    and the oracle."""
    wl = W.canvas4096_2000stamps(seed=8, size=1024, n_stamps=400, dry_every=130, foot=(17, 41))
    monkeypatch.setenv("PAINTY_FUSED_SCHED", sched)
    monkeypatch.setenv("PAINTY_FUSED_CHAINS", str(chains))
    a, ab = run_gpu(wl, mode=FUSED)
    b, bb = _fresh(hip, wl)
    bb.imprintStrokes(_strokes(wl), b, FUSED)
    _same(a, ab, b, bb)
    ocv, ob = run_oracle(wl)
    compare_canvas("strokes%d" % chains, b, ocv, pickup=(bb, ob))


@pytest.mark.parametrize("sched", ["waves", "streams"])
@pytest.mark.parametrize("snapshot", [True, False])
def test_edge_strokes_equal_plain_bitwise(hip, oracle, sched, snapshot, monkeypatch):
    """Strokes along the top and left edges: most stamps straddle canvas
    column/row 0 (pixels visited twice, some on the window border), and the
    fused schedulers run each such stamp in one launch (window tiles, ring
    scan and the shared-pixel levels side by side).  The result equals the
    plain per-stamp path, which splits them, to the bit, and the oracle."""
    wl = W.small(72, 90, 0, 11, 5, pickup=True, noisy=True)
    st = np.zeros(60, dtype=W.STAMP_DTYPE)
    t = np.arange(60, dtype=np.float64)
    st["cx"][:30] = 2.3 + 2.1 * t[:30]          # along row 0, centres above and below it
    st["cy"][:30] = 0.35 + 3.0 * np.sin(t[:30] / 3.0)
    # down column 0; with the centre in (10, 11) the window's left border is column 0 itself
    st["cx"][30:] = 10.45 + 0.5 * np.cos(t[30:] / 4.0)
    st["cy"][30:] = 1.4 + 1.9 * (t[30:] - 30)
    st["theta"] = 0.37 * t
    st["K"], st["S"] = PA
    st["flags"][0] = st["flags"][30] = W.DIP
    st["flags"][44] |= W.DRY_AFTER
    monkeypatch.setenv("PAINTY_FUSED_SCHED", sched)
    monkeypatch.setenv("PAINTY_FUSED_CHAINS", "4")
    a, ab = run_gpu(wl, st, use_snapshot=snapshot, mode=PLAIN)
    b, bb = run_gpu(wl, st, use_snapshot=snapshot, mode=FUSED)
    _same(a, ab, b, bb)
    c, cb = _fresh(hip, wl)
    cb.setUseSnapshotBuffer(snapshot)
    cb.imprintStrokes([(wl.radii[0], wl.cores[0], st[:30]), (wl.radii[0], wl.cores[0], st[30:])],
                      c, FUSED)
    _same(a, ab, c, cb)
    ocv, ob = run_oracle(wl, st, use_snapshot=snapshot)
    compare_canvas("edges/%s/snap%d" % (sched, snapshot), c, ocv, pickup=(cb, ob))


@pytest.mark.parametrize("mode", [PLAIN, FUSED])
def test_imprint_strokes_geometry_rules(hip, oracle, mode):
    """Radius changes within 0.5 keep the footprint and the pickup map; a
    stroke without stamps still changes the geometry (and cleans the map);
    the brush ends with the last stroke's geometry."""
    wl = W.small(70, 80, 30, 5, 12, pickup=True, noisy=True, reach=0.3)
    st = wl.stamps.copy()
    st["flags"] = 0
    st["flags"][0] = W.DIP
    c5, c9 = W.gaussian_core(11, 3.0), W.gaussian_core(19, 5.0)
    strokes = [(5.0, c5, st[:8]), (5.3, c5, st[8:14]), (9.0, c9, st[:0]), (9.2, c9, st[14:22]),
               (5.0, c5, st[22:26]), (9.0, c9, st[:0])]
    a, ab = _fresh(hip, wl)
    for radius, core, part in strokes:  # FootprintBrush::setRadius + imprint, one by one
        ab.setRadius(radius, core)
        ab.imprintStamps(part, a, PLAIN)
    b, bb = _fresh(hip, wl)
    bb.imprintStrokes(strokes, b, mode)
    _same(a, ab, b, bb)
