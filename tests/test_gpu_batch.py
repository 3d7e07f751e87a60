"""pnt_batch_run: many (canvas, brush, stamps) jobs in one launch, each job
checked against the FP64 oracle running that job alone (1e-5 relative per
channel): stamped wet layer, R0/h after drying, pickup map and composite."""
import numpy as np
import pytest

from painty_amd import workload as W
from tests.parity import assert_close, compare_canvas, run_oracle

pytestmark = pytest.mark.gpu


def _setup(hip, wl, use_snapshot=True):
    cv = hip.Canvas(wl.rows, wl.cols)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    dry = hip.PaintLayer(wl.rows, wl.cols)
    dry.setBuffers(wl.dry.K, wl.dry.S, wl.dry.V)
    dry.composeOntoCanvas(cv)
    br = hip.FootprintBrush(wl.radii[0], wl.cores[0])
    br.setPickupRate(wl.pickup_rate)
    br.setDepositionRate(wl.deposition_rate)
    br.setUseSnapshotBuffer(use_snapshot)
    out = hip.Canvas(wl.rows, wl.cols)
    return cv, br, out


def _mixed_jobs(seed):
    """Canvases of assorted sizes and brushes: pickup on/off, snapshot
    on/off, re-dips, a dry in the middle, strokes near the edges."""
    rng = W.Stream(seed, 77)
    jobs = []
    for k in range(14):
        rows, cols = (int(x) for x in rng.integers(2, 24, 200))
        radius = int(rng.integers(1, 2, 40)[0])
        n = int(rng.integers(1, 0, 30)[0])
        wl = W.small(rows, cols, n, radius, seed * 100 + k, pickup=bool(k % 3), noisy=bool(k % 2),
                     reach=0.4)
        if n > 4 and k % 4 == 1:
            wl.stamps["flags"][n // 2] |= W.DRY_AFTER
        jobs.append((wl, k % 5 != 2))
    return jobs


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("seed", [1, 2])
def test_batch_matches_oracle_per_job(hip, oracle, mode, seed):
    jobs = _mixed_jobs(seed)
    gpu = []
    for wl, snap in jobs:
        cv, br, out = _setup(hip, wl, snap)
        gpu.append((cv, br, wl.stamps, out))
    hip.batch_run(gpu, mode=mode)
    for k, ((wl, snap), (cv, br, _, out)) in enumerate(zip(jobs, gpu)):
        ocv, ob = run_oracle(wl, use_snapshot=snap)
        tag = "batch%d/s%d/job%d(%dx%d,r%g,n%d)" % (mode, seed, k, wl.rows, wl.cols, wl.radii[0],
                                                   len(wl.stamps))
        compare_canvas(tag, cv, ocv, pickup=(br, ob) if ob is not None else None)
        assert_close(tag + ".out", out.getR0(), ocv.compose())


def test_batch_big_and_small_jobs(hip, oracle):
    """A few large jobs among small ones, so the large ones get groups of
    several workgroups (group barriers) and the small ones share."""
    jobs = []
    for k in range(6):
        big = k < 2
        wl = W.small(300 if big else 60, 280 if big else 70, 60 if big else 8, 40 if big else 6,
                     40 + k, pickup=True, noisy=True, reach=0.3)
        jobs.append(wl)
    gpu = []
    for wl in jobs:
        cv, br, out = _setup(hip, wl)
        gpu.append((cv, br, wl.stamps, out))
    hip.batch_run(gpu, mode=0)
    for k, (wl, (cv, br, _, out)) in enumerate(zip(jobs, gpu)):
        ocv, ob = run_oracle(wl)
        compare_canvas("bigsmall/job%d" % k, cv, ocv, pickup=(br, ob))
        assert_close("bigsmall/job%d.out" % k, out.getR0(), ocv.compose())


def test_batch_reduced(hip, oracle):
    """The batched configuration at reduced size: 24 canvases 128..384 px,
    10..60 stamps, footprints 32..128 px, pixels from the shared pool."""
    bw = W.batch256(n_canvas=24, seed=5, sides=(128, 384), stamps=(10, 60), foot=(32, 128))
    gpu = []
    for wl in bw.jobs:
        cv, br, out = _setup(hip, wl)
        gpu.append((cv, br, wl.stamps, out))
    hip.batch_run(gpu, mode=0)
    for k, (wl, (cv, br, _, out)) in enumerate(zip(bw.jobs, gpu)):
        ocv, ob = run_oracle(wl)
        compare_canvas("batch-reduced/job%d" % k, cv, ocv, pickup=(br, ob))
        assert_close("batch-reduced/job%d.out" % k, out.getR0(), ocv.compose())


def test_batch_then_single_calls_continue_state(hip, oracle):
    """After a batch, the same brush and canvas continue with ordinary calls
    (snapshot, dirty bytes and paint carried over)."""
    wl = W.small(90, 100, 30, 7, 61, pickup=True, noisy=True, reach=0.3)
    first, second = wl.stamps[:18], wl.stamps[18:]
    cv, br, out = _setup(hip, wl)
    hip.batch_run([(cv, br, first, out)], mode=0)
    br.imprintStamps(second, cv, 1)
    ocv, ob = run_oracle(wl)
    compare_canvas("batch+persistent", cv, ocv, pickup=(br, ob))


def test_batch_rejects_shared_brush(hip):
    wl = W.small(40, 40, 5, 4, 3)
    cv1, br, out = _setup(hip, wl)
    cv2 = hip.Canvas(40, 40)
    with pytest.raises(hip.InvalidArgument):
        hip.batch_run([(cv1, br, wl.stamps, None), (cv2, br, wl.stamps, None)])


def _big_small(seed0=40):
    jobs = []
    for k in range(6):
        big = k < 2
        jobs.append(W.small(300 if big else 60, 280 if big else 70, 60 if big else 8,
                            40 if big else 6, seed0 + k, pickup=True, noisy=True, reach=0.3))
    return jobs


@pytest.fixture
def restore_config(hip):
    saved = hip.batch_get_config()
    yield
    hip.batch_set_config(*saved)


@pytest.mark.parametrize("tile,max_group,waves,chunk",
                         [(16, 64, 0, 1), (16, 4, 0, 1), (16, 1, 0, 1), (32, 1, 0, 1),
                          (32, 4, 0, 1), (16, 1, 1, 1), (16, 1, 1, 2), (16, 1, 1, 4)])
def test_batch_configs_match_oracle(hip, oracle, restore_config, tile, max_group, waves, chunk):
    """Every configuration the autotuner may pick -- one persistent launch
    with 16- or 32-pixel window tiles and one to 64 workgroups per job, or
    waves with one, two or four strokes of a canvas in flight -- gives the
    oracle's result."""
    hip.batch_set_config(tile, max_group, waves, chunk)
    assert hip.batch_get_config() == (tile, max_group, waves, chunk)
    jobs = _big_small()
    gpu = []
    for wl in jobs:
        cv, br, out = _setup(hip, wl)
        gpu.append((cv, br, wl.stamps, out))
    hip.batch_run(gpu, mode=0)
    for k, (wl, (cv, br, _, out)) in enumerate(zip(jobs, gpu)):
        ocv, ob = run_oracle(wl)
        tag = "cfg%d.%d/job%d" % (tile, max_group, k)
        compare_canvas(tag, cv, ocv, pickup=(br, ob))
        assert_close(tag + ".out", out.getR0(), ocv.compose())


def test_batch_config_rejects_bad_shapes(hip, restore_config):
    for cfg in [(8, 4), (64, 4), (32, 0), (16, 257), (16, 4, 2, 1), (16, 4, 1, 3),
                (16, 4, 1, 0)]:
        with pytest.raises(hip.InvalidArgument):
            hip.batch_set_config(*cfg)


def test_batch_tune_keeps_state_and_config(hip, oracle, restore_config):
    """pnt_batch_tune runs the batch many times but leaves every canvas and
    brush as it found them; the eager run afterwards uses the tuned shape and
    still equals the oracle."""
    jobs = _mixed_jobs(3)[:8]
    gpu = []
    for wl, snap in jobs:
        cv, br, out = _setup(hip, wl, snap)
        gpu.append((cv, br, wl.stamps, out))
    (tile, g, waves, chunk), ms = hip.batch_tune(gpu)
    assert tile in (16, 32) and 1 <= g <= 256 and waves in (0, 1) and chunk in (1, 2, 4)
    assert ms > 0
    assert hip.batch_get_config() == (tile, g, waves, chunk)
    hip.batch_run(gpu, mode=0)
    for k, ((wl, snap), (cv, br, _, out)) in enumerate(zip(jobs, gpu)):
        ocv, ob = run_oracle(wl, use_snapshot=snap)
        compare_canvas("tuned/job%d" % k, cv, ocv, pickup=(br, ob) if ob is not None else None)
        assert_close("tuned/job%d.out" % k, out.getR0(), ocv.compose())


def _layers(cv):
    lay = cv.getPaintLayer()
    return [lay.getK_buffer(), lay.getS_buffer(), lay.getV_buffer(), cv.getR0()]


@pytest.mark.parametrize("waves", [0, 1])
def test_batch_graph_replays_equal_eager_and_oracle(hip, oracle, restore_config, waves):
    """A captured batch launched twice equals the eager batch run twice (to
    the bit: same kernels, same plan) and the oracle running each stamp list
    twice -- for one persistent launch and for waves."""
    hip.batch_set_config(16, 16, waves, 4)
    jobs = _big_small(70)
    eager, graph = [], []
    for wl in jobs:
        cv, br, out = _setup(hip, wl)
        eager.append((cv, br, wl.stamps, out))
        cv, br, out = _setup(hip, wl)
        graph.append((cv, br, wl.stamps, out))
    g = hip.BatchGraph(graph)
    for _ in range(2):
        hip.batch_run(eager, mode=0)
        g.launch()
    g.close()
    for k, (wl, e, q) in enumerate(zip(jobs, eager, graph)):
        for a, b in zip(_layers(e[0]) + [e[3].getR0()] + list(e[1].getPickupMap()),
                        _layers(q[0]) + [q[3].getR0()] + list(q[1].getPickupMap())):
            assert np.array_equal(a, b), "graph/job%d" % k
        ocv, ob = run_oracle(wl, np.concatenate([wl.stamps, wl.stamps]))
        compare_canvas("graph/job%d" % k, q[0], ocv, pickup=(q[1], ob))
        assert_close("graph/job%d.out" % k, q[3].getR0(), ocv.compose())


def test_batch_profile_runs_the_batch_and_reports_its_waves(hip, restore_config):
    """pnt_batch_profile runs the batch as pnt_batch_run does in the waves
    configuration (the same bits) and reports a schedule that covers it:
    every job's first wave before its last, the last wave belongs to some
    job, wave end times rise, and the two long jobs end after the small
    ones."""
    hip.batch_set_config(16, 1, 1, 1)
    jobs = _big_small(90)
    ref, prof = [], []
    for wl in jobs:
        cv, br, out = _setup(hip, wl)
        ref.append((cv, br, wl.stamps, out))
        cv, br, out = _setup(hip, wl)
        prof.append((cv, br, wl.stamps, out))
    hip.batch_run(ref, mode=0)
    jf, jl, wb, we = hip.batch_profile(prof)
    for k, (e, q) in enumerate(zip(ref, prof)):
        for a, b in zip(_layers(e[0]) + [e[3].getR0()] + list(e[1].getPickupMap()),
                        _layers(q[0]) + [q[3].getR0()] + list(q[1].getPickupMap())):
            assert np.array_equal(a, b), "profile/job%d" % k
    assert len(wb) > 0 and len(we) == len(wb)
    assert (jf >= 0).all() and (jf <= jl).all()
    assert jl.max() == len(wb) - 1
    assert (wb > 0).all() and we[0] > 0 and np.all(np.diff(we) >= 0)
    assert jl[2:].max() < jl[:2].min()


def test_batch_graph_refuses_stale_state(hip):
    wl = W.small(50, 60, 12, 5, 81)
    cv, br, out = _setup(hip, wl)
    g = hip.BatchGraph([(cv, br, wl.stamps, out)])
    g.launch()
    br.imprintStamps(wl.stamps[:3], cv, 0)  # an ordinary call changes the canvas
    with pytest.raises(hip.InvalidArgument):
        g.launch()
    g.close()
    st = wl.stamps.copy()
    st["flags"][0] = 0  # no leading dip: paint would depend on the brush at launch time
    with pytest.raises(hip.InvalidArgument):
        hip.BatchGraph([(cv, br, st, out)])


def test_objects_outlive_their_context(hip, oracle):
    """pnt_ctx_destroy while canvases, brushes and a captured batch still
    use the context: they keep working, and destroying them afterwards (in
    any order) releases the context last."""
    wl = W.small(48, 52, 14, 5, 91, pickup=True)
    ctx = hip.Context(0)
    cv = hip.Canvas(wl.rows, wl.cols, ctx)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    br = hip.FootprintBrush(wl.radii[0], wl.cores[0], ctx)
    br.setPickupRate(wl.pickup_rate)
    br.setDepositionRate(wl.deposition_rate)
    g = hip.BatchGraph([(cv, br, wl.stamps[:7], None)], ctx)
    ctx.close()  # the caller's reference goes first
    g.launch()
    br.imprintStamps(wl.stamps[7:], cv, 2)
    ocv, ob = run_oracle(wl, bake_dry=False)
    compare_canvas("outlive", cv, ocv, pickup=(br, ob))
    del cv
    g.close()
    del br


def test_batch_submit_then_wait(hip):
    """pnt_batch_submit returns with the batch enqueued; after pnt_batch_wait
    every canvas equals the same jobs run by pnt_batch_run, bit for bit."""
    bw = W.batch256(n_canvas=12, seed=9, sides=(128, 320), stamps=(10, 50), foot=(32, 96))
    a = [(_setup(hip, wl)) for wl in bw.jobs]
    b = [(_setup(hip, wl)) for wl in bw.jobs]
    hip.batch_run([(cv, br, wl.stamps, out) for (cv, br, out), wl in zip(a, bw.jobs)], mode=0)
    hip.batch_submit([(cv, br, wl.stamps, out) for (cv, br, out), wl in zip(b, bw.jobs)])
    hip.batch_wait()
    for (ca, _, oa), (cb, _, ob) in zip(a, b):
        La, Lb = ca.getPaintLayer(), cb.getPaintLayer()
        np.testing.assert_array_equal(La.getK_buffer(), Lb.getK_buffer())
        np.testing.assert_array_equal(La.getV_buffer(), Lb.getV_buffer())
        np.testing.assert_array_equal(oa.getR0(), ob.getR0())


def _setup_on(hip, wl, ctx):
    cv = hip.Canvas(wl.rows, wl.cols, ctx)
    cv.getPaintLayer().setBuffers(wl.wet.K, wl.wet.S, wl.wet.V)
    cv.setR0(wl.R0)
    br = hip.FootprintBrush(wl.radii[0], wl.cores[0], ctx)
    br.setPickupRate(wl.pickup_rate)
    br.setDepositionRate(wl.deposition_rate)
    return cv, br, hip.Canvas(wl.rows, wl.cols, ctx)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_jobs_on_another_context_read_back_without_wait(hip, oracle, graph):
    """Canvases and brushes made on a second context, the batch submitted on
    the default one, and the results read back through the jobs' own context
    with no pnt_batch_wait: the batch's end orders the other context's stream,
    so the read sees the finished batch."""
    bw = W.batch256(n_canvas=6, seed=13, sides=(200, 420), stamps=(40, 90), foot=(48, 96))
    other = hip.Context(0)
    jobs = [_setup_on(hip, wl, other) for wl in bw.jobs]
    other.synchronize()
    arg = [(cv, br, wl.stamps, out) for (cv, br, out), wl in zip(jobs, bw.jobs)]
    if graph:
        g = hip.BatchGraph(arg)
        g.launch()
    else:
        hip.batch_submit(arg)
    for k, ((cv, br, out), wl) in enumerate(zip(jobs, bw.jobs)):
        ocv, ob = run_oracle(wl, bake_dry=False)
        compare_canvas("otherctx%d/job%d" % (graph, k), cv, ocv, pickup=(br, ob))
        assert_close("otherctx%d/job%d.out" % (graph, k), out.getR0(), ocv.compose())
    if graph:
        g.close()
    hip.batch_wait()
