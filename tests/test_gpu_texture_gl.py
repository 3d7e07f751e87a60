"""The GL texture brush (TextureBrushGpu, renderer/src/TextureBrushGpu.cxx:
17-305) on the device against the oracle's restatement: the same FP32
arithmetic in the same order, so the canvas planes, the lighting plane and
the smudge maps agree to 1e-6 relative (in practice bit for bit)."""
import numpy as np
import pytest

from tests.test_oracle_texture import sample_thickness

pytestmark = pytest.mark.gpu


def _pair(hip, oracle, rows, cols, seed, radius, tex, smudge, scale=1.0,
          paint=((0.2, 0.3, 0.4), (0.1, 0.23, 0.14))):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0.05, 0.9, (rows, cols, 3)).astype(np.float32)
    S = rng.uniform(0.05, 0.9, (rows, cols, 3)).astype(np.float32)
    V = rng.uniform(0.0, 0.6, (rows, cols)).astype(np.float32)
    gc = hip.Canvas(rows, cols)
    gc.getPaintLayer().setBuffers(K, S, V)
    oc = oracle.GlCanvas(K, S, V)
    gb, ob = hip.TextureBrushGpu(), oracle.TextureBrushGpu()
    for b, gpu in ((gb, True), (ob, False)):
        if gpu:
            b.setRadius(radius)
            b.setTexture(tex)
            b.setThicknessScale(scale)
            b.enableSmudge(smudge)
            b.dip(paint)
        else:
            b.set_radius(radius)
            b.set_texture(tex)
            b.set_thickness_scale(scale)
            b.enable_smudge(smudge)
            b.dip(*paint)
    return gc, oc, gb, ob


def _check(tag, hip, gc, oc, gb, ob):
    L = gc.getPaintLayer()
    got = {"K": L.getK_buffer(), "S": L.getS_buffer(), "V": L.getV_buffer(),
           "light": hip.canvas_lighting(gc)}
    want = {"K": oc.K, "S": oc.S, "V": oc.V, "light": oc.light}
    gK, gS = gb.getSmudge()
    oK, oS = ob.smudge()
    got.update(smK=gK, smS=gS)
    want.update(smK=oK, smS=oS)
    for k in got:
        a, b = np.asarray(got[k], np.float64), np.asarray(want[k], np.float64)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
        print("parity %s.%s max_rel=%.3e exact=%s" % (tag, k, err.max() if err.size else 0.0,
                                                      np.array_equal(a, b)), flush=True)
        assert err.max() <= 1e-6 if err.size else True, (tag, k)


STROKES = [np.array([[6.3, 30.4], [40.2, 22.7], [70.8, 35.1], [90.6, 28.3]]),  # curved
           np.array([[10.2, 12.4], [80.7, 14.1], [12.6, 20.3]]),                 # hairpin
           np.array([[-8.4, 50.2], [30.1, 58.9], [60.3, 70.5]])]                 # off canvas


@pytest.mark.parametrize("smudge", [False, True])
def test_strokes_match_oracle(hip, oracle, smudge):
    gc, oc, gb, ob = _pair(hip, oracle, 64, 96, 3, 7.3, sample_thickness(), smudge, 0.8)
    for k, xy in enumerate(STROKES):
        gb.paintStroke(xy, gc)
        ob.paint_stroke(xy, oc)
        _check("gl/smudge%d/stroke%d" % (smudge, k), hip, gc, oc, gb, ob)


def test_radius_change_and_middle_rows(hip, oracle):
    """A stroke along the middle rows (where the smudge's unflipped window
    meets it), a radius change (new, cleared smudge maps), a second stroke."""
    gc, oc, gb, ob = _pair(hip, oracle, 64, 80, 5, 4.2, sample_thickness(), True)
    xy = np.array([[5.4, 31.6], [40.3, 32.2], [75.1, 31.1]])
    gb.paintStroke(xy, gc)
    ob.paint_stroke(xy, oc)
    _check("gl/middle", hip, gc, oc, gb, ob)
    gb.setRadius(9.7)
    ob.set_radius(9.7)
    xy2 = np.array([[70.2, 20.3], [40.1, 40.7], [8.8, 44.4]])
    gb.paintStroke(xy2, gc)
    ob.paint_stroke(xy2, oc)
    _check("gl/radius", hip, gc, oc, gb, ob)
    assert gb.getSmudge()[0].shape[0] == 20


def test_short_smudge_strokes_in_a_row(hip, oracle):
    """Strokes of one smudge step (all spine samples in one integer cell),
    two steps, and a few, back to back: each step's claims are applied by the
    next step's launch or by the flush after the last, and every stroke
    starts from cleared claims."""
    gc, oc, gb, ob = _pair(hip, oracle, 64, 80, 11, 6.1, sample_thickness(), True)
    strokes = [np.array([[20.2, 30.3], [20.7, 30.6]]),   # one step
               np.array([[20.6, 30.3], [21.4, 30.3]]),   # two
               np.array([[21.3, 30.2], [20.8, 31.7]]),   # two, the other way
               np.array([[18.1, 28.4], [23.9, 33.2]]),   # a few
               np.array([[20.5, 30.5], [20.6, 30.5]])]   # one again
    for k, xy in enumerate(strokes):
        gb.paintStroke(xy, gc)
        ob.paint_stroke(xy, oc)
        _check("gl/short/stroke%d" % k, hip, gc, oc, gb, ob)
    assert np.asarray(gb.getSmudge()[0])[..., 3].max() > 0  # the map took paint up


def test_no_texture_is_an_error(hip):
    from painty_amd import _lib
    gb = hip.TextureBrushGpu()
    gc = hip.Canvas(16, 16)
    gb.paintStroke(np.array([[3.0, 3.0]]), gc)  # one vertex: nothing
    with pytest.raises(_lib.InvalidArgument):
        gb.paintStroke(np.array([[2.0, 2.0], [9.0, 9.0]]), gc)


def test_reference_gputest_smudge_strokes(hip, oracle):
    """The reference's own GL case (renderer/test/src/GpuTest.cxx:29-60): a
    1024 x 768 canvas, smudging on, radius 40, a horizontal stroke
    (50, 400) -> (800, 400) of the paint ComputeScatteringAndAbsorption
    makes from (0.01, 0.02) per channel, then a vertical stroke
    (500, 50) -> (500, 800) of the paint from (0.8, 0.9).  The dictionary's
    random texture is replaced by the reference's brush stroke sample."""
    rows, cols = 768, 1024
    Ka, Sa = oracle.compute_scattering_absorption([0.01] * 3, [0.02] * 3)
    Kb, Sb = oracle.compute_scattering_absorption([0.8] * 3, [0.9] * 3)
    gc = hip.Canvas(rows, cols)
    gc.clear()
    oc = oracle.GlCanvas(np.zeros((rows, cols, 3)), np.zeros((rows, cols, 3)),
                         np.zeros((rows, cols)))
    gb, ob = hip.TextureBrushGpu(), oracle.TextureBrushGpu()
    tex = sample_thickness()
    gb.setTexture(tex)
    ob.set_texture(tex)
    gb.enableSmudge(True)
    ob.enable_smudge(True)
    gb.setRadius(40.0)
    ob.set_radius(40.0)
    for (K, S), xy in [((Ka, Sa), [[50.0, 400.0], [800.0, 400.0]]),
                       ((Kb, Sb), [[500.0, 50.0], [500.0, 800.0]])]:
        gb.dip((K, S))
        ob.dip(K, S)
        gb.paintStroke(np.array(xy), gc)
        ob.paint_stroke(np.array(xy), oc)
    _check("gl/gputest", hip, gc, oc, gb, ob)
    assert (oc.V > 0).sum() > 50000
