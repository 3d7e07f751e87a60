"""TextureBrush with Smudge (renderer/TextureBrush.hxx:52-205,
renderer/Smudge.hxx:40-198) on the GPU against the FP64 oracle: wet layer,
substrate, height and composite within 1e-5 relative per channel (the
canvas is FP32 on the GPU), the smudge's pickup map (FP64 on both sides)
within the same bound, time maps exactly.

Textures: the reference's brush stroke sample (data/sample_0, every 4th row
and column, as tests/golden/texture holds it) and synthetic maps."""
import numpy as np
import pytest

from painty_amd import _lib
from tests.parity import assert_close, compare_canvas
from tests.test_oracle_texture import sample_thickness

pytestmark = pytest.mark.gpu


def _canvases(hip, oracle, rows, cols, seed, V0=None):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0.05, 0.9, (rows, cols, 3)).astype(np.float32)
    S = rng.uniform(0.05, 0.9, (rows, cols, 3)).astype(np.float32)
    V = (rng.uniform(0.2, 1.2, (rows, cols)) if V0 is None
         else np.full((rows, cols), V0)).astype(np.float32)
    R0 = rng.uniform(0.3, 1.0, (rows, cols, 3)).astype(np.float32)
    oc = oracle.Canvas.from_arrays(K.astype(np.float64), S.astype(np.float64),
                                   V.astype(np.float64), R0.astype(np.float64))
    gc = hip.Canvas(rows, cols)
    gc.getPaintLayer().setBuffers(K, S, V)
    gc.setR0(R0)
    return oc, gc


def _brushes(hip, oracle, radius, tex, smudge=True, scale=1.0,
             paint=((0.2, 0.3, 0.4), (0.1, 0.23, 0.14))):
    ob, gb = oracle.TextureBrush(), hip.TextureBrush()
    ob.enable_smudge(smudge)
    gb.enableSmudge(smudge)
    ob.set_radius(radius)
    gb.setRadius(radius)
    ob.set_thickness_map(tex)
    gb.setThicknessMap(tex)
    ob.set_thickness_scale(scale)
    gb.setThicknessScale(scale)
    ob.dip(*paint)
    gb.dip(paint)
    return ob, gb


def _compare(tag, gc, oc, gb, ob):
    compare_canvas(tag, gc, oc)
    gK, gS, gV = gb.getSmudgeMap()
    oK, oS, oV = ob.smudge_map()
    assert gV.shape == oV.shape, tag
    assert_close(tag + ".smudge.K", gK, oK)
    assert_close(tag + ".smudge.S", gS, oS)
    assert_close(tag + ".smudge.V", gV, oV)


def test_straight_stroke_band(hip, oracle):
    """The closed-form case of the oracle's pins: a straight stroke whose
    frame is an affine (u, v) map of the pixel grid."""
    tex = np.tile(0.05 + np.arange(10) / 10.0, (5, 1))
    oc, gc = _canvases(hip, oracle, 40, 40, 1, V0=0.5)
    ob, gb = _brushes(hip, oracle, 4.0, tex, smudge=False)
    xy = np.array([[10.25 + 4 * k, 20.5] for k in range(4)])
    ob.paint_stroke(xy, oc)
    gb.paintStroke(xy, gc)
    _compare("texture/band", gc, oc, gb, ob)


@pytest.mark.parametrize("smudge", [False, True])
def test_curved_strokes(hip, oracle, smudge):
    """Two curved strokes in a row over a noisy canvas: the smudge map and
    its rotation carry over from the first stroke to the second."""
    oc, gc = _canvases(hip, oracle, 64, 80, 2)
    ob, gb = _brushes(hip, oracle, 5.6, sample_thickness(), smudge, scale=0.8)
    strokes = [np.array([[5.3, 8.1], [20.7, 14.2], [38.4, 12.9], [55.6, 25.3], [70.2, 40.8]]),
               np.array([[74.1, 55.2], [50.6, 45.3], [30.2, 52.7], [6.9, 41.1]])]
    for k, xy in enumerate(strokes):
        ob.paint_stroke(xy, oc)
        gb.paintStroke(xy, gc)
        _compare("texture/curved/smudge%d/stroke%d" % (smudge, k), gc, oc, gb, ob)


def test_stroke_leaving_the_canvas(hip, oracle):
    """Bounding box clamped to the canvas, smudge map cells off the canvas,
    a stroke that doubles back."""
    oc, gc = _canvases(hip, oracle, 48, 56, 3)
    ob, gb = _brushes(hip, oracle, 7.2, sample_thickness(), True)
    xy = np.array([[-6.3, 4.4], [14.8, -3.1], [40.2, 10.6], [61.7, 30.2], [30.3, 50.9],
                   [12.1, 30.6]])
    ob.paint_stroke(xy, oc)
    gb.paintStroke(xy, gc)
    _compare("texture/off-canvas", gc, oc, gb, ob)


def test_radius_change_and_scale(hip, oracle):
    """setRadius rebuilds the smudge map only on a change of 0.5 or more;
    the thickness scale multiplies the texture."""
    oc, gc = _canvases(hip, oracle, 48, 64, 4)
    ob, gb = _brushes(hip, oracle, 3.2, sample_thickness(), True, scale=1.7)
    xy1 = np.array([[4.4, 6.6], [30.1, 20.2], [58.3, 12.9]])
    xy2 = np.array([[8.2, 40.3], [33.3, 30.1], [60.6, 44.4]])
    ob.paint_stroke(xy1, oc)
    gb.paintStroke(xy1, gc)
    for r in (3.5, 9.1):  # 3.5: kept (same map), 9.1: a new 18 x 18 map
        ob.set_radius(r)
        gb.setRadius(r)
        assert gb.getSmudgeMap()[2].shape == ob.smudge_map()[2].shape
    ob.paint_stroke(xy2, oc)
    gb.paintStroke(xy2, gc)
    _compare("texture/radius", gc, oc, gb, ob)
    assert gb.getSmudgeMap()[2].shape == (18, 18)


@pytest.mark.parametrize("age_ms,dry_ms", [(30, 20), (10, 20)])
def test_check_dry_on_textured_pixels(hip, oracle, age_ms, dry_ms):
    rows, cols = 40, 48
    now = 90_000_000
    oc, gc = _canvases(hip, oracle, rows, cols, 5)
    tmap = np.full((rows, cols), now - age_ms * 1000, np.int64)
    oc.tmap[...] = tmap
    gc.setDryingTime(dry_ms)
    gc.setClock(_lib.CLOCK_MANUAL)
    gc.setTimeMap(tmap)
    gc.setTime(now)
    ob, gb = _brushes(hip, oracle, 4.3, sample_thickness(), True)
    xy = np.array([[6.1, 9.9], [22.4, 20.3], [41.2, 31.7]])
    ob.paint_stroke(xy, oc, now_us=now, dry_ms=dry_ms)
    gb.paintStroke(xy, gc)
    tag = "texture/clock/age%d/dry%d" % (age_ms, dry_ms)
    _compare(tag, gc, oc, gb, ob)
    assert np.array_equal(gc.getTimeMap(), oc.tmap), tag + ".tmap"
    assert (oc.tmap == now).any() and (oc.tmap != now).any()


def test_reference_test_stroke(hip, oracle):
    """The reference's own case (renderer/test/src/TextureBrushTest.cxx:16-
    36): a 768 x 1024 cleared canvas, radius 40, the path (50, 250) ->
    (400, 250) -> (650, 250), paint K (0.2, 0.3, 0.4), S (0.1, 0.23, 0.14),
    smudging on (the default)."""
    rows, cols = 768, 1024
    oc = oracle.Canvas(rows, cols)
    gc = hip.Canvas(rows, cols)
    gc.clear()
    ob, gb = _brushes(hip, oracle, 40.0, sample_thickness(), True)
    xy = np.array([[50.0, 250.0], [400.0, 250.0], [650.0, 250.0]])
    ob.paint_stroke(xy, oc)
    gb.paintStroke(xy, gc)
    _compare("texture/reference-test", gc, oc, gb, ob)
    assert (oc.V > 0).sum() > 20000


def test_short_stroke_and_errors(hip, oracle):
    oc, gc = _canvases(hip, oracle, 16, 16, 6)
    gb = hip.TextureBrush()
    gb.setRadius(3.0)
    gb.paintStroke(np.array([[5.0, 5.0]]), gc)  # fewer than 2 vertices: nothing
    with pytest.raises(_lib.InvalidArgument):
        gb.paintStroke(np.array([[2.0, 2.0], [9.0, 9.0]]), gc)  # no thickness map
    with pytest.raises(_lib.InvalidArgument):
        gb.setThicknessMap(np.full((4, 4), np.nan))
    gb.setThicknessMap(np.ones((4, 4)))
    compare_canvas("texture/no-op", gc, oc)
