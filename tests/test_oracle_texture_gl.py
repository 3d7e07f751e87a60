"""Pins the oracle's restatement of the GL texture brush (TextureBrushGpu,
renderer/src/TextureBrushGpu.cxx:17-305, and its shaders) on the CPU.

Closed forms on straight strokes, where the triangle strip is a rectangle:
the covered pixels are those whose centres lie inside it; the texture
coordinate runs linearly along the stroke and across it from the right side
(v = 0) to the left (v = 1); the texture is normalised and flipped as the
texture dictionary uploads it; PaintTextureFootprintOnCanvas sums volumes,
stores the lighting thickness, and (its dispatch rounded up to 16 x 16)
writes the lighting of uncovered pixels of the rounded box too.  The smudge:
canvas plus pickup volume is conserved along a straight rightward stroke
(no rotation), and a fresh pickup map turns the canvas colour it meets to
zero (the shader mixes the remaining canvas volume with the pickup colour,
Smudge.compute.glsl:91-97) -- the reference's behaviour, kept.
GL's own choices (edge ties, subpixel snapping, sampler state, the order of
racing smudge writes) are fixed in the restatement and not checked against
a GL implementation: parity unpinned at that level.
"""
import numpy as np

P_K, P_S = np.array([0.2, 0.3, 0.4]), np.array([0.1, 0.2, 0.3])


def _canvas(oracle, rows, cols, V0=0.0, K0=0.0):
    return oracle.GlCanvas(np.full((rows, cols, 3), K0), np.full((rows, cols, 3), 0.25),
                           np.full((rows, cols), V0))


def _brush(oracle, tex, r, smudge=False, scale=1.0):
    b = oracle.TextureBrushGpu()
    b.set_radius(r)
    b.set_texture(tex)
    b.dip(P_K, P_S)
    b.set_thickness_scale(scale)
    b.enable_smudge(smudge)
    return b


def _rect_expect(rows, cols, x0, x1, y, r, tcols):
    """Coverage and value of a straight rightward stroke (x0, y) -> (x1, y)
    over a texture whose value is its column / (tcols - 1)."""
    x0, x1 = round(x0 * 256) / 256, round(x1 * 256) / 256  # vertices snap to 1/256 px
    Y, X = np.mgrid[0:rows, 0:cols].astype(float)
    cx, cy = X + 0.5, Y + 0.5
    cov = (cx > x0) & (cx < x1) & (cy > y - r) & (cy < y + r)
    s = (cx - x0) / (x1 - x0)
    val = np.clip(s * tcols - 0.5, 0, tcols - 1) / (tcols - 1)
    return cov, val


def test_straight_stroke_coverage_and_value(oracle):
    rows, cols, r = 48, 64, 6.0
    x0, x1, y = 10.3, 50.7, 20.4
    tex = np.tile(np.arange(12, dtype=float), (5, 1))  # column index, normalised by the brush
    b = _brush(oracle, tex, r)
    cv = _canvas(oracle, rows, cols)
    b.paint_stroke([[x0, y], [x1, y]], cv)
    cov, val = _rect_expect(rows, cols, x0, x1, y, r, 12)
    assert cov.sum() == 41 * 12  # centres 10.5 .. 50.5 by 14.5 .. 25.5
    np.testing.assert_allclose(cv.V[cov], val[cov], rtol=0, atol=1e-5)
    assert (cv.V[~cov] == 0).all()
    paint = cov & (val > 1e-5)  # vSum below 1e-5 leaves a pixel alone
    np.testing.assert_allclose(cv.K[paint], np.broadcast_to(P_K, (paint.sum(), 3)), atol=1e-6)
    assert (cv.K[cov & ~paint] == 0).all()
    np.testing.assert_array_equal(cv.light[paint], cv.V[paint])


def test_texture_is_flipped_and_v_runs_right_to_left(oracle):
    """Rows of the height map go 0..1 downwards; the dictionary flips it on
    upload, and v = 0 is the stroke's right side (c + r d, below a rightward
    stroke in canvas rows): there the sample comes from the image's last
    row, so the values fall from the bottom of the band to its top."""
    rows, cols, r = 48, 64, 8.0
    tex = np.tile(np.arange(9, dtype=float)[:, None], (1, 30))
    b = _brush(oracle, tex, r)
    cv = _canvas(oracle, rows, cols)
    b.paint_stroke([[5.3, 24.2], [58.9, 24.2]], cv)
    col = cv.V[:, 30]
    band = np.arange(17, 32)  # centres 16.5 .. 31.5 are covered; row 16 samples 0
    assert col[16] == 0 and (col[:16] == 0).all() and (col[32:] == 0).all()
    assert (col[band] > 0).all()
    assert col[band[-1]] > 0.9 and col[band[0]] < 0.1
    assert np.all(np.diff(col[band]) > 0)


def test_imprint_sums_volume_and_rounds_its_box(oracle):
    rows, cols, r = 64, 80, 5.0
    x0, x1, y = 12.6, 40.2, 30.3
    tex = np.tile(np.arange(12, dtype=float), (5, 1))
    b = _brush(oracle, tex, r, scale=0.5)
    Kc = np.array([0.7, 0.5, 0.1])
    cv = _canvas(oracle, rows, cols, V0=0.5, K0=0.0)
    cv.K[...] = Kc
    b.paint_stroke([[x0, y], [x1, y]], cv)
    cov, val = _rect_expect(rows, cols, x0, x1, y, r, 12)
    vt = 0.5 * val[cov]
    np.testing.assert_allclose(cv.V[cov], 0.5 + vt, atol=1e-5)
    want = (0.5 * Kc + vt[:, None] * P_K) / (0.5 + vt)[:, None]
    np.testing.assert_allclose(cv.K[cov], want, atol=1e-5)
    np.testing.assert_allclose(cv.light[cov], np.where(vt > 0, vt, 0.5), atol=1e-5)
    # the box: bounds +- r, from (int) min; the dispatch rounds it up to 16
    bx0, by0 = int(x0 - r), int(y - r)
    bw = int((x1 + r) - (x0 - r)) + 1
    bh = int((y + r) - (y - r)) + 1
    gw, gh = (bw + 15) // 16 * 16, (bh + 15) // 16 * 16
    box = np.zeros((rows, cols), bool)
    box[by0:by0 + gh, bx0:bx0 + gw] = True
    np.testing.assert_array_equal(cv.light[box & ~cov], 0.5)  # vTex = 0: lighting = V
    assert (cv.light[~box] == 0).all()
    assert (cv.V[~cov] == 0.5).all()


def test_smudge_conserves_volume_on_a_rightward_stroke(oracle):
    """theta = 0: the pickup map is read where it is written, so each step
    trades volume between canvas and map.  The stroke runs along the middle
    rows, where the smudge's unflipped window meets it.  r = 7.6 makes the
    window 16 wide, the dispatch's own width: with a narrower window the
    invocations past it pick paint up into cells outside the map, whose
    stores GL drops, and that volume is lost (as in the reference)."""
    rows, cols, r = 64, 96, 7.6
    tex = np.full((6, 20), 1.0)
    tex[0, 0] = 0.0  # min-max: all but one texel 1
    xy = [[8.2, 31.6], [80.4, 31.6]]
    plain = _canvas(oracle, rows, cols, V0=0.4, K0=0.3)
    _brush(oracle, tex, r, scale=0.05).paint_stroke(xy, plain)
    sm = _canvas(oracle, rows, cols, V0=0.4, K0=0.3)
    b = _brush(oracle, tex, r, smudge=True, scale=0.05)
    b.paint_stroke(xy, sm)
    K4, _ = b.smudge()
    assert K4[..., 3].sum() > 1.0  # the map took paint up
    total = sm.V.astype(np.float64).sum() + K4[..., 3].astype(np.float64).sum()
    np.testing.assert_allclose(total, plain.V.astype(np.float64).sum(), rtol=2e-5)


def test_smudge_fresh_map_zeroes_canvas_colour(oracle):
    """First contact of an empty pickup cell (pV = 0): the canvas keeps
    cVr = 0.3 cV of its volume but takes the pickup colour (zero) --
    Smudge.compute.glsl:91-97 as written.  Then the imprint mixes in the
    brush paint by volume."""
    rows, cols, r = 64, 96, 6.0
    tex = np.full((6, 20), 1.0)
    tex[0, 0] = 0.0
    cv = _canvas(oracle, rows, cols, V0=0.4, K0=0.3)
    b = _brush(oracle, tex, r, smudge=True, scale=0.05)
    b.paint_stroke([[8.2, 31.6], [80.4, 31.6]], cv)
    hit = (cv.K[..., 0] < 0.3 - 1e-3) & (cv.V < 0.4 - 1e-3)
    assert hit.sum() > 50
    assert (cv.K[hit][:, 0] < 0.3).all() and (cv.K[hit] >= 0).all()
