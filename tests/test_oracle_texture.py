"""Pins the oracle's TextureBrush and Smudge (CPU only).

Checks that do not rest on the oracle's own code: the reference's golden
values for the mean value interpolation (core/test/src/MathTest.cxx:14-51),
linear precision and Floater's tan-half-angle form of the same coordinates, a
closed form for a straight stroke, closed forms of the smudge's rotation,
conservation and invariance properties of the smudge, and an independent
numpy restatement (tests/brute_texture.py) on small strokes.
"""
import math
import os

import numpy as np
import pytest

from tests.brute_texture import BruteTextureBrush

FIX = os.path.join(os.path.dirname(__file__), "golden", "texture", "sample0_thickness_q4.npy")


def sample_thickness() -> np.ndarray:
    """io::imRead(thickness_map.png, Mat<double>&, true) of the fixture:
    16-bit values / 65535, then sRGB -> linear (core/Color.hxx:189-195)."""
    s = np.load(FIX).astype(np.float64) / 65535.0
    return np.where(s <= 0.0404482362771082, s / 12.92, ((s + 0.055) / 1.055) ** 2.4)


# --------------------------------------------------- mean value coords -----
def test_gbc_reference_golden(oracle):
    sq = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]])
    vals = np.array([[0.0, 0.0], [0.0, 0.0], [1.0, 1.0], [1.0, 1.0]])
    assert abs(oracle.gbc(sq, 0.0, 0.0, vals)[0] - 0.5) < 1e-6   # MathTest.cxx:14-24
    one = oracle.gbc(np.array([[-0.5, -0.5]]), 0.0, 0.0, np.array([[0.563, 0.563]]))
    assert abs(one[0] - 0.563) < 1e-6                             # MathTest.cxx:43-51


def _convex(rng, n):
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(3.0, 5.0)
    return np.stack([rad * np.cos(ang) + 1.3, 0.7 * rad * np.sin(ang) - 0.4], 1)


def test_gbc_linear_precision(oracle):
    """Mean value coordinates reproduce affine functions everywhere off the
    boundary (Hormann and Floater 2006), inside and outside the polygon."""
    rng = np.random.default_rng(5)
    for _ in range(40):
        P = _convex(rng, int(rng.integers(3, 12)))
        M = rng.normal(size=(2, 2))
        b = rng.normal(size=2)
        vals = P @ M.T + b
        for p in rng.uniform(-8, 8, size=(10, 2)):
            got = oracle.gbc(P, p[0], p[1], vals)
            np.testing.assert_allclose(got, M @ p + b, rtol=0, atol=1e-6)  # outside: W may be small


def test_gbc_floater_half_angle(oracle):
    """Inside a convex polygon the weights are Floater's
    (tan(a_{i-1}/2) + tan(a_i/2)) / |v_i - p|, a_i the angle at p between
    v_i and v_{i+1}: an independent form, checked on non-affine values."""
    rng = np.random.default_rng(6)
    for _ in range(30):
        P = _convex(rng, int(rng.integers(3, 10)))
        vals = rng.normal(size=(len(P), 2))
        c = P.mean(0)
        for lam in rng.uniform(0.05, 0.9, size=5):
            p = c + lam * (P[int(rng.integers(len(P)))] - c) * 0.9
            d = P - p
            r = np.hypot(d[:, 0], d[:, 1])
            d1 = np.roll(d, -1, 0)
            a = np.arctan2(d[:, 0] * d1[:, 1] - d[:, 1] * d1[:, 0], (d * d1).sum(1))
            w = (np.tan(np.roll(a, 1) / 2) + np.tan(a / 2)) / r
            want = (w[:, None] * vals).sum(0) / w.sum()
            np.testing.assert_allclose(oracle.gbc(P, p[0], p[1], vals), want, rtol=1e-10,
                                       atol=1e-10)


def test_gbc_vertex_and_edge(oracle):
    P = np.array([[0.0, 0.0], [4.0, 0.0], [4.0, 2.0], [0.0, 2.0]])
    vals = np.array([[1.0, 2.0], [3.0, 5.0], [7.0, 11.0], [13.0, 17.0]])
    for i in range(4):
        np.testing.assert_array_equal(oracle.gbc(P, *P[i], vals), vals[i])
    # a quarter of the way along edge 1 -> 2 (non-affine values)
    got = oracle.gbc(P, 4.0, 0.5, vals)
    np.testing.assert_allclose(got, 0.75 * vals[1] + 0.25 * vals[2], rtol=1e-14)


# ------------------------------------------------------ straight stroke -----
class _Cv:
    def __init__(self, oracle, rows, cols, V0=0.0, K0=(0.0, 0.0, 0.0), S0=(0.0, 0.0, 0.0)):
        self.c = oracle.Canvas(rows, cols)
        self.c.V[...] = V0
        self.c.K[...] = K0
        self.c.S[...] = S0


def _band_case(oracle, V0, K0):
    """Horizontal stroke, vertices r apart so that with the radius
    extensions every frame vertex sits on an affine (u, v) map:
    u = (x - 6.25) / 20, v = (y - 16.5) / 8.  Texture thick[row, col] =
    0.05 + col / 10 (10 columns), so Vtex = 0.05 + min(u, 0.9)."""
    r = 4.0
    xy = np.array([[10.25 + 4 * k, 20.5] for k in range(4)])
    tex = np.tile(0.05 + np.arange(10) / 10.0, (5, 1))
    b = oracle.TextureBrush()
    b.enable_smudge(False)
    b.set_radius(r)
    b.set_thickness_map(tex)
    Kp, Sp = np.array([0.3, 0.2, 0.7]), np.array([0.5, 0.9, 0.4])
    b.dip(Kp, Sp)
    cv = _Cv(oracle, 40, 40, V0, K0, (0.25, 0.25, 0.25)).c
    b.paint_stroke(xy, cv)
    Y, X = np.mgrid[0:40, 0:40].astype(float)
    u, v = (X - 6.25) / 20.0, (Y - 16.5) / 8.0
    cov = (u >= 0) & (u <= 1) & (v >= 0) & (v <= 1)
    vt = 0.05 + np.minimum(u, 0.9)
    return cv, cov, vt, Kp, Sp


def test_straight_stroke_closed_form(oracle):
    cv, cov, vt, Kp, Sp = _band_case(oracle, 0.0, (0.0, 0.0, 0.0))
    assert cov.sum() == 20 * 8  # x 7..26, y 17..24
    np.testing.assert_allclose(cv.V[cov], vt[cov], rtol=0, atol=1e-12)
    np.testing.assert_allclose(cv.K[cov], np.broadcast_to(Kp, (cov.sum(), 3)), atol=1e-12)
    np.testing.assert_allclose(cv.S[cov], np.broadcast_to(Sp, (cov.sum(), 3)), atol=1e-12)
    assert (cv.V[~cov] == 0).all() and (cv.K[~cov] == 0).all()


def test_straight_stroke_mixes_by_volume(oracle):
    Kc = np.array([0.9, 0.1, 0.4])
    cv, cov, vt, Kp, Sp = _band_case(oracle, 0.5, Kc)
    want = (0.5 * Kc + vt[cov][:, None] * Kp) / (0.5 + vt[cov][:, None])
    np.testing.assert_allclose(cv.K[cov], want, atol=1e-12)
    np.testing.assert_allclose(cv.V[cov], np.maximum(vt[cov], 0.5), atol=1e-12)
    np.testing.assert_array_equal(cv.V[~cov], 0.5)
    np.testing.assert_array_equal(cv.K[~cov], np.broadcast_to(Kc, ((~cov).sum(), 3)))


# --------------------------------------------------------------- smudge -----
def test_smudge_turn_closed_form(oracle):
    """updateOrientation: a heading along +x from rotation 0 leaves the map
    as it is; a reversed heading is the same direction (angles modulo pi);
    a heading along +y turns by pi/2 about (max/2, max/2): cell (x, y)
    takes the value at (max - y, x), and keeps its own where that leaves
    the map.  Affine maps make bilinear sampling exact."""
    b = oracle.TextureBrush()
    b.set_radius(4.0)  # 8 x 8 map, max size 9
    K, S, V = b.smudge_map()
    assert V.shape == (8, 8)
    y, x = np.mgrid[0:8, 0:8].astype(float)
    f = lambda xx, yy: 1.0 + 0.1 * xx + 0.01 * yy
    V0 = f(x, y)
    K0 = np.stack([f(x, y), 2 * f(x, y), 3 + f(x, y)], -1)
    b.set_smudge_map(K0, K0 * 0.5, V0)
    b.turn(1.0, 0.0)
    np.testing.assert_array_equal(b.smudge_map()[2], V0)
    b.turn(-1.0, 0.0)  # theta = pi, pi - 0 normalises to 0
    np.testing.assert_array_equal(b.smudge_map()[2], V0)
    b2 = oracle.TextureBrush()
    b2.set_radius(4.0)
    b2.set_smudge_map(K0, K0 * 0.5, V0)
    b2.turn(0.0, 1.0)
    K1, S1, V1 = b2.smudge_map()
    want = np.where(y >= 2, f(9.0 - y, x), V0)
    np.testing.assert_allclose(V1, want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(K1[..., 1], 2 * want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(S1[..., 2], 0.5 * (3 + want), rtol=0, atol=1e-12)


def _smudge_canvas(oracle, rows=48, cols=64):
    cv = oracle.Canvas(rows, cols)
    cv.V[...] = 1.0
    cv.K[:, : cols // 2] = [0.8, 0.2, 0.1]
    cv.K[:, cols // 2:] = [0.1, 0.3, 0.9]
    cv.S[...] = [0.4, 0.5, 0.6]
    return cv


def test_smudge_conserves_volume_on_a_straight_stroke(oracle):
    """A straight stroke never turns the map, and each exchange moves
    volume between canvas and map; with Vtex below the canvas volume the
    deposit keeps V = max(Vtex, Vcan) = Vcan, so canvas plus map volume is
    what the canvas held."""
    b = oracle.TextureBrush()
    b.set_radius(5.0)
    b.set_thickness_map(np.full((6, 9), 0.5))
    b.set_thickness_scale(0.1)
    cv = _smudge_canvas(oracle)
    before = cv.V.sum()
    b.paint_stroke(np.array([[8.3, 20.6], [30.3, 20.6], [50.3, 20.6]]), cv)
    _, _, mV = b.smudge_map()
    assert mV.sum() > 1.0  # the map picked paint up
    np.testing.assert_allclose(cv.V.sum() + mV.sum(), before, rtol=1e-12)


def test_smudge_keeps_uniform_paint(oracle):
    """Mixing a colour with itself gives it back: a canvas of one paint,
    smudged and painted with that paint along a straight (diagonal) stroke,
    stays it.  (Not so on a curved stroke: the map's turns resample K, S and
    V apart, so K of filled cells blends with the zeros of empty ones.)"""
    b = oracle.TextureBrush()
    b.set_radius(4.5)
    b.set_thickness_map(sample_thickness())
    Kp, Sp = np.array([0.3, 0.6, 0.2]), np.array([0.7, 0.1, 0.5])
    b.dip(Kp, Sp)
    cv = oracle.Canvas(48, 64)
    cv.V[...] = 0.8
    cv.K[...] = Kp
    cv.S[...] = Sp
    b.paint_stroke(np.array([[6.2, 8.7], [26.2, 23.7], [46.2, 38.7]]), cv)
    np.testing.assert_allclose(cv.K, np.broadcast_to(Kp, cv.K.shape), rtol=1e-12)
    np.testing.assert_allclose(cv.S, np.broadcast_to(Sp, cv.S.shape), rtol=1e-12)
    mK, _, mV = b.smudge_map()
    held = mV > 0.001  # cells that took paint (turns leak ulps into empty ones)
    assert held.sum() > 10
    np.testing.assert_allclose(mK[held], np.broadcast_to(Kp, (held.sum(), 3)), rtol=1e-12)


@pytest.mark.parametrize("smudge", [False, True])
def test_against_brute(oracle, smudge):
    """The oracle and the independent numpy restatement on curved strokes
    over a two-colour canvas, the reference's sample texture, two strokes in
    a row (the smudge map and its rotation carry over)."""
    tex = sample_thickness()
    b = oracle.TextureBrush()
    ref = BruteTextureBrush(tex)
    for x in (b, ref):
        if x is b:
            x.enable_smudge(smudge)
            x.set_radius(4.6)
            x.set_thickness_map(tex)
            x.set_thickness_scale(0.7)
            x.dip([0.2, 0.3, 0.4], [0.1, 0.23, 0.14])
        else:
            x.smudge_on = smudge
            x.set_radius(4.6)
            x.scale = 0.7
            x.K, x.S = np.array([0.2, 0.3, 0.4]), np.array([0.1, 0.23, 0.14])
    cv = _smudge_canvas(oracle)

    class Cv:
        pass
    bc = Cv()
    bc.K, bc.S, bc.V = cv.K.copy(), cv.S.copy(), cv.V.copy()
    K_before = cv.K.copy()
    strokes = [np.array([[5.3, 8.1], [20.7, 14.2], [38.4, 12.9], [55.6, 25.3]]),
               np.array([[58.1, 40.2], [40.6, 33.3], [22.2, 38.7], [3.9, 30.1]])]
    for xy in strokes:
        b.paint_stroke(xy, cv)
        ref.paint_stroke(xy, bc)
    np.testing.assert_allclose(cv.V, bc.V, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(cv.K, bc.K, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(cv.S, bc.S, rtol=1e-9, atol=1e-12)
    if smudge:
        mK, mS, mV = b.smudge_map()
        np.testing.assert_allclose(mV, ref.mV, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(mK, ref.mK, rtol=1e-9, atol=1e-12)
        assert (mV > 0).any()
    assert (np.abs(cv.K - K_before) > 1e-2).sum() > 100


def test_short_stroke_is_a_no_op(oracle):
    b = oracle.TextureBrush()
    b.set_radius(3.0)
    b.set_thickness_map(sample_thickness())
    cv = _smudge_canvas(oracle)
    K0 = cv.K.copy()
    b.paint_stroke(np.array([[10.0, 10.0]]), cv)
    np.testing.assert_array_equal(cv.K, K0)
    assert (cv.V == 1.0).all()


def test_check_dry_on_textured_pixels(oracle):
    """With a clock, paint older than the drying time dries into R0 on the
    textured pixels only (Canvas::checkDry, TextureBrush.hxx:164); their
    time becomes now."""
    b = oracle.TextureBrush()
    b.enable_smudge(False)
    b.set_radius(4.0)
    b.set_thickness_map(np.full((5, 10), 0.3))
    cv = oracle.Canvas(40, 40)
    cv.V[...] = 0.6
    cv.K[...] = [0.5, 0.2, 0.3]
    cv.S[...] = [0.4, 0.4, 0.4]
    cv.tmap[...] = 0
    R0 = cv.R0.copy()
    b.paint_stroke(np.array([[10.25 + 4 * k, 20.5] for k in range(4)]), cv, now_us=20_000_000,
                   dry_ms=15_000)
    Y, X = np.mgrid[0:40, 0:40].astype(float)
    cov = ((X - 6.25) / 20 >= 0) & ((X - 6.25) / 20 <= 1) & ((Y - 16.5) / 8 >= 0) & \
          ((Y - 16.5) / 8 <= 1)
    assert (cv.tmap[cov] == 20_000_000).all() and (cv.tmap[~cov] == 0).all()
    assert (cv.h[cov] == 0.6).all() and (cv.h[~cov] == 0).all()
    # dried pixels then take the brush's paint alone
    np.testing.assert_allclose(cv.V[cov], 0.3)
    np.testing.assert_allclose(cv.K[cov], 0.1)
    assert (np.abs(cv.R0[cov] - R0[cov]) > 1e-3).any()
    np.testing.assert_array_equal(cv.R0[~cov], R0[~cov])
