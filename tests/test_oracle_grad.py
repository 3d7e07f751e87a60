"""The FP64 oracle gradient of composite + stamping (orc_grad_run) against
FP64 central finite differences (h = 1e-6) of the oracle's own forward, on
>= 64 sampled coordinates per case: wet-layer K, S, V, substrate R0 and the
mixing weights W (64-paint x 14-pigment style, reduced).  A coordinate whose
central differences at h and 2h disagree sits on a branch threshold (volume
thresholds, max(0, .), coth clamp) and is skipped -- the gradient there is
one-sided by definition.  Loss: L = sum (R - T)^2 over the composite."""
import json
import os

import numpy as np
import pytest

from painty_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H = 1e-6


def palette():
    d = json.load(open(os.path.join(ROOT, "painty_amd", "data", "lindemeier-measured.json")))
    return (np.array([p["K"] for p in d], np.float64), np.array([p["S"] for p in d], np.float64))


def grad_case(seed, snapshot=True, pickup=True, rows=26, cols=30, radius=4, n=18, n_paint=4,
              reach=0.4):
    wl = W.small(rows, cols, n, radius, seed, pickup=pickup, noisy=True, reach=reach)
    palK, palS = palette()
    rng = np.random.default_rng(seed)
    Wt = rng.dirichlet(np.ones(len(palK)), size=n_paint)
    st = wl.stamps.copy()
    st["flags"][0] |= W.DIP
    dips = np.flatnonzero(st["flags"] & W.DIP)
    paint_of = np.full(len(st), -1, np.int32)
    paint_of[dips] = np.arange(len(dips)) % n_paint
    T = rng.uniform(0.0, 1.0, (rows, cols, 3))
    R0 = wl.R0.astype(np.float64)
    return dict(radius=wl.radii[0], core=wl.cores[0].astype(np.float64),
                rates=(wl.pickup_rate, wl.deposition_rate), snapshot=snapshot,
                K=wl.wet.K.astype(np.float64), S=wl.wet.S.astype(np.float64),
                V=wl.wet.V.astype(np.float64), R0=R0, stamps=st, paint_of=paint_of,
                palK=palK, palS=palS, W=Wt, T=T)


def run(oracle, c, **over):
    a = dict(c, **over)
    out, final = oracle.grad_run(a["radius"], a["core"], a["rates"], a["snapshot"], a["K"], a["S"],
                                 a["V"], a["R0"], a["stamps"], a["paint_of"], a["palK"],
                                 a["palS"], a["W"], np.zeros_like(a["R0"]))
    return out[0]


def loss(oracle, c, **over):
    R = run(oracle, c, **over)
    return float(((R - c["T"]) ** 2).sum())


def analytic(oracle, c):
    R = run(oracle, c)
    gR = 2.0 * (R - c["T"])
    out, final = oracle.grad_run(c["radius"], c["core"], c["rates"], c["snapshot"], c["K"], c["S"],
                                 c["V"], c["R0"], c["stamps"], c["paint_of"], c["palK"],
                                 c["palS"], c["W"], gR)
    _, gK, gS, gV, gR0, gW = out
    return {"K": gK, "S": gS, "V": gV, "R0": gR0, "W": gW}, final


def central(oracle, c, name, idx, h):
    """(L(x+h) - L(x-h)) / 2h as sum (R+ - R-)(R+ + R- - 2T) / 2h, which
    keeps the round-off of the large sum L out of the difference; and the
    round-off floor of that estimate (pixels whose R moved)."""
    def at(delta):
        arr = c[name].copy()
        arr[idx] += delta
        return run(oracle, c, **{name: arr})
    Rp, Rm = at(h), at(-h)
    w = Rp + Rm - 2.0 * c["T"]
    d = float(((Rp - Rm) * w).sum()) / (2 * h)
    moved = Rp != Rm
    floor = 8 * np.finfo(np.float64).eps * float(np.abs(w[moved]).sum()) / (2 * h)
    return d, floor


def fd_check(oracle, c, coords, need=64):
    g, _ = analytic(oracle, c)
    ok = 0
    worst = 0.0
    for name, idx in coords:
        d1, f1 = central(oracle, c, name, idx, H)
        d2, f2 = central(oracle, c, name, idx, 2 * H)
        floor = f1 + 1e-12
        if abs(d1 - d2) > 1e-6 * abs(d1) + 4 * floor:  # a branch threshold inside [-2h, 2h]
            continue
        err = abs(g[name][idx] - d1)
        worst = max(worst, err / max(abs(d1), floor))
        assert err < 1e-6 * abs(d1) + floor, (name, idx, g[name][idx], d1, floor)
        ok += 1
    assert ok >= need, "only %d smooth coordinates" % ok
    print("fd: %d coordinates, worst rel %.2e" % (ok, worst))


def sample_coords(c, final, seed, n_px=14):
    rng = np.random.default_rng(seed + 99)
    rows, cols = c["V"].shape
    touched = np.argwhere(np.abs(final[2] - c["V"]) > 0)  # pixels the stamps reached
    pick = touched[rng.choice(len(touched), size=min(n_px, len(touched)), replace=False)]
    coords = []
    for (y, x) in pick:
        ch = int(rng.integers(0, 3))
        coords += [("K", (y, x, ch)), ("S", (y, x, ch)), ("V", (y, x)), ("R0", (y, x, ch))]
    used = sorted(set(int(p) for p in c["paint_of"] if p >= 0))
    for p in used:
        for l in rng.choice(c["W"].shape[1], size=4, replace=False):
            coords.append(("W", (p, int(l))))
    return coords


@pytest.mark.parametrize("snapshot,pickup", [(True, True), (False, True), (True, False)])
@pytest.mark.parametrize("seed", [3, 8])
def test_oracle_gradient_matches_central_differences(oracle, seed, snapshot, pickup):
    c = grad_case(seed, snapshot=snapshot, pickup=pickup)
    _, final = analytic(oracle, c)
    fd_check(oracle, c, sample_coords(c, final, seed))


def test_oracle_gradient_edge_stamps(oracle):
    """Stamps with fractional centres near column/row 0 (shared pixels)."""
    c = grad_case(5, rows=20, cols=22, radius=3, n=6)
    st = c["stamps"][:6].copy()
    st["cx"] = [0.4, 1.7, 2.6, 0.3, 4.2, 1.1]
    st["cy"] = [0.6, 0.2, 3.3, 5.5, 0.7, 1.5]
    st["theta"] = [0.9, 1.1, 2.0, 0.2, 0.77, 1.4]
    st["flags"] = 0
    st["flags"][0] = W.DIP
    st["flags"][3] = W.DIP
    c["stamps"] = st
    c["paint_of"] = np.array([0, -1, -1, 1, -1, -1], np.int32)
    _, final = analytic(oracle, c)
    fd_check(oracle, c, sample_coords(c, final, 5, n_px=16), need=48)


def test_composite_adjoint_matches_central_differences(oracle):
    rng = np.random.default_rng(1)
    n = 0
    for _ in range(200):
        K = np.exp(rng.uniform(np.log(0.01), np.log(10), 3))
        S = np.exp(rng.uniform(np.log(0.1), np.log(5), 3))
        R0 = rng.uniform(0.7, 0.95, 3)
        d = rng.uniform(0.0, 0.5)
        gR = rng.normal(size=3)
        gK, gS, gd, gR0 = oracle.compute_reflectance_adj(K, S, R0, d, gR)
        f = lambda K=K, S=S, R0=R0, d=d: float(gR @ oracle.compute_reflectance(K, S, R0, d))
        for name, arr, g in (("K", K, gK), ("S", S, gS), ("R0", R0, gR0)):
            for i in range(3):
                e = np.zeros(3)
                e[i] = H
                kw = {name: arr + e}
                kw2 = {name: arr - e}
                fd = (f(**kw) - f(**kw2)) / (2 * H)
                assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[i])
                n += 1
        fd = (f(d=d + H) - f(d=d - H)) / (2 * H)
        assert abs(fd - gd) < 1e-6 * max(1.0, abs(fd))
    assert n == 200 * 9


def test_gradient_is_linear_in_upstream(oracle):
    """<gR, R> is linear in gR: grad(a gR1 + b gR2) = a grad(gR1) + b grad(gR2)."""
    c = grad_case(4)
    rng = np.random.default_rng(2)
    gs = []
    for gR in (rng.normal(size=c["R0"].shape), rng.normal(size=c["R0"].shape)):
        out, _ = oracle.grad_run(c["radius"], c["core"], c["rates"], True, c["K"], c["S"], c["V"],
                                 c["R0"], c["stamps"], c["paint_of"], c["palK"], c["palS"], c["W"],
                                 gR)
        gs.append((gR, out[1:]))
    (g1, a1), (g2, a2) = gs
    out, _ = oracle.grad_run(c["radius"], c["core"], c["rates"], True, c["K"], c["S"], c["V"],
                             c["R0"], c["stamps"], c["paint_of"], c["palK"], c["palS"], c["W"],
                             2.0 * g1 - 0.5 * g2)
    for x, y1, y2 in zip(out[1:], a1, a2):
        np.testing.assert_allclose(x, 2.0 * y1 - 0.5 * y2, rtol=1e-9, atol=1e-12)
