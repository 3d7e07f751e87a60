"""An independent restatement of TextureBrush::paintStroke with Smudge
(renderer/TextureBrush.hxx:52-205, renderer/Smudge.hxx:40-198), vectorised in
numpy, used only to pin the C oracle on small cases.

It is written from the reference's definitions, not from the oracle's code:
the Catmull-Rom basis is the textbook tau = 1/2 form expanded per power of t
(core/Spline.hxx:22-73 gives the same polynomials), the mean value
coordinates are evaluated for all pixels at once from the signed-area form
of Hormann and Floater (core/Math.hxx:73-133), and the smudge's rotation and
exchange act on whole maps per spine step (each map cell meets a distinct
canvas pixel, so the per-cell loops of Smudge.hxx are order-free).
"""
import math

import numpy as np


def cr(P: np.ndarray, u: float, deriv: bool) -> np.ndarray:
    n = len(P)
    x = (n - 1) * u
    i = int(x)
    t = x - math.floor(x)
    idx = np.clip([i - 1, i, i + 1, i + 2], 0, n - 1)
    if deriv:
        b = 0.5 * np.array([-1 + 4 * t - 3 * t * t, -10 * t + 9 * t * t,
                            1 + 8 * t - 9 * t * t, -2 * t + 3 * t * t])
    else:
        b = 0.5 * np.array([-t + 2 * t * t - t ** 3, 2 - 5 * t * t + 3 * t ** 3,
                            t + 4 * t * t - 3 * t ** 3, -t * t + t ** 3])
    return b @ P[idx]


def unit(v: np.ndarray) -> np.ndarray:
    n = math.hypot(v[0], v[1])
    return v / n if n > 0 else v


def mvc(poly: np.ndarray, vals: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """Mean value interpolation of vals (np, 2) at pts (m, 2)."""
    eps = np.finfo(float).eps * 100
    s = poly[None, :, :] - pts[:, None, :]
    s1 = np.roll(s, -1, axis=1)
    r = np.hypot(s[..., 0], s[..., 1])
    r1 = np.roll(r, -1, axis=1)
    A = 0.5 * (s[..., 0] * s1[..., 1] - s1[..., 0] * s[..., 1])
    D = (s * s1).sum(-1)
    Am, rm, Dm = np.roll(A, 1, axis=1), np.roll(r, 1, axis=1), np.roll(D, 1, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(Am != 0, (rm - Dm / r) / Am, 0.0) + np.where(A != 0, (r1 - D / r) / A, 0.0)
        W = w.sum(1)
        out = (w @ vals) / W[:, None]
    out[np.abs(W) < eps] = vals[0]
    # a point on a vertex or on an edge: the first such vertex / edge in order
    for m in range(len(pts)):
        for i in range(len(poly)):
            if abs(r[m, i]) < eps:
                out[m] = vals[i]
                break
            if abs(A[m, i]) < eps and D[m, i] < 0:
                j = (i + 1) % len(poly)
                out[m] = (r1[m, i] * vals[i] + r[m, i] * vals[j]) / (r[m, i] + r1[m, i])
                break
    return out


def reflect(p: np.ndarray, n: int) -> np.ndarray:
    p = np.asarray(p)
    if n == 1:
        return np.zeros_like(p)
    period = 2 * n
    q = np.mod(p, period)
    return np.where(q < n, q, period - 1 - q)


def bilinear(img: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """img (rows, cols[, c]) at (x, y), reflected borders (cv BORDER_REFLECT)."""
    rows, cols = img.shape[:2]
    xi, yi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    a, c = x - xi, y - yi
    x0, x1 = reflect(xi, cols), reflect(xi + 1, cols)
    y0, y1 = reflect(yi, rows), reflect(yi + 1, rows)
    if img.ndim == 3:
        a, c = a[..., None], c[..., None]
    return ((img[y0, x0] * (1 - a) + img[y0, x1] * a) * (1 - c) +
            (img[y1, x0] * (1 - a) + img[y1, x1] * a) * c)


class BruteTextureBrush:
    def __init__(self, thick: np.ndarray):
        self.radius = 0.0
        self.K = np.full(3, 0.1)
        self.S = np.full(3, 0.1)
        self.thick = np.asarray(thick, float)
        self.scale = 1.0
        self.smudge_on = True
        self._reset(0)

    def _reset(self, size: int) -> None:
        self.n = max(size, 0)
        self.max_size = size + 1 if size % 2 == 0 else size
        self.mK = np.zeros((self.n, self.n, 3))
        self.mS = np.zeros((self.n, self.n, 3))
        self.mV = np.zeros((self.n, self.n))
        self.rot = 0.0

    def set_radius(self, r: float) -> None:
        if not abs(self.radius - r) < 0.5:
            self.radius = r
            if self.smudge_on:
                self._reset(int(2.0 * r))

    def _turn(self, heading: np.ndarray) -> None:
        theta = math.atan2(heading[1], heading[0])
        d = theta - self.rot
        while d <= -math.pi / 2:
            d += math.pi
        while d > math.pi / 2:
            d -= math.pi
        self.rot = theta
        n = self.n
        c = self.max_size / 2.0
        y, x = np.mgrid[0:n, 0:n].astype(float)
        rx = (x - c) * math.cos(d) - (y - c) * math.sin(d) + c
        ry = (x - c) * math.sin(d) + (y - c) * math.cos(d) + c
        inside = (rx >= 0) & (ry >= 0) & (rx < n) & (ry < n)
        for name in ("mK", "mS", "mV"):
            m = getattr(self, name)
            r = bilinear(m, np.where(inside, rx, 0), np.where(inside, ry, 0))
            msk = inside[..., None] if m.ndim == 3 else inside
            setattr(self, name, np.where(msk, r, m))

    def paint_stroke(self, xy, cv) -> None:
        """cv: object with K, S (rows, cols, 3) and V (rows, cols) float64."""
        P = np.asarray(xy, float).reshape(-1, 2)
        if len(P) < 2:
            return
        r = self.radius
        rows, cols = cv.V.shape
        v = np.vstack([P[0] - unit(P[1] - P[0]) * r, P, P[-1] + unit(P[-1] - P[-2]) * r])
        bmin = np.maximum(v.min(0) - r, 0.0)
        bmax = np.minimum(v.max(0) + r, [cols - 1.0, rows - 1.0])
        length = float(np.hypot(*np.diff(v, axis=0).T).sum())
        nv = len(v)
        up, down = [], []
        for i in range(nv):
            u = i / (nv - 1)
            c, t = cr(v, u, False), unit(cr(v, u, True))
            nrm = np.array([-t[1], t[0]])
            up.append((c - r * nrm, (u, 0.0)))
            down.append((c + r * nrm, (u, 1.0)))
        ring = down[::-1] + up
        poly = np.array([p for p, _ in ring])
        uv = np.array([q for _, q in ring])
        xs = np.arange(int(bmin[0]), int(bmax[0]) + 1)
        ys = np.arange(int(bmin[1]), int(bmax[1]) + 1)
        X, Y = np.meshgrid(xs, ys)
        X, Y = X.ravel(), Y.ravel()
        ok = (X >= 0) & (X < cols) & (Y >= 0) & (Y < rows)
        X, Y = X[ok], Y[ok]
        tp = mvc(poly, uv, np.stack([X, Y], 1).astype(float))
        ins = (tp >= 0).all(1) & (tp <= 1).all(1)
        X, Y, tp = X[ins], Y[ins], tp[ins]
        tr, tc = self.thick.shape
        vt = self.scale * bilinear(self.thick, tp[:, 0] * tc, tp[:, 1] * tr)
        trows, tcols = int(bmax[1] - bmin[1] + 1), int(bmax[0] - bmin[0] + 1)
        tmap = np.zeros((trows, tcols))
        s, t = X - int(bmin[0]), Y - int(bmin[1])
        keep = (vt > 0) & (s >= 0) & (t >= 0) & (s < tcols) & (t < trows)
        tmap[t[keep], s[keep]] = vt[keep]
        if self.smudge_on and self.n > 0 and tmap.max() > 0:
            self._smudge(cv, bmin, v, length, tmap)
        ys_, xs_ = np.nonzero(tmap > 0)
        Xp, Yp = xs_ + int(bmin[0]), ys_ + int(bmin[1])
        okp = (Xp < cols) & (Yp < rows)
        Xp, Yp, Vt = Xp[okp], Yp[okp], tmap[ys_[okp], xs_[okp]]
        Vc = cv.V[Yp, Xp]
        Vs = Vc + Vt
        m = Vs > 0
        Xp, Yp, Vt, Vc, Vs = Xp[m], Yp[m], Vt[m], Vc[m], Vs[m]
        cv.K[Yp, Xp] = (Vc[:, None] * cv.K[Yp, Xp] + Vt[:, None] * self.K) / Vs[:, None]
        cv.S[Yp, Xp] = (Vc[:, None] * cv.S[Yp, Xp] + Vt[:, None] * self.S) / Vs[:, None]
        cv.V[Yp, Xp] = np.maximum(Vt, Vc)

    def _smudge(self, cv, bmin, v, length, tmap) -> None:
        rows, cols = cv.V.shape
        maxD = tmap.max()
        n = self.n
        y, x = np.mgrid[0:n, 0:n]
        u = 0.0
        while u <= 1.0:
            center = cr(v, u, False)
            self._turn(unit(cr(v, u, True)))
            cx = x + int(center[0] - n / 2.0)
            cy = y + int(center[1] - n / 2.0)
            tx = (cx - bmin[0]).astype(np.int64)
            ty = (cy - bmin[1]).astype(np.int64)
            ok = (cx >= 0) & (cy >= 0) & (cx < cols) & (cy < rows)
            ok &= (tx >= 0) & (ty >= 0) & (tx < tmap.shape[1]) & (ty < tmap.shape[0])
            ok &= np.hypot(center[0] - cx, center[1] - cy) <= self.max_size * 0.5
            Dv = np.where(ok, tmap[np.clip(ty, 0, tmap.shape[0] - 1),
                                   np.clip(tx, 0, tmap.shape[1] - 1)], 0.0)
            ok &= Dv > 0
            my, mx = np.nonzero(ok)
            py, px = cy[ok], cx[ok]
            f = Dv[ok] / maxD
            cV, pV = cv.V[py, px], self.mV[my, mx]
            cVl, pVl = 0.1 * cV * f, 0.1 * pV * f
            cVr, pVr = cV - cVl, pV - pVl
            cK, cS = cv.K[py, px], cv.S[py, px]
            pK, pS = self.mK[my, mx], self.mS[my, mx]
            pn, cn = pVr + cVl, cVr + pVl
            a = pn > 0.001
            self.mK[my[a], mx[a]] = ((pVr[:, None] * pK + cVl[:, None] * cK) / pn[:, None])[a]
            self.mS[my[a], mx[a]] = ((pVr[:, None] * pS + cVl[:, None] * cS) / pn[:, None])[a]
            self.mV[my[a], mx[a]] = pn[a]
            b = cn > 0.001
            cv.K[py[b], px[b]] = ((cVr[:, None] * cK + pVl[:, None] * pK) / cn[:, None])[b]
            cv.S[py[b], px[b]] = ((cVr[:, None] * cS + pVl[:, None] * pS) / cn[:, None])[b]
            cv.V[py[b], px[b]] = cn[b]
            u += 1.0 / length
