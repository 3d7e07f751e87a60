"""Second, independent restatement of FootprintBrush::imprint for tiny inputs.

Pure Python over dicts and lists, written from renderer/FootprintBrush.hxx
(imprint :73-143, updateSnapshot :278-319, blend :331-340, pickupPaint
:349-384, depositPaint :393-431) without looking at the C oracle, so that a
transcription mistake in either shows up as a disagreement.  Slow by design:
only for canvases up to ~64x64 and a handful of stamps.
"""
import math

MIN_VOLUME = 0.001


def _blend(va, a, vb, b):
    vt = va + vb
    if vt > MIN_VOLUME:
        return [(va * x + vb * y) / vt for x, y in zip(a, b)]
    return list(a)


class BruteBrush:
    def __init__(self, radius, core):
        width = int(2.0 * math.ceil(radius) + 1.0)
        assert len(core) == width
        self.radius = radius
        self.size_map = int(math.ceil(math.sqrt(2.0) * width))
        pad = (self.size_map - width) // 2
        self.fp = {}
        for i in range(self.size_map):
            for j in range(self.size_map):
                ii, jj = i - pad, j - pad
                inside = 0 <= ii < width and 0 <= jj < width
                self.fp[(i, j)] = float(core[ii][jj]) if inside else 0.0
        self.pick = {}  # (row, col) -> [K(3), S(3), V]
        self.clean()
        self.snap = None  # dict (row, col) -> [K, S, V]
        self.snap_shape = None
        self.use_snapshot = True
        self.cap = 1.0
        self.pickup_rate = 0.9
        self.deposition_rate = 0.05
        self.paint = ([0.0] * 3, [0.0] * 3)

    def clean(self):
        for i in range(self.size_map):
            for j in range(self.size_map):
                self.pick[(i, j)] = [[0.0] * 3, [0.0] * 3, 0.0]

    def dip(self, K, S):
        self.clean()
        self.paint = ([float(x) for x in K], [float(x) for x in S])

    def _copy_full(self, canvas):
        self.snap = {k: [list(v[0]), list(v[1]), v[2]] for k, v in canvas.cells.items()}
        self.snap_shape = (canvas.rows, canvas.cols)

    def imprint(self, cx, cy, theta, canvas, clock=None):
        """clock: (tmap, now_us, dry_ms, R0, h, km) -- dicts keyed (row, col)
        for the time map (microseconds), substrate and height, and a KM
        function km(K, S, R0, d) -> R; checkDry then runs on every visited
        pixel before its pickup (FootprintBrush.hxx:115)."""
        hr = (self.size_map - 1) // 2
        if self.use_snapshot:
            if self.snap_shape != (canvas.rows, canvas.cols):
                self._copy_full(canvas)
            tl = (int(cx - hr), int(cy - hr))
            br = (int(cx + hr), int(cy + hr))
            a0 = (max(int(cx - hr - self.radius), 0), max(int(cy - hr - self.radius), 0))
            a1 = (min(int(cx + hr + self.radius), canvas.cols - 1),
                  min(int(cy + hr + self.radius), canvas.rows - 1))
            for row in range(a0[1], a1[1] + 1):
                for col in range(a0[0], a1[0] + 1):
                    if tl[1] < row < br[1] and tl[0] < col < br[0]:
                        continue
                    v = canvas.cells[(row, col)]
                    self.snap[(row, col)] = [list(v[0]), list(v[1]), v[2]]
            src = self.snap
        else:
            src = canvas.cells
        ct, st = math.cos(-theta), math.sin(-theta)
        for row in range(-hr, hr + 1):
            for col in range(-hr, hr + 1):
                x, y = int(col + cx), int(row + cy)
                mx = int(round_half_away(col * ct - row * st + hr))
                my = int(round_half_away(col * st + row * ct + hr))
                if not (0 <= x < canvas.cols and 0 <= y < canvas.rows):
                    continue
                if not (0 <= mx < self.size_map and 0 <= my < self.size_map):
                    continue
                if clock is not None:
                    _check_dry(clock, canvas.cells[(y, x)], (y, x))
                fh = self.fp[(my, mx)]
                if fh <= 0.0:
                    continue
                p = self.pick[(my, mx)]
                s = src[(y, x)]
                # pickup
                v_leave = self.pickup_rate * s[2] * fh
                if v_leave > MIN_VOLUME:
                    s[2] = s[2] - v_leave
                    p[0] = _blend(p[2], p[0], v_leave, s[0])
                    p[1] = _blend(p[2], p[1], v_leave, s[1])
                    p[2] = p[2] + v_leave
                # deposit
                v_free = max(0.0, self.cap - p[2])
                k_src = _blend(p[2], p[0], v_free, self.paint[0])
                s_src = _blend(p[2], p[1], v_free, self.paint[1])
                p[2] = p[2] - self.deposition_rate * p[2] * fh
                c = canvas.cells[(y, x)]
                vb = self.cap * fh
                c[0] = _blend(vb, k_src, c[2], c[0])
                c[1] = _blend(vb, s_src, c[2], c[1])
                c[2] = vb + c[2]


def _check_dry(clock, cell, key):
    """Canvas::checkDry (Canvas.hxx:123-155): a pixel whose paint is older
    than the drying time dries into the substrate; its time becomes now."""
    tmap, now_us, dry_ms, R0, h, km = clock
    v = cell[2]
    if dry_ms > 0 and v > 0.001:
        elapsed_ms = int((now_us - tmap[key]) / 1000)  # truncation toward zero
        if elapsed_ms >= dry_ms:
            h[key] += v
            R0[key] = list(km(cell[0], cell[1], R0[key], v))
            cell[0] = [0.0, 0.0, 0.0]
            cell[1] = [0.0, 0.0, 0.0]
            cell[2] = 0.0
        # otherwise the reference's partial branch needs elapsed_ms / dry_ms
        # (an integer quotient, 0 here) above 0.01: never
    tmap[key] = now_us


def round_half_away(v):
    """C round(): nearest integer, halves away from zero (exact: v - floor(v)
    is exact for |v| < 2**52)."""
    r = math.floor(v)
    frac = v - r
    if frac > 0.5 or (frac == 0.5 and v >= 0):
        return r + 1
    return r


class BruteCanvas:
    def __init__(self, K, S, V):
        self.rows, self.cols = len(V), len(V[0])
        self.cells = {}
        for r in range(self.rows):
            for c in range(self.cols):
                self.cells[(r, c)] = [[float(x) for x in K[r][c]], [float(x) for x in S[r][c]],
                                      float(V[r][c])]

    def arrays(self):
        import numpy as np
        K = np.zeros((self.rows, self.cols, 3))
        S = np.zeros((self.rows, self.cols, 3))
        V = np.zeros((self.rows, self.cols))
        for (r, c), v in self.cells.items():
            K[r, c], S[r, c], V[r, c] = v[0], v[1], v[2]
        return K, S, V
