"""ctypes view of the FP64 CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker / the timed CPU
baseline.  The product (``painty_amd``) never imports this module.

Arrays follow the reference's ``cv::Mat_`` storage: K, S, R0 are ``(rows,
cols, 3)`` float64 (AoS ``vec3``), V and h are ``(rows, cols)`` float64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PAINTY_ORACLE_LIB selects another build of the same sources (sanitize.sh
# points it at the AddressSanitizer / UndefinedBehaviorSanitizer build)
_LIB_PATH = os.environ.get("PAINTY_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class Stamp(C.Structure):
    """orc_stamp: one imprint with optional dip before and dry after."""

    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("theta", C.c_double),
                ("K", C.c_double * 3), ("S", C.c_double * 3),
                ("flags", C.c_uint32)]


class Clock(C.Structure):
    """orc_clock: time map (microseconds), time point, drying time (ms), R0, h."""

    _fields_ = [("tmap", C.POINTER(C.c_int64)), ("now_us", C.c_int64), ("dry_ms", C.c_int64),
                ("R0", C.POINTER(C.c_double)), ("h", C.POINTER(C.c_double))]


DIP = 1
DRY_AFTER = 2


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_coth.restype = C.c_double
        L.orc_coth.argtypes = [C.c_double]
        L.orc_acoth.restype = C.c_double
        L.orc_acoth.argtypes = [C.c_double]
        L.orc_compute_reflectance.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
        L.orc_compute_reflectance_glsl.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
        L.orc_compute_scattering_absorption.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_compute_scattering_absorption.restype = C.c_int
        L.orc_compose.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.orc_compose_onto.argtypes = [C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_dry_canvas.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.orc_dry_step.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double]
        L.orc_brush_create.restype = C.c_void_p
        L.orc_brush_create.argtypes = [C.c_double, _dp, C.c_int]
        L.orc_brush_destroy.argtypes = [C.c_void_p]
        L.orc_brush_set_radius.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int]
        L.orc_brush_set_radius.restype = C.c_int
        L.orc_brush_size_map.argtypes = [C.c_void_p]
        L.orc_brush_size_map.restype = C.c_int
        L.orc_brush_footprint.argtypes = [C.c_void_p]
        L.orc_brush_footprint.restype = _dp
        L.orc_brush_set_rates.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_brush_set_use_snapshot.argtypes = [C.c_void_p, C.c_int]
        L.orc_brush_dip.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_brush_clean.argtypes = [C.c_void_p]
        L.orc_brush_update_snapshot.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_brush_get_pickup.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_brush_imprint.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp,
                                        C.c_double, C.c_double, C.c_double]
        L.orc_run_stamps.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                     C.POINTER(Stamp), C.c_int64]
        L.orc_render.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_render_pixel_normal.argtypes = [C.c_double] * 4 + [_dp]
        L.orc_rgb2srgb.restype = C.c_double
        L.orc_rgb2srgb.argtypes = [C.c_double]
        L.orc_encode.argtypes = [C.c_int64, _dp, C.c_int, C.c_int, C.POINTER(C.c_uint16)]
        L.orc_brush_imprint_timed.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp,
                                              C.c_double, C.c_double, C.c_double,
                                              C.POINTER(Clock)]
        L.orc_tbrush_create.restype = C.c_void_p
        L.orc_tbrush_destroy.argtypes = [C.c_void_p]
        L.orc_tbrush_set_radius.argtypes = [C.c_void_p, C.c_double]
        L.orc_tbrush_dip.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tbrush_set_thickness_map.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
        L.orc_tbrush_set_thickness_scale.argtypes = [C.c_void_p, C.c_double]
        L.orc_tbrush_set_use_smudge.argtypes = [C.c_void_p, C.c_int]
        L.orc_tbrush_get_smudge_map.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp, _dp, _dp]
        L.orc_tbrush_paint_stroke.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp,
                                              _dp, C.c_int, C.POINTER(Clock)]
        L.orc_gbc.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_tbrush_set_smudge_map.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_tbrush_turn.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_srgb2rgb.restype = C.c_double
        L.orc_srgb2rgb.argtypes = [C.c_double]
        L.orc_gray_from_u16.argtypes = [C.c_int64, C.POINTER(C.c_uint16), C.c_int, C.c_int, _dp]
        L.orc_lanczos4_coeffs.argtypes = [C.c_float, C.POINTER(C.c_float)]
        L.orc_resize_lanczos4.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_tbgl_create.restype = C.c_void_p
        L.orc_tbgl_destroy.argtypes = [C.c_void_p]
        L.orc_tbgl_set_radius.argtypes = [C.c_void_p, C.c_double]
        L.orc_tbgl_dip.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tbgl_set_texture.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
        L.orc_tbgl_set_thickness_scale.argtypes = [C.c_void_p, C.c_double]
        L.orc_tbgl_enable_smudge.argtypes = [C.c_void_p, C.c_int]
        _f = C.POINTER(C.c_float)
        L.orc_tbgl_get_smudge.argtypes = [C.c_void_p, C.POINTER(C.c_int), _f, _f]
        L.orc_tbgl_paint_stroke.argtypes = [C.c_void_p, C.c_int, C.c_int, _f, _f, _f, _f, _dp,
                                            C.c_int]
        L.orc_mix_single.restype = C.c_int
        L.orc_mix_single.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_mixed.argtypes = [_dp, _dp, C.c_double, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_catmull_rom.restype = C.c_double
        L.orc_catmull_rom.argtypes = [C.c_double] * 5
        L.orc_catmull_rom_d1.restype = C.c_double
        L.orc_catmull_rom_d1.argtypes = [C.c_double] * 5
        L.orc_stroke_stamps.restype = C.c_int64
        L.orc_stroke_stamps.argtypes = [_dp, C.c_int64, _dp, C.c_int64]
        _ip = C.POINTER(C.c_int)
        L.orc_grad_run.restype = C.c_int
        L.orc_grad_run.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                   C.POINTER(Stamp), C.c_int64, _ip, C.c_int, _dp, _dp, C.c_int,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_compute_reflectance_adj.argtypes = [_dp, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                                  _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- KM -------
def coth(x: float) -> float:
    return lib().orc_coth(float(x))


def acoth(x: float) -> float:
    return lib().orc_acoth(float(x))


def compute_reflectance(K, S, R0, d) -> np.ndarray:
    K, S, R0 = _f64(K), _f64(S), _f64(R0)
    R = np.empty(3)
    lib().orc_compute_reflectance(_p(K), _p(S), _p(R0), float(d), _p(R))
    return R


def compute_reflectance_glsl(K, S, R0, d) -> np.ndarray:
    K, S, R0 = _f64(K), _f64(S), _f64(R0)
    R = np.empty(3)
    lib().orc_compute_reflectance_glsl(_p(K), _p(S), _p(R0), float(d), _p(R))
    return R


def compute_scattering_absorption(Rb, Rw):
    Rb, Rw = _f64(Rb), _f64(Rw)
    K, S = np.empty(3), np.empty(3)
    if lib().orc_compute_scattering_absorption(_p(Rb), _p(Rw), _p(K), _p(S)) != 0:
        raise ValueError("on black or white inputs violate one of the following "
                         "conditions: 0 < black < white < 1")
    return K, S


def compose(K, S, V, R0) -> np.ndarray:
    K, S, V, R0 = _f64(K), _f64(S), _f64(V), _f64(R0)
    n = V.size
    assert K.size == S.size == R0.size == 3 * n
    R = np.empty_like(R0)
    lib().orc_compose(n, _p(K), _p(S), _p(V), _p(R0), _p(R))
    return R


def compose_onto(K, S, V, R0: np.ndarray) -> None:
    """In place on R0 (float64, C-contiguous)."""
    K, S, V = _f64(K), _f64(S), _f64(V)
    lib().orc_compose_onto(V.size, _p(K), _p(S), _p(V), _p(R0))


def dry_canvas(K, S, V, R0, h) -> None:
    """In place on all five arrays (float64, C-contiguous)."""
    lib().orc_dry_canvas(V.size, _p(K), _p(S), _p(V), _p(R0), _p(h))


def dry_step(K, S, V, R0, portion: float) -> None:
    lib().orc_dry_step(V.size, _p(_f64(K)), _p(_f64(S)), _p(V), _p(R0), float(portion))


# ------------------------------------------------------------- brush -------
class Brush:
    """FootprintBrush with an injected footprint core (width x width)."""

    def __init__(self, radius: float, core: np.ndarray):
        core = _f64(core)
        self._core = core
        self.h = lib().orc_brush_create(float(radius), _p(core), int(core.shape[0]))
        if not self.h:
            raise ValueError("footprint core width must be 2*ceil(radius)+1")
        self.radius = radius

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_brush_destroy(self.h)
            self.h = None

    def set_radius(self, radius: float, core: np.ndarray) -> None:
        core = _f64(core)
        if lib().orc_brush_set_radius(self.h, float(radius), _p(core), int(core.shape[0])) != 0:
            raise ValueError("footprint core width must be 2*ceil(radius)+1")
        self._core = core
        self.radius = radius

    @property
    def size_map(self) -> int:
        return lib().orc_brush_size_map(self.h)

    def footprint(self) -> np.ndarray:
        m = self.size_map
        ptr = lib().orc_brush_footprint(self.h)
        return np.ctypeslib.as_array(ptr, shape=(m, m)).copy()

    def set_rates(self, pickup: float, deposition: float) -> None:
        lib().orc_brush_set_rates(self.h, pickup, deposition)

    def set_use_snapshot(self, use: bool) -> None:
        lib().orc_brush_set_use_snapshot(self.h, int(bool(use)))

    def dip(self, K, S) -> None:
        lib().orc_brush_dip(self.h, _p(_f64(K)), _p(_f64(S)))

    def clean(self) -> None:
        lib().orc_brush_clean(self.h)

    def update_snapshot(self, canvas: "Canvas") -> None:
        lib().orc_brush_update_snapshot(self.h, canvas.rows, canvas.cols,
                                        _p(canvas.K), _p(canvas.S), _p(canvas.V))

    def pickup_map(self):
        m = self.size_map
        K, S, V = np.empty((m, m, 3)), np.empty((m, m, 3)), np.empty((m, m))
        lib().orc_brush_get_pickup(self.h, _p(K), _p(S), _p(V))
        return K, S, V

    def imprint(self, cx: float, cy: float, theta: float, canvas: "Canvas") -> None:
        lib().orc_brush_imprint(self.h, canvas.rows, canvas.cols, _p(canvas.K),
                                _p(canvas.S), _p(canvas.V), float(cx), float(cy),
                                float(theta))

    def imprint_timed(self, cx: float, cy: float, theta: float, canvas: "Canvas",
                      now_us: int, dry_ms: int) -> None:
        """imprint with Canvas::checkDry on every visited pixel; canvas.tmap
        (int64 microseconds) is the time map."""
        assert canvas.tmap.dtype == np.int64 and canvas.tmap.flags.c_contiguous
        clk = Clock(canvas.tmap.ctypes.data_as(C.POINTER(C.c_int64)), int(now_us), int(dry_ms),
                    _p(canvas.R0), _p(canvas.h))
        lib().orc_brush_imprint_timed(self.h, canvas.rows, canvas.cols, _p(canvas.K),
                                      _p(canvas.S), _p(canvas.V), float(cx), float(cy),
                                      float(theta), C.byref(clk))

    def run(self, canvas: "Canvas", stamps) -> None:
        arr = stamps_to_c(stamps)
        lib().orc_run_stamps(self.h, canvas.rows, canvas.cols, _p(canvas.K), _p(canvas.S),
                             _p(canvas.V), _p(canvas.R0), _p(canvas.h), arr, len(arr))


def stamps_to_c(stamps):
    """stamps: structured numpy array with fields cx, cy, theta, K(3), S(3), flags."""
    n = len(stamps)
    arr = (Stamp * n)()
    for i in range(n):
        s = stamps[i]
        arr[i].cx, arr[i].cy, arr[i].theta = float(s["cx"]), float(s["cy"]), float(s["theta"])
        for c in range(3):
            arr[i].K[c] = float(s["K"][c])
            arr[i].S[c] = float(s["S"][c])
        arr[i].flags = int(s["flags"])
    return arr


class Canvas:
    """Canvas state in float64 (K, S, V wet layer, R0 substrate, h height)."""

    def __init__(self, rows: int, cols: int):
        self.rows, self.cols = rows, cols
        self.K = np.zeros((rows, cols, 3))
        self.S = np.zeros((rows, cols, 3))
        self.V = np.zeros((rows, cols))
        self.R0 = np.ones((rows, cols, 3))
        self.h = np.zeros((rows, cols))
        self.tmap = np.zeros((rows, cols), np.int64)  # time map, microseconds

    @classmethod
    def from_arrays(cls, K, S, V, R0, h=None) -> "Canvas":
        rows, cols = V.shape
        c = cls(rows, cols)
        c.K[...] = K
        c.S[...] = S
        c.V[...] = V
        c.R0[...] = R0
        if h is not None:
            c.h[...] = h
        return c

    def compose(self) -> np.ndarray:
        return compose(self.K, self.S, self.V, self.R0)

    def dry(self) -> None:
        dry_canvas(self.K, self.S, self.V, self.R0, self.h)


# ------------------------------------------------------------- mixer -------
def mix_single(baseK, baseS, weights):
    baseK, baseS, w = _f64(baseK), _f64(baseS), _f64(weights)
    K, S = np.empty(3), np.empty(3)
    if lib().orc_mix_single(baseK.shape[0], _p(baseK), _p(baseS), w.size, _p(w),
                            _p(K), _p(S)) != 0:
        raise ValueError("Palette size does not match underlying size.")
    return K, S


def mixed(K1, S1, v1, K2, S2, v2):
    K, S = np.empty(3), np.empty(3)
    lib().orc_mixed(_p(_f64(K1)), _p(_f64(S1)), float(v1), _p(_f64(K2)), _p(_f64(S2)),
                    float(v2), _p(K), _p(S))
    return K, S


# --------------------------------------------------------- gradient -------
def compute_reflectance_adj(K, S, R0, d, gR):
    """Adjoint of ComputeReflectance for one pixel: (gK, gS, gd, gR0)."""
    K, S, R0, gR = _f64(K), _f64(S), _f64(R0), _f64(gR)
    gK, gS, gR0, gd = np.empty(3), np.empty(3), np.empty(3), C.c_double()
    lib().orc_compute_reflectance_adj(_p(K), _p(S), _p(R0), float(d), _p(gR), _p(gK), _p(gS),
                                      C.cast(C.pointer(gd), _dp), _p(gR0))
    return gK, gS, gd.value, gR0


def grad_run(radius, core, rates, use_snapshot, K, S, V, R0, stamps, paint_of, palK, palS, W,
             gR):
    """Forward composite+stamp with dip paints mixSinglePaint(W[paint_of[i]])
    and the gradient of <gR, R>.  Returns (R, gK, gS, gV, gR0, gW) and the
    final wet layer (K, S, V)."""
    b = Brush(radius, core)
    b.set_rates(*rates)
    b.set_use_snapshot(use_snapshot)
    K, S, V, R0, gR = (np.array(a, dtype=np.float64, order="C") for a in (K, S, V, R0, gR))
    palK, palS, W = _f64(palK), _f64(palS), _f64(W)
    rows, cols = V.shape
    pi = np.ascontiguousarray(paint_of, dtype=np.int32)
    R, gK, gS, gR0 = (np.empty((rows, cols, 3)) for _ in range(4))
    gV = np.empty((rows, cols))
    gW = np.empty(W.shape)
    arr = stamps_to_c(stamps)
    rc = lib().orc_grad_run(b.h, rows, cols, _p(K), _p(S), _p(V), _p(R0), arr, len(arr),
                            pi.ctypes.data_as(C.POINTER(C.c_int)), palK.shape[0], _p(palK),
                            _p(palS), W.shape[0], _p(W), _p(gR), _p(R), _p(gK), _p(gS), _p(gV),
                            _p(gR0), _p(gW))
    if rc != 0:
        raise ValueError("unsupported stamp list for the gradient run")
    return (R, gK, gS, gV, gR0, gW), (K, S, V)


# ------------------------------------------------------------ spline -------
def catmull_rom(p_1, p0, p1, p2, t) -> float:
    return lib().orc_catmull_rom(p_1, p0, p1, p2, t)


def catmull_rom_d1(p_1, p0, p1, p2, t) -> float:
    return lib().orc_catmull_rom_d1(p_1, p0, p1, p2, t)


def stroke_stamps(path) -> np.ndarray:
    path = _f64(path).reshape(-1, 2)
    n = lib().orc_stroke_stamps(_p(path), len(path), None, 0)
    out = np.empty((max(n, 1), 3))
    lib().orc_stroke_stamps(_p(path), len(path), _p(out), n)
    return out[:n]


# ------------------------------------------------------------ display -------
def render(K, S, V, R0) -> np.ndarray:
    """Renderer::render of the wet layer (K, S (rows, cols, 3), V (rows, cols))
    over R0 -> (rows, cols, 3) in [0, 1]."""
    K, S, V, R0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (K, S, V, R0))
    rows, cols = V.shape
    out = np.empty((rows, cols, 3))
    lib().orc_render(rows, cols, _p(K), _p(S), _p(V), _p(R0), _p(out))
    return out


def render_pixel_normal(s01, s21, s10, s12) -> np.ndarray:
    n = np.empty(3)
    lib().orc_render_pixel_normal(float(s01), float(s21), float(s10), float(s12), _p(n))
    return n


def rgb2srgb(l: float) -> float:
    return lib().orc_rgb2srgb(float(l))


def encode(rgb, to_srgb: bool = True, bits: int = 8) -> np.ndarray:
    a = np.ascontiguousarray(rgb, dtype=np.float64)
    out = np.empty(a.shape, np.uint16)
    lib().orc_encode(a.size // 3, _p(a), 1 if to_srgb else 0, int(bits),
                     out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return out.astype(np.uint8) if bits == 8 else out



# ------------------------------------------------------- texture brush -----
def gbc(poly, px: float, py: float, vals) -> np.ndarray:
    """Mean value coordinate interpolation (core/Math.hxx:73-133) of the 2-d
    values `vals` (np x 2) at the vertices of polygon `poly` (np x 2)."""
    poly, vals = _f64(poly), _f64(vals)
    out = np.empty(2)
    lib().orc_gbc(_p(poly), int(poly.shape[0]), float(px), float(py), _p(vals), _p(out))
    return out


class TextureBrush:
    """TextureBrush with Smudge (renderer/TextureBrush.hxx:19-238,
    renderer/Smudge.hxx:22-229) over an FP64 Canvas."""

    def __init__(self):
        self.h = lib().orc_tbrush_create()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tbrush_destroy(self.h)
            self.h = None

    def set_radius(self, radius: float) -> None:
        lib().orc_tbrush_set_radius(self.h, float(radius))

    def dip(self, K, S) -> None:
        lib().orc_tbrush_dip(self.h, _p(_f64(K)), _p(_f64(S)))

    def set_thickness_map(self, m) -> None:
        m = _f64(m)
        lib().orc_tbrush_set_thickness_map(self.h, _p(m), int(m.shape[0]), int(m.shape[1]))

    def set_thickness_scale(self, s: float) -> None:
        lib().orc_tbrush_set_thickness_scale(self.h, float(s))

    def enable_smudge(self, on: bool) -> None:
        lib().orc_tbrush_set_use_smudge(self.h, int(bool(on)))

    def smudge_map(self):
        side = C.c_int(0)
        lib().orc_tbrush_get_smudge_map(self.h, C.byref(side), None, None, None)
        n = side.value
        K, S, V = np.zeros((n, n, 3)), np.zeros((n, n, 3)), np.zeros((n, n))
        lib().orc_tbrush_get_smudge_map(self.h, C.byref(side), _p(K), _p(S), _p(V))
        return K, S, V

    def set_smudge_map(self, K, S, V) -> None:
        n = self.smudge_map()[2].shape[0]
        K, S, V = _f64(K), _f64(S), _f64(V)
        assert K.shape == (n, n, 3) and S.shape == (n, n, 3) and V.shape == (n, n)
        lib().orc_tbrush_set_smudge_map(self.h, _p(K), _p(S), _p(V))

    def turn(self, hx: float, hy: float) -> None:
        """Smudge::updateOrientation to the heading (hx, hy)."""
        lib().orc_tbrush_turn(self.h, float(hx), float(hy))

    def paint_stroke(self, xy, canvas: "Canvas", now_us=None, dry_ms=None) -> None:
        """paintStroke; with now_us / dry_ms, Canvas::checkDry runs on the
        textured pixels against canvas.tmap."""
        xy = _f64(xy).reshape(-1, 2)
        clk = None
        if now_us is not None:
            assert canvas.tmap.dtype == np.int64 and canvas.tmap.flags.c_contiguous
            clk = C.byref(Clock(canvas.tmap.ctypes.data_as(C.POINTER(C.c_int64)), int(now_us),
                                int(dry_ms), _p(canvas.R0), _p(canvas.h)))
        lib().orc_tbrush_paint_stroke(self.h, canvas.rows, canvas.cols, _p(canvas.K),
                                      _p(canvas.S), _p(canvas.V), _p(xy), int(xy.shape[0]), clk)


# ----------------------------------------------------- footprint ingest -----
def srgb2rgb(s: float) -> float:
    return lib().orc_srgb2rgb(float(s))


def gray_from_u16(v, maxval: int, srgb: bool) -> np.ndarray:
    """imRead's pixel transform of raw grey samples (uint16 array)."""
    a = np.ascontiguousarray(v, dtype=np.uint16)
    out = np.empty(a.shape)
    lib().orc_gray_from_u16(a.size, a.ctypes.data_as(C.POINTER(C.c_uint16)), int(maxval),
                            1 if srgb else 0, _p(out))
    return out


def lanczos4_coeffs(x: float) -> np.ndarray:
    c = (C.c_float * 8)()
    lib().orc_lanczos4_coeffs(float(x), c)
    return np.array(c[:], dtype=np.float32)


def resize_lanczos4(img, out_rows: int, out_cols: int) -> np.ndarray:
    """ScaledMat(img, out_rows, out_cols): cv::resize INTER_LANCZOS4, FP64."""
    a = _f64(img)
    out = np.empty((out_rows, out_cols))
    lib().orc_resize_lanczos4(_p(a), a.shape[0], a.shape[1], int(out_rows), int(out_cols),
                              _p(out))
    return out


# --------------------------------------------------- GL texture brush -----
class GlCanvas:
    """FP32 canvas planes as the GL renderer holds them (canvas row order):
    K, S (rows, cols, 3), V and the lighting thickness (rows, cols)."""

    def __init__(self, K, S, V, light=None):
        self.K = np.ascontiguousarray(K, np.float32).copy()
        self.S = np.ascontiguousarray(S, np.float32).copy()
        self.V = np.ascontiguousarray(V, np.float32).copy()
        self.light = (np.zeros(self.V.shape, np.float32) if light is None
                      else np.ascontiguousarray(light, np.float32).copy())
        self.rows, self.cols = self.V.shape


def _pf32(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class TextureBrushGpu:
    """TextureBrushGpu (renderer/src/TextureBrushGpu.cxx:17-305)."""

    def __init__(self):
        self.h = lib().orc_tbgl_create()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tbgl_destroy(self.h)
            self.h = None

    def set_radius(self, r: float) -> None:
        lib().orc_tbgl_set_radius(self.h, float(r))

    def dip(self, K, S) -> None:
        lib().orc_tbgl_dip(self.h, _p(_f64(K)), _p(_f64(S)))

    def set_texture(self, img) -> None:
        a = _f64(img)
        lib().orc_tbgl_set_texture(self.h, _p(a), a.shape[0], a.shape[1])

    def set_thickness_scale(self, s: float) -> None:
        lib().orc_tbgl_set_thickness_scale(self.h, float(s))

    def enable_smudge(self, on: bool) -> None:
        lib().orc_tbgl_enable_smudge(self.h, int(bool(on)))

    def smudge(self):
        side = C.c_int(0)
        lib().orc_tbgl_get_smudge(self.h, C.byref(side), None, None)
        n = side.value
        K, S = np.zeros((n, n, 4), np.float32), np.zeros((n, n, 4), np.float32)
        lib().orc_tbgl_get_smudge(self.h, C.byref(side), _pf32(K), _pf32(S))
        return K, S

    def paint_stroke(self, xy, cv: GlCanvas) -> None:
        xy = _f64(xy).reshape(-1, 2)
        lib().orc_tbgl_paint_stroke(self.h, cv.rows, cv.cols, _pf32(cv.K), _pf32(cv.S),
                                    _pf32(cv.V), _pf32(cv.light), _p(xy), int(xy.shape[0]))
