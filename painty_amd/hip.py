"""Python mirror of the reference's renderer interface over the C ABI.

Class and method names follow painty's C++ templates (instantiated with vec3)
so that code and tests read like the reference's own:

  PaintLayer      renderer/PaintLayer.hxx:21-145
  Canvas          renderer/Canvas.hxx:20-198
  Renderer        renderer/Renderer.hxx:14-53
  FootprintBrush  renderer/FootprintBrush.hxx:22-502
  TextureBrush    renderer/TextureBrush.hxx:19-238 (with Smudge.hxx:22-229)
  PaintMixer      mixer/PaintMixer.hxx:22-57 (mixSinglePaint, mixed)
  ComputeReflectance  core/KubelkaMunk.hxx:25-83

The reference hands out mutable host ``Mat&`` buffers; here getters return
host copies (numpy float32, (rows, cols, 3) for K/S/R0, (rows, cols) for V/h)
and ``set*`` methods write them back.  Everything runs through
libpainty_hip.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import AOS, PLANAR, InvalidArgument, Owned, PaintyError, PntStamp, call, load

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError("expected shape %s, got %s" % (tuple(shape), a.shape))
    return a


def _pf(a: np.ndarray):
    return a.ctypes.data_as(_fp) if a is not None else None


class Context(Owned):
    """A device and one HIP stream on it."""

    _kind, _destroy = "ctx", "pnt_ctx_destroy"

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("pnt_ctx_create", device, C.byref(h))
        self._own(h)
        self.device = device

    def synchronize(self) -> None:
        call("pnt_ctx_synchronize", self.h)

    @property
    def stream(self) -> int:
        return load().pnt_ctx_stream(self.h) or 0


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class PaintLayer(Owned):
    _kind, _destroy = "layer", "pnt_layer_destroy"

    def __init__(self, rows: int, cols: int, ctx: Context | None = None, _handle=None,
                 _parent=None):
        self.ctx = ctx or default_context()
        self._parent = _parent  # keeps a canvas alive behind a view
        if _handle is None:
            h = C.c_void_p()
            call("pnt_layer_create", self.ctx.h, rows, cols, C.byref(h))
            self._own(h)
            self._owner = True
        else:
            self.h = _handle
            self._owner = False

    def close(self) -> None:
        if getattr(self, "_owner", False):
            super().close()
        else:  # a view of a canvas's wet layer: the canvas owns it
            self.h = None
            self.h = None

    def getRows(self) -> int:
        r, c = C.c_int(), C.c_int()
        call("pnt_layer_shape", self.h, C.byref(r), C.byref(c))
        return r.value

    def getCols(self) -> int:
        r, c = C.c_int(), C.c_int()
        call("pnt_layer_shape", self.h, C.byref(r), C.byref(c))
        return c.value

    @property
    def shape(self):
        return self.getRows(), self.getCols()

    def clear(self) -> None:
        call("pnt_layer_clear", self.h)

    def _download(self, which: str) -> np.ndarray:
        rows, cols = self.shape
        out = np.empty((rows, cols, 3) if which in "KS" else (rows, cols), np.float32)
        args = [None, None, None]
        args["KSV".index(which)] = _pf(out)
        call("pnt_layer_download", self.h, *args, AOS)
        return out

    def getK_buffer(self) -> np.ndarray:
        return self._download("K")

    def getS_buffer(self) -> np.ndarray:
        return self._download("S")

    def getV_buffer(self) -> np.ndarray:
        """V as the device holds it: FP64 (rows, cols)."""
        rows, cols = self.shape
        out = np.empty((rows, cols), np.float64)
        call("pnt_layer_download_volume", self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def setBuffers(self, K=None, S=None, V=None) -> None:
        """K, S go to FP32 planes; V to the FP64 plane (a float64 V exactly,
        a float32 V widened)."""
        rows, cols = self.shape
        K = None if K is None else _f32(K, (rows, cols, 3))
        S = None if S is None else _f32(S, (rows, cols, 3))
        v64 = V is not None and np.asarray(V).dtype == np.float64
        V32 = None if V is None or v64 else _f32(V, (rows, cols))
        call("pnt_layer_upload", self.h, _pf(K), _pf(S), _pf(V32), AOS)
        if v64:
            Vd = np.ascontiguousarray(V, dtype=np.float64)
            assert Vd.shape == (rows, cols)
            call("pnt_layer_upload_volume", self.h, Vd.ctypes.data_as(C.POINTER(C.c_double)))

    def composeOnto(self, R0: np.ndarray) -> np.ndarray:
        """In place on R0 (float32 (rows, cols, 3)); returns it.  A
        mismatched R0 is replaced by a ones image of the layer's size, as
        PaintLayer.hxx:82-87 does; the replacement is returned."""
        rows, cols = self.shape
        R0 = np.asarray(R0)
        if R0.shape != (rows, cols, 3) or R0.dtype != np.float32 or not R0.flags.c_contiguous:
            if R0.shape == (rows, cols, 3):
                raise ValueError("R0 must be a C-contiguous float32 array")
            R0 = np.empty((rows, cols, 3), np.float32)
            call("pnt_layer_compose_onto_host", self.h, _pf(R0), -1, -1, AOS)
            return R0
        call("pnt_layer_compose_onto_host", self.h, _pf(R0), rows, cols, AOS)
        return R0

    def composeOntoCanvas(self, canvas: "Canvas") -> None:
        call("pnt_layer_compose_onto_canvas", self.h, canvas.h)

    def copyTo(self, other: "PaintLayer") -> None:
        call("pnt_layer_copy_to", self.h, other.h)

    def set(self, i: int, j: int, k, s, v: float) -> None:
        k, s = _f32(k, (3,)), _f32(s, (3,))
        call("pnt_layer_set", self.h, i, j, _pf(k), _pf(s), float(v))


class Canvas(Owned):
    _kind, _destroy = "canvas", "pnt_canvas_destroy"

    def __init__(self, rows: int, cols: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        call("pnt_canvas_create", self.ctx.h, rows, cols, C.byref(h))
        self._own(h)
        self.rows, self.cols = rows, cols

    def clear(self) -> None:
        call("pnt_canvas_clear", self.h)

    def getPaintLayer(self) -> PaintLayer:
        v = C.c_void_p()
        call("pnt_canvas_layer", self.h, C.byref(v))
        return PaintLayer(self.rows, self.cols, self.ctx, _handle=v, _parent=self)

    def getR0(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols, 3), np.float32)
        call("pnt_canvas_download_R0", self.h, _pf(out), AOS)
        return out

    def setR0(self, R0) -> None:
        R0 = _f32(R0, (self.rows, self.cols, 3))
        call("pnt_canvas_upload_R0", self.h, _pf(R0), AOS)

    def get_h(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), np.float32)
        call("pnt_canvas_download_h", self.h, _pf(out))
        return out

    def set_h(self, h) -> None:
        h = _f32(h, (self.rows, self.cols))
        call("pnt_canvas_upload_h", self.h, _pf(h))

    def getReflectanceLayerDry(self) -> np.ndarray:
        return self.getR0()

    def setBackground(self, background) -> None:
        bg = _f32(background, (self.rows, self.cols, 3))
        call("pnt_canvas_set_background", self.h, _pf(bg), AOS)

    def dryCanvas(self) -> None:
        call("pnt_canvas_dry", self.h)

    def dryStep(self, portion: float) -> None:
        call("pnt_canvas_dry_step", self.h, float(portion))

    def setDryTimer(self, period_ms: int = 250, portion: float = 0.01) -> None:
        """The GPU render thread's drying timer (SbrRenderThread.cxx:37-41):
        dryStep(portion) every period_ms of canvas clock time; 0 stops it."""
        call("pnt_canvas_set_dry_timer", self.h, int(period_ms), float(portion))

    def setDryingTime(self, ms: int) -> None:
        call("pnt_canvas_set_drying_time_ms", self.h, int(ms))

    def getDryingTime(self) -> int:
        v = C.c_int64()
        call("pnt_canvas_get_drying_time_ms", self.h, C.byref(v))
        return v.value

    def setClock(self, clock: int) -> None:
        """_lib.CLOCK_OFF / CLOCK_SYSTEM / CLOCK_MANUAL (Canvas::checkDry)."""
        call("pnt_canvas_set_clock", self.h, int(clock))

    def getClock(self) -> int:
        v = C.c_int()
        call("pnt_canvas_get_clock", self.h, C.byref(v))
        return v.value

    def setTime(self, now_us: int) -> None:
        """The time point of the next imprints under CLOCK_MANUAL (us)."""
        call("pnt_canvas_set_time_us", self.h, int(now_us))

    def checkDry(self, x: int, y: int, now_us: int) -> None:
        """Canvas::checkDry of one pixel at time point now_us (needs a clock)."""
        call("pnt_canvas_check_dry", self.h, int(x), int(y), int(now_us))

    def getTimeMap(self) -> np.ndarray:
        t = np.empty((self.rows, self.cols), np.int64)
        call("pnt_canvas_download_time_map", self.h, t.ctypes.data_as(C.POINTER(C.c_int64)))
        return t

    def setTimeMap(self, tmap) -> None:
        t = np.ascontiguousarray(tmap, dtype=np.int64)
        assert t.shape == (self.rows, self.cols)
        call("pnt_canvas_upload_time_map", self.h, t.ctypes.data_as(C.POINTER(C.c_int64)))

    def device_planes(self):
        """(wet layer as a DevWet -- K/S FP32 planes, V FP64 --, R0 planes[3],
        h) as device addresses."""
        L, R, h = _lib.DevWet(), _lib.DevRgb(), C.c_void_p()
        call("pnt_canvas_device_planes", self.h, C.byref(L), C.byref(R), C.byref(h))
        return L, list(R.p), h.value


class Renderer:
    @staticmethod
    def compose(*args) -> np.ndarray:
        """compose(canvas) or compose(paintLayer, R0) -> R (float32 (rows, cols, 3))."""
        if len(args) == 1:
            cv = args[0]
            out = np.empty((cv.rows, cv.cols, 3), np.float32)
            call("pnt_compose", cv.h, _pf(out), AOS)
            return out
        layer, R0 = args
        rows, cols = layer.shape
        R0 = _f32(R0, (rows, cols, 3))
        out = np.empty_like(R0)
        call("pnt_compose_layer_host", layer.h, _pf(R0), _pf(out), AOS)
        return out

    @staticmethod
    def render(canvas) -> np.ndarray:
        """Renderer::render(canvas): the lit composite, float32 (rows, cols, 3)."""
        out = np.empty((canvas.rows, canvas.cols, 3), np.float32)
        call("pnt_render", canvas.h, _pf(out), AOS)
        return out


def compute_scattering_absorption(Rb, Rw, ctx: Context | None = None):
    """ComputeScatteringAndAbsorption (Curtis inversion) of (..., 3) colours
    over black and over white -> (K, S), float64, on the device."""
    ctx = ctx or default_context()
    b = np.ascontiguousarray(Rb, dtype=np.float64)
    w = np.ascontiguousarray(Rw, dtype=np.float64)
    assert b.shape == w.shape and b.shape[-1] == 3
    K, S = np.empty_like(b), np.empty_like(b)
    dp = C.POINTER(C.c_double)
    call("pnt_compute_scattering_absorption", ctx.h, b.size // 3, b.ctypes.data_as(dp),
         w.ctypes.data_as(dp), K.ctypes.data_as(dp), S.ctypes.data_as(dp))
    return K, S


def palette_from_json(text: str):
    """LoadPalette: JSON text -> (K, S), each (n, 3) float64."""
    raw = text.encode()
    n = C.c_int64()
    cap = max(1, raw.count(b"{"))  # an upper bound on the entries
    K, S = np.empty((cap, 3)), np.empty((cap, 3))
    dp = C.POINTER(C.c_double)
    call("pnt_palette_from_json", raw, len(raw), K.ctypes.data_as(dp), S.ctypes.data_as(dp), cap,
         C.byref(n))
    return K[: n.value].copy(), S[: n.value].copy()


def palette_to_json(K, S) -> str:
    """SavePalette: (n, 3) K and S -> compact JSON text."""
    K = np.ascontiguousarray(K, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    n = len(K)
    ln = C.c_int64()
    cap = 128 + 320 * n
    buf = C.create_string_buffer(cap)
    dp = C.POINTER(C.c_double)
    call("pnt_palette_to_json", K.ctypes.data_as(dp), S.ctypes.data_as(dp), n, buf, cap, C.byref(ln))
    return buf.value.decode()


def thinning_medium():
    K, S = (C.c_double * 3)(), (C.c_double * 3)()
    load().pnt_thinning_medium(K, S)
    return np.array(K), np.array(S)


def encode(rgb, to_srgb: bool = True, bits: int = 8, ctx: Context | None = None) -> np.ndarray:
    """imSave's pixel transform: (..., 3) linear float RGB -> uint8 / uint16
    (optionally through sRGB), interleaved RGB of the same shape."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(rgb, dtype=np.float32)
    assert a.shape[-1] == 3
    n = a.size // 3
    out = np.empty(a.shape, np.uint16 if bits == 16 else np.uint8)
    call("pnt_encode", ctx.h, n, _pf(a), AOS, 1 if to_srgb else 0, int(bits),
         out.ctypes.data_as(C.c_void_p))
    return out


class FootprintBrush(Owned):
    """Imprint brush with an explicit footprint (width = 2*ceil(radius)+1)."""

    _kind, _destroy = "brush", "pnt_brush_destroy"

    def __init__(self, radius: float, footprint, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        fp = _f32(footprint)
        h = C.c_void_p()
        call("pnt_brush_create", self.ctx.h, float(radius), _pf(fp), int(fp.shape[0]),
             C.byref(h))
        self._own(h)

    def setRadius(self, radius: float, footprint) -> None:
        fp = _f32(footprint)
        call("pnt_brush_set_radius", self.h, float(radius), _pf(fp), int(fp.shape[0]))

    def getSizeMap(self) -> int:
        v = C.c_int()
        call("pnt_brush_size_map", self.h, C.byref(v))
        return v.value

    @property
    def sizeMap(self) -> int:
        v = C.c_int()
        call("pnt_brush_size_map", self.h, C.byref(v))
        return v.value

    def getFootprint(self) -> np.ndarray:
        m = self.sizeMap
        out = np.empty((m, m), np.float32)
        call("pnt_brush_get_footprint", self.h, _pf(out))
        return out

    def getPickupMap(self):
        m = self.sizeMap
        K, S, V = (np.empty((m, m, 3), np.float32), np.empty((m, m, 3), np.float32),
                   np.empty((m, m), np.float32))
        call("pnt_brush_get_pickup_map", self.h, _pf(K), _pf(S), _pf(V), AOS)
        return K, S, V

    def setPickupRate(self, r: float) -> None:
        call("pnt_brush_set_pickup_rate", self.h, float(r))

    def getPickupRate(self) -> float:
        v = C.c_double()
        call("pnt_brush_get_pickup_rate", self.h, C.byref(v))
        return v.value

    def setDepositionRate(self, r: float) -> None:
        call("pnt_brush_set_deposition_rate", self.h, float(r))

    def getDepositionRate(self) -> float:
        v = C.c_double()
        call("pnt_brush_get_deposition_rate", self.h, C.byref(v))
        return v.value

    def setUseSnapshotBuffer(self, use: bool) -> None:
        call("pnt_brush_set_use_snapshot", self.h, int(bool(use)))

    def getUseSnapshotBuffer(self) -> bool:
        v = C.c_int()
        call("pnt_brush_get_use_snapshot", self.h, C.byref(v))
        return bool(v.value)

    def dip(self, paint) -> None:
        K, S = paint
        K, S = _f32(K, (3,)), _f32(S, (3,))
        call("pnt_brush_dip", self.h, _pf(K), _pf(S))

    def clean(self) -> None:
        call("pnt_brush_clean", self.h)

    def updateSnapshot(self, canvas: Canvas) -> None:
        call("pnt_brush_update_snapshot", self.h, canvas.h)

    def imprint(self, center, theta: float, canvas: Canvas) -> None:
        call("pnt_imprint", self.h, canvas.h, float(center[0]), float(center[1]), float(theta))

    def imprintStamps(self, stamps, canvas: Canvas, mode: int = _lib.EXEC_PLAIN) -> None:
        arr = stamps_to_c(stamps)
        call("pnt_imprint_stamps", self.h, canvas.h, arr, len(arr), int(mode))

    def imprintStrokes(self, strokes, canvas: Canvas, mode: int = _lib.EXEC_FUSED) -> None:
        """pnt_imprint_strokes over [(radius, footprint, stamps), ...]: per
        stroke setRadius(radius, footprint), then its stamps."""
        n = len(strokes)
        arr = (_lib.PntStroke * n)()
        keep = []
        for k, (radius, fp, stamps) in enumerate(strokes):
            f = _f32(fp)
            st = stamps_to_c(stamps)
            keep.append((f, st))
            arr[k].radius = float(radius)
            arr[k].footprint = _pf(f)
            arr[k].width = int(f.shape[0])
            arr[k].stamps = C.cast(st, C.POINTER(PntStamp)) if len(st) else None
            arr[k].n_stamps = len(st)
        call("pnt_imprint_strokes", self.h, canvas.h, arr, n, int(mode))

    def paintStroke(self, path, canvas: Canvas) -> None:
        p = np.ascontiguousarray(path, dtype=np.float64).reshape(-1, 2)
        call("pnt_paint_stroke", self.h, canvas.h, p.ctypes.data_as(_dp), len(p))


def read_image_gray(data: bytes, srgb: bool = True) -> np.ndarray:
    """io::imRead(buffer, Mat<double>&, convertFrom_sRGB) of a PNG file's
    bytes (io/src/ImageIO.cxx:115-150): grey samples / maxval, sRGB ->
    linear when srgb.  Host work; no device needed."""
    data = bytes(data)
    r, c = C.c_int(), C.c_int()
    call("pnt_image_read_gray", data, len(data), int(bool(srgb)), C.byref(r), C.byref(c), None, 0)
    out = np.empty((r.value, c.value))
    call("pnt_image_read_gray", data, len(data), int(bool(srgb)), C.byref(r), C.byref(c),
         out.ctypes.data_as(_dp), out.size)
    return out


def resize_lanczos4(img, rows: int, cols: int, ctx: Context | None = None) -> np.ndarray:
    """ScaledMat(img, rows, cols) (image/Mat.hxx:141-148): cv::resize with
    INTER_LANCZOS4, FP64, on the device."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(img, np.float64)
    out = np.empty((int(rows), int(cols)))
    call("pnt_resize_lanczos4", ctx.h, a.ctypes.data_as(_dp), a.shape[0], a.shape[1], int(rows),
         int(cols), out.ctypes.data_as(_dp))
    return out


def footprint_from_image(img, radius: float, ctx: Context | None = None) -> np.ndarray:
    """The footprint FootprintBrush::setRadius builds (FootprintBrush.hxx:
    48-55) from the full-size image (read_image_gray of footprint.png):
    ScaledMat to width x width, width = 2*ceil(radius)+1, as float32 for
    FootprintBrush(radius, footprint)."""
    import math
    w = int(2.0 * math.ceil(radius) + 1.0)
    return resize_lanczos4(img, w, w, ctx).astype(np.float32)


class TextureBrush(Owned):
    """Brush stroke texture warped along a stroke, with smudging
    (renderer/TextureBrush.hxx:19-238).  The reference loads the sample's
    thickness map from ``sampleDir/thickness_map.png``
    (renderer/src/BrushStrokeSample.cxx:164); here the caller passes the
    map (rows x cols, float64) to ``setThicknessMap``."""

    _kind, _destroy = "tbrush", "pnt_tbrush_destroy"

    def __init__(self, thickness_map=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        call("pnt_tbrush_create", self.ctx.h, C.byref(h))
        self._own(h)
        if thickness_map is not None:
            self.setThicknessMap(thickness_map)

    def setRadius(self, radius: float) -> None:
        call("pnt_tbrush_set_radius", self.h, float(radius))

    def dip(self, paint) -> None:
        """paint = (K, S), 3 values each (TextureBrush.hxx:48-50)."""
        K = np.ascontiguousarray(paint[0], np.float64).reshape(3)
        S = np.ascontiguousarray(paint[1], np.float64).reshape(3)
        call("pnt_tbrush_dip", self.h, K.ctypes.data_as(_dp), S.ctypes.data_as(_dp))

    def setThicknessMap(self, m) -> None:
        m = np.ascontiguousarray(m, np.float64)
        if m.ndim != 2:
            raise ValueError("thickness map must be 2-d")
        call("pnt_tbrush_set_thickness_map", self.h, m.ctypes.data_as(_dp), int(m.shape[0]),
             int(m.shape[1]))

    def setThicknessScale(self, scale: float) -> None:
        call("pnt_tbrush_set_thickness_scale", self.h, float(scale))

    def enableSmudge(self, enable: bool) -> None:
        call("pnt_tbrush_set_use_smudge", self.h, int(bool(enable)))

    def getSmudgeMap(self):
        """The smudge's pickup map: K, S (side, side, 3) and V (side, side)."""
        side = C.c_int()
        call("pnt_tbrush_get_smudge_map", self.h, C.byref(side), None, None, None)
        n = side.value
        K, S, V = np.zeros((n, n, 3)), np.zeros((n, n, 3)), np.zeros((n, n))
        call("pnt_tbrush_get_smudge_map", self.h, C.byref(side), K.ctypes.data_as(_dp),
             S.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
        return K, S, V

    def paintStroke(self, vertices, canvas: Canvas) -> None:
        """TextureBrush.hxx:52-205; vertices: (n, 2) x, y."""
        xy = np.ascontiguousarray(vertices, np.float64).reshape(-1, 2)
        call("pnt_tbrush_paint_stroke", self.h, canvas.h, xy.ctypes.data_as(_dp),
             int(xy.shape[0]))


class TextureBrushGpu(Owned):
    """The reference's GL texture brush (renderer/src/TextureBrushGpu.cxx:
    17-305) on the device; the brush texture is set by the caller
    (setTexture) instead of drawn from the texture dictionary."""

    _kind, _destroy = "tbrush", "pnt_tbgl_destroy"

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        call("pnt_tbgl_create", self.ctx.h, C.byref(h))
        self._own(h)

    def setRadius(self, radius: float) -> None:
        call("pnt_tbgl_set_radius", self.h, float(radius))

    def dip(self, paint) -> None:
        K = np.ascontiguousarray(paint[0], np.float64).reshape(3)
        S = np.ascontiguousarray(paint[1], np.float64).reshape(3)
        call("pnt_tbgl_dip", self.h, K.ctypes.data_as(_dp), S.ctypes.data_as(_dp))

    def setTexture(self, img) -> None:
        a = np.ascontiguousarray(img, np.float64)
        call("pnt_tbgl_set_texture", self.h, a.ctypes.data_as(_dp), a.shape[0], a.shape[1])

    def setThicknessScale(self, scale: float) -> None:
        call("pnt_tbgl_set_thickness_scale", self.h, float(scale))

    def enableSmudge(self, enable: bool) -> None:
        call("pnt_tbgl_enable_smudge", self.h, int(bool(enable)))

    def getSmudge(self):
        side = C.c_int()
        call("pnt_tbgl_get_smudge", self.h, C.byref(side), None, None)
        n = side.value
        K, S = np.zeros((n, n, 4), np.float32), np.zeros((n, n, 4), np.float32)
        call("pnt_tbgl_get_smudge", self.h, C.byref(side), _pf(K), _pf(S))
        return K, S

    def paintStroke(self, vertices, canvas: Canvas) -> None:
        xy = np.ascontiguousarray(vertices, np.float64).reshape(-1, 2)
        call("pnt_tbgl_paint_stroke", self.h, canvas.h, xy.ctypes.data_as(_dp), int(xy.shape[0]))


def canvas_lighting(canvas: Canvas) -> np.ndarray:
    """The GL renderer's lighting thickness plane (S.a) of a canvas."""
    out = np.empty((canvas.rows, canvas.cols), np.float32)
    call("pnt_canvas_download_lighting", canvas.h, _pf(out))
    return out


def stamps_to_c(stamps):
    """A workload.STAMP_DTYPE array (or a C array already) -> PntStamp array."""
    if isinstance(stamps, C.Array):
        return stamps
    n = len(stamps)
    arr = (PntStamp * n)()
    if n:
        buf = np.frombuffer(arr, dtype=np.dtype([("cx", "f8"), ("cy", "f8"), ("theta", "f8"),
                                                   ("K", "f4", 3), ("S", "f4", 3),
                                                   ("flags", "u4"), ("reserved", "i4")]))
        for f in ("cx", "cy", "theta", "K", "S", "flags"):
            buf[f] = stamps[f]
    return arr


def _batch_jobs(jobs):
    n = len(jobs)
    arr = (_lib.BatchJob * n)()
    keep = []
    for k, (cv, br, stamps, out) in enumerate(jobs):
        st = stamps_to_c(stamps)
        keep.append(st)
        arr[k].canvas = cv.h.value if isinstance(cv.h, C.c_void_p) else cv.h
        arr[k].brush = br.h.value if isinstance(br.h, C.c_void_p) else br.h
        arr[k].stamps = C.cast(st, C.POINTER(PntStamp)) if len(st) else None
        arr[k].n_stamps = len(st)
        if out is not None:
            _, R, _ = out.device_planes()
            for c in range(3):
                arr[k].out[c] = R[c]
    return arr, keep


def batch_run(jobs, mode: int = _lib.BATCH_ONE_LAUNCH, ctx: Context | None = None) -> None:
    """pnt_batch_run over [(canvas, brush, stamps, out_canvas_or_None), ...]:
    the composite of job k lands in out_canvas's R0 planes."""
    ctx = ctx or default_context()
    arr, _keep = _batch_jobs(jobs)
    call("pnt_batch_run", ctx.h, arr, len(jobs), int(mode))


def batch_submit(jobs, ctx: Context | None = None) -> None:
    """pnt_batch_submit: the one-launch batch, enqueued; batch_wait joins it."""
    ctx = ctx or default_context()
    arr, _keep = _batch_jobs(jobs)
    call("pnt_batch_submit", ctx.h, arr, len(jobs))


def batch_wait(ctx: Context | None = None) -> None:
    call("pnt_batch_wait", (ctx or default_context()).h)


def batch_set_config(tile: int, max_group: int, waves: int = 0, chunk: int = 1,
                     ctx: Context | None = None) -> None:
    """pnt_batch_set_config: waves=0 -- one persistent launch with tile x tile
    window tiles and at most max_group workgroups per job; waves=1 -- the
    batch as a sequence of wave launches, `chunk` strokes of a canvas in
    flight at once."""
    ctx = ctx or default_context()
    c = _lib.BatchConfig(tile, max_group, waves, chunk)
    call("pnt_batch_set_config", ctx.h, C.byref(c))


def batch_get_config(ctx: Context | None = None):
    """(tile, max_group, waves, chunk) of the context's batches."""
    ctx = ctx or default_context()
    c = _lib.BatchConfig()
    call("pnt_batch_get_config", ctx.h, C.byref(c))
    return c.tile, c.max_group, c.waves, c.chunk


def batch_tune(jobs, ctx: Context | None = None):
    """pnt_batch_tune: returns ((tile, max_group, waves, chunk), ms) of the
    fastest configuration, which the context keeps; the jobs' state is
    unchanged."""
    ctx = ctx or default_context()
    arr, _keep = _batch_jobs(jobs)
    c, ms = _lib.BatchConfig(), C.c_double()
    call("pnt_batch_tune", ctx.h, arr, len(jobs), C.byref(c), C.byref(ms))
    return (c.tile, c.max_group, c.waves, c.chunk), ms.value


def batch_profile(jobs, ctx: Context | None = None):
    """pnt_batch_profile: runs the batch once as waves, timed wave by wave.
    Returns (job_first, job_last, wave_blocks, wave_end_ms) as numpy arrays:
    each job's first and last wave, and each wave's workgroups and end time
    (ms from the batch's start)."""
    ctx = ctx or default_context()
    arr, _keep = _batch_jobs(jobs)
    n = len(jobs)
    # every wave holds at least one phase of some job: a stamp, a unit's
    # map clear, a dry, the pickup copy or the composite
    cap = 3 * sum(int(arr[k].n_stamps) for k in range(n)) + 8 * n + 64
    jf, jl = np.zeros(n, np.int32), np.zeros(n, np.int32)
    wb, we = np.zeros(cap, np.int32), np.zeros(cap, np.float32)
    nw = C.c_int()
    ip = C.POINTER(C.c_int)
    call("pnt_batch_profile", ctx.h, arr, n, cap, jf.ctypes.data_as(ip), jl.ctypes.data_as(ip),
         wb.ctypes.data_as(ip), we.ctypes.data_as(C.POINTER(C.c_float)), C.byref(nw))
    k = min(nw.value, cap)
    return jf, jl, wb[:k], we[:k]


class BatchGraph(Owned):
    """pnt_batch_capture: the batch captured once as a HIP graph;
    ``launch()`` replays it on the jobs' current state."""

    _kind, _destroy = "graph", "pnt_batch_graph_destroy"

    def __init__(self, jobs, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        arr, _keep = _batch_jobs(jobs)
        h = C.c_void_p()
        call("pnt_batch_capture", self.ctx.h, arr, len(jobs), C.byref(h))
        self._own(h)

    def launch(self) -> None:
        call("pnt_batch_graph_launch", self.h)


class PaintMixer:
    """mixSinglePaint / mixed over a base palette [(K[3], S[3]), ...]."""

    def __init__(self, palette):
        self.setUnderlyingPalette(palette)

    def setUnderlyingPalette(self, palette) -> None:
        self._K = np.ascontiguousarray([p[0] for p in palette], dtype=np.float64).reshape(-1, 3)
        self._S = np.ascontiguousarray([p[1] for p in palette], dtype=np.float64).reshape(-1, 3)

    def getUnderlyingPalette(self):
        return [(self._K[i].copy(), self._S[i].copy()) for i in range(len(self._K))]

    def mixSinglePaint(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        K, S = np.empty(3), np.empty(3)
        call("pnt_mix_single", len(self._K), self._K.ctypes.data_as(_dp),
             self._S.ctypes.data_as(_dp), len(w), w.ctypes.data_as(_dp),
             K.ctypes.data_as(_dp), S.ctypes.data_as(_dp))
        return K, S

    @staticmethod
    def mixed(paint, volume, other, other_volume):
        K1, S1 = (np.ascontiguousarray(x, dtype=np.float64) for x in paint)
        K2, S2 = (np.ascontiguousarray(x, dtype=np.float64) for x in other)
        K, S = np.empty(3), np.empty(3)
        call("pnt_mixed", K1.ctypes.data_as(_dp), S1.ctypes.data_as(_dp), float(volume),
             K2.ctypes.data_as(_dp), S2.ctypes.data_as(_dp), float(other_volume),
             K.ctypes.data_as(_dp), S.ctypes.data_as(_dp))
        return K, S


def ComputeReflectance(K, S, R0, d, ctx: Context | None = None) -> np.ndarray:
    """core/KubelkaMunk.hxx:25-83 over arrays: K, S, R0 (..., 3), d (...)."""
    ctx = ctx or default_context()
    K, S, R0 = _f32(K), _f32(S), _f32(R0)
    d = _f32(d)
    n = d.size
    if K.size != 3 * n or S.size != 3 * n or R0.size != 3 * n:
        raise ValueError("K, S, R0 must hold 3 values per thickness")
    out = np.empty_like(R0)
    call("pnt_compute_reflectance", ctx.h, n, _pf(K), _pf(S), _pf(R0), _pf(d), _pf(out))
    return out


__all__ = ["Context", "PaintLayer", "Canvas", "Renderer", "FootprintBrush", "PaintMixer",
           "ComputeReflectance", "PaintyError", "InvalidArgument", "default_context",
           "stamps_to_c", "PLANAR", "AOS"]
