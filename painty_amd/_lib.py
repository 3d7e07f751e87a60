"""ctypes binding of include/painty_hip.h (libpainty_hip.so, built in-tree).

There is no fallback: importing a compute entry point without the library
raises, and every call that returns a non-zero status raises PaintyError with
the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpainty_hip.so")

OK, EINVAL, EHIP, ENOMEM, ESTATE, EKERNEL = 0, -1, -2, -3, -4, -5
AOS, PLANAR = 0, 1
STAMP_DIP, STAMP_DRY_AFTER = 1, 2
EXEC_PLAIN, EXEC_PERSISTENT, EXEC_FUSED, EXEC_CHAINED = 0, 1, 2, 3
CLOCK_OFF, CLOCK_SYSTEM, CLOCK_MANUAL = 0, 1, 2
BATCH_ONE_LAUNCH, BATCH_PLAIN = 0, 1


class PaintyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("painty_hip error %d: %s" % (code, msg))
        self.code = code


class InvalidArgument(PaintyError, ValueError):
    """Where the reference throws std::invalid_argument."""


class PntStamp(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("theta", C.c_double),
                ("K", C.c_float * 3), ("S", C.c_float * 3), ("flags", C.c_uint32),
                ("reserved", C.c_int32)]


class PntStroke(C.Structure):
    _fields_ = [("radius", C.c_double), ("footprint", C.POINTER(C.c_float)), ("width", C.c_int),
                ("stamps", C.POINTER(PntStamp)), ("n_stamps", C.c_int64)]


class BatchJob(C.Structure):
    _fields_ = [("canvas", C.c_void_p), ("brush", C.c_void_p), ("stamps", C.POINTER(PntStamp)),
                ("n_stamps", C.c_int64), ("out", C.c_void_p * 3)]


class BatchConfig(C.Structure):
    _fields_ = [("tile", C.c_int), ("max_group", C.c_int), ("waves", C.c_int), ("chunk", C.c_int)]


class PntRequest(C.Structure):
    _fields_ = [("canvas_id", C.c_int64), ("rows", C.c_int), ("cols", C.c_int),
                ("background", C.c_float * 3), ("radius", C.c_double),
                ("footprint", C.POINTER(C.c_float)), ("width", C.c_int),
                ("pickup_rate", C.c_double), ("deposition_rate", C.c_double),
                ("use_snapshot", C.c_int), ("stamps", C.POINTER(PntStamp)),
                ("n_stamps", C.c_int64), ("out", C.POINTER(C.c_float))]


class DevLayer(C.Structure):
    _fields_ = [("p", C.c_void_p * 7)]


class DevRgb(C.Structure):
    _fields_ = [("p", C.c_void_p * 3)]


class DevWet(C.Structure):
    """pnt_dev_wet: a canvas' wet layer, K/S FP32 planes and the FP64 V plane."""
    _fields_ = [("ks", C.c_void_p * 6), ("v", C.c_void_p)]


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "pnt_last_error": [],
    "pnt_version": [],
    "pnt_ctx_create": [C.c_int, C.POINTER(_vp)],
    "pnt_ctx_destroy": [_vp],
    "pnt_ctx_synchronize": [_vp],
    "pnt_ctx_stream": [_vp],
    "pnt_layer_create": [_vp, C.c_int, C.c_int, C.POINTER(_vp)],
    "pnt_layer_destroy": [_vp],
    "pnt_layer_shape": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pnt_layer_clear": [_vp],
    "pnt_layer_upload": [_vp, _fp, _fp, _fp, C.c_int],
    "pnt_layer_download": [_vp, _fp, _fp, _fp, C.c_int],
    "pnt_layer_upload_volume": [_vp, _dp],
    "pnt_layer_download_volume": [_vp, _dp],
    "pnt_layer_copy_to": [_vp, _vp],
    "pnt_layer_set": [_vp, C.c_int, C.c_int, _fp, _fp, C.c_float],
    "pnt_layer_compose_onto_host": [_vp, _fp, C.c_int, C.c_int, C.c_int],
    "pnt_layer_compose_onto_canvas": [_vp, _vp],
    "pnt_canvas_create": [_vp, C.c_int, C.c_int, C.POINTER(_vp)],
    "pnt_canvas_destroy": [_vp],
    "pnt_canvas_shape": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pnt_canvas_layer": [_vp, C.POINTER(_vp)],
    "pnt_canvas_clear": [_vp],
    "pnt_canvas_set_background": [_vp, _fp, C.c_int],
    "pnt_canvas_upload_R0": [_vp, _fp, C.c_int],
    "pnt_canvas_download_R0": [_vp, _fp, C.c_int],
    "pnt_canvas_upload_h": [_vp, _fp],
    "pnt_canvas_download_h": [_vp, _fp],
    "pnt_canvas_dry": [_vp],
    "pnt_canvas_dry_step": [_vp, C.c_float],
    "pnt_canvas_set_drying_time_ms": [_vp, C.c_int64],
    "pnt_canvas_set_clock": [_vp, C.c_int],
    "pnt_canvas_get_clock": [_vp, C.POINTER(C.c_int)],
    "pnt_canvas_set_time_us": [_vp, C.c_int64],
    "pnt_canvas_check_dry": [_vp, C.c_int, C.c_int, C.c_int64],
    "pnt_canvas_upload_time_map": [_vp, C.POINTER(C.c_int64)],
    "pnt_canvas_download_time_map": [_vp, C.POINTER(C.c_int64)],
    "pnt_canvas_get_drying_time_ms": [_vp, C.POINTER(C.c_int64)],
    "pnt_compose": [_vp, _fp, C.c_int],
    "pnt_compose_layer_host": [_vp, _fp, _fp, C.c_int],
    "pnt_brush_create": [_vp, C.c_double, _fp, C.c_int, C.POINTER(_vp)],
    "pnt_brush_destroy": [_vp],
    "pnt_brush_set_radius": [_vp, C.c_double, _fp, C.c_int],
    "pnt_brush_size_map": [_vp, C.POINTER(C.c_int)],
    "pnt_brush_get_footprint": [_vp, _fp],
    "pnt_brush_get_pickup_map": [_vp, _fp, _fp, _fp, C.c_int],
    "pnt_brush_set_pickup_rate": [_vp, C.c_double],
    "pnt_brush_get_pickup_rate": [_vp, _dp],
    "pnt_brush_set_deposition_rate": [_vp, C.c_double],
    "pnt_brush_get_deposition_rate": [_vp, _dp],
    "pnt_brush_set_use_snapshot": [_vp, C.c_int],
    "pnt_brush_get_use_snapshot": [_vp, C.POINTER(C.c_int)],
    "pnt_brush_dip": [_vp, _fp, _fp],
    "pnt_brush_clean": [_vp],
    "pnt_brush_update_snapshot": [_vp, _vp],
    "pnt_imprint": [_vp, _vp, C.c_double, C.c_double, C.c_double],
    "pnt_imprint_stamps": [_vp, _vp, C.POINTER(PntStamp), C.c_int64, C.c_int],
    "pnt_imprint_stamps_noalloc": [_vp, _vp, C.POINTER(PntStamp), C.c_int64],
    "pnt_imprint_strokes": [_vp, _vp, C.POINTER(PntStroke), C.c_int64, C.c_int],
    "pnt_paint_stroke": [_vp, _vp, _dp, C.c_int64],
    "pnt_batch_run": [_vp, C.POINTER(BatchJob), C.c_int, C.c_int],
    "pnt_batch_submit": [_vp, C.POINTER(BatchJob), C.c_int],
    "pnt_batch_wait": [_vp],
    "pnt_batch_set_config": [_vp, C.POINTER(BatchConfig)],
    "pnt_batch_get_config": [_vp, C.POINTER(BatchConfig)],
    "pnt_batch_tune": [_vp, C.POINTER(BatchJob), C.c_int, C.POINTER(BatchConfig), _dp],
    "pnt_batch_profile": [_vp, C.POINTER(BatchJob), C.c_int, C.c_int, C.POINTER(C.c_int),
                          C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float),
                          C.POINTER(C.c_int)],
    "pnt_batch_capture": [_vp, C.POINTER(BatchJob), C.c_int, C.POINTER(_vp)],
    "pnt_batch_graph_launch": [_vp],
    "pnt_batch_graph_destroy": [_vp],
    "pnt_brush_last_launches": [_vp, C.POINTER(C.c_int64)],
    "pnt_brush_set_kernel_timing": [_vp, C.c_int],
    "pnt_brush_kernel_time": [_vp, _dp, C.POINTER(C.c_int64)],
    "pnt_canvas_set_dry_timer": [_vp, C.c_int64, C.c_float],
    "pnt_brush_stamp_traffic": [_vp, C.c_int, C.c_int, C.POINTER(PntStamp), C.c_int64,
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "pnt_mix_single": [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp],
    "pnt_compute_scattering_absorption": [_vp, C.c_int64, _dp, _dp, _dp, _dp],
    "pnt_palette_from_json": [C.c_char_p, C.c_int64, _dp, _dp, C.c_int64,
                              C.POINTER(C.c_int64)],
    "pnt_palette_to_json": [_dp, _dp, C.c_int64, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)],
    "pnt_thinning_medium": [_dp, _dp],
    "pnt_image_read_gray": [C.c_char_p, C.c_int64, C.c_int, C.POINTER(C.c_int),
                            C.POINTER(C.c_int), _dp, C.c_int64],
    "pnt_resize_lanczos4": [_vp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
    "pnt_tbgl_create": [_vp, C.POINTER(_vp)],
    "pnt_tbgl_destroy": [_vp],
    "pnt_tbgl_set_radius": [_vp, C.c_double],
    "pnt_tbgl_dip": [_vp, _dp, _dp],
    "pnt_tbgl_set_texture": [_vp, _dp, C.c_int, C.c_int],
    "pnt_tbgl_set_thickness_scale": [_vp, C.c_double],
    "pnt_tbgl_enable_smudge": [_vp, C.c_int],
    "pnt_tbgl_get_smudge": [_vp, C.POINTER(C.c_int), _fp, _fp],
    "pnt_tbgl_paint_stroke": [_vp, _vp, _dp, C.c_int64],
    "pnt_canvas_download_lighting": [_vp, _fp],
    "pnt_tbrush_create": [_vp, C.POINTER(_vp)],
    "pnt_tbrush_destroy": [_vp],
    "pnt_tbrush_set_radius": [_vp, C.c_double],
    "pnt_tbrush_dip": [_vp, _dp, _dp],
    "pnt_tbrush_set_thickness_map": [_vp, _dp, C.c_int, C.c_int],
    "pnt_tbrush_set_thickness_scale": [_vp, C.c_double],
    "pnt_tbrush_set_use_smudge": [_vp, C.c_int],
    "pnt_tbrush_get_smudge_map": [_vp, C.POINTER(C.c_int), _dp, _dp, _dp],
    "pnt_tbrush_paint_stroke": [_vp, _vp, _dp, C.c_int64],
    "pnt_mixed": [_dp, _dp, C.c_double, _dp, _dp, C.c_double, _dp, _dp],
    "pnt_compute_reflectance": [_vp, C.c_int64, _fp, _fp, _fp, _fp, _fp],
    "pnt_dev_compose": [_vp, C.c_int64, DevLayer, DevRgb, DevRgb, _vp],
    "pnt_render": [_vp, _fp, C.c_int],
    "pnt_dev_render": [_vp, DevRgb],
    "pnt_encode": [_vp, C.c_int64, _fp, C.c_int, C.c_int, C.c_int, _vp],
    "pnt_dev_encode": [_vp, C.c_int64, DevRgb, C.c_int, C.c_int, _vp, _vp],
    "pnt_dev_compose2": [_vp, C.c_int64, DevLayer, DevLayer, DevRgb, DevRgb, _vp],
    "pnt_dev_compose_wet": [_vp, C.c_int64, DevWet, DevRgb, DevRgb, _vp],
    "pnt_dev_compose2_wet": [_vp, C.c_int64, DevWet, DevWet, DevRgb, DevRgb, _vp],
    "pnt_canvas_device_planes": [_vp, C.POINTER(DevWet), C.POINTER(DevRgb),
                                 C.POINTER(_vp)],
    "pnt_grad_create": [_vp, C.c_int, C.c_int, _vp, C.POINTER(PntStamp), C.c_int64,
                        C.POINTER(C.c_int32), C.c_int, _dp, _dp, C.c_int, C.c_int,
                        C.POINTER(_vp)],
    "pnt_grad_destroy": [_vp],
    "pnt_server_create": [C.c_int, C.c_int, C.c_int64, C.POINTER(_vp)],
    "pnt_server_submit": [_vp, C.POINTER(PntRequest), C.POINTER(C.c_int64)],
    "pnt_server_wait": [_vp, C.c_int64, C.POINTER(C.c_double)],
    "pnt_server_canvas": [_vp, C.c_int64, C.POINTER(_vp), C.POINTER(DevRgb)],
    "pnt_server_release": [_vp, C.c_int64],
    "pnt_server_stats": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                         C.POINTER(C.c_int64)],
    "pnt_server_destroy": [_vp],
    "pnt_grad_forward": [_vp, DevLayer, DevRgb, _dp, DevRgb],
    "pnt_grad_backward": [_vp, DevRgb, DevLayer, DevRgb, _dp],
}
_RESTYPE = {"pnt_last_error": C.c_char_p, "pnt_ctx_stream": _vp}

_lib = None


def load() -> C.CDLL:
    """Load libpainty_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("painty_amd: %s is missing -- build it with "
                              "`python -m painty_amd.build` (no CPU fallback exists)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = load().pnt_last_error().decode(errors="replace")
        if status == EINVAL:
            raise InvalidArgument(status, msg)
        raise PaintyError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


# ------------------------------------------------------------- lifetimes --
# Objects that own a library handle release it exactly once: through their
# close(), through the garbage collector, or -- for whatever is still alive
# when the interpreter exits -- through shutdown(), which runs first among
# the exit handlers and releases in dependency order: captured batches,
# servers and gradient runners (they hold brushes and canvases), texture
# brushes, brushes, canvases, layers, and the contexts last.  Nothing is
# left for the finalizers of a half-torn-down interpreter.
import atexit  # noqa: E402
import threading  # noqa: E402
import weakref  # noqa: E402

_ORDER = ("graph", "server", "grad", "tbrush", "brush", "canvas", "layer", "ctx")
_live = {k: weakref.WeakSet() for k in _ORDER}
_live_lock = threading.Lock()
_shut = False


class Owned:
    """Base of the Python objects that own a handle: subclasses set `_kind`
    (one of _ORDER) and `_destroy` (the pnt_*_destroy name) and call
    `_own(handle)` once the handle exists."""

    _kind = ""
    _destroy = ""
    h = None

    def _own(self, h) -> None:
        self.h = h
        with _live_lock:
            _live[self._kind].add(self)

    def close(self) -> None:
        h, self.h = self.h, None
        if h and _lib is not None:
            status = getattr(_lib, self._destroy)(h)
            if status != OK and not _shut:
                check(status)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def shutdown() -> None:
    """Releases every handle still owned by a Python object, in dependency
    order (registered with atexit; safe to call more than once)."""
    global _shut
    with _live_lock:
        groups = [list(_live[k]) for k in _ORDER]
    for objs in groups:
        for o in objs:
            try:
                o.close()
            except Exception:
                pass
    _shut = True


atexit.register(shutdown)
