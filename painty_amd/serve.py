"""Python view of the request server (pnt_server_*, include/painty_hip.h).

Requests are submitted from Python and run on the library's worker slots;
``wait`` returns the request's latency (submit to done, in ms)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from painty_amd import _lib
from painty_amd import hip as H


class Server(_lib.Owned):
    _kind, _destroy = "server", "pnt_server_destroy"

    def __init__(self, n_slots: int = 4, hbm_budget_bytes: int = 64 << 30, device: int = 0):
        self._keep = {}  # ticket -> host buffers that must outlive the request
        h = C.c_void_p()
        _lib.call("pnt_server_create", device, n_slots, int(hbm_budget_bytes), C.byref(h))
        self._own(h)

    def close(self) -> None:
        super().close()  # waits for the queued requests
        self._keep.clear()

    def submit(self, canvas_id: int, rows: int, cols: int, stamps, radius: float, footprint,
               background=(1.0, 1.0, 1.0), pickup_rate: float = 0.9,
               deposition_rate: float = 0.05, use_snapshot: bool = True,
               out: np.ndarray | None = None) -> int:
        """Queue stamps + composite on canvas `canvas_id`; `out` (3, rows, cols)
        float32 receives the composite if given.  Returns a ticket."""
        fp = np.ascontiguousarray(footprint, dtype=np.float32)
        arr = H.stamps_to_c(stamps)
        r = _lib.PntRequest()
        r.canvas_id, r.rows, r.cols = int(canvas_id), int(rows), int(cols)
        for c in range(3):
            r.background[c] = float(background[c])
        r.radius, r.width = float(radius), int(fp.shape[0])
        r.footprint = fp.ctypes.data_as(C.POINTER(C.c_float))
        r.pickup_rate, r.deposition_rate = float(pickup_rate), float(deposition_rate)
        r.use_snapshot = 1 if use_snapshot else 0
        r.stamps, r.n_stamps = arr, len(arr)
        if out is not None:
            assert out.dtype == np.float32 and out.shape == (3, rows, cols) and \
                out.flags.c_contiguous
            r.out = out.ctypes.data_as(C.POINTER(C.c_float))
        t = C.c_int64()
        _lib.call("pnt_server_submit", self.h, C.byref(r), C.byref(t))
        if out is not None:
            self._keep[t.value] = out
        return t.value

    def wait(self, ticket: int) -> float:
        ms = C.c_double()
        try:
            _lib.call("pnt_server_wait", self.h, int(ticket), C.byref(ms))
        finally:
            self._keep.pop(ticket, None)
        return ms.value

    def canvas_state(self, canvas_id: int):
        """Wet layer K, S, V, R0 and the last composite of a resident canvas
        (host arrays, reference layouts)."""
        cvh, comp = C.c_void_p(), _lib.DevRgb()
        _lib.call("pnt_server_canvas", self.h, int(canvas_id), C.byref(cvh), C.byref(comp))
        rows, cols = C.c_int(), C.c_int()
        _lib.call("pnt_canvas_shape", cvh, C.byref(rows), C.byref(cols))
        n = rows.value * cols.value
        K = np.empty((rows.value, cols.value, 3), np.float32)
        S = np.empty_like(K)
        V = np.empty((rows.value, cols.value), np.float64)  # FP64, as the device holds it
        R0 = np.empty_like(K)
        lay = C.c_void_p()
        _lib.call("pnt_canvas_layer", cvh, C.byref(lay))
        _lib.call("pnt_layer_download", lay, H._pf(K), H._pf(S), None, _lib.AOS)
        _lib.call("pnt_layer_download_volume", lay, V.ctypes.data_as(C.POINTER(C.c_double)))
        _lib.call("pnt_canvas_download_R0", cvh, H._pf(R0), _lib.AOS)
        R = np.empty((3, n), np.float32)
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for c in range(3):
            assert hip.hipMemcpy(R[c].ctypes.data, comp.p[c], n * 4, 2) == 0  # device -> host
        return K, S, V, R0, R.reshape(3, rows.value, cols.value)

    def release_canvas(self, canvas_id: int) -> None:
        """Ends a canvas: released once no request on it is queued or
        running (a request submitted on the id before then cancels this)."""
        _lib.call("pnt_server_release", self.h, int(canvas_id))

    def stats(self):
        b, c, d = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.call("pnt_server_stats", self.h, C.byref(b), C.byref(c), C.byref(d))
        return {"resident_bytes": b.value, "resident_canvases": c.value, "completed": d.value}
