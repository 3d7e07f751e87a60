"""Differentiable composite + stamping behind ``torch.autograd.Function``.

The forward is a stamp list on a canvas (FootprintBrush::imprint,
renderer/FootprintBrush.hxx:73-143) whose dips take their paint from
PaintMixer::mixSinglePaint (mixer/src/PaintMixer.cxx:327-355) of one row of
a mixing-weight matrix W, followed by Renderer::compose
(renderer/Renderer.hxx:26-41).  The backward is the library's hand-written
adjoint (pnt_grad_backward): gradients with respect to the initial wet-layer
K, S and thickness V, the substrate R0 and W.  Torch provides device memory;
every computation runs in libpainty_hip.so.

Tensor layout (cuda, float32, contiguous): K and S (3, rows, cols) -- one
plane per channel --, V (rows, cols), R0 (3, rows, cols); the result R is
(3, rows, cols).  W is (n_paint, n_pig) float64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from painty_amd import _lib
from painty_amd import hip as H


def _planes(*tensors):
    out = []
    for t in tensors:
        assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), t.dtype
        if t.dim() == 3:
            out += [t[c].data_ptr() for c in range(t.shape[0])]
        else:
            out.append(t.data_ptr())
    return out


class GradRunner(_lib.Owned):
    """One differentiable stamp list (pnt_grad).  Holds the checkpoints of its
    last forward; backward applies to that forward."""

    _kind, _destroy = "grad", "pnt_grad_destroy"

    def __init__(self, rows: int, cols: int, brush: H.FootprintBrush, stamps, paint_of, palK,
                 palS, n_paint: int, seg_len: int = 32, ctx: H.Context | None = None):
        self.ctx = ctx or H.default_context()
        self.rows, self.cols = rows, cols
        self.brush = brush
        self._arr = H.stamps_to_c(stamps)
        po = np.ascontiguousarray(paint_of, dtype=np.int32)
        self.palK = np.ascontiguousarray(palK, dtype=np.float64)
        self.palS = np.ascontiguousarray(palS, dtype=np.float64)
        self.n_pig, self.n_paint = self.palK.shape[0], int(n_paint)
        h = C.c_void_p()
        _lib.call("pnt_grad_create", self.ctx.h, rows, cols, brush.h, self._arr, len(self._arr),
                  po.ctypes.data_as(C.POINTER(C.c_int32)), self.n_pig,
                  self.palK.ctypes.data_as(_lib._dp), self.palS.ctypes.data_as(_lib._dp),
                  self.n_paint, int(seg_len), C.byref(h))
        self._own(h)

    def forward(self, K, S, V, R0, W) -> torch.Tensor:
        rc = (self.rows, self.cols)
        assert K.shape == S.shape == R0.shape == (3,) + rc and V.shape == rc, "plane shapes"
        W = np.ascontiguousarray(W.detach().cpu().numpy() if torch.is_tensor(W) else W,
                                 dtype=np.float64)
        assert W.shape == (self.n_paint, self.n_pig)
        R = torch.empty((3, self.rows, self.cols), device=K.device, dtype=torch.float32)
        torch.cuda.synchronize(K.device)  # inputs written on torch's stream
        _lib.call("pnt_grad_forward", self.h, _lib.DevLayer((C.c_void_p * 7)(*_planes(K, S, V))),
                  _lib.DevRgb((C.c_void_p * 3)(*_planes(R0))), W.ctypes.data_as(_lib._dp),
                  _lib.DevRgb((C.c_void_p * 3)(*_planes(R))))
        self.ctx.synchronize()
        return R

    def backward(self, gR: torch.Tensor):
        assert gR.shape == (3, self.rows, self.cols), "gradient shape"
        gR = gR.contiguous()
        dev = gR.device
        gK = torch.empty((3, self.rows, self.cols), device=dev, dtype=torch.float32)
        gS = torch.empty_like(gK)
        gV = torch.empty((self.rows, self.cols), device=dev, dtype=torch.float32)
        gR0 = torch.empty_like(gK)
        gW = np.zeros((self.n_paint, self.n_pig), np.float64)
        torch.cuda.synchronize(dev)
        _lib.call("pnt_grad_backward", self.h, _lib.DevRgb((C.c_void_p * 3)(*_planes(gR))),
                  _lib.DevLayer((C.c_void_p * 7)(*_planes(gK, gS, gV))),
                  _lib.DevRgb((C.c_void_p * 3)(*_planes(gR0))), gW.ctypes.data_as(_lib._dp))
        self.ctx.synchronize()
        return gK, gS, gV, gR0, gW


class PaintStamps(torch.autograd.Function):
    """R = compose(stamps(K, S, V over R0; paints = mix(W))).

    ``PaintStamps.apply(K, S, V, R0, W, runner)``; W is a float64 tensor
    (n_paint, n_pig) on any device, its gradient is returned on W's device."""

    @staticmethod
    def forward(ctx, K, S, V, R0, W, runner: GradRunner):
        ctx.runner = runner
        ctx.w_device = W.device
        return runner.forward(K, S, V, R0, W)

    @staticmethod
    def backward(ctx, gR):
        gK, gS, gV, gR0, gW = ctx.runner.backward(gR)
        return gK, gS, gV, gR0, torch.from_numpy(gW).to(ctx.w_device), None
