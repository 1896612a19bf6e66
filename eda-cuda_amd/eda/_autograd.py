# SPDX-License-Identifier: Apache-2.0
"""The differentiable decomposition: ``eda.differentiable_decompose`` and ``eda.DecomposeFunction``
(a ``torch.autograd.Function``) over the library's hand-written adjoint (``eda_backward``).

The forward is ``eda.decompose`` of a batch of ROCm device fields with the call's decisions kept
(its tape: windows, rung ladders and the rungs used, the scale fields, the lower rung of every
voxel, the sifts run and accepted, the amplitude floors). The backward differentiates the rows with
those decisions frozen: the gradient reaches each max / min through the position it selected, every
box, gather, ramp and blend through its transpose, and (unless ``orientation_grad=False``) the
orientation's direction and coherence through the transposed Riesz transforms. The reference has
no adjoint; its forward steps are cpp/src/method/driver.cpp:540-689."""

from __future__ import annotations

import ctypes
import math
from typing import Any

from . import _core_hip, _native

_CODES = {"float32": 0, "float64": 1}


class TapedDecomposition:
    """One decomposition of a device batch ``x`` (B, *spatial) with its tape: ``rows`` (B, P,
    *spatial), ``info`` as ``eda._core_hip`` returns it, ``backward(grad_rows)`` and
    ``replay(x)``. Holds the library's result (and its device tape) until deleted."""

    def __init__(self, x, *, sifts: int = 3, max_imf: int = -1, dct: str = "exact"):
        import torch  # noqa: PLC0415

        if not isinstance(x, torch.Tensor) or x.device.type != "cuda":
            raise TypeError("eda: differentiable_decompose needs a ROCm device tensor")
        if x.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"eda: differentiable_decompose needs float32 or float64, got {x.dtype}")
        if not 2 <= x.dim() <= 4:
            raise ValueError(f"eda: expected (B, *spatial) with 1 to 3 spatial axes, got {x.dim()} axes")
        self._torch = torch
        x = x.contiguous()
        self.shape = tuple(x.shape)
        self.dtype = x.dtype
        self.device = x.device
        name = "float32" if x.dtype == torch.float32 else "float64"
        spatial = self.shape[1:]
        strides = [math.prod(spatial[a + 1:]) for a in range(len(spatial))]
        desc, keep = _core_hip._device_desc(x.data_ptr(), x.device.index or 0, name, name, (self.shape[0],),
                                            (math.prod(spatial),), spatial, strides)
        p = _core_hip._params(max_imf, None, sifts, True, False, 0,
                              int(torch.cuda.current_stream(x.device).cuda_stream))
        p.keep_tape = 1
        p.dct_mode = {"exact": 0, "bf16x3": 1}[dct]
        self._res = _core_hip._run(desc, p)
        del keep
        self.info = _core_hip._info(self._res)
        self.P = int(self._res.rows)
        self.rows = torch.empty((self.shape[0], self.P) + spatial, dtype=x.dtype, device=x.device)
        if self.rows.numel():
            err = _native.errbuf()
            _native.check(_native.core().eda_collect(ctypes.byref(self._res), ctypes.c_void_p(self.rows.data_ptr()),
                                                     self.device.index or 0, err, len(err)), err)

    def backward(self, grad_rows, orientation_grad: bool = True):
        """grad_x (B, *spatial): the rows' Jacobian transposed applied to grad_rows (B, P, *spatial)."""
        torch = self._torch
        g = grad_rows.to(device=self.device, dtype=self.dtype).contiguous()
        if tuple(g.shape) != tuple(self.rows.shape):
            raise ValueError(f"eda: grad_rows must have the rows' shape {tuple(self.rows.shape)}, got {tuple(g.shape)}")
        gx = torch.empty(self.shape, dtype=self.dtype, device=self.device)
        if gx.numel():
            torch.cuda.current_stream(self.device).synchronize()
            err = _native.errbuf()
            _native.check(_native.core().eda_backward(ctypes.byref(self._res), ctypes.c_void_p(g.data_ptr()),
                                                      ctypes.c_void_p(gx.data_ptr()),
                                                      0 if orientation_grad else 1, err, len(err)), err)
        return gx

    def replay(self, x):
        """The rows of the frozen map at another input x (the shape and dtype of the original)."""
        torch = self._torch
        xx = x.to(device=self.device, dtype=self.dtype).contiguous()
        if tuple(xx.shape) != self.shape:
            raise ValueError(f"eda: replay needs the input shape {self.shape}")
        rows = torch.empty_like(self.rows)
        if rows.numel():
            torch.cuda.current_stream(self.device).synchronize()
            err = _native.errbuf()
            _native.check(_native.core().eda_replay(ctypes.byref(self._res), ctypes.c_void_p(xx.data_ptr()),
                                                    ctypes.c_void_p(rows.data_ptr()), err, len(err)), err)
        return rows

    def mode_tape(self, element: int, mode: int) -> dict:
        """The frozen decisions of one mode of one element, in the oracle's tape layout: w, W,
        rung_w, used, pts, ran, accepted, tol, thr, and the residual r at the mode's start, the
        scale field L and the lower rung lo of every voxel as host arrays."""
        import numpy as np  # noqa: PLC0415

        spatial = self.shape[1:]
        ints = np.zeros(40, np.int32)
        dbls = np.zeros(34, np.float64)
        dt = np.float32 if self.dtype == self._torch.float32 else np.float64
        r = np.zeros(spatial, dt)
        L = np.zeros(spatial, dt)
        lo = np.zeros(spatial, np.uint8)
        err = _native.errbuf()
        _native.check(_native.core().eda_tape_mode(
            ctypes.byref(self._res), int(element), int(mode), ints.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            dbls.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_void_p(r.ctypes.data),
            ctypes.c_void_p(L.ctypes.data), ctypes.c_void_p(lo.ctypes.data), err, len(err)), err)
        K = int(ints[2])
        ran = int(ints[37])
        return {"r": r, "L": L, "lo": lo, "w": int(ints[0]), "W": int(ints[1]),
                "rung_w": [int(v) for v in ints[5:5 + K]], "used": int(ints[3]) & 0xffffffff,
                "pts": int(ints[4]) & 0xffffffff, "ran": ran,
                "accepted": [bool((int(ints[38]) >> j) & 1) for j in range(1, ran + 1)],
                "tol": float(dbls[0]), "thr": [float(dbls[1 + j]) for j in range(1, ran + 1)]}

    def __del__(self):
        res = getattr(self, "_res", None)
        if res is not None:
            try:
                _native.core().eda_release(ctypes.byref(res))
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._res = None


def _function():
    import torch  # noqa: PLC0415

    class DecomposeFunction(torch.autograd.Function):
        """rows = eda.decompose(x) for a device batch x (B, *spatial); backward by the library's
        adjoint with the forward's decisions frozen."""

        @staticmethod
        def forward(ctx: Any, x, sifts: int = 3, max_imf: int = -1, orientation_grad: bool = True):
            td = TapedDecomposition(x.detach(), sifts=sifts, max_imf=max_imf)
            ctx.td = td
            ctx.orientation_grad = bool(orientation_grad)
            return td.rows

        @staticmethod
        def backward(ctx: Any, grad_rows):
            return ctx.td.backward(grad_rows, ctx.orientation_grad), None, None, None

    return DecomposeFunction


_FN = None


def decompose_function():
    """The ``torch.autograd.Function`` class (built on first use: importing eda needs no torch)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN


def differentiable_decompose(x, *, sifts: int = 3, max_imf: int = -1, orientation_grad: bool = True):
    """The rows (B, P, *spatial) of eda.decompose for a ROCm device tensor x (B, *spatial) of
    float32 or float64, differentiable in x through the library's adjoint (decisions frozen at
    the forward's). ``orientation_grad=False`` holds the orientation fixed as well."""
    return decompose_function().apply(x, sifts, max_imf, orientation_grad)
