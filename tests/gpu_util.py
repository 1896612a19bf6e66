# SPDX-License-Identifier: Apache-2.0
"""Helpers of the GPU tests: device tensors and calls of the C ABI's primitive entry points."""

from __future__ import annotations

import ctypes

import numpy as np


def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "the GPU tests need an AMD GPU"
    return torch


def shape_arr(sh):
    return (ctypes.c_int64 * 3)(*(list(sh) + [1] * (3 - len(sh))))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def code(dtype) -> int:
    return 0 if np.dtype(dtype) == np.float32 else 1


def call(name, *args):
    from eda import _native

    lib = _native.core()
    err = _native.errbuf()
    _native.check(getattr(lib, name)(*args, err, len(err)), err)


def dev(a: np.ndarray):
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def cube(x: np.ndarray, op: str, wins, inplace: bool = False) -> np.ndarray | tuple[np.ndarray, np.ndarray]:
    """x: (B, *spatial) host array; op max / min / maxmin / box; one window per element; inplace:
    the first output is the input's own memory."""
    torch = torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    xt = dev(x)
    o0 = xt if inplace else torch.empty_like(xt)
    o1 = torch.empty_like(xt)
    w = (ctypes.c_int32 * B)(*wins)
    opc = {"max": 0, "min": 1, "maxmin": 2, "box": 3}[op]
    call("eda_prim_cube", code(x.dtype), len(sh), B, shape_arr(sh), opc, w, ptr(xt), ptr(o0), ptr(o1))
    if op == "maxmin":
        return o0.cpu().numpy(), o1.cpu().numpy()
    return o0.cpu().numpy()


def select(x: np.ndarray, mask: np.ndarray | None, mode: int, q: float = 0.5):
    B = x.shape[0]
    sh = x.shape[1:]
    xt = dev(x)
    mt = dev(mask.astype(np.uint8)) if mask is not None else None
    out = np.zeros(B)
    cnt = np.zeros(B, np.int64)
    call("eda_prim_select", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt), ptr(mt) if mt is not None else None,
         1, mode, q, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
         cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out, cnt


def noise(x: np.ndarray, per_axis: bool):
    B = x.shape[0]
    sh = x.shape[1:]
    out = np.zeros(B)
    xt = dev(x)  # kept alive across the call (a temporary's block can be handed out again)
    call("eda_prim_noise", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt), int(per_axis),
         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def stats(x: np.ndarray):
    B = x.shape[0]
    sh = x.shape[1:]
    e = np.zeros(B)
    a = np.zeros(B)
    nf = np.zeros(B, np.uint64)
    xt = dev(x)
    call("eda_prim_stats", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt),
         e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
         nf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return e, a, nf


def tau(x: np.ndarray, Wd, Wl, sigma, tol):
    torch = torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    xt = dev(x)
    out = torch.empty_like(xt)
    nf = np.zeros(B, np.int32)
    i32 = ctypes.c_int32 * B
    f64 = ctypes.c_double * B
    call("eda_prim_tau", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt), i32(*Wd), i32(*Wl), f64(*sigma), f64(*tol),
         ptr(out), nf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out.cpu().numpy(), nf


def flags(x: np.ndarray, t: np.ndarray):
    torch = torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    fl = torch.empty(x.shape, dtype=torch.uint8, device="cuda")
    ext = np.zeros(3 * B, np.uint64)
    xt = dev(x)
    tt = dev(t.astype(x.dtype))
    call("eda_prim_flags", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt), ptr(tt), ptr(fl),
         ext.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return fl.cpu().numpy(), ext.reshape(B, 3)


def raw_window(fl: np.ndarray, wins, dtype):
    torch = torch_cuda()
    B = fl.shape[0]
    sh = fl.shape[1:]
    om = torch.empty(fl.shape, dtype=torch.float32 if code(dtype) == 0 else torch.float64, device="cuda")
    ft = dev(fl.astype(np.uint8))
    call("eda_prim_raw_window", code(dtype), len(sh), B, shape_arr(sh), ptr(ft),
         (ctypes.c_int32 * B)(*wins), ptr(om))
    return om.cpu().numpy()


def ccl(mask: np.ndarray):
    torch = torch_cuda()
    B = mask.shape[0]
    sh = mask.shape[1:]
    lab = torch.empty(mask.shape, dtype=torch.int32, device="cuda")
    mt = dev(mask.astype(np.uint8))
    call("eda_prim_ccl", len(sh), B, shape_arr(sh), ptr(mt), 1, ptr(lab))
    return lab.cpu().numpy()


def dct(x: np.ndarray, inverse: int = 0, odd: int = 0):
    torch = torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    xt = dev(x)
    out = torch.empty_like(xt)
    call("eda_prim_dct", code(x.dtype), len(sh), B, shape_arr(sh), ptr(xt), ptr(out), inverse, odd)
    return out.cpu().numpy()


def orientation(x: np.ndarray, wins):
    torch = torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    D = len(sh)
    xt = dev(x)
    nv = torch.empty((D,) + x.shape, dtype=xt.dtype, device="cuda")
    c = torch.empty_like(xt)
    call("eda_prim_orientation", code(x.dtype), D, B, shape_arr(sh), ptr(xt), (ctypes.c_int32 * B)(*wins), ptr(nv),
         ptr(c))
    return nv.cpu().numpy(), c.cpu().numpy()
