# SPDX-License-Identifier: Apache-2.0
"""ctypes wrapper of the oracle (liboracle.so, built by oracle/Makefile).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
it as the checker. The product package (eda-cuda_amd/eda) never imports it.
"""

from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"the oracle is not built: run make -C {HERE}")
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_rows_cap.argtypes = [ctypes.c_int, _i64p, ctypes.c_int32]
        L.oracle_rows_cap.restype = ctypes.c_int32
        L.oracle_decompose.argtypes = [ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int64, _i64p, ctypes.c_int32,
                                       ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                                       _i32p, _i32p, _i32p, _i32p, _f64p, _f64p]
        L.oracle_decompose.restype = ctypes.c_int
        L.oracle_cube.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int32, _vp]
        L.oracle_vhgw_line.argtypes = [ctypes.c_int, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int, _vp]
        L.oracle_median.argtypes = [_f64p, ctypes.c_int64]
        L.oracle_median.restype = ctypes.c_double
        L.oracle_quantile.argtypes = [_f64p, ctypes.c_int64, ctypes.c_double]
        L.oracle_quantile.restype = ctypes.c_double
        L.oracle_noise_sigma.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, ctypes.c_int]
        L.oracle_noise_sigma.restype = ctypes.c_double
        L.oracle_energy.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, _f64p, _f64p]
        L.oracle_sig_flags.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int, _i64p, _u8p, _u64p]
        L.oracle_tau.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_double, ctypes.c_double, _vp]
        L.oracle_tau.restype = ctypes.c_int
        L.oracle_raw_window.argtypes = [ctypes.c_int, _u8p, ctypes.c_int, _i64p, ctypes.c_int32, _vp]
        L.oracle_ccl.argtypes = [_u8p, ctypes.c_int, _i64p, _i32p]
        L.oracle_dct.argtypes = [_f64p, ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int, _f64p]
        L.oracle_orientation.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, ctypes.c_int32, _vp, _vp]
        L.oracle_orientation.restype = ctypes.c_int
        L.oracle_riesz.argtypes = [_f64p, ctypes.c_int, _i64p, ctypes.c_int32, _f64p, _f64p]
        L.oracle_eigen.argtypes = [ctypes.c_int, _f64p, ctypes.c_int64, _f64p, _f64p]
        L.oracle_monogenic.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _i64p, ctypes.c_void_p]
        L.oracle_monogenic.restype = None
        L.oracle_rung_ladder.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _i32p]
        L.oracle_rung_ladder.restype = ctypes.c_int
        L.oracle_interp.argtypes = [_f64p, ctypes.c_int, _i64p, _f64p, ctypes.c_int64, _f64p]
        L.oracle_decompose_tape.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64p, ctypes.c_int32, _vp, _i32p, _i32p,
                                            ctypes.c_int32, _vp, _vp, _u8p, _i32p, _f64p]
        L.oracle_decompose_tape.restype = ctypes.c_int
        _lib = L
    return _lib


def _code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise TypeError(f"oracle: float32 or float64, got {dt}")


def _shape(shape) -> ctypes.Array:
    return (ctypes.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))


def _p(a: np.ndarray, t=_vp):
    return ctypes.cast(a.ctypes.data, t)


def rows_cap(shape, max_imf: int = -1) -> int:
    return int(lib().oracle_rows_cap(len(shape), _shape(shape), max_imf))


def decompose(x: np.ndarray, dim: int | None = None, *, max_imf: int = -1, tie_tolerance: float | None = None,
              sifts: int = 3, check_finite: bool = True, threads: int = 0):
    """The oracle's decomposition of x (B, *spatial) or (*spatial) in x's dtype: (rows, info).

    rows has shape (B, P, *spatial) with P = max n_imfs + 1 over the elements without an error
    (NaN rows for a failed element); info holds n_imfs, error, windows, sifts, spacing, energy
    with (B,) and (B, K) shapes as the reference's binding returns them."""
    x = np.ascontiguousarray(x)
    if dim is None:
        dim = x.ndim
        x = x[None]
    spatial = tuple(x.shape[1:])
    if len(spatial) != dim:
        raise ValueError("oracle: x must be (B, *spatial)")
    B = x.shape[0]
    count = math.prod(spatial)
    cap = rows_cap(spatial, max_imf)
    rows = np.zeros((B, cap) + spatial, dtype=x.dtype)
    n_imfs = np.zeros(B, np.int32)
    error = np.zeros(B, np.int32)
    hw = np.zeros((B, cap - 1), np.int32)
    hs = np.zeros((B, cap - 1), np.int32)
    hsp = np.zeros((B, cap - 1), np.float64)
    he = np.zeros((B, cap - 1), np.float64)
    tie = -1.0 if tie_tolerance is None else float(tie_tolerance)
    rc = lib().oracle_decompose(_code(x.dtype), _p(x), dim, B, _shape(spatial), max_imf, tie, sifts,
                                1 if check_finite else 0, threads, _p(rows), _p(n_imfs, _i32p),
                                _p(error, _i32p), _p(hw, _i32p), _p(hs, _i32p), _p(hsp, _f64p), _p(he, _f64p))
    if rc != 0:
        raise RuntimeError("oracle: the decomposition raised (an internal invariant failed)")
    ok = error == 0
    P = int(n_imfs[ok].max(initial=-1)) + 1
    if not ok.all():
        P = max(P, 1)
    K = int(n_imfs.max(initial=0))
    out = rows[:, :P].copy()
    for b in range(B):
        if error[b]:
            out[b] = np.nan
        else:
            out[b, n_imfs[b] + 1:] = 0
    info = {"n_imfs": n_imfs, "error": error, "windows": hw[:, :K], "sifts": hs[:, :K],
            "spacing": hsp[:, :K], "energy": he[:, :K]}
    del count
    return out, info


def cube(x: np.ndarray, op: str, w: int) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    code = {"max": 0, "min": 1, "box": 3}[op]
    lib().oracle_cube(_code(x.dtype), _p(x), x.ndim, _shape(x.shape), code, w, _p(out))
    return out


def vhgw_line(x: np.ndarray, w: int, op: str) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    lib().oracle_vhgw_line(_code(x.dtype), _p(x), x.size, w, {"max": 0, "min": 1, "sum": 2}[op], _p(out))
    return out


def median(v) -> float:
    a = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().oracle_median(_p(a, _f64p), a.size))


def quantile(v, q: float) -> float:
    a = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().oracle_quantile(_p(a, _f64p), a.size, q))


def noise_sigma(r: np.ndarray, per_axis: bool) -> float:
    r = np.ascontiguousarray(r)
    return float(lib().oracle_noise_sigma(_code(r.dtype), _p(r), r.ndim, _shape(r.shape), int(per_axis)))


def energy(f: np.ndarray) -> tuple[float, float]:
    f = np.ascontiguousarray(f)
    e, a = ctypes.c_double(), ctypes.c_double()
    lib().oracle_energy(_code(f.dtype), _p(f), f.ndim, _shape(f.shape), ctypes.byref(e), ctypes.byref(a))
    return e.value, a.value


def sig_flags(r: np.ndarray, tau: np.ndarray):
    r = np.ascontiguousarray(r)
    tau = np.ascontiguousarray(tau, dtype=r.dtype)
    flags = np.zeros(r.shape, np.uint8)
    ext = np.zeros(3, np.uint64)
    lib().oracle_sig_flags(_code(r.dtype), _p(r), _p(tau), r.ndim, _shape(r.shape), _p(flags, _u8p), _p(ext, _u64p))
    return flags, ext[: r.ndim]


def tau(r: np.ndarray, Wd: int, Wl: int, sigma_n: float, tol: float):
    r = np.ascontiguousarray(r)
    out = np.empty_like(r)
    ok = lib().oracle_tau(_code(r.dtype), _p(r), r.ndim, _shape(r.shape), Wd, Wl, sigma_n, tol, _p(out))
    return out, bool(ok)


def raw_window(flags: np.ndarray, win: int, dtype) -> np.ndarray:
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    out = np.empty(flags.shape, dtype=dtype)
    lib().oracle_raw_window(_code(dtype), _p(flags, _u8p), flags.ndim, _shape(flags.shape), win, _p(out))
    return out


def ccl(mask: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    labels = np.empty(m.shape, np.int32)
    lib().oracle_ccl(_p(m, _u8p), m.ndim, _shape(m.shape), _p(labels, _i32p))
    return labels


def dct(x: np.ndarray, inverse: bool = False, odd: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().oracle_dct(_p(x, _f64p), x.ndim, _shape(x.shape), int(inverse), odd, _p(out, _f64p))
    return out


def orientation(r: np.ndarray, w: int):
    r = np.ascontiguousarray(r)
    n = np.empty((r.ndim,) + r.shape, r.dtype)
    c = np.empty(r.shape, r.dtype)
    ok = lib().oracle_orientation(_code(r.dtype), _p(r), r.ndim, _shape(r.shape), w, _p(n), _p(c))
    return n, c, bool(ok)


def monogenic(x: np.ndarray) -> dict:
    """The monogenic signal of one field x (*spatial) in x's dtype: amplitude, phase, frequency (the
    shape of x) and, in 2-D and 3-D, direction (D, *spatial) and coherence (None in 1-D)."""
    x = np.ascontiguousarray(x)
    D = x.ndim
    planes = 3 if D == 1 else 4 + D
    out = np.empty((planes,) + x.shape, x.dtype)
    lib().oracle_monogenic(_code(x.dtype), _p(x), D, _shape(x.shape), _p(out))
    res = {"amplitude": out[0], "phase": out[1], "frequency": out[2], "direction": None, "coherence": None}
    if D >= 2:
        res["direction"] = out[3:3 + D]
        res["coherence"] = out[3 + D]
    return res


def riesz(r: np.ndarray, w: int):
    r = np.ascontiguousarray(r, dtype=np.float64)
    D = r.ndim
    r1 = np.empty((D,) + r.shape)
    r2 = np.empty((D * (D + 1) // 2,) + r.shape)
    lib().oracle_riesz(_p(r, _f64p), D, _shape(r.shape), w, _p(r1, _f64p), _p(r2, _f64p))
    return r1, r2


def eigen(t: np.ndarray, dim: int):
    t = np.ascontiguousarray(t, dtype=np.float64)
    count = t.shape[0]
    n = np.empty((count, dim))
    c = np.empty(count)
    lib().oracle_eigen(dim, _p(t, _f64p), count, _p(n, _f64p), _p(c, _f64p))
    return n, c


def rung_ladder(d_min: float, d_max: float, w: int, n_max: int):
    out = (ctypes.c_int32 * 32)()
    k = lib().oracle_rung_ladder(d_min, d_max, w, n_max, out)
    if k < 0:
        raise ValueError("oracle: the ladder exceeds 32 rungs")
    return list(out[:k])


def interp(f: np.ndarray, y: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, f.ndim)
    out = np.empty(y.shape[0])
    lib().oracle_interp(_p(f, _f64p), f.ndim, _shape(f.shape), _p(y, _f64p), y.shape[0], _p(out, _f64p))
    return out


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


TAPE_INTS, TAPE_DBLS = 40, 34


def decompose_tape(x: np.ndarray, *, sifts: int = 3) -> dict:
    """One element's decomposition (x: the spatial array, float32 or float64) with its tape: the
    frozen decisions of every mode (the residual r at the mode's start, the scale field L, the lower
    rung lo per voxel, the windows w and W, the rung ladder with its used / segment-point bits, the
    tie tolerance, the sifts run and accepted, the amplitude floors thr_j) and the rows (P,
    *spatial), n_imfs, error, kexp."""
    x = np.ascontiguousarray(x)
    spatial = tuple(x.shape)
    cap = rows_cap(spatial)
    modes = cap - 1
    rows = np.zeros((cap,) + spatial, dtype=x.dtype)
    r = np.zeros((modes,) + spatial, dtype=x.dtype)
    L = np.zeros((modes,) + spatial, dtype=x.dtype)
    lo = np.zeros((modes,) + spatial, dtype=np.uint8)
    ints = np.zeros((modes, TAPE_INTS), np.int32)
    dbls = np.zeros((modes, TAPE_DBLS), np.float64)
    n_imfs = np.zeros(1, np.int32)
    error = np.zeros(1, np.int32)
    rc = lib().oracle_decompose_tape(_code(x.dtype), _p(x), len(spatial), _shape(spatial), sifts, _p(rows),
                                     _p(n_imfs, _i32p), _p(error, _i32p), modes, _p(r), _p(L), _p(lo, _u8p),
                                     _p(ints, _i32p), _p(dbls, _f64p))
    if rc != 0:
        raise RuntimeError("oracle: the decomposition raised (an internal invariant failed)")
    k = int(n_imfs[0])
    mode_list = []
    for m in range(k):
        iv = ints[m]
        K = int(iv[2])
        ran = int(iv[37])
        mode_list.append({
            "r": r[m], "L": L[m], "lo": lo[m], "w": int(iv[0]), "W": int(iv[1]),
            "rung_w": [int(v) for v in iv[5:5 + K]], "used": int(iv[3]) & 0xffffffff,
            "pts": int(iv[4]) & 0xffffffff, "ran": ran,
            "accepted": [bool((int(iv[38]) >> j) & 1) for j in range(1, ran + 1)],
            "tol": float(dbls[m, 0]), "r_inf": float(dbls[m, 1]),
            "thr": [float(dbls[m, 1 + j]) for j in range(1, ran + 1)],
        })
    kexp = int(ints[0, 39]) if k > 0 else 0
    return {"rows": rows[: k + 1].copy(), "n_imfs": k, "error": int(error[0]), "kexp": kexp, "modes": mode_list,
            "sifts": sifts}
