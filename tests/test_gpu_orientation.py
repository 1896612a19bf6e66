# SPDX-License-Identifier: Apache-2.0
"""``eda.decompose(..., return_orientation=True)`` on the GPU: the direction n and the coherence c
that the method computes at the start of each mode, kept beside the rows (reference
src/eda/_api.py:64-83, cpp/src/method/orientation.cu:495-549; the reference's own checks are
tests/python/test_orientation_output.py).

- Asking for them changes nothing else: the rows and the information are the same bits.
- Layout: ``direction`` ``x.shape[:m] + (K, D) + x.shape[m:]``, ``coherence``
  ``x.shape[:m] + (K,) + x.shape[m:]``; rows beyond n_imfs and every row of a failed element are 0;
  |n| = 1 where c > 0, c in [0, 1], and the canonical sign (the first component of magnitude above
  the dtype's default tie tolerance is positive).
- Values: row k equals the oracle's orientation of the residual of mode k (x minus the earlier
  modes, rounded to the dtype after each subtraction) at the mode's window, and not that of a
  neighbouring row; on a fast oblique plane wave over a slow one, the first mode's direction is the
  fast wave's unit wave vector (a closed form, independent of the oracle).
- Bitwise: host and ROCm input, an element alone and in a batch, any chunking, max_imf.
"""

from __future__ import annotations

import numpy as np
import pytest

import eda
import oracle

pytestmark = pytest.mark.gpu

ZERO = {"float32": 2.0**-14, "float64": 2.0**-30}
EPS = {"float32": float(np.finfo(np.float32).eps), "float64": float(np.finfo(np.float64).eps)}
# The oracle comparison: c within C_TOL, and |n . n_ref| within N_TOL of 1 where the oracle's
# coherence exceeds 0.2 (n is defined there). Measured on the 2-D batch, float64 / float32: |c -
# c_ref| 1.2e-14 / 1.3e-6, 1 - |n . n_ref| 2.2e-16 / 4.1e-8; a neighbouring row gives |c - c_ref|
# 0.84.
C_TOL = {"float32": 1e-4, "float64": 1e-11}
N_TOL = {"float32": 1e-5, "float64": 1e-11}


def bits(a) -> bytes:
    a = np.ascontiguousarray(a)
    return a.tobytes() + str(a.shape).encode() + str(a.dtype).encode()


def waves(shape, seed, noise=0.02):
    """Plane waves of periods 5, 13 and 34 (2-D) or 4, 9 and 20 (3-D) in random directions plus a
    little noise: fields with one or more modes, each mode a clearly oriented pattern."""
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    out = rng.normal(0.0, noise, shape)
    periods = (5.0, 13.0, 34.0) if len(shape) <= 2 else (4.0, 9.0, 20.0)
    for period, amp in zip(periods, (1.0, 1.5, 2.0)):
        k = rng.normal(size=len(shape))
        k *= 2 * np.pi / period / np.linalg.norm(k)
        out += amp * np.cos(sum(ki * gi for ki, gi in zip(k, g)) + rng.uniform(0, 2 * np.pi))
    return out


def canonical_first(n, zero):
    """The first component of magnitude above zero of every direction (n: (..., D)), 0 if none."""
    first = np.zeros(n.shape[:-1])
    for a in range(n.shape[-1]):
        first = np.where((first == 0) & (np.abs(n[..., a]) > zero), n[..., a], first)
    return first


def check_rows(o, info, axes):
    """The layout and the invariants of one call's orientation (batch axes before the spatial
    axes)."""
    m, D = min(axes), len(axes)
    K = info.windows.shape[-1]
    n = np.moveaxis(np.asarray(o.direction), (m, m + 1), (-2, -1))  # (*batch, *spatial, K, D)
    c = np.moveaxis(np.asarray(o.coherence), m, -1)  # (*batch, *spatial, K)
    assert n.shape[-2:] == (K, D) and c.shape[-1] == K
    B = max(info.n_imfs.size, 1)
    flat_n = n.reshape(B, -1, K, D)
    flat_c = c.reshape(B, -1, K)
    dtype = info.dtype
    for b, (kb, err) in enumerate(zip(info.n_imfs.reshape(-1), info.error.reshape(-1))):
        kb = 0 if err else int(kb)
        assert not flat_n[b, :, kb:].any() and not flat_c[b, :, kb:].any()
        cb, nb = flat_c[b, :, :kb], flat_n[b, :, :kb].astype(np.float64)
        assert (cb >= 0).all() and (cb <= 1).all()
        norm = np.sqrt((nb**2).sum(axis=-1))
        assert np.abs(norm[cb > 0] - 1).max(initial=0) <= 8 * EPS[dtype]
        assert (canonical_first(nb, ZERO[dtype]) > 0).all(), "the canonical sign"


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("shape", [(3, 48, 40), (2, 24, 22, 20)], ids=["2d", "3d"])
def test_rows_and_info_do_not_change(shape, dtype):
    print("orientation: rows and info", shape, dtype, flush=True)
    x = np.stack([waves(shape[1:], 30 + b) for b in range(shape[0])]).astype(dtype)
    axes = tuple(range(1, len(shape)))
    y, info = eda.decompose(x, axes=axes, return_info=True)
    y2, info2, o = eda.decompose(x, axes=axes, return_info=True, return_orientation=True)
    assert bits(y) == bits(y2)
    for name in ("n_imfs", "error", "windows", "sifts", "spacing", "energy"):
        assert bits(getattr(info, name)) == bits(getattr(info2, name)), name
    y3, o3 = eda.decompose(x, axes=axes, return_orientation=True)
    assert bits(y3) == bits(y) and bits(o3.direction) == bits(o.direction)
    assert bits(o3.coherence) == bits(o.coherence)
    assert int(info.n_imfs.min()) >= 1
    K, D = info.windows.shape[-1], len(axes)
    assert o.direction.shape == (shape[0], K, D) + shape[1:] and o.coherence.shape == (shape[0], K) + shape[1:]
    assert o.direction.dtype == np.dtype(dtype) and o.coherence.dtype == np.dtype(dtype)
    check_rows(o, info, axes)


def test_layout_with_batch_axes_in_between():
    print("orientation: layout", flush=True)
    x = np.stack([waves((40, 36), 40 + b) for b in range(3)], axis=1)  # (40, 3, 36)
    y, info, o = eda.decompose(x, axes=(2, 0), return_info=True, return_orientation=True)
    K = info.windows.shape[-1]
    assert o.direction.shape == (K, 2, 40, 3, 36) and o.coherence.shape == (K, 40, 3, 36)
    one, info_one, o_one = eda.decompose(np.ascontiguousarray(x[:, 1, :].T), return_info=True,
                                         return_orientation=True)
    k = int(info_one.n_imfs)
    assert int(info.n_imfs[1]) == k
    # element 1 in the order of axes (2, 0) is the transposed field: component 0 along axis 2
    assert bits(o.direction[:k, :, :, 1, :]) == bits(o_one.direction[:k].transpose(0, 1, 3, 2))
    assert bits(o.coherence[:k, :, 1, :]) == bits(o_one.coherence[:k].transpose(0, 2, 1))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("shape", [(4, 64, 56), (3, 24, 22, 20)], ids=["2d", "3d"])
def test_every_row_matches_the_oracle(shape, dtype):
    """Row k of element b against the oracle's orientation (float64) of the residual of mode k.
    Element 1 is constant (no mode); element 3 of the 2-D batch holds a NaN (every row 0)."""
    print("orientation: oracle rows", shape, dtype, flush=True)
    x = np.stack([waves(shape[1:], 100 + b) for b in range(shape[0])]).astype(dtype)
    x[1] = 5.0
    if shape[0] > 3:
        x[(3,) + (2,) * (len(shape) - 1)] = np.nan
    axes = tuple(range(1, len(shape)))
    D = len(axes)
    y, info, o = eda.decompose(x, axes=axes, errors="status", return_info=True, return_orientation=True)
    assert info.n_imfs[1] == 0
    if shape[0] > 3:
        assert info.error[3] == 1
    modes, worst_c, worst_n, nearest = 0, 0.0, 0.0, np.inf
    for b in range(shape[0]):
        kb = 0 if info.error[b] else int(info.n_imfs[b])
        assert not o.direction[b, kb:].any() and not o.coherence[b, kb:].any()
        r = x[b].copy()
        for k in range(kb):
            rn, rc, ok = oracle.orientation(r.astype(np.float64), int(info.windows[b, k]))
            assert ok
            c = o.coherence[b, k].astype(np.float64)
            n = o.direction[b, k].astype(np.float64)
            worst_c = max(worst_c, float(np.abs(c - rc).max()))
            strong = rc > 0.2
            dots = np.abs(sum(n[a] * rn[a] for a in range(D)))
            if strong.any():
                worst_n = max(worst_n, float(1 - dots[strong].min()))
            for other in (k - 1, k + 1):
                if 0 <= other < kb:
                    nearest = min(nearest, float(np.abs(o.coherence[b, other].astype(np.float64) - rc).max()))
            modes += 1
            r = (r - y[b, k]).astype(dtype)
    print(f"orientation rows {shape} {dtype}: modes {modes} |c - c_ref| {worst_c:.3g} 1 - |n.n_ref| {worst_n:.3g} "
          f"neighbour |c - c_ref| {nearest:.3g}", flush=True)
    assert modes >= (4 if D == 2 else 3)
    assert worst_c <= C_TOL[dtype] and worst_n <= N_TOL[dtype]
    assert nearest >= 0.1, "a neighbouring row matches the mode's residual as well"


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_the_first_mode_follows_the_fast_wave(dtype):
    """A plane wave cos(2 pi k . p) is oriented along k: with a fast wave (period 8.5) over a slow
    one (period 50), the first mode's direction is k_fast / |k_fast| (first component positive)
    away from the borders, with coherence near 1."""
    print("orientation: fast wave", dtype, flush=True)
    k_fast, k_slow = np.array([0.07, 0.095]), np.array([0.016, -0.012])
    g = np.meshgrid(np.arange(160.0), np.arange(150.0), indexing="ij")
    x = (np.cos(2 * np.pi * (k_fast[0] * g[0] + k_fast[1] * g[1]) + 0.3)
         + np.cos(2 * np.pi * (k_slow[0] * g[0] + k_slow[1] * g[1]) + 1.1)).astype(dtype)
    y, info, o = eda.decompose(x, return_info=True, return_orientation=True)
    assert int(info.n_imfs) >= 2
    margin = int(np.ceil(2 / np.linalg.norm(k_fast)))
    inner = (slice(margin, -margin), slice(margin, -margin))
    khat = k_fast / np.linalg.norm(k_fast)
    err = max(float(np.abs(o.direction[0, a][inner] - khat[a]).max()) for a in range(2))
    print(f"orientation fast wave {dtype}: |n - k/|k|| {err:.3g} min c {o.coherence[0][inner].min():.6f}", flush=True)
    assert err <= 2.5e-3
    assert o.coherence[0][inner].min() >= 0.99


def test_failed_elements_and_max_imf():
    print("orientation: failed elements, max_imf", flush=True)
    x = np.stack([waves((40, 36), 50 + b) for b in range(3)])
    x[1, 3, 3] = np.nan
    y, info, o = eda.decompose(x, axes=(1, 2), errors="status", return_info=True, return_orientation=True)
    assert info.error[1] == 1 and not o.direction[1].any() and not o.coherence[1].any()
    check_rows(o, info, (1, 2))
    full, info_full, o_full = eda.decompose(x[0], return_info=True, return_orientation=True)
    one, o_one = eda.decompose(x[0], max_imf=1, return_orientation=True)
    assert o_one.direction.shape[0] == 1 and int(info_full.n_imfs) >= 2
    assert bits(o_one.direction) == bits(o_full.direction[:1])
    assert bits(o_one.coherence) == bits(o_full.coherence[:1])


@pytest.mark.parametrize(("shape", "dtype", "batch", "positions"),
                         [((16, 51), "float32", 70, (0, 63, 64, 69)), ((12, 23, 19), "float64", 52, (0, 48, 49, 51))],
                         ids=["2d", "3d"])
def test_an_element_does_not_see_the_batch(shape, dtype, batch, positions):
    print("orientation: batch invariance", shape, dtype, flush=True)
    x = waves(shape, 7).astype(dtype)
    _, info_alone, alone = eda.decompose(x, return_info=True, return_orientation=True)
    k = int(info_alone.n_imfs)
    stack = np.stack([waves(shape, 20 + b).astype(dtype) for b in range(batch)])
    for p in positions:
        stack[p] = x
    axes = tuple(range(1, len(shape) + 1))
    _, info, o = eda.decompose(stack, axes=axes, return_info=True, return_orientation=True)
    for p in positions:
        assert int(info.n_imfs[p]) == k
        assert bits(o.direction[p, :k]) == bits(alone.direction[:k])
        assert bits(o.coherence[p, :k]) == bits(alone.coherence[:k])
        assert not o.direction[p, k:].any()


def test_chunks_do_not_change_the_orientation():
    """The reference's native interface with max_chunk 3: the same bits as one chunk."""
    from eda import _core_hip

    print("orientation: chunks", flush=True)
    x = np.stack([waves((24, 20), 80 + b) for b in range(7)]).astype(np.float32)
    _, ref = _core_hip.decompose_host(x, -1, None, 3, True, True, 0)
    _, info = _core_hip.decompose_host(x, -1, None, 3, True, True, 3)
    assert info["chunk"] == 3 and ref["chunk"] == 7
    assert bits(info["direction"]) == bits(ref["direction"]) and bits(info["coherence"]) == bits(ref["coherence"])
    _, plain = _core_hip.decompose_host(x, -1, None, 3, True, False, 0)
    assert "direction" not in plain and "coherence" not in plain


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_device_input_gives_the_host_bits(dtype):
    import torch

    from eda import _core_hip

    print("orientation: device input", dtype, flush=True)
    for shape in [(3, 48, 40), (2, 20, 18, 16)]:
        x = np.stack([waves(shape[1:], 60 + b) for b in range(shape[0])]).astype(dtype)
        axes = tuple(range(1, len(shape)))
        y, info, o = eda.decompose(x, axes=axes, return_info=True, return_orientation=True)
        xd = torch.from_numpy(x).cuda()
        yd, info_d, od = eda.decompose(xd, axes=axes, return_info=True, return_orientation=True)
        assert bits(info_d.n_imfs) == bits(info.n_imfs) and bits(yd.cpu().numpy()) == bits(y)
        for name in ("direction", "coherence"):
            d = getattr(od, name)
            assert isinstance(d, torch.Tensor) and d.is_cuda and tuple(d.shape) == getattr(o, name).shape
            assert bits(d.cpu().numpy()) == bits(getattr(o, name))
        # the reference's native interface: one device buffer, the direction then the coherence
        owner, ptr, raw = _core_hip.decompose_device(
            xd.data_ptr(), 0, dtype, dtype, (shape[0],), (int(np.prod(shape[1:])),), shape[1:],
            tuple(s // x.itemsize for s in x.strides[1:]), (), -1, None, 3, True, True, 0)
        buf, optr = raw["orientation"]
        assert optr == buf.data_ptr()
        nd = o.direction.size
        flat = buf.cpu().numpy()
        assert bits(flat[:nd].reshape(o.direction.shape)) == bits(o.direction)
        assert bits(flat[nd:].reshape(o.coherence.shape)) == bits(o.coherence)
    _, od = eda.decompose(torch.zeros((0, 16, 16), dtype=getattr(torch, dtype), device="cuda"), axes=(1, 2),
                          return_orientation=True)
    assert tuple(od.direction.shape) == (0, 0, 2, 16, 16) and tuple(od.coherence.shape) == (0, 0, 16, 16)
