# SPDX-License-Identifier: Apache-2.0
"""eda_prim_select_noalloc (the allocation-free order statistic, whose workspace is a static device
array) against eda_prim_select (which allocates its workspace) through the C ABI (GPU): bit-equal
medians, quantiles and counts on two calls in a row, the second with a smaller input, so that the
second call must clear what the first left in the shared array; plus the size limit."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import gpu_util as G

pytestmark = pytest.mark.gpu


def noalloc(x, mask, mode, q, stream):
    torch = G.torch_cuda()
    B = x.shape[0]
    sh = x.shape[1:]
    xt = G.dev(x)
    mt = G.dev(mask.astype(np.uint8)) if mask is not None else None
    out = torch.full((B,), -1.0, dtype=torch.float64, device="cuda")
    cnt = torch.full((B,), -1, dtype=torch.int64, device="cuda")
    G.call("eda_prim_select_noalloc", G.code(x.dtype), len(sh), B, G.shape_arr(sh), G.ptr(xt),
           G.ptr(mt) if mt is not None else None, 1, mode, q, G.ptr(out), G.ptr(cnt), ctypes.c_uint64(stream))
    torch.cuda.synchronize()
    return out.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("mode,q", [(0, 0.5), (1, 0.25), (1, 0.9)])
def test_noalloc_select_equals_the_allocating_entry(dtype, mode, q):
    torch = G.torch_cuda()
    stream = int(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(7 + mode)
    cases = [
        (rng.standard_normal((6, 96, 80)).astype(dtype), rng.random((6, 96, 80)) < 0.6),
        (rng.standard_normal((3, 40, 33)).astype(dtype), None),  # smaller, after the larger one
    ]
    for x, mask in cases:
        want, want_n = G.select(x, mask, mode, q)
        got, got_n = noalloc(x, mask, mode, q, stream)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
        assert np.array_equal(got_n, want_n)


def test_noalloc_select_empty_sets_and_ties():
    torch = G.torch_cuda()
    stream = int(torch.cuda.current_stream().cuda_stream)
    x = np.zeros((3, 50), np.float32)
    x[1] = np.repeat(np.arange(5, dtype=np.float32), 10)
    mask = np.ones((3, 50), bool)
    mask[0] = False
    got, n = noalloc(x, mask, 0, 0.5, stream)
    want, wn = G.select(x, mask, 0, 0.5)
    assert np.isnan(got[0]) and n[0] == 0
    assert np.array_equal(got[1:], want[1:]) and np.array_equal(n, wn)


def test_noalloc_select_refuses_batches_beyond_the_scratch():
    from eda import _native

    torch = G.torch_cuda()
    x = torch.zeros((2000, 4), dtype=torch.float32, device="cuda")
    out = torch.zeros(2000, dtype=torch.float64, device="cuda")
    lib = _native.core()
    err = _native.errbuf()
    rc = lib.eda_prim_select_noalloc(0, 1, 2000, G.shape_arr((4,)), G.ptr(x), None, 1, 0, 0.5, G.ptr(out), None,
                                     ctypes.c_uint64(0), err, len(err))
    assert rc == _native.EDA_EINVAL and b"scratch" in err.value
