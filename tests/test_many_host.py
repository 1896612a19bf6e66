# SPDX-License-Identifier: Apache-2.0
"""eda.decompose_many's host logic on the CPU (no GPU): the batched call is replaced by a stand-in
that records its batches, so that the grouping by shape / dtype, the order of the results, the
trimming of each field's rows to its own count and the error policy are checked without the
library."""

from __future__ import annotations

import threading

import numpy as np
import pytest

import eda
from eda import _many
from eda._api import DecompositionInfo


class FakeBatch:
    """Stands in for eda.decompose on a stacked batch: field b gets n_imfs = its length mod 4, the
    rows are the field times (1 + row index), a NaN in a field fails it with code 1."""

    def __init__(self):
        self.calls = []
        self.lock = threading.Lock()

    def __call__(self, batch, *, axes, errors, return_info, **kw):
        assert errors == "status" and return_info and axes == tuple(range(1, batch.ndim))
        with self.lock:
            self.calls.append((batch.shape, batch.dtype))
        B = batch.shape[0]
        bad = np.array([np.isnan(batch[b]).any() for b in range(B)])
        k = np.array([batch[b].shape[-1] % 4 for b in range(B)], np.int32)
        K = int(k.max())
        P = max([int(k[b]) + 1 for b in range(B) if not bad[b]] + [1])
        y = np.zeros((B, P) + batch.shape[1:], batch.dtype)
        for b in range(B):
            for p in range(P):
                y[b, p] = np.nan if bad[b] else (batch[b] * (p + 1) if p <= k[b] else 0)
        win = np.zeros((B, K), np.int32)
        for b in range(B):
            win[b, : k[b]] = np.arange(1, k[b] + 1) * 2 + 1
        info = DecompositionInfo(n_imfs=k, error=bad.astype(np.int32), windows=win, spacing=win * 0.5,
                                 sifts=win * 0 + 3, energy=win * 1.0, dtype=str(batch.dtype), axes=axes)
        return y, info


@pytest.fixture
def fake(monkeypatch):
    f = FakeBatch()
    monkeypatch.setattr(_many, "decompose", f)
    return f


def fields():
    rng = np.random.default_rng(0)
    return [rng.standard_normal(n).astype(np.float32) for n in (10, 11, 10, 13, 11, 10)] + [
        rng.standard_normal((6, 7)).astype(np.float32), rng.standard_normal(10)]


@pytest.mark.parametrize("workers", [1, 3])
def test_grouping_order_and_trimming(fake, workers):
    fs = fields()
    out = eda.decompose_many(fs, return_info=True, workers=workers)
    # one batch per (shape, dtype): float32 lengths 10, 11, 13, the image, and the float64 signal
    assert sorted((s, str(d)) for s, d in fake.calls) == sorted(
        [((3, 10), "float32"), ((2, 11), "float32"), ((1, 13), "float32"), ((1, 6, 7), "float32"),
         ((1, 10), "float64")])
    for f, (rows, info) in zip(fs, out):
        k = f.shape[-1] % 4
        assert rows.shape == (k + 1,) + f.shape
        for p in range(k + 1):
            assert np.array_equal(rows[p], f * (p + 1))
        assert int(info.n_imfs) == k and info.windows.shape == (k,) and info.axes == tuple(range(f.ndim))


def test_errors_by_position(fake):
    fs = fields()
    fs[3] = fs[3].copy()
    fs[3][0] = np.nan
    with pytest.raises(eda.DecompositionError) as ei:
        eda.decompose_many(fs)
    assert list(ei.value.indices) == [3] and list(ei.value.codes) == [1]
    out = eda.decompose_many(fs, errors="status", return_info=True)
    assert out[3][0].shape == (1, 13) and int(out[3][1].error) == 1


def test_arguments_are_checked(fake):
    with pytest.raises(ValueError):
        eda.decompose_many(fields(), errors="ignore")
    with pytest.raises(ValueError):
        eda.decompose_many([np.zeros((2, 2, 2, 2), np.float32)])
    assert eda.decompose_many([]) == []
