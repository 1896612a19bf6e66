# SPDX-License-Identifier: Apache-2.0
"""The ragged batch (eda.decompose_many) on the GPU, on a reduced batching workload: 24 chirp
signals of lengths log-uniform in [1024, 262144] (two of each length drawn again, so that groups
hold several fields) and 6 images of sides 128 / 256 / 384.

- Every field's rows and information are bitwise those of eda.decompose of that field alone, host
  and device input alike, with 1 and with 4 worker threads (streams).
- Float64 fields of the workload against the FP64 oracle: decisions exact, rows within 1e-11
  relative L2.
- errors="raise" names the failed fields by their position; errors="status" returns their codes.
"""

from __future__ import annotations

import numpy as np
import pytest

import eda
import oracle
import workloads

pytestmark = pytest.mark.gpu


def fields(dtype=np.float32):
    out = []
    for i in range(12):
        s = workloads.chirp_signal(i, dtype)
        out += [s, workloads.chirp_signal(i + 500, dtype, n=s.size)]
    for j, side in enumerate((128, 256, 384, 128, 256, 384)):
        out.append(workloads.rings_stripes_blobs(1, side, dtype, first=20000 + j)[0])
    return out


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return np.array_equal(np.atleast_1d(a).view(np.uint8), np.atleast_1d(b).view(np.uint8))


@pytest.mark.parametrize("workers", [1, 4])
def test_each_field_is_its_own_decomposition(workers):
    fs = fields()
    print(f"batching: {len(fs)} fields, workers={workers}", flush=True)
    many = eda.decompose_many(fs, return_info=True, workers=workers)
    for i, f in enumerate(fs):
        y, info = eda.decompose(f, return_info=True)
        rows, inf = many[i]
        assert same(rows, y), i
        for key in ("n_imfs", "error", "windows", "spacing", "sifts", "energy"):
            assert same(getattr(inf, key), getattr(info, key)), (i, key)


def test_device_fields():
    import torch

    fs = fields()[::3]
    dev = [torch.from_numpy(f).to("cuda") for f in fs]
    many = eda.decompose_many(dev, workers=3)
    for i, f in enumerate(dev):
        assert torch.equal(many[i], eda.decompose(f)), i


def test_float64_fields_against_the_oracle():
    fs = [workloads.chirp_signal(i, np.float64, n=n) for i, n in ((3, 1500), (4, 3000), (5, 3000))]
    fs.append(workloads.rings_stripes_blobs(1, 96, np.float64, first=7)[0])
    many = eda.decompose_many(fs, return_info=True)
    for f, (rows, info) in zip(fs, many):
        ref, ri = oracle.decompose(f[None], dim=f.ndim)
        assert int(info.n_imfs) == int(ri["n_imfs"][0])
        assert info.windows.tolist() == ri["windows"][0].tolist()
        assert info.sifts.tolist() == ri["sifts"][0].tolist()
        for k in range(rows.shape[0]):
            err = np.linalg.norm(rows[k] - ref[0, k]) / np.linalg.norm(ref[0, k])
            assert err <= 1e-11, (f.shape, k, err)


def test_failed_fields_by_position():
    fs = fields()[:6]
    fs[4] = fs[4].copy()
    fs[4][10] = np.nan
    with pytest.raises(eda.DecompositionError) as ei:
        eda.decompose_many(fs)
    assert list(ei.value.indices) == [4]
    out = eda.decompose_many(fs, errors="status", return_info=True)
    assert int(out[4][1].error) == 1 and np.isnan(out[4][0]).all() and out[4][0].shape[0] == 1
    assert all(int(out[i][1].error) == 0 for i in (0, 1, 2, 3, 5))
