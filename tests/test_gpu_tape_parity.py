# SPDX-License-Identifier: Apache-2.0
"""Per-mode parity of the stages between the method's decisions and the rows (GPU): the tape of
every mode (eda.TapedDecomposition.mode_tape) against the oracle's tape of the same input.

Per mode: the windows w and W (G1, and G2 / G3 where the fine path runs: the golden g3blobs image),
the rung ladder with its used and segment-point bits (the element's rungs), the sifts run and
accepted, exactly; the sift amplitude floors thr_j and the tie tolerance, the scale field L (the
raw window, absent field, components and floor stages), and the residual r at the mode's start
(the previous mode's element gathers, envelope mean and sifts) within 1e-12 relative in float64
and 1e-5 in float32; the lower rung lo of every voxel (the hat weights) equal in float64 and at
99.9 % of the voxels in float32 (a scale landing on a rung can take either neighbour there, with
the same result). 1-D tones, 2-D rings / stripes / blobs, the golden g3blobs image and a 3-D
plane-wave volume."""

from __future__ import annotations

import numpy as np
import pytest

import eda
import oracle
import workloads

pytestmark = pytest.mark.gpu

CASES = ["tones", "rings96", "g3blobs", "vol24"]


def field(name):
    if name == "tones":
        return workloads.tones_2048(np.float64)
    if name == "rings96":
        return workloads.rings_stripes_blobs(1, 96, np.float64, first=3)[0]
    if name == "g3blobs":
        return np.load("tests/golden/corpus/g3blobs_2d_64.npz")["x"].astype(np.float64)
    return workloads.plane_wave_volumes(1, 24, np.float64)[0]


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("name", CASES)
def test_mode_tapes_match_the_oracle(name, dtype, tol):
    import torch

    x = field(name).astype(dtype)
    print(f"tape parity: {name} {np.dtype(dtype).name}", flush=True)
    td = eda.TapedDecomposition(torch.from_numpy(x[None]).to("cuda"))
    ref = oracle.decompose_tape(x)
    assert int(td.info["n_imfs"][0]) == ref["n_imfs"] >= 1
    for m, om in enumerate(ref["modes"]):
        t = td.mode_tape(0, m)
        for key in ("w", "W", "rung_w", "used", "pts", "ran", "accepted"):
            assert t[key] == om[key], (m, key, t[key], om[key])
        assert abs(t["tol"] - om["tol"]) <= tol * abs(om["tol"]), m
        assert all(abs(a - b) <= tol * abs(b) for a, b in zip(t["thr"], om["thr"])), m
        assert float(np.max(np.abs(t["L"].astype(np.float64) - om["L"]) / om["L"])) <= tol, m
        assert rel(t["r"], om["r"]) <= tol, m
        same_lo = float(np.mean(t["lo"] == om["lo"]))
        assert same_lo == 1.0 if dtype == np.float64 else same_lo >= 0.999, (m, same_lo)
