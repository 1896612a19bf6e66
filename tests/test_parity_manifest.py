# SPDX-License-Identifier: Apache-2.0
"""The reference's parity manifest (tests/data/parity_v2.json; decisions in
tests/golden/parity_decisions.json) on the CPU: 287 cases over 147 recipes, 1-D to 3-D, float32 and
float64, batches and layouts, chunking, NaN / inf with and without check_finite, gains from 1e-300
to 1e300, offsets, ties and quantized inputs, max_imf.

- Every input regenerates bit for bit (tests/parity_inputs.py, checked against input_sha256).
- The oracle's decisions (n_imfs, error, windows, sifts) equal the manifest's in every case and
  both dtypes: the float64 oracle end to end, and the float32 oracle (T = float), whose decisions
  no other fixture pins. The two 64^3 float32 volumes (about 40 s of oracle time) are left to the
  GPU test of the same manifest.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
import parity_inputs

MANIFEST = json.loads((Path(__file__).resolve().parent / "golden" / "parity_decisions.json").read_text())
CASES = MANIFEST["cases"]
SLOW_ON_CPU = {"d3-gabor-64x64x64.f32", "d3-mixture-64x64x64.f32"}


def batch_layout(x: np.ndarray, axes) -> tuple[np.ndarray, tuple[int, ...]]:
    """x as (B, *spatial) with the spatial axes in the order of axes and the batch in C order."""
    axes = tuple(axes) if axes is not None else tuple(range(x.ndim))
    batch = [i for i in range(x.ndim) if i not in axes]
    spatial = tuple(x.shape[a] for a in axes)
    return np.ascontiguousarray(x.transpose(batch + list(axes))).reshape((-1,) + spatial), axes


def decisions_equal(c: dict, n_imfs, error, windows, sifts) -> list[str]:
    bad = []
    if list(np.asarray(n_imfs).reshape(-1)) != c["n_imfs"]:
        bad.append("n_imfs")
    if list(np.asarray(error).reshape(-1)) != c["error"]:
        bad.append("error")
    if c["windows"] is not None and list(np.asarray(windows).reshape(-1)) != c["windows"]:
        bad.append("windows")
    if c["sifts"] is not None and list(np.asarray(sifts).reshape(-1)) != c["sifts"]:
        bad.append("sifts")
    return bad


def test_the_manifest_covers_every_recipe():
    assert MANIFEST["count"] == len(CASES) == 287
    assert {c["recipe"] for c in CASES} == set(parity_inputs.BY_NAME)


def test_every_input_regenerates_bitwise():
    bad = [c["id"] for c in CASES
           if parity_inputs.sha256(parity_inputs.build(c["recipe"], c["dtype"])) != c["input_sha256"]]
    assert not bad, bad


@pytest.mark.parametrize("case", [c for c in CASES if c["id"] not in SLOW_ON_CPU], ids=lambda c: c["id"])
def test_oracle_decisions_match_the_manifest(case):
    x = parity_inputs.build(case["recipe"], case["dtype"])
    xb, axes = batch_layout(x, case["axes"])
    _, info = oracle.decompose(xb, dim=len(axes), max_imf=case["max_imf"], tie_tolerance=case["tie_tolerance"],
                               check_finite=case["check_finite"])
    assert not decisions_equal(case, info["n_imfs"], info["error"], info["windows"], info["sifts"])
