# SPDX-License-Identifier: Apache-2.0
"""The reference's parity manifest (tests/data/parity_v2.json; decisions in
tests/golden/parity_decisions.json) on the CPU: 287 cases over 147 recipes, 1-D to 3-D, float32 and
float64, batches and layouts, chunking, NaN / inf with and without check_finite, gains from 1e-300
to 1e300, offsets, ties and quantized inputs, max_imf.

- Every input regenerates bit for bit (tests/golden/parity_inputs.py, checked against input_sha256),
  and the committed inputs (tests/golden/parity_inputs.npz) are those bits.
- The oracle's decisions (n_imfs, error, windows, sifts) equal the manifest's in every case and
  both dtypes: the float64 oracle end to end, and the float32 oracle (T = float), whose decisions
  no other fixture pins. The two 64^3 float32 volumes (about 40 s of oracle time) are left to the
  GPU test of the same manifest.
"""

from __future__ import annotations

import sys

import pytest

import oracle
from parity_cases import CASES, GOLDEN, MANIFEST, batch_layout, decisions_equal, load_input

SLOW_ON_CPU = {"d3-gabor-64x64x64.f32", "d3-mixture-64x64x64.f32"}


def generator():
    """The restated input generator (tests/golden/parity_inputs.py; CPU tests only)."""
    if str(GOLDEN) not in sys.path:
        sys.path.insert(0, str(GOLDEN))
    import parity_inputs  # noqa: PLC0415

    return parity_inputs


def test_the_manifest_covers_every_recipe():
    assert MANIFEST["count"] == len(CASES) == 287
    assert {c["recipe"] for c in CASES} == set(generator().BY_NAME)


def test_every_input_regenerates_bitwise():
    pi = generator()
    bad = [c["id"] for c in CASES if pi.sha256(pi.build(c["recipe"], c["dtype"])) != c["input_sha256"]]
    assert not bad, bad


def test_the_committed_inputs_are_the_manifest_inputs():
    pi = generator()
    bad = [c["id"] for c in CASES if pi.sha256(load_input(c["id"])) != c["input_sha256"]]
    assert not bad, bad


@pytest.mark.parametrize("case", [c for c in CASES if c["id"] not in SLOW_ON_CPU], ids=lambda c: c["id"])
def test_oracle_decisions_match_the_manifest(case):
    x = load_input(case["id"])
    xb, axes = batch_layout(x, case["axes"])
    _, info = oracle.decompose(xb, dim=len(axes), max_imf=case["max_imf"], tie_tolerance=case["tie_tolerance"],
                               check_finite=case["check_finite"])
    assert not decisions_equal(case, info["n_imfs"], info["error"], info["windows"], info["sifts"])
