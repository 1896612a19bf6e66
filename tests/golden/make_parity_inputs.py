# SPDX-License-Identifier: Apache-2.0
"""Writes tests/golden/parity_inputs.npz: the input array of every case of the reference's parity
manifest (tests/golden/parity_decisions.json), keyed by case id, regenerated by parity_inputs.py
(the restated input algorithm of the reference's tests/python/cases.py) and checked against the
manifest's input_sha256 before it is written. The GPU tests read only this file.

Usage: python tests/golden/make_parity_inputs.py
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import parity_inputs  # noqa: E402


def main() -> int:
    manifest = json.loads((HERE / "parity_decisions.json").read_text())
    arrays = {}
    for c in manifest["cases"]:
        x = parity_inputs.build(c["recipe"], c["dtype"])
        if parity_inputs.sha256(x) != c["input_sha256"]:
            raise SystemExit(f"{c['id']}: the regenerated input does not match input_sha256")
        arrays[c["id"]] = x
    np.savez_compressed(HERE / "parity_inputs.npz", **arrays)
    total = sum(a.nbytes for a in arrays.values())
    print(f"wrote {len(arrays)} inputs ({total / 2**20:.1f} MiB) to {HERE / 'parity_inputs.npz'}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
