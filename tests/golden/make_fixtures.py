# SPDX-License-Identifier: Apache-2.0
"""Writes the known-answer fixtures of tests/golden from the reference's own test data (read as
text and data only; nothing of the reference is imported or run):

- rungs_table.json: the 66-row KAT of the element's rung ladder, cpp/tests/rungs_table.inc
  (d_min, d_max as float.hex, w, n_max, the rungs);
- quantile_table.json: the NumPy quantile table of cpp/tests/test_stats.cu:282-300
  (numpy.quantile(v, q) for v = [i * i + 0.5 for i in range(n)]);
- corpus/: a copy of tests/data/golden (9 FP64 cases: inputs, modes, windows, sifts, sensitivity,
  and its index.json);
- parity_decisions.json: the decision part of tests/data/parity_v2.json (287 cases): per case its
  recipe, dtype, call arguments, input_sha256, y_shape, n_imfs, error, windows and sifts (values,
  or null where the manifest stores only a hash), chunks, conditioning class and quick flag. The
  manifest's row hashes are CUDA bits, which no other build reproduces, and are left out.

Run in a container that mounts the reference at /root/reference:
    python tests/golden/make_fixtures.py
"""

from __future__ import annotations

import json
import re
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent


def rungs_table() -> list[dict]:
    text = (REF / "cpp/tests/rungs_table.inc").read_text()
    rows = []
    for m in re.finditer(r"\{(0x[0-9a-fp.+-]+),\s*(0x[0-9a-fp.+-]+),\s*(\d+),\s*(\d+),\s*\{([0-9,\s]+)\}\}", text):
        rows.append({
            "d_min": m.group(1),
            "d_max": m.group(2),
            "w": int(m.group(3)),
            "n_max": int(m.group(4)),
            "rungs": [int(v) for v in m.group(5).split(",") if v.strip()],
        })
    return rows


def quantile_table() -> dict:
    text = (REF / "cpp/tests/test_stats.cu").read_text()
    block = text[text.index("void quantile_table()"):]
    qs = [float(v) for v in re.search(r"const double qs\[7\] = \{([^}]*)\}", block).group(1).split(",")]
    rows = []
    for m in re.finditer(r"\{(\d+), \{([^}]*)\}\}", block[: block.index("const eda::Geom")]):
        rows.append({"n": int(m.group(1)), "values": [float(v) for v in m.group(2).split(",")]})
    return {"q": qs, "rows": rows}


PARITY_KEYS = ("id", "recipe", "dtype", "quick", "shape", "axes", "max_imf", "tie_tolerance", "check_finite",
               "errors", "max_chunk", "input_sha256", "y_shape", "n_imfs", "error", "info_shape", "windows",
               "sifts", "chunks", "conditioning_class")


def parity_decisions() -> dict:
    src = json.loads((REF / "tests/data/parity_v2.json").read_text())
    cases = []
    for c in src["cases"]:
        d = {k: c.get(k) for k in PARITY_KEYS}
        for k in ("windows", "sifts"):
            if not isinstance(d[k], list):
                d[k] = None
        cases.append(d)
    return {"format": "parity_v2 decisions", "source": "tests/data/parity_v2.json of the reference",
            "count": len(cases), "cases": cases}


def main() -> None:
    rt = rungs_table()
    assert len(rt) == 66, len(rt)
    (HERE / "rungs_table.json").write_text(json.dumps(rt, indent=1) + "\n")
    qt = quantile_table()
    assert len(qt["rows"]) == 7
    (HERE / "quantile_table.json").write_text(json.dumps(qt, indent=1) + "\n")
    dst = HERE / "corpus"
    dst.mkdir(exist_ok=True)
    for f in sorted((REF / "tests/data/golden").iterdir()):
        shutil.copy(f, dst / f.name)
    pd = parity_decisions()
    assert pd["count"] == 287
    (HERE / "parity_decisions.json").write_text(json.dumps(pd) + "\n")
    print("wrote", len(rt), "ladder rows,", len(qt["rows"]), "quantile rows, the golden corpus and",
          pd["count"], "parity cases")


if __name__ == "__main__":
    main()
