"""Writes tests/golden/pflog_alpha_fixtures.npz: the data of the reference's
pinned fixtures for PFlog's overdispersion estimator
(scx-accel/src/pflog_reference_values.rs), read from that file as text.

Each fixture is a raw-count matrix simulated from a known negative-binomial
dispersion, plus the constants the reference's tests assert against
(scx-accel/src/pflog_reference_tests.rs):

* A (50 x 120, alpha_true 0.02): the pool holds every gene with a mean; the
  estimate is A_SEPARATION times closer to alpha_true than the pre-fix
  (truncated) estimator's answer; within A_REL_BAR of alpha_true.
* B (40 x 40, alpha_true 0.5): no fallback, every gene pooled, within
  B_REL_BAR of alpha_true.
* C (30 x 24): 20 under-dispersed genes + 4 over-dispersed ones; the
  estimator falls back (alpha 0.25) with all 24 genes pooled.

Stored per fixture: ``<X>_counts`` (uint16, cells x genes) and every scalar
constant of the fixture under its reference name. Only data is kept: the
counts and the expected values.

Run: python tests/golden/make_pflog_alpha_fixtures.py [path to the reference checkout]
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC_REL = "scx-accel/src/pflog_reference_values.rs"


def parse(text: str) -> dict:
    out = {}
    for name, ty, val in re.findall(r"pub const ([A-Z_]+): (f64|usize) = ([0-9.eE+-]+);", text):
        out[name] = float(val) if ty == "f64" else int(val)
    for fx in ("A", "B", "C"):
        m = re.search(rf"const {fx}_COUNTS: \[\[u16; (\d+)\]; (\d+)\] = \[(.*?)\];", text, re.S)
        if m is None:
            raise SystemExit(f"fixture {fx}: counts array not found")
        n_genes, n_cells = int(m.group(1)), int(m.group(2))
        rows = re.findall(r"\[([^\[\]]*)\]", m.group(3))
        mat = np.array([[int(v) for v in r.replace("\n", " ").split(",") if v.strip()] for r in rows],
                       dtype=np.uint16)
        if mat.shape != (n_cells, n_genes):
            raise SystemExit(f"fixture {fx}: parsed {mat.shape}, declared {(n_cells, n_genes)}")
        if (out[f"{fx}_N_CELLS"], out[f"{fx}_N_GENES"]) != (n_cells, n_genes):
            raise SystemExit(f"fixture {fx}: shape constants disagree with the array")
        out[f"{fx}_counts"] = mat
    return out


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, SRC_REL)) as f:
        data = parse(f.read())
    dst = os.path.join(HERE, "pflog_alpha_fixtures.npz")
    np.savez_compressed(dst, **{k: np.asarray(v) for k, v in data.items()})
    print("wrote", dst, sorted(data))


if __name__ == "__main__":
    main()
