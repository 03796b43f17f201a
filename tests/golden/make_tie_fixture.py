#!/usr/bin/env python3
"""Write tests/golden/seed42_tie_matrices.npz: the matrices of the 1M x n=32
seed-42 U[-1,1) workload whose pivots differ between the first-max rule and
parallel_pivot's tree rule (an exact |a| tie at a column maximum, which the
two rules break differently).  Keys: "index" (int64, position in the
workload), "a" (float32, k x 32 x 32).

Inputs come from matrixinversion_amd.workloads.uniform_matrices(42, 1000000,
32); the selection uses the oracle's two pivot rules.  The tests recompute
the oracle's pivots themselves, so the file holds inputs only.

Usage:  python tests/golden/make_tie_fixture.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from matrixinversion_amd import workloads  # noqa: E402


def main() -> None:
    a = workloads.uniform_matrices(42, 1_000_000, 32)
    fm = oracle.batch_pivots(a, oracle.PIV_FIRSTMAX)
    tr = oracle.batch_pivots(a, oracle.PIV_TREE)
    idx = np.nonzero((fm != tr).any(axis=1))[0].astype(np.int64)
    out = os.path.join(HERE, "seed42_tie_matrices.npz")
    np.savez(out, index=idx, a=np.ascontiguousarray(a[idx]))
    print(f"wrote {out}: {idx.size} matrices, indices {idx.tolist()}")


if __name__ == "__main__":
    main()
