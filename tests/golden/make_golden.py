#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ (run in the build container).

Inputs are the reference's own data and host code:
  * the matrix fixtures the reference's drivers read (mtrand32_new1.txt, read by
    every variant, e.g. parallel_pivot/luBatchedInplace.cu:44-53; mtrand32_new.txt
    and mtrand32.txt, the commented alternatives at :41-43) are copied verbatim;
  * the hand-written matrices of parallel_pivot/luBatchedInplace.cu:20-25 get
    exact rational inverses (Gauss-Jordan over fractions.Fraction);
  * pivot vectors, verifyInv counts, verifyLUwithPivoting counts and
    calc_cond_num values come from the reference's verify.hpp, compiled from the
    reference sources by oracle/Makefile (target `ref`) and called through
    oracle/_ref/libref_verify.so.

Usage:  python tests/golden/make_golden.py  (requires /root/reference)
"""
from __future__ import annotations

import json
import os
import shutil
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

REF = "/root/reference"
FIXTURES = ["mtrand32_new1.txt", "mtrand32_new.txt", "mtrand32.txt"]

# parallel_pivot/luBatchedInplace.cu:20-25 (the 2x2, 3x3, two 4x4 and 10*I_4 inputs)
HAND = {
    "m2": (2, [2, 3, 4, 5]),
    "m3": (3, [4, 11, 3, 40, 10, 4, 2, 40, 2]),
    "m4a": (4, [4, 11, 3, 7, 4, 10, 4, 9, 2, 4, 2, 1, 7, 9, 18, 30]),
    "m4b": (4, [2, 7, 1, 5, 3, -2, 0, 1, 1, 5, 3, 4, 7, 3, 2, 8]),
    "ten_eye4": (4, [10, 0, 0, 0, 0, 10, 0, 0, 0, 0, 10, 0, 0, 0, 0, 10]),
}


def exact_inverse(n, vals):
    m = [[Fraction(vals[i * n + j]) for j in range(n)] + [Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if m[r][c] != 0)
        m[c], m[p] = m[p], m[c]
        piv = m[c][c]
        m[c] = [v / piv for v in m[c]]
        for r in range(n):
            if r != c and m[r][c] != 0:
                f = m[r][c]
                m[r] = [a - f * b for a, b in zip(m[r], m[c])]
    return [[str(m[i][n + j]) for j in range(n)] for i in range(n)]


def load_fixture(name):
    return np.loadtxt(os.path.join(HERE, name), dtype=np.float32).ravel()


def main():
    if not os.path.isdir(REF):
        sys.exit("the reference sources are needed to regenerate the fixtures")
    oracle.build(ref=True)
    for f in FIXTURES:
        shutil.copyfile(os.path.join(REF, "parallel_pivot", f), os.path.join(HERE, f))

    kat = {k: {"n": n, "a": v, "inverse": exact_inverse(n, v)} for k, (n, v) in HAND.items()}
    with open(os.path.join(HERE, "kat_exact_inverses.json"), "w") as fh:
        json.dump(kat, fh, indent=1)

    # pivotedA (first maximum |a| on the un-updated column) and calc_cond_num for
    # every n x n prefix of each fixture, as the drivers read it (first n*n values).
    piv = {}
    cond = {}
    for f in FIXTURES:
        t = load_fixture(f)
        piv[f] = {}
        cond[f] = {}
        for n in range(1, 33):
            a = t[: n * n].reshape(n, n)
            p, _ = oracle.ref_pivotedA(a)
            piv[f][str(n)] = p.tolist()
            cond[f][str(n)] = float(oracle.ref_calc_cond_num(a))
    rng = np.random.RandomState(20240917)
    rand_cases = []
    for i in range(64):
        n = int(rng.randint(1, 33))
        if i % 4 == 0:   # integer entries: many exact |a| ties
            a = rng.randint(-4, 5, size=(n, n)).astype(np.float32)
        else:
            a = rng.uniform(-1, 1, size=(n, n)).astype(np.float32)
        p, _ = oracle.ref_pivotedA(a)
        rand_cases.append({"n": n, "a": a.ravel().tolist(), "pivots": p.tolist()})
    with open(os.path.join(HERE, "pivots_pivotedA.json"), "w") as fh:
        json.dump({"fixture_prefixes": piv, "random": rand_cases}, fh)
    with open(os.path.join(HERE, "calc_cond_num.json"), "w") as fh:
        json.dump(cond, fh, indent=1)

    # verifyInv counts on crafted inverses straddling its 1e-3 threshold.
    t = load_fixture("mtrand32_new1.txt")
    cases = []
    for n in (4, 8, 16):
        a = np.stack([t[: n * n].reshape(n, n)] * 6)
        x64 = np.linalg.inv(a[0].astype(np.float64))
        x = np.stack([x64] * 6).astype(np.float32)
        # perturb X[0][0] so that the largest entry change in column 0 of A X is d
        col = float(np.abs(a[0, :, 0]).max())
        for i, d in enumerate([0.0, 5e-4, 2e-3, -2e-3, 1e-2, 0.0]):
            x[i, 0, 0] += np.float32(d / col)
        x[5, :, :] = np.nan
        correct, wrong = oracle.ref_verify_inv(a, x)
        cases.append({"n": n, "a": a.ravel().tolist(), "x": x.ravel().tolist(),
                      "correct": correct, "wrong": wrong})
    with open(os.path.join(HERE, "verifyInv_counts.json"), "w") as fh:
        json.dump(cases, fh)

    # verifyLUwithPivoting on the oracle's Doolittle factors of P*A.
    lu_cases = []
    for f in FIXTURES[:2]:
        t = load_fixture(f)
        for n in (4, 8, 12, 16, 32):
            a = t[: n * n].reshape(n, n)
            p, pa = oracle.ref_pivotedA(a)
            lu, _ = oracle.lu_factor_f32(a, oracle.PIV_FIRSTMAX)
            lu_cases.append({"fixture": f, "n": n, "correct": oracle.ref_verify_lu(pa, lu[None])})
    with open(os.path.join(HERE, "verifyLU_counts.json"), "w") as fh:
        json.dump(lu_cases, fh, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
