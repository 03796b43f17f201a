"""Accuracy helpers shared by the parity tests (CPU side, numpy + oracle).

Error measure, per matrix: e(X) = max_ij |X - X64| / max_ij |X64|, with X64
the FP64 inverse of the FP32 input (oracle.inverse_f64).  The contract stated
by the library for FP32 results is

    e(X_gpu) <= C_REL * n * kappa_1(A) * 2^-24        (well-conditioned, any pivoting)

and, for every input (the REF pivot rule is not GEPP, so growth can push
both the reference's arithmetic and ours past the first bound), that the GPU
stays within a small factor of the larger of that bound and the error of the
FP32 restatement of the reference's own arithmetic:

    e(X_gpu) <= C_ORC * max(e(X_oracle), n * kappa_1(A) * 2^-24)
"""
from __future__ import annotations

import numpy as np

import oracle

EPS32 = 2.0 ** -24
C_REL = 1.0      # multiplier of n * kappa_1 * 2^-24 (diagonally dominant inputs)
C_ORC = 8.0      # GPU vs max(oracle error, n * kappa_1 * 2^-24)


def rel_err64(a: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Per-matrix e(X) against the FP64 inverse."""
    a = a.reshape(-1, a.shape[-1], a.shape[-1])
    x = x.reshape(a.shape)
    out = np.empty(a.shape[0])
    for i in range(a.shape[0]):
        x64, _, info = oracle.inverse_f64(a[i])
        if info != 0:
            out[i] = np.inf
            continue
        num = np.abs(x[i].astype(np.float64) - x64)
        num = np.where(np.isnan(num), np.inf, num).max()
        out[i] = num / np.abs(x64).max()
    return out


def kappa1(a: np.ndarray) -> np.ndarray:
    a = a.reshape(-1, a.shape[-1], a.shape[-1])
    return np.array([oracle.cond1(m) for m in a])


def max_abs_axmi(a: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Per-matrix max |A X - I| in FP64 (the verify.hpp residual, as a number)."""
    a = a.reshape(-1, a.shape[-1], a.shape[-1]).astype(np.float64)
    x = x.reshape(a.shape).astype(np.float64)
    r = a @ x - np.eye(a.shape[-1])
    r = np.where(np.isnan(r), np.inf, np.abs(r))
    return r.reshape(r.shape[0], -1).max(axis=1)


def contract_bound(a: np.ndarray, e_oracle: np.ndarray) -> np.ndarray:
    """Per-matrix C_ORC * max(e_oracle, n * kappa_1 * 2^-24)."""
    n = a.shape[-1]
    return C_ORC * np.maximum(e_oracle, n * kappa1(a) * EPS32)
