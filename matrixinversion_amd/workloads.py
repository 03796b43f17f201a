"""Seeded synthetic inputs of the benchmark workloads.

All entries come from std::mt19937(seed) (numpy's MT19937 with legacy seeding
is the same generator) with a pinned float mapping a = 2*((x >> 8)*2^-24) - 1,
so they are reproducible bit for bit on any host.  Fill order is matrix-major,
row-major: the reference's own inputs are row-major n*n blocks
(parallel_pivot/luBatchedInplace.cu:44-53).
"""
from __future__ import annotations

import numpy as np


def mt19937_raw(seed: int, count: int) -> np.ndarray:
    """The raw 32-bit output stream of std::mt19937(seed)."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return bg.random_raw(count).astype(np.uint32)


def mt19937_uniform(seed: int, count: int) -> np.ndarray:
    """`count` FP32 values U[-1,1) from mt19937(seed), pinned mapping."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    out = np.empty(count, dtype=np.float32)
    step = 1 << 24
    for s in range(0, count, step):
        raw = bg.random_raw(min(step, count - s)).astype(np.uint32)
        out[s:s + raw.size] = (raw >> 8).astype(np.float32) * np.float32(2.0 ** -23) - np.float32(1.0)
    return out


def uniform_matrices(seed: int, batch: int, n: int) -> np.ndarray:
    return mt19937_uniform(seed, batch * n * n).reshape(batch, n, n)


def dominant_matrices(seed: int, batch: int, n: int) -> np.ndarray:
    """U[-1,1) + n*I: diagonally dominant, safe without pivoting."""
    a = uniform_matrices(seed, batch, n)
    a[:, np.arange(n), np.arange(n)] += np.float32(n)
    return a


def make_inputs(kind: str, n: int, batch: int, seed: int) -> np.ndarray:
    if kind == "random":
        return uniform_matrices(seed, batch, n)
    if kind == "dominant":
        return dominant_matrices(seed, batch, n)
    if kind.startswith("template:"):
        t = np.loadtxt(kind.split(":", 1)[1], dtype=np.float32).ravel()[: n * n].reshape(n, n)
        return np.broadcast_to(t, (batch, n, n)).copy()
    raise ValueError(f"unknown input kind {kind!r}")


def ragged_sizes(seed: int, batch: int, n_lo: int = 2, n_hi: int = 32) -> np.ndarray:
    """n_i = n_lo + (x_i mod (n_hi - n_lo + 1)) over mt19937(seed)'s raw stream."""
    raw = mt19937_raw(seed, batch)
    return (n_lo + raw % np.uint32(n_hi - n_lo + 1)).astype(np.int32)


def ragged_batch(seed: int, batch: int, n_lo: int = 2, n_hi: int = 32, dominant: bool = False):
    """A packed ragged batch: sizes from mt19937(seed + 1), entries U[-1,1)
    from mt19937(seed) in matrix order (+ n*I per matrix when `dominant`).
    Returns (flat A, offsets int64, ns int32, pivot offsets int64); matrix i
    occupies A[offsets[i] : offsets[i] + ns[i]**2] (row-major)."""
    ns = ragged_sizes(seed + 1, batch, n_lo, n_hi)
    sq = ns.astype(np.int64) ** 2
    offsets = np.zeros(batch, np.int64)
    np.cumsum(sq[:-1], out=offsets[1:])
    poffs = np.zeros(batch, np.int64)
    np.cumsum(ns[:-1].astype(np.int64), out=poffs[1:])
    flat = mt19937_uniform(seed, int(sq.sum()))
    if dominant and batch:
        n64 = ns.astype(np.int64)
        k = np.arange(int(n64.sum())) - np.repeat(poffs, n64)       # 0..n_i-1 within matrix i
        diag = np.repeat(offsets, n64) + k * (np.repeat(n64, n64) + 1)
        flat[diag] += np.repeat(ns, n64).astype(np.float32)
    return flat, offsets, ns, poffs


def ragged_bytes(ns: np.ndarray, pivots: bool = True) -> int:
    """Algorithmic HBM bytes of one ragged launch: each matrix read and written
    once (8 n^2), its pivots written (4 n), and its offset (8 B), size (4 B)
    and pivot offset (8 B) read once."""
    n = ns.astype(np.int64)
    per = 8 * n * n + 8 + 4
    if pivots:
        per = per + 4 * n + 8
    return int(per.sum())


def conditioned_matrices(seed: int, batch: int, n: int, k: float) -> np.ndarray:
    """A = Q1 diag(sigma) Q2^T in FP64, rounded to FP32: Q1, Q2 from the QR
    factorisation of seeded Gaussian matrices (RandomState(seed)), sigma
    log-spaced from 1 to 10^k (2-norm condition number 10^k before rounding)."""
    rng = np.random.RandomState(seed)
    q1, _ = np.linalg.qr(rng.standard_normal((batch, n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((batch, n, n)))
    sigma = np.logspace(0.0, k, n)
    a = np.einsum("bij,j,bkj->bik", q1, sigma, q2)
    return a.astype(np.float32)
