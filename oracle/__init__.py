"""CPU oracle for batched small-matrix LU inversion -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker.  The product path
(``matrixinversion_amd``) never imports it and has no CPU fallback.

The arithmetic lives in ``lu_oracle.c`` (plain C, built into ``liboracle.so``
by ``oracle/Makefile``); this module is a thin numpy/ctypes wrapper.  Each
function names the reference code it restates.  ``ref`` exposes the
reference's own host checker (``verify.hpp``) when ``oracle/_ref`` was built
from the reference sources in this container.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_verify.so")

PIV_NONE = 0
PIV_FIRSTMAX = 1
PIV_TREE = 3
PIV_GEPP = 4  # partial pivoting on the updated column (gepp_inverse_f32)

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)

_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and the reference shim when asked and possible)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference/parallel_pivot"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.orc_mt19937_raw.argtypes = [ctypes.c_uint32, _c_u32_p, ctypes.c_size_t]
        lib.orc_mt19937_uniform.argtypes = [ctypes.c_uint32, _c_float_p, ctypes.c_size_t]
        lib.orc_pivots.argtypes = [_c_float_p, ctypes.c_int, ctypes.c_int, _c_int_p]
        lib.orc_pivots.restype = ctypes.c_int
        lib.orc_lu_factor_f32.argtypes = [_c_float_p, _c_float_p, ctypes.c_int, ctypes.c_int, _c_int_p]
        lib.orc_lu_factor_f32.restype = ctypes.c_int
        lib.orc_lu_inverse_f32.argtypes = [_c_float_p, ctypes.c_int, ctypes.c_int, _c_int_p]
        lib.orc_lu_inverse_f32.restype = ctypes.c_int
        lib.orc_batch_inverse_f32.argtypes = [_c_float_p, _c_int_p, ctypes.c_int, ctypes.c_longlong,
                                              ctypes.c_int, ctypes.c_int]
        lib.orc_batch_inverse_f32.restype = ctypes.c_int
        lib.orc_openmp_threads.restype = ctypes.c_int
        lib.orc_inverse_f64.argtypes = [_c_float_p, _c_double_p, ctypes.c_int, _c_int_p]
        lib.orc_inverse_f64.restype = ctypes.c_int
        lib.orc_cond1.argtypes = [_c_float_p, ctypes.c_int]
        lib.orc_cond1.restype = ctypes.c_double
        lib.orc_verify_inv.argtypes = [_c_float_p, _c_float_p, ctypes.c_int, ctypes.c_longlong,
                                       ctypes.c_float, _c_double_p]
        lib.orc_verify_inv.restype = ctypes.c_longlong
        lib.orc_verify_inv_threads.argtypes = [_c_float_p, _c_float_p, ctypes.c_int, ctypes.c_longlong,
                                               ctypes.c_float, _c_double_p, ctypes.c_int]
        lib.orc_verify_inv_threads.restype = ctypes.c_longlong
        lib.orc_batch_gepp_f32.argtypes = [_c_float_p, _c_int_p, _c_int_p, ctypes.c_int, ctypes.c_longlong,
                                           ctypes.c_int]
        lib.orc_batch_gepp_f32.restype = ctypes.c_int
        _lib = lib
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_c_float_p)


# ------------------------------------------------------------------ inputs

def mt19937_raw(seed: int, count: int) -> np.ndarray:
    """Raw std::mt19937(seed) outputs."""
    out = np.empty(count, dtype=np.uint32)
    _load().orc_mt19937_raw(seed, out.ctypes.data_as(_c_u32_p), count)
    return out


def uniform_matrices(seed: int, batch: int, n: int) -> np.ndarray:
    """batch x n x n FP32 matrices, i.i.d. U[-1,1) from mt19937(seed), filled
    matrix-major then row-major with a = 2*((x>>8)*2^-24) - 1."""
    out = np.empty((batch, n, n), dtype=np.float32)
    _load().orc_mt19937_uniform(seed, out.ctypes.data_as(_c_float_p), out.size)
    return out


def dominant_matrices(seed: int, batch: int, n: int) -> np.ndarray:
    """A = U[-1,1) + n*I (diagonally dominant), the no-pivot sweep inputs."""
    a = uniform_matrices(seed, batch, n)
    idx = np.arange(n)
    a[:, idx, idx] += np.float32(n)
    return a


# ------------------------------------------------------------------ pivots

def pivots(a: np.ndarray, rule: int = PIV_FIRSTMAX) -> np.ndarray:
    """Pivot vector (perm[p] = original row at position p) of one matrix."""
    n = a.shape[-1]
    a, ap = _f32(a)
    perm = np.empty(n, dtype=np.int32)
    rc = _load().orc_pivots(ap, n, rule, perm.ctypes.data_as(_c_int_p))
    if rc != 0:
        raise ValueError(f"orc_pivots failed for n={n} rule={rule}")
    return perm


def batch_pivots(a: np.ndarray, rule: int = PIV_FIRSTMAX) -> np.ndarray:
    return np.stack([pivots(m, rule) for m in a.reshape(-1, a.shape[-1], a.shape[-1])])


# --------------------------------------------------------------- inversion

def lu_inverse_f32(a: np.ndarray, rule: int = PIV_FIRSTMAX, threads: int = 0):
    """FP32 restatement of the reference kernel over a batch (b, n, n).
    Returns (inverses, pivots)."""
    x = np.array(a, dtype=np.float32, copy=True, order="C")
    n = x.shape[-1]
    batch = x.size // (n * n)
    piv = np.empty((batch, n), dtype=np.int32)
    rc = _load().orc_batch_inverse_f32(x.ctypes.data_as(_c_float_p), piv.ctypes.data_as(_c_int_p),
                                       n, batch, rule, threads)
    if rc != 0:
        raise ValueError("orc_batch_inverse_f32 failed")
    return x, piv


def gepp_inverse_f32(a: np.ndarray, threads: int = 0):
    """FP32 Gauss-Jordan with partial pivoting on the updated column (the
    library's LUB_PIVOT_GEPP arithmetic; lu_oracle.c:orc_gepp_inverse_f32)
    over a batch (b, n, n).  Returns (inverses, pivots, info)."""
    x = np.array(a, dtype=np.float32, copy=True, order="C")
    n = x.shape[-1]
    batch = x.size // (n * n)
    piv = np.empty((batch, n), dtype=np.int32)
    info = np.empty(batch, dtype=np.int32)
    rc = _load().orc_batch_gepp_f32(x.ctypes.data_as(_c_float_p), piv.ctypes.data_as(_c_int_p),
                                    info.ctypes.data_as(_c_int_p), n, batch, threads)
    if rc != 0:
        raise ValueError("orc_batch_gepp_f32 failed")
    return x, piv, info


def lu_factor_f32(a: np.ndarray, rule: int = PIV_FIRSTMAX):
    """Doolittle LU of P*A for one matrix in the reference's order.
    Returns (LU, perm) with unit-lower L below the diagonal and U on/above."""
    n = a.shape[-1]
    a, ap = _f32(a)
    lu = np.empty((n, n), dtype=np.float32)
    perm = np.empty(n, dtype=np.int32)
    rc = _load().orc_lu_factor_f32(ap, lu.ctypes.data_as(_c_float_p), n, rule, perm.ctypes.data_as(_c_int_p))
    if rc != 0:
        raise ValueError("orc_lu_factor_f32 failed")
    return lu, perm


def openmp_threads() -> int:
    return int(_load().orc_openmp_threads())


def batch_inverse_inplace(x: np.ndarray, piv, rule: int, threads: int) -> None:
    """In-place batched FP32 inversion (used to time the CPU baseline)."""
    n = x.shape[-1]
    batch = x.size // (n * n)
    pp = piv.ctypes.data_as(_c_int_p) if piv is not None else None
    rc = _load().orc_batch_inverse_f32(x.ctypes.data_as(_c_float_p), pp, n, batch, rule, threads)
    if rc != 0:
        raise ValueError("orc_batch_inverse_f32 failed")


def batch_gepp_inplace(x: np.ndarray, piv, threads: int) -> None:
    """In-place batched FP32 GEPP inversion (times the CPU baseline of the
    partial-pivoting mode)."""
    n = x.shape[-1]
    batch = x.size // (n * n)
    pp = piv.ctypes.data_as(_c_int_p) if piv is not None else None
    rc = _load().orc_batch_gepp_f32(x.ctypes.data_as(_c_float_p), pp, None, n, batch, threads)
    if rc != 0:
        raise ValueError("orc_batch_gepp_f32 failed")


def inverse_f64(a: np.ndarray):
    """FP64 GEPP inverse of one FP32 matrix.  Returns (X, gepp_perm, info)."""
    n = a.shape[-1]
    a, ap = _f32(a)
    x = np.zeros((n, n), dtype=np.float64)
    perm = np.empty(n, dtype=np.int32)
    info = _load().orc_inverse_f64(ap, x.ctypes.data_as(_c_double_p), n, perm.ctypes.data_as(_c_int_p))
    return x, perm, int(info)


def cond1(a: np.ndarray) -> float:
    a, ap = _f32(a)
    return float(_load().orc_cond1(ap, a.shape[-1]))


def verify_inv(a: np.ndarray, x: np.ndarray, thr: float = 1e-3, threads: int = 1):
    """verifyInv semantics: (correct_count, max|AX - I| in FP64); threads > 1
    (or 0 = all) splits the matrices over OpenMP threads."""
    n = a.shape[-1]
    a, ap = _f32(a)
    x, xp = _f32(x)
    batch = a.size // (n * n)
    worst = ctypes.c_double(0.0)
    if threads == 1:
        correct = _load().orc_verify_inv(ap, xp, n, batch, thr, ctypes.byref(worst))
    else:
        correct = _load().orc_verify_inv_threads(ap, xp, n, batch, thr, ctypes.byref(worst), threads)
    return int(correct), float(worst.value)


# ----------------------------------------------- reference's own checker

def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def _load_ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(_REF_PATH)
        lib.ref_pivotedA.argtypes = [_c_float_p, ctypes.c_int, _c_int_p, _c_float_p]
        lib.ref_verifyInv.argtypes = [_c_float_p, _c_float_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_longlong)]
        lib.ref_verifyInv.restype = ctypes.c_longlong
        lib.ref_verifyLUwithPivoting.argtypes = [_c_float_p, _c_float_p, ctypes.c_int, ctypes.c_int]
        lib.ref_verifyLUwithPivoting.restype = ctypes.c_longlong
        lib.ref_calc_cond_num.argtypes = [_c_float_p, ctypes.c_int]
        lib.ref_calc_cond_num.restype = ctypes.c_float
        _ref = lib
    return _ref


def ref_pivotedA(a: np.ndarray):
    n = a.shape[-1]
    a, ap = _f32(a)
    piv = np.empty(n, dtype=np.int32)
    pa = np.empty((n, n), dtype=np.float32)
    _load_ref().ref_pivotedA(ap, n, piv.ctypes.data_as(_c_int_p), pa.ctypes.data_as(_c_float_p))
    return piv, pa


def ref_verify_inv(a: np.ndarray, x: np.ndarray):
    n = a.shape[-1]
    a, ap = _f32(a)
    x, xp = _f32(x)
    wrong = ctypes.c_longlong(0)
    correct = _load_ref().ref_verifyInv(ap, xp, n, a.size // (n * n), ctypes.byref(wrong))
    return int(correct), int(wrong.value)


def ref_verify_lu(pa: np.ndarray, lu: np.ndarray) -> int:
    n = pa.shape[-1]
    pa, pap = _f32(pa)
    lu, lup = _f32(lu)
    return int(_load_ref().ref_verifyLUwithPivoting(pap, lup, n, lu.size // (n * n)))


def ref_calc_cond_num(a: np.ndarray) -> float:
    a, ap = _f32(a)
    return float(_load_ref().ref_calc_cond_num(ap, a.shape[-1]))


def ipiv_from_perm(perm: np.ndarray) -> np.ndarray:
    """LAPACK getrf-style pivots (1-based row interchanges) describing the
    permutation `perm` (row i of P*A is row perm[i] of A): at step k the row
    now holding perm[k] is swapped into position k.  Restates no reference
    code (the reference never exports its `pivots`,
    parallel_pivot/luBatchedInplace.cuh:145-148); LAPACK's dgetrf documents
    the format."""
    perm = np.asarray(perm)
    n = perm.shape[-1]
    flat = perm.reshape(-1, n)
    out = np.empty_like(flat, dtype=np.int32)
    for b in range(flat.shape[0]):
        cur = list(range(n))
        for k in range(n):
            p = cur.index(int(flat[b, k]), k)
            out[b, k] = p + 1
            cur[k], cur[p] = cur[p], cur[k]
    return out.reshape(perm.shape)


def lu_info(a: np.ndarray, rule: int = PIV_FIRSTMAX) -> int:
    """LAPACK-style info of the oracle's LU of P*A: 0, or k + 1 for the first
    exactly zero U(k,k)."""
    lu, _ = lu_factor_f32(a, rule)
    z = np.nonzero(np.diag(lu) == 0.0)[0]
    return int(z[0]) + 1 if z.size else 0
