"""CPU checks of the accuracy contract itself (tests/accuracy.py): the
distribution gate on e_gpu / e_oracle must pass the oracle's own result and
reject an inverse whose error is a uniform multiple of the oracle's -- the
kind of result a dropped rounding step or a wrong operand order in the
elimination would give while every per-matrix bound still passes.  The
mistakes are planted in the oracle's result, on the CPU."""
import numpy as np
import pytest

import oracle
from tests import accuracy as acc


def _inflated(a, x, factor):
    """x64 + factor * (x - x64): the same error pattern, `factor` times larger."""
    out = np.empty_like(x)
    for i in range(a.shape[0]):
        x64, _, info = oracle.inverse_f64(a[i])
        assert info == 0
        out[i] = (x64 + factor * (x[i].astype(np.float64) - x64)).astype(np.float32)
    return out


@pytest.fixture(scope="module")
def c1_like():
    """300 of the C1 inputs (1,000 x n=32 U[-1,1), mt19937(42)) and their
    oracle inverses under the reference's first-max rule."""
    a = oracle.uniform_matrices(42, 300, 32)
    x, _ = oracle.lu_inverse_f32(a, oracle.PIV_FIRSTMAX)
    return a, x, acc.rel_err64(a, x)


def test_gate_accepts_the_oracle_itself(c1_like):
    a, x, e = c1_like
    m, p, mx, cnt = acc.ratio_stats(e, e)
    assert (m, p, mx, cnt) == (1.0, 1.0, 1.0, 300)
    assert acc.ratio_gate_ok(e, e)


@pytest.mark.parametrize("factor", [10.0, 3.0])
def test_gate_rejects_uniformly_inflated_error(c1_like, factor):
    a, x, e = c1_like
    e_bad = acc.rel_err64(a, _inflated(a, x, factor))
    # every per-matrix guard bound still passes at 3x: only the gate catches it
    if factor == 3.0:
        assert np.all(e_bad <= acc.contract_bound(a, e))
    for kind, (med, p90) in acc.RatioGate.LIMITS.items():
        if factor == 3.0 and kind == "dominant":
            continue  # the dominant-input pair is checked on dominant inputs below
        assert not acc.ratio_gate_ok(e_bad, e, med, p90), kind
        with pytest.raises(AssertionError):
            g = acc.RatioGate(f"planted x{factor}", kind)
            g.add(e_bad, e)
            g.check()


def test_dominant_gate_rejects_threefold_error_of_the_gpu_order():
    """On dominant inputs the GPU's own errors sit at about 2.5x the oracle's
    (median; Gauss-Jordan rounds more often).  An inverse 3x worse than that
    -- 7.5x the oracle's -- fails the dominant-input pair."""
    a = oracle.dominant_matrices(2024, 300, 24)
    x, _ = oracle.lu_inverse_f32(a, oracle.PIV_NONE)
    e = acc.rel_err64(a, x)
    e_bad = acc.rel_err64(a, _inflated(a, x, 7.5))
    assert acc.ratio_gate_ok(acc.rel_err64(a, _inflated(a, x, 2.5)), e, acc.R_MED_DOM, acc.R_P90_DOM)
    assert not acc.ratio_gate_ok(e_bad, e, acc.R_MED_DOM, acc.R_P90_DOM)


def test_gate_rejects_half_the_matrices_inflated(c1_like):
    """A mistake that hits only some matrices (e.g. one lane layout or one
    tile position): 10x error on every other matrix moves the 90th
    percentile past R_P90."""
    a, x, e = c1_like
    xb = x.copy()
    xb[::2] = _inflated(a[::2], x[::2], 10.0)
    assert not acc.ratio_gate_ok(acc.rel_err64(a, xb), e)


def test_dominant_bound_rejects_inflated_error():
    """The dominant-input contract e <= min(n, C_DOM) * kappa_1 * 2^-24 holds
    for the oracle at every n and fails for 10x its error at n = 32."""
    for n in (2, 8, 17, 32):
        a = oracle.dominant_matrices(2000 + n, 64, n)
        x, _ = oracle.lu_inverse_f32(a, oracle.PIV_NONE)
        assert np.all(acc.rel_err64(a, x) <= acc.dominant_bound(a)), f"n={n}"
    a = oracle.dominant_matrices(2032, 64, 32)
    x, _ = oracle.lu_inverse_f32(a, oracle.PIV_NONE)
    assert not np.all(acc.rel_err64(a, _inflated(a, x, 10.0)) <= acc.dominant_bound(a))


def test_floor_treats_exact_results_as_equal():
    """n = 1: both the oracle's division and an exact reciprocal give e = 0
    or a fraction of an ulp; the floor makes their ratio 1."""
    e_gpu = np.array([0.0, 1e-9, 2.0 ** -27])
    e_orc = np.array([0.0, 0.0, 1e-12])
    r = acc.error_ratios(e_gpu, e_orc)
    assert np.all(r == 1.0)
    # non-finite errors are left to the per-matrix guard
    assert acc.error_ratios(np.array([np.inf, 1e-6]), np.array([1e-6, np.nan])).size == 0
