"""GPU tests of the LAPACK-shaped entry (lu_batched_inverse_lapack):
1-based interchange pivots and per-matrix info, checked against the oracle,
against LAPACK's own getrf (scipy.linalg.lu_factor) where both pivot rules
must agree, and against the plain entry point's inverse."""
import numpy as np
import pytest
import scipy.linalg

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 5, 8, 13, 16, 24, 31, 32])
def test_ipiv_info_against_oracle(gpu, n):
    a = oracle.uniform_matrices(700 + n, 37, n)
    print(f"[case] lapack n={n}", flush=True)
    x, ipiv, info = gpu.invert_lapack(a, 0x1)
    xp, perm = gpu.invert(a, 0x1)
    assert np.array_equal(x.view(np.uint32), xp.view(np.uint32)), "inverse differs from the plain entry"
    assert (ipiv == oracle.ipiv_from_perm(oracle.batch_pivots(a, oracle.PIV_FIRSTMAX))).all()
    assert (ipiv == oracle.ipiv_from_perm(perm)).all()
    assert (info == [oracle.lu_info(m) for m in a]).all()


def test_ipiv_matches_lapack_getrf_when_rules_agree(gpu):
    """Row-permuted, strongly diagonally dominant matrices: the reference's
    first-max rule and LAPACK's partial pivoting pick the same rows."""
    rng = np.random.RandomState(3)
    for n in (3, 8, 17, 32):
        d = rng.uniform(-1, 1, size=(20, n, n)).astype(np.float32)
        d[:, np.arange(n), np.arange(n)] = np.float32(10 * n)
        a = np.stack([m[rng.permutation(n)] for m in d])
        print(f"[case] lapack vs getrf n={n}", flush=True)
        _, ipiv, info = gpu.invert_lapack(a, 0x1)
        for i in range(a.shape[0]):
            _, piv = scipy.linalg.lu_factor(a[i].astype(np.float64))
            assert (ipiv[i] == piv + 1).all(), f"n={n} matrix {i}: {ipiv[i]} vs getrf {piv + 1}"
        assert (info == 0).all()


def test_info_flags_exact_zero_pivots(gpu):
    n = 3
    a = np.stack([np.zeros((n, n), np.float32),
                  np.array([[1, 1, 2], [1, 1, 3], [0, 0, 1]], np.float32),
                  np.eye(n, dtype=np.float32),
                  np.array([[2, 1, 0], [4, 2, 5], [1, 0, 1]], np.float32)])
    print("[case] lapack info", flush=True)
    for flags, rule in ((0x1, oracle.PIV_FIRSTMAX), (0x0, oracle.PIV_NONE)):
        _, _, info = gpu.invert_lapack(a, flags)
        assert info.tolist() == [oracle.lu_info(m, rule) for m in a], (flags, info)
    _, _, info = gpu.invert_lapack(a, 0x1)
    assert info[0] == 1 and info[1] == 2 and info[2] == 0
    z = np.zeros((5, 32, 32), np.float32)
    z[1] = np.eye(32)
    _, _, info = gpu.invert_lapack(z, 0x1)
    assert info.tolist() == [1, 0, 1, 1, 1]
