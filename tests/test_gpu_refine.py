"""GPU tests of one refinement step X <- X + X (I - A X) (lu_batched_refine):
the kernel against a CPU restatement of the step (FP64 residual rounded to
FP32, then X + X R), for every n (padding to 32), and the accuracy gain on
the precision workload (A = Q1 diag(sigma) Q2^T, sigma log-spaced 1..10^k)
measured against the FP64 inverse."""
import numpy as np
import pytest

import oracle
from matrixinversion_amd import workloads

pytestmark = pytest.mark.gpu


def _step_reference(a, x):
    """X + X fp32(I - A X), the residual exact up to FP64 summation."""
    a64, x64 = a.astype(np.float64), x.astype(np.float64)
    n = a.shape[-1]
    r = (np.eye(n) - a64 @ x64).astype(np.float32).astype(np.float64)
    return x64 + x64 @ r


def _rel_err(a, x):
    inv = np.linalg.inv(a.astype(np.float64))
    return np.linalg.norm(x - inv, axis=(1, 2)) / np.linalg.norm(inv, axis=(1, 2))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 16, 17, 24, 31, 32])
def test_refine_step_matches_reference(gpu, n):
    a = oracle.dominant_matrices(300 + n, 65, n)
    print(f"[case] refine step n={n}", flush=True)
    x0, x1 = gpu.invert_refined(a, 0x1)
    ref = _step_reference(a, x0)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(x1 - ref) <= 4 * 2.0 ** -24 * (np.abs(ref) + 1e-3 * scale)), "refinement step differs"


@pytest.mark.parametrize("k", [2, 4, 6])
def test_refinement_reduces_error(gpu, k):
    a = workloads.conditioned_matrices(11 + k, 256, 32, k)
    print(f"[case] refine conditioned k={k}", flush=True)
    x0, x1 = gpu.invert_refined(a, 0x1)
    e0, e1 = _rel_err(a, x0), _rel_err(a, x1)
    print(f"  k={k}: median rel err plain {np.median(e0):.2e} refined {np.median(e1):.2e}; "
          f"max {e0.max():.2e} -> {e1.max():.2e}", flush=True)
    if k <= 4:
        assert np.median(e1) < np.median(e0) / 10
    else:
        assert np.median(e1) <= np.median(e0)


def test_inverse_refined_entry_keeps_a(gpu):
    n, batch = 32, 40
    a = workloads.conditioned_matrices(5, batch, n, 3)
    A = gpu.DeviceBuffer.from_array(a)
    X = gpu.DeviceBuffer(a.nbytes)
    P = gpu.DeviceBuffer(batch * n * 4)
    print("[case] lu_batched_inverse_refined", flush=True)
    try:
        from matrixinversion_amd._native import check, lib
        check(lib().lu_batched_inverse_refined(A.ptr, X.ptr, P.ptr, n, batch, 0x1, None), "refined")
        gpu.synchronize()
        x = X.download(a.shape, np.float32)
        assert np.array_equal(A.download(a.shape, np.float32), a), "A was modified"
        _, x1 = gpu.invert_refined(a, 0x1)
        assert np.array_equal(x.view(np.uint32), x1.view(np.uint32))
        p = P.download((batch, n), np.int32)
        assert (p == oracle.batch_pivots(a, oracle.PIV_FIRSTMAX)).all()
    finally:
        for b in (A, X, P):
            b.close()
