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


def _inverse_refined(gpu, a, flags, offset=0, x_offset=0):
    """lu_batched_inverse_refined through the C ABI (A at `offset` bytes into
    its buffer, X at `x_offset` bytes into its own): (refined inverse,
    pivots, A after the call)."""
    from matrixinversion_amd._native import check, lib

    batch, n, _ = a.shape
    A = gpu.DeviceBuffer(a.nbytes + offset)
    X = gpu.DeviceBuffer(a.nbytes + x_offset)
    P = gpu.DeviceBuffer(batch * n * 4)
    try:
        A.upload(a, offset=offset)
        check(lib().lu_batched_inverse_refined(A.ptr + offset, X.ptr + x_offset, P.ptr, n, batch, flags, None),
              "refined")
        gpu.synchronize()
        return (X.download(a.shape, np.float32, offset=x_offset), P.download((batch, n), np.int32),
                A.download(a.shape, np.float32, offset=offset))
    finally:
        for b in (A, X, P):
            b.close()


def _within_step(a, x0, x1, ulps=4):
    """x1 equals one refinement step of x0 (FP64 residual) within `ulps`."""
    ref = _step_reference(a, x0)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True)
    return np.all(np.abs(x1 - ref) <= ulps * 2.0 ** -24 * (np.abs(ref) + 1e-3 * scale))


def test_inverse_refined_entry_keeps_a(gpu):
    n, batch = 32, 40
    a = workloads.conditioned_matrices(5, batch, n, 3)
    print("[case] lu_batched_inverse_refined", flush=True)
    x, p, a_after = _inverse_refined(gpu, a, 0x1)
    assert np.array_equal(a_after, a), "A was modified"
    x0, x1 = gpu.invert_refined(a, 0x1)
    assert _within_step(a, x0, x), "fused refinement differs from one step of the plain inverse"
    assert np.all(np.abs(x - x1) <= 4 * 2.0 ** -24 * (np.abs(x1) + 1e-3 * np.abs(x1).max())), \
        "fused and three-step refinement disagree"
    assert (p == oracle.batch_pivots(a, oracle.PIV_FIRSTMAX)).all()


@pytest.mark.parametrize("flags", [0x0, 0x1, 0x2])
def test_fused_refinement_every_size(gpu, flags):
    """n = 17..32 take the fused kernel (inverse + refinement with A read
    once): its result is one refinement step of the plain inverse, within
    4 ulp, for batches ending in a partial tile, for an unaligned A (the
    element-wise tile load) and for an unaligned X (the element-wise tile
    store); pivots equal the oracle's; A is untouched."""
    rule = oracle.PIV_NONE if flags == 0 else oracle.PIV_FIRSTMAX
    for n in range(17, 33):
        for batch, offset, x_offset in ((23, 0, 0), (9, 4, 0), (12, 0, 4)):
            a = oracle.dominant_matrices(700 + n, batch, n) if flags == 0 else \
                workloads.conditioned_matrices(800 + n, batch, n, 2)
            print(f"[case] fused refine n={n} flags={flags:#x} batch={batch} offset={offset} x_offset={x_offset}",
                  flush=True)
            x, p, a_after = _inverse_refined(gpu, a, flags, offset, x_offset)
            assert np.array_equal(a_after, a), f"n={n}: A was modified"
            x0, _ = gpu.invert(a, flags)
            assert _within_step(a, x0, x), f"n={n}: not one refinement step of the plain inverse"
            assert (p == oracle.batch_pivots(a, rule)).all(), f"n={n}: pivots"


def test_fused_refinement_more_tiles_than_the_grid(gpu):
    """2^22 tiles of four n = 32 matrices plus 1,000 tiles and 3 matrices
    (16.8M matrices; A and X 68.7 GB each): the grid is capped at 2^22
    workgroups, so the first 1,001 stride to a second tile and fetch it into
    registers during their first tile's refinement, and the batch ends in a
    partial tile.  The batch repeats a block of 16,384 matrices; windows at
    the start, across the grid wrap and at the end equal the fused kernel's
    result on the block itself, bit for bit, and A is untouched there."""
    from matrixinversion_amd._native import check, lib

    n, block_count = 32, 16_384
    batch = (1 << 22) * 4 + 4 * 1000 + 3
    block = workloads.conditioned_matrices(820, block_count, n, 2)
    xb, pb, _ = _inverse_refined(gpu, block, 0x1)
    per = n * n * 4
    print(f"[case] fused refine n={n} batch={batch} ({2 * batch * per / 1e9:.1f} GB)", flush=True)
    A = gpu.DeviceBuffer(batch * per)
    X = gpu.DeviceBuffer(batch * per)
    P = gpu.DeviceBuffer(batch * n * 4)
    src = gpu.DeviceBuffer.from_array(block)
    s = gpu.Stream()
    try:
        done = 0
        while done < batch:
            k = min(block_count, batch - done)
            check(lib().lubx_memcpy_dtod_async(A.ptr + done * per, src.ptr, k * per, s.handle), "d2d")
            done += k
        check(lib().lu_batched_inverse_refined(A.ptr, X.ptr, P.ptr, n, batch, 0x1, s.handle), "refined")
        s.synchronize()
        wrap = (1 << 22) * 4  # first matrix of tile 2^22: the first strided iteration
        for start in (0, wrap - 2048, batch - 4096):
            x = X.download((4096, n, n), np.float32, offset=start * per)
            p = P.download((4096, n), np.int32, offset=start * n * 4)
            a = A.download((4096, n, n), np.float32, offset=start * per)
            idx = (start + np.arange(4096)) % block_count
            assert np.array_equal(x.view(np.uint32), xb[idx].view(np.uint32)), f"window at {start}"
            assert (p == pb[idx]).all(), f"pivots at {start}"
            assert np.array_equal(a, block[idx]), f"A changed at {start}"
    finally:
        s.close()
        for b in (A, X, P, src):
            b.close()
