"""GPU checks at the benchmark configurations' full sizes, through properties
that need no per-matrix oracle run:

  * 1,000,000 x n=32, U[-1,1), pivoted (the headline configuration): every
    pivot vector is a permutation; pivots of a sample equal the oracle's; the
    device verifyInv count equals the oracle's count on the distinct inputs
    (the batch repeats a 65,536-matrix block, so each distinct matrix appears
    16 times and must be inverted identically every time);
  * 1,000,000 per n in {4, 8, ..., 32}, A = U[-1,1) + n*I, no pivoting (the
    no-pivot sweep): every matrix meets max|AX - I| <= 1e-4 on the device, and
    inverting twice returns A to FP32 accuracy.

Inputs are generated on the host once per block and replicated on the device
(cheap), so the tests run in seconds.  Each case prints its name first.
"""
import numpy as np
import pytest

import oracle
from matrixinversion_amd import workloads
from tests import accuracy as acc

pytestmark = pytest.mark.gpu

BATCH = 1_000_000
BLOCK = 65_536


def _replicated(mia, block: np.ndarray, batch: int):
    """A device batch of `batch` matrices repeating `block` (matrix i =
    block[i % len(block)])."""
    per = block[0].nbytes
    buf = mia.DeviceBuffer(batch * per)
    src = mia.DeviceBuffer.from_array(block)
    done = 0
    while done < batch:
        k = min(block.shape[0], batch - done)
        from matrixinversion_amd._native import check, lib
        check(lib().lubx_memcpy_dtod_async(buf.ptr + done * per, src.ptr, k * per, None), "d2d")
        done += k
    mia.synchronize()
    src.close()
    return buf


def _range(buf, start: int, count: int, n: int, dtype=np.float32, per_matrix=None):
    """Matrices start .. start+count-1 of a device batch (n x n each, or
    per_matrix items of dtype)."""
    per = per_matrix if per_matrix is not None else n * n
    shape = (count, n, n) if per_matrix is None else (count, per)
    return buf.download(shape, dtype, offset=start * per * np.dtype(dtype).itemsize)


def test_headline_configuration_full_size(gpu):
    n = 32
    block = workloads.uniform_matrices(42, BLOCK, n)
    print(f"[case] full size {BATCH} x n={n} pivoted", flush=True)
    A = _replicated(gpu, block, BATCH)
    X = gpu.DeviceBuffer(A.nbytes)
    P = gpu.DeviceBuffer(BATCH * n * 4)
    try:
        X.copy_from(A)
        gpu.lu_batched_inverse(X, n, BATCH, gpu.PIVOT_PARALLEL, P)
        piv = P.download((BATCH, n), np.int32)
        assert (np.sort(piv, axis=1) == np.arange(n)).all(), "a pivot vector is not a permutation"
        reps = BATCH // BLOCK
        assert (piv[:reps * BLOCK].reshape(reps, BLOCK, n) == piv[:BLOCK]).all(), "repeats pivoted differently"
        x = _range(X, 0, BLOCK, n)
        mid = _range(X, 7 * BLOCK, BLOCK, n)
        tail = BATCH - reps * BLOCK
        last = _range(X, reps * BLOCK, tail, n)
        assert np.array_equal(mid.view(np.uint32), x.view(np.uint32)), "a repeat was inverted differently"
        assert np.array_equal(last.view(np.uint32), x[:tail].view(np.uint32)), "the tail was inverted differently"
        sample = np.random.RandomState(0).choice(BLOCK, 512, replace=False)
        assert (piv[sample] == oracle.batch_pivots(block[sample], oracle.PIV_FIRSTMAX)).all()
        xo, _ = oracle.lu_inverse_f32(block[sample], oracle.PIV_FIRSTMAX)
        e = acc.rel_err64(block[sample], x[sample])
        e_orc = acc.rel_err64(block[sample], xo)
        assert np.all(e <= acc.contract_bound(block[sample], e_orc))
        gate = acc.RatioGate("headline sample of 512")
        gate.add(e, e_orc)
        gate.check()
        correct, _ = gpu.residual(A, X, n, BATCH, 1e-3)
        correct_block = oracle.verify_inv(block, x)[0]
        correct_tail = oracle.verify_inv(block[:tail], x[:tail])[0]
        assert correct == reps * correct_block + correct_tail, "device verifyInv count differs from the oracle's"
    finally:
        for b in (A, X, P):
            b.close()


@pytest.mark.parametrize("n", [4, 8, 12, 16, 20, 24, 28, 32])
def test_nopivot_sweep_full_size(gpu, n):
    block = workloads.dominant_matrices(42 + n, 16_384, n)
    print(f"[case] full size {BATCH} x n={n} no pivot, dominant", flush=True)
    A = _replicated(gpu, block, BATCH)
    X = gpu.DeviceBuffer(A.nbytes)
    try:
        X.copy_from(A)
        gpu.lu_batched_inverse(X, n, BATCH, gpu.PIVOT_NONE)
        correct, worst = gpu.residual(A, X, n, BATCH, 1e-4)
        assert correct == BATCH and worst <= 1e-4, f"max|AX-I| = {worst}"
        gpu.lu_batched_inverse(X, n, BATCH, gpu.PIVOT_NONE)  # back to A
        starts = (0, BATCH // 2, BATCH - 1000)
        back = np.concatenate([_range(X, s0, 1000, n) for s0 in starts])
        a = np.concatenate([block[(np.arange(s0, s0 + 1000)) % block.shape[0]] for s0 in starts])
        rel = np.abs(back - a).max(axis=(1, 2)) / np.abs(a).max(axis=(1, 2))
        assert rel.max() <= 64 * n * 2.0 ** -24, f"inverse of the inverse is off by {rel.max():.2e}"
    finally:
        for b in (A, X):
            b.close()


def test_batch_beyond_int32(gpu):
    """2^31 + 5 matrices of n = 1 (8.6 GB): the int64 batch path of
    lu_batched_inverse_async64.  The reference's int32 offsets stop at
    batch * n^2 < 2^31.  Every byte is 0x40, so every matrix is
    [[3.0039062]]; the first, middle and last matrices (and the tail past
    2^31) must hold its FP32 reciprocal, to 1 ulp, and nothing past the
    batch may be touched."""
    batch = 2**31 + 5
    a = np.frombuffer(bytes([0x40] * 4), np.float32)[0]
    expect = np.float32(1) / a
    print(f"[case] n=1 batch={batch} (int64 path)", flush=True)
    X = gpu.DeviceBuffer((batch + 64) * 4)
    s = gpu.Stream()
    try:
        X.fill_bytes(0x40)
        gpu.lu_batched_inverse(X, 1, batch, gpu.PIVOT_PARALLEL, None, s)
        s.synchronize()
        for start in (0, batch // 2 - 3, batch - 16):  # the last window spans index 2^31
            got = X.download((16,), np.float32, offset=start * 4)
            assert np.all(np.abs(got - expect) <= np.spacing(expect)), f"matrices {start}..: {got}"
        beyond = X.download((64,), np.float32, offset=batch * 4)
        assert np.all(beyond == a), "the kernel wrote past the batch"
    finally:
        s.close()
        X.close()
