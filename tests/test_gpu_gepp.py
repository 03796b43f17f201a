"""GPU parity of the partial-pivoting rule, LUB_PIVOT_GEPP (LAPACK getrf's
rule: largest |a| of the updated column).  It has no counterpart in the
reference, whose rules search the un-updated column.

Its pivot vectors must equal, bit for bit, those of the oracle's FP32 GEPP
restatement (oracle.gepp_inverse_f32), of LAPACK dgetrf (scipy) and of
rocSOLVER sgetrf on the device.  Its inverses must meet the accuracy
contract of tests/accuracy.py against that restatement (the per-matrix
guard and the distribution gate with the partial-pivoting constants).  On U[-1,1] inputs,
where the reference rule loses accuracy, it must be clearly more accurate.
Every case prints its name before touching the GPU.
"""
import os

import numpy as np
import pytest

import oracle
from tests import accuracy as acc

pytestmark = pytest.mark.gpu

GEPP = 0x4
LAYOUTS = {"row16": (0x10, 32), "quad4": (0x20, 16)}


def _check(mia, a, tag, layout_flag=0, gate=None):
    print(f"[case] gepp {tag} shape={a.shape}", flush=True)
    x, piv = mia.invert(a, GEPP | layout_flag)
    xo, po, info = oracle.gepp_inverse_f32(a)
    bad = int((piv != po).any(axis=1).sum())
    assert bad == 0, f"{tag}: {bad} pivot vectors differ from the oracle"
    e_gpu = acc.rel_err64(a, x)
    e_orc = acc.rel_err64(a, xo)
    bound = acc.contract_bound(a, e_orc)
    worst = int(np.argmax(e_gpu / bound))
    assert np.all(e_gpu <= bound), (f"{tag}: matrix {worst} e_gpu={e_gpu[worst]:.3e} "
                                    f"e_oracle={e_orc[worst]:.3e} bound={bound[worst]:.3e}")
    if gate is not None:
        gate.add(e_gpu, e_orc)
    return x, piv, e_gpu, e_orc


def _gate(tag):
    return acc.RatioGate(tag, "gepp")


@pytest.mark.parametrize("layout", ["row16", "quad4"])
def test_gepp_all_sizes(gpu, layout):
    """Every n of the layout; 37 matrices end in a partial tile in both."""
    lf, nmax = LAYOUTS[layout]
    gate = _gate(f"gepp all sizes {layout}")
    for n in range(1, nmax + 1):
        a = oracle.uniform_matrices(3000 + n, 37, n)
        _check(gpu, a, f"n={n} {layout}", lf, gate)
    gate.check()


@pytest.mark.parametrize("batch", [1, 3, 5, 17, 1000])
def test_gepp_batch_tails_n32(gpu, batch):
    a = oracle.uniform_matrices(3100 + batch, batch, 32)
    gate = _gate(f"gepp n=32 batch={batch}")
    _check(gpu, a, f"batch={batch}", gate=gate)
    if batch >= 100:
        gate.check()


def test_gepp_ipiv_and_info_equal_lapack_getrf(gpu):
    """The LAPACK-shaped entry under GEPP returns getrf's own interchanges."""
    import scipy.linalg as sla

    for n in (5, 17, 32):
        a = oracle.uniform_matrices(3200 + n, 100, n)
        print(f"[case] gepp lapack ipiv n={n}", flush=True)
        x, ipiv, info = gpu.invert_lapack(a, GEPP)
        assert (info == 0).all()
        for i in range(len(a)):
            _, ref = sla.lu_factor(a[i].astype(np.float64))
            assert (ipiv[i] == ref + 1).all(), f"n={n} matrix {i}: ipiv differs from dgetrf"


def test_gepp_ipiv_equals_rocsolver_getrf(gpu):
    """rocSOLVER's batched sgetrf (an independent implementation on the same
    device) chooses the same interchanges.  rocSOLVER is column-major, so it
    is handed A^T: its view of that buffer is A."""
    import ctypes

    from matrixinversion_amd import _native
    from matrixinversion_amd._native import check, cmp_lib

    if not os.path.exists(_native.CMP_LIB_PATH):
        pytest.skip("rocSOLVER comparator not built")
    n, batch = 32, 200
    a = oracle.uniform_matrices(3300, batch, n)
    print("[case] gepp vs rocsolver ipiv n=32", flush=True)
    A = gpu.DeviceBuffer.from_array(np.ascontiguousarray(a.transpose(0, 2, 1)))
    P = gpu.DeviceBuffer(batch * n * 4)
    I = gpu.DeviceBuffer(batch * 4)
    h = ctypes.c_void_p()
    lib = cmp_lib()
    check(lib.lubref_rocsolver_create(ctypes.byref(h), None), "rocsolver create")
    try:
        check(lib.lubref_rocsolver_inverse(h, A.ptr, P.ptr, I.ptr, n, batch), "rocsolver inverse")
        gpu.synchronize()
        ipiv_rs = P.download((batch, n), np.int32)
        info_rs = I.download((batch,), np.int32)
    finally:
        lib.lubref_rocsolver_destroy(h)
        for b in (A, P, I):
            b.close()
    _, ipiv, info = gpu.invert_lapack(a, GEPP)
    assert (info_rs == 0).all() and (info == 0).all()
    assert (ipiv == ipiv_rs).all()


def test_gepp_is_more_accurate_than_the_reference_rule(gpu):
    """1000 x n=32 U[-1,1] (seed 42, the C1 inputs): the partial-pivoting
    inverse is an order of magnitude closer to the FP64 inverse, and passes
    the reference's verifyInv check (|AX - I| < 1e-3) more often.  (A few of
    these matrices are ill-conditioned enough to fail it under any rule.)"""
    a = oracle.uniform_matrices(42, 1000, 32)
    print("[case] gepp vs parallel accuracy C1", flush=True)
    x_g, _ = gpu.invert(a, GEPP)
    x_r, _ = gpu.invert(a, 0x1)
    e_g, e_r = acc.rel_err64(a, x_g), acc.rel_err64(a, x_r)
    print(f"  median e gepp={np.median(e_g):.2e} ref={np.median(e_r):.2e}; max {e_g.max():.2e} {e_r.max():.2e}")
    assert np.median(e_g) * 10 < np.median(e_r)
    assert e_g.max() < e_r.max()
    correct_g, _ = oracle.verify_inv(a, x_g)
    correct_r, _ = oracle.verify_inv(a, x_r)
    print(f"  verifyInv correct gepp={correct_g} ref={correct_r}")
    assert correct_g > correct_r


def test_gepp_singular_inputs(gpu):
    """Zero matrix: info 1 and a NaN inverse; a zero column k (the first k
    pivots regular): info k + 1; NaN inputs give NaN, never a hang."""
    n = 32
    rng = np.random.RandomState(77)
    a = rng.uniform(-1, 1, size=(4, n, n)).astype(np.float32)
    a[0] = 0.0
    a[1][:, 9] = 0.0
    a[2, 5, 7] = np.nan
    print("[case] gepp singular / NaN", flush=True)
    x, ipiv, info = gpu.invert_lapack(a, GEPP)
    _, _, info_o = oracle.gepp_inverse_f32(a)
    assert info.tolist()[:2] == [1, 10] and info[3] == 0
    assert (info[:2] == info_o[:2]).all()
    assert np.isnan(x[0]).all() and np.isnan(x[2]).any()
    assert np.isfinite(x[3]).all()
