"""GPU parity: liblubatch (through the C ABI) against the CPU oracle.

Pivot vectors must match the oracle bit for bit; inverses must meet the FP32
accuracy contract of tests/accuracy.py: per matrix the guard bound, and per
test set the distribution gate on e_gpu / e_oracle.  Inputs are seeded mt19937 streams
(oracle.uniform_matrices / dominant_matrices) at sizes the oracle finishes in
seconds.  Every case prints its name before touching the GPU.
"""
import numpy as np
import pytest

import oracle
from tests import accuracy as acc

pytestmark = pytest.mark.gpu

MODES = {
    "none": (0x0, oracle.PIV_NONE),
    "parallel": (0x1, oracle.PIV_FIRSTMAX),
    "serial": (0x2, oracle.PIV_FIRSTMAX),
    "tree": (0x3, oracle.PIV_TREE),
}


def _inputs(mode, n, batch, seed):
    if mode == "none":
        return oracle.dominant_matrices(seed, batch, n)
    return oracle.uniform_matrices(seed, batch, n)


LAYOUTS = {"row16": (0x10, 32), "quad4": (0x20, 16)}


def _check(mia, a, mode, tag, layout_flag=0, gate=None):
    flags, rule = MODES[mode]
    flags |= layout_flag
    print(f"[case] {tag} mode={mode} shape={a.shape}", flush=True)
    x, piv = mia.invert(a, flags)
    xo, po = oracle.lu_inverse_f32(a, rule)
    bad_piv = int((piv != po).any(axis=1).sum())
    assert bad_piv == 0, f"{tag}: {bad_piv} pivot vectors differ from the oracle"
    e_gpu = acc.rel_err64(a, x)
    e_orc = acc.rel_err64(a, xo)
    bound = acc.contract_bound(a, e_orc)
    worst = int(np.argmax(e_gpu / bound))
    assert np.all(e_gpu <= bound), (f"{tag}: matrix {worst} e_gpu={e_gpu[worst]:.3e} "
                                    f"e_oracle={e_orc[worst]:.3e} bound={bound[worst]:.3e}")
    if gate is not None:
        gate.add(e_gpu, e_orc)
    return x, piv, e_gpu, e_orc


@pytest.mark.parametrize("layout", ["row16", "quad4"])
@pytest.mark.parametrize("mode", ["parallel", "none", "serial", "tree"])
def test_all_sizes(gpu, mode, layout):
    """Every n of the layout; 37 matrices is a partial last tile in both layouts
    (9 tiles + 1 of 4, or 2 tiles + 5 of 16)."""
    lf, nmax = LAYOUTS[layout]
    rows = []
    gate = acc.RatioGate(f"all sizes mode={mode} {layout}", "dominant" if mode == "none" else "ref")
    for n in range(1, nmax + 1):
        a = _inputs(mode, n, 37, seed=1000 + n)
        _, _, e_gpu, e_orc = _check(gpu, a, mode, f"n={n} {layout}", lf, gate)
        rows.append((n, e_gpu.max(), e_orc.max()))
    print("\n".join(f"  n={n:2d} e_gpu={g:.2e} e_orc={o:.2e}" for n, g, o in rows))
    gate.check()


@pytest.mark.parametrize("batch", [1, 2, 3, 4, 5, 7, 8, 9, 63, 64, 65, 1000])
def test_batch_tails_n32(gpu, batch):
    a = oracle.uniform_matrices(42, batch, 32)
    gate = acc.RatioGate(f"n=32 batch={batch}")
    _check(gpu, a, "parallel", f"batch={batch}", gate=gate)
    if batch >= 100:
        gate.check()


@pytest.mark.parametrize("batch", [1, 15, 16, 17, 33, 1000])
def test_batch_tails_quads(gpu, batch):
    gate = acc.RatioGate(f"quad4 batch={batch}")
    for n in (4, 7, 16):
        a = oracle.uniform_matrices(43, batch, n)
        _check(gpu, a, "parallel", f"batch={batch} n={n} quad4", 0x20, gate)
    if batch >= 100:
        gate.check()


def test_config1_n32_1000(gpu):
    """1,000 x n=32, U[-1,1] from mt19937(42), pivoted: the reference's C1 case."""
    a = oracle.uniform_matrices(42, 1000, 32)
    gate = acc.RatioGate("C1 1000 x n=32")
    x, piv, e_gpu, e_orc = _check(gpu, a, "parallel", "C1", gate=gate)
    correct, worst = oracle.verify_inv(a, x)
    correct_o, worst_o = oracle.verify_inv(a, oracle.lu_inverse_f32(a, oracle.PIV_FIRSTMAX)[0])
    print(f"  C1 verifyInv correct gpu={correct} oracle={correct_o}; max|AX-I| gpu={worst:.3e} oracle={worst_o:.3e}; "
          f"median e_gpu={np.median(e_gpu):.2e} e_orc={np.median(e_orc):.2e}")
    gate.check()


def test_ties_use_exact_rule(gpu):
    """Integer matrices (reference fixture mtrand32_new.txt style) have many exact
    |a| ties: the DPP fast path must hand them to the exact first-max scan."""
    rng = np.random.RandomState(7)
    for n in (2, 5, 8, 13, 16, 24, 32):
        a = rng.randint(0, 10, size=(64, n, n)).astype(np.float32)
        a += np.float32(0.5) * (np.arange(n) == np.arange(n)[:, None])  # keep most of them invertible
        for mode in ("parallel", "serial", "tree"):
            for layout, (lf, nmax) in LAYOUTS.items():
                if n > nmax:
                    continue
                flags, rule = MODES[mode]
                print(f"[case] ties n={n} mode={mode} {layout}", flush=True)
                _, piv = gpu.invert(a, flags | lf)
                po = oracle.batch_pivots(a, rule)
                assert (piv == po).all(), f"ties n={n} {mode} {layout}: pivots differ"


def test_real_valued_ties_of_the_1m_workload(gpu):
    """The 17 matrices of the 1M x n=32 seed-42 workload on which the tree rule
    and first-max disagree: each has an exact |a| tie at a column maximum
    (tests/golden/seed42_tie_matrices.npz).  The DPP fast paths must hand them
    to the exact rule, so every flag still matches its oracle rule."""
    import os

    from tests.conftest import GOLDEN

    a = np.load(os.path.join(GOLDEN, "seed42_tie_matrices.npz"))["a"]
    gate = acc.RatioGate("seed42 1M ties")
    for mode in ("parallel", "serial", "tree"):
        _check(gpu, a, mode, "seed42 1M ties", gate=gate)
    gate.check()


def test_nan_and_zero_pivot_rules(gpu):
    """NaN at position k is never beaten, NaN elsewhere never wins; all-zero
    columns keep position k (strict '>' of the reference)."""
    rng = np.random.RandomState(11)
    n = 32
    a = rng.uniform(-1, 1, size=(16, n, n)).astype(np.float32)
    a[0, 0, 0] = np.nan
    a[1, 5, 0] = np.nan
    a[2, :, 3] = 0.0
    a[3, 7, :] = np.nan
    a[4] = 0.0
    for mode in ("parallel", "serial", "tree"):
        flags, rule = MODES[mode]
        print(f"[case] nan/zero mode={mode}", flush=True)
        _, piv = gpu.invert(a, flags)
        po = oracle.batch_pivots(a, rule)
        assert (piv == po).all(), f"{mode}: pivots differ on NaN/zero inputs"


def test_unaligned_batch(gpu):
    """A not 16-byte aligned takes the element-wise tile path."""
    n, batch = 20, 23
    a = oracle.uniform_matrices(5, batch, n)
    buf = gpu.DeviceBuffer(a.nbytes + 4)
    piv = gpu.DeviceBuffer(batch * n * 4)
    try:
        buf.upload(a, offset=4)

        class _Off:  # a device pointer 4 bytes into the allocation
            ptr = buf.ptr + 4
            nbytes = a.nbytes

        print("[case] unaligned", flush=True)
        gpu.lu_batched_inverse(_Off, n, batch, gpu.PIVOT_PARALLEL, piv)
        x = buf.download(a.shape, np.float32, offset=4)
        p = piv.download((batch, n), np.int32)
    finally:
        buf.close()
        piv.close()
    xo, po = oracle.lu_inverse_f32(a, oracle.PIV_FIRSTMAX)
    assert (p == po).all()
    assert np.all(acc.rel_err64(a, x) <= acc.contract_bound(a, acc.rel_err64(a, xo)))


@pytest.mark.parametrize("mode", ["none", "parallel"])
def test_well_conditioned_bounds(gpu, mode):
    """Diagonally dominant inputs (A = U[-1,1] + n*I): e <= min(n, C_DOM)*kappa_1*2^-24
    and max|A X - I| <= 1e-4 for every n, and the distribution gate against
    the oracle over all of them."""
    flags, rule = MODES[mode]
    gate = acc.RatioGate(f"dominant mode={mode}", "dominant")
    worst = 0.0
    for n in range(1, 33):
        a = oracle.dominant_matrices(2000 + n, 64, n)
        print(f"[case] dominant n={n} mode={mode}", flush=True)
        x = gpu.invert(a, flags, return_pivots=False)
        e = acc.rel_err64(a, x)
        b = acc.dominant_bound(a)
        worst = max(worst, float((e / (acc.kappa1(a) * acc.EPS32)).max()))
        assert np.all(e <= b), f"n={n}: e={e.max():.3e} bound={b[np.argmax(e / b)]:.3e}"
        r = acc.max_abs_axmi(a, x)
        assert r.max() <= 1e-4, f"n={n}: max|AX-I|={r.max():.3e}"
        gate.add(e, acc.rel_err64(a, oracle.lu_inverse_f32(a, rule)[0]))
    print(f"[gate] dominant mode={mode}: max e / (kappa_1 2^-24) = {worst:.3f} (<= min(n, {acc.C_DOM}))",
          flush=True)
    gate.check()


@pytest.mark.parametrize("n,batch", [(16, 64), (32, 4096)])
def test_device_residual_matches_verify_inv(gpu, n, batch):
    """lu_batched_residual counts exactly what the oracle's verifyInv
    restatement counts (same FP32 operations and summation order), NaN
    included; 4096 random n=32 matrices hold borderline cases that an FMA
    would count differently."""
    a = oracle.uniform_matrices(77, batch, n)
    x, _ = oracle.lu_inverse_f32(a, oracle.PIV_FIRSTMAX)
    x[3, 2, 2] += np.float32(0.5)
    x[5, :, :] = np.nan
    A = gpu.DeviceBuffer.from_array(a)
    X = gpu.DeviceBuffer.from_array(x)
    try:
        print("[case] residual", flush=True)
        correct, worst = gpu.residual(A, X, n, batch, 1e-3)
    finally:
        A.close()
        X.close()
    correct_o, worst_o = oracle.verify_inv(a, x)
    assert correct == correct_o
    assert np.isnan(worst) or worst >= 0.4
