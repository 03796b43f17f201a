"""The sweep harness (run.py replacement): inputs, keys and sizes (CPU only)."""
import os

import numpy as np

import oracle
from matrixinversion_amd import harness
from tests.conftest import GOLDEN


def test_random_inputs_match_the_oracle_generator():
    """numpy's MT19937 with legacy seeding and the oracle's C MT19937 give the
    same matrices under the pinned float mapping."""
    for n, batch, seed in ((3, 5, 1), (32, 7, 42)):
        assert (harness.make_inputs("random", n, batch, seed) == oracle.uniform_matrices(seed, batch, n)).all()
        assert (harness.make_inputs("dominant", n, batch, seed) == oracle.dominant_matrices(seed, batch, n)).all()


def test_template_inputs_replicate_the_fixture_prefix():
    """parallel_pivot/luBatchedInplace.cu:44-53, 88-101: the first n*n numbers of
    the fixture, row-major, copied into every matrix."""
    path = os.path.join(GOLDEN, "mtrand32_new1.txt")
    a = harness.make_inputs(f"template:{path}", 5, 3, 0)
    t = np.loadtxt(path, dtype=np.float32).ravel()[:25].reshape(5, 5)
    assert a.shape == (3, 5, 5) and all((m == t).all() for m in a)


def test_sizes_and_labels():
    assert harness.parse_sizes("1-4,8,30-32") == [1, 2, 3, 4, 8, 30, 31, 32]
    assert harness._batch_label(100000) == "100k"
    assert harness._batch_label(1000000) == "1M"
    assert harness._batch_label(1234) == "1234"


def test_workload_generators_match_the_oracle():
    import oracle
    from matrixinversion_amd import workloads

    assert workloads.mt19937_raw(42, 3).tolist() == [1608637542, 3421126067, 4083286876]
    assert (workloads.mt19937_raw(7, 1000) == oracle.mt19937_raw(7, 1000)).all()
    for n in (1, 5, 32):
        assert (workloads.uniform_matrices(11, 3, n) == oracle.uniform_matrices(11, 3, n)).all()
        assert (workloads.dominant_matrices(11, 3, n) == oracle.dominant_matrices(11, 3, n)).all()


def test_ragged_batch_layout():
    from matrixinversion_amd import workloads

    flat, offs, ns, poffs = workloads.ragged_batch(42, 5000)
    assert ns.min() == 2 and ns.max() == 32 and abs(ns.mean() - 17) < 0.5
    sq = ns.astype(np.int64) ** 2
    assert (offs[1:] == offs[:-1] + sq[:-1]).all() and offs[0] == 0
    assert flat.size == sq.sum()
    assert (poffs[1:] == poffs[:-1] + ns[:-1]).all()
    assert (flat == workloads.mt19937_uniform(42, flat.size)).all()
    assert workloads.ragged_bytes(ns) == int((8 * sq + 4 * ns + 20).sum())


def test_ragged_batch_dominant():
    from matrixinversion_amd import workloads

    flat, offs, ns, _ = workloads.ragged_batch(3, 50, 4, 9)
    dflat, _, _, _ = workloads.ragged_batch(3, 50, 4, 9, dominant=True)
    for o, n in zip(offs, ns):
        a = flat[o:o + n * n].reshape(n, n)
        d = dflat[o:o + n * n].reshape(n, n)
        assert (d == a + np.float32(n) * np.eye(n, dtype=np.float32)).all()


def test_host_pivot_check_equals_the_oracle_first_max_rule():
    """harness.pivoted_a (the reference driver's pivotedA, vectorised) equals
    the oracle's first-max rule on random, tied (integer) and NaN inputs, and
    the reference's own pivotedA where oracle/_ref is built."""
    rng = np.random.RandomState(3)
    cases = [oracle.uniform_matrices(5, 50, 17),
             rng.randint(0, 4, size=(50, 9, 9)).astype(np.float32),
             oracle.uniform_matrices(6, 20, 32)]
    cases[2][0, 0, 0] = np.nan      # NaN at position k: kept
    cases[2][1, 4, 0] = np.nan      # NaN elsewhere: never wins
    cases[2][2, :, 3] = 0.0          # all-zero column: position k
    for a in cases:
        got = harness.pivoted_a(a)
        want = oracle.batch_pivots(a, oracle.PIV_FIRSTMAX)
        assert (got == want).all()
        if oracle.ref_available():
            for m, g in zip(a[:10], got[:10]):
                if not np.isnan(m).any():
                    assert (oracle.ref_pivotedA(m)[0] == g).all()


def test_host_checks_count_mismatches_and_errors():
    from matrixinversion_amd import api

    a = oracle.uniform_matrices(9, 64, 12)
    x, piv = oracle.lu_inverse_f32(a, oracle.PIV_FIRSTMAX)
    rel, mism = harness.host_checks(a, x, piv, api.PIVOT_PARALLEL)
    assert mism == 0 and 0 < rel < 1e-2
    bad = piv.copy()
    bad[[3, 40], :2] = bad[[3, 40], 1::-1]
    assert harness.host_checks(a, x, bad, api.PIVOT_PARALLEL)[1] == 2
    assert harness.host_checks(a, x, piv, api.PIVOT_TREE)[1] is None
    x2 = x.copy()
    x2[7] *= 1.01
    assert harness.host_checks(a, x2, piv, api.PIVOT_PARALLEL)[0] > 5e-3


def test_serving_drivers_share_one_schedule():
    """The native and the Python serving drivers replay the same stream:
    schedule() is what both build their requests from, and it is a pure
    function of its seed."""
    import numpy as np

    from matrixinversion_amd import serve

    class FakePool:
        def __init__(self, size):
            self.ns = np.full(size, 8, np.int32)

    pools = (FakePool(5000), FakePool(5000))
    a = serve.schedule(pools, 3000.0, 0.2, seed=11)
    b = serve.schedule(pools, 3000.0, 0.2, seed=11)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    arrivals, counts, starts, pivot, flags = a
    assert np.all(np.diff(arrivals) >= 0) and arrivals.max() <= 0.2
    assert counts.min() >= 1 and counts.max() <= 4096 and np.all(starts + counts <= 5000)
    assert set(np.unique(flags)) <= {1, 0}
