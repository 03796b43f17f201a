"""GPU tests of the tools around the library: the sweep harness (the run.py
replacement) and the autotuner.  Every case prints its name first."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RUN_PY_KEYS = ("matrix_size", "num_matrices", "num_threads", "runtimes", "runtime_avg", "variance", "std_dev",
               "incorrect_inversions")  # parallel_pivot/run.py:237-247


@pytest.mark.parametrize("pivot,inputs", [("parallel", "random"), ("none", "dominant")])
def test_harness_run_config(gpu, pivot, inputs):
    """One sweep record: the reference's keys, incorrect_inversions equal to
    the oracle's verifyInv count on the same inputs, no pivot mismatches."""
    from matrixinversion_amd import api, harness, workloads

    n, batch = 24, 3000
    print(f"[case] harness run_config n={n} batch={batch} pivot={pivot} inputs={inputs}", flush=True)
    s = gpu.Stream()
    rec = harness.run_config(n, batch, api.PIVOT_NAMES[pivot], 3, inputs, 42 + n, s)
    s.close()
    for k in RUN_PY_KEYS + ("max_rel_err_vs_fp64", "pivot_mismatches", "cpu_cores", "gbps", "pct_hbm"):
        assert k in rec, k
    assert rec["matrix_size"] == n and rec["num_matrices"] == batch and len(rec["runtimes"]) == 3
    a = workloads.make_inputs(inputs, n, batch, 42 + n)
    x = gpu.invert(a, api.PIVOT_NAMES[pivot], return_pivots=False)
    correct, _ = oracle.verify_inv(a, x)
    assert rec["incorrect_inversions"] == [batch - correct] * 3
    assert rec["pivot_mismatches"] == 0
    assert rec["max_rel_err_vs_fp64"] < (1e-5 if inputs == "dominant" else 1.0)
    json.dumps(rec)


def test_harness_main_resumes(gpu, tmp_path):
    """main() writes run.py's layout and key names and skips finished keys."""
    from matrixinversion_amd import harness

    out = str(tmp_path)
    print("[case] harness main + resume", flush=True)
    assert harness.main(["--sizes", "3,8", "--batches", "1000", "--runs", "2", "--out", out]) == 0
    path = os.path.join(out, "runtime_results", "1k", "benchmark_results_parallel.json")
    recs = json.load(open(path))
    assert sorted(recs) == ["m3_n1000_t64", "m8_n1000_t64"]
    recs["m3_n1000_t64"]["runtime_avg"] = -1.0  # a finished key is not measured again
    json.dump(recs, open(path, "w"))
    assert harness.main(["--sizes", "3,8", "--batches", "1000", "--runs", "2", "--out", out]) == 0
    assert json.load(open(path))["m3_n1000_t64"]["runtime_avg"] == -1.0


def test_autotuner_installs_a_table_with_identical_results(gpu, tmp_path):
    """A small autotuner run writes a table; installed, it changes only launch
    shapes: plain and ragged results stay bit-identical."""
    from matrixinversion_amd import autotune

    print("[case] autotune small", flush=True)
    layout0, grab0 = gpu.tuning_get()
    pad0 = gpu.lds_pad_get()
    path = str(tmp_path / "table.json")
    try:
        table = autotune.autotune(sizes=[4, 12, 20, 32], plain_batch=20_000, ragged_batch=20_000)
        with open(path, "w") as f:
            json.dump(table, f)
        assert len(table["layout"]) == 33 and len(table["grab"]) == 33 and len(table["pad"]) == 33
        assert all(table["pad"][n] == 0 for n in range(33) if n % 4 != 0)
        assert all(set(table["detail"][n]["pad_ms"]) == {"0", "1"} for n in (4, 12, 20, 32))
        a = {n: oracle.uniform_matrices(70 + n, 300, n) for n in (4, 12, 20, 32)}
        before = {n: gpu.invert(m, 0x1) for n, m in a.items()}
        mats = [m for n in sorted(a) for m in a[n][:40]]
        rag0 = gpu.invert_ragged(mats, 0x1)
        autotune.load_table(path)
        after = {n: gpu.invert(m, 0x1) for n, m in a.items()}
        rag1 = gpu.invert_ragged(mats, 0x1)
        for n in a:
            assert np.array_equal(before[n][0].view(np.uint32), after[n][0].view(np.uint32)), f"n={n}"
            assert (before[n][1] == after[n][1]).all()
        for x0, x1 in zip(rag0[0], rag1[0]):
            assert np.array_equal(x0.view(np.uint32), x1.view(np.uint32))
    finally:
        gpu.tuning_set(layout0, grab0)
        gpu.lds_pad_set(pad0)


@pytest.mark.parametrize("flags", [0x1, 0x0, 0x4, 0x3])
def test_lds_padding_does_not_change_results(gpu, flags):
    """Every n % 4 == 0 kernel with its LDS row padding flipped from the
    default gives the default's inverses and pivots bit for bit, in both
    lane layouts; partial last tiles included."""
    pad0 = gpu.lds_pad_get()
    try:
        for n in (4, 8, 12, 16, 20, 24, 28, 32):
            a = (oracle.dominant_matrices if flags == 0 else oracle.uniform_matrices)(900 + n, 45, n)
            for lf in ((0x10, 0x20) if n <= 16 else (0x10,)):
                print(f"[case] lds pad n={n} flags={flags | lf:#x}", flush=True)
                gpu.lds_pad_set(pad0)
                x0, p0 = gpu.invert(a, flags | lf)
                p2 = pad0.copy()
                p2[n] = 1 - pad0[n]
                gpu.lds_pad_set(p2)
                x1, p1 = gpu.invert(a, flags | lf)
                assert np.array_equal(x0.view(np.uint32), x1.view(np.uint32)), f"n={n} layout={lf:#x}"
                assert (p0 == p1).all()
    finally:
        gpu.lds_pad_set(pad0)
