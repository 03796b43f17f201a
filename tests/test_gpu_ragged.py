"""GPU parity of the ragged (mixed-n) entry points.

A packed batch of matrices of mixed n, in any order, at any float offset, is
inverted by lu_batched_inverse_ragged[_plan].  Each matrix's inverse and
pivot vector must equal, bit for bit, what the plain entry point returns for
that matrix in a batch of its own n with the same lane layout (the ragged
kernel runs the same tile routines), and must meet the oracle parity bar of
the plain path (pivots equal to the oracle's, accuracy contract of
tests/accuracy.py).  Every case prints its name before touching the GPU.
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
    "gepp": (0x4, oracle.PIV_GEPP),
}
RAGGED_QUAD_MAX_N = 12


def _mixed(mode, ns, seed):
    """One seeded input per matrix (dominant for the unpivoted rule)."""
    mats = []
    for i, n in enumerate(ns):
        gen = oracle.dominant_matrices if mode == "none" else oracle.uniform_matrices
        mats.append(gen(seed + i, 1, int(n))[0])
    return mats


def _pack_scattered(mats, rng):
    """Pack with random gaps (odd float offsets) and a shuffled order in memory."""
    order = rng.permutation(len(mats))
    offs = np.zeros(len(mats), np.int64)
    pos = 0
    for i in order:
        pos += int(rng.randint(0, 4))
        offs[i] = pos
        pos += mats[i].size
    flat = np.full(pos + 3, np.float32(-7.25), np.float32)
    for i, m in enumerate(mats):
        flat[offs[i]:offs[i] + m.size] = m.ravel()
    ns = np.array([m.shape[0] for m in mats], np.int32)
    poffs = np.zeros(len(mats), np.int64)
    poffs[1:] = np.cumsum(ns[:-1].astype(np.int64))
    return flat, offs, ns, poffs


def _run_ragged(mia, flat, offs, ns, poffs, flags, plan=None, with_piv=True):
    A = mia.DeviceBuffer.from_array(flat)
    O = mia.DeviceBuffer.from_array(offs)
    NS = mia.DeviceBuffer.from_array(ns)
    total = int(np.clip(ns, 0, None).sum()) if len(ns) else 0
    P = mia.DeviceBuffer(max(total * 4, 4)) if with_piv else None
    PO = mia.DeviceBuffer.from_array(poffs) if with_piv else None
    try:
        if P is not None:
            P.fill_bytes(0xFF)
        mia.lu_batched_inverse_ragged(A, O, NS, len(ns), flags, P, PO, plan=plan)
        mia.synchronize()
        x = A.download(flat.shape, np.float32)
        p = P.download((total,), np.int32) if with_piv else None
    finally:
        for b in (A, O, NS, P, PO):
            if b is not None:
                b.close()
    return x, p


def _plain_layout(n, ragged_layout_flag):
    """The plain-path layout flag that runs the same tile routine as the ragged
    kernel under `ragged_layout_flag` (quads only up to n = 12 there)."""
    if ragged_layout_flag == 0x20 and n <= RAGGED_QUAD_MAX_N:
        return 0x20
    if ragged_layout_flag == 0x00 and n <= RAGGED_QUAD_MAX_N:
        return 0x00  # default table: quads for n <= 12 in both paths
    return 0x10


def _same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("layout", [0x10, 0x20, 0x00])
@pytest.mark.parametrize("mode", ["parallel", "none", "serial", "tree", "gepp"])
def test_ragged_matches_plain_and_oracle(gpu, mode, layout):
    rng = np.random.RandomState(100 + layout + len(mode))
    ns = rng.randint(1, 33, size=300)
    ns[:40] = 32  # several full tiles of the largest class
    ns[40:60] = 3
    mats = _mixed(mode, ns, seed=5000)
    flat, offs, ns32, poffs = _pack_scattered(mats, rng)
    flags, rule = MODES[mode]
    print(f"[case] ragged mode={mode} layout={layout:#x} batch={len(ns)}", flush=True)
    x, p = _run_ragged(gpu, flat, offs, ns32, poffs, flags | layout)

    # untouched gaps
    mask = np.ones(flat.size, bool)
    for o, n in zip(offs, ns32):
        mask[o:o + n * n] = False
    assert _same_bits(x[mask], flat[mask]), "ragged kernel wrote outside its matrices"

    kind = {"gepp": "gepp", "none": "dominant"}.get(mode, "ref")
    gate = acc.RatioGate(f"ragged mode={mode} layout={layout:#x}", kind)
    for n in np.unique(ns32):
        sel = np.nonzero(ns32 == n)[0]
        a = np.stack([mats[i] for i in sel])
        print(f"[case] ragged mode={mode} layout={layout:#x} plain n={n} count={len(sel)}", flush=True)
        xp, pp = gpu.invert(a, flags | _plain_layout(int(n), layout))
        xr = np.stack([x[offs[i]:offs[i] + n * n].reshape(n, n) for i in sel])
        pr = np.stack([p[poffs[i]:poffs[i] + n] for i in sel])
        assert _same_bits(xr, xp), f"n={n}: ragged inverse differs from the plain launch"
        assert (pr == pp).all(), f"n={n}: ragged pivots differ from the plain launch"
        if rule == oracle.PIV_GEPP:
            xo, po, _ = oracle.gepp_inverse_f32(a)
        else:
            xo, po = oracle.lu_inverse_f32(a, rule)
        assert (pr == po).all(), f"n={n}: pivots differ from the oracle"
        e = acc.rel_err64(a, xr)
        e_orc = acc.rel_err64(a, xo)
        bound = acc.contract_bound(a, e_orc)
        assert np.all(e <= bound), f"n={n}: accuracy contract"
        gate.add(e, e_orc)
    gate.check()


def test_ragged_invalid_sizes_are_skipped_and_counted(gpu):
    rng = np.random.RandomState(3)
    good = _mixed("parallel", [5, 17, 32, 2], seed=77)
    mats = [good[0], np.ones((1, 1), np.float32), good[1], good[2], good[3]]
    flat, offs, ns, poffs = _pack_scattered(mats, rng)
    ns = ns.copy()
    ns[1] = 0          # n = 0
    bad_ns = ns.copy()
    print("[case] ragged invalid n", flush=True)
    A = gpu.DeviceBuffer.from_array(flat)
    extra = np.array([33, -4], np.int32)
    ns_all = np.concatenate([bad_ns, extra])
    offs_all = np.concatenate([offs, np.array([0, 0], np.int64)])
    O = gpu.DeviceBuffer.from_array(offs_all)
    NS = gpu.DeviceBuffer.from_array(ns_all)
    with gpu.RaggedPlan(64) as plan:
        try:
            gpu.lu_batched_inverse_ragged(A, O, NS, len(ns_all), 0x1, plan=plan)
            gpu.synchronize()
            assert plan.invalid_count() == 3
            x = A.download(flat.shape, np.float32)
        finally:
            for b in (A, O, NS):
                b.close()
    assert x[offs[1]] == 1.0, "a matrix with n = 0 was written"
    for k, i in enumerate([0, 2, 3, 4]):
        n = mats[i].shape[0]
        xp = gpu.invert(mats[i][None], 0x1, return_pivots=False)[0]
        assert _same_bits(x[offs[i]:offs[i] + n * n].reshape(n, n), xp), f"matrix {i}"


def test_ragged_large_batch_planning(gpu):
    """2.2M small matrices (n = 1..3) with invalid sizes sprinkled in.  Past
    2M matrices each lane of the planning scan walks more than 8 chunks of
    1024 (its general path); every class must still land in its slots:
    results and pivots equal the plain launch of each n bit for bit, the gaps
    left by invalid entries stay untouched, and the invalid entries are
    counted."""
    rng = np.random.RandomState(6)
    batch = 2_200_000
    ns = rng.randint(1, 4, size=batch).astype(np.int32)
    bad = rng.choice(batch, size=1000, replace=False)
    sizes = ns.astype(np.int64) ** 2
    offs = np.zeros(batch, np.int64)
    offs[1:] = np.cumsum(sizes)[:-1]
    poffs = np.zeros(batch, np.int64)
    poffs[1:] = np.cumsum(ns[:-1].astype(np.int64))
    flat = rng.uniform(-1, 1, size=int(sizes.sum())).astype(np.float32)
    ns_run = ns.copy()
    ns_run[bad] = np.where(np.arange(bad.size) % 2 == 0, 0, 40)
    total = int(ns.sum())
    print(f"[case] ragged large batch={batch} invalid={bad.size}", flush=True)
    A = gpu.DeviceBuffer.from_array(flat)
    O, NS, PO = (gpu.DeviceBuffer.from_array(v) for v in (offs, ns_run, poffs))
    P = gpu.DeviceBuffer(total * 4)
    with gpu.RaggedPlan(batch) as plan:
        try:
            P.fill_bytes(0xFF)
            gpu.lu_batched_inverse_ragged(A, O, NS, batch, 0x1, P, PO, plan=plan)
            gpu.synchronize()
            assert plan.invalid_count() == bad.size
            x = A.download(flat.shape, np.float32)
            p = P.download((total,), np.int32)
        finally:
            for b in (A, O, NS, PO, P):
                b.close()
    skipped = np.zeros(batch, bool)
    skipped[bad] = True
    for n in range(1, 4):
        sel = np.nonzero((ns == n) & ~skipped)[0]
        idx = offs[sel][:, None] + np.arange(n * n)[None, :]
        a = flat[idx].reshape(-1, n, n)
        xp, pp = gpu.invert(a, 0x1)
        assert _same_bits(x[idx].reshape(-1, n, n), xp), f"n={n}: ragged inverse differs from the plain launch"
        pidx = poffs[sel][:, None] + np.arange(n)[None, :]
        assert (p[pidx] == pp).all(), f"n={n}: ragged pivots differ from the plain launch"
        gone = np.nonzero((ns == n) & skipped)[0]
        gidx = offs[gone][:, None] + np.arange(n * n)[None, :]
        assert _same_bits(x[gidx], flat[gidx]), f"n={n}: a skipped matrix was written"


def test_ragged_without_pivots_and_single_class(gpu):
    rng = np.random.RandomState(4)
    for n in (1, 4, 13, 32):
        mats = _mixed("parallel", [n] * 70, seed=900 + n)
        flat, offs, ns, poffs = _pack_scattered(mats, rng)
        print(f"[case] ragged single class n={n} no pivots", flush=True)
        x, _ = _run_ragged(gpu, flat, offs, ns, poffs, 0x1, with_piv=False)
        xp = gpu.invert(np.stack(mats), 0x1, return_pivots=False)
        xr = np.stack([x[o:o + n * n].reshape(n, n) for o in offs])
        assert _same_bits(xr, xp)


def test_ragged_grab_table_does_not_change_results(gpu):
    rng = np.random.RandomState(5)
    ns = rng.randint(1, 33, size=400)
    mats = _mixed("parallel", ns, seed=321)
    flat, offs, ns32, poffs = _pack_scattered(mats, rng)
    print("[case] ragged default tuning", flush=True)
    x0, p0 = _run_ragged(gpu, flat, offs, ns32, poffs, 0x1)
    layout, grab = gpu.tuning_get()
    try:
        g2 = grab.copy()
        g2[1:] = rng.randint(1, 9, size=32)
        l2 = layout.copy()
        l2[1:] = rng.randint(0, 2, size=32)
        gpu.tuning_set(l2, g2)
        print(f"[case] ragged tuning grab={g2[1:].tolist()} layout={l2[1:].tolist()}", flush=True)
        x1, p1 = _run_ragged(gpu, flat, offs, ns32, poffs, 0x10)
        x2, p2 = _run_ragged(gpu, flat, offs, ns32, poffs, 0x1)
    finally:
        gpu.tuning_set(layout, grab)
    x3, p3 = _run_ragged(gpu, flat, offs, ns32, poffs, 0x10)
    assert _same_bits(x1, x3) and (p1 == p3).all(), "grab changed row-layout results"
    assert (p2 == p0).all(), "layout choice changed pivots"
    for i, n in enumerate(ns32):
        e = acc.rel_err64(mats[i][None], x2[offs[i]:offs[i] + n * n].reshape(1, n, n))
        assert e[0] <= acc.contract_bound(mats[i][None], acc.rel_err64(
            mats[i][None], oracle.lu_inverse_f32(mats[i][None], oracle.PIV_FIRSTMAX)[0]))[0]


def test_ragged_graph_replays_match_eager(gpu):
    """Capture (restore input; ragged inverse) on a plan once, replay it three
    times with eager work and synchronisations in between; every replay must
    reproduce the eager result bit for bit."""
    rng = np.random.RandomState(6)
    ns = rng.randint(1, 33, size=500)
    mats = _mixed("parallel", ns, seed=4242)
    flat, offs, ns32, poffs = _pack_scattered(mats, rng)
    print("[case] ragged eager reference", flush=True)
    x_ref, p_ref = _run_ragged(gpu, flat, offs, ns32, poffs, 0x1)

    src = gpu.DeviceBuffer.from_array(flat)
    A = gpu.DeviceBuffer(flat.nbytes)
    O = gpu.DeviceBuffer.from_array(offs)
    NS = gpu.DeviceBuffer.from_array(ns32)
    total = int(ns32.sum())
    P = gpu.DeviceBuffer(total * 4)
    PO = gpu.DeviceBuffer.from_array(poffs)
    other = oracle.uniform_matrices(9, 64, 8)
    s = gpu.Stream()
    g = gpu.Graph(s)
    with gpu.RaggedPlan(len(ns32)) as plan:
        try:
            print("[case] ragged graph capture", flush=True)
            g.begin()
            A.copy_from(src, stream=s)
            gpu.lu_batched_inverse_ragged(A, O, NS, len(ns32), 0x1, P, PO, plan=plan, stream=s)
            g.end()
            for r in range(3):
                print(f"[case] ragged graph replay {r}", flush=True)
                P.fill_bytes(0)
                g.launch()
                s.synchronize()
                x = A.download(flat.shape, np.float32)
                p = P.download((total,), np.int32)
                assert _same_bits(x, x_ref), f"replay {r}: inverses differ from eager"
                assert (p == p_ref).all(), f"replay {r}: pivots differ from eager"
                gpu.invert(other, 0x1)  # eager work between replays
        finally:
            g.close()
            s.close()
            for b in (src, A, O, NS, P, PO):
                b.close()


def test_ragged_plan_capacity_is_checked(gpu):
    mats = _mixed("parallel", [4] * 10, seed=1)
    flat, offs, ns, poffs = _pack_scattered(mats, np.random.RandomState(0))
    A = gpu.DeviceBuffer.from_array(flat)
    O = gpu.DeviceBuffer.from_array(offs)
    NS = gpu.DeviceBuffer.from_array(ns)
    print("[case] ragged plan capacity", flush=True)
    with gpu.RaggedPlan(8) as plan:
        try:
            with pytest.raises(Exception, match="EINVAL_ARG|argument"):
                gpu.lu_batched_inverse_ragged(A, O, NS, 10, 0x1, plan=plan)
        finally:
            for b in (A, O, NS):
                b.close()


def test_gather_scatter_utilities(gpu):
    """lubx_gather_f32 / lubx_scatter_f32 (the per-n baseline's data movement)
    against numpy, for every n, at odd offsets with gaps between matrices."""
    from matrixinversion_amd._native import check, lib

    rng = np.random.RandomState(8)
    for n in (1, 2, 3, 5, 8, 13, 31, 32):
        count = 97
        mats = [rng.uniform(-1, 1, (n, n)).astype(np.float32) for _ in range(count)]
        flat, offs, _, _ = _pack_scattered(mats, rng)
        print(f"[case] gather/scatter n={n}", flush=True)
        P = gpu.DeviceBuffer.from_array(flat)
        O = gpu.DeviceBuffer.from_array(offs)
        C = gpu.DeviceBuffer(count * n * n * 4)
        try:
            check(lib().lubx_gather_f32(C.ptr, P.ptr, O.ptr, n * n, count, None), "gather")
            gpu.synchronize()
            c = C.download((count, n, n), np.float32)
            assert np.array_equal(c, np.stack(mats))
            C.upload(-c)
            check(lib().lubx_scatter_f32(P.ptr, C.ptr, O.ptr, n * n, count, None), "scatter")
            gpu.synchronize()
            back = P.download(flat.shape, np.float32)
            expect = flat.copy()
            for i, o in enumerate(offs):
                expect[o:o + n * n] = -mats[i].ravel()
            assert np.array_equal(back, expect), "scatter wrote outside or wrong values"
        finally:
            for b in (P, O, C):
                b.close()
