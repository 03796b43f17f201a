"""GPU tests of the launch shapes other than one workgroup per tile.

The plain kernel strides over tiles and prefetches the next full tile into
registers while it factors the current one.  That loop runs when a batch has
more than 2^22 tiles (the grid cap) and under the tuning variables
LUB_TILES_PER_WG (grid = tiles / T) and LUB_WAVES_PER_CU (a persistent grid);
LUB_RAGGED_WAVES_PER_CU makes the ragged kernel persistent.  Results must not
depend on the shape: each case is compared bit for bit with the default
launch.  The variables are read once per process, so the shaped runs happen
in a child process that sets them before its first GPU call.  Every case
prints its name first.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

# (n, flags, batch): every pivot rule's kernel, both lane layouts, partial tiles
PLAIN_CASES = [(32, 0x1, 5003), (32, 0x3, 1203), (32, 0x4, 1203), (32, 0x0, 1203), (27, 0x1, 1203),
               (16, 0x11, 2003), (16, 0x4 | 0x10, 1203), (8, 0x21, 4003), (12, 0x24, 4003), (5, 0x0, 4003)]
RAGGED_COUNT = 20_000

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[2])
import matrixinversion_amd as mia
import oracle
from tests.test_gpu_launch_shapes import PLAIN_CASES, inputs, ragged_inputs
out = {}
for n, flags, batch in PLAIN_CASES:
    print(f"[child] plain n={n} flags={flags:#x} batch={batch}", flush=True)
    x, p = mia.invert(inputs(n, flags, batch), flags)
    out[f"x{n}_{flags}"], out[f"p{n}_{flags}"] = x, p
mats = ragged_inputs()
for flags in (0x1, 0x0, 0x4):
    print(f"[child] ragged flags={flags:#x}", flush=True)
    xs, ps = mia.invert_ragged(mats, flags)
    out[f"rx{flags}"] = np.concatenate([x.ravel() for x in xs])
    out[f"rp{flags}"] = np.concatenate(ps)
np.savez(sys.argv[1], **out)
print("[child] done", flush=True)
"""


def inputs(n, flags, batch):
    if flags & 0xF == 0:
        return oracle.dominant_matrices(600 + n, batch, n)
    return oracle.uniform_matrices(600 + n, batch, n)


def ragged_inputs():
    rng = np.random.RandomState(17)
    ns = rng.randint(1, 33, size=RAGGED_COUNT)
    return [oracle.dominant_matrices(700 + i % 50, 1, int(n))[0] for i, n in enumerate(ns)]


def _child(env_extra, path):
    env = dict(os.environ)
    env.update(env_extra)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    print(f"[case] child process with {env_extra}", flush=True)
    r = subprocess.run([sys.executable, "-u", "-c", CHILD, path, ROOT], env=env, timeout=240,
                       capture_output=True, text=True)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, f"child failed with {r.returncode}"
    return np.load(path)


@pytest.mark.parametrize("env_extra", [{"LUB_TILES_PER_WG": "3"},
                                       {"LUB_WAVES_PER_CU": "1", "LUB_RAGGED_WAVES_PER_CU": "1"}],
                         ids=["tiles_per_wg_3", "persistent_1_wave_per_cu"])
def test_launch_shapes_match_default(gpu, env_extra, tmp_path):
    got = _child(env_extra, str(tmp_path / "shaped.npz"))
    for n, flags, batch in PLAIN_CASES:
        print(f"[case] default plain n={n} flags={flags:#x} batch={batch}", flush=True)
        x, p = gpu.invert(inputs(n, flags, batch), flags)
        assert np.array_equal(got[f"x{n}_{flags}"].view(np.uint32), x.view(np.uint32)), f"n={n} flags={flags:#x}"
        assert (got[f"p{n}_{flags}"] == p).all(), f"pivots n={n} flags={flags:#x}"
    mats = ragged_inputs()
    for flags in (0x1, 0x0, 0x4):
        print(f"[case] default ragged flags={flags:#x}", flush=True)
        xs, ps = gpu.invert_ragged(mats, flags)
        assert np.array_equal(got[f"rx{flags}"].view(np.uint32),
                              np.concatenate([x.ravel() for x in xs]).view(np.uint32)), f"ragged flags={flags:#x}"
        assert (got[f"rp{flags}"] == np.concatenate(ps)).all()


@pytest.mark.parametrize("flags", [0x11, 0x14], ids=["parallel", "gepp"])
def test_more_tiles_than_the_grid(gpu, flags):
    """2 * 2^22 tiles of 4 matrices + 5 matrices, n = 16 in 16-lane rows
    (33.6M matrices, 34 GB): every workgroup runs the strided loop two or
    three times, with the register prefetch of its next tile, and the batch
    ends in a partial tile.  The batch repeats a block of 16,384 matrices; windows at the start,
    across the first grid wrap (tile 2^22) and at the end equal the block's
    own inverses bit for bit."""
    from matrixinversion_amd._native import check, lib

    n, block_count = 16, 16_384
    batch = 2 * (1 << 22) * 4 + 5
    block = oracle.uniform_matrices(810, block_count, n)
    xb, pb = gpu.invert(block, flags)
    per = n * n * 4
    print(f"[case] n={n} flags={flags:#x} batch={batch} ({batch * per / 1e9:.1f} GB)", flush=True)
    A = gpu.DeviceBuffer(batch * per)
    P = gpu.DeviceBuffer(batch * n * 4)
    src = gpu.DeviceBuffer.from_array(block)
    s = gpu.Stream()
    try:
        done = 0
        while done < batch:
            k = min(block_count, batch - done)
            check(lib().lubx_memcpy_dtod_async(A.ptr + done * per, src.ptr, k * per, s.handle), "d2d")
            done += k
        gpu.lu_batched_inverse(A, n, batch, flags, P, s)
        s.synchronize()
        wrap = (1 << 22) * 4  # first matrix of tile 2^22: the first strided iteration
        for start in (0, wrap - 2048, batch - 4096):
            x = A.download((4096, n, n), np.float32, offset=start * per)
            p = P.download((4096, n), np.int32, offset=start * n * 4)
            idx = (start + np.arange(4096)) % block_count
            assert np.array_equal(x.view(np.uint32), xb[idx].view(np.uint32)), f"window at {start}"
            assert (p == pb[idx]).all(), f"pivots at {start}"
    finally:
        s.close()
        for b in (A, P, src):
            b.close()
