"""The C-ABI library loads and exports every symbol that include/*.h declares,
and its host-side argument checks return the documented status codes.  No
compute calls: these run without a GPU."""
import os
import re

import pytest

from tests.conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def _declared_functions(compare: bool = False):
    """Functions declared by include/*.h: the liblubatch ABI, or (compare=True)
    the comparator library's header."""
    names = []
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h") or (h == "lubatch_compare.h") != compare:
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(lu\w+|lubx_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_the_abi():
    names = _declared_functions()
    for required in ("lu_batched_inverse", "lu_batched_inverse_async", "lu_status_string"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from matrixinversion_amd import _native

    lib = _native.lib()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    for n in _native.ABI:
        assert n in _declared_functions(), f"{n} bound by ctypes but not declared in include/"


def test_comparator_exports_its_header():
    from matrixinversion_amd import _native

    if not os.path.exists(_native.CMP_LIB_PATH):
        pytest.skip("rocSOLVER comparator not built (rocSOLVER absent)")
    lib = _native.cmp_lib()
    names = _declared_functions(compare=True)
    assert names and all(hasattr(lib, n) for n in names)
    assert sorted(_native.COMPARE) == names


def test_host_side_validation_codes():
    from matrixinversion_amd import _native

    lib = _native.lib()
    fake = 0x1000  # never dereferenced: validation fails or batch == 0 first
    assert lib.lu_batched_inverse(fake, None, 0, 1, 1) == 1          # LUB_EINVAL_N
    assert lib.lu_batched_inverse(fake, None, 33, 1, 1) == 1
    assert lib.lu_batched_inverse(fake, None, 4, -1, 1) == 2         # LUB_EINVAL_BATCH
    assert lib.lu_batched_inverse(None, None, 4, 5, 1) == 3          # LUB_ENULL
    assert lib.lu_batched_inverse(fake, None, 4, 1, 0x40) == 4       # LUB_EFLAGS
    assert lib.lu_batched_inverse(fake, None, 4, 1, 0x31) == 4       # both layouts
    assert lib.lu_batched_inverse(fake, None, 17, 1, 0x21) == 4      # quads hold n <= 16
    assert lib.lu_batched_inverse(fake, None, 4, 1, 0x7) == 4
    assert lib.lu_batched_inverse(fake, None, 4, 1, 0x5) == 4       # rules end at LUB_PIVOT_GEPP (4)
    assert lib.lu_batched_inverse(fake + 2, None, 4, 1, 1) == 7      # LUB_EALIGN
    assert lib.lu_batched_inverse(None, None, 4, 0, 1) == 0          # empty batch: nothing to do
    assert lib.lu_status_string(1).decode().startswith("n outside")
    assert lib.lu_version().decode().count(".") == 2


def test_ragged_and_tuning_validation_codes():
    import ctypes

    import numpy as np

    from matrixinversion_amd import _native, api

    lib = _native.lib()
    fake = 0x1000
    h = ctypes.c_void_p()
    assert lib.lu_ragged_plan_create(None, 16) == 3                          # LUB_ENULL
    assert lib.lu_ragged_plan_create(ctypes.byref(h), -1) == 2               # LUB_EINVAL_BATCH
    assert lib.lu_ragged_plan_create(ctypes.byref(h), 2**31) == 2            # int32 indices on the device
    assert lib.lu_ragged_plan_destroy(None) == 3
    assert lib.lu_batched_inverse_ragged_plan(None, fake, fake, fake, None, None, 4, 1, None) == 3
    assert lib.lu_batched_inverse_ragged(fake, fake, fake, None, None, -1, 1, None) == 2
    assert lib.lu_batched_inverse_ragged(None, None, None, None, None, 0, 1, None) == 0
    assert lib.lu_batched_inverse_ragged(fake, fake, fake, None, None, 0, 0x31, None) == 0  # empty first
    h2 = ctypes.c_void_p(0x2000)  # flags are checked before the plan is touched
    assert lib.lu_batched_inverse_ragged_plan(h2, fake, fake, fake, None, None, 4, 0x31, None) == 4
    assert lib.lu_batched_inverse_ragged_plan(h2, fake, fake, fake, None, None, 4, 0x40, None) == 4
    assert lib.lu_batched_inverse_ragged_plan(h2, fake, fake, fake, None, None, -2, 1, None) == 2
    assert lib.lu_ragged_invalid_count(None, None) == 3

    layout, grab = api.tuning_get()
    assert layout.shape == (33,) and grab.shape == (33,)
    assert set(np.unique(layout)) <= {0, 1} and (grab >= 1).all()
    bad = grab.copy()
    bad[5] = 0
    assert lib.lu_tuning_set(layout.tobytes(), bad.tobytes()) == 8         # LUB_EINVAL_ARG
    bad = layout.copy()
    bad[3] = 2
    assert lib.lu_tuning_set(bad.tobytes(), grab.tobytes()) == 8
    assert lib.lu_tuning_set(None, grab.tobytes()) == 3
    try:
        g2 = grab.copy()
        g2[7] = 5
        api.tuning_set(grab=g2)
        assert api.tuning_get()[1][7] == 5
    finally:
        api.tuning_set(layout, grab)
    assert (api.tuning_get()[1] == grab).all()

    pad = api.lds_pad_get()
    assert pad.shape == (33,) and pad.tolist() == [1 if (n % 4 == 0 and n not in (0, 28)) else 0 for n in range(33)]
    bad = pad.copy()
    bad[5] = 1                                                         # only n % 4 == 0 can be padded
    assert lib.lu_tuning_set_lds_pad(bad.tobytes()) == 8
    bad = pad.copy()
    bad[8] = 2
    assert lib.lu_tuning_set_lds_pad(bad.tobytes()) == 8
    assert lib.lu_tuning_set_lds_pad(None) == 3 and lib.lu_tuning_get_lds_pad(None) == 3
    try:
        p2 = pad.copy()
        p2[12] = 1
        p2[32] = 0
        api.lds_pad_set(p2)
        assert api.lds_pad_get()[12] == 1 and api.lds_pad_get()[32] == 0
        assert (api.tuning_get()[1] == grab).all()                     # layout/grab untouched
    finally:
        api.lds_pad_set(pad)


def test_refine_validation_codes():
    from matrixinversion_amd import _native

    lib = _native.lib()
    fake = 0x1000
    assert lib.lu_batched_refine(fake, fake, 0, 1, None) == 1
    assert lib.lu_batched_refine(fake, fake, 33, 1, None) == 1
    assert lib.lu_batched_refine(fake, fake, 4, -1, None) == 2
    assert lib.lu_batched_refine(None, fake, 4, 1, None) == 3
    assert lib.lu_batched_refine(fake + 2, fake, 4, 1, None) == 7
    assert lib.lu_batched_refine(None, None, 4, 0, None) == 0
    assert lib.lu_batched_inverse_refined(fake, fake, None, 4, 1, 1, None) == 8   # A == X


def test_server_validation_codes():
    import ctypes

    from matrixinversion_amd import _native

    lib = _native.lib()
    h = ctypes.c_void_p()
    t = ctypes.c_uint64()
    assert lib.lu_server_create(None, 1 << 20, 16, 16) == 3
    assert lib.lu_server_create(ctypes.byref(h), 100, 16, 16) == 8        # budget below one block
    assert lib.lu_server_create(ctypes.byref(h), 1 << 20, 16, 8) == 8     # ring smaller than a batch
    assert lib.lu_server_destroy(None) == 3
    assert lib.lu_server_submit(None, 0x1000, 0x1000, 1, 1, ctypes.byref(t)) == 3
    assert lib.lu_server_progress(None, None, None) == 3
    assert lib.lu_server_release(None, 1) == 3
    assert lib.lu_server_poll(None, None, 0, None) == 3
    cnt = ctypes.c_int()
    assert lib.lu_server_poll(None, None, 0, ctypes.byref(cnt)) == 3
    assert lib.lu_server_counters(None, None, None) == 3
    assert lib.lu_status_string(9).decode().startswith("no room")
    assert lib.lu_status_string(10).decode().startswith("batch above")


def test_build_id_is_the_source_hash():
    """lu_build_id() of the loaded library is the hash of the sources in this
    tree (so counter profiles can be matched to the code they measured)."""
    from matrixinversion_amd import _native, build

    assert _native.lib().lu_build_id().decode() == build.source_id()


def test_gather_scatter_validation_codes():
    from matrixinversion_amd import _native

    lib = _native.lib()
    fake = 0x1000
    assert lib.lubx_gather_f32(fake, fake, fake, 0, 4, None) == 8
    assert lib.lubx_gather_f32(None, fake, fake, 4, 4, None) == 3
    assert lib.lubx_gather_f32(fake + 4, fake, fake, 4, 4, None) == 7
    assert lib.lubx_scatter_f32(fake, fake + 4, fake, 4, 4, None) == 7
    assert lib.lubx_scatter_f32(None, None, None, 4, 0, None) == 0


def test_pack_ragged_host_layout():
    import numpy as np

    from matrixinversion_amd import api

    mats = [np.full((n, n), n, np.float32) for n in (3, 1, 32, 2)]
    flat, offs, ns, poffs = api.pack_ragged(mats)
    assert ns.tolist() == [3, 1, 32, 2]
    assert offs.tolist() == [0, 9, 10, 1034]
    assert poffs.tolist() == [0, 3, 4, 36]
    assert flat.size == 1038 and flat[9] == 1 and flat[1034] == 2


def test_no_cpu_fallback_in_product():
    """The product path never reaches the oracle."""
    pkg = os.path.join(ROOT, "matrixinversion_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text, f
