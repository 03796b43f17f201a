"""The C-ABI library loads and exports every symbol that include/*.h declares,
and its host-side argument checks return the documented status codes.  No
compute calls: these run without a GPU."""
import os
import re

import pytest

from tests.conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = []
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
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
    assert lib.lu_batched_inverse(fake + 2, None, 4, 1, 1) == 7      # LUB_EALIGN
    assert lib.lu_batched_inverse(None, None, 4, 0, 1) == 0          # empty batch: nothing to do
    assert lib.lu_status_string(1).decode().startswith("n outside")
    assert lib.lu_version().decode().count(".") == 2


def test_no_cpu_fallback_in_product():
    """The product path never reaches the oracle."""
    pkg = os.path.join(ROOT, "matrixinversion_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text, f
