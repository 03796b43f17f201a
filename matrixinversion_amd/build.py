"""Build liblubatch.so for gfx950 in-tree (matrixinversion_amd/_lib/).

    python -m matrixinversion_amd.build [--jobs J]

Steps: generate csrc/gj_asm.inc, compile the kernel translation units (one per
size range, in parallel) and the C-ABI / utility units with hipcc, link one
shared library.  Every HIP unit is compiled with packed-FP32 code generation
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import argparse
import hashlib
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "_lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "liblubatch.so")
CMP_LIB = os.path.join(LIBDIR, "liblubref_rocsolver.so")

ARCH = "gfx950"
SIZE_RANGES = [(1, 8), (9, 16), (17, 20), (21, 24), (25, 28), (29, 32)]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd):
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is noise.
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{err}")
    return time.time() - t0, err


def source_id() -> str:
    """Hash of every source that goes into liblubatch.so (csrc/, include/):
    compiled into the library as lu_build_id(), and recorded with every
    counter profile, so a profile taken on other sources shows as stale."""
    h = hashlib.sha256()
    for d in (CSRC, INCLUDE):
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".hpp", ".h", ".py", ".cpp")):
                h.update(f.encode() + b"\0")
                with open(os.path.join(d, f), "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()[:16]


def _stale(cmd, deps_mtime: float) -> bool:
    obj = cmd[-1]
    return not os.path.exists(obj) or os.path.getmtime(obj) < deps_mtime


def build(jobs: int = 8, verbose: bool = False, incremental: bool = False) -> str:
    """Compile every unit and link; with `incremental`, units whose object is
    newer than every source and header are kept (a developer shortcut: the
    default always rebuilds)."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    inc = os.path.join(CSRC, "gj_asm.inc")
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_gj_asm.py"), inc + ".new"], check=True)
    if os.path.exists(inc) and open(inc).read() == open(inc + ".new").read():
        os.remove(inc + ".new")  # unchanged: keep its timestamp
    else:
        os.replace(inc + ".new", inc)
    units = []
    for lo, hi in SIZE_RANGES:
        obj = os.path.join(OBJDIR, f"lu_kernels_{lo}_{hi}.o")
        units.append([HIPCC, *CFLAGS, f"-DLUB_N_LO={lo}", f"-DLUB_N_HI={hi}", "-c",
                      os.path.join(CSRC, "lu_kernels_inst.hip"), "-o", obj])
    for lo, hi in ((17, 24), (25, 32)):
        units.append([HIPCC, *CFLAGS, f"-DLUB_RF_LO={lo}", f"-DLUB_RF_HI={hi}", "-c",
                      os.path.join(CSRC, "lu_refined_inst.hip"), "-o", os.path.join(OBJDIR, f"lu_refined_{lo}_{hi}.o")])
    for piv in (0, 1, 3, 4):  # the serial rule (2) runs the first-max kernel (1)
        units.append([HIPCC, *CFLAGS, f"-DLUB_RAGGED_PIV={piv}", "-c", os.path.join(CSRC, "lu_ragged_inst.hip"),
                      "-o", os.path.join(OBJDIR, f"lu_ragged_{piv}.o")])
    for name in ("lub_api", "lub_util", "lub_verify", "lub_ragged", "lub_serve", "lub_refine", "lub_lapack"):
        extra = [f'-DLUB_BUILD_ID="{source_id()}"'] if name == "lub_api" else []
        units.append([HIPCC, *CFLAGS, *extra, "-c", os.path.join(CSRC, f"{name}.hip"),
                      "-o", os.path.join(OBJDIR, f"{name}.o")])
    todo = units
    if incremental:
        deps = [os.path.join(d, f) for d in (CSRC, INCLUDE) for f in os.listdir(d)
                if f.endswith((".hip", ".hpp", ".h", ".inc"))]
        newest_header = max(os.path.getmtime(f) for f in deps if not f.endswith(".hip"))
        todo = [u for u in units
                if _stale(u, max(newest_header, os.path.getmtime(next(a for a in u if a.endswith(".hip")))))]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(_run, todo))
    units_run = todo
    if verbose:
        for cmd, (dt, err) in zip(units_run, results):
            print(f"[{dt:6.1f}s] {os.path.basename(cmd[-1])}" + (f"\n{err}" if err.strip() else ""))
    objs = [u[-1] for u in units]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_comparator()
    return LIB


def build_comparator() -> str | None:
    """The benchmark-only rocSOLVER comparator (its own library, so that
    liblubatch.so never loads rocSOLVER); skipped when rocSOLVER is absent."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    if not os.path.exists(os.path.join(rocm, "include", "rocsolver", "rocsolver.h")):
        return None
    tmp = CMP_LIB + ".tmp"
    _run([HIPCC, *CFLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "cmp_rocsolver.cpp"),
          f"-L{os.path.join(rocm, 'lib')}", "-lrocsolver", "-lrocblas", f"-Wl,-rpath,{os.path.join(rocm, 'lib')}"])
    os.replace(tmp, CMP_LIB)
    return CMP_LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="keep objects newer than their sources")
    args = ap.parse_args()
    t0 = time.time()
    lib = build(args.jobs, args.verbose, args.incremental)
    print(f"built {lib} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
