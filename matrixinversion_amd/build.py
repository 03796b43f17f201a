"""Build liblubatch.so for gfx950 in-tree (matrixinversion_amd/_lib/).

    python -m matrixinversion_amd.build [--jobs J]

Steps: generate csrc/gj_asm.inc, compile the kernel translation units (one per
size range, in parallel) and the C-ABI / utility units with hipcc, link one
shared library.  Every HIP unit is compiled with packed-FP32 code generation
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import argparse
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


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_gj_asm.py"), os.path.join(CSRC, "gj_asm.inc")],
                   check=True)
    units = []
    for lo, hi in SIZE_RANGES:
        obj = os.path.join(OBJDIR, f"lu_kernels_{lo}_{hi}.o")
        units.append([HIPCC, *CFLAGS, f"-DLUB_N_LO={lo}", f"-DLUB_N_HI={hi}", "-c",
                      os.path.join(CSRC, "lu_kernels_inst.hip"), "-o", obj])
    for piv in range(4):
        units.append([HIPCC, *CFLAGS, f"-DLUB_RAGGED_PIV={piv}", "-c", os.path.join(CSRC, "lu_ragged_inst.hip"),
                      "-o", os.path.join(OBJDIR, f"lu_ragged_{piv}.o")])
    for name in ("lub_api", "lub_util", "lub_verify", "lub_ragged"):
        units.append([HIPCC, *CFLAGS, "-c", os.path.join(CSRC, f"{name}.hip"),
                      "-o", os.path.join(OBJDIR, f"{name}.o")])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(_run, units))
    if verbose:
        for cmd, (dt, err) in zip(units, results):
            print(f"[{dt:6.1f}s] {os.path.basename(cmd[-1])}" + (f"\n{err}" if err.strip() else ""))
    objs = [u[-1] for u in units]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    t0 = time.time()
    lib = build(args.jobs, args.verbose)
    print(f"built {lib} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
