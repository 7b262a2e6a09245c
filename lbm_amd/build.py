"""Builds lbm_amd/liblbm_hip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Each translation unit is compiled separately and in parallel; the EXACT64
unit is built without FMA contraction because the fp64 parity mode must keep
the reference's separately rounded products and sums.

    python -m lbm_amd.build [--force] [-j N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJDIR = PKG / "_build"
LIB = PKG / "liblbm_hip.so"

ARCH = os.environ.get("LBM_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function",
]

# translation unit -> extra flags. Kernels are built without implicit FMA
# contraction: every product and sum is rounded as written, so a cell's result
# does not depend on how many cells a thread handles or which kernel (single,
# batched) handles it.
NOFMA = ["-ffp-contract=off"]
UNITS = {
    "capi.hip": [],
    "capi_noalloc.hip": [],
    "capi_adjoint.hip": [],
    "mode_adjoint.hip": NOFMA,
    "capi_batch.hip": [],
    "mode_batch.hip": NOFMA,
    "mode_batch_adjoint.hip": NOFMA,
    "mode_exact.hip": NOFMA,
    "mode_fast_q15.hip": NOFMA,
    "mode_fast_q19.hip": NOFMA,
    "mode_fast_q27.hip": NOFMA,
}

HEADERS = sorted(CSRC.glob("*.hpp")) + sorted(INCLUDE.rglob("*.h*"))


def _stale(obj: Path, src: Path) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(p.stat().st_mtime > t for p in [src, *HEADERS])


def _run(cmd: list[str]) -> tuple[int, str]:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    return p.returncode, out


def build(force: bool = True, jobs: int | None = None, verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    todo = []
    for unit, extra in UNITS.items():
        src = CSRC / unit
        obj = OBJDIR / (unit + ".o")
        if force or _stale(obj, src):
            todo.append([HIPCC, *COMMON, *extra, "-c", str(src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_run, todo))
    for cmd, (rc, out) in zip(todo, results):
        if verbose or rc:
            print(" ".join(cmd), file=sys.stderr)
            if out.strip():
                print(out, file=sys.stderr)
        if rc:
            raise RuntimeError(f"hipcc failed for {cmd[-3]}")
    objs = [str(OBJDIR / (u + ".o")) for u in UNITS]
    if todo or force or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        rc, out = _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs])
        if rc:
            print(out, file=sys.stderr)
            raise RuntimeError("link of liblbm_hip.so failed")
        os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER = PKG / "lbm"


def build_driver() -> Path:
    """The drop-in `lbm` executable (host C++ over the C ABI, built with g++)."""
    src = PKG / "apps" / "lbm_main.cpp"
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", str(src),
           f"-L{PKG}", "-llbm_hip", "-Wl,-rpath,$ORIGIN", "-o", str(DRIVER)]
    rc, out = _run(cmd)
    if rc:
        print(" ".join(cmd), file=sys.stderr)
        print(out, file=sys.stderr)
        raise RuntimeError("build of the lbm driver failed")
    return DRIVER


def build_oracle() -> Path:
    """CPU oracle (test infrastructure) with its own Makefile."""
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return ROOT / "oracle" / "liblbm_oracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j, verbose=a.v))


if __name__ == "__main__":
    main()
