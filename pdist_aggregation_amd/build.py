"""Build libpdist_knn.so (the HIP kernels and the C ABI) for gfx950, in-tree,
and libpdist_probe.so (the matrix-core rate probe bench.py reports beside the
datasheet peak; not part of the C ABI).

Usage: ``python -m pdist_aggregation_amd.build`` (or call :func:`build`).
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libpdist_knn.so"
PROBE = PKG / "probe" / "mfma_ceiling.hip"
PROBE_LIB = PKG / "libpdist_probe.so"
ARCH = os.environ.get("PDIST_KNN_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # Packed FP32 VALU forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpdist_knn.so)")


def _compile(srcs, out: pathlib.Path, verbose: bool) -> None:
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(CSRC), "-I", str(PKG.parent / "include"),
           *map(str, srcs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, out)


def build(verbose: bool = False) -> pathlib.Path:
    """Compile every csrc/*.hip into one shared library next to this file, and
    the rate probe into its own."""
    _compile(sources(), LIB, verbose)
    _compile([PROBE], PROBE_LIB, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
