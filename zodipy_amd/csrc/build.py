"""Build the gfx950 shared library ``zodipy_amd/libzodi_hip.so`` with hipcc.

Usage: ``python -m zodipy_amd.csrc.build`` (also called by ``__graft_entry__.build()``).
The library is built in-tree so it travels with the source snapshot.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB = PKG / "libzodi_hip.so"
SOURCES = [CSRC / "zodi_forward.hip", CSRC / "zodi_adjoint.hip", CSRC / "zodi_adjoint_generic.hip",
           CSRC / "zodi_table_adjoint.hip", CSRC / "zodi_pointing.hip", CSRC / "zodi_spectral.hip"]
HEADERS = [
    CSRC / "zodi_adjoint.hpp",
    CSRC / "zodi_common.hpp",
    CSRC / "zodi_device.hpp",
    CSRC / "zodi_grad.hpp",
    CSRC / "zodi_math.hpp",
    CSRC / "zodi_trig.hpp",
    CSRC / "zodi_trig_coeffs.inc",
    INCLUDE / "zodi_hip.h",
]

ARCH = os.environ.get("ZODI_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # Disable packed FP32 instruction selection (all arithmetic here is FP64).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    if verbose and proc.stderr.strip():
        sys.stderr.write(proc.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into the in-tree shared library; return its path.

    Each source compiles to an object of its own, in parallel, in a temporary
    directory; one link step then writes the library.
    """
    compile_flags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="zodi_build_") as tmpdir:
        objs = [Path(tmpdir) / (src.stem + ".o") for src in SOURCES]
        cmds = [[hipcc(), *compile_flags, "-c", f"-I{INCLUDE}", f"-I{CSRC}", str(src), "-o",
                 str(obj)] for src, obj in zip(SOURCES, objs)]
        workers = max(1, min(len(cmds), os.cpu_count() or 1))
        with concurrent.futures.ThreadPoolExecutor(workers) as pool:
            for fut in [pool.submit(_run, c, verbose) for c in cmds]:
                fut.result()
        tmp = LIB.with_suffix(".so.tmp")
        _run([hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *map(str, objs), "-o",
              str(tmp)], verbose)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
