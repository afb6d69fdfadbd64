"""Build recipe of the gfx950 library (libgoferro.so, in-tree).

Every HIP source is compiled for gfx950 with packed-FP32 instructions off
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).  The host-side Newell
tensor file is compiled separately with ``-ffp-contract=off`` so its values
match the reference's evaluation bit for bit; device code keeps hipcc's
default contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
LIB = PKG / "libgoferro.so"
OBJDIR = PKG.parent / "_build"

HIP_SOURCES = ["gf_api.hip", "gf_llg.hip", "gf_fft.hip", "gf_fft_any.hip", "gf_fe.hip", "gf_elastic.hip", "gf_em.hip", "gf_batch.hip", "gf_adjoint.hip", "gf_fp32.hip", "gf_textio.hip", "gf_elastostatic.hip", "gf_noalloc.hip"]
HOST_SOURCES = ["gf_newell.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the gfx950 library)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout)


def build(verbose: bool = False) -> Path:
    """Compile every source and link libgoferro.so next to this file."""
    cc = hipcc()
    OBJDIR.mkdir(exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
    objs = []
    for src in HOST_SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", f"-I{INCLUDE}",
               "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        cmd = [cc, "--offload-arch=gfx950", *common,
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
