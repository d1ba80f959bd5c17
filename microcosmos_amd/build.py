"""Builds the HIP library ``libmicrocosmos_hip.so`` for gfx950 with hipcc.

Every kernel is compiled with -ffp-contract=off (each source-level float
operation rounds once, in the written order) and with packed-FP32 instruction
selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "libmicrocosmos_hip.so"
SOURCES = ["mc_api.hip", "mc_lbm.hip", "mc_ib.hip", "mc_rod.hip", "mc_steric.hip", "mc_track.hip",
           "mc_adjoint.hip", "mc_losses.hip", "mc_record.hip",
           "mc_fluid_adj.hip"]
ARCH = os.environ.get("MC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def lib_path() -> Path:
    return LIBDIR / LIBNAME


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library (always rebuilds)."""
    obj = PKG / "_build"
    obj.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)

    def compile_one(src):
        out = obj / (src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(out)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
        if verbose and msg.strip():
            print(msg, file=sys.stderr)
        return out

    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == "__main__":
    print(build(verbose=True))
