"""Builds the HIP engine library ``libaerhip.so`` in-tree for gfx950.

Usage: ``python -m qiskit_aer_amd.build`` (also called by the repository's
``__graft_entry__.build()``).  hipcc cross-compiles without a GPU present.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "lib" / "libaerhip.so"
OBJ = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AERHIP_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code.  MFMA results
# are kept in ordinary VGPRs (no AGPR copies, which would halve occupancy).
# SimplifyCFG's sinking of common code is off: it merges the per-register-bit
# gate variants into one path with computed register offsets, which forces
# the register-resident amplitudes into scratch memory.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-pthread",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-mllvm", "-simplifycfg-sink-common=false",
    f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-pthread", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        l for l in (r.stdout + r.stderr).splitlines()
        if "is not a recognized feature for this target" not in l
    )
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose: bool = True) -> pathlib.Path:
    """Compiles every HIP source of the engine and links ``libaerhip.so``."""
    if not shutil.which(HIPCC) and not pathlib.Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(parents=True, exist_ok=True)
    OUT.parent.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OBJ / (src.stem + ".hip.o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)], verbose)
        objs.append(str(obj))
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJ / (src.stem + ".cpp.o")
        _run(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)], verbose)
        objs.append(str(obj))
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    build_controller_wrappers(verbose)
    return OUT


def build_controller_wrappers(verbose: bool = True) -> pathlib.Path:
    """The controller_wrappers-compatible pybind11 module, linked against
    libaerhip.so (found next to it through an $ORIGIN run path)."""
    import sysconfig

    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"controller_wrappers{ext}"
    tmp = out.with_name(out.name + ".tmp")
    _run(["g++", *CXX_FLAGS, "-shared", f"-I{pybind11.get_include()}",
          f"-I{sysconfig.get_paths()['include']}", str(PKG / "pybind" / "controller_wrappers.cpp"),
          f"-L{OUT.parent}", "-laerhip", "-Wl,-rpath,$ORIGIN/lib", "-o", str(tmp)], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose="-q" not in sys.argv)
    print(p)
