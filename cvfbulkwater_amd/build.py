"""Build the engine's native artefacts in-tree.

    libcvf.so      HIP kernels + C ABI (include/cvf_api.h), gfx950
    cvf_water3d    drop-in executable (input_data/input_config -> data.out ...)
    cvf_thermo     drop-in for the thermodynamic analysis scripts (data.out -> isobars)
    cvf_reweight   drop-in for the histogram reweighting program (data_L* -> Binder, Confront, ...)

Run ``python -m cvfbulkwater_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcvf.so"
EXE = PKG / "cvf_water3d"

ARCH = os.environ.get("CVF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 VALU forms; -ffp-contract=off: no fused
# multiply-add anywhere, so host and device FP64 volume chains agree bit for bit.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function",
]

LIB_SOURCES = ["cvf_kernels.hip", "cvf_resident.hip", "cvf_rw_kernels.hip", "cvf_noalloc.hip", "cvf_api.cpp", "cvf_host.cpp", "cvf_corr.cpp",
               "cvf_thermo.cpp", "cvf_reweight.cpp"]
# executable -> sources (each linked against libcvf.so)
EXES = {"cvf_water3d": ["cvf_water3d.cpp"], "cvf_thermo": ["cvf_thermo_main.cpp"],
        "cvf_reweight": ["cvf_reweight_main.cpp"]}


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _compile(src: Path, obj: Path) -> None:
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    _run([HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(obj)])


def build(verbose: bool = False) -> Path:
    """Compile every HIP/C++ source for gfx950 and link libcvf.so + cvf_water3d."""
    BUILD.mkdir(exist_ok=True)
    objs = []
    for name in LIB_SOURCES:
        src = CSRC / name
        obj = BUILD / (src.stem + ".o")
        if verbose:
            print("hipcc", name, flush=True)
        _compile(src, obj)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    for exe, sources in EXES.items():
        exe_objs = []
        for name in sources:
            src = CSRC / name
            obj = BUILD / (src.stem + ".o")
            _compile(src, obj)
            exe_objs.append(str(obj))
        _run([HIPCC, f"--offload-arch={ARCH}", "-o", str(PKG / exe), *exe_objs,
              f"-L{PKG}", "-lcvf", f"-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
