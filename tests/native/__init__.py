"""CPU test probe of the engine's shared core (built by build_probe())."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "_build" / "libcvf_core_probe.so"


def build_probe() -> Path:
    (HERE / "_build").mkdir(exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           f"-I{ROOT / 'cvfbulkwater_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
           str(HERE / "cvf_core_probe.cpp"), str(ROOT / "cvfbulkwater_amd" / "csrc" / "cvf_host.cpp"),
           "-o", str(OUT)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    return OUT


_lib = None


def probe() -> C.CDLL:
    global _lib
    if _lib is None:
        srcs = [HERE / "cvf_core_probe.cpp", ROOT / "cvfbulkwater_amd" / "csrc" / "cvf_core.h",
                ROOT / "cvfbulkwater_amd" / "csrc" / "cvf_host.cpp", ROOT / "cvfbulkwater_amd" / "csrc" / "cvf_host.h"]
        if not OUT.exists() or any(f.stat().st_mtime > OUT.stat().st_mtime for f in srcs):
            build_probe()
        L = C.CDLL(str(OUT))
        U32P, U64P, DP, FP, I32P = (C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                    C.POINTER(C.c_float), C.POINTER(C.c_int32))
        sig = {
            "probe_philox": (None, [U32P, U32P, U32P]),
            "probe_philox_site": (None, [U32P, U32P, U32P]),
            "probe_rocrand_philox": (None, [U32P, U32P, U32P]),
            "probe_det_exp": (C.c_double, [C.c_double]),
            "probe_det_log": (C.c_double, [C.c_double]),
            "probe_count_eq": (C.c_int, [C.c_uint32, C.c_uint32]),
            "probe_sol_mask": (C.c_uint32, [C.c_uint32]),
            "probe_corner_edge": (C.c_int, [C.c_int, C.c_int]),
            "probe_edge_start": (C.c_int, [C.c_int]),
            "probe_edge_axis": (C.c_int, [C.c_int]),
            "probe_fdiv": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "probe_uniform_int": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "probe_u53": (C.c_double, [C.c_uint32, C.c_uint32]),
            "probe_shell_table": (C.c_int, [C.c_int, C.c_float, FP, I32P, C.c_int]),
            "probe_thresholds": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, U64P, U64P,
                                        U64P, DP]),
            "probe_vdw_fp64": (C.c_double, [C.c_double, C.c_float, FP, I32P, C.c_int]),
            "probe_volume_chain": (None, [C.c_uint64, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int64,
                                          C.c_int64, DP]),
        }
        for n, (res, args) in sig.items():
            fn = getattr(L, n)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
