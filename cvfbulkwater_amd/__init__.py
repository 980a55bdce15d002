"""MI355X-native CVF bulk-water Monte Carlo engine (Python binding).

A thin ctypes mirror of the C ABI in ``include/cvf_api.h`` (libcvf.so: HIP
kernels for gfx950 + C++ host code).  The reference (lcoronas/CVFBulkWater)
has no library interface -- its hot path is the body of the ``gpu_water3D``
executable -- so this binding mirrors the reference's *file contract*
operations: set-up from ``input_data`` values, ``input_config`` load,
``MonteCarlo_Step``, ``measure``, cluster statistics and ``config.out``.

The native library is required: importing works without it, but every
operation raises :class:`CVFError` if ``libcvf.so`` is missing.  There is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

__all__ = [
    "CVFError", "Params", "Obs", "ClusterStats", "Simulation", "load_library",
    "FP64", "FP32_COMPENSATED", "VOLUME_HOST", "VOLUME_DEVICE", "vdw_energy", "vdw_pair_energy",
    "LIB_PATH", "EXE_PATH", "THERMO_EXE_PATH", "Correlator", "KT_NOT_VALUE", "thermo_averages",
    "thermo_derivative", "thermo_kt", "REWEIGHT_EXE_PATH", "SCHED_REFERENCE", "SCHED_FAST",
]

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcvf.so"
EXE_PATH = PKG / "cvf_water3d"
THERMO_EXE_PATH = PKG / "cvf_thermo"
REWEIGHT_EXE_PATH = PKG / "cvf_reweight"

FP64, FP32_COMPENSATED = 0, 1
VOLUME_HOST, VOLUME_DEVICE = 0, 1
SCHED_REFERENCE, SCHED_FAST = 0, 1


class CVFError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [
        ("L", C.c_int32), ("P", C.c_float), ("T", C.c_float), ("seed", C.c_uint64),
        ("n_metropolis", C.c_int32), ("n_cluster", C.c_int32), ("r_cutoff", C.c_float),
        ("flag_config", C.c_int32), ("flag_chessboard", C.c_int32), ("device_id", C.c_int32),
        ("J_hb", C.c_float), ("J_sig", C.c_float), ("v_hb", C.c_float),
        ("precision", C.c_int32), ("volume_mode", C.c_int32), ("schedule", C.c_int32),
    ]


class _Obs(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("E", C.c_double), ("H", C.c_double), ("V", C.c_double),
        ("n_hb", C.c_double), ("m", C.c_double), ("r_cell", C.c_double), ("E_vdw", C.c_double),
        ("nhb2", C.c_int64), ("n_sigma", C.c_int64),
    ]


class _Tuning(C.Structure):
    _fields_ = [("streams", C.c_int32), ("block_sigma", C.c_int32), ("block_eta", C.c_int32),
                ("block_cell", C.c_int32), ("graph_steps", C.c_int32), ("fused", C.c_int32),
                ("slab_planes", C.c_int32), ("block_chunk", C.c_int32), ("chunk_sweeps", C.c_int32)]


class _CStat(C.Structure):
    _fields_ = [
        ("largest", C.c_int32), ("second", C.c_int32), ("n_perc", C.c_int32),
        ("n_clusters", C.c_int32), ("mean", C.c_float), ("var", C.c_float),
    ]


@dataclass
class Params:
    """input_data fields (gpu_water3D.cu:292-386) + lattice size and modes."""
    L: int = 32
    P: float = 0.45
    T: float = 0.815
    seed: int = 44648604
    n_metropolis: int = 1
    n_cluster: int = 0
    r_cutoff: float = 6.0
    flag_config: int = 0
    flag_chessboard: int = 1
    device_id: int = 0
    J_hb: float = 0.5
    J_sig: float = 0.08
    v_hb: float = 0.6
    precision: int = FP64
    volume_mode: int = VOLUME_HOST
    schedule: int = SCHED_REFERENCE

    def _c(self) -> _Params:
        return _Params(self.L, self.P, self.T, self.seed, self.n_metropolis, self.n_cluster,
                       self.r_cutoff, self.flag_config, self.flag_chessboard, self.device_id,
                       self.J_hb, self.J_sig, self.v_hb, self.precision, self.volume_mode, self.schedule)


@dataclass
class Obs:
    step: int
    E: float
    H: float
    V: float
    n_hb: float
    m: float
    r_cell: float
    E_vdw: float
    nhb2: int
    n_sigma: int


@dataclass
class ClusterStats:
    largest: int
    second: int
    n_perc: int
    n_clusters: int
    mean: float
    var: float


_LIB = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libcvf.so (raises CVFError if it has not been built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise CVFError(f"{p} not found: build it with `python -m cvfbulkwater_amd.build`")
    lib = C.CDLL(str(p))
    P, I32, I64, U8P, I32P, F32P, DP = (C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_uint8),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                       C.POINTER(C.c_double))
    sig = {
        "cvf_params_default": (None, [C.POINTER(_Params)]),
        "cvf_create": (C.c_int, [C.POINTER(_Params), C.POINTER(P)]),
        "cvf_create_batch": (C.c_int, [C.POINTER(_Params), I32, C.POINTER(P)]),
        "cvf_set_config": (C.c_int, [P, I32, U8P, U8P]),
        "cvf_get_config": (C.c_int, [P, I32, U8P, U8P]),
        "cvf_step": (C.c_int, [P, I64]),
        "cvf_observe": (C.c_int, [P, I32, C.POINTER(_Obs)]),
        "cvf_cluster_stats": (C.c_int, [P, I32, C.POINTER(_CStat)]),
        "cvf_get_labels": (C.c_int, [P, I32, I32P]),
        "cvf_get_volume": (C.c_int, [P, I32, DP, DP, DP]),
        "cvf_get_shells": (C.c_int, [P, I32, F32P, I32P, I32, I32P]),
        "cvf_set_shells": (C.c_int, [P, I32, F32P, I32P, I32]),
        "cvf_info": (C.c_int, [P, I32, I32P, I32P, C.POINTER(C.c_int64)]),
        "cvf_get_stream": (C.c_int, [P, C.POINTER(C.c_void_p)]),
        "cvf_sync": (C.c_int, [P]),
        "cvf_last_error": (C.c_char_p, []),
        "cvf_destroy": (None, [P]),
        "cvf_vdw_energy": (C.c_int, [I32, I32, F32P, I32P, I32, C.c_float, DP, I32, I32, DP]),
        "cvf_profile_steps": (C.c_int, [P, I64, DP, C.POINTER(C.c_int64), I32]),
        "cvf_get_tuning": (C.c_int, [P, C.POINTER(_Tuning)]),
        "cvf_autotune": (C.c_int, [P, I32, C.POINTER(_Tuning)]),
        "cvf_sigma_histogram": (C.c_int, [P, I32, C.POINTER(C.c_int64)]),
        "cvf_step_histograms": (C.c_int, [P, I32, I64, C.POINTER(C.c_int64)]),
        "cvf_corr_create": (C.c_void_p, []),
        "cvf_corr_add": (None, [C.c_void_p, C.c_double]),
        "cvf_corr_result": (C.c_int, [C.c_void_p, DP, DP, I32]),
        "cvf_corr_destroy": (None, [C.c_void_p]),
        "cvf_thermo_averages": (C.c_int, [DP, C.c_int64, I32, DP]),
        "cvf_thermo_derivative": (C.c_int, [DP, DP, DP, C.POINTER(C.c_int32), I32, I32, DP, DP, DP]),
        "cvf_thermo_kt": (C.c_int, [DP, DP, DP, DP, I32, C.c_double, DP, DP]),
        "cvf_set_tuning": (C.c_int, [P, C.POINTER(_Tuning)]),
        "cvf_get_plan": (C.c_int, [P, I32, I32P, I32P, I32P, I32P]),
        "cvf_vdw_energy_noalloc": (C.c_int, [I32, I32, F32P, I32P, I32, C.c_float, DP, I32, I32, DP]),
        "cvf_vdw_pair_energy": (C.c_int, [I32, I32, C.c_float, DP, I32, I32, DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _check(lib, rc: int, what: str) -> int:
    if rc < 0:
        raise CVFError(f"{what}: {lib.cvf_last_error().decode(errors='replace')} (status {rc})")
    return rc


def _u8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class Simulation:
    """One handle: a single state point or a batch of independent points."""

    def __init__(self, params: Params | Sequence[Params]):
        self._lib = load_library()
        pts = [params] if isinstance(params, Params) else list(params)
        if not pts:
            raise CVFError("no state points")
        arr = (_Params * len(pts))(*[p._c() for p in pts])
        h = C.c_void_p()
        if len(pts) == 1:
            _check(self._lib, self._lib.cvf_create(arr, C.byref(h)), "cvf_create")
        else:
            _check(self._lib, self._lib.cvf_create_batch(arr, len(pts), C.byref(h)), "cvf_create_batch")
        self._h = h
        self.params = pts

    # -- lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cvf_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- MC
    def step(self, n: int = 1) -> None:
        _check(self._lib, self._lib.cvf_step(self._h, n), "cvf_step")

    def sync(self) -> None:
        _check(self._lib, self._lib.cvf_sync(self._h), "cvf_sync")

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        _check(self._lib, self._lib.cvf_get_stream(self._h, C.byref(s)), "cvf_get_stream")
        return s.value or 0

    @property
    def steps(self) -> int:
        n, s = C.c_int32(), C.c_int64()
        _check(self._lib, self._lib.cvf_info(self._h, 0, C.byref(n), None, C.byref(s)), "cvf_info")
        return s.value

    # -- state
    def N(self, pt: int = 0) -> int:
        return self.params[pt].L ** 3

    def set_config(self, sigma: np.ndarray, eta: np.ndarray, pt: int = 0) -> int:
        n = 6 * self.N(pt)
        s = np.ascontiguousarray(sigma, dtype=np.uint8).reshape(-1)
        e = np.ascontiguousarray(eta, dtype=np.uint8).reshape(-1)
        if s.size != n or e.size != n:
            raise CVFError(f"config arrays must hold 6N = {n} values")
        return _check(self._lib, self._lib.cvf_set_config(self._h, pt, _u8(s), _u8(e)), "cvf_set_config")

    def get_config(self, pt: int = 0) -> tuple[np.ndarray, np.ndarray]:
        n = 6 * self.N(pt)
        s = np.empty(n, np.uint8)
        e = np.empty(n, np.uint8)
        _check(self._lib, self._lib.cvf_get_config(self._h, pt, _u8(s), _u8(e)), "cvf_get_config")
        return s, e

    def labels(self, pt: int = 0) -> np.ndarray:
        out = np.empty(6 * self.N(pt), np.int32)
        _check(self._lib, self._lib.cvf_get_labels(self._h, pt, out.ctypes.data_as(C.POINTER(C.c_int32))),
               "cvf_get_labels")
        return out

    def volume(self, pt: int = 0) -> tuple[float, float, float]:
        r, v, e = C.c_double(), C.c_double(), C.c_double()
        _check(self._lib, self._lib.cvf_get_volume(self._h, pt, C.byref(r), C.byref(v), C.byref(e)),
               "cvf_get_volume")
        return r.value, v.value, e.value

    def shells(self, pt: int = 0) -> tuple[np.ndarray, np.ndarray]:
        d = np.empty(512, np.float32)
        f = np.empty(512, np.int32)
        n = C.c_int32()
        _check(self._lib, self._lib.cvf_get_shells(self._h, pt, d.ctypes.data_as(C.POINTER(C.c_float)),
                                                   f.ctypes.data_as(C.POINTER(C.c_int32)), 512,
                                                   C.byref(n)), "cvf_get_shells")
        return d[: n.value].copy(), f[: n.value].copy()

    def set_shells(self, dist: np.ndarray, freq: np.ndarray, pt: int = 0) -> None:
        d = np.ascontiguousarray(dist, np.float32)
        f = np.ascontiguousarray(freq, np.int32)
        _check(self._lib, self._lib.cvf_set_shells(self._h, pt, d.ctypes.data_as(C.POINTER(C.c_float)),
                                                   f.ctypes.data_as(C.POINTER(C.c_int32)), d.size),
               "cvf_set_shells")

    TUNING_FIELDS = ("streams", "block_sigma", "block_eta", "block_cell", "graph_steps", "fused",
                     "slab_planes", "block_chunk", "chunk_sweeps")

    @property
    def tuning(self) -> dict:
        t = _Tuning()
        _check(self._lib, self._lib.cvf_get_tuning(self._h, C.byref(t)), "cvf_get_tuning")
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    @tuning.setter
    def tuning(self, value: dict) -> None:
        cur = self.tuning
        cur.update(value)
        t = _Tuning(*[int(cur[k]) for k in self.TUNING_FIELDS])
        _check(self._lib, self._lib.cvf_set_tuning(self._h, C.byref(t)), "cvf_set_tuning")

    PLAN_KINDS = ("unfused", "fused", "resident")

    def plan(self, pt: int = 0) -> dict:
        """Launch plan of point pt: stream group, planes per z-slab (0:
        unfused kernels), sigma half-sweeps per fused launch and the kind of
        Metropolis launch (unfused, fused, resident)."""
        g, t, k, kind = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(self._lib, self._lib.cvf_get_plan(self._h, pt, C.byref(g), C.byref(t), C.byref(k), C.byref(kind)),
               "cvf_get_plan")
        return {"group": g.value, "slab_planes": t.value, "chunk_sweeps": k.value,
                "kind": self.PLAN_KINDS[kind.value]}

    def autotune(self, steps_per_trial: int = 20) -> dict:
        """Time candidate launch configurations and keep the fastest (state unchanged)."""
        t = _Tuning()
        _check(self._lib, self._lib.cvf_autotune(self._h, steps_per_trial, C.byref(t)), "cvf_autotune")
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    KCLASSES = ("begin", "eta", "sigma", "sw_local", "sw_union", "sw_flip", "chunk", "resident")

    def profile_steps(self, n: int) -> dict:
        """Instrumented steps: mean ms per launch and launch count per kernel class."""
        k = len(self.KCLASSES)
        ms = (C.c_double * k)()
        cnt = (C.c_int64 * k)()
        _check(self._lib, self._lib.cvf_profile_steps(self._h, n, ms, cnt, k), "cvf_profile_steps")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KCLASSES)}

    def sigma_histogram(self, pt: int = 0) -> np.ndarray:
        out = (C.c_int64 * 6)()
        _check(self._lib, self._lib.cvf_sigma_histogram(self._h, pt, out), "cvf_sigma_histogram")
        return np.array(list(out), np.int64)

    def step_histograms(self, n: int, pt: int = 0) -> np.ndarray:
        """Advance n steps and return point pt's sigma histogram after each:
        shape (n, 6)."""
        out = np.zeros((max(n, 0), 6), np.int64)
        _check(self._lib, self._lib.cvf_step_histograms(self._h, pt, n, out.ctypes.data_as(C.POINTER(C.c_int64))),
               "cvf_step_histograms")
        return out

    # -- measurement
    def observe(self, pt: int = 0) -> Obs:
        o = _Obs()
        _check(self._lib, self._lib.cvf_observe(self._h, pt, C.byref(o)), "cvf_observe")
        return Obs(o.step, o.E, o.H, o.V, o.n_hb, o.m, o.r_cell, o.E_vdw, o.nhb2, o.n_sigma)

    def cluster_stats(self, pt: int = 0) -> ClusterStats:
        c = _CStat()
        _check(self._lib, self._lib.cvf_cluster_stats(self._h, pt, C.byref(c)), "cvf_cluster_stats")
        return ClusterStats(c.largest, c.second, c.n_perc, c.n_clusters, c.mean, c.var)


def vdw_energy(L: int, dist: np.ndarray, freq: np.ndarray, r_cutoff: float, r: np.ndarray,
               precision: int = FP64, device_id: int = 0, noalloc: bool = False) -> np.ndarray:
    """E_vdW(r) of a shell table evaluated on the GPU (FP64 or FP32-compensated).
    noalloc: through cvf_vdw_energy_noalloc (fixed device scratch, no allocation)."""
    lib = load_library()
    d = np.ascontiguousarray(dist, np.float32)
    f = np.ascontiguousarray(freq, np.int32)
    rr = np.ascontiguousarray(r, np.float64)
    out = np.empty_like(rr)
    fn, name = (lib.cvf_vdw_energy_noalloc, "cvf_vdw_energy_noalloc") if noalloc else (lib.cvf_vdw_energy,
                                                                                      "cvf_vdw_energy")
    _check(lib, fn(device_id, L, d.ctypes.data_as(C.POINTER(C.c_float)),
                                   f.ctypes.data_as(C.POINTER(C.c_int32)), d.size, r_cutoff,
                                   rr.ctypes.data_as(C.POINTER(C.c_double)), rr.size, precision,
                                   out.ctypes.data_as(C.POINTER(C.c_double))), name)
    return out


def vdw_pair_energy(L: int, r_cutoff: float, r: np.ndarray, precision: int = FP64, device_id: int = 0) -> np.ndarray:
    """E_vdW(r) as the direct minimum-image pair sum over the L^3 lattice on
    the GPU (no shell table): FP64 or FP32-compensated."""
    lib = load_library()
    rr = np.ascontiguousarray(r, np.float64)
    out = np.empty_like(rr)
    _check(lib, lib.cvf_vdw_pair_energy(device_id, L, r_cutoff, rr.ctypes.data_as(C.POINTER(C.c_double)), rr.size,
                                        precision, out.ctypes.data_as(C.POINTER(C.c_double))), "cvf_vdw_pair_energy")
    return out


class Correlator:
    """Multi-tau correlator of the Flag_correlation mode (host code in libcvf.so)."""

    def __init__(self):
        self._lib = load_library()
        self._h = self._lib.cvf_corr_create()

    def add(self, w: float) -> None:
        self._lib.cvf_corr_add(self._h, float(w))

    def result(self) -> tuple[np.ndarray, np.ndarray]:
        t = np.empty(1024)
        f = np.empty(1024)
        n = self._lib.cvf_corr_result(self._h, t.ctypes.data_as(C.POINTER(C.c_double)),
                                      f.ctypes.data_as(C.POINTER(C.c_double)), 1024)
        if n < 0:
            raise CVFError("cvf_corr_result failed")
        return t[:n].copy(), f[:n].copy()

    def __del__(self):
        try:
            self._lib.cvf_corr_destroy(self._h)
        except Exception:
            pass


# ----------------------------------------------------- thermodynamic analysis
KT_NOT_VALUE = -10.0


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def thermo_averages(rows) -> np.ndarray:
    """Means and population variances of e, h, v, N_HB, m, r_cell over data.out
    rows (columns 1..6 of each row after the step), as script_analisi.sh:
    returns [mean_e, var_e, mean_h, var_h, ..., mean_r, var_r]."""
    a = np.ascontiguousarray(np.asarray(rows, np.float64)[:, :6])
    out = np.empty(12)
    lib = load_library()
    rc = lib.cvf_thermo_averages(_dp(a), a.shape[0], 6, _dp(out))
    if rc:
        raise CVFError(f"cvf_thermo_averages: {rc}")
    return out


def thermo_derivative(T, X, var, n, mode: int):
    """d X/dT (mode 0, CPderivate) or (1/X) dX/dT (mode 1, AlphaPderivate)
    with their error estimates, rows 2 .. len-3 of the T-sorted input."""
    T, X, var = (np.ascontiguousarray(np.asarray(v, np.float64)) for v in (T, X, var))
    n = np.ascontiguousarray(np.asarray(n, np.int32))
    m = len(T)
    to, d, e = np.empty(max(m, 1)), np.empty(max(m, 1)), np.empty(max(m, 1))
    k = load_library().cvf_thermo_derivative(_dp(T), _dp(X), _dp(var), n.ctypes.data_as(C.POINTER(C.c_int32)),
                                             m, mode, _dp(to), _dp(d), _dp(e))
    if k < 0:
        raise CVFError(f"cvf_thermo_derivative: {k}")
    return to[:k].copy(), d[:k].copy(), e[:k].copy()


def thermo_kt(T, P, rho, rho_err, gas_rho: float = 500.0):
    """K_T = d ln(rho)/dP along isotherms with its error (KTderivate.c);
    KT_NOT_VALUE where undefined."""
    T, P, rho, rho_err = (np.ascontiguousarray(np.asarray(v, np.float64)) for v in (T, P, rho, rho_err))
    kt, ekt = np.empty(max(len(T), 1)), np.empty(max(len(T), 1))
    rc = load_library().cvf_thermo_kt(_dp(T), _dp(P), _dp(rho), _dp(rho_err), len(T), gas_rho, _dp(kt), _dp(ekt))
    if rc:
        raise CVFError(f"cvf_thermo_kt: {rc}")
    return kt[:len(T)].copy(), ekt[:len(T)].copy()
