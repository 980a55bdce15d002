"""Histogram reweighting of (energy, density) samples (include/cvf_reweight.h).

The engine's form of the reference's HistogramReweighting program
(histogram_reweighting_mixed_paramSI_norm_x.f90): the simulations'
(e, rho) histograms, the reweighting constants, the reweighted density at
any (T, P) batch, its Binder cumulants and its mixed order-parameter
distribution compared with the 3D Ising one.  The per-cell sums over
simulations run on the GPU (libcvf.so); there is no CPU fallback.

:func:`from_series` feeds it straight from a batched scan
(:func:`cvfbulkwater_amd.scan.scan_series`), with h = H/N and rho = N/V as
the driver script's awk takes them from data.out.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import CVFError, load_library

__all__ = ["Reweighting", "NOT_CONVERGED", "ising_scores", "ising2d", "ising3d", "from_series"]

NOT_CONVERGED = 1
DP = C.POINTER(C.c_double)


class _Info(C.Structure):
    _fields_ = [
        ("n_sim", C.c_int32), ("N", C.c_int32), ("n_bin", C.c_int32), ("n_bin2d", C.c_int32),
        ("n_bin_mixed", C.c_int32), ("pad", C.c_int32), ("n_occupied", C.c_int64),
        ("e_min", C.c_double), ("rho_min", C.c_double), ("de", C.c_double), ("drho", C.c_double),
    ]


_SIGS = {
    "cvf_rw_create": (C.c_int, [C.c_int32, C.c_int32, DP, DP, C.POINTER(C.c_int64), DP, C.c_double, C.c_int32,
                                C.POINTER(C.c_void_p)]),
    "cvf_rw_destroy": (None, [C.c_void_p]),
    "cvf_rw_get_info": (C.c_int, [C.c_void_p, C.POINTER(_Info)]),
    "cvf_rw_sim_points": (C.c_int, [C.c_void_p, DP, DP]),
    "cvf_rw_solve_constants": (C.c_int, [C.c_void_p, C.c_int32, DP, DP]),
    "cvf_rw_set_constants": (C.c_int, [C.c_void_p, DP]),
    "cvf_rw_reweight": (C.c_int, [C.c_void_p, C.c_int32, DP, DP, DP]),
    "cvf_rw_density": (C.c_int, [C.c_void_p, C.c_int32, DP]),
    "cvf_rw_binder_ed": (C.c_int, [C.c_void_p, C.c_int32, DP, DP]),
    "cvf_rw_order_parameter": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, DP, DP, DP]),
    "cvf_rw_ising_scores": (C.c_int, [DP, DP, C.c_int32, DP]),
    "cvf_rw_ising2d": (C.c_double, [C.c_double]),
    "cvf_rw_ising3d": (C.c_double, [C.c_double]),
    "cvf_rw_errors": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, DP, DP]),
    "cvf_rw_scan": (C.c_int, [C.c_void_p, C.c_int32, DP, DP, C.c_int32, DP, DP, DP, DP, DP, DP]),
}


_BOUND = None


def _lib():
    global _BOUND
    lib = load_library()
    if _BOUND is not lib:
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _BOUND = lib
    return lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(DP)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise CVFError(f"{what}: {_lib().cvf_last_error().decode(errors='replace')} (status {rc})")
    return rc


class Reweighting:
    """Histograms of a set of simulations and their reweighting.

    samples[j] is an (n_j, 2) array of (h, rho) in internal units taken at
    internal temperature T_hat[j] and pressure P_hat[j]."""

    def __init__(self, L: int, T_hat: Sequence[float], P_hat: Sequence[float], samples: Sequence[np.ndarray],
                 bin_factor: float = 1.0, device: int = 0):
        lib = _lib()
        T_hat, P_hat = _f64(T_hat), _f64(P_hat)
        if len(samples) != len(T_hat) or len(P_hat) != len(T_hat):
            raise CVFError("one sample array per (T, P) is needed")
        counts = np.array([len(s) for s in samples], dtype=np.int64)
        h_rho = _f64(np.concatenate([np.asarray(s, dtype=np.float64).reshape(-1, 2) for s in samples]))
        h = C.c_void_p()
        _check(lib.cvf_rw_create(L, len(T_hat), _d(T_hat), _d(P_hat), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                 _d(h_rho), bin_factor, device, C.byref(h)), "cvf_rw_create")
        self._h = h
        self.n_sim = len(T_hat)
        self.n_targets = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().cvf_rw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        i = _Info()
        _check(_lib().cvf_rw_get_info(self._h, C.byref(i)), "cvf_rw_get_info")
        return {k: getattr(i, k) for k, _ in _Info._fields_ if k != "pad"}

    def sim_points(self) -> tuple[np.ndarray, np.ndarray]:
        T, P = np.zeros(self.n_sim), np.zeros(self.n_sim)
        _check(_lib().cvf_rw_sim_points(self._h, _d(T), _d(P)), "cvf_rw_sim_points")
        return T, P

    def solve_constants(self, steps: int) -> tuple[np.ndarray, np.ndarray, bool]:
        """(constants, constants of the sweep before, converged)."""
        c, prev = np.zeros(self.n_sim), np.zeros(self.n_sim)
        rc = _check(_lib().cvf_rw_solve_constants(self._h, steps, _d(c), _d(prev)), "cvf_rw_solve_constants")
        return c, prev, rc == 0

    def set_constants(self, ctes) -> None:
        ctes = _f64(ctes)
        _check(_lib().cvf_rw_set_constants(self._h, _d(ctes)), "cvf_rw_set_constants")

    def reweight(self, T_K, P_bar) -> np.ndarray:
        """Reweight to the (T K, P bar) targets; returns log_Z per target."""
        T_K, P_bar = _f64(np.atleast_1d(T_K)), _f64(np.atleast_1d(P_bar))
        lz = np.zeros(len(T_K))
        _check(_lib().cvf_rw_reweight(self._h, len(T_K), _d(T_K), _d(P_bar), _d(lz)), "cvf_rw_reweight")
        self.n_targets = len(T_K)
        return lz

    def density(self, t: int = 0) -> np.ndarray:
        nb = self.info()["n_bin2d"]
        g = np.zeros(nb * nb)
        _check(_lib().cvf_rw_density(self._h, t, _d(g)), "cvf_rw_density")
        return g.reshape(nb, nb)

    def binder_ed(self, t: int = 0) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        _check(_lib().cvf_rw_binder_ed(self._h, t, C.byref(a), C.byref(b)), "cvf_rw_binder_ed")
        return a.value, b.value

    def order_parameter(self, t: int, alpha: float):
        """(x, p, (media, a1, binder)) of target t at angle alpha."""
        nm = self.info()["n_bin_mixed"]
        x, p, st = np.zeros(nm), np.zeros(nm), np.zeros(3)
        _check(_lib().cvf_rw_order_parameter(self._h, t, alpha, _d(x), _d(p), _d(st)), "cvf_rw_order_parameter")
        return x, p, tuple(st)

    def scan(self, T_K, P_bar, alpha) -> dict:
        """The scan over (T K, P bar) targets and angles with the projections
        reduced on the device: binder_ed (n_t, 2); x, p (n_t, n_a, nm);
        stats (n_t, n_a, 3) = {media, a1, binder}; scores (n_t, n_a, 6)."""
        T_K, P_bar = _f64(np.atleast_1d(T_K)), _f64(np.atleast_1d(P_bar))
        alpha = _f64(np.atleast_1d(alpha))
        nt, na, nm = len(T_K), len(alpha), self.info()["n_bin_mixed"]
        out = {"binder_ed": np.zeros((nt, 2)), "x": np.zeros((nt, na, nm)), "p": np.zeros((nt, na, nm)),
               "stats": np.zeros((nt, na, 3)), "scores": np.zeros((nt, na, 6))}
        _check(_lib().cvf_rw_scan(self._h, nt, _d(T_K), _d(P_bar), na, _d(alpha), _d(out["binder_ed"]),
                                  _d(out["x"]), _d(out["p"]), _d(out["stats"]), _d(out["scores"])), "cvf_rw_scan")
        self.n_targets = 0
        return out

    def errors(self, t: int, sim: int, alpha: float) -> tuple[np.ndarray, np.ndarray]:
        nm = self.info()["n_bin_mixed"]
        eh, ex = np.zeros(nm), np.zeros(nm)
        _check(_lib().cvf_rw_errors(self._h, t, sim, alpha, _d(eh), _d(ex)), "cvf_rw_errors")
        return eh, ex


def ising_scores(x, p) -> np.ndarray:
    """{kl_2D, kl_3D, liu_2D, liu_3D, zero, height} of a distribution."""
    x, p = _f64(x), _f64(p)
    out = np.zeros(6)
    _check(_lib().cvf_rw_ising_scores(_d(x), _d(p), len(x), _d(out)), "cvf_rw_ising_scores")
    return out


def ising2d(z: float) -> float:
    return _lib().cvf_rw_ising2d(z)


def ising3d(z: float) -> float:
    return _lib().cvf_rw_ising3d(z)


def from_series(L: int, T_hat: Sequence[float], P_hat: Sequence[float], series: np.ndarray,
                bin_factor: float = 1.0, device: int = 0) -> Reweighting:
    """Reweighting of a batched scan's samples: series is (points, samples, 6)
    with columns (e, h, v, N_HB, m, r_cell) as scan_series returns them; the
    samples are h / N and N / v (the driver's awk over data.out)."""
    N = float(L) ** 3
    s = np.asarray(series, dtype=np.float64)
    samples = [np.stack([p[:, 1] / N, N / p[:, 2]], axis=1) for p in s]
    return Reweighting(L, T_hat, P_hat, samples, bin_factor, device)
