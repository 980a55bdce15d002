"""ctypes binding of the histogram reweighting oracle (oracle/cvf_rw_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the engine's reweighting library
and executable, never used by the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import py as _py

DP = C.POINTER(C.c_double)


class Grid(C.Structure):
    _fields_ = [
        ("L", C.c_int32), ("N", C.c_int32), ("n_sim", C.c_int32), ("n_bin", C.c_int32),
        ("n_bin2d", C.c_int32), ("n_bin_mixed", C.c_int32),
        ("Emin", C.c_double), ("Dmin", C.c_double), ("dE", C.c_double), ("dD", C.c_double),
        ("T", DP), ("P", DP), ("log_n", DP), ("datos", DP),
    ]


_SIGS = {
    "ora_rw_grid_build": (C.c_int, [C.c_int32, C.c_int32, DP, DP, C.POINTER(C.c_int32), DP, C.c_double,
                                    C.POINTER(Grid)]),
    "ora_rw_grid_from_datos": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                         C.c_double, DP, DP, DP, DP, C.POINTER(Grid)]),
    "ora_rw_grid_free": (None, [C.POINTER(Grid)]),
    "ora_rw_calcula": (C.c_int, [C.POINTER(Grid), DP, C.c_double, C.c_double, DP, C.POINTER(C.c_uint8), DP]),
    "ora_rw_constants": (C.c_int, [C.POINTER(Grid), C.c_int32, DP, DP]),
    "ora_rw_distribution": (C.c_int, [C.POINTER(Grid), DP, C.c_double, C.c_double, DP]),
    "ora_rw_binder_ed": (None, [C.POINTER(Grid), DP, DP, DP]),
    "ora_rw_normx": (C.c_int, [C.POINTER(Grid), DP, C.c_double, DP, DP, DP, DP]),
    "ora_rw_binder": (C.c_double, [DP, DP, C.c_int32, C.c_double, C.c_double]),
    "ora_rw_ising2d": (C.c_double, [C.c_double]),
    "ora_rw_ising3d": (C.c_double, [C.c_double]),
    "ora_rw_peaks": (C.c_int, [DP, DP]),
    "ora_rw_scores": (None, [DP, DP, C.c_int32, C.c_double, C.c_double, DP]),
    "ora_rw_errors": (C.c_int, [C.POINTER(Grid), DP, C.c_double, DP, C.c_int32, DP, DP]),
    "ora_rw_main": (C.c_int, [C.c_char_p, C.c_char_p]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _py.lib()
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(DP)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class OraGrid:
    """The program's histogram state (datos_originales and bin geometry)."""

    def __init__(self, g: Grid):
        self.g = g
        self._keep = []

    @classmethod
    def from_samples(cls, L: int, T_hat, P_hat, counts, h_rho, bin_factor: float) -> "OraGrid":
        T_hat, P_hat, h_rho = _f64(T_hat), _f64(P_hat), _f64(h_rho)
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        g = Grid()
        rc = lib().ora_rw_grid_build(L, len(T_hat), _d(T_hat), _d(P_hat),
                                     counts.ctypes.data_as(C.POINTER(C.c_int32)), _d(h_rho), bin_factor,
                                     C.byref(g))
        if rc != 0:
            raise ValueError("ora_rw_grid_build failed")
        return cls(g)

    @classmethod
    def from_datos(cls, N: int, datos: np.ndarray, Emin: float, Dmin: float, dE: float, dD: float, T_K, P_bar,
                   log_n) -> "OraGrid":
        datos = _f64(datos)
        n_sim, nb, nb2 = datos.shape
        assert nb == nb2
        T_K, P_bar, log_n = _f64(T_K), _f64(P_bar), _f64(log_n)
        g = Grid()
        rc = lib().ora_rw_grid_from_datos(N, n_sim, nb, Emin, Dmin, dE, dD, _d(T_K), _d(P_bar), _d(log_n),
                                          _d(datos), C.byref(g))
        if rc != 0:
            raise ValueError("ora_rw_grid_from_datos failed")
        return cls(g)

    def close(self):
        if self.g is not None:
            lib().ora_rw_grid_free(C.byref(self.g))
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nb(self) -> int:
        return self.g.n_bin2d

    @property
    def n_sim(self) -> int:
        return self.g.n_sim

    def datos(self) -> np.ndarray:
        n = self.g.n_sim * self.nb * self.nb
        return np.ctypeslib.as_array(self.g.datos, shape=(n,)).reshape(self.g.n_sim, self.nb, self.nb).copy()

    def T(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.T, shape=(self.g.n_sim,)).copy()

    def P(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.P, shape=(self.g.n_sim,)).copy()

    def calcula(self, ctes, temp: float, pres: float):
        ctes = _f64(ctes)
        lp = np.zeros(self.nb * self.nb)
        memo = np.zeros(self.nb * self.nb, dtype=np.uint8)
        lz = C.c_double()
        rc = lib().ora_rw_calcula(C.byref(self.g), _d(ctes), temp, pres, _d(lp),
                                  memo.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(lz))
        if rc != 0:
            raise ValueError("empty histogram")
        return lp.reshape(self.nb, self.nb), memo.reshape(self.nb, self.nb), lz.value

    def constants(self, steps: int):
        c = np.zeros(self.n_sim)
        prev = np.zeros(self.n_sim)
        rc = lib().ora_rw_constants(C.byref(self.g), steps, _d(c), _d(prev))
        return c, prev, rc

    def distribution(self, ctes, temp: float, pres: float) -> np.ndarray:
        ctes = _f64(ctes)
        grid = np.zeros(self.nb * self.nb)
        if lib().ora_rw_distribution(C.byref(self.g), _d(ctes), temp, pres, _d(grid)) != 0:
            raise ValueError("empty histogram")
        return grid.reshape(self.nb, self.nb)

    def binder_ed(self, grid: np.ndarray):
        grid = _f64(grid)
        ue, ur = C.c_double(), C.c_double()
        lib().ora_rw_binder_ed(C.byref(self.g), _d(grid), C.byref(ue), C.byref(ur))
        return ue.value, ur.value

    def normx(self, grid: np.ndarray, alpha: float):
        grid = _f64(grid)
        nm = self.g.n_bin_mixed
        x, p = np.zeros(nm), np.zeros(nm)
        media, a1 = C.c_double(), C.c_double()
        rc = lib().ora_rw_normx(C.byref(self.g), _d(grid), alpha, _d(x), _d(p), C.byref(media), C.byref(a1))
        return x, p, media.value, a1.value, rc

    def errors(self, grid: np.ndarray, alpha: float, samples: np.ndarray):
        grid, samples = _f64(grid), _f64(samples)
        nm = self.g.n_bin_mixed
        eh, ex = np.zeros(nm), np.zeros(nm)
        rc = lib().ora_rw_errors(C.byref(self.g), _d(grid), alpha, _d(samples), len(samples) // 2
                                 if samples.ndim == 1 else samples.shape[0], _d(eh), _d(ex))
        return eh, ex, rc


def binder(x, p, media: float, a1: float) -> float:
    x, p = _f64(x), _f64(p)
    return lib().ora_rw_binder(_d(x), _d(p), len(x), media, a1)


def ising2d(z: float) -> float:
    return lib().ora_rw_ising2d(z)


def ising3d(z: float) -> float:
    return lib().ora_rw_ising3d(z)


def peaks() -> tuple[float, float]:
    a, b = C.c_double(), C.c_double()
    if lib().ora_rw_peaks(C.byref(a), C.byref(b)) != 0:
        raise ValueError("Ising_Peak stopped")
    return a.value, b.value


def scores(x, p, peak2d: float, peak3d: float) -> np.ndarray:
    x, p = _f64(x), _f64(p)
    out = np.zeros(6)
    lib().ora_rw_scores(_d(x), _d(p), len(x), peak2d, peak3d, _d(out))
    return out


def run_program(directory: str | Path, summary: str | Path) -> int:
    return lib().ora_rw_main(str(directory).encode(), str(summary).encode())
