"""ctypes binding of the CPU oracle (oracle/cvf_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libcvf_oracle.so"

ST_SCHED, ST_SIGMA, ST_ETA, ST_VOLUME, ST_SWBOND, ST_SWFLIP, ST_INIT_SIGMA, ST_INIT_CUBE, ST_SIGMA_F = range(1, 10)


class _Params(C.Structure):
    _fields_ = [
        ("L", C.c_int32), ("P", C.c_float), ("T", C.c_float), ("seed", C.c_uint64),
        ("n_metropolis", C.c_int32), ("n_cluster", C.c_int32), ("r_cutoff", C.c_float),
        ("J_hb", C.c_float), ("J_sig", C.c_float), ("v_hb", C.c_float), ("flag_config", C.c_int32),
        ("flag_chessboard", C.c_int32), ("schedule", C.c_int32),
    ]


class _Obs(C.Structure):
    _fields_ = [
        ("E", C.c_double), ("H", C.c_double), ("V", C.c_double), ("n_hb", C.c_double),
        ("m", C.c_double), ("r_cell", C.c_double), ("nhb2", C.c_int64), ("n_sigma", C.c_int64),
    ]


class _CStat(C.Structure):
    _fields_ = [
        ("largest", C.c_int32), ("second", C.c_int32), ("n_perc", C.c_int32),
        ("n_clusters", C.c_int32), ("mean", C.c_float), ("var", C.c_float),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, U8P, I32P, U32P, U64P, DP, FP = (C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_double), C.POINTER(C.c_float))
        sig = {
            "ora_philox": (None, [U32P, U32P, U32P]),
            "ora_shell_table": (C.c_int, [C.c_int, C.c_float, FP, I32P, C.c_int]),
            "ora_vdw_energy": (C.c_double, [C.c_double, C.c_float, FP, I32P, C.c_int]),
            "ora_create": (P, [C.POINTER(_Params)]),
            "ora_destroy": (None, [P]),
            "ora_set_config": (C.c_int, [P, U8P, U8P]),
            "ora_get_config": (None, [P, U8P, U8P]),
            "ora_get_cube_states": (None, [P, I32P]),
            "ora_get_nhb": (None, [P, U8P]),
            "ora_step": (None, [P]),
            "ora_steps": (None, [P, C.c_int64]),
            "ora_eta_sweep": (None, [P, C.c_uint64]),
            "ora_sigma_halfsweep": (None, [P, C.c_uint64, C.c_uint32, C.c_int, C.c_int]),
            "ora_sigma_colour": (None, [P, C.c_uint64, C.c_uint32]),
            "ora_schedule": (None, [C.c_uint64, C.c_uint64, C.c_uint32, I32P, I32P]),
            "ora_volume_move": (None, [P, C.c_uint64]),
            "ora_sw_step": (None, [P, C.c_uint64, C.c_uint32]),
            "ora_get_step": (C.c_uint64, [P]),
            "ora_set_step": (None, [P, C.c_uint64]),
            "ora_measure": (None, [P, C.POINTER(_Obs)]),
            "ora_get_labels": (None, [P, I32P]),
            "ora_cluster_stats": (C.c_int, [P, C.POINTER(_CStat)]),
            "ora_cluster_stats_labels": (C.c_int, [C.c_int, I32P, C.POINTER(_CStat)]),
            "ora_get_volume_state": (None, [P, DP, DP, DP]),
            "ora_get_thresholds": (None, [P, U64P, U64P, U64P]),
            "ora_format_data_line": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_Obs)]),
            "ora_format_cluster_line": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_CStat)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().ora_philox(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def draw(seed: int, stream: int, site: int, step: int, sub: int) -> np.ndarray:
    return philox([site, step & 0xFFFFFFFF, (stream << 24) | sub, step >> 32],
                  [seed & 0xFFFFFFFF, seed >> 32])


def shell_table(L: int, r_cutoff: float) -> tuple[np.ndarray, np.ndarray]:
    d = np.zeros(256, np.float32)
    f = np.zeros(256, np.int32)
    n = lib().ora_shell_table(L, r_cutoff, d.ctypes.data_as(C.POINTER(C.c_float)),
                              f.ctypes.data_as(C.POINTER(C.c_int32)), 256)
    return d[:n].copy(), f[:n].copy()


def vdw_energy(r: float, r_cutoff: float, dist: np.ndarray, freq: np.ndarray) -> float:
    d = np.ascontiguousarray(dist, np.float32)
    f = np.ascontiguousarray(freq, np.int32)
    return lib().ora_vdw_energy(r, r_cutoff, d.ctypes.data_as(C.POINTER(C.c_float)),
                                f.ctypes.data_as(C.POINTER(C.c_int32)), d.size)


@dataclass
class OraObs:
    E: float
    H: float
    V: float
    n_hb: float
    m: float
    r_cell: float
    nhb2: int
    n_sigma: int


def cluster_stats_labels(L: int, labels: np.ndarray):
    lab = np.ascontiguousarray(labels, np.int32)
    c = _CStat()
    rc = lib().ora_cluster_stats_labels(L, lab.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(c))
    if rc < 0:
        raise ValueError("label out of range")
    return (c.largest, c.second, c.n_perc, c.n_clusters, c.mean, c.var)


class OracleSim:
    def __init__(self, L=32, P=0.45, T=0.815, seed=44648604, n_metropolis=1, n_cluster=0,
                 r_cutoff=6.0, J_hb=0.5, J_sig=0.08, v_hb=0.6, flag_config=0, flag_chessboard=1,
                 schedule=0):
        self.L = L
        self.N = L ** 3
        p = _Params(L, P, T, seed, n_metropolis, n_cluster, r_cutoff, J_hb, J_sig, v_hb, flag_config,
                    flag_chessboard, schedule)
        self._h = lib().ora_create(C.byref(p))
        if not self._h:
            raise ValueError("ora_create failed")

    def close(self):
        if self._h:
            lib().ora_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, sigma, eta) -> int:
        s = np.ascontiguousarray(sigma, np.uint8).reshape(-1)
        e = np.ascontiguousarray(eta, np.uint8).reshape(-1)
        return lib().ora_set_config(self._h, s.ctypes.data_as(C.POINTER(C.c_uint8)),
                                    e.ctypes.data_as(C.POINTER(C.c_uint8)))

    def get_config(self):
        s = np.empty(6 * self.N, np.uint8)
        e = np.empty(6 * self.N, np.uint8)
        lib().ora_get_config(self._h, s.ctypes.data_as(C.POINTER(C.c_uint8)),
                             e.ctypes.data_as(C.POINTER(C.c_uint8)))
        return s, e

    def cube_states(self):
        out = np.empty(self.N // 4, np.int32)
        lib().ora_get_cube_states(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def nhb(self):
        out = np.empty(self.N, np.uint8)
        lib().ora_get_nhb(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def step(self, n: int = 1):
        lib().ora_steps(self._h, n)

    @property
    def steps(self) -> int:
        return lib().ora_get_step(self._h)

    @steps.setter
    def steps(self, v: int):
        lib().ora_set_step(self._h, v)

    def eta_sweep(self, step: int):
        lib().ora_eta_sweep(self._h, step)

    def sigma_halfsweep(self, step: int, sub: int, arm: int, indices: int):
        lib().ora_sigma_halfsweep(self._h, step, sub, arm, indices)

    def sigma_colour(self, step: int, pass_: int):
        lib().ora_sigma_colour(self._h, step, pass_)

    def volume_move(self, step: int):
        lib().ora_volume_move(self._h, step)

    def sw_step(self, step: int, jc: int = 0):
        lib().ora_sw_step(self._h, step, jc)

    def measure(self) -> OraObs:
        o = _Obs()
        lib().ora_measure(self._h, C.byref(o))
        return OraObs(o.E, o.H, o.V, o.n_hb, o.m, o.r_cell, o.nhb2, o.n_sigma)

    def labels(self) -> np.ndarray:
        out = np.empty(6 * self.N, np.int32)
        lib().ora_get_labels(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def cluster_stats(self):
        c = _CStat()
        lib().ora_cluster_stats(self._h, C.byref(c))
        return (c.largest, c.second, c.n_perc, c.n_clusters, c.mean, c.var)

    def volume(self):
        r, v, e = C.c_double(), C.c_double(), C.c_double()
        lib().ora_get_volume_state(self._h, C.byref(r), C.byref(v), C.byref(e))
        return r.value, v.value, e.value

    def thresholds(self):
        s = (C.c_uint64 * 33)()
        e = (C.c_uint64 * 9)()
        w = (C.c_uint64 * 2)()
        lib().ora_get_thresholds(self._h, s, e, w)
        return np.array(list(s), np.uint64), np.array(list(e), np.uint64), np.array(list(w), np.uint64)

    def data_line(self, t: int) -> str:
        o = _Obs()
        lib().ora_measure(self._h, C.byref(o))
        buf = C.create_string_buffer(256)
        lib().ora_format_data_line(buf, 256, t, C.byref(o))
        return buf.value.decode()


def schedule(seed: int, step: int, sub: int) -> tuple[int, int]:
    a, i = C.c_int32(), C.c_int32()
    lib().ora_schedule(seed, step, sub, C.byref(a), C.byref(i))
    return a.value, i.value
