"""Batched (T, P) scans feeding the thermodynamic analysis directly.

The reference runs one gpu_water3D per L<L>_T<T>_P<P>/ folder and then
averages the data.out files with CVF_ThermodynamicAnalysis/script_analisi.sh.
Here every state point of the grid advances in one batched handle, the
samples stay in memory, and the averages table and isobar tables are written
from them (cvf_thermo_averages, in the formats script_analisi.sh and
filter_isobars.sh produce); the isobars and convert stages of the cvf_thermo
executable then write the per-observable, derivative and SI files.
"""
from __future__ import annotations

import subprocess
from pathlib import Path
from typing import Sequence

import numpy as np

from . import (FP32_COMPENSATED, SCHED_REFERENCE, THERMO_EXE_PATH, VOLUME_DEVICE, Params, Simulation,
               thermo_averages)


def awk_num(v: float) -> str:
    """A number as awk prints it: integers as integers, otherwise "%.6g"."""
    if np.isfinite(v) and v == np.floor(v) and abs(v) < 1e15:
        return "%d" % int(v)
    return "%.6g" % v


def label(x: float) -> str:
    """Folder-name form of T or P (0.815, 0.45, -0.1)."""
    return "%.6g" % x


def scan_series(L: int, points: Sequence[tuple[float, float]], equil_steps: int, samples: int, interval: int,
                seed: int = 44648604, precision: int = FP32_COMPENSATED, n_cluster: int = 0,
                tuning: dict | None = None, schedule: int = SCHED_REFERENCE) -> np.ndarray:
    """Observables (e, h, v, N_HB, m, r_cell) of every point after each
    sampling step: shape (points, samples, 6).  Sampling follows the
    reference's production loop (gpu_water3D.cu:1204-1243): a sample after
    production step i whenever i % interval == 0, i.e. after steps 1,
    1 + interval, 1 + 2 interval, ...  schedule=SCHED_FAST runs the
    colour-pass schedule (same model, another trajectory)."""
    pts = [Params(L=L, T=T, P=P, seed=seed + 1000003 * i, precision=precision, volume_mode=VOLUME_DEVICE,
                  n_cluster=n_cluster, schedule=schedule) for i, (T, P) in enumerate(points)]
    sim = Simulation(pts)
    if tuning:
        sim.tuning = tuning
    sim.step(equil_steps)
    out = np.empty((len(pts), samples, 6))
    for s in range(samples):
        sim.step(1 if s == 0 else interval)
        for i in range(len(pts)):
            o = sim.observe(i)
            out[i, s] = (o.E, o.H, o.V, o.n_hb, o.m, o.r_cell)
    sim.close()
    return out


def write_data_out(path: Path, series: np.ndarray, interval: int) -> None:
    """data.out rows in the reference format (gpu_water3D.cu:2159)."""
    with open(path, "w") as f:
        for s, r in enumerate(series):
            f.write("%d %s\n" % (s * interval, " ".join("%1.7e" % v for v in r)))


def isobar_scan(out_dir: str | Path, L: int, temperatures: Sequence[float], pressures: Sequence[float],
                equil_steps: int, samples: int, interval: int = 10, seed: int = 44648604,
                precision: int = FP32_COMPENSATED, n_cluster: int = 0, data_out: bool = False,
                tuning: dict | None = None, schedule: int = SCHED_REFERENCE) -> Path:
    """Scan the (T, P) grid in one batch and write the analysis files of
    CVF_ThermodynamicAnalysis into out_dir.  With data_out=True the
    L<L>_T<T>_P<P>/data.out files are written as well."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    points = [(T, P) for P in pressures for T in temperatures]
    series = scan_series(L, points, equil_steps, samples, interval, seed, precision, n_cluster, tuning, schedule)
    rows = []
    for (T, P), s in zip(points, series):
        a = thermo_averages(s)
        rows.append([str(L), label(P), label(T)] + [awk_num(v) for v in a] + [awk_num(len(s))])
        if data_out:
            d = out / f"L{L}_T{label(T)}_P{label(P)}"
            d.mkdir(exist_ok=True)
            write_data_out(d / "data.out", s, interval)
    (out / f"L{L}_average_data.dat").write_text("".join("  ".join(r) + "\n" for r in rows))
    for P in pressures:
        iso = sorted((r for r in rows if r[1] == label(P)), key=lambda r: float(r[2]))
        (out / f"L{L}_P{label(P)}.dat").write_text("".join("  ".join(r) + "\n" for r in iso))
    if not THERMO_EXE_PATH.exists():
        raise FileNotFoundError(f"{THERMO_EXE_PATH} is not built")
    r = subprocess.run([str(THERMO_EXE_PATH), "isobars", str(out)], capture_output=True, text=True)
    if r.returncode == 0:
        r = subprocess.run([str(THERMO_EXE_PATH), "convert", str(out)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cvf_thermo failed: {r.stderr}")
    return out
