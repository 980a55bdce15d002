"""Synthetic inputs of the histogram reweighting program.

(h, rho) samples in internal units shaped like the reference's runs near
the liquid-liquid critical point: two overlapping (energy, density) blobs
whose weights shift with T, one simulation per temperature of an isobar.
The files follow HistogramReweighting/histogram_reweighting_mixed_paramSI_
norm_x.sh:167-206: `input` with thirteen "KEY value" lines and one
"T P count" line per simulation (T padded to 9 characters, P to 6), and
data_L<L>_T<T>_P<P> files holding "h rho" as awk prints them ("%.6g").
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def samples(T_hat, n: int, seed: int = 5, Tc: float = 0.1063, width: float = 0.0004) -> list[np.ndarray]:
    """One (n, 2) array of (h, rho) per temperature, rounded to 6
    significant digits like the driver's awk."""
    rng = np.random.default_rng(seed)
    out = []
    for T in T_hat:
        w_hi = 1.0 / (1.0 + np.exp((T - Tc) / width))  # high-density weight
        hi = rng.random(n) < w_hi
        rho = np.where(hi, 0.690, 0.640) + 0.006 * rng.standard_normal(n)
        h = np.where(hi, -2.80, -2.72) + 0.012 * rng.standard_normal(n) + 0.5 * (rho - 0.665) \
            + 0.8 * (T - Tc)
        out.append(np.array([[float("%.6g" % a), float("%.6g" % b)] for a, b in zip(h, rho)]))
    return out


def pad(s: str, width: int) -> str:
    """The driver's zero padding of T and P labels (…_x.sh:184-198)."""
    while len(s) < width:
        s += "0"
    return s


def write_case(d: Path, L: int, T_labels, P_label: str, n: int, steps: int = 300,
               T_range=("0.106000", "0.106700", "0.0002"), P_range=("0.500", "0.501", "0.1"),
               S_range=("1.5500", "1.6000", "0.02"), bin_fct: str = "0.5", seed: int = 5,
               counts=None) -> list[np.ndarray]:
    d.mkdir(parents=True, exist_ok=True)
    T_hat = [float(t) for t in T_labels]
    data = samples(T_hat, n, seed)
    if counts is not None:
        data = [a[:c] for a, c in zip(data, counts)]
    lines = [f"N_SIMULACIONES {len(T_labels)}", f"STEP {steps}", f"L {L}",
             f"MIN_T {T_range[0]}", f"MAX_T {T_range[1]}", f"DELTA_T {T_range[2]}",
             f"MIN_P {P_range[0]}", f"MAX_P {P_range[1]}", f"DELTA_P {P_range[2]}",
             f"MIN_S {S_range[0]}", f"MAX_S {S_range[1]}", f"DELTA_S {S_range[2]}", f"BIN_FCT {bin_fct}"]
    for t, a in zip(T_labels, data):
        t9, p6 = pad(t, 9), pad(P_label, 6)
        lines.append(f"{t9} {p6} {len(a)}")
        (d / f"data_L{L}_T{t9}_P{p6}").write_text("".join("%.6g %.6g\n" % (h, r) for h, r in a))
    (d / "input").write_text("\n".join(lines) + "\n")
    return data


def read_table(path: Path) -> np.ndarray:
    rows = [list(map(float, ln.split())) for ln in Path(path).read_text().splitlines() if ln.strip()]
    return np.array(rows)
