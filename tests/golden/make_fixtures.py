#!/usr/bin/env python3
"""Write the golden fixtures from the data files the reference ships.

Reads (as data, nothing is executed) from the reference checkout:
  SOURCE/distances, SOURCE/data.out, SOURCE/cluster.out, SOURCE/water3D.log,
  SOURCE/input_data, SOURCE/input_config, SOURCE/config.out,
  PROCESSED_DATA/CVF_isobars.tar.gz (L32_P*_density.dat, .dat.SI,
  _AlphaPderivate.dat.SI, _KTderivate.dat.SI)
and writes into tests/golden/:
  distances_L32, input_data, data.out.xz, input_config.xz, config.out.xz,
  reference_values.json (log summary lines, cluster.out first row, SHA-256 of
  the configs, N_HB / N_sigma of both configs computed here with numpy, and
  the published L=32 density isobars), isobars_L32.tar.xz (the processed
  isobar files above, byte for byte, for the thermodynamic post-processing).

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""
from __future__ import annotations

import hashlib
import io
import json
import lzma
import sys
import tarfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def config_counts(text: str, L: int = 32) -> dict:
    """N_HB (both-ends eta average, as measure() gpu_water3D.cu:2037-2063) and
    N_sigma of a config file, vectorised; independent of the engine/oracle."""
    a = np.loadtxt(io.StringIO(text), dtype=np.int64)
    N = L ** 3
    assert a.shape == (N, 12)
    s = a[:, :6].T.reshape(6, L, L, L)  # [arm, z, y, x]
    e = a[:, 6:].T.reshape(6, L, L, L)
    nhb2 = 0
    for ax, axis in ((0, 2), (1, 1), (2, 0)):  # x -> last axis
        a0, a1 = 2 * ax, 2 * ax + 1
        s1 = np.roll(s[a1], -1, axis=axis)
        e1 = np.roll(e[a1], -1, axis=axis)
        eq = s[a0] == s1
        nhb2 += int((eq * (e[a0] + e1)).sum())
    nsig = 0
    for j in range(6):
        for k in range(j + 1, 6):
            nsig += int((s[j] == s[k]).sum())
    active = e.sum(axis=0)
    consistent = all(
        np.array_equal(e[2 * ax], np.roll(e[2 * ax + 1], -1, axis=axis))
        for ax, axis in ((0, 2), (1, 1), (2, 0)))
    return {"nhb2": nhb2, "n_hb": nhb2 / 2, "n_sigma": nsig,
            "eta_active_min": int(active.min()), "eta_active_max": int(active.max()),
            "eta_ends_consistent": bool(consistent)}


def main(ref: Path) -> None:
    src = ref / "SOURCE"
    (HERE / "distances_L32").write_bytes((src / "distances").read_bytes())
    (HERE / "input_data").write_bytes((src / "input_data").read_bytes())
    for name in ("data.out", "input_config", "config.out"):
        raw = (src / name).read_bytes()
        (HERE / f"{name}.xz").write_bytes(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME))
    log = (src / "water3D.log").read_text().splitlines()
    cl = (src / "cluster.out").read_text().splitlines()
    vals = {
        "water3D_log": log,
        "log_magnetization_line": next(l for l in log if l.startswith("Magnetization")),
        "log_energy_line": next(l for l in log if l.startswith("Energy")),
        "cluster_out_first_row": cl[0],
        "cluster_out_rows": len(cl),
        "cluster_out_all_rows_same_but_time": len({" ".join(r.split()[1:]) for r in cl}) == 1,
        "sha256": {n: hashlib.sha256((src / n).read_bytes()).hexdigest()
                   for n in ("input_config", "config.out", "data.out")},
        "input_config_counts": config_counts((src / "input_config").read_text()),
        "config_out_counts": config_counts((src / "config.out").read_text()),
    }
    iso = {}
    with tarfile.open(ref / "PROCESSED_DATA" / "CVF_isobars.tar.gz") as tf:
        for m in tf.getmembers():
            n = Path(m.name).name
            if n.startswith("L32_P") and n.endswith("_density.dat"):
                P = n[len("L32_P"):-len("_density.dat")]
                rows = np.loadtxt(io.TextIOWrapper(tf.extractfile(m)), ndmin=2)
                iso[P] = rows.tolist()
    vals["published_density_isobars_L32"] = dict(sorted(iso.items(), key=lambda kv: float(kv[0])))
    keep = ("_density.dat", "_density.dat.SI", "_AlphaPderivate.dat.SI", "_KTderivate.dat.SI")
    buf = io.BytesIO()
    with tarfile.open(ref / "PROCESSED_DATA" / "CVF_isobars.tar.gz") as tf, \
            tarfile.open(fileobj=buf, mode="w") as out:
        for m in sorted(tf.getmembers(), key=lambda m: m.name):
            n = Path(m.name).name
            if m.isfile() and n.startswith("L32_P") and n.endswith(keep):
                data = tf.extractfile(m).read()
                info = tarfile.TarInfo(n)
                info.size = len(data)
                out.addfile(info, io.BytesIO(data))
    (HERE / "isobars_L32.tar.xz").write_bytes(lzma.compress(buf.getvalue(), preset=9 | lzma.PRESET_EXTREME))
    (HERE / "reference_values.json").write_text(json.dumps(vals, indent=1) + "\n")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
