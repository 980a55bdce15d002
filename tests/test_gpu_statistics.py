"""Statistical parity with the reference's published output.

The reference's trajectories come from cuRAND streams that cannot be
reproduced, so parity with its results is statistical: long-run densities
against the published L=32 isobars (PROCESSED_DATA/CVF_isobars.tar.gz,
L32_P<P>_density.dat: T, rho = N/<V>, error = N sqrt(3 var(V)/n)/<V>^2,
CVF_ThermodynamicAnalysis/calculadora.sh).  The engine's own standard error
is a 20-block standard error of <V> carried to N/<V>; z = (ours - published)
/ hypot(ours_se, published_err).  The z values are printed (pytest -rP).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import bench
from cvfbulkwater_amd import Params, Simulation

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
PUB = json.loads((GOLD / "reference_values.json").read_text())["published_density_isobars_L32"]


def published(T, P):
    for row in PUB[P]:
        if abs(row[0] - T) < 1e-9:
            return row[1], row[2]
    raise KeyError((T, P))


def density_z(V, N, want, err, nblocks=20):
    V = np.asarray(V, float)
    b = V[: len(V) // nblocks * nblocks].reshape(nblocks, -1).mean(axis=1)
    mV = V.mean()
    seV = b.std(ddof=1) / math.sqrt(nblocks)
    rho, se = N / mV, N * seV / mV ** 2
    return rho, se, (rho - want) / math.hypot(se, err)


def test_config3_density_vs_published_isobar():
    """Config 3: L=32, T=0.70, P=0.45, 1 Metropolis + 1 Swendsen-Wang per
    step, 2000 equilibration + 50,000 sampling steps sampled every 10:
    N/<V> within 2 sigma of the published 0.669334 +- 2.0e-5."""
    sim = Simulation(Params(L=32, T=0.70, P=0.45, seed=7, n_cluster=1))
    sim.step(2000)
    V = []
    for _ in range(5000):
        sim.step(10)
        V.append(sim.observe().V)
    want, err = published(0.70, "0.45")
    rho, se, z = density_z(V, 32 ** 3, want, err)
    print(f"config 3 density: ours {rho:.6f} +- {se:.1e}, published {want:.6f} +- {err:.1e}, z = {z:+.2f}")
    assert abs(z) < 2.0, (rho, se, want, err, z)
    c = sim.cluster_stats()
    assert c.n_clusters > 1 and c.largest >= c.second > 0


# 32 published (T, P) grid points for the batch's L=32 replicas: four isobars,
# eight temperatures each, away from the low-T region where the published
# runs would need the longest equilibration
PUB_T = (0.62, 0.66, 0.70, 0.74, 0.80, 0.86, 0.94, 1.00)
PUB_P = ("0.0", "0.3", "0.5", "0.7")


def test_batched_L32_replicas_vs_published_isobars():
    """The headline batch (128 points, L = 16/24/32/48, FP32-compensated vdW,
    device volume move, the benchmark's launch plan) with its 32 L=32
    replicas moved onto published grid points; 10^4 equilibration + 10^5
    sampling steps sampled every 10 (the reference's production run).
    32 comparisons: every |z| < 3.5 (family-wise about 1.5 % under a correct
    engine), at least 28 of 32 within 2 sigma (30.5 expected), and
    sum z^2 below the 0.999 quantile of chi^2(32) (62.5)."""
    pts = bench.batch_params()
    l32 = [i for i, p in enumerate(pts) if p.L == 32]
    grid = [(T, P) for P in PUB_P for T in PUB_T]
    assert len(l32) == len(grid) == 32
    for i, (T, P) in zip(l32, grid):
        pts[i].T, pts[i].P = T, float(P)
    sim = Simulation(pts)
    sim.tuning = {"streams": 2, "fused": 1, "slab_planes": 0, "block_chunk": 1024, "chunk_sweeps": 6}
    sim.step(10000)
    V = {i: [] for i in l32}
    for _ in range(10000):
        sim.step(10)
        for i in l32:
            V[i].append(sim.observe(i).V)
    zs = []
    for i, (T, P) in zip(l32, grid):
        want, err = published(T, P)
        rho, se, z = density_z(V[i], 32 ** 3, want, err)
        zs.append(z)
        print(f"T={T:.2f} P={P:>3}: ours {rho:.6f} +- {se:.1e}, published {want:.6f} +- {err:.1e}, z = {z:+.2f}")
    zs = np.array(zs)
    print(f"within 2 sigma: {(np.abs(zs) < 2).sum()}/32, max |z| = {np.abs(zs).max():.2f}, "
          f"sum z^2 = {(zs ** 2).sum():.1f}")
    assert np.abs(zs).max() < 3.5, zs
    assert (np.abs(zs) < 2).sum() >= 28, zs
    assert (zs ** 2).sum() < 62.5, zs
