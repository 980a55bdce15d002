"""The fast schedule (two-colour cell passes, CVF_SCHED_FAST) on the GPU:
bit-exact against the oracle's fast schedule (sigma/eta words, SW labels,
volume chains), independent of the launch configuration, and sampling the
reference's distribution (config 2 long-run means against SOURCE/data.out).
"""
import io
import json
import lzma
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cvfbulkwater_amd import (EXE_PATH, FP32_COMPENSATED, SCHED_FAST, VOLUME_DEVICE, VOLUME_HOST, CVFError,
                              Params, Simulation)
from oracle import py as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def fast_points(Ls, nc=0, nm=1, volume_mode=VOLUME_DEVICE, precision=0):
    return [Params(L=L, T=0.55 + 0.07 * i, P=-0.05 + 0.1 * i, seed=500 + 11 * i, n_metropolis=nm,
                   n_cluster=nc, schedule=SCHED_FAST, volume_mode=volume_mode, precision=precision)
            for i, L in enumerate(Ls)]


@pytest.mark.parametrize("Ls,nm,nc,steps", [([8], 1, 0, 40), ([8, 16, 12], 1, 1, 25), ([12, 24], 2, 0, 15),
                                            ([4, 20], 1, 1, 30)])
def test_fast_schedule_matches_oracle_every_few_steps(Ls, nm, nc, steps):
    pts = fast_points(Ls, nc=nc, nm=nm)
    eng = Simulation(pts)
    oras = [O.OracleSim(L=p.L, T=p.T, P=p.P, seed=p.seed, n_metropolis=nm, n_cluster=nc, schedule=1)
            for p in pts]
    for chunk in (1, steps // 2, steps - steps // 2 - 1):
        eng.step(chunk)
        for i, o in enumerate(oras):
            o.step(chunk)
            a, b = eng.get_config(i), o.get_config()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (Ls[i], chunk)
            if nc:
                assert np.array_equal(eng.labels(i), o.labels())
            assert eng.volume(i) == o.volume()
            m, n = eng.observe(i), o.measure()
            assert (m.nhb2, m.n_sigma) == (n.nhb2, n.n_sigma)


@pytest.mark.parametrize("tuning", [
    {"streams": 1, "block_sigma": 64, "block_eta": 64, "graph_steps": 1},
    {"streams": 2, "block_sigma": 512, "block_eta": 256, "graph_steps": 7},
    {"streams": 3, "block_sigma": 1024, "block_eta": 1024, "graph_steps": 16, "fused": 1},
])
def test_fast_schedule_tuning_does_not_change_results(tuning):
    pts = fast_points([8, 16, 12, 24], nc=1)
    ref, alt = Simulation(pts), Simulation(pts)
    alt.tuning = tuning
    ref.step(23)
    alt.step(23)
    for i in range(len(pts)):
        a, b = ref.get_config(i), alt.get_config(i)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(ref.labels(i), alt.labels(i))
        assert ref.volume(i) == alt.volume(i)


@pytest.mark.parametrize("Ls,nm,graph_steps", [([8, 16, 12], 1, 16), ([24], 2, 3), ([4, 20], 1, 5)])
def test_fast_schedule_slab_eta_matches_oracle(Ls, nm, graph_steps):
    """fused = 1 with the fast schedule: the eta sweep of each step as one
    LDS slab launch into the other word array (the colour passes then run on
    that array); sequences with an odd step count end with the per-family
    eta kernels.  Bit-exact with the oracle after calls of 1, 2, 3 and 9
    steps."""
    pts = fast_points(Ls, nc=0, nm=nm)
    eng = Simulation(pts)
    eng.tuning = {"fused": 1, "graph_steps": graph_steps}
    oras = [O.OracleSim(L=p.L, T=p.T, P=p.P, seed=p.seed, n_metropolis=nm, n_cluster=0, schedule=1)
            for p in pts]
    for n in (1, 2, 3, 9):
        eng.step(n)
        for i, o in enumerate(oras):
            o.step(n)
            a, b = eng.get_config(i), o.get_config()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (Ls[i], n)
            assert eng.volume(i) == o.volume()


def test_fast_schedule_autotune_restores_state():
    pts = fast_points([16, 24, 32], precision=FP32_COMPENSATED)
    a, b = Simulation(pts), Simulation(pts)
    a.step(5)
    b.step(5)
    chosen = b.autotune(4)
    assert b.tuning == chosen
    a.step(12)
    b.step(12)
    for i in range(len(pts)):
        x, y = a.get_config(i), b.get_config(i)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_fast_schedule_rejects_the_non_chessboard_mode():
    with pytest.raises(CVFError):
        Simulation(Params(L=8, schedule=SCHED_FAST, flag_chessboard=0))


def block_mean_se(x, nblocks=10):
    x = np.asarray(x, float)
    b = x[: len(x) // nblocks * nblocks].reshape(nblocks, -1).mean(axis=1)
    return x.mean(), b.std(ddof=1) / math.sqrt(nblocks)


def test_fast_schedule_config2_long_run_matches_reference_data_out():
    """Config 2 (L=32, T=0.815, P=0.45, the shipped input_config, 10^4 +
    10^5 steps sampled every 10) with the fast schedule: density, enthalpy
    and N_HB means against SOURCE/data.out.  The schedule samples the same
    distribution by another path, so the bar is 3 combined standard errors."""
    a = np.loadtxt(io.StringIO(lzma.decompress((GOLD / "input_config.xz").read_bytes()).decode()),
                   dtype=np.int64)
    sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=7, flag_config=1, schedule=SCHED_FAST,
                            volume_mode=VOLUME_HOST))
    assert sim.set_config(a[:, :6].T.astype(np.uint8).reshape(-1), a[:, 6:].T.astype(np.uint8).reshape(-1)) == 0
    sim.step(10000)
    rows = []
    for i in range(10000):
        sim.step(1 if i == 0 else 10)
        o = sim.observe()
        rows.append((o.H, o.V, o.n_hb))
    ours = np.array(rows)
    ref = np.loadtxt(io.StringIO(lzma.decompress((GOLD / "data.out.xz").read_bytes()).decode()))[:, 1:]
    N = 32 ** 3
    for name, f, g in (("rho", lambda r: N / r[:, 1], lambda r: N / r[:, 2]),
                       ("H/N", lambda r: r[:, 0] / N, lambda r: r[:, 1] / N),
                       ("N_HB/N", lambda r: r[:, 2] / N, lambda r: r[:, 3] / N)):
        m1, s1 = block_mean_se(f(ours))
        m2, s2 = block_mean_se(g(ref))
        z = (m1 - m2) / math.hypot(s1, s2)
        print(name, m1, m2, z)
        assert abs(z) < 3.0, (name, m1, s1, m2, s2)


def test_dropin_fast_schedule_matches_oracle(tmp_path):
    """cvf_water3d -schedule fast: data.out rows equal the oracle's fast
    schedule run from the same input."""
    from test_dropin import INPUT_DATA

    v = dict(P=0.45, T=0.70, seed=321, nc=0, neq=30, nloops=60, interval=10, flag_config=0, flag_dist=0,
             flag_corr=0)
    (tmp_path / "input_data").write_text(INPUT_DATA.format(**v))
    r = subprocess.run([str(EXE_PATH), "-schedule", "fast", "-L", "16"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ora = O.OracleSim(L=16, T=0.70, P=0.45, seed=321, schedule=1)
    ora.step(30)
    want = []
    for i in range(60):
        ora.step(1)
        if i % 10 == 0:
            want.append(ora.data_line(i).rstrip("\n"))
    assert (tmp_path / "data.out").read_text().splitlines() == want
    bad = subprocess.run([str(EXE_PATH), "-schedule", "quick"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=60)
    assert bad.returncode == 1 and "-schedule" in bad.stderr
