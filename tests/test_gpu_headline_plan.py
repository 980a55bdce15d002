"""The benchmark's headline batch at the launch plan the benchmark runs.

Workload: the 128 state points of the batched scan (T = 0.55 + 0.45 k/7,
P = -0.10 + 0.06 j, L = 16, 24, 32, 48 for point i % 4, own seed each,
FP32-compensated vdW, device volume move), tuned as the autotuner picks on
an MI355X: two stream groups, 1024-thread fused launches of six half-sweeps,
automatic slabs.  That gives 16-plane slabs at L = 16 (the whole lattice, so
a window repeats its own planes), 12 at L = 24, 16 at L = 32 and 8 at
L = 48.  Bar: every point bit-exact (sigma/eta words, volume chain) with its
own single-point run under the default plan, and one point of each lattice
size bit-exact with the oracle after every step (rules of
gpu_water3D.cu:1346-1417, 1890-2008, 3374-3466).
"""
import numpy as np
import pytest

import bench
from cvfbulkwater_amd import Simulation
from oracle import py as O

pytestmark = pytest.mark.gpu

HEADLINE_TUNING = {"streams": 2, "fused": 1, "slab_planes": 0, "block_chunk": 1024, "chunk_sweeps": 6}
EXPECTED_SLABS = {16: 16, 24: 12, 32: 16, 48: 8}


def headline_batch():
    pts = bench.batch_params()
    sim = Simulation(pts)
    sim.tuning = HEADLINE_TUNING
    assert {k: sim.tuning[k] for k in HEADLINE_TUNING} == HEADLINE_TUNING
    return pts, sim


def test_headline_plan_is_the_benchmarked_one():
    pts, sim = headline_batch()
    slabs = {}
    groups = set()
    for i, p in enumerate(pts):
        pl = sim.plan(i)
        slabs.setdefault(p.L, set()).add(pl["slab_planes"])
        groups.add(pl["group"])
        assert pl["chunk_sweeps"] == 6, (i, pl)
    assert slabs == {L: {T} for L, T in EXPECTED_SLABS.items()}
    assert groups == {0, 1}


def test_headline_batch_equals_single_runs():
    """20 steps in calls of 1, 3 and 16 (graphs of 1, 2 and 16 steps), then
    every one of the 128 points against its own default-plan run."""
    pts, sim = headline_batch()
    for n in (1, 3, 16):
        sim.step(n)
    assert sim.steps == 20
    for i, p in enumerate(pts):
        single = Simulation(p)
        assert single.plan(0)["slab_planes"] == 1  # one-plane slabs (the default for one small point)
        single.step(20)
        a, b = sim.get_config(i), single.get_config()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"point {i} (L={p.L})"
        assert sim.volume(i) == single.volume(), f"point {i} (L={p.L})"
        assert sim.observe(i).nhb2 == single.observe().nhb2, f"point {i}"
        single.close()


@pytest.mark.parametrize("points", [(0, 1, 2, 3), (124, 125, 126, 127)])
def test_headline_batch_points_match_oracle_every_step(points):
    """One point of each lattice size (whole-lattice L=16 windows, 12-plane
    L=24, 16-plane L=32 and 8-plane L=48 slabs) against the oracle after
    every one of 12 steps."""
    pts, sim = headline_batch()
    oras = {}
    for i in points:
        p = pts[i]
        assert sim.plan(i)["slab_planes"] == EXPECTED_SLABS[p.L]
        oras[i] = O.OracleSim(L=p.L, T=p.T, P=p.P, seed=p.seed, n_metropolis=1, n_cluster=0)
    for t in range(12):
        sim.step(1)
        for i, ora in oras.items():
            ora.step(1)
            s1, e1 = sim.get_config(i)
            s2, e2 = ora.get_config()
            bad_s = np.flatnonzero(s1 != s2)
            bad_e = np.flatnonzero(e1 != e2)
            assert bad_s.size == 0 and bad_e.size == 0, \
                f"step {t} point {i} (L={pts[i].L}): sigma differs at {bad_s[:8]}, eta at {bad_e[:8]}"
    for i, ora in oras.items():
        a, b = sim.observe(i), ora.measure()
        assert (a.nhb2, a.n_sigma) == (b.nhb2, b.n_sigma), i


def test_tune_cache_is_read_back_and_keyed_by_rule(tmp_path, monkeypatch):
    """CVF_TUNE_CACHE: the first autotune stores its choice, a second handle
    of the same shape reads it back without timing (no new line), the state
    and trajectory are those of an untuned run, and a batch of the same
    lattice sizes under the non-chessboard rule gets its own entry."""
    from cvfbulkwater_amd import VOLUME_DEVICE, Params

    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("CVF_TUNE_CACHE", str(cache))

    def pts(chess):
        return [Params(L=L, T=0.6 + 0.05 * i, P=0.3, seed=70 + i, flag_chessboard=chess, volume_mode=VOLUME_DEVICE)
                for i, L in enumerate((8, 16, 12))]

    a = Simulation(pts(1))
    a.step(5)
    ta = a.autotune(3)
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "|ncb0|" in lines[0]
    b = Simulation(pts(1))
    b.step(5)
    tb = b.autotune(3)
    assert tb == ta and b.tuning == ta
    assert cache.read_text().splitlines() == lines
    ref = Simulation(pts(1))
    ref.step(25)
    a.step(20)
    b.step(20)
    for i in range(3):
        r = ref.get_config(i)
        for sim in (a, b):
            c = sim.get_config(i)
            assert np.array_equal(r[0], c[0]) and np.array_equal(r[1], c[1]), i
            assert sim.volume(i) == ref.volume(i)
    nc = Simulation(pts(0))
    nc.autotune(3)
    lines2 = cache.read_text().splitlines()
    assert len(lines2) == 2 and "|ncb3|" in lines2[1]
