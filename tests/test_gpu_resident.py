"""The resident Metropolis kernel (tuning fused = 2) against the oracle and
against the other launch plans.

The kernel keeps each state point's z-slabs in LDS for a whole graph of MC
steps and exchanges the slabs' boundary planes through versioned rings, so
the cases cover every form of the exchange: whole lattices in one window
(no exchange), two slabs per point (the neighbour above is also the one
below), one-plane slabs (a plane is both the top and the bottom), slabs of
unequal thickness, several slabs of several points per workgroup (each on
its own waves), Swendsen-Wang steps between one-step launches, and graphs
of 1 to 64 steps.  Bar: sigma/eta words,
labels and volume chains bit-exact (gpu_water3D.cu:1346-1417, 1890-2008,
3374-3466).
"""
import numpy as np
import pytest

from cvfbulkwater_amd import FP32_COMPENSATED, VOLUME_DEVICE, Params, Simulation
from oracle import py as O

pytestmark = pytest.mark.gpu


def points(Ls, nc=0, seed0=300):
    return [Params(L=L, T=0.55 + 0.4 * (i % 9) / 8, P=-0.1 + 0.9 * (i % 7) / 6, seed=seed0 + 17 * i, n_cluster=nc,
                   precision=FP32_COMPENSATED, volume_mode=VOLUME_DEVICE) for i, L in enumerate(Ls)]


def resident(pts, **tuning):
    sim = Simulation(pts)
    sim.tuning = {"fused": 2, **tuning}
    for i in range(len(pts)):
        assert sim.plan(i)["kind"] == "resident", (i, sim.plan(i))
    return sim


def same(a, b, what):
    sa, ea = a
    sb, eb = b
    bad_s, bad_e = np.flatnonzero(sa != sb), np.flatnonzero(ea != eb)
    assert bad_s.size == 0 and bad_e.size == 0, f"{what}: sigma differs at {bad_s[:8]}, eta at {bad_e[:8]}"


@pytest.mark.parametrize("L,max_T", [(4, 0), (8, 0), (12, 0), (16, 0), (32, 0), (8, 1), (32, 1), (24, 5), (8, 5),
                                     (16, 5)])
def test_single_point_matches_oracle_every_step(L, max_T):
    """One point under the automatic plan (a whole lattice in one window
    for L <= 12, two 8-plane slabs at L = 16, sixteen 2-plane slabs at
    L = 32), one-plane slabs (slab_planes = 1: a plane is both the top and
    the bottom of its slab), and slabs of at most five planes (unequal
    thickness: 3 + 5 planes at L = 8, five slabs of 2-5 planes in three
    workgroups at L = 16), each exchanging with its neighbours every step."""
    p = points([L])[0]
    sim = resident([p], slab_planes=max_T)
    T = sim.plan(0)["slab_planes"]
    assert 1 <= T <= (max_T or L)
    ora = O.OracleSim(L=L, T=p.T, P=p.P, seed=p.seed)
    for t in range(25):
        sim.step(1)
        ora.step(1)
        same(sim.get_config(), ora.get_config(), f"L={L} T={T} step {t}")
    a, b = sim.observe(), ora.measure()
    assert (a.nhb2, a.n_sigma) == (b.nhb2, b.n_sigma)


@pytest.mark.parametrize("Ls,steps", [
    ([16] * 256, (1, 7, 16, 40)),       # whole lattices, no exchange
    ([24] * 128, (3, 16, 64)),          # two slabs per point
    ([16, 24, 32, 48] * 8, (1, 2, 16)),  # mixed lattices, several slabs per workgroup
    ([8, 12] * 20, (5, 64, 1)),
])
def test_batches_match_single_runs(Ls, steps):
    """Batches whose automatic plan gives different slab forms, stepped in
    graphs of several lengths, against each point's default-plan run."""
    pts = points(Ls)
    sim = resident(pts, graph_steps=64)
    Ts = {sim.plan(i)["slab_planes"] for i in range(len(pts))}
    total = 0
    for n in steps:
        sim.step(n)
        total += n
    idx = sorted(set(range(0, len(pts), max(1, len(pts) // 12))) | {len(pts) - 1})
    for i in idx:
        single = Simulation(pts[i])
        single.step(total)
        same(sim.get_config(i), single.get_config(), f"point {i} L={pts[i].L} slabs {Ts}")
        assert sim.volume(i) == single.volume(), i
        single.close()


@pytest.mark.parametrize("block,gs", [(1024, 16), (256, 3), (512, 64)])
def test_mixed_batch_matches_oracle_every_step(block, gs):
    pts = points([16, 24, 32, 48, 8, 12] * 3)
    sim = resident(pts, block_chunk=block, graph_steps=gs)
    check = (0, 3, 4, 9, 17)
    oras = {i: O.OracleSim(L=pts[i].L, T=pts[i].T, P=pts[i].P, seed=pts[i].seed) for i in check}
    for t in range(12):
        sim.step(1)
        for i, o in oras.items():
            o.step(1)
            same(sim.get_config(i), o.get_config(), f"step {t} point {i} (L={pts[i].L}, T={sim.plan(i)['slab_planes']})")
    sim.step(30)
    for i, o in oras.items():
        o.step(30)
        same(sim.get_config(i), o.get_config(), f"after 42 steps point {i}")


@pytest.mark.parametrize("L", [8, 16, 32])
def test_cluster_steps_between_resident_launches(L):
    """With Swendsen-Wang steps each MC step is one resident launch followed
    by the SW kernels on the written-back lattice: labels and lattice equal
    to the oracle after every step."""
    p = points([L], nc=1)[0]
    sim = resident([p])
    ora = O.OracleSim(L=L, T=p.T, P=p.P, seed=p.seed, n_cluster=1)
    for t in range(20):
        sim.step(1)
        ora.step(1)
        same(sim.get_config(), ora.get_config(), f"step {t}")
        assert np.array_equal(sim.labels(), ora.labels()), f"labels step {t}"


def test_headline_batch_resident_equals_fused_plan():
    """The benchmark's 128-point batch under the resident kernel and under
    the fused-slab plan: identical lattices and volume chains after 40 steps
    (graphs of 16, 16, 8)."""
    import bench

    pts = bench.batch_params()
    a = resident(pts, graph_steps=16, block_chunk=1024)
    b = Simulation(pts)
    b.tuning = {"streams": 2, "fused": 1, "slab_planes": 0, "block_chunk": 1024, "chunk_sweeps": 6}
    a.step(40)
    b.step(40)
    for i in range(len(pts)):
        same(a.get_config(i), b.get_config(i), f"point {i} (L={pts[i].L})")
        assert a.volume(i) == b.volume(i), i


def test_resident_profile_and_histograms():
    """The instrumented steps and the per-step histogram record run the
    resident kernel one step per launch and leave the same state."""
    pts = points([16, 24, 8])
    a, b, c = resident(pts), resident(pts), resident(pts)
    prof = a.profile_steps(6)
    assert prof["resident"][1] == 1 and prof["resident"][0] > 0  # the 6 steps in one launch
    assert prof["begin"][1] == 1
    rec = b.step_histograms(6, pt=1)
    c.step(6)
    for i in range(3):
        same(a.get_config(i), c.get_config(i), f"profile point {i}")
        same(b.get_config(i), c.get_config(i), f"record point {i}")
    s, _ = c.get_config(1)
    assert np.array_equal(rec[-1], np.bincount(s, minlength=6))


@pytest.mark.parametrize("block", [512, 256])
def test_two_workgroups_per_cu(block):
    """streams = 2 under the resident kernel: two workgroups per CU, each
    with half the LDS, so slabs are thinner and more of them exchange;
    lattices equal to each point's default-plan run."""
    pts = points([16, 24, 32, 48] * 4)
    sim = resident(pts, streams=2, block_chunk=block, graph_steps=16)
    sim.step(21)
    for i in (0, 3, 7, 11, 15):
        single = Simulation(pts[i])
        single.step(21)
        same(sim.get_config(i), single.get_config(), f"point {i} L={pts[i].L}")
        assert sim.volume(i) == single.volume(), i
        single.close()


@pytest.mark.parametrize("max_T", [0, 1])
def test_large_lattice_L96(max_T):
    """One L=96 point (an L^2 plane is 36 KB of LDS): 2-plane slabs on 48
    workgroups, or one-plane slabs on 96, against the unfused kernels for
    8 steps (the oracle takes the point at its first step)."""
    p = points([96])[0]
    sim = resident([p], slab_planes=max_T)
    ref = Simulation(p)
    ref.tuning = {"fused": 0}
    ora = O.OracleSim(L=96, T=p.T, P=p.P, seed=p.seed)
    sim.step(1)
    ref.step(1)
    ora.step(1)
    same(sim.get_config(), ora.get_config(), "after step 1 (oracle)")
    sim.step(7)
    ref.step(7)
    same(sim.get_config(), ref.get_config(), "after 8 steps (unfused)")
    assert sim.volume() == ref.volume()
