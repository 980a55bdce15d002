"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bar: bit-exact sigma/eta words and Swendsen-Wang labels, exact integer
N_HB / N_sigma, identical volume chains (r_cell, V, E_vdW compared with ==),
E/H/V within 1e-12 relative (the oracle sums per cell like measure(), the
engine from integer totals).
"""
import numpy as np
import pytest

from cvfbulkwater_amd import FP32_COMPENSATED, VOLUME_DEVICE, VOLUME_HOST, CVFError, Params, Simulation
from oracle import py as O

pytestmark = pytest.mark.gpu


def pair(L, T=0.815, P=0.45, seed=44648604, nm=1, nc=0, flag_config=0, volume_mode=VOLUME_HOST):
    eng = Simulation(Params(L=L, T=T, P=P, seed=seed, n_metropolis=nm, n_cluster=nc,
                            flag_config=flag_config, volume_mode=volume_mode))
    ora = O.OracleSim(L=L, T=T, P=P, seed=seed, n_metropolis=nm, n_cluster=nc,
                      flag_config=flag_config)
    return eng, ora


def assert_same_state(eng, ora, what=""):
    s1, e1 = eng.get_config()
    s2, e2 = ora.get_config()
    if not np.array_equal(s1, s2):
        bad = np.flatnonzero(s1 != s2)
        raise AssertionError(f"{what}: sigma differs at {bad.size} sites, first {bad[:8]}")
    if not np.array_equal(e1, e2):
        bad = np.flatnonzero(e1 != e2)
        raise AssertionError(f"{what}: eta differs at {bad.size} sites, first {bad[:8]}")


def assert_same_obs(eng, ora, what=""):
    a, b = eng.observe(), ora.measure()
    assert a.nhb2 == b.nhb2, what
    assert a.n_sigma == b.n_sigma, what
    assert a.r_cell == b.r_cell, what
    for k in ("E", "H", "V", "n_hb", "m"):
        x, y = getattr(a, k), getattr(b, k)
        assert abs(x - y) <= 1e-12 * max(1.0, abs(y)), (what, k, x, y)


@pytest.mark.parametrize("flag_config", [0, 2])
def test_initial_state_matches(flag_config):
    eng, ora = pair(8, flag_config=flag_config)
    assert_same_state(eng, ora, "init")
    assert_same_obs(eng, ora, "init")


def test_config1_L8_every_step():
    """Config 1 shape: L=8, T=0.815, P=0.45, 1 Metropolis + 1 volume move,
    1000 + 2000 steps; state compared after every step."""
    eng, ora = pair(8)
    for t in range(3000):
        eng.step(1)
        ora.step(1)
        if t < 300 or t % 10 == 0:
            assert_same_state(eng, ora, f"step {t}")
    assert_same_obs(eng, ora, "end")
    assert eng.volume() == ora.volume()


def test_config2_L32_first_steps():
    """Config 2 shape (L=32), first 200 steps compared every step."""
    eng, ora = pair(32)
    for t in range(200):
        eng.step(1)
        ora.step(1)
        assert_same_state(eng, ora, f"step {t}")
        if t % 20 == 0:
            assert_same_obs(eng, ora, f"step {t}")
    assert eng.volume() == ora.volume()


@pytest.mark.parametrize("L,T", [(8, 0.70), (16, 0.70), (12, 0.55)])
def test_cluster_form_labels_every_step(L, T):
    """Config 3 shape (1 Metropolis + 1 SW per step) at small L: labels,
    sigma and eta identical after every step."""
    eng, ora = pair(L, T=T, nc=1)
    for t in range(60):
        eng.step(1)
        ora.step(1)
        assert np.array_equal(eng.labels(), ora.labels()), f"labels step {t}"
        assert_same_state(eng, ora, f"step {t}")
        if t % 10 == 0:
            a = eng.cluster_stats()
            b = ora.cluster_stats()
            assert (a.largest, a.second, a.n_perc, a.n_clusters) == tuple(b[:4]), t
            assert a.mean == b[4] and a.var == b[5], t
    assert_same_obs(eng, ora)


@pytest.mark.parametrize("T,P,rounded", [(0.05, 0.3, True), (0.2, 2.0, False), (0.815, 0.45, False)])
def test_cluster_row_exact_and_rounded_sums(T, P, rounded):
    """cluster.out's float sums (print_cluster_statistics, gpu_water3D.cu:
    2691-2718, sizes and squares added as floats in label order): with
    every partial sum at most 2^24 each addition is exact and the engine
    takes the row from device integer totals; a percolating cluster
    (T=0.05: sum of squares about 6e8) rounds, and the host adds the sizes
    in label order.  Both rows equal the oracle's, float bits included."""
    eng, ora = pair(16, T=T, P=P, seed=91, nc=1)
    for t in range(40):
        eng.step(1)
        ora.step(1)
        if t % 10 == 9:
            lab = eng.labels()
            assert np.array_equal(lab, ora.labels()), t
            sz = np.bincount(lab)
            sq = int((sz[sz > 0].astype(np.int64) ** 2).sum())
            assert (sq > 1 << 24) == rounded, (t, sq)
            a, b = eng.cluster_stats(), ora.cluster_stats()
            assert (a.largest, a.second, a.n_perc, a.n_clusters, a.mean, a.var) == tuple(b), (t, sq)


@pytest.fixture(scope="module")
def config3_run():
    """One engine and one oracle run of config 3 shared by its segments, and
    the number of steps compared so far."""
    eng, ora = pair(32, T=0.70, nc=1)
    run = {"eng": eng, "ora": ora, "done": 0}
    yield run
    eng.close()


CONFIG3_SEGMENT = 1000


@pytest.mark.parametrize("segment", range(7))
def test_config3_L32_labels_every_step(config3_run, segment):
    """Config 3 over its whole schedule, 2000 equilibration + 5000 production
    steps, in seven segments of 1000 steps on one run: L=32, T=0.70, P=0.45,
    1 Metropolis + 1 Swendsen-Wang step per MC step (gpu_water3D.cu:1444-1603,
    labels 2440-2529).  Sigma, eta and the 6N canonical labels identical to
    the oracle after every step, the cluster.out row every 50 steps, the
    observables and the volume chain at the end of each segment.  A segment
    run on its own first steps (and compares) the segments before it."""
    eng, ora = config3_run["eng"], config3_run["ora"]
    end = (segment + 1) * CONFIG3_SEGMENT
    while config3_run["done"] < end:
        t = config3_run["done"]
        eng.step(1)
        ora.step(1)
        config3_run["done"] = t + 1
        la, lb = eng.labels(), ora.labels()
        if not np.array_equal(la, lb):
            bad = np.flatnonzero(la != lb)
            raise AssertionError(f"labels step {t}: {bad.size} nodes differ, first {bad[:8]}")
        assert_same_state(eng, ora, f"step {t}")
        if t % 50 == 49:
            a, b = eng.cluster_stats(), ora.cluster_stats()
            assert (a.largest, a.second, a.n_perc, a.n_clusters, a.mean, a.var) == tuple(b), t
    assert_same_obs(eng, ora, f"after {end} steps")
    assert eng.volume() == ora.volume()


def test_config5_L96_labels_first_steps():
    """Config 5 shape: L=96, T=0.65, P=0.30, 1 Metropolis + 1 Swendsen-Wang
    + 1 volume move, FP32-compensated vdW: the lattice (sigma, eta) and the
    5.3 M labels identical to the oracle after each of the first 30 steps,
    the cluster.out row every 10 (the volume chain does not enter the
    lattice rules)."""
    kw = dict(L=96, T=0.65, P=0.30, seed=44648604, n_cluster=1)
    eng = Simulation(Params(precision=FP32_COMPENSATED, volume_mode=VOLUME_DEVICE, **kw))
    ora = O.OracleSim(**kw)
    for t in range(30):
        eng.step(1)
        ora.step(1)
        assert np.array_equal(eng.labels(), ora.labels()), f"labels step {t}"
        assert_same_state(eng, ora, f"step {t}")
        if t % 10 == 9:
            a, b = eng.cluster_stats(), ora.cluster_stats()
            assert (a.largest, a.second, a.n_perc, a.n_clusters, a.mean, a.var) == tuple(b), t
    a, b = eng.observe(), ora.measure()
    assert (a.nhb2, a.n_sigma) == (b.nhb2, b.n_sigma)
    eng.close()


def test_device_volume_chain_equals_host_chain():
    """The device volume move (batched form) and the host one (plain form)
    share the project exp/log: identical r_cell chains."""
    eng_h, ora = pair(16)
    eng_d = Simulation(Params(L=16, volume_mode=VOLUME_DEVICE))
    for _ in range(5):
        eng_h.step(200)
        eng_d.step(200)
        ora.step(200)
        assert eng_h.volume() == eng_d.volume() == ora.volume()


def test_config_roundtrip_and_normalisation():
    rng = np.random.default_rng(5)
    L, N = 8, 512
    eng, ora = pair(L)
    s = rng.integers(0, 6, 6 * N, dtype=np.uint8)
    e = rng.integers(0, 2, 6 * N, dtype=np.uint8)
    nb1 = eng.set_config(s, e)
    nb2 = ora.set_config(s, e)
    assert nb1 == nb2 > 0
    assert_same_state(eng, ora, "after set_config")
    for t in range(50):
        eng.step(1)
        ora.step(1)
    assert_same_state(eng, ora, "after 50 steps")


@pytest.mark.parametrize("L,nc,flag_config", [(8, 0, 0), (16, 0, 2), (12, 1, 0), (8, 1, 2), (32, 0, 0)])
def test_nonchessboard_mode_every_step(L, nc, flag_config):
    """Flag_chessboard = 0: eta fixed, HB gated by the per-cell counters;
    sigma, eta, labels and the capped N_HB identical to the oracle."""
    kw = dict(L=L, T=0.6, P=0.45, seed=77, n_cluster=nc, flag_config=flag_config)
    eng = Simulation(Params(flag_chessboard=0, **kw))
    ora = O.OracleSim(flag_chessboard=0, **kw)
    assert_same_obs(eng, ora, "init")
    for t in range(60 if L < 32 else 20):
        eng.step(1)
        ora.step(1)
        assert_same_state(eng, ora, f"step {t}")
        if nc:
            assert np.array_equal(eng.labels(), ora.labels()), f"labels step {t}"
        if t % 10 == 0:
            assert_same_obs(eng, ora, f"step {t}")
    assert eng.volume() == ora.volume()


@pytest.mark.parametrize("tuning", [{"fused": 1, "slab_planes": 1}, {"fused": 1, "slab_planes": 2, "block_chunk": 256},
                                    {"fused": 1, "slab_planes": 1, "chunk_sweeps": 12, "graph_steps": 3},
                                    {"fused": 0}])
@pytest.mark.parametrize("L", [4, 12])
def test_nonchessboard_fused_slabs_match_oracle(tuning, L):
    """The non-chessboard rule in the fused slab kernel (its z half-sweeps
    write the facing neighbour's HB counter, so each widens the window on
    both sides): sigma, eta and N_HB equal to the oracle after calls of 1, 2,
    5 and 12 steps, for thin slabs (the window repeats planes at L=4),
    whole-pass launches and the unfused kernels."""
    kw = dict(L=L, T=0.6, P=0.45, seed=78, n_cluster=0)
    eng = Simulation(Params(flag_chessboard=0, volume_mode=VOLUME_DEVICE, **kw))
    eng.tuning = tuning
    ora = O.OracleSim(flag_chessboard=0, **kw)
    done = 0
    for n in (1, 2, 5, 12):
        eng.step(n)
        ora.step(n)
        done += n
        assert_same_state(eng, ora, f"after {done} steps")
        assert_same_obs(eng, ora, f"after {done} steps")


def test_nonchessboard_input_config_keeps_eta():
    L, N = 8, 512
    rng = np.random.default_rng(9)
    s = rng.integers(0, 6, 6 * N).astype(np.uint8)
    e = rng.integers(0, 2, 6 * N).astype(np.uint8)
    kw = dict(L=L, T=0.6, P=0.45, seed=3, n_cluster=1, flag_config=1)
    eng = Simulation(Params(flag_chessboard=0, **kw))
    ora = O.OracleSim(flag_chessboard=0, **kw)
    assert eng.set_config(s, e) == 0 and ora.set_config(s, e) == 0
    assert_same_state(eng, ora, "set")
    for t in range(40):
        eng.step(1)
        ora.step(1)
        assert_same_state(eng, ora, f"step {t}")
        assert np.array_equal(eng.labels(), ora.labels()), f"labels step {t}"
    assert np.array_equal(eng.get_config()[1], e)
    assert_same_obs(eng, ora, "end")


def test_batch_mixing_chessboard_modes_equals_single_runs():
    pts = [Params(L=L, T=0.6 + 0.05 * i, P=0.45, seed=500 + i, n_cluster=1, flag_chessboard=i % 2,
                  volume_mode=VOLUME_DEVICE) for i, L in enumerate([8, 16, 12, 8])]
    batch = Simulation(pts)
    singles = [Simulation(p) for p in pts]
    batch.step(30)
    for i, sim in enumerate(singles):
        sim.step(30)
        a, b = batch.get_config(i), sim.get_config()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
        assert batch.observe(i).nhb2 == sim.observe().nhb2, i


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("L,nm,nc,steps", [(4, 1, 1, 40), (8, 2, 0, 30), (12, 3, 1, 20), (64, 1, 0, 3)])
def test_fused_and_unfused_paths_edge_cases(fused, L, nm, nc, steps):
    """Both Metropolis paths against the oracle at the edges of the fused
    kernel: L=4 (window planes repeat around the lattice), several sweeps per
    step (2 nm fused launches), and L=64 (one-plane slabs, the largest L
    whose window fits in LDS)."""
    kw = dict(L=L, T=0.65, P=0.3, seed=4242, n_metropolis=nm, n_cluster=nc)
    eng = Simulation(Params(volume_mode=VOLUME_DEVICE, **kw))
    eng.tuning = {"fused": fused}
    ora = O.OracleSim(**kw)
    for t in range(steps):
        eng.step(1)
        ora.step(1)
        assert_same_state(eng, ora, f"step {t}")
        if nc:
            assert np.array_equal(eng.labels(), ora.labels()), f"labels step {t}"
    assert_same_obs(eng, ora, "end")


@pytest.mark.parametrize("nm,nc,graph_steps,steps", [(1, 0, 16, 37), (1, 0, 3, 20), (2, 0, 5, 13), (3, 0, 4, 9),
                                                     (1, 1, 16, 21), (3, 1, 4, 9)])
def test_whole_pass_launches_match_oracle(nm, nc, graph_steps, steps):
    """chunk_sweeps = 12: one fused launch per Metropolis pass.  The state
    alternates between the two word arrays (the SW kernels follow it) and a
    sequence with an odd launch count runs its last step in 6-sweep launches;
    bit-exact with the oracle (and the SW labels) after calls of 1, 2, 3 and
    `steps` steps (several graph sizes each)."""
    kw = dict(L=16, T=0.7, P=0.4, seed=2024, n_metropolis=nm, n_cluster=nc)
    eng = Simulation(Params(volume_mode=VOLUME_DEVICE, **kw))
    eng.tuning = {"fused": 1, "chunk_sweeps": 12, "graph_steps": graph_steps, "slab_planes": 1}
    assert eng.tuning["chunk_sweeps"] == 12
    ora = O.OracleSim(**kw)
    done = 0
    for n in (1, 2, 3, steps):
        eng.step(n)
        ora.step(n)
        done += n
        assert_same_state(eng, ora, f"after {done} steps")
        if nc:
            assert np.array_equal(eng.labels(), ora.labels()), f"labels after {done} steps"
    assert_same_obs(eng, ora, "end")


def test_whole_pass_launches_in_a_batch_with_a_wide_point():
    """A whole-pass tuning on a batch whose L=64 point cannot hold the wider
    window: that stream group keeps 6-sweep launches, the other takes whole
    passes, and every point stays equal to its default-tuning run."""
    pts = [Params(L=L, T=0.6 + 0.05 * i, P=0.2, seed=31 + i, volume_mode=VOLUME_DEVICE)
           for i, L in enumerate((64, 8, 12, 16))]
    ref, alt = Simulation(pts), Simulation(pts)
    alt.tuning = {"fused": 1, "chunk_sweeps": 12, "streams": 2, "slab_planes": 0}
    ref.step(11)
    alt.step(11)
    for i in range(len(pts)):
        a, b = ref.get_config(i), alt.get_config(i)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
        assert ref.volume(i) == alt.volume(i)
    with pytest.raises(CVFError):
        alt.tuning = {"chunk_sweeps": 8}


def test_lattice_above_the_fused_limit_runs_unfused():
    """L=68: one plane plus the worst-case halo exceeds the LDS, so the
    handle runs the unfused path even when the fused one is requested."""
    kw = dict(L=68, T=0.7, P=0.45, seed=99)
    eng = Simulation(Params(volume_mode=VOLUME_DEVICE, **kw))
    eng.tuning = {"fused": 1}
    ora = O.OracleSim(**kw)
    for t in range(2):
        eng.step(1)
        ora.step(1)
        assert_same_state(eng, ora, f"step {t}")
    assert eng.volume() == ora.volume()


def test_many_point_batch_equals_single_runs():
    pts = [Params(L=(8, 12)[i % 2], T=0.5 + 0.01 * i, P=0.1 * (i % 5), seed=77 + i, volume_mode=VOLUME_DEVICE)
           for i in range(40)]
    batch = Simulation(pts)
    batch.step(25)
    for i in (0, 1, 17, 38, 39):
        single = Simulation(pts[i])
        single.step(25)
        a, b = batch.get_config(i), single.get_config()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
        assert batch.volume(i) == single.volume()
