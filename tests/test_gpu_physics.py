"""GPU tests of the physics and of the batched / precision forms."""
import io
import json
import lzma
import math
from pathlib import Path

import numpy as np
import pytest

from cvfbulkwater_amd import (FP32_COMPENSATED, FP64, VOLUME_DEVICE, VOLUME_HOST, Params, Simulation,
                              vdw_energy, vdw_pair_energy)
from oracle import py as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_values.json").read_text())


def load_config(name):
    a = np.loadtxt(io.StringIO(lzma.decompress((GOLD / name).read_bytes()).decode()), dtype=np.int64)
    return a[:, :6].T.astype(np.uint8).reshape(-1), a[:, 6:].T.astype(np.uint8).reshape(-1)


def block_mean_se(x, nblocks=10):
    x = np.asarray(x, float)
    b = x[: len(x) // nblocks * nblocks].reshape(nblocks, -1).mean(axis=1)
    return x.mean(), b.std(ddof=1) / math.sqrt(nblocks)


def test_config2_long_run_matches_reference_data_out():
    """Config 2: L=32, T=0.815, P=0.45 from the shipped input_config,
    10^4 + 10^5 steps sampled every 10 (the reference's own run).  Long-run
    density, enthalpy and N_HB within 2 sigma of SOURCE/data.out (sigma =
    combined 10-block standard errors)."""
    s, e = load_config("input_config.xz")
    sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, flag_config=1))
    assert sim.set_config(s, e) == 0
    sim.step(10000)
    rows = []
    for i in range(100000 // 10):
        sim.step(1 if i == 0 else 10)
        o = sim.observe()
        rows.append((o.E, o.H, o.V, o.n_hb, o.m, o.r_cell))
    ours = np.array(rows)
    ref = np.loadtxt(io.StringIO(lzma.decompress((GOLD / "data.out.xz").read_bytes()).decode()))[:, 1:]
    N = 32 ** 3
    report = {}
    for name, f in (("rho", lambda a: N / a[:, 2]), ("H/N", lambda a: a[:, 1] / N),
                    ("N_HB/N", lambda a: a[:, 3] / N)):
        m1, s1 = block_mean_se(f(ours))
        m2, s2 = block_mean_se(f(ref))
        z = (m1 - m2) / math.hypot(s1, s2)
        report[name] = (m1, m2, z)
        assert abs(z) < 2.0, (name, m1, m2, s1, s2)
    print(report)


def test_cluster_out_degenerate_row_engine():
    sim = Simulation(Params(L=32))
    sim.step(3)
    c = sim.cluster_stats()
    line = "%d %d %d %d %f %f %d" % (0, c.largest, c.second, c.n_perc, c.mean, c.var, c.n_clusters)
    assert line == REF["cluster_out_first_row"]


def test_shipped_configs_through_engine():
    for name, key in (("input_config.xz", "input_config_counts"), ("config.out.xz", "config_out_counts")):
        s, e = load_config(name)
        sim = Simulation(Params(L=32, flag_config=1))
        assert sim.set_config(s, e) == 0
        o = sim.observe()
        assert (o.nhb2, o.n_sigma) == (REF[key]["nhb2"], REF[key]["n_sigma"])


def batch_points(Ls, precision=FP64, nc=0):
    return [Params(L=L, T=0.55 + 0.05 * i, P=-0.1 + 0.1 * i, seed=1000 + 7 * i, n_cluster=nc,
                   precision=precision, volume_mode=VOLUME_DEVICE) for i, L in enumerate(Ls)]


@pytest.mark.parametrize("nc", [0, 1])
def test_batch_equals_sequential_single_runs(nc):
    """A batch of uneven-L state points advanced in one launch sequence is
    bit-identical to the same points run one at a time."""
    Ls = [8, 16, 12, 24, 8, 20]
    pts = batch_points(Ls, nc=nc)
    batch = Simulation(pts)
    singles = [Simulation(p) for p in pts]
    for _ in range(3):
        batch.step(40)
        for s in singles:
            s.step(40)
        for i, s in enumerate(singles):
            a, b = batch.get_config(i), s.get_config()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i
            assert batch.volume(i) == s.volume()
            if nc:
                assert np.array_equal(batch.labels(i), s.labels())
            assert batch.observe(i).nhb2 == s.observe().nhb2


def test_batch_points_match_oracle():
    pts = batch_points([8, 12, 16], nc=1)
    batch = Simulation(pts)
    oras = [O.OracleSim(L=p.L, T=p.T, P=p.P, seed=p.seed, n_cluster=1) for p in pts]
    batch.step(30)
    for i, o in enumerate(oras):
        o.step(30)
        a = batch.get_config(i)
        b = o.get_config()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(batch.labels(i), o.labels())
        assert batch.volume(i) == o.volume()


@pytest.mark.parametrize("L", [32, 96])
def test_fp32_compensated_vdw_within_1e6(L):
    d, f = O.shell_table(L, 6.0)
    r = np.linspace(0.92, 1.35, 2000)
    e64 = vdw_energy(L, d, f, 6.0, r, FP64)
    e32 = vdw_energy(L, d, f, 6.0, r, FP32_COMPENSATED)
    ref = np.array([O.vdw_energy(x, 6.0, d, f) for x in r])
    assert np.array_equal(e64, ref)  # device FP64 == host FP64, bit for bit
    rel = np.abs(e32 - e64) / np.abs(e64)
    assert rel.max() <= 1e-6, rel.max()


@pytest.mark.parametrize("L", [12, 32, 96])
def test_vdw_pair_form_equals_shell_form(L):
    """The GPU pair form (every minimum-image pair of the lattice, no shell
    table; calculate_energy_vdW + distance_PBC, gpu_water3D.cu:1784-1805,
    1323-1337) equals the shell form: FP64 within 1e-12, FP32-compensated
    within 1e-6 relative, for r below 1 (where the table's d < R_cutoff
    filter binds) and above."""
    d, f = O.shell_table(L, 6.0)
    r = np.array([0.90, 0.95, 0.999, 1.0001, 1.05, 1.12, 1.2, 1.3])
    shell = np.array([O.vdw_energy(x, 6.0, d, f) for x in r])
    p64 = vdw_pair_energy(L, 6.0, r, FP64)
    p32 = vdw_pair_energy(L, 6.0, r, FP32_COMPENSATED)
    rel64 = np.abs(p64 - shell) / np.abs(shell)
    rel32 = np.abs(p32 - shell) / np.abs(shell)
    print(L, rel64.max(), rel32.max())
    assert rel64.max() <= 1e-12, rel64
    assert rel32.max() <= 1e-6, rel32


def test_vdw_pair_form_L8_keeps_the_odd_classes():
    """At L=8 the shell table halves odd class counts with integer division
    (gpu_water3D.cu:912-915): the pair form exceeds it by N/2 times the
    terms of the odd classes (126,976 pairs in the table, 127,488 pairs)."""
    L, N, rc = 8, 512, 6.0
    d, f = O.shell_table(L, rc)
    o = np.arange(-3, 5)
    X, Y, Z = np.meshgrid(o, o, o, indexing="ij")
    dist = np.sqrt((X ** 2 + Y ** 2 + Z ** 2).astype(np.float64)).astype(np.float32).ravel()
    dist = dist[(dist > 0) & (dist < rc)]
    cls, cnt = np.unique(dist, return_counts=True)
    assert (f // N).sum() * N == 126976 and cnt.sum() * N // 2 == 127488
    r = np.array([0.95, 1.0001, 1.1])
    shell = np.array([O.vdw_energy(x, rc, d, f) for x in r])
    pair = vdw_pair_energy(L, rc, r, FP64)

    def g(x):
        t = (1.0 / x) ** 3
        a = t * t
        t0 = (1.0 / rc) ** 3
        return a * a - a - (t0 ** 4 - t0 ** 2)

    for k, x in enumerate(r):
        keep = (cnt % 2 == 1) & (x * cls.astype(np.float64) < rc)
        want = 0.5 * N * g(x * cls[keep].astype(np.float64)).sum()
        assert abs((pair[k] - shell[k]) - want) <= 1e-9 * abs(shell[k]), (x, pair[k] - shell[k], want)


def test_precision_form_keeps_lattice_trajectory():
    """FP32-compensated vdW changes only the volume chain: sigma/eta
    trajectories are identical to the FP64 run from the same seeds."""
    Ls = [16, 24]
    a = Simulation(batch_points(Ls, FP64))
    b = Simulation(batch_points(Ls, FP32_COMPENSATED))
    a.step(300)
    b.step(300)
    for i in range(len(Ls)):
        x, y = a.get_config(i), b.get_config(i)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
        ra, rb = a.volume(i)[0], b.volume(i)[0]
        assert abs(ra - rb) < 0.05


def test_host_and_device_volume_modes_agree():
    a = Simulation(Params(L=24, volume_mode=VOLUME_HOST))
    b = Simulation(Params(L=24, volume_mode=VOLUME_DEVICE))
    a.step(500)
    b.step(500)
    assert a.volume() == b.volume()
    assert a.observe().E == b.observe().E


@pytest.mark.parametrize("tuning", [
    {"streams": 1, "block_sigma": 256, "block_eta": 256, "block_cell": 256, "graph_steps": 1,
     "fused": 0},
    {"streams": 3, "block_sigma": 64, "block_eta": 128, "block_cell": 64, "graph_steps": 5,
     "fused": 0},
    {"streams": 8, "block_sigma": 1024, "block_eta": 512, "block_cell": 1024, "graph_steps": 16,
     "fused": 0},
    {"streams": 2, "fused": 1, "slab_planes": 4, "block_chunk": 256},
    {"streams": 1, "fused": 1, "slab_planes": 2, "block_chunk": 1024, "graph_steps": 3},
    {"streams": 4, "fused": 1, "slab_planes": 1, "block_chunk": 64},
    # thick slabs, as the automatic plan gives large batches; T = L at L = 8,
    # 12 and 16 (a window repeats its own planes)
    {"streams": 2, "fused": 1, "slab_planes": 8, "block_chunk": 1024},
    {"streams": 1, "fused": 1, "slab_planes": 12, "block_chunk": 512},
    {"streams": 2, "fused": 1, "slab_planes": 16, "block_chunk": 1024, "chunk_sweeps": 12},
    # the resident kernel (SW steps between one-step launches here)
    {"fused": 2, "block_chunk": 1024},
    {"fused": 2, "block_chunk": 256, "graph_steps": 5},
])
def test_tuning_does_not_change_results(tuning):
    """Launch configuration (streams, workgroup sizes, steps per graph) never
    changes the trajectory: random numbers are keyed on sites."""
    pts = batch_points([8, 16, 12, 24, 8, 20], nc=1)
    ref = Simulation(pts)
    alt = Simulation(pts)
    alt.tuning = tuning
    assert {k: alt.tuning[k] for k in tuning} == tuning
    ref.step(37)
    alt.step(20)
    alt.tuning = {"streams": 2, "fused": 0 if alt.tuning["fused"] else 1}
    alt.step(17)
    for i in range(len(pts)):
        a, b = ref.get_config(i), alt.get_config(i)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(ref.labels(i), alt.labels(i))
        assert ref.volume(i) == alt.volume(i)


def test_autotune_keeps_state_and_trajectory():
    """cvf_autotune times candidate tunings from a snapshot, restores it and
    applies the winner: the run continues exactly as without tuning."""
    pts = batch_points([8, 16, 24, 12], nc=1)
    ref = Simulation(pts)
    alt = Simulation(pts)
    ref.step(15)
    alt.step(15)
    chosen = alt.autotune(4)
    assert alt.steps == 15 and alt.tuning == chosen
    for i in range(len(pts)):
        assert np.array_equal(ref.get_config(i)[0], alt.get_config(i)[0])
        assert ref.volume(i) == alt.volume(i)
    ref.step(20)
    alt.step(20)
    for i in range(len(pts)):
        a, b = ref.get_config(i), alt.get_config(i)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(ref.labels(i), alt.labels(i))
        assert ref.volume(i) == alt.volume(i)


def test_sigma_histogram_matches_config():
    sim = Simulation(batch_points([8, 16, 12]))
    sim.step(25)
    for i in range(3):
        s, _e = sim.get_config(i)
        assert np.array_equal(sim.sigma_histogram(i), np.bincount(s, minlength=6))


@pytest.mark.parametrize("n_cluster,whole", [(0, False), (1, False), (1, True)])
def test_step_histograms_equal_per_step_histograms(n_cluster, whole):
    # the device-recorded per-step histograms equal stepping one at a time and
    # reading the histogram after each step, for every point of a batch and
    # across the copy-back span (4096 steps); with whole-pass launches the
    # state alternates between the word arrays and the histogram and SW
    # kernels follow it
    pts = batch_points([8, 12], nc=n_cluster)
    a, b = Simulation(pts), Simulation(pts)
    if whole:
        a.tuning = {"fused": 1, "chunk_sweeps": 12, "slab_planes": 1, "graph_steps": 5}
    n = 4100 if n_cluster == 0 else 40
    rec = a.step_histograms(n, pt=1)
    ref = []
    for _ in range(n):
        b.step(1)
        ref.append(b.sigma_histogram(1))
    assert np.array_equal(rec, np.array(ref))
    assert np.all(rec.sum(axis=1) == 6 * 12 ** 3)
    assert a.steps == b.steps == n
    for i in range(2):
        ca, cb = a.get_config(i), b.get_config(i)
        assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[1], cb[1])
    assert a.step_histograms(0).shape == (0, 6)


@pytest.mark.parametrize("schedule,tuning,classes", [
    (0, {"fused": 1, "chunk_sweeps": 12}, {"chunk": 2, "begin": 1}),
    (0, {"fused": 0}, {"eta": 2, "sigma": 12, "begin": 1}),
    (1, {"fused": 1}, {"eta": 2, "sigma": 2, "begin": 1}),
])
def test_profile_steps_matches_plain_steps(schedule, tuning, classes):
    """cvf_profile_steps (every launch timed by its own dispatch events) runs
    the plain form of each step, whatever the tuning's variant forms, and
    leaves the same state as cvf_step; launch counts per class as expected."""
    from cvfbulkwater_amd import SCHED_FAST

    pts = [Params(L=L, T=0.7, P=0.3, seed=90 + L, schedule=SCHED_FAST if schedule else 0,
                  volume_mode=VOLUME_DEVICE) for L in (8, 16)]
    a, b = Simulation(pts), Simulation(pts)
    a.tuning = tuning
    b.tuning = tuning
    prof = a.profile_steps(5)
    b.step(5)
    for i in range(len(pts)):
        x, y = a.get_config(i), b.get_config(i)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
        assert a.volume(i) == b.volume(i)
    for cls, per_step in classes.items():
        ms, n = prof[cls]
        assert n == 5 * per_step and ms > 0, (cls, prof[cls])


def test_config5_precision_form_errors():
    """L=96, T=0.65, P=0.30 with SW over the config's 5000 steps sampled
    every 10, FP32-compensated vdW against FP64 everywhere from the same
    seed: the lattice trajectory is identical (N_HB error exactly 0); E, H,
    V and density agree within 1e-6 relative per sample, and so do their
    long-run means."""
    import bench

    res = bench.precision_form(steps=5000, interval=10)
    err = res["errors"]
    print(json.dumps(err))
    assert err["N_HB"]["max_rel"] == 0.0
    for k in ("E", "H", "V", "density"):
        assert err[k]["max_rel"] <= 1e-6, (k, err[k])
        assert err[k]["mean_rel"] <= 1e-6, (k, err[k])


def test_isobar_scan_feeds_the_analysis(tmp_path):
    """A small isobar scan in one batch: the averages written from memory
    equal those cvf_thermo computes from the data.out files of the same run
    (to the 6 printed digits), and the analysis files exist."""
    import subprocess

    from cvfbulkwater_amd import THERMO_EXE_PATH
    from cvfbulkwater_amd.scan import isobar_scan

    a = tmp_path / "mem"
    isobar_scan(a, 8, [0.6, 0.65, 0.7, 0.75, 0.8, 0.85], [0.45, 0.3], equil_steps=50, samples=40,
                interval=5, data_out=True)
    b = tmp_path / "files"
    b.mkdir()
    for d in a.glob("L8_T*_P*"):
        (b / d.name).mkdir()
        (b / d.name / "data.out").write_text((d / "data.out").read_text())
    r = subprocess.run([str(THERMO_EXE_PATH), str(b)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x = np.loadtxt(a / "L8_average_data.dat")
    y = np.loadtxt(b / "L8_average_data.dat")
    key = lambda m: m[np.lexsort((m[:, 2], m[:, 1]))]
    x, y = key(x), key(y)
    assert x.shape == (12, 16)
    # means to the printed digits; variances lose more to data.out's 8 digits
    assert np.array_equal(x[:, [0, 1, 2, 15]], y[:, [0, 1, 2, 15]])
    assert np.allclose(x[:, 3:15:2], y[:, 3:15:2], rtol=2e-5)
    assert np.allclose(x[:, 4:15:2], y[:, 4:15:2], rtol=5e-2, atol=1e-9)
    for P in ("0.45", "0.3"):
        for suffix in ("density.dat", "AlphaP.dat", "CP.dat", "density.dat.SI", "AlphaPderivate.dat.SI"):
            assert (a / f"L8_P{P}_{suffix}").stat().st_size > 0, suffix
