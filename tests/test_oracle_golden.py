"""CPU tests: pin the oracle (and the engine's shared core through the CPU
probe) against the reference's shipped data, known-answer vectors,
library routines and the mathematics.  No GPU needed."""
import ctypes as C
import io
import json
import lzma
import math
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import py as O
from tests.native import probe

GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_values.json").read_text())

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key, expected)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def _probe_philox(fn, ctr, key):
    o = (C.c_uint32 * 4)()
    fn((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
    return list(o)


def load_config(name):
    a = np.loadtxt(io.StringIO(lzma.decompress((GOLD / name).read_bytes()).decode()), dtype=np.int64)
    return a[:, :6].T.astype(np.uint8).reshape(-1), a[:, 6:].T.astype(np.uint8).reshape(-1)


def load_data_out():
    return np.loadtxt(io.StringIO(lzma.decompress((GOLD / "data.out.xz").read_bytes()).decode()))


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(ctr, key, want):
    assert list(O.philox(ctr, key)) == want
    assert _probe_philox(probe().probe_philox, ctr, key) == want
    assert _probe_philox(probe().probe_philox_site, ctr, key) == want
    assert _probe_philox(probe().probe_rocrand_philox, ctr, key) == want


def test_philox_matches_rocrand_random_inputs():
    rng = np.random.default_rng(1)
    for _ in range(300):
        ctr = [int(x) for x in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
        key = [int(x) for x in rng.integers(0, 2 ** 32, 2, dtype=np.uint64)]
        a = list(O.philox(ctr, key))
        assert a == _probe_philox(probe().probe_philox, ctr, key)
        assert a == _probe_philox(probe().probe_philox_site, ctr, key)
        assert a == _probe_philox(probe().probe_rocrand_philox, ctr, key)


def test_uniform_helpers():
    p = probe()
    for x in (0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0x2AAAAAAA, 0x2AAAAAAB):
        for n in (2, 6, 9):
            assert p.probe_uniform_int(x, n) == (x * n) >> 32
    assert p.probe_u53(0, 0) == 0.0
    assert p.probe_u53(0xFFFFFFFF, 0xFFFFFFFF) == 1 - 2 ** -53


# ------------------------------------------------------------ shell table
def test_shell_table_L32_equals_reference_distances_file():
    want = (GOLD / "distances_L32").read_text().splitlines()
    assert want[0] == "32 32 32"
    d, f = O.shell_table(32, 6.0)
    assert ["%f %d" % (a, b) for a, b in zip(d, f)] == want[1:]
    dd = np.zeros(256, np.float32)
    ff = np.zeros(256, np.int32)
    n = probe().probe_shell_table(32, 6.0, dd.ctypes.data_as(C.POINTER(C.c_float)),
                                  ff.ctypes.data_as(C.POINTER(C.c_int32)), 256)
    assert ["%f %d" % (a, b) for a, b in zip(dd[:n], ff[:n])] == want[1:]
    assert int(f.astype(np.int64).sum()) == 14647296  # 447 neighbours x N/2 x 2


def brute_shells(L, rc):
    """Brute-force minimum-image classes: float32(sqrt(int)) < rc, first
    occurrence in cell-index order, multiplicity count."""
    N = L ** 3
    j = np.arange(1, N)
    dx, dy, dz = j % L, (j // L) % L, j // (L * L)
    m = lambda v: np.where(v > L // 2, L - v, v)
    d2 = m(dx) ** 2 + m(dy) ** 2 + m(dz) ** 2
    d = np.sqrt(d2.astype(np.float64)).astype(np.float32)
    keep = d < np.float32(rc)
    vals, first, counts = np.unique(d[keep], return_index=True, return_counts=True)
    order = np.argsort(first)
    return vals[order], counts[order]


@pytest.mark.parametrize("L", [8, 12, 16, 20])
def test_shell_table_bruteforce(L):
    vals, counts = brute_shells(L, 6.0)
    d, f = O.shell_table(L, 6.0)
    assert np.array_equal(d, vals)
    assert np.array_equal(f, (counts // 2) * L ** 3)


def test_shell_table_L8_integer_division_quirk():
    """L=8: the (4,0,0) and (4,4,0) classes have odd counts, so integer /2
    drops half a pair each: 126976 pairs instead of 127488."""
    vals, counts = brute_shells(8, 6.0)
    d, f = O.shell_table(8, 6.0)
    assert len(d) == 28
    assert int(f.astype(np.int64).sum()) == 126976
    assert int((counts.sum() / 2) * 512) == 127488
    odd = vals[counts % 2 == 1]
    assert sorted(np.round(odd.astype(float) ** 2).astype(int).tolist()) == [16, 32]


# -------------------------------------------------------------------- vdW
def lj(x, rc):
    s = (1.0 / rc) ** 6
    return x ** -12 - x ** -6 - (s * s - s)


@pytest.mark.parametrize("L", [12, 16])
def test_vdw_shell_sum_equals_pair_sum(L):
    """E_vdW from the shell table equals the direct minimum-image pair sum
    (1/2 sum_i sum_j) over the lattice for every r."""
    N = L ** 3
    j = np.arange(1, N)
    m = lambda v: np.where(v > L // 2, L - v, v)
    d2 = m(j % L) ** 2 + m((j // L) % L) ** 2 + m(j // (L * L)) ** 2
    d = np.sqrt(d2.astype(np.float64)).astype(np.float32).astype(np.float64)
    d = d[d < 6.0]
    dist, freq = O.shell_table(L, 6.0)
    for r in (0.95, 1.0001, 1.0762, 1.2, 1.5):
        x = r * d
        pair = 0.5 * N * lj(x[x < 6.0], 6.0).sum()
        assert abs(O.vdw_energy(r, 6.0, dist, freq) - pair) <= 1e-11 * abs(pair)
        e2 = probe().probe_vdw_fp64(r, 6.0, np.ascontiguousarray(dist).ctypes.data_as(C.POINTER(C.c_float)),
                                    np.ascontiguousarray(freq).ctypes.data_as(C.POINTER(C.c_int32)), len(dist))
        assert e2 == O.vdw_energy(r, 6.0, dist, freq)


def test_vdw_single_shell_closed_form():
    d = np.array([1.0], np.float32)
    f = np.array([3 * 512], np.int32)
    for r in (0.9, 1.0, 1.1):
        want = 3 * 512 * lj(r, 6.0)
        assert abs(O.vdw_energy(r, 6.0, d, f) - want) <= 1e-13 * abs(want)
    assert O.vdw_energy(6.5, 6.0, d, f) == 0.0  # beyond the cutoff


# ------------------------------------------------ data.out / log identities
def test_data_out_row_identities():
    """Every row of the shipped data.out satisfies e = E_vdW(r) - J N_HB -
    Jsig 15N m, v = N r^3 + v_HB N_HB, h = e + P v (measure(),
    gpu_water3D.cu:2138-2159) with the oracle's shell table and vdW."""
    D = load_data_out()
    assert D.shape == (10000, 7)
    dist, freq = O.shell_table(32, 6.0)
    N = 32 ** 3
    vhb, jsig, P = struct.unpack("3f", struct.pack("3f", 0.6, 0.08, 0.45))
    worst = 0.0
    for t, e, h, v, nhb, m, r in D[::7]:
        ev = O.vdw_energy(r, 6.0, dist, freq)
        e2 = ev - 0.5 * nhb - jsig * (m * 15 * N)
        v2 = N * r ** 3 + vhb * nhb
        h2 = e2 + P * v2
        worst = max(worst, abs(e2 - e) / abs(e), abs(v2 - v) / v, abs(h2 - h) / abs(h))
    assert worst < 4e-7, worst


def test_water3d_log_summary_formula():
    """Magnetization/Energy lines = mean and population std of data.out
    columns 6 and 2 divided by n_loops/interval (gpu_water3D.cu:1277-1293)."""
    D = load_data_out()
    n = 100000 / 10
    M, M2 = D[:, 5].sum() / n, (D[:, 5] ** 2).sum() / n
    E, E2 = D[:, 1].sum() / n, (D[:, 1] ** 2).sum() / n
    mline = REF["log_magnetization_line"].split("\t")
    eline = [x for x in REF["log_energy_line"].split("\t") if x]
    assert abs(M - float(mline[1])) < 2e-6 and abs(math.sqrt(M2 - M * M) - float(mline[3])) < 2e-6
    assert abs(E - float(eline[1])) < 1e-3 and abs(math.sqrt(E2 - E * E) - float(eline[3])) < 1e-3


def test_cluster_out_degenerate_row():
    """Cluster_Steps=0: labels all zero -> the shipped cluster.out row,
    including the int32-overflowing variance."""
    lab = np.zeros(6 * 32 ** 3, np.int32)
    lg, sc, npc, ncl, mean, var = O.cluster_stats_labels(32, lab)
    line = "%d %d %d %d %f %f %d" % (0, lg, sc, npc, mean, var, ncl)
    assert line == REF["cluster_out_first_row"]
    assert REF["cluster_out_all_rows_same_but_time"]


def test_cluster_stats_bruteforce():
    rng = np.random.default_rng(3)
    L = 8
    NN = 6 * L ** 3
    lab = np.minimum(np.arange(NN), rng.integers(0, NN, NN)).astype(np.int32)
    lab = lab[lab]  # roots are fixed points of a 2-step map
    lab[:200] = 0  # one cluster of >= 2L nodes
    st = O.cluster_stats_labels(L, lab)
    sizes = np.bincount(lab, minlength=NN)
    nz = sizes[sizes > 0]
    assert st[3] == nz.size
    assert st[0] == nz.max()
    assert st[1] == np.sort(nz)[-2]
    # percolation: clusters >= 2L covering every x, y or z coordinate
    cell = np.arange(NN) % L ** 3
    xyz = [cell % L, (cell // L) % L, cell // (L * L)]
    perc = 0
    for c in np.flatnonzero(sizes >= 2 * L):
        idx = lab == c
        perc += any(np.unique(k[idx]).size == L for k in xyz)
    assert st[2] == perc


# ---------------------------------------------------------- configurations
@pytest.mark.parametrize("name,key", [("input_config.xz", "input_config_counts"),
                                      ("config.out.xz", "config_out_counts")])
def test_shipped_config_invariants(name, key):
    """The shipped configs are valid chessboard states (every cube one of the
    9 solutions) and the oracle's N_HB / N_sigma match a vectorised count."""
    s, e = load_config(name)
    sim = O.OracleSim(L=32)
    assert sim.set_config(s, e) == 0
    s2, e2 = sim.get_config()
    assert np.array_equal(s, s2) and np.array_equal(e, e2)
    o = sim.measure()
    want = REF[key]
    assert o.nhb2 == want["nhb2"] and o.n_sigma == want["n_sigma"]


def eta_invariants(L, s, e):
    N = L ** 3
    E6 = e.reshape(6, L, L, L)  # [arm, z, y, x]
    assert np.all(E6.sum(axis=0) == 4)
    for ax, axis in ((0, 2), (1, 1), (2, 0)):
        assert np.array_equal(E6[2 * ax], np.roll(E6[2 * ax + 1], -1, axis=axis))
    assert s.max() <= 5


def test_oracle_keeps_chessboard_invariants():
    sim = O.OracleSim(L=8, T=0.6, n_cluster=1)
    for _ in range(30):
        sim.step(1)
        s, e = sim.get_config()
        eta_invariants(8, s, e)
        # eta re-derived through the decision tree reproduces itself
        chk = O.OracleSim(L=8)
        assert chk.set_config(s, e) == 0


# --------------------------------------------------- thresholds / exp / log
def _thr_probe(P, T):
    s = (C.c_uint64 * 33)()
    e = (C.c_uint64 * 9)()
    w = (C.c_uint64 * 2)()
    jb = (C.c_double * 2)()
    probe().probe_thresholds(P, T, 0.5, 0.08, 0.6, s, e, w, jb)
    return np.array(list(s), np.uint64), np.array(list(e), np.uint64), np.array(list(w), np.uint64)


@pytest.mark.parametrize("T,P", [(0.815, 0.45), (0.55, -0.10), (1.0, 0.8), (0.7, 0.45)])
def test_thresholds_oracle_equals_engine(T, P):
    sim = O.OracleSim(L=4, T=T, P=P)
    a = sim.thresholds()
    b = _thr_probe(P, T)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    sig = a[0].reshape(3, 11)
    assert sig[1, 5] == 2 ** 32  # dE = 0
    assert np.all(np.diff(sig.astype(np.float64), axis=1) >= 0)  # more equal arms -> more likely


def test_det_exp_log_accuracy():
    p = probe()
    rng = np.random.default_rng(7)

    def ulp_err(a, b):
        return abs(a - b) / math.ulp(b)

    for x in np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-2, 2, 2000), [0.0, 1e-300]]):
        assert ulp_err(p.probe_det_exp(float(x)), math.exp(float(x))) <= 2.0, x
    for x in np.concatenate([10 ** rng.uniform(-300, 300, 2000), rng.uniform(0.5, 2, 2000), [1.0]]):
        got, want = p.probe_det_log(float(x)), math.log(float(x))
        assert abs(got - want) <= 2 * math.ulp(want) + 1e-300, x
    assert p.probe_det_exp(-800.0) == 0.0
    assert math.isinf(p.probe_det_exp(710.0))


def test_fastdiv_exact():
    p = probe()
    rng = np.random.default_rng(11)
    for d in list(range(2, 130)) + [252]:
        L = 2 * d if 2 * d <= 252 else d
        tmax = L ** 3
        ts = np.concatenate([rng.integers(0, tmax, 400), [0, tmax - 1, d - 1, d, 2 * d - 1]])
        for t in ts:
            assert p.probe_fdiv(int(t), d) == int(t) // d


def test_packed_word_helpers():
    p = probe()
    rng = np.random.default_rng(2)
    for _ in range(2000):
        sig = rng.integers(0, 6, 6)
        w = int(sum(int(v) << (3 * a) for a, v in enumerate(sig))) | (int(rng.integers(0, 64)) << 18)
        v = int(rng.integers(0, 6))
        assert p.probe_count_eq(w, v) == int((sig == v).sum())
    for j in range(12):
        A, ax = p.probe_edge_start(j), p.probe_edge_axis(j)
        assert not (A >> ax) & 1
        assert p.probe_corner_edge(A, ax) == j
        assert p.probe_corner_edge(A | (1 << ax), ax) == j
    for s in range(9):
        assert bin(p.probe_sol_mask(s)).count("1") == 8


# ------------------------------------------- transition probabilities (sigma)
def hamiltonian_counts(L, s, e):
    """(2*N_HB, N_sigma) of an arm-major config, vectorised."""
    S = s.reshape(6, L, L, L).astype(np.int64)
    E = e.reshape(6, L, L, L).astype(np.int64)
    nhb2 = 0
    for ax, axis in ((0, 2), (1, 1), (2, 0)):
        eq = S[2 * ax] == np.roll(S[2 * ax + 1], -1, axis=axis)
        nhb2 += int((eq * (E[2 * ax] + np.roll(E[2 * ax + 1], -1, axis=axis))).sum())
    nsig = sum(int((S[j] == S[k]).sum()) for j in range(6) for k in range(j + 1, 6))
    return nhb2, nsig


def test_sigma_acceptance_is_boltzmann():
    """Every proposal of the oracle's sigma half-sweep is accepted with
    probability min(1, exp(-beta dH)), dH = -Jeff dN_HB - Jsig dN_sigma
    computed from whole-lattice counts before/after the single move."""
    L, N, T, P, seed = 4, 64, 0.6, 0.45, 99
    sim = O.OracleSim(L=L, T=T, P=P, seed=seed)
    jeff = float(np.float32(0.5) - np.float32(P) * np.float32(0.6))
    jsig = float(np.float32(0.08))
    stats = {}
    rng = np.random.default_rng(0)
    for step in range(1, 1500):
        arm, ind = int(rng.integers(0, 6)), int(rng.integers(0, 2))
        s0, e0 = sim.get_config()
        sim.sigma_halfsweep(step, 0, arm, ind)
        s1, _ = sim.get_config()
        base = hamiltonian_counts(L, s0, e0)
        for c in range(N):
            x, y, z = c % L, (c // L) % L, c // (L * L)
            coord = (x, y, z)[arm // 2]
            if (ind and coord % 2) or (not ind and not coord % 2):
                assert s1[arm * N + c] == s0[arm * N + c]
                continue
            # (x + b L/2, y + a L/2, z) share one Philox block, word a + 2b;
            # proposal floor(6w/2^32), acceptance (6w) mod 2^32
            a_, b_ = int(y >= L // 2), int(x >= L // 2)
            r = O.draw(seed, O.ST_SIGMA, c - a_ * (L // 2) * L - b_ * (L // 2), step, 0)
            prop = (int(r[a_ + 2 * b_]) * 6) >> 32
            if prop == s0[arm * N + c]:
                continue
            t = s0.copy()
            t[arm * N + c] = prop
            n1 = hamiltonian_counts(L, t, e0)
            dH = -jeff * (n1[0] - base[0]) / 2 - jsig * (n1[1] - base[1])
            key = round(dH, 9)
            acc = s1[arm * N + c] == prop
            assert acc or s1[arm * N + c] == s0[arm * N + c]
            a, n = stats.get(key, (0, 0))
            stats[key] = (a + acc, n + 1)
    assert len(stats) >= 8
    for dH, (a, n) in stats.items():
        p = min(1.0, math.exp(-dH / T))
        if n < 30:
            continue
        sd = math.sqrt(max(p * (1 - p), 1e-12) / n)
        assert abs(a / n - p) <= 5 * sd + 1e-12, (dH, a, n, p)


# cube geometry restated from gpu_set_cubes (gpu_water3D.cu:3180-3257) and
# the solution table of setSolutions_Typewriter (chessboard.h:211-262)
CUBE_VERT = {1: (0, 0, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (0, 0, 1), 5: (1, 1, 1), 6: (0, 1, 1),
             7: (1, 0, 1), 8: (1, 1, 0)}
CUBE_EDGES = [(1, 2, 0), (1, 3, 2), (1, 4, 4), (2, 7, 4), (2, 8, 2), (3, 6, 4), (3, 8, 0), (4, 6, 2),
              (4, 7, 0), (6, 5, 0), (7, 5, 2), (8, 5, 4)]
SOL_OFF = [(2, 3, 6, 9), (2, 3, 5, 11), (0, 6, 8, 9), (0, 5, 8, 11), (1, 3, 7, 11), (1, 4, 7, 10),
           (2, 4, 5, 10), (1, 4, 8, 9), (0, 6, 7, 10)]


def cube_list(L):
    out = []
    for z in range(L):
        for y in range(L):
            for x in range(L):
                if (z % 2 == 0 and x % 2 == 0 and y % 2 == 0) or (z % 2 == 1 and x % 2 == 1 and y % 2 == 1):
                    cube = x // 2 + (L // 2) * (y // 2) + (L * L // 4) * z
                    cell = lambda v: ((x + CUBE_VERT[v][0]) % L) + L * ((y + CUBE_VERT[v][1]) % L) \
                        + L * L * ((z + CUBE_VERT[v][2]) % L)
                    out.append((cube, [(cell(a), cell(b), arm) for a, b, arm in CUBE_EDGES]))
    return sorted(out)


def test_eta_acceptance_is_boltzmann():
    """Cube moves (gpu_water3D.cu:3374-3466): a proposal to solution s' is
    accepted with min(1, exp(beta Jeff dN_HB)), dN_HB counted over the whole
    lattice with only that cube's 12 edges switched to s'."""
    L, N, T, P, seed = 8, 512, 0.55, 0.45, 5
    sim = O.OracleSim(L=L, T=T, P=P, seed=seed)
    jeff = float(np.float32(0.5) - np.float32(P) * np.float32(0.6))
    cubes = cube_list(L)
    stats = {}
    for step in range(120):
        s0, e0 = sim.get_config()
        st0 = sim.cube_states()
        sim.eta_sweep(step)
        s1, e1 = sim.get_config()
        st1 = sim.cube_states()
        assert np.array_equal(s0, s1)
        eta_invariants(L, s1, e1)
        base = hamiltonian_counts(L, s0, e0)[0]
        for cube, edges in cubes:
            hL = L // 2  # cubes (ox + b L/2, oy + a L/2, oz) share one Philox block
            a_, b_ = int((cube // hL) % hL >= L // 4), int(cube % hL >= L // 4)
            site = cube - a_ * hL * (L // 4) - b_ * (L // 4)
            prop = (int(O.draw(seed, O.ST_ETA, site, step, 0)[a_ + 2 * b_]) * 9) >> 32
            acc = st1[cube] == prop
            assert acc or st1[cube] == st0[cube]
            e = e0.copy()
            for j, (va, vb, arm) in enumerate(edges):
                on = 0 if j in SOL_OFF[prop] else 1
                e[arm * N + va] = on
                e[(arm + 1) * N + vb] = on
                if acc:
                    assert e1[arm * N + va] == on and e1[(arm + 1) * N + vb] == on
            dn = (hamiltonian_counts(L, s0, e)[0] - base) // 2
            a, n = stats.get(dn, (0, 0))
            stats[dn] = (a + acc, n + 1)
    assert len(stats) >= 5
    for dn, (a, n) in stats.items():
        p = min(1.0, math.exp(jeff * dn / T))
        if n < 30:
            continue
        sd = math.sqrt(max(p * (1 - p), 1e-12) / n)
        assert abs(a / n - p) <= 5 * sd + 1e-12, (dn, a, n, p)


# ------------------------------------------------------------ volume move
def test_volume_chain_oracle_equals_engine_host_code():
    for L, T, P in ((8, 0.815, 0.45), (16, 0.6, 0.3)):
        sim = O.OracleSim(L=L, T=T, P=P, seed=17)
        for t in range(5000):
            sim.volume_move(t)
        out = (C.c_double * 4)()
        probe().probe_volume_chain(17, L, T, P, 6.0, 0, 5000, out)
        assert (out[0], out[1], out[2]) == sim.volume()
        assert 100 < out[3] < 4900


def test_volume_chain_stationary_mean():
    """The r chain samples p(r) ~ V(r)^N exp(-beta (E_vdW(r) + P V(r))) (the
    NPT weight of gpu_water3D.cu:1772-1774 with proposals symmetric in r);
    its long-run mean matches direct numerical integration."""
    L, T, P, seed = 4, 0.8, 0.45, 23
    N = L ** 3
    sim = O.OracleSim(L=L, T=T, P=P, seed=seed)
    dist, freq = O.shell_table(L, 6.0)
    rs = []
    for t in range(200000):
        sim.volume_move(t)
        if t >= 2000 and t % 5 == 0:
            rs.append(sim.volume()[0])
    rs = np.array(rs)
    grid = np.linspace(0.85, 1.4, 4001)
    Tf, Pf = float(np.float32(T)), float(np.float32(P))
    logw = np.array([N * math.log(N * r ** 3) - (O.vdw_energy(r, 6.0, dist, freq) + Pf * N * r ** 3) / Tf
                     for r in grid])
    w = np.exp(logw - logw.max())
    mean = (grid * w).sum() / w.sum()
    sd = math.sqrt(((grid - mean) ** 2 * w).sum() / w.sum())
    blocks = rs[: len(rs) // 40 * 40].reshape(40, -1).mean(axis=1)
    se = blocks.std(ddof=1) / math.sqrt(40)
    assert abs(rs.mean() - mean) < 4 * se + 1e-4, (rs.mean(), mean, se)
    assert abs(rs.std() - sd) < 0.1 * sd


# ------------------------------------------------------- Swendsen-Wang CCL
def test_sw_labels_are_minimum_node_components():
    """Re-derive the 18 links per cell of gpu_create_cluster
    (gpu_water3D.cu:2383-2438) from the Philox stream, label components with
    scipy's connected_components, and compare with the oracle: labels are
    the minimum node id of each component and sigma rotates by floor(6u(label))."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    L, N, T, P, seed = 8, 512, 0.6, 0.45, 31
    sim = O.OracleSim(L=L, T=T, P=P, seed=seed, n_cluster=1)
    sim.step(5)
    thr_hb, thr_s = [int(x) for x in sim.thresholds()[2]]
    for step in (5, 6, 7):
        s0, e0 = sim.get_config()
        sim.sw_step(step, 0)
        lab = sim.labels()
        s1, _ = sim.get_config()
        rows, cols = [], []
        pairs = [(i, j) for i in range(5) for j in range(i + 1, 6)]
        for c in range(N):
            x, y, z = c % L, (c // L) % L, c // (L * L)
            nbs = [((x + 1) % L) + L * y + L * L * z, x + L * ((y + 1) % L) + L * L * z,
                   x + L * y + L * L * ((z + 1) % L)]
            w = np.concatenate([O.draw(seed, O.ST_SWBOND, c, step, g) for g in range(5)])
            for ax in range(3):
                a0, a1 = 2 * ax, 2 * ax + 1
                if e0[a0 * N + c] and s0[a0 * N + c] == s0[a1 * N + nbs[ax]] and int(w[ax]) < thr_hb:
                    rows.append(a0 * N + c)
                    cols.append(a1 * N + nbs[ax])
            for k, (i, j) in enumerate(pairs):
                if s0[i * N + c] == s0[j * N + c] and int(w[3 + k]) < thr_s:
                    rows.append(i * N + c)
                    cols.append(j * N + c)
        g = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(6 * N, 6 * N))
        ncomp, comp = connected_components(g, directed=False)
        minnode = np.full(ncomp, 6 * N)
        np.minimum.at(minnode, comp, np.arange(6 * N))
        assert np.array_equal(lab, minnode[comp])
        shift = np.array([(int(O.draw(seed, O.ST_SWFLIP, int(l), step, 0)[0]) * 6) >> 32 for l in lab])
        assert np.array_equal(s1, (s0.astype(int) + shift) % 6)


# ------------------------------------------------------------ C ABI symbols
def test_library_exports_every_declared_symbol():
    import re

    from cvfbulkwater_amd import LIB_PATH

    if not LIB_PATH.exists():
        pytest.skip("libcvf.so not built")
    names = set()
    for hdr in sorted((Path(__file__).resolve().parent.parent / "include").glob("*.h")):
        names |= set(re.findall(r"\b(cvf_[a-z0-9_]+)\s*\(", hdr.read_text()))
    lib = C.CDLL(str(LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(names) >= 21 and "cvf_thermo_kt" in names
