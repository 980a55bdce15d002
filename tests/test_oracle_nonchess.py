"""The oracle's non-chessboard mode (Flag_chessboard = 0): every bond active,
no cube moves, HB energy gated by per-cell counters nhb <= NHBMAX = 4
(reference gpu_water3D.cu:721-724, 845-862, 1981-1984, 2004-2007,
2064-2099)."""
import math

import numpy as np
import pytest

from oracle import py as O

NHBMAX = 4


def recount_nhb(L, s, e):
    """nhb[c] = sum over arms of eta * delta(sigma_a(c), sigma_{a^1}(nbr_a(c))),
    from the arm-major config, vectorised over the lattice."""
    S = s.reshape(6, L, L, L).astype(np.int64)  # [arm, z, y, x]
    E = e.reshape(6, L, L, L).astype(np.int64)
    n = np.zeros((L, L, L), np.int64)
    for a in range(6):
        axis = (2, 2, 1, 1, 0, 0)[a]
        shift = -1 if a % 2 == 0 else 1
        nb = np.roll(S[a ^ 1], shift, axis=axis)
        n += E[a] * (S[a] == nb)
    return n.reshape(-1)


def greedy_nhb2(L, s):
    """Sum of the capped counters of measure(): cells in index order, forward
    x, y, z bonds, an HB counted for both cells while both are below 4."""
    N = L ** 3
    cnt = [0] * N
    for c in range(N):
        x, y, z = c % L, (c // L) % L, c // (L * L)
        nbs = [(x + 1) % L + L * y + L * L * z, x + L * ((y + 1) % L) + L * L * z,
               x + L * y + L * L * ((z + 1) % L)]
        for ax in range(3):
            nb = nbs[ax]
            if s[2 * ax * N + c] == s[(2 * ax + 1) * N + nb]:
                if cnt[c] < NHBMAX and cnt[nb] < NHBMAX:
                    cnt[c] += 1
                    cnt[nb] += 1
    return sum(cnt), cnt


@pytest.mark.parametrize("flag_config", [0, 2])
def test_start_has_all_bonds_active_and_counted(flag_config):
    sim = O.OracleSim(L=8, seed=5, flag_config=flag_config, flag_chessboard=0)
    s, e = sim.get_config()
    assert np.all(e == 1)
    assert np.array_equal(sim.nhb(), recount_nhb(8, s, e))
    if flag_config == 2:
        assert np.all(sim.nhb() == 6)


def test_set_config_keeps_eta_and_counts_with_it():
    """input_config eta is kept as read in this mode and weights the initial
    counters (the cube decoding applies only with the chessboard)."""
    L, N = 8, 512
    rng = np.random.default_rng(3)
    s = rng.integers(0, 6, 6 * N).astype(np.uint8)
    e = rng.integers(0, 2, 6 * N).astype(np.uint8)
    sim = O.OracleSim(L=L, flag_chessboard=0, flag_config=1)
    assert sim.set_config(s, e) == 0
    s1, e1 = sim.get_config()
    assert np.array_equal(s1, s) and np.array_equal(e1, e)
    assert np.array_equal(sim.nhb(), recount_nhb(L, s, e))


@pytest.mark.parametrize("flag_config", [0, 2])
def test_nhb_bookkeeping_follows_the_lattice(flag_config):
    """With every eta = 1 and no cluster moves the counters always equal the
    recount from sigma; eta never changes."""
    L = 8
    sim = O.OracleSim(L=L, T=0.6, seed=11, flag_config=flag_config, flag_chessboard=0)
    for _ in range(30):
        sim.step(1)
        s, e = sim.get_config()
        assert np.all(e == 1)
        assert np.array_equal(sim.nhb(), recount_nhb(L, s, e))


def test_nonchess_acceptance_follows_gated_rule():
    """Each proposal is accepted with probability min(1, exp(-beta dE)),
    dE = -Jeff h - Jsig d_sigma, where h is the gated HB change computed from
    recounted counters: h = [n_i+dN<=4][n_j+dN<=4] delta(new,nb)
    - [n_i<=4][n_j<=4] delta(old,nb).  Starting from the ordered lattice
    (every counter 6) makes the gate decide most classes."""
    L, N, T, P, seed = 4, 64, 0.5, 0.45, 21
    sim = O.OracleSim(L=L, T=T, P=P, seed=seed, flag_config=2, flag_chessboard=0)
    jeff = float(np.float32(0.5) - np.float32(P) * np.float32(0.6))
    jsig = float(np.float32(0.08))
    stats = {}
    gated = edge = 0
    rng = np.random.default_rng(1)
    for step in range(1, 1200):
        arm, ind = int(rng.integers(0, 6)), int(rng.integers(0, 2))
        s0, e0 = sim.get_config()
        n0 = recount_nhb(L, s0, e0)
        sim.sigma_halfsweep(step, 0, arm, ind)
        s1, _ = sim.get_config()
        for c in range(N):
            x, y, z = c % L, (c // L) % L, c // (L * L)
            coord = (x, y, z)[arm // 2]
            if (ind and coord % 2) or (not ind and not coord % 2):
                continue
            a_, b_ = int(y >= L // 2), int(x >= L // 2)
            r = O.draw(seed, O.ST_SIGMA, c - a_ * (L // 2) * L - b_ * (L // 2), step, 0)
            prop = (int(r[a_ + 2 * b_]) * 6) >> 32
            old = int(s0[arm * N + c])
            if prop == old:
                continue
            d3 = [(x + 1) % L, (x - 1) % L, x, x, x, x]
            nbx = d3[arm]
            nby = [y, y, (y + 1) % L, (y - 1) % L, y, y][arm]
            nbz = [z, z, z, z, (z + 1) % L, (z - 1) % L][arm]
            nb = nbx + L * nby + L * L * nbz
            snb = int(s0[(arm ^ 1) * N + nb])
            dn = int(prop == snb) - int(old == snb)
            ni, nj = int(n0[c]), int(n0[nb])
            h = int(ni + dn <= 4 and nj + dn <= 4) * int(prop == snb) - \
                int(ni <= 4 and nj <= 4) * int(old == snb)
            if dn != h:
                gated += 1
            if prop == snb and ni <= 4 and nj <= 4 and (ni + dn > 4 or nj + dn > 4):
                edge += 1  # the gate on the new counters alone closes the HB
            ds = sum(int(prop == s0[k * N + c]) - int(old == s0[k * N + c]) for k in range(6) if k != arm)
            dE = -jeff * h - jsig * ds
            acc = s1[arm * N + c] == prop
            assert acc or s1[arm * N + c] == old
            # the decision itself: low word of 6w below floor(p 2^32)
            p_acc = min(1.0, math.exp(-dE / T))
            thr = 1 << 32 if p_acc >= 1.0 else int(p_acc * 2.0 ** 32)
            assert acc == (((int(r[a_ + 2 * b_]) * 6) & 0xFFFFFFFF) < thr), (step, c, h, ds)
            key = round(dE, 9)
            a, n = stats.get(key, (0, 0))
            stats[key] = (a + acc, n + 1)
    assert gated > 200 and edge > 5  # the gate changed the HB term of many proposals
    assert len(stats) >= 6
    for dE, (a, n) in stats.items():
        p = min(1.0, math.exp(-dE / T))
        if n < 30:
            continue
        sd = math.sqrt(max(p * (1 - p), 1e-12) / n)
        assert abs(a / n - p) <= 5 * sd + 1e-12, (dE, a, n, p)


@pytest.mark.parametrize("flag_config,steps", [(2, 0), (0, 0), (0, 40), (2, 40)])
def test_measure_counts_capped_hbs_greedily(flag_config, steps):
    L = 4
    sim = O.OracleSim(L=L, T=0.5, seed=8, flag_config=flag_config, flag_chessboard=0)
    sim.step(steps)
    s, _ = sim.get_config()
    o = sim.measure()
    want, cnt = greedy_nhb2(L, s)
    assert o.nhb2 == want and max(cnt) <= NHBMAX
    assert o.n_hb == want / 2
    if flag_config == 2 and steps == 0:
        # every bond is an HB; the cap leaves at most 4 of each cell's 6
        assert want <= 4 * L ** 3 and want > 2 * L ** 3
