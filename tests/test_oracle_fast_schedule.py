"""The oracle's fast schedule (two-colour cell passes, each cell's six arms
in order), CPU only.

It is this engine's own schedule, not the reference's, so nothing of the
reference pins it.  Its single-move rule is the reference schedule's
(shared code, pinned by the reference-schedule tests); what is specific is
pinned here:
  * the randoms: at T -> infinity every proposal is accepted, so after one
    colour pass every arm of every cell of that colour equals its proposal,
    computed here from Philox and the quad / word / sub contract, and the
    other colour is untouched;
  * the sampling: both schedules sample the same Boltzmann distribution, so
    long-run means of N_HB, N_sigma and r_cell agree within their errors.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import py as O


def test_colour_pass_randoms_at_infinite_temperature():
    L, seed, step = 8, 97, 5
    sim = O.OracleSim(L=L, T=1e30, P=0.0, seed=seed, schedule=1)
    before, _ = sim.get_config()
    N = L ** 3
    for pass_ in (0, 3):
        colour = pass_ & 1
        prev, _ = sim.get_config()
        sim.sigma_colour(step, pass_)
        after, _ = sim.get_config()
        for c in range(N):
            x, y, z = c % L, (c // L) % L, c // (L * L)
            for arm in range(6):
                if (x + y + z) % 2 != colour:
                    assert after[arm * N + c] == prev[arm * N + c]
                    continue
                ay, bx = int(y >= L // 2), int(x >= L // 2)
                site = c - ay * (L // 2) * L - bx * (L // 2)
                w = int(O.draw(seed, O.ST_SIGMA_F, site, step, 8 * pass_ + arm)[ay + 2 * bx])
                assert after[arm * N + c] == (w * 6) >> 32, (c, arm)
    assert not np.array_equal(before, sim.get_config()[0])


def block_stats(x, nb=10):
    x = np.asarray(x, float)
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return x.mean(), b.std(ddof=1) / math.sqrt(nb)


@pytest.mark.parametrize("T,P", [(0.815, 0.45), (0.60, 0.10)])
def test_fast_and_reference_schedules_sample_the_same_distribution(T, P):
    L, equil, n = 8, 400, 4000
    out = {}
    for sched in (0, 1):
        sim = O.OracleSim(L=L, T=T, P=P, seed=2024 + sched, schedule=sched)
        for _ in range(equil):
            sim.step(1)
        rows = []
        for _ in range(n):
            sim.step(1)
            m = sim.measure()
            rows.append((m.nhb2 / 2 / L ** 3, m.n_sigma / L ** 3, m.r_cell))
        out[sched] = np.array(rows)
    for k, name in enumerate(("N_HB/N", "N_sigma/N", "r_cell")):
        (m0, e0), (m1, e1) = block_stats(out[0][:, k]), block_stats(out[1][:, k])
        z = (m1 - m0) / math.hypot(e0, e1)
        assert abs(z) < 4.0, (name, m0, e0, m1, e1)


def test_fast_schedule_needs_the_chessboard():
    with pytest.raises(ValueError):
        O.OracleSim(L=8, schedule=1, flag_chessboard=0)
