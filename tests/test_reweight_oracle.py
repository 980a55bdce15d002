"""Pins of the histogram reweighting oracle (oracle/cvf_rw_oracle.c) and the
host-only parts of the engine's reweighting library, CPU only.

The reference program (HistogramReweighting/histogram_reweighting_mixed_
paramSI_norm_x.f90) needs gfortran, which is absent, and ships no outputs;
the oracle is pinned by what the mathematics fixes:
  * multiple-histogram reweighting of exact Boltzmann histograms has the
    log partition-function ratios as its fixed point (closed form);
  * single-histogram reweighting to the simulation's own point is the
    identity, and to another point multiplies by exp(-(beta' H' - beta H));
  * the Binder cumulant of a two-point distribution, the unit norm and
    variance of the 3D Ising distribution, its maximum (scipy);
  * the order-parameter histogram's norm, mean and variance and the mean of
    the rotated coordinate e sin(a) + rho cos(a) under the density.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import minimize_scalar

from oracle import rw
import rw_synth

SL = lambda x: float(np.float32(x))  # noqa: E731  (a default-real literal held in real*8)


def exact_grid(T, P, e_range=0.02, rho_range=40.0, N=8, nb=24, seed=1):
    """Histograms proportional to g(e, rho) exp(-beta (E + P V)) on a grid,
    with their log partition functions."""
    rng = np.random.default_rng(seed)
    Emin, Dmin = -1.0, 950.0
    dE, dD = e_range / nb, rho_range / nb
    logg = rng.normal(0, 0.5, (nb, nb))
    E = (Emin + (np.arange(nb) + 0.5) * dE)[:, None] * N
    V = (SL(0.018) / (Dmin + (np.arange(nb) + 0.5) * dD))[None, :] * N
    beta = 1000 / (SL(0.2293) * np.asarray(T))
    logZ, datos = [], []
    for j in range(len(T)):
        a = logg - beta[j] * (E + P[j] * 100 * V)
        m = a.max()
        z = np.log(np.exp(a - m).sum()) + m
        logZ.append(z)
        datos.append(np.exp(a - z) / (dE * dD))
    g = rw.OraGrid.from_datos(N, np.array(datos), Emin, Dmin, dE, dD, T, P, np.zeros(len(T)))
    return g, np.array(logZ), (E, V, beta, logg, dE, dD)


T4 = np.array([250.0, 250.7, 251.3, 252.0])
P4 = np.array([1.0, 300.0, -200.0, 50.0])


def test_constants_fixed_point_is_log_z_ratio():
    g, logZ, _ = exact_grid(T4, P4)
    c, prev, rc = g.constants(300)
    assert rc == 0
    np.testing.assert_allclose(c, logZ - logZ[0], rtol=0, atol=1e-10)
    assert np.max(np.abs(c - prev)) < 1e-7


@pytest.mark.parametrize("e_range", [0.2, 1.0])
def test_constants_with_cutoffs_stay_close_to_closed_form(e_range):
    # wide grids send arguments past +-10 and cells below max - 10, where
    # the reference drops terms; each dropped term is below e^-10
    g, logZ, _ = exact_grid(T4, P4, e_range=e_range)
    c, _, rc = g.constants(600)
    assert rc == 0
    np.testing.assert_allclose(c, logZ - logZ[0], rtol=0, atol=5e-4)


def test_constants_not_converged_flag():
    g, _, _ = exact_grid(T4, P4)
    _, _, rc = g.constants(2)
    assert rc == 1


def test_single_histogram_identity_and_reweighting():
    g, logZ, (E, V, beta, logg, dE, dD) = exact_grid(T4[:1], P4[:1], e_range=0.05)
    own = g.distribution([0.0], T4[0], P4[0])
    np.testing.assert_allclose(own, g.datos()[0], rtol=1e-12, atol=0)
    # to another (T', P'): p' ~ p exp(-(beta' H' - beta H))
    T2, P2 = 251.0, 120.0
    b2 = 1000 / (SL(0.2293) * T2)
    w = g.datos()[0] * np.exp(-(b2 * (E + P2 * 100 * V) - beta[0] * (E + P4[0] * 100 * V)))
    w = w / (w.sum() * dE * dD)
    np.testing.assert_allclose(g.distribution([0.0], T2, P2), w, rtol=1e-9, atol=0)


def test_binder_of_two_energies():
    g, _, _ = exact_grid(T4[:1], P4[:1])
    nb = g.nb
    grid = np.zeros((nb, nb))
    grid[3, 5], grid[10, 7] = 0.25, 0.75
    e = g.g.Emin + (np.array([3, 10]) + 0.5) * g.g.dE
    r = g.g.Dmin + (np.array([5, 7]) + 0.5) * g.g.dD
    w = np.array([0.25, 0.75])
    ue, ur = g.binder_ed(grid)
    assert ue == pytest.approx(1 - (w * e**4).sum() / (3 * (w * e**2).sum() ** 2), rel=1e-12)
    assert ur == pytest.approx(1 - (w * r**4).sum() / (3 * (w * r**2).sum() ** 2), rel=1e-12)
    grid[:] = 0
    grid[4, 4] = 1.0
    assert g.binder_ed(grid) == pytest.approx((2 / 3, 2 / 3), rel=1e-12)


def test_ising3d_has_unit_norm_and_variance():
    assert quad(rw.ising3d, -8, 8)[0] == pytest.approx(1.0, abs=1e-6)
    assert quad(lambda z: z * z * rw.ising3d(z), -8, 8)[0] == pytest.approx(1.0, abs=1e-6)
    # the 2D fit of Liu and Panagiotopoulos is normalized to its fit accuracy
    assert quad(rw.ising2d, -3, 3)[0] == pytest.approx(1.0, abs=2e-3)
    assert quad(lambda z: z * z * rw.ising2d(z), -3, 3)[0] == pytest.approx(1.0, abs=2e-3)


def test_ising_peaks():
    p2, p3 = rw.peaks()
    for f, p, tol in ((rw.ising2d, p2, 1e-9), (rw.ising3d, p3, 2e-5)):
        r = minimize_scalar(lambda z: -f(z), bounds=(0.1, 2), method="bounded", options={"xatol": 1e-10})
        assert p == pytest.approx(-r.fun, abs=tol)
        assert p <= -r.fun + 1e-12


def test_scores_closed_forms():
    p2, p3 = rw.peaks()
    x = np.linspace(-2.5, 2.5, 21)
    p = np.array([rw.ising3d(v) for v in x])
    s = rw.scores(x, p, p2, p3)
    assert abs(s[1]) < 1e-12 and abs(s[3]) < 1e-12  # KL and Liu distances to itself
    # zero test: the bin nearest x = 0 is exact
    assert s[4] == pytest.approx(0.0, abs=1e-24)
    # height test: maxima at the grid points next to -+1.14
    i1, i2 = np.argmax(np.where(x < 0, p, -1)), np.argmax(np.where(x >= 0, p, -1))
    k = 7.5
    h = sum((p[i] - rw.ising3d(c)) ** 2 * k * k + (x[i] - c) ** 2 for i, c in ((i1, SL(-1.14)), (i2, SL(1.14))))
    assert s[5] == pytest.approx(h, rel=1e-12)


def synth_grid(n=3000):
    Ts = [0.1060, 0.1062, 0.1064]
    data = rw_synth.samples(Ts, n)
    g = rw.OraGrid.from_samples(8, Ts, [0.5] * 3, [len(d) for d in data], np.concatenate(data), 1.0)
    return g, Ts, data


def si(h, rho, T_hat, P_hat):
    """e and rho in SI from the physical relations: rho = 1527.3 rho^ - 23.1017
    kg/m^3, h_SI = 11.272132 h^ + 9.028914 T^ kJ/mol, P = 4694.59 P^ - 2111.5552
    bar, and e = h_SI - P v with P v = 100 P[bar] * 0.018/rho (kPa m^3/mol)."""
    rho_si = 1527.3 * rho - 23.1017
    P = 4694.59 * P_hat - 2111.5552
    return 11.272132 * h + 9.028914 * T_hat - 100 * P * 0.018 / rho_si, rho_si


def test_grid_bins_every_sample_in_its_cell():
    g, Ts, data = synth_grid()
    n_bin = int(np.sqrt(np.float32(3000)) / 2) + 20
    assert g.g.n_bin == n_bin and g.nb == int(n_bin * 10 * 1.0)
    assert g.g.n_bin_mixed == int(np.sqrt(g.nb))
    datos = g.datos()
    for j, d in enumerate(data):
        # per-simulation normalization: sum * de * drho = 1
        assert datos[j].sum() * g.g.dE * g.g.dD == pytest.approx(1.0, rel=1e-12)
        e, r = si(d[:, 0], d[:, 1], Ts[j], 0.5)
        ie = np.floor((e - g.g.Emin) / g.g.dE).astype(int)
        ir = np.floor((r - g.g.Dmin) / g.g.dD).astype(int)
        counts = np.zeros((g.nb, g.nb))
        np.add.at(counts, (ie, ir), 1)
        # a sample within 1e-5 of a bin edge may sit on either side
        assert np.abs(counts - datos[j] * len(d) * g.g.dE * g.g.dD).sum() <= 2 * 0.01 * len(d)
    # the extended range keeps ten empty bins below the smallest sample
    e_all = np.concatenate([si(d[:, 0], d[:, 1], Ts[j], 0.5)[0] for j, d in enumerate(data)])
    assert (e_all.min() - g.g.Emin) / g.g.dE == pytest.approx(10 * g.nb / n_bin, rel=1e-6)


def test_order_parameter_moments_and_rotation():
    g, Ts, _ = synth_grid()
    c, _, rc = g.constants(300)
    assert rc == 0
    T = 0.1062 * SL(140.5678) + SL(185.4676)
    P = 0.5 * SL(4694.59) + SL(-2111.5552)
    grid = g.distribution(c, T, P)
    nb = g.nb
    e = g.g.Emin + (np.arange(nb) + 0.5) * g.g.dE
    r = g.g.Dmin + (np.arange(nb) + 0.5) * g.g.dD
    for alpha in (0.3, 1.2, 1.57):
        x, p, media, a1, rc = g.normx(grid, alpha)
        assert rc == 0
        X = e[:, None] * np.sin(alpha) + r[None, :] * np.cos(alpha)
        mean = (X * grid).sum() / grid.sum()
        width = (X[grid > 0].max() - X[grid > 0].min()) / g.g.n_bin_mixed
        assert abs(media - mean) < width
        rx = np.gradient(x)
        rx[0], rx[-1] = x[1] - x[0], x[-1] - x[-2]
        m = p > 0
        assert (p * rx)[m].sum() == pytest.approx(1.0, abs=1e-9)
        assert (x * p * rx)[m].sum() == pytest.approx(0.0, abs=1e-9)
        assert (x * x * p * rx)[m].sum() == pytest.approx(1.0, abs=1e-9)


def test_errors_of_a_simulation_against_its_own_samples():
    g, Ts, data = synth_grid(6000)
    c, _, _ = g.constants(300)
    k = 1
    grid = g.distribution(c, g.T()[k], g.P()[k])
    e, r = si(data[k][:, 0], data[k][:, 1], Ts[k], 0.5)
    eh, ex, rc = g.errors(grid, 1.57, np.stack([e, r], axis=1).ravel())
    assert rc == 0
    assert np.all(ex >= 0) and np.all(eh >= 0)
    assert np.median(ex) < 0.2 and np.median(eh) < 0.1


def test_program_runs_and_checks_its_input(tmp_path):
    d = tmp_path / "case"
    rw_synth.write_case(d, 8, ["0.1058", "0.106", "0.1062", "0.1064", "0.1066", "0.1068"], "0.5", 1500)
    assert rw.run_program(d, d / "summary.txt") == 0
    names = sorted(p.name for p in d.iterdir())
    assert "Constants_L8" in names
    assert sum(n.startswith("Binder_L8_P235.7bar_Tm200.36779K_TM200.46619K") for n in names) == 2
    assert sum(n.startswith("Confront_L8") for n in names) == 2
    assert sum(n.startswith("Histogram_Reweighted_OP_L8_T") for n in names) == 1
    ct = rw_synth.read_table(d / "Constants_L8")
    assert ct.shape == (6, 3) and ct[0, 2] == 0.0
    # a second run reads the constants back and gives the same files
    before = {n: (d / n).read_text() for n in names if n.startswith(("Binder", "Confront", "Histogram"))}
    assert rw.run_program(d, d / "summary2.txt") == 0
    for n, text in before.items():
        assert (d / n).read_text() == text
    # too few sweeps: not converged, exit status 1
    d2 = tmp_path / "short"
    rw_synth.write_case(d2, 8, ["0.1058", "0.106", "0.1062"], "0.5", 800, steps=2)
    assert rw.run_program(d2, d2 / "s.txt") == 1
    # empty scan ranges are rejected before anything is read
    d3 = tmp_path / "bad"
    rw_synth.write_case(d3, 8, ["0.1058", "0.106"], "0.5", 500, T_range=("0.1065", "0.1060", "0.0002"))
    assert rw.run_program(d3, d3 / "s.txt") == 1


def test_engine_scores_match_oracle():
    rwl = pytest.importorskip("cvfbulkwater_amd.reweight")
    try:
        rwl._lib()
    except Exception as e:  # noqa: BLE001
        pytest.skip(f"libcvf.so not built: {e}")
    p2, p3 = rw.peaks()
    rng = np.random.default_rng(3)
    for _ in range(5):
        x = np.sort(rng.normal(0, 1, 40))
        p = np.abs(rng.normal(0.3, 0.2, 40))
        p[rng.random(40) < 0.2] = 0
        np.testing.assert_allclose(rwl.ising_scores(x, p), rw.scores(x, p, p2, p3), rtol=1e-13, atol=1e-15)
        for v in x:
            assert rwl.ising3d(v) == rw.ising3d(v) and rwl.ising2d(v) == rw.ising2d(v)


def test_order_parameter_axis_is_the_sine_stretch_of_a_uniform_grid():
    # Calcula_Histograma_Rotado_normX stretches the standardized uniform axis
    # by x + b^2/pi^2 sin(pi x / b) (b = 1.53) and standardizes again: the
    # returned axis must be an affine image of that map on a uniform grid
    from scipy.optimize import least_squares

    g, _, _ = synth_grid()
    c, _, _ = g.constants(300)
    grid = g.distribution(c, 0.1062 * SL(140.5678) + SL(185.4676), 0.5 * SL(4694.59) + SL(-2111.5552))
    z, _, _, _, rc = g.normx(grid, 1.2)
    assert rc == 0
    b = 1.53
    f = lambda u: b * b * np.sin(np.pi * u / b) / np.pi**2 + u  # noqa: E731
    i = np.arange(len(z))

    def resid(q):
        a, m, c0, w = q
        return a * (f(c0 + i * w) - m) - z

    w0 = (z[-1] - z[0]) / (len(z) - 1)
    fit = least_squares(resid, [1.0, 0.0, z[0], w0], xtol=1e-15, ftol=1e-15, gtol=1e-15)
    assert np.max(np.abs(fit.fun)) < 1e-5
    # the plain (unstretched) map does not fit
    lin = np.polyfit(i, z, 1)
    assert np.max(np.abs(np.polyval(lin, i) - z)) > 1e-3


# ------------------------------------------------------------ cut-offs, closed forms
def two_sim_grid(hist_cells, nb=4, T=250.0, P=1.0):
    """Two simulations at the same (T, P) holding the same normalized cell
    values; reweighted at that (T, P) every cell's arguments reduce to
    arg_j = log_n_j - ctes_j (f90:1045), the same for all cells."""
    datos = np.zeros((2, nb, nb))
    for (i, w), v in hist_cells.items():
        datos[:, i, w] = v
    return rw.OraGrid.from_datos(8, datos, -1.0, 950.0, 0.01, 2.0, [T, T], [P, P], np.zeros(2))


@pytest.mark.parametrize("ctes,xx", [
    # |arg| <= 10 everywhere: the plain sum over both simulations
    ((0.0, 10.0), lambda: np.log1p(np.exp(-10.0))),
    ((-5.0, 4.0), lambda: np.log(np.exp(5.0) + np.exp(-4.0))),
    # some |arg| > 10: terms within 10 of the largest kept, the rest dropped
    ((0.0, 15.0), lambda: 0.0),
    ((0.0, 10.0 + 1e-9), lambda: 0.0),
    ((-12.0, -3.0), lambda: 12.0 + np.log1p(np.exp(-9.0))),
    ((-12.0, -1.5), lambda: 12.0),
])
def test_calcula_switch_to_max_shifted_sum(ctes, xx):
    """f90:1047-1078: with any |arg_j| > 10 the cell sum keeps only the terms
    within 10 of the largest (strict <), otherwise it sums all terms; log_P =
    log(hist) - xx.  Dropping the switch or the cut, or making either test
    non-strict, changes log_P in one of these cells."""
    hist = {(0, 0): 0.3, (1, 2): 0.7, (3, 3): 2.0}
    g = two_sim_grid(hist)
    lp, memo, _ = g.calcula(np.array(ctes), 250.0, 1.0)
    for (i, w), v in hist.items():
        assert memo[i, w] == 0
        assert lp[i, w] == pytest.approx(np.log(2 * v) - xx(), rel=0, abs=1e-14), (ctes, i, w)
    assert memo.sum() == 16 - len(hist)


@pytest.mark.parametrize("depths,kept", [
    ((9.0, 11.0), (9.0,)),          # a cell 11 below the maximum is dropped
    ((9.999, 10.0 + 1e-6), (9.999,)),  # strict < 10
    ((3.0, 7.0), (3.0, 7.0)),
])
def test_log_z_drops_cells_ten_below_the_maximum(depths, kept):
    """f90:1106-1118: log_Z sums exp(log_P - max) over the non-empty cells
    less than 10 below the maximum, then adds the maximum."""
    hist = {(0, 1): 1.0}
    for k, d in enumerate(depths):
        hist[(2, k)] = float(np.exp(-d))
    g = two_sim_grid(hist)
    lp, memo, lz = g.calcula(np.zeros(2), 250.0, 1.0)
    top = lp[0, 1]
    assert all(lp[2, k] - top == pytest.approx(-d, abs=1e-12) for k, d in enumerate(depths))
    want = top + np.log1p(sum(np.exp(-d) for d in kept))
    assert lz == pytest.approx(want, rel=0, abs=1e-14)
    every = top + np.log1p(sum(np.exp(-d) for d in depths))
    if kept != depths:
        assert abs(lz - every) > 1e-7


def test_error_bars_of_a_two_point_distribution():
    """Err_Calcula_Histograma_Rotado (f90:1622-2039) for one simulation whose
    samples sit at two grid-cell centres, reweighted to its own point, at
    alpha = 0 (x = rho).  Both histograms (the reweighted grid projected on
    n_bin_mixed = 7 bins, the samples on n_bin = 49 bins, each over its range
    widened by two empty bins a side) are two equal spikes; standardized
    (mean 0, norm and variance 1) they sit at x = -1 and +1 with height
    (k2 - k1) / 4, where k1 = int(2 n / (n + 4)) and k2 = int(n (n + 2) /
    (n + 4)) are the spikes' bins.  Each reweighted bin i is matched to the
    nearest sample bin: error_X = that distance, error_hist = the height
    difference."""
    nb, N = 49, 8
    Emin, Dmin, dE, dD = -1.0, 950.0, 0.01, 2.0
    datos = np.zeros((1, nb, nb))
    cells = [(10, 5), (10, 30)]
    for i, w in cells:
        datos[0, i, w] = 1.0
    g = rw.OraGrid.from_datos(N, datos, Emin, Dmin, dE, dD, [250.0], [1.0], np.zeros(1))
    grid = g.distribution([0.0], 250.0, 1.0)
    assert np.count_nonzero(grid) == 2
    centre = lambda i, w: (Emin + (i + 0.5) * dE, Dmin + (w + 0.5) * dD)  # noqa: E731
    samples = np.array([centre(i, w) for i, w in cells for _ in range(500)])
    eh, ex, rc = g.errors(grid, 0.0, samples.ravel())
    assert rc == 0

    def spikes(n):
        k1, k2 = int(2 * n / (n + 4)), int(n * (n + 2) / (n + 4))
        t = 2.0 / (k2 - k1)  # standardized bin width
        x = -1.0 + (np.arange(n) - k1) * t
        h = np.zeros(n)
        h[[k1, k2]] = (k2 - k1) / 4.0
        return x, h

    xa, ha = spikes(7)
    xb, hb = spikes(nb)
    j = np.array([np.argmin(np.abs(xb - v)) for v in xa])
    np.testing.assert_allclose(ex, np.abs(xb[j] - xa), rtol=0, atol=1e-12)
    np.testing.assert_allclose(eh, np.abs(hb[j] - ha), rtol=0, atol=1e-12)
    # the spikes: exact positions, heights 46/4 against 4/4
    assert eh[1] == pytest.approx(10.5, abs=1e-12) and eh[5] == pytest.approx(10.5, abs=1e-12)
    assert ex[1] == pytest.approx(0.0, abs=1e-12) and ex[3] == pytest.approx(0.0, abs=1e-12)
