#!/usr/bin/env python3
"""Benchmark of the CVF bulk-water MC hot path on one MI355X.

Headline workload (the metric's "batched mixed-prec" form): 128 independent
(T, P) state points in one launch sequence, T = 0.55 + 0.45 k/7 (k = 0..7),
P = -0.10 + 0.06 j (j = 0..15), L = 16, 24, 32, 48 for point i % 4 (32 points
each, 5,160,960 cells), seed_i = 44648604 + 1000003 i, random start, 1
Metropolis sweep (12 half-lattice sigma sweeps) + 1 eta cube sweep + 1
volume move per MC step, FP32-compensated vdW with FP64 acceptance, the
step captured in a hipGraph.  A step = one MC step of every state point.

Beside it, the plain drop-in form is timed: one state point, L = 32,
T = 0.815, P = 0.45, FP64 vdW, host volume move (the reference's
production shape).  value = site-updates/s of the batched form
(site-update = one molecule advanced one full MC step).

The roofline is for the dominant kernel of the chosen launch configuration:
k_chunk (fused, six half-sweeps per launch) at 8 B per cell of the launch
(every word read and written once), or k_sigma (one half-lattice sigma
sweep) at 12 B per active cell; divided by the mean launch duration recorded
by the dispatch itself (hipExtLaunchKernelGGL start/stop events).

Further legs: the plain form run as the drop-in executable samples it, the
cluster form, the 128 points as sequential plain runs, the non-chessboard
and correlation modes, the L = 96 precision form, and the histogram
reweighting at its production shape.  The CPU legs (the oracle on one core
and on all usable cores) run before the GPU is touched.

Inputs are synthetic (random initial lattices, synthetic (h, rho) samples);
nothing is read from disk.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SIGMA_BYTES_PER_ACTIVE_CELL = 12
CHUNK_BYTES_PER_CELL = 8


def batch_params(fp32: bool = True):
    from cvfbulkwater_amd import FP32_COMPENSATED, FP64, VOLUME_DEVICE, Params

    pts = []
    for i in range(128):
        k, j = i % 8, i // 8
        pts.append(Params(L=(16, 24, 32, 48)[i % 4], T=0.55 + 0.45 * k / 7, P=-0.10 + 0.06 * j,
                          seed=44648604 + 1000003 * i, n_metropolis=1, n_cluster=0,
                          precision=FP32_COMPENSATED if fp32 else FP64, volume_mode=VOLUME_DEVICE))
    return pts


class HipEvents:
    """hipEventRecord on an explicit stream (the engine's own)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [C.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.lib.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.lib.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.lib.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def time_steps(sim, steps: int, warmup: int, ev: HipEvents) -> float:
    """ms per MC step over exactly `steps` steps, HIP events on the engine stream."""
    sim.step(warmup)
    sim.sync()
    a, b = ev.event(), ev.event()
    ev.record(a, sim.stream)
    sim.step(steps)
    ev.record(b, sim.stream)
    ms = ev.elapsed_ms(a, b)
    sim.sync()
    return ms / steps


def cpu_baseline(budget_s: float) -> dict:
    """The CPU oracle (same rules, serial C) on a bounded sample of the batch:
    one state point of each lattice size, stepped until the budget is used."""
    from oracle import py as O

    O.build()
    cells_steps = 0
    t0 = time.perf_counter()
    per = budget_s / 4
    used = []
    for i, L in enumerate((16, 24, 32, 48)):
        k, j = i % 8, i // 8
        sim = O.OracleSim(L=L, T=0.55 + 0.45 * k / 7, P=-0.10 + 0.06 * j, seed=44648604 + 1000003 * i)
        n, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < per:
            sim.step(1)
            n += 1
        cells_steps += n * L ** 3
        used.append(f"L={L}x{n}")
        sim.close()
    dt = time.perf_counter() - t0
    return {"value": cells_steps / dt, "unit": "site-updates/s", "cores": 1, "kind": "port",
            "sample": "CPU oracle, serial C, one state point of each L of the batch (" + ", ".join(used)
                      + f" MC steps), {dt:.1f} s on {platform.processor() or platform.machine()}"}


def _cpu_point(args) -> tuple[int, int, int]:
    """One batch state point on the CPU oracle for `budget` seconds (a
    worker of cpu_baseline_all_cores): (cells x steps, L, steps)."""
    i, budget = args
    from oracle import py as O

    L = (16, 24, 32, 48)[i % 4]
    k, j = (i % 128) % 8, (i % 128) // 8
    sim = O.OracleSim(L=L, T=0.55 + 0.45 * k / 7, P=-0.10 + 0.06 * j, seed=44648604 + 1000003 * i)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget:
        sim.step(1)
        n += 1
    sim.close()
    return n * L ** 3, L, n


def host_info() -> dict:
    model = platform.processor() or platform.machine()
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0))}


def cpu_baseline_all_cores(budget_s: float, max_workers: int = 16) -> dict:
    """The CPU oracle on every usable host core (at most 16, the GPU box's
    share): one independent batch state point per worker process, all
    running for the same budget; value = summed site-updates / wall time."""
    import multiprocessing as mp

    from oracle import py as O

    O.build()
    n = max(1, min(max_workers, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(n) as pool:
        res = pool.map(_cpu_point, [(i, budget_s) for i in range(n)])
    dt = time.perf_counter() - t0
    total = sum(r[0] for r in res)
    return {"value": total / dt, "unit": "site-updates/s", "cores": n, "kind": "port",
            "sample": f"CPU oracle, serial C, {n} worker processes each stepping one batch state point "
                      f"(L = 16/24/32/48 by index) for {budget_s:.0f} s; wall {dt:.1f} s incl. start-up"}


def calibrate_copy(mib: float, reps: int = 10) -> int:
    """hipMemcpyDtoD of a known size with the runtime's copy kernel: under
    rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch give the scale
    between the counters and bytes actually moved."""
    hip = C.CDLL("libamdhip64.so")
    n = int(mib * 1024 * 1024)
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpyDtoD(b, a, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": n, "reps": reps}))
    return 0


def tuning_signature(kname: str, t: dict) -> str:
    if kname == "k_resident":
        # instructions per step do not depend on the graph length
        return f"per_cu:{t['streams']},block:{t['block_chunk']},fused:2"
    if kname == "k_chunk":
        whole = ",chunk:12" if t.get("chunk_sweeps") == 12 else ""
        return f"streams:{t['streams']},block:{t['block_chunk']},fused:1,slab:{t['slab_planes']}{whole}"
    return f"streams:{t['streams']},block:{t['block_sigma']},fused:0"


def _profile_line(tag: str, kname: str, tuning: dict):
    """The first line "<tag> kernel=<kname> tuning=<signature> ..." of the
    counter summaries kept in profiles/, as a dict, or None."""
    sig = tuning_signature(kname, tuning)
    for path in sorted((ROOT / "profiles").glob("*.txt")):
        for line in path.read_text().splitlines():
            if not line.startswith(tag + " "):
                continue
            kv = dict(f.split("=", 1) for f in line.split()[1:] if "=" in f)
            if kv.get("kernel") == kname and kv.get("tuning") == sig:
                return kv
    return None


def measured_traffic(kname: str, tuning: dict):
    """HBM bytes per dispatch of the roofline kernel under this launch
    configuration, from the FETCH_SIZE / WRITE_SIZE summaries kept in
    profiles/ (lines "traffic kernel=... tuning=... bytes=..."), or None
    when no counter run covers the configuration."""
    kv = _profile_line("traffic", kname, tuning)
    return float(kv["bytes"]) if kv else None


def measured_valu_per_step(kname: str, tuning: dict, dispatches_per_step: float):
    """VALU wave instructions per MC step of the roofline kernel under this
    launch configuration, from the SQ counter summaries kept in profiles/
    (lines "valu kernel=... tuning=... insts=<per dispatch>" or
    "insts_per_step=..."), or None."""
    kv = _profile_line("valu", kname, tuning)
    if not kv:
        return None
    if "insts_per_step" in kv:
        return float(kv["insts_per_step"])
    return float(kv["insts"]) * dispatches_per_step


def precision_form(steps: int = 5000, interval: int = 10, ev: HipEvents | None = None) -> dict:
    """Mixed-precision large lattice: L = 96, T = 0.65, P = 0.30, 1 Metropolis
    + 1 Swendsen-Wang + 1 volume move per step, run once with FP32-compensated
    vdW and once FP64 everywhere from the same seed, sampled every `interval`
    steps.  Returns, per observable (E, H, V, density, N_HB), the largest
    per-sample relative difference and the relative differences of the
    long-run mean and variance (FP64 accumulators in both runs)."""
    import numpy as np

    from cvfbulkwater_amd import FP32_COMPENSATED, FP64, VOLUME_DEVICE, Params, Simulation

    L = 96
    N = L ** 3
    series, ms = {}, {}
    for name, prec in (("fp32c", FP32_COMPENSATED), ("fp64", FP64)):
        sim = Simulation(Params(L=L, T=0.65, P=0.30, seed=20240611, n_metropolis=1, n_cluster=1,
                                precision=prec, volume_mode=VOLUME_DEVICE))
        rows = []
        for _ in range(steps // interval):
            sim.step(interval)
            o = sim.observe()
            rows.append((o.E, o.H, o.V, N / o.V, o.n_hb))
        if ev is not None:
            ms[name] = time_steps(sim, 200, 0, ev)
            if name == "fp32c":
                # kernel classes of a step (HIP events around each launch)
                prof = sim.profile_steps(50)
                kernels = {k: {"ms_per_launch": v[0], "launches_per_step": v[1] / 50}
                           for k, v in prof.items() if v[1]}
        sim.close()
        series[name] = np.array(rows)
    a, b = series["fp32c"], series["fp64"]
    errors = {}
    for k, name in enumerate(("E", "H", "V", "density", "N_HB")):
        x, y = a[:, k], b[:, k]
        errors[name] = {
            "max_rel": float(np.max(np.abs(x - y) / np.abs(y))),
            "mean_rel": float(abs(x.mean() - y.mean()) / abs(y.mean())),
            "var_rel": float(abs(x.var() - y.var()) / y.var()) if y.var() > 0 else 0.0,
            "mean_fp64": float(y.mean()),
        }
    out = {"workload": f"L={L}, T=0.65, P=0.30, 1 Metropolis + 1 Swendsen-Wang + 1 volume move "
                       f"per step, {steps} steps sampled every {interval}, FP32-compensated vdW vs "
                       "FP64 everywhere, same seed", "errors": errors}
    if ms:
        out["ms_per_step"] = ms
        out["kernels_fp32c"] = kernels
    return out


def plain_sampled(steps: int, warmup: int, ev: HipEvents, interval: int = 10) -> dict:
    """The drop-in executable's production loop on the plain form: steps in
    calls of `interval` with the observables read after each (L = 32,
    T = 0.815, P = 0.45, FP64, host volume move)."""
    from cvfbulkwater_amd import FP64, VOLUME_HOST, Params, Simulation

    sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, precision=FP64, volume_mode=VOLUME_HOST))
    sim.step(warmup)
    sim.observe()
    a, b = ev.event(), ev.event()
    n = max(1, steps // interval)
    t0 = time.perf_counter()
    ev.record(a, sim.stream)
    for _ in range(n):
        sim.step(interval)
        sim.observe()
    ev.record(b, sim.stream)
    ev.elapsed_ms(a, b)
    wall = (time.perf_counter() - t0) * 1e3
    sim.close()
    ms = wall / (n * interval)
    return {"workload": f"plain form sampled every {interval} steps as the drop-in executable does "
                        "(observables read back each sample), wall clock", "ms_per_step": ms,
            "value": 32 ** 3 / (ms * 1e-3)}


def cluster_form(steps: int, warmup: int, ev: HipEvents) -> dict:
    """Cluster form: one state point, L = 32, T = 0.70, P = 0.45, 1 Metropolis
    sweep + 1 Swendsen-Wang step + 1 volume move per step (FP64, host volume
    move), labels kept canonical every step."""
    from cvfbulkwater_amd import FP64, VOLUME_HOST, Params, Simulation

    sim = Simulation(Params(L=32, T=0.70, P=0.45, seed=7, n_cluster=1, precision=FP64,
                            volume_mode=VOLUME_HOST))
    ms = time_steps(sim, steps, warmup, ev)
    sim.close()
    return {"workload": "cluster form: L=32, T=0.70, P=0.45, 1 Metropolis + 1 Swendsen-Wang + 1 volume "
                        "move per step, FP64, host volume move", "ms_per_step": ms,
            "value": 32 ** 3 / (ms * 1e-3)}


def fast_schedule_form(steps: int, warmup: int, ev: HipEvents, tune_steps: int) -> dict:
    """The engine's fast schedule (two-colour cell passes, CVF_SCHED_FAST) on
    the headline batch and on the plain L = 32 point: the same model and
    work per step (6 sigma proposals per molecule, eta sweep, volume move),
    another order of the sigma moves."""
    from cvfbulkwater_amd import FP64, SCHED_FAST, VOLUME_HOST, Params, Simulation

    pts = batch_params(fp32=True)
    for p in pts:
        p.schedule = SCHED_FAST
    cells = sum(p.L ** 3 for p in pts)
    sim = Simulation(pts)
    tuning = sim.autotune(tune_steps) if tune_steps else sim.tuning
    ms = time_steps(sim, steps, warmup, ev)
    sim.close()
    plain = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, precision=FP64, volume_mode=VOLUME_HOST,
                              schedule=SCHED_FAST))
    ptuning = plain.autotune(tune_steps) if tune_steps else plain.tuning
    pms = time_steps(plain, max(steps, 1000), warmup, ev)
    plain.close()
    return {"workload": "fast schedule: the headline batch and the plain L=32 point with two-colour "
                        "cell passes instead of the reference's 12 half-sweeps",
            "batch": {"value": cells / (ms * 1e-3), "ms_per_step": ms, "tuning": tuning},
            "plain": {"value": 32 ** 3 / (pms * 1e-3), "ms_per_step": pms, "tuning": ptuning}}


def nonchess_form(steps: int, warmup: int, ev: HipEvents, cpu_seconds: float = 0.0) -> dict:
    """Flag_chessboard = 0 at the plain shape: L = 32, T = 0.815, P = 0.45,
    eta fixed, hydrogen bonds gated by the per-molecule counters, FP64,
    host volume move; with cpu_seconds, the CPU oracle runs the same mode."""
    from cvfbulkwater_amd import FP64, VOLUME_HOST, Params, Simulation

    sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, flag_chessboard=0, precision=FP64,
                            volume_mode=VOLUME_HOST))
    ms = time_steps(sim, steps, warmup, ev)
    sim.close()
    res = {"workload": "non-chessboard mode (Flag_chessboard=0): L=32, T=0.815, P=0.45, 1 Metropolis + "
                       "1 volume move per step, FP64, host volume move", "ms_per_step": ms,
           "value": 32 ** 3 / (ms * 1e-3)}
    if cpu_seconds > 0:
        from oracle import py as O

        o = O.OracleSim(L=32, T=0.815, P=0.45, seed=44648604, flag_chessboard=0)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            o.step(1)
            n += 1
        dt = time.perf_counter() - t0
        o.close()
        res["cpu_baseline"] = {"value": n * 32 ** 3 / dt, "unit": "site-updates/s", "cores": 1, "kind": "port",
                               "sample": f"CPU oracle, serial C, the same state point, {n} MC steps"}
    return res


def correlation_form(steps: int, warmup: int) -> dict:
    """Flag_correlation = 1 as the drop-in executable runs it: after every
    step the sigma histogram (one device reduction, six counters back to the
    host) feeds the multi-tau correlator; L = 32 plain form, wall clock."""
    from cvfbulkwater_amd import FP64, VOLUME_HOST, Correlator, Params, Simulation

    sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, precision=FP64, volume_mode=VOLUME_HOST))
    if sim.tuning["fused"] == 2:
        # as the executable does for recorded steps: the fused slab plan
        sim.tuning = {"fused": 1, "slab_planes": 0, "block_chunk": 512, "graph_steps": 16}
    corr = Correlator()
    sim.step(warmup)
    sim.step_histograms(10)  # the recording graphs are captured here, not in the timed loop
    sim.sync()
    t0 = time.perf_counter()
    # as the drop-in runs it: the histograms of the steps between two samples
    # (interval 10) collect on the device and come back together
    for _ in range(steps // 10):
        for h in sim.step_histograms(10):
            corr.add(h.max() / (6.0 * 32 ** 3))
    ms = (time.perf_counter() - t0) * 1e3 / (steps // 10 * 10)
    sim.close()
    return {"workload": "Flag_correlation=1 on the plain form: the sigma histogram after every step into "
                        "the multi-tau correlator, histograms copied back every 10 steps, wall clock",
            "ms_per_step": ms, "value": 32 ** 3 / (ms * 1e-3)}


def plain_sequential(steps: int, warmup: int, ev: HipEvents) -> dict:
    """The headline batch's 128 state points run one after another in the
    plain drop-in form (one point per handle, its own T, P and seed, FP64,
    host volume move, default launch plan): `steps` timed steps each, HIP
    events on each handle's stream; value = all cells x steps / summed time."""
    from cvfbulkwater_amd import FP64, VOLUME_HOST, Simulation

    total_ms, cells, by_L = 0.0, 0, {}
    for p in batch_params(fp32=False):
        p.precision, p.volume_mode = FP64, VOLUME_HOST
        sim = Simulation(p)
        ms = time_steps(sim, steps, warmup, ev)
        sim.close()
        total_ms += ms * steps
        cells += p.L ** 3 * steps
        by_L.setdefault(p.L, []).append(ms)
    return {"workload": f"the 128 batch state points as sequential plain runs (FP64, host volume move), "
                        f"{steps} timed steps each", "value": cells / (total_ms * 1e-3),
            "ms_per_step_by_L": {L: sum(v) / len(v) for L, v in by_L.items()}}


DROPIN_INPUT = """Pressure 0.45
Temperature 0.815
Seed 44648604
Metropolis_Steps 1
Cluster_Steps {nc}
Equilibrium_Steps 0
Sampling_Steps {n}
Sampling_Interval 10
R_cutoff 6.0
Flag_config 0
Flag_distances 0
Flag_correlation 0
Flag_chessboard 1
Device_ID 0
"""


def dropin_form(steps: int) -> dict:
    """The drop-in executable end to end at the production shape (L = 32,
    T = 0.815, P = 0.45, sampled every 10 steps with data.out and
    cluster.out rows), Cluster_Steps 0 and 1: wall clock of a run of
    `steps` + 1000 sampling steps minus a run of 1000, per step, so process
    start, set-up and config.out are not counted."""
    import tempfile

    from cvfbulkwater_amd import EXE_PATH

    res = {"workload": "cvf_water3d end to end, L=32, T=0.815, P=0.45, data.out + cluster.out row every "
                       "10 steps, wall clock per step (difference of two run lengths)"}
    for nc in (0, 1):
        times = {}
        for n in (1000, steps + 1000):
            with tempfile.TemporaryDirectory() as d:
                Path(d, "input_data").write_text(DROPIN_INPUT.format(nc=nc, n=n))
                t0 = time.perf_counter()
                subprocess.run([str(EXE_PATH)], cwd=d, check=True, capture_output=True, timeout=600)
                times[n] = time.perf_counter() - t0
        ms = (times[steps + 1000] - times[1000]) * 1e3 / steps
        res[f"cluster_steps_{nc}"] = {"ms_per_step": ms, "value": 32 ** 3 / (ms * 1e-3)}
    return res


def reweight_samples(n_sim: int = 11, n: int = 100000, seed: int = 9):
    """(h, rho) samples in internal units for the histogram reweighting leg:
    an isobar of n_sim temperatures 0.0001 apart near T^ = 0.1063, each a
    T-dependent mixture of a low- and a high-density blob, n samples each."""
    import numpy as np

    rng = np.random.default_rng(seed)
    T_hat = [0.1058 + 0.0001 * k for k in range(n_sim)]
    out = []
    for T in T_hat:
        hi = rng.random(n) < 1.0 / (1.0 + np.exp((T - 0.1063) / 0.0004))
        rho = np.where(hi, 0.690, 0.640) + 0.006 * rng.standard_normal(n)
        h = np.where(hi, -2.80, -2.72) + 0.012 * rng.standard_normal(n) + 0.5 * (rho - 0.665) + 0.8 * (T - 0.1063)
        out.append(np.round(np.stack([h, rho], axis=1), 5))
    return T_hat, out


def reweight_form(sweeps: int = 300, cpu_sweeps: int = 2, with_cpu: bool = True) -> dict:
    """Histogram reweighting at the reference's production shape: 11
    simulations x 100000 samples, BIN_FCT 2 (3560 x 3560 grid), L = 8.
    Times `sweeps` Jacobi sweeps of the constants (the program's STEP), a
    64-point x 3-angle scan with the projections reduced on the device, and
    the same 64 densities reweighted and copied to the host, wall clock
    around the library calls; the CPU oracle (the Fortran loops restated in C, serial) times a
    few sweeps of the same input."""
    from cvfbulkwater_amd.reweight import Reweighting

    T_hat, data = reweight_samples()
    eng = Reweighting(8, T_hat, [0.5] * len(T_hat), data, 2.0)
    info = eng.info()
    eng.solve_constants(3)  # warm-up
    t0 = time.perf_counter()
    c, prev, ok = eng.solve_constants(sweeps)
    sweep_ms = (time.perf_counter() - t0) * 1e3 / sweeps
    TK = [(0.1058 + 0.00002 * k) * 140.5678 + 185.4676 for k in range(64)]
    PB = [0.5 * 4694.59 - 2111.5552] * 64
    eng.reweight(TK[:4], PB[:4])
    t0 = time.perf_counter()
    eng.reweight(TK, PB)
    copy_ms = (time.perf_counter() - t0) * 1e3
    angles = [1.55, 1.57, 1.59]
    eng.scan(TK[:4], PB[:4], angles)
    t0 = time.perf_counter()
    eng.scan(TK, PB, angles)
    scan_ms = (time.perf_counter() - t0) * 1e3
    eng.close()
    terms = info["n_occupied"] * len(T_hat) ** 2
    res = {"workload": "histogram reweighting: 11 simulations x 100000 (h, rho) samples, BIN_FCT 2, "
                       f"{info['n_bin2d']}^2 grid, {info['n_occupied']} occupied cells, {sweeps} constant "
                       "sweeps (wall clock)", "ms_per_sweep": sweep_ms, "converged": ok,
           "terms_per_s": terms / (sweep_ms * 1e-3),
           "scan_64_points_3_angles_ms": scan_ms,
           "reweight_64_points_to_host_ms": copy_ms,
           "densities_copied_mb": 64 * info["n_occupied"] * 8 / 2**20}
    if with_cpu:
        import numpy as np
        from oracle import rw as ORW

        g = ORW.OraGrid.from_samples(8, T_hat, [0.5] * len(T_hat), [len(d) for d in data],
                                     np.concatenate(data), 2.0)
        t0 = time.perf_counter()
        g.constants(cpu_sweeps)
        cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_sweeps
        g.close()
        res["cpu_baseline"] = {"value": terms / (cpu_ms * 1e-3), "unit": "cell-simulation terms/s",
                               "ms_per_sweep": cpu_ms, "cores": 1, "kind": "port",
                               "sample": f"CPU oracle, serial C over the full grid as the Fortran loops run, "
                                         f"{cpu_sweeps} sweeps of the same input"}
        res["speedup_per_sweep"] = cpu_ms / sweep_ms
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # step counts are whole numbers of the longest captured step graph (64
    # steps), so that no shorter tail graph's launch cost lands in a figure
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=2048)
    ap.add_argument("--profile-steps", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tune-steps", type=int, default=20,
                    help="MC steps per autotuner trial (0: default launch configuration)")
    ap.add_argument("--tuning", type=json.loads, default=None, metavar="JSON",
                    help="launch configuration of the batch leg, e.g. '{\"fused\": 0}' "
                         "(skips the autotuner)")
    ap.add_argument("--plain-tuning", type=json.loads, default=None, metavar="JSON",
                    help="launch configuration of the plain leg (skips its autotuner)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-steps", type=int, default=5000)
    ap.add_argument("--only", choices=["batch", "plain", "plain1", "precision", "cluster", "reweight", "fast"],
                    default=None,
                    help="run one leg only (for profiling)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(state points shard across GPUs with no exchange, "
                                   "but no multi-GPU run is provided)"}))
        return 2

    from cvfbulkwater_amd import FP64, VOLUME_HOST, Params, Simulation

    out = {}
    if not args.no_cpu_baseline and args.only is None:
        # the CPU legs run before this process touches the GPU
        out["cpu"] = cpu_baseline(args.cpu_seconds)
        out["cpu_all"] = cpu_baseline_all_cores(args.cpu_seconds / 2)
    ev = HipEvents()
    if args.only in (None, "batch"):
        pts = batch_params(fp32=True)
        cells = sum(p.L ** 3 for p in pts)
        sim = Simulation(pts)
        if args.tuning is not None:
            sim.tuning = args.tuning
            tuning = sim.tuning
        else:
            tuning = sim.autotune(args.tune_steps) if args.tune_steps else sim.tuning
        ms = time_steps(sim, args.steps, args.warmup, ev)
        prof = sim.profile_steps(args.profile_steps)
        plan = sim.plan(0)
        sim.close()
        if prof["resident"][1]:
            # resident kernel: every word staged once and written back once
            # per launch; the profiled steps run in one launch
            kname, k_ms = "k_resident", prof["resident"][0]
            per_step = prof["resident"][1] / args.profile_steps
            k_bytes = cells * CHUNK_BYTES_PER_CELL
        elif prof["chunk"][1]:
            # fused slab kernel: one read and one write of every word per launch
            kname, k_ms = "k_chunk", prof["chunk"][0]
            per_step = prof["chunk"][1] / args.profile_steps
            k_bytes = cells * CHUNK_BYTES_PER_CELL * 2 / per_step
        else:
            # one launch sweeps one group's points; the groups split the cells
            kname, k_ms = "k_sigma", prof["sigma"][0]
            per_step = prof["sigma"][1] / args.profile_steps
            k_bytes = (cells // 2) * SIGMA_BYTES_PER_ACTIVE_CELL * 12 / per_step
        achieved = k_bytes / (k_ms * 1e-3) / 1e9
        out.update(batch_ms=ms, batch_cells=cells, batch_rate=cells / (ms * 1e-3), prof=prof, tuning=tuning,
                   kname=kname, sig_ms=k_ms, sig_bytes=k_bytes, achieved=achieved, per_step=per_step, plan=plan)
    if args.only in (None, "plain", "plain1"):
        sim = Simulation(Params(L=32, T=0.815, P=0.45, seed=44648604, precision=FP64,
                                volume_mode=VOLUME_HOST))
        if args.plain_tuning is not None:
            sim.tuning = args.plain_tuning
            ptuning = sim.tuning
        else:
            ptuning = sim.autotune(args.tune_steps) if args.tune_steps else sim.tuning
        pms = time_steps(sim, args.plain_steps, args.warmup, ev)
        sim.close()
        out.update(plain_ms=pms, plain_rate=32 ** 3 / (pms * 1e-3), plain_tuning=ptuning)
    if args.only in (None, "plain"):
        out.update(seq=plain_sequential(max(args.plain_steps // 10, 100), args.warmup, ev),
                   dropin=dropin_form(20000),
                   cluster=cluster_form(max(args.plain_steps // 4, 100), args.warmup, ev),
                   sampled=plain_sampled(args.plain_steps, args.warmup, ev),
                   nonchess=nonchess_form(max(args.plain_steps // 4, 100), args.warmup, ev,
                                          0.0 if args.no_cpu_baseline else 3.0),
                   correlation=correlation_form(max(args.plain_steps // 4, 100), args.warmup))
    if args.only == "cluster":
        out["cluster_only"] = cluster_form(args.plain_steps, args.warmup, ev)
    if args.only in (None, "precision"):
        out["precision"] = precision_form(args.precision_steps, 10, ev)
    if args.only in (None, "fast"):
        out["fast"] = fast_schedule_form(args.steps, args.warmup, ev, args.tune_steps)
    if args.only in (None, "reweight"):
        out["reweight"] = reweight_form(with_cpu=not args.no_cpu_baseline)

    line = {"metric": "site-updates/s per GPU, batched mixed-prec vs plain; achieved HBM GB/s",
            "unit": "site-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic"}
    if "batch_rate" in out:
        line.update(value=out["batch_rate"], ms_per_step=out["batch_ms"])
        line["config"] = {"workload": "batched scan: 128 state points, L in {16,24,32,48} x32, "
                                      "T in [0.55,1.00] x P in [-0.10,0.80], 1 Metropolis + 1 eta "
                                      "sweep + 1 volume move per step, FP32-compensated vdW, "
                                      "hipGraph step", "cells": out["batch_cells"],
                          "parallelism": "one GPU, state points batched",
                          "tuning": out["tuning"], "plan_point0": out["plan"]}
        hbm = {"achieved": out["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": out["achieved"] / HBM_PEAK_GBS, "traffic": measured_traffic(out["kname"], out["tuning"]),
               "bytes_per_launch": out["sig_bytes"], "launch_ms": out["sig_ms"]}
        line["kernel_ms"] = {k: v[0] for k, v in out["prof"].items() if v[1]}
        valu = measured_valu_per_step(out["kname"], out["tuning"], out["per_step"])
        if valu is not None:
            # the kernel is bound by integer VALU issue, not HBM (its state is
            # cache- or LDS-resident): its wave64 instructions per MC step
            # (rocprofv3 SQ_INSTS_VALU of the same launch configuration)
            # against the chip's issue peak, 256 CUs x 4 SIMDs x 2.4 GHz / 4
            # cycles per wave64 instruction (v_mad_u64_u32 and v_mul_hi_u32
            # issue at that rate on gfx950); HBM kept beside it
            peak = 256 * 4 * 2.4e9 / 4
            rate = valu / (out["batch_ms"] * 1e-3)
            line["roofline"] = {"bound": "valu", "kernel": out["kname"], "achieved": rate / 1e9,
                                "peak": peak / 1e9, "unit": "G wave-instr/s", "frac": rate / peak,
                                "traffic": hbm["traffic"], "insts_per_step": valu, "hbm": hbm}
        else:
            line["roofline"] = {"bound": "hbm", "kernel": out["kname"], **hbm}
    if "plain_rate" in out:
        line["plain"] = {"workload": "plain drop-in form: L=32, T=0.815, P=0.45, FP64 vdW, host "
                                     "volume move", "value": out["plain_rate"],
                         "ms_per_step": out["plain_ms"], "steps": args.plain_steps,
                         "tuning": out["plain_tuning"]}
    if "seq" in out:
        line["cluster_form"] = out["cluster"]
        line["plain_sampled"] = out["sampled"]
        line["nonchess_form"] = out["nonchess"]
        line["correlation_form"] = out["correlation"]
        line["dropin_form"] = out["dropin"]
        line["plain_sequential_128"] = out["seq"]
    if "plain_rate" in out:
        if "batch_rate" in out:
            line["batched_over_plain"] = out["batch_rate"] / out["plain_rate"]
            line["batched_over_sequential_plain"] = out["batch_rate"] / out["seq"]["value"]
        else:
            line.update(value=out["plain_rate"], ms_per_step=out["plain_ms"])
    if "cluster_only" in out:
        line.update(value=out["cluster_only"]["value"], ms_per_step=out["cluster_only"]["ms_per_step"],
                    cluster_form=out["cluster_only"])
    if "precision" in out:
        line["precision_form"] = out["precision"]
        if "value" not in line:
            line.update(value=None, ms_per_step=out["precision"].get("ms_per_step", {}).get("fp32c"))
    if "fast" in out:
        line["fast_schedule"] = out["fast"]
        if "value" not in line:
            line.update(value=out["fast"]["batch"]["value"], ms_per_step=out["fast"]["batch"]["ms_per_step"])
    if "reweight" in out:
        line["reweight_form"] = out["reweight"]
        if "value" not in line:
            line.update(value=out["reweight"]["terms_per_s"], ms_per_step=out["reweight"]["ms_per_sweep"])
    if "batch_rate" in out:
        # 6 arms x NMETROPOLIS (= 1) sigma proposals per molecule-step
        line["sigma_attempts_per_s"] = 6 * out["batch_rate"]
    if "cpu" in out:
        line["cpu_baseline"] = out["cpu"]
        line["cpu_baseline_all_cores"] = out["cpu_all"]
        line["host"] = host_info()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
