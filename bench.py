"""Benchmark: FP64 pair-interactions/s of the BIA-ALG step on one MI355X.

Workload (default): N = 65536 equal-mass gravitating particles, Plummer sphere (a = 1, M = 1,
G = 1, m = 1/N) with 20 % of the particles in hard binaries (a in [1e-3, 1e-2], thermal e),
dt = 1e-3, eps = -2Gm, seed 65536.  One "step" = one BIA step of every unordered pair
(N(N-1)/2 pair interactions) followed by the state update, state resident in HBM.

Modes measured on the same input: plain (the reference's conic solvers, guesses and iteration
control), plain_learned (the same solvers and iteration control, the hyperbola and ellipse Newton
started from network 1's anomaly, the reference's guess the fallback and the judge), guess
(universal-variable Kepler drift, every pair from its series start, no network) and learned (the
same drift; a pair whose series start is far off takes its second point from network 0's table).
Each network is trained on the GPU first, from samples the solvers record on other systems
(training time reported separately).

Printed: one JSON line (rank 0).  ``value`` is the pair-interactions/s of the learned mode
(guess if training is skipped with --no-learn).  ``roofline`` prices the dominant kernel (the
pair kernel of that mode) with the algorithmic FP64 flop count of its own algorithm (stats
counters of an untimed step on the same input; bia_3d_gpu_amd.flops_per_step) over its average
duration, measured with HIP events recorded on the kernel's own stream; ``roofline.useful_work``
gives the reference algorithm's flops for the same pairs (the plain mode's count) over the same
duration, a rate of reference-equivalent work that can exceed the FP64 peak.  ``energy`` gives
|dE/E0| over the timed steps of each mode from the library's double-double invariants.
``cpu_baseline`` times the reference's own Fortran CPU pair loop (oracle/_ref) on bounded row
samples, on one core and on one process per host core (at most 16), and its RKF path on one core
(CPU model, compiler and flags given).  ``agreement_vs_rkf45`` compares every mode's first step
with an untimed per-pair RKF45 step of the same system (the reference's own integrator, an
independent solution of the same pairs).  ``reductions`` gives the HBM rate of the step's
tile-partial reduction kernel (its own HIP events) and of the invariants.  Side measurements:
``learned_vs_guess`` (headline, longer steps, the reference fixture), ``plain_learned_vs_plain``
(time, evaluations, caps and fallbacks of the reference solvers with and without the learned
starts), ``reference_fixture``
(with each mode's normwise agreement with the reference's own .BIA output), ``config1`` (1000
steps against the DIM), ``energy_order`` (|dE/E0| over a fixed time at dt, dt/2, dt/4), and the
two ensembles.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 pair-interactions/s, BIA-ALG (learned init) vs plain, 1 MI355X; % FP64 peak"
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_state(n: int, seed: int):
    from bia_3d_gpu_amd import ic
    return ic.plummer_with_binaries(n, seed)


def run_mode(ctx, st_host, eps, dt, itr, mode, steps, warmup, autotune=True):
    import torch

    from bia_3d_gpu_amd import flops_per_step

    dev = "cuda:0"
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    s = torch.from_numpy(st_host).to(dev)
    d = torch.empty_like(s)
    # untimed stats step on the initial state: branch mix and iteration counts
    stats = torch.zeros(16, dtype=torch.int64, device=dev)
    ctx.step_(s, d, eps, dt, mode=mode, itr=itr, stats=stats)
    torch.cuda.synchronize()
    st = stats.cpu().numpy().astype(np.uint64)
    first_delta = d.cpu().numpy()  # the first step's deltas, compared across modes
    flops = flops_per_step(st, mode)
    tuned = None
    if autotune:  # persistent grid size and register-occupancy variant, timed on this input
        tms, grid, occ = ctx.autotune_(s, d, [s.shape[1]], eps=[eps], dt=dt, mode=mode, itr=itr)
        tuned = {"grid": grid, "occupancy_variant": occ, "ms": tms}
    for _ in range(warmup):
        ctx.step_(s, d, eps, dt, mode=mode, itr=itr)
        s.add_(d)
    inv0 = torch.zeros(11, dtype=torch.float64, device=dev)
    ctx.invariants_(s, inv0, eps)
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(steps):  # enqueued back to back: no host synchronisation inside the timed steps
        ctx.step_(s, d, eps, dt, mode=mode, itr=itr)
        s.add_(d)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the pair kernel and the tile-partial reduction of each timed step (HIP events on the
    # library's stream), summed over the stages of a step
    kms, red_ms = step_kernel_ms(ctx)
    nstages, wbytes = ctx.plan_info()
    ctx.set_profiling(False)
    ms = e0.elapsed_time(e1) / steps
    inv1 = torch.zeros(11, dtype=torch.float64, device=dev)
    ctx.invariants_(s, inv1, eps)
    # the invariants' kernels (on the same stream), timed over five calls
    i0, i1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    i0.record(stream)
    for _ in range(5):
        ctx.invariants_(s, inv1, eps)
    i1.record(stream)
    torch.cuda.synchronize()
    inv_ms = i0.elapsed_time(i1) / 5
    a, b = inv0.cpu().numpy(), inv1.cpu().numpy()
    e_0, e_1 = a[0] + a[1], b[0] + b[1]  # KE + PE = 2E/m
    n = s.shape[1]
    from bia_3d_gpu_amd import partial_doubles
    # the reduction reads every partial once (bia_3d_gpu_amd.partial_doubles: the column partials
    # of each block of four row tiles and the row partials of its items) and the velocities, and
    # writes the deltas (once per stage, but the last, for the running sums)
    red_bytes = 8 * partial_doubles(n) + 24 * n + 48 * n + 2 * 48 * n * max(0, nstages - 1)
    return dict(ms_per_step=ms, wall_s=wall, kernel_ms=kms, stats=st.tolist(), flops=flops, tuned=tuned,
                tflops=flops / (kms / 1e3) / 1e12, dE_rel=abs(e_1 - e_0) / abs(e_0), dE_steps=steps,
                first_delta=first_delta,
                reduction={"kernel_ms": red_ms, "bytes_per_step": red_bytes, "stages": nstages,
                           "workspace_bytes": wbytes, "GB_per_s": red_bytes / (red_ms * 1e-3) / 1e9},
                invariants={"ms": inv_ms, "state_bytes": 48 * n, "GB_per_s_of_state": 48 * n / (inv_ms * 1e-3) / 1e9})


def step_kernel_ms(ctx):
    """(pair kernel, reduction) milliseconds per step, summed over the step's stages, from the most
    recent whole steps in the library's timing rings."""
    kms, rms = ctx.kernel_ms(), ctx.reduce_kernel_ms()
    nstages, _ = ctx.plan_info()
    whole = (min(len(kms), len(rms)) // nstages) * nstages
    if whole == 0:
        return float("nan"), float("nan")
    return float(np.sum(kms[-whole:]) / (whole / nstages)), float(np.sum(rms[-whole:]) / (whole / nstages))


def agreement(st_host, d, ref):
    """Normwise agreement of one step's deltas with a reference step: max |(x + d) - (x + ref)| /
    max |x + ref| over positions and over velocities, and the particles whose deltas differ by
    more than 1e-6 and 1e-8 of the largest delta of that component."""
    a, b = st_host + d, st_host + ref
    off = (np.abs(d - ref) / np.abs(ref).max(axis=1, keepdims=True)).max(axis=0)
    return {"positions": float(np.abs(a[:3] - b[:3]).max() / np.abs(b[:3]).max()),
            "velocities": float(np.abs(a[3:] - b[3:]).max() / np.abs(b[3:]).max()),
            "particles_off_1e-6": int((off > 1e-6).sum()), "particles_off_1e-8": int((off > 1e-8).sum())}


def run_ensemble(ctx, mode, steps: int, warmup: int, seed: int = 512):
    """Side measurement: 512 independent Plummer systems (a = 1, M = 1, G = 1, m = 1/N_s, each
    from its own seed), N_s log-uniform in [64, 4096], dt = 1e-3, in one persistent launch per
    step; the steps are replayed from a captured hipGraph.  Then the same number of steps through
    the monitored driver (statistics every step, invariants of all 512 systems four times)."""
    import torch

    from bia_3d_gpu_amd import driver, ic

    sizes = ic.ensemble_sizes(512, 64, 4096, seed).astype(int)
    st = np.ascontiguousarray(np.concatenate([ic.plummer_system(int(n), seed + 1 + k) for k, n in enumerate(sizes)],
                                             axis=1))
    eps = [-2.0 / n for n in sizes]
    s = torch.from_numpy(st).to("cuda:0")
    d = torch.empty_like(s)
    ctx.autotune_(s, d, sizes, eps=eps, dt=1e-3, mode=mode)
    ctx.run_steps_(s, d, sizes, max(warmup, 1), eps=eps, dt=1e-3, mode=mode, use_graph=True)
    torch.cuda.synchronize()
    # at least 20 timed steps: a call's fixed cost (graph launch, synchronisation, about 0.5 ms)
    # would otherwise weigh on a 3.4 ms step
    timed = max(steps, 20)
    t0 = time.perf_counter()
    ctx.run_steps_(s, d, sizes, timed, eps=eps, dt=1e-3, mode=mode, use_graph=True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / timed
    pairs = float(sum(int(n) * (int(n) - 1) // 2 for n in sizes))
    every = max(1, steps // 4)
    rep = driver.run(ctx, s, sizes, steps, eps=eps, dt=1e-3, mode=mode, monitor_every=every)
    lean = driver.run(ctx, s, sizes, steps, eps=eps, dt=1e-3, mode=mode, monitor_every=every, collect_stats=False)
    h = rep.histogram()
    return {"systems": 512, "particles": int(sizes.sum()), "pairs_per_step": pairs, "ms_per_step": sec * 1e3,
            "pairs_per_s": pairs / sec, "steps": timed, "graph": True,
            "monitored": {"ms_per_step": rep.seconds * 1e3 / steps, "monitor_every": every,
                          "ms_per_step_without_statistics": lean.seconds * 1e3 / steps,
                          "max_abs_energy_error": float(np.abs(rep.energy_error()).max()),
                          "evaluations_per_pair": h["evaluations_per_pair"], "iteration_cap": h["iteration cap"]}}


def run_mass_ensemble(ctx, mode, steps: int, warmup: int, seed: int = 4096):
    """Side measurement: 64 Plummer systems, N log-uniform in [4096, 32768] (9.6e9 pairs per
    step), masses log-uniform in [0.1, 10] normalised to 1 per system, gravity coupling
    eps_ij = -(m_i + m_j) with the centre-of-mass split, dt = 5e-4; steps replayed from a captured
    hipGraph."""
    import torch

    from bia_3d_gpu_amd import COUPLING_GRAVITY, ic

    sizes = [int(n) for n in ic.ensemble_sizes(64, 4096, 32768, seed)]
    rng = np.random.default_rng(seed + 1)
    states, masses = [], []
    for k, n in enumerate(sizes):
        states.append(ic.plummer_system(n, seed + 904 + k))
        m = np.exp(rng.uniform(np.log(0.1), np.log(10), n))
        masses.append(m / m.sum())
    s = torch.from_numpy(np.ascontiguousarray(np.concatenate(states, axis=1))).to("cuda:0")
    mc = torch.from_numpy(np.concatenate(masses)).to("cuda:0")
    d = torch.empty_like(s)
    kw = dict(dt=5e-4, mode=mode, coupling=COUPLING_GRAVITY, coup=1.0, mass=mc)
    ctx.autotune_(s, d, sizes, **kw)
    ctx.run_steps_(s, d, sizes, max(warmup, 1), use_graph=True, **kw)
    inv = torch.zeros(2, len(sizes), 11, dtype=torch.float64, device="cuda:0")
    ctx.invariants_coupled_(s, sizes, inv[0], COUPLING_GRAVITY, 1.0, mass=mc)  # mass-weighted, KE + PE = 2E
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.run_steps_(s, d, sizes, steps, use_graph=True, **kw)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    ctx.invariants_coupled_(s, sizes, inv[1], COUPLING_GRAVITY, 1.0, mass=mc)
    e = (inv[:, :, 0] + inv[:, :, 1]).cpu().numpy()
    de = np.abs(e[1] / e[0] - 1)
    pairs = float(sum(n * (n - 1) // 2 for n in sizes))
    return {"systems": 64, "particles": int(sum(sizes)), "pairs_per_step": pairs, "ms_per_step": sec * 1e3,
            "pairs_per_s": pairs / sec, "steps": steps, "graph": True, "coupling": "gravity, mass spectrum",
            "dE_rel_median": float(np.median(de)), "dE_rel_max": float(de.max())}


def run_fixture(ctx, steps: int, learned: bool, plain_learned: bool = False):
    """Side measurement on the reference's own N = 1332 input (tests/golden/0001332_10^0.bbb), the
    only case with published reference timings (one BIA step, dt = 1, eps = -6e-7, TITAN V: ALG
    0.193 s and per-pair RKF 0.293 s of GPU solver time; the DIM RKF45 run to t = 1: 2.946 s
    total).  Times the plain ALG mode (the reference's algorithm), the reference solvers with
    learned starts, the guess and learned modes, the per-pair RKF method and the DIM run, and gives each step's normwise agreement with the
    reference's own output on this input (OUT/0001332_10^0.BIA, its CUDA ALG run)."""
    import torch

    from bia_3d_gpu_amd import MODE_GUESS, MODE_LEARN, MODE_PLAIN, MODE_PLAIN_LEARN, io

    st = io.read_bbb(os.path.join(ROOT, "tests", "golden", "0001332_10^0.bbb"))[0]
    (_, bia_ref), = io.read_timed(os.path.join(ROOT, "tests", "golden", "0001332_10^0.BIA"))
    n = st.shape[1]
    pairs = n * (n - 1) / 2
    out = {"n": n, "pairs_per_step": pairs, "reference_published": {"bia_alg_gpu_solver_s": 0.193,
                                                                     "bia_rkf_gpu_solver_s": 0.293,
                                                                     "dim_total_s": 2.946}}
    s = torch.from_numpy(np.ascontiguousarray(st)).to("cuda:0")
    d = torch.empty_like(s)
    runs = [("plain", 1, MODE_PLAIN, steps), ("guess", 1, MODE_GUESS, steps), ("rkf", 0, MODE_PLAIN, 2)]
    if learned:
        runs.insert(2, ("learned", 1, MODE_LEARN, steps))
    if plain_learned:
        runs.insert(1, ("plain_learned", 1, MODE_PLAIN_LEARN, steps))
    stats = torch.zeros(16, dtype=torch.int64, device="cuda:0")
    for name, method, mode, k in runs:
        stats.zero_()
        ctx.step_(s, d, -6e-7, 1.0, method=method, mode=mode, stats=stats)
        torch.cuda.synchronize()
        c = stats.cpu().numpy()
        # normwise agreement with the reference's own output for this input (OUT/0001332_10^0.BIA)
        agree = agreement(st, d.cpu().numpy(), bia_ref)
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(k):
            ctx.step_(s, d, -6e-7, 1.0, method=method, mode=mode)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / k
        kms = step_kernel_ms(ctx)[0]
        ctx.set_profiling(False)
        out[name] = {"s_per_step": sec, "pairs_per_s": pairs / sec, "kernel_ms": kms,
                     "evaluations_per_pair": float(c[4:8].sum() / max(1, c[0:4].sum())), "learned_starts": int(c[14]),
                     "fallbacks": int(c[15]),
                     "vs_reference_BIA": {"positions": agree["positions"], "velocities": agree["velocities"]}}
    if learned:
        out["learned_vs_guess"] = {"kernel_ms_ratio": out["learned"]["kernel_ms"] / out["guess"]["kernel_ms"],
                                   "s_per_step_ratio": out["learned"]["s_per_step"] / out["guess"]["s_per_step"]}
    if plain_learned:
        out["plain_learned_vs_plain"] = {
            "kernel_ms_ratio": out["plain_learned"]["kernel_ms"] / out["plain"]["kernel_ms"],
            "s_per_step_ratio": out["plain_learned"]["s_per_step"] / out["plain"]["s_per_step"]}
    y = np.ascontiguousarray(st.T).reshape(-1)
    t0 = time.perf_counter()
    r = ctx.dim_rkf45(y, -3e-7, 0.0, 1.0)
    out["dim"] = {"s": time.perf_counter() - t0, "accepted_steps": int(r["naccept"]), "ier": int(r["ier"])}
    return out


def run_config1(ctx, steps: int = 1000, plain_learned: bool = False):
    """Side measurement, config 1: N = 1332 uniform in the unit ball, virial ratio 0.5, G = 1,
    m = 1/N, seed 1332; `steps` graph-replayed BIA steps of dt = 1e-3 in the plain (reference
    solvers) and guess modes, against the DIM RKF45 (abserr 1e-16, the reference's controller) to
    t = steps dt on the GPU: ms per step, the median and worst particle's position error, |dE/E0|
    from the invariants.  The reference's ellipse Newton fails on near-radial bound pairs of this
    system (from step 20), so the plain trajectory ends far from the DIM; with the learned starts
    (plain_learned, when network 1 is trained) the same solvers are run again."""
    import torch

    from bia_3d_gpu_amd import MODE_GUESS, MODE_PLAIN, MODE_PLAIN_LEARN, ic

    n, dt = 1332, 1e-3
    st = ic.uniform_ball_virial(n, seed=1332)
    eps = -2.0 / n
    E = lambda inv: inv["KE"] + inv["PE"]
    e0 = E(ctx.invariants(st, eps))
    y = np.ascontiguousarray(st.T).copy()
    h, init, acc = 0.0, 1, 0
    t0 = time.perf_counter()
    t_end, chunks = steps * dt, 20
    for k in range(chunks):  # the controller's 10000-step budget is per call; h carries over
        r = ctx.dim_rkf45(y, -1.0 / n, t_end * k / chunks, t_end * (k + 1) / chunks, h=h, init=init)
        y, h, init, acc = r["y"].reshape(-1, 6), r["h"], 0, acc + r["naccept"]
    dim_s = time.perf_counter() - t0
    dim = np.ascontiguousarray(y.T)
    out = {"n": n, "dt": dt, "steps": steps, "dim": {"s": dim_s, "accepted_steps": acc,
                                                      "dE_rel": abs(E(ctx.invariants(dim, eps)) - e0) / abs(e0)}}
    runs = [("plain", MODE_PLAIN), ("guess", MODE_GUESS)] + ([("plain_learned", MODE_PLAIN_LEARN)] if plain_learned else [])
    for name, mode in runs:
        s = torch.from_numpy(st).to("cuda:0")
        d = torch.empty_like(s)
        ctx.run_steps_(s, d, [n], 1, eps=eps, dt=dt, mode=mode, use_graph=True)  # capture
        s.copy_(torch.from_numpy(st))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run_steps_(s, d, [n], steps, eps=eps, dt=dt, mode=mode, use_graph=True)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        x = s.cpu().numpy()
        dp = np.linalg.norm(x[:3] - dim[:3], axis=0)
        out[name] = {"ms_per_step": sec * 1e3, "pairs_per_s": n * (n - 1) / 2 / sec,
                     "position_error_median": float(np.median(dp)), "position_error_max": float(dp.max()),
                     "dE_rel": abs(E(ctx.invariants(x, eps)) - e0) / abs(e0)}
    return out


def run_energy_order(ctx, st, eps, dt: float, learned: bool, nsteps: int = 4):
    """Side measurement: |dE/E0| of the benchmark system over the fixed time T = nsteps dt with
    steps dt, dt/2 and dt/4 (guess and learned modes) and with the per-pair RKF45 at dt, from the
    library's double-double invariants (KE + PE = 2E/m); the observed order of the energy error,
    log2(e(h) / e(h/2))."""
    import torch

    from bia_3d_gpu_amd import METHOD_RKF, MODE_GUESS, MODE_LEARN, MODE_PLAIN

    n = st.shape[1]
    inv = torch.zeros(11, dtype=torch.float64, device="cuda:0")
    s0 = torch.from_numpy(st).to("cuda:0")
    ctx.invariants_(s0, inv, eps)
    e0 = float(inv[0] + inv[1])
    out = {"T": nsteps * dt}
    runs = [("guess", 1, MODE_GUESS)] + ([("learned", 1, MODE_LEARN)] if learned else [])
    for name, method, mode in runs:
        errs = []
        for k in (1, 2, 4):
            s = s0.clone()
            d = torch.empty_like(s)
            ctx.run_steps_(s, d, [n], nsteps * k, eps=eps, dt=dt / k, mode=mode, use_graph=True)
            ctx.invariants_(s, inv, eps)
            errs.append(abs(float(inv[0] + inv[1]) - e0) / abs(e0))
        out[name] = {"dE_rel": dict(zip(("dt", "dt/2", "dt/4"), errs)),
                     "order": [float(np.log2(errs[0] / errs[1])), float(np.log2(errs[1] / errs[2]))]}
    s = s0.clone()
    d = torch.empty_like(s)
    ctx.run_steps_(s, d, [n], nsteps, eps=eps, dt=dt, method=METHOD_RKF, mode=MODE_PLAIN, use_graph=False)
    ctx.invariants_(s, inv, eps)
    out["rkf45"] = {"dE_rel": {"dt": abs(float(inv[0] + inv[1]) - e0) / abs(e0)}}
    return out


def run_learned_gain(ctx, st, eps, dts=(1e-2, 1e-1), steps: int = 5):
    """Side measurement: the guess and learned modes on the benchmark system at longer steps,
    where more pairs are strong over dt and the learned table is consulted: pair-kernel ms,
    Kepler-equation evaluations per pair and learned starts."""
    import torch

    from bia_3d_gpu_amd import MODE_GUESS, MODE_LEARN

    n = st.shape[1]
    s = torch.from_numpy(st).to("cuda:0")
    d = torch.empty_like(s)
    c = torch.zeros(16, dtype=torch.int64, device="cuda:0")
    out = {}
    for dt in dts:
        row = {}
        for name, mode in (("guess", MODE_GUESS), ("learned", MODE_LEARN)):
            ctx.autotune_(s, d, [n], eps=eps, dt=dt, mode=mode)
            c.zero_()
            ctx.step_(s, d, eps, dt, mode=mode, stats=c)
            torch.cuda.synchronize()
            cs = c.cpu().numpy()
            ctx.set_profiling(True)
            for _ in range(steps):
                ctx.step_(s, d, eps, dt, mode=mode)
            kms = step_kernel_ms(ctx)[0]
            ctx.set_profiling(False)
            row[name] = {"kernel_ms": kms, "evaluations_per_pair": float(cs[4:8].sum() / cs[0:4].sum()),
                         "learned_starts": int(cs[14])}
        row["learned_vs_guess_kernel"] = row["learned"]["kernel_ms"] / row["guess"]["kernel_ms"]
        out[f"dt={dt:g}"] = row
    return out


def measured_traffic(mode: str, n: int):
    """HBM bytes per pair-kernel launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (profiles/hbm_traffic_plummer65536.txt), for the N=65536 workload only."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plummer65536.txt")
    if n != 65536 or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and not line.startswith("#") and f[0] == mode:
            return int(f[1]) + int(f[2])
    return None


def _cpu_rows(args):
    """One CPU worker: the reference's Fortran pair loop (method 1 ALG, 0 RKF) on the given row
    ranges, until budget_s."""
    st_host, eps, dt, itr, ranges, budget_s = args[:6]
    method = args[6] if len(args) > 6 else 1
    import oracle as O

    n = st_host.shape[1]
    kind = "reference" if O.ref_available() else "port"
    pairs, t = 0, 0.0
    for r0, r1 in ranges:
        if t >= budget_s:
            break
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_cpu_rows(st_host, eps, dt, method, itr, rows=(r0, r1))
        else:
            O.bia_step(st_host, eps, dt, method, itr, rows=(r0, r1))
        t += time.perf_counter() - t0
        pairs += sum(n - 1 - i for i in range(r0, r1))
    return pairs, t, kind


def _host_info():
    """CPU model, and the compiler and flags of the reference CPU build (oracle/Makefile)."""
    import subprocess
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        comp = subprocess.run(["amdflang", "--version"], capture_output=True, text=True, timeout=30).stdout.splitlines()[0]
    except Exception:
        comp = "amdflang"
    flags = None
    for line in open(os.path.join(ROOT, "oracle", "Makefile")):
        if line.startswith("FFLAGS"):
            flags = line.split(":=", 1)[1].strip()
    return model, comp, flags


def cpu_baseline(st_host, eps, dt, itr, budget_s: float):
    """Reference Fortran CPU pair loop (bia.f90:142-196 with the reference's find_* routines) on
    row chunks of the same system: one thread for budget_s / 2 seconds, then one process per host
    core (at most 16) on interleaved chunks for another budget_s / 2.  Run before the GPU is
    initialised (the workers are forked)."""
    import multiprocessing as mp

    n = st_host.shape[1]
    chunk = max(1, int(2e6 / n))
    ranges = [(r, min(n, r + chunk)) for r in range(0, n, chunk)]
    p1, t1, kind = _cpu_rows((st_host, eps, dt, itr, ranges, budget_s / 2))
    cores = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)))
    work = [(st_host, eps, dt, itr, ranges[k::cores], budget_s / 2) for k in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        out = pool.map(_cpu_rows, work)
    wall = time.perf_counter() - t0
    pairs = sum(o[0] for o in out)
    # the reference's CPU RKF path (bia.f90:147-154, find_dXYUV_RKF.f90) on one core, rows of two
    rk_ranges = [(r, min(n, r + 2)) for r in range(0, n, max(2, n // 64))]
    pr, tr, _ = _cpu_rows((st_host, eps, dt, itr, rk_ranges, budget_s / 4, 0))
    what = "amdflang -O3 build of the reference Fortran CPU routines" if kind == "reference" else "C oracle"
    model, comp, flags = _host_info()
    return dict(value=pairs / wall, unit="pair-interactions/s", cores=cores, kind=kind,
                single_core_value=p1 / t1, rkf_single_core_value=pr / tr, cpu_model=model, compiler=comp,
                flags=flags,
                sample=f"row chunks of the same N={n} system, one step, {what}: ALG {p1} pairs on one core in "
                       f"{t1:.1f} s; ALG {pairs} pairs on {cores} processes in {wall:.1f} s wall; RKF {pr} pairs "
                       f"on one core in {tr:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=65536)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--itr", type=int, default=18)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learn", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble side measurements")
    ap.add_argument("--no-fixture", action="store_true", help="skip the N=1332 reference-fixture side measurement")
    ap.add_argument("--modes", default="plain,guess,learned,plain_learned",
                    help="comma-separated subset of plain,guess,learned,plain_learned (profiling runs)")
    ap.add_argument("--copy-mb", type=float, default=0.0,
                    help="also run one device copy of this many MB (HBM counter calibration)")
    ap.add_argument("--workspace-mb", type=float, default=None,
                    help="tile-partial workspace limit of a step in MB (0 = none; default: the library's 1 GiB)")
    ap.add_argument("--train-samples", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=3000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the pair hot path of one system per GPU); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from bia_3d_gpu_amd import METHOD_RKF, Context, MODE_GUESS, MODE_LEARN, MODE_PLAIN, MODE_PLAIN_LEARN, learn

    n = args.n
    eps = -2.0 / n
    log(f"building state N={n} seed={args.seed}")
    st = make_state(n, args.seed)
    pairs = n * (n - 1) / 2
    # CPU baseline first: its worker processes are forked before this process touches the GPU
    cpu = None if args.no_cpu else cpu_baseline(st, eps, args.dt, args.itr, args.cpu_budget)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s on {cpu['cores']} cores "
            f"({cpu['single_core_value']:.3e} on one; RKF {cpu['rkf_single_core_value']:.3e} on one; "
            f"{cpu['cpu_model']}, {cpu['compiler']} {cpu['flags']})")
    assert torch.cuda.is_available()
    ctx = Context(0)
    if args.workspace_mb is not None:
        ctx.set_workspace_limit(int(args.workspace_mb * 1e6))
    modes = set(args.modes.split(","))
    if args.copy_mb > 0:  # a plain copy of known size for scaling HBM byte counters
        nb = int(args.copy_mb * 1e6) // 8
        a = torch.empty(nb, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb * 8} bytes read, {nb * 8} bytes written")
        del a, b
    res = {}
    for name, mode in (("plain", MODE_PLAIN), ("guess", MODE_GUESS)):
        if name not in modes:
            continue
        r = run_mode(ctx, st, eps, args.dt, args.itr, mode, args.steps, args.warmup, not args.no_autotune)
        r["pairs_per_s"] = pairs / (r["ms_per_step"] / 1e3)
        res[name] = r
        log(f"{name}: {r['ms_per_step']:.2f} ms/step, kernel {r['kernel_ms']:.2f} ms ({r['tflops']:.1f} TFLOP/s), "
            f"|dE/E0| {r['dE_rel']:.2e} over {r['dE_steps']} steps, tuned {r['tuned']}")
    train = None
    if not args.no_learn and "learned" in modes:
        t0 = time.perf_counter()
        w, curve, rows = learn.train_default(ctx, nsamples=args.train_samples, steps=args.train_steps, log=log)
        torch.cuda.synchronize()
        train = {"seconds": time.perf_counter() - t0, "samples": int(rows.shape[0]), "steps": args.train_steps,
                 "loss_first": float(curve[0]), "loss_last": float(np.mean(curve[-20:]))}
        del rows
        r = run_mode(ctx, st, eps, args.dt, args.itr, MODE_LEARN, args.steps, args.warmup, not args.no_autotune)
        r["pairs_per_s"] = pairs / (r["ms_per_step"] / 1e3)
        res["learned"] = r
        log(f"learned: {r['ms_per_step']:.2f} ms/step, kernel {r['kernel_ms']:.2f} ms ({r['tflops']:.1f} TFLOP/s), "
            f"|dE/E0| {r['dE_rel']:.2e}, tuned {r['tuned']}")
    train_ref = None
    if not args.no_learn and "plain_learned" in modes:
        # network 1: the starting anomaly of the reference's own Newton iterations
        t0 = time.perf_counter()
        w, curve, rows = learn.train_reference(ctx, nsamples=args.train_samples, steps=args.train_steps, log=log)
        torch.cuda.synchronize()
        train_ref = {"seconds": time.perf_counter() - t0, "samples": int(rows.shape[0]), "steps": args.train_steps,
                     "loss_first": float(curve[0]), "loss_last": float(np.mean(curve[-20:]))}
        del rows
        r = run_mode(ctx, st, eps, args.dt, args.itr, MODE_PLAIN_LEARN, args.steps, args.warmup, not args.no_autotune)
        r["pairs_per_s"] = pairs / (r["ms_per_step"] / 1e3)
        res["plain_learned"] = r
        sr = np.asarray(r["stats"], dtype=np.float64)
        log(f"plain_learned: {r['ms_per_step']:.2f} ms/step, kernel {r['kernel_ms']:.2f} ms "
            f"({r['tflops']:.1f} TFLOP/s), evaluations/pair {sr[4:8].sum() / sr[0:4].sum():.3f}, "
            f"fallbacks {int(sr[15])}, tuned {r['tuned']}")
    ensemble = None
    if not args.no_ensemble:
        emode = MODE_LEARN if "learned" in res else MODE_GUESS
        ensemble = run_ensemble(ctx, emode, max(args.steps, 3), args.warmup)
        ensemble["mode"] = "learned" if emode == MODE_LEARN else "guess"
        log(f"ensemble (512 systems, {ensemble['particles']} particles): {ensemble['ms_per_step']:.2f} ms/step, "
            f"{ensemble['pairs_per_s']:.3e} pairs/s; monitored driver {ensemble['monitored']['ms_per_step']:.2f} "
            f"ms/step ({ensemble['monitored']['ms_per_step_without_statistics']:.2f} without statistics), "
            f"max |dE/E0| {ensemble['monitored']['max_abs_energy_error']:.2e}")
    mass_ensemble = None
    if not args.no_ensemble:
        emode = MODE_LEARN if "learned" in res else MODE_GUESS
        mass_ensemble = run_mass_ensemble(ctx, emode, max(args.steps, 3), args.warmup)
        mass_ensemble["mode"] = "learned" if emode == MODE_LEARN else "guess"
        log(f"mass-spectrum ensemble (64 systems, {mass_ensemble['particles']} particles, gravity coupling): "
            f"{mass_ensemble['ms_per_step']:.2f} ms/step, {mass_ensemble['pairs_per_s']:.3e} pairs/s")
    fixture = None if args.no_fixture else run_fixture(ctx, 20, "learned" in res, "plain_learned" in res)
    if fixture:
        log(f"reference fixture N=1332: plain {fixture['plain']['s_per_step'] * 1e3:.3f} ms/step, guess "
            f"{fixture['guess']['s_per_step'] * 1e3:.3f} ms/step, RKF {fixture['rkf']['s_per_step'] * 1e3:.1f} "
            f"ms/step, DIM {fixture['dim']['s']:.3f} s ({fixture['dim']['accepted_steps']} steps)")
    config1 = None if args.no_fixture else run_config1(ctx, plain_learned="plain_learned" in res)
    if config1:
        log(f"config 1 (N=1332 ball, 1000 steps of 1e-3 vs DIM, {config1['dim']['accepted_steps']} DIM steps in "
            f"{config1['dim']['s']:.1f} s): plain {config1['plain']['ms_per_step']:.3f} ms/step, median error "
            f"{config1['plain']['position_error_median']:.2e}, |dE/E0| {config1['plain']['dE_rel']:.2e}; guess "
            f"{config1['guess']['ms_per_step']:.3f} ms/step, median error {config1['guess']['position_error_median']:.2e}, "
            f"|dE/E0| {config1['guess']['dE_rel']:.2e}")
        if "plain_learned" in config1:
            c1 = config1["plain_learned"]
            log(f"config 1 with learned starts: {c1['ms_per_step']:.3f} ms/step, median error "
                f"{c1['position_error_median']:.2e}, max {c1['position_error_max']:.2e}, |dE/E0| {c1['dE_rel']:.2e}")
    energy_order = None if args.no_ensemble else run_energy_order(ctx, st, eps, args.dt, "learned" in res)
    if energy_order:
        log(f"energy over T={energy_order['T']}: " + "; ".join(
            f"{k} {v['dE_rel']}" + (f" order {v['order']}" if 'order' in v else "") for k, v in energy_order.items()
            if k != "T"))
    learned_gain = run_learned_gain(ctx, st, eps) if "learned" in res and not args.no_ensemble else None
    if learned_gain:
        for k, v in learned_gain.items():
            log(f"learned vs guess at {k}: kernel {v['learned']['kernel_ms']:.2f} vs {v['guess']['kernel_ms']:.2f} ms, "
                f"evaluations/pair {v['learned']['evaluations_per_pair']:.5f} vs {v['guess']['evaluations_per_pair']:.5f}, "
                f"{v['learned']['learned_starts']} learned starts")
    plain = res.get("plain")
    # every mode's first step against an independent solution of the same pairs: the per-pair
    # RKF45 (the reference's FEHL integrator and step control), one untimed step at full size
    s0 = torch.from_numpy(st).to("cuda:0")
    d0 = torch.empty_like(s0)
    ctx.step_(s0, d0, eps, args.dt, method=METHOD_RKF, itr=args.itr)
    torch.cuda.synchronize()
    rkf_delta = d0.cpu().numpy()
    del s0, d0
    parity = {k: agreement(st, v["first_delta"], rkf_delta) for k, v in res.items()}
    for k, v in parity.items():
        log(f"{k} vs per-pair RKF45, first step: normwise positions {v['positions']:.2e}, velocities "
            f"{v['velocities']:.2e}; particles off by more than 1e-6 (1e-8) of the largest delta: "
            f"{v['particles_off_1e-6']} ({v['particles_off_1e-8']})")
    head_name = "learned" if "learned" in res else "guess" if "guess" in res else "plain"
    head = res[head_name]
    # roofline of the dominant kernel (the pair kernel of the reported mode), its own algorithm's flops
    achieved = head["tflops"]
    roof = {"bound": "fp64-valu", "pipe": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(head_name, n),
            "flops_per_launch": head["flops"], "kernel_ms": head["kernel_ms"], "kernel": f"pair kernel, {head_name} mode"}
    if plain is not None and head_name != "plain":
        # useful-work normalisation: the reference algorithm's flops for the same pairs (the plain
        # mode's count from its own iteration statistics) over this mode's kernel time; a rate of
        # reference-equivalent work, not of executed arithmetic, so it may exceed the FP64 peak
        roof["useful_work"] = {"flops_per_launch": plain["flops"],
                               "tflops": plain["flops"] / (head["kernel_ms"] * 1e-3) / 1e12,
                               "flops_per_pair": plain["flops"] / (n * (n - 1) / 2)}
    out = {
        "metric": METRIC, "value": head["pairs_per_s"], "unit": "pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Plummer N={n} with 20% hard binaries, G=1, m=1/N, dt={args.dt}, one BIA-ALG "
                               f"step per timed step, seed {args.seed}",
                   "n": n, "dt": args.dt, "eps": eps, "itr": args.itr,
                   "mode": head_name},
        "roofline": roof,
        "cpu_baseline": cpu,
        "modes": {k: {q: v[q] for q in ("pairs_per_s", "ms_per_step", "kernel_ms", "tflops", "dE_rel", "dE_steps",
                                         "stats", "tuned", "reduction", "invariants")} for k, v in res.items()},
        "learned_vs_guess": None if not {"learned", "guess"} <= set(res) else {
            "ms_per_step_ratio": res["learned"]["ms_per_step"] / res["guess"]["ms_per_step"],
            "kernel_ms_ratio": res["learned"]["kernel_ms"] / res["guess"]["kernel_ms"],
            "learned_starts": int(res["learned"]["stats"][14]),
            "longer_steps": learned_gain,
            "reference_fixture": None if not fixture else fixture.get("learned_vs_guess")},
        "reductions": {"tile_partials": head["reduction"], "invariants": head["invariants"]},
        "plain_learned_vs_plain": None if not {"plain_learned", "plain"} <= set(res) else {
            "ms_per_step_ratio": res["plain_learned"]["ms_per_step"] / res["plain"]["ms_per_step"],
            "kernel_ms_ratio": res["plain_learned"]["kernel_ms"] / res["plain"]["kernel_ms"],
            "evaluations_per_pair": {k: float(np.sum(res[k]["stats"][4:8]) / np.sum(res[k]["stats"][0:4]))
                                     for k in ("plain", "plain_learned")},
            "iteration_caps": {k: int(res[k]["stats"][10]) for k in ("plain", "plain_learned")},
            "learned_starts": int(res["plain_learned"]["stats"][14]),
            "fallbacks": int(res["plain_learned"]["stats"][15]),
            "reference_fixture": None if not fixture else fixture.get("plain_learned_vs_plain")},
        "config1": config1,
        "energy_order": energy_order,
        "speedup_vs_plain": None if plain is None else {k: v["pairs_per_s"] / plain["pairs_per_s"]
                                                        for k, v in res.items()},
        "agreement_vs_rkf45": parity,
        "training": train,
        "training_reference_starts": train_ref,
        "ensemble": ensemble,
        "mass_spectrum_ensemble": mass_ensemble,
        "reference_fixture": fixture,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
