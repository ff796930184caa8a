"""The benchmark configurations at their full size, and one system four times the benchmark's N,
checked through properties that do not need the oracle at that size: agreement of independent
solvers pair by pair, conservation of momentum, invariance under reordering, determinism of the
multi-step driver, the invariants it monitors, and the oracle on a few particles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _coulomb_energy(s, q):
    """KE + PE of unit masses with charges q (k = 1), direct sum on the device in float64."""
    import torch
    x = s[:3].T.contiguous()
    ke = 0.5 * (s[3:] * s[3:]).sum().item()
    pe, n = 0.0, x.shape[0]
    for a in range(0, n, 2048):
        r = torch.cdist(x[a:a + 2048], x)
        idx = torch.arange(a, min(a + 2048, n), device=x.device)
        r[idx - a, idx] = float("inf")
        pe += 0.5 * ((q[a:a + 2048, None] * q[None, :]) / r).sum().item()
    return ke + pe


def test_coulomb_box_16384(ctx):
    """Charges +-1 (n/2 each), unit masses, k = 1, uniform in the unit cube, velocities with about
    40 % of pairs bound, dt = 5e-4, N = 16384 (134 million pairs).  One step: the Kepler drift and
    the per-pair RKF45 (the reference's FEHL integrator and step control) agree on every particle
    except the two of each pair the RKF45 reports as failed; the total momentum change is
    rounding.  Then 500 Kepler-drift steps (hipGraph): momentum is kept to rounding."""
    import torch
    from bia_3d_gpu_amd import COUPLING_COULOMB, METHOD_ALG, METHOD_RKF, MODE_GUESS, MODE_PLAIN, STATS_LEN, ic
    n, dt = 16384, 5e-4
    st, q = ic.coulomb_box(n, seed=16384)
    s = torch.from_numpy(st).cuda()
    qc = torch.from_numpy(q).cuda()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    out, stats = {}, {}
    for name, method, mode in (("kepler", METHOD_ALG, MODE_GUESS), ("plain", METHOD_ALG, MODE_PLAIN),
                               ("rkf", METHOD_RKF, MODE_PLAIN)):
        d = torch.empty_like(s)
        c = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        ctx.step_batched_(s, d, [n], dt=dt, method=method, mode=mode, coupling=COUPLING_COULOMB, coup=1.0,
                          charge=qc, stats=c)
        torch.cuda.synchronize()
        out[name], stats[name] = d.cpu().numpy(), c.cpu().numpy()
    npairs = n * (n - 1) // 2
    for name in out:
        assert stats[name][0:4].sum() == npairs
        dv = out[name][3:]
        print(f"{name}: branches {stats[name][:4]}, flags {stats[name][8:14]}, "
              f"|sum dv| {np.abs(dv.sum(1)).max():.2e} of sum|dv| {np.abs(dv).sum(1).max():.3e}")
        assert np.abs(dv.sum(1)).max() <= 1e-14 * np.abs(dv).sum(1).max()
    k, r = out["kepler"], out["rkf"]
    scale = np.abs(r).max(axis=1, keepdims=True)
    differs = (np.abs(k - r) > 1e-9 * scale).any(axis=0)
    rkf_failed = int(stats["rkf"][12])  # status bit 16: RKF failure
    ok = ~differs
    ep = np.abs(k[:3, ok] - r[:3, ok]).max() / np.abs(st[:3]).max()
    print(f"Kepler drift vs per-pair RKF45: {differs.sum()} particles differ (RKF45 failed on {rkf_failed} "
          f"pairs); elsewhere max |delta position difference| / max |x| = {ep:.2e}; the reference solvers "
          f"hit their iteration cap on {stats['plain'][10]} pairs")
    assert stats["kepler"][10] == 0 and stats["kepler"][13] == 0
    assert differs.sum() <= 2 * rkf_failed
    assert ep < 1e-12
    d = torch.empty_like(s)
    p0 = s[3:].sum(1)
    ctx.run_steps_(s, d, [n], 500, dt=dt, mode=MODE_GUESS, coupling=COUPLING_COULOMB, coup=1.0, charge=qc)
    torch.cuda.synchronize()
    dp = (s[3:].sum(1) - p0).abs().max().item()
    assert torch.isfinite(s).all()
    print(f"500 steps: momentum change {dp:.2e} (sum |v0| {np.abs(st[3:]).sum():.3e}); energy "
          f"{_coulomb_energy(torch.from_numpy(st).cuda(), qc):.6e} -> {_coulomb_energy(s, qc):.6e}")
    assert dp <= 1e-16 * np.abs(st[3:]).sum() * 500


def test_plummer_ensemble_512_for_200_steps(ctx):
    """512 independent Plummer systems (a = 1, M = 1, G = 1, each from its own seed), N log-uniform
    in [64, 4096], dt = 1e-3, 200 steps through the monitored driver (hipGraph replays, statistics
    and the invariants of every system every 50 steps): the final state is bitwise that of 200
    plain graph replays, momentum is kept to rounding in every system, and the energy error of
    each system is the method's (no solver flags, one evaluation per pair)."""
    import torch
    from bia_3d_gpu_amd import MODE_GUESS, driver, ic
    sizes = ic.ensemble_sizes(512, 64, 4096, 512).astype(int)
    st = np.ascontiguousarray(np.concatenate([ic.plummer_system(int(m), 513 + k) for k, m in enumerate(sizes)],
                                             axis=1))
    eps = [-2.0 / m for m in sizes]
    a = torch.from_numpy(st).cuda()
    rep = driver.run(ctx, a, sizes, 200, eps=eps, dt=1e-3, mode=MODE_GUESS, monitor_every=50)
    b = torch.from_numpy(st).cuda()
    d = torch.empty_like(b)
    ctx.run_steps_(b, d, sizes, 200, eps=eps, dt=1e-3, mode=MODE_GUESS)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    de = np.abs(rep.energy_error())
    h = rep.histogram()
    print(f"200 steps of {sizes.sum()} particles in 512 systems: {rep.seconds:.2f} s; |dE/E0| median "
          f"{np.median(de[-1]):.2e}, max {de.max():.2e}; momentum {np.abs(rep.momentum_error()).max():.1e}; "
          f"{h['evaluations_per_pair']:.7f} evaluations per pair")
    assert rep.steps == [0, 50, 100, 150, 200]
    assert h["pairs"] == 200 * sum(int(m) * (int(m) - 1) // 2 for m in sizes)
    assert h["iteration cap"] == 0 and h["non-finite"] == 0
    assert np.abs(rep.momentum_error()).max() < 1e-12
    assert np.median(de[-1]) < 1e-7 and de.max() < 1e-4


def _closest_pairs(st, k):
    """The k closest pairs of a system (its hardest binaries), as index pairs."""
    from scipy.spatial import cKDTree
    dist, idx = cKDTree(st[:3].T).query(st[:3].T, k=2)
    order = np.argsort(dist[:, 1])
    out, seen = [], set()
    for i in order:
        j = int(idx[i, 1])
        if i not in seen and j not in seen:
            out.append((int(i), j))
            seen.update((int(i), j))
        if len(out) == k:
            break
    return out


def test_benchmark_system_65536_against_rkf45_and_quad(ctx):
    """The benchmark system itself (Plummer N = 65536 with 20 % of the particles in hard binaries,
    eps = -2/N, dt = 1e-3; 2.1e9 pairs), one step in every mode.

    1. Against the per-pair RKF45 of the same pairs (the reference's FEHL integrator and step
       control, an independent solution) on every particle: no particle's delta is off by more
       than 1e-6 of the largest delta of its component; positions agree normwise to 1e-11, the
       RKF45's own floor here (it advances the whole relative state and subtracts the straight-line
       motion, so each weak distant pair carries ~1.5e-15 of rounding, which 65535 pairs add up).
    2. Against the two-body solution in quadruple precision (oracle orc_rows_quad) on 16 whole
       rows: both members of the three hardest binaries, the particles where the Kepler drift and
       the reference solvers disagree most with RKF45, the first and the last particle and two
       more.  The Kepler drift, with and without the learned start, is within 1e-12 normwise in
       positions and velocities (max |x - x_quad| / max |x_quad| over the system, the
       normwise error of the accuracy bar) and within 1e-11 of each row's Kepler correction (the sum of its pair deltas,
       the part of the delta that is not drift).  The reference solvers (plain mode) are
       reported: they stop at their iteration cap on some binary pairs and are off there."""
    import torch
    from bia_3d_gpu_amd import METHOD_ALG, METHOD_RKF, MODE_GUESS, MODE_LEARN, MODE_PLAIN, STATS_LEN, ic, learn
    import oracle as O
    n = 65536
    st = ic.plummer_with_binaries(n, 65536)
    eps, dt = -2.0 / n, 1e-3
    s = torch.from_numpy(st).cuda()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    # a network for the learned mode, trained on the recorder's usual systems (small run)
    raw = learn.record(ctx, learn.training_systems(29), per_system=100_000, seed=29)
    rows_t = learn.balance(raw, 500_000, seed=29)
    ctx.mlp_init(29)
    ctx.mlp_train_(rows_t, steps=600, batch=8192, lr=3e-3, seed=29)
    ctx.mlp_build_table(learn.TABLE_N)
    out, stats = {}, {}
    for name, method, mode in (("guess", METHOD_ALG, MODE_GUESS), ("learned", METHOD_ALG, MODE_LEARN),
                               ("plain", METHOD_ALG, MODE_PLAIN), ("rkf", METHOD_RKF, MODE_PLAIN)):
        d = torch.empty_like(s)
        c = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        ctx.step_(s, d, eps, dt, method=method, mode=mode, stats=c)
        torch.cuda.synchronize()
        out[name], stats[name] = d.cpu().numpy(), c.cpu().numpy()
    ref = out["rkf"]
    assert stats["rkf"][12] == 0  # no RKF failure on this system
    scale = np.abs(ref).max(axis=1, keepdims=True)
    res, off = {}, {}
    for name in ("guess", "learned", "plain"):
        off[name] = (np.abs(out[name] - ref) / scale).max(axis=0)
        a, b = st + out[name], st + ref
        res[name] = (int((off[name] > 1e-6).sum()), int((off[name] > 1e-8).sum()),
                     np.abs(a[:3] - b[:3]).max() / np.abs(b[:3]).max(), np.abs(a[3:] - b[3:]).max() / np.abs(b[3:]).max())
        print(f"{name} vs RKF45: particles off by more than 1e-6 / 1e-8 of the largest delta {res[name][0]} / "
              f"{res[name][1]}; normwise positions {res[name][2]:.2e}, velocities {res[name][3]:.2e}; "
              f"iteration caps {stats[name][10]}, learned starts {stats[name][14]}")
    for name in ("guess", "learned"):
        assert res[name][0] == 0 and res[name][2] < 1e-11 and res[name][3] < 1e-12
        assert stats[name][10] == 0 and stats[name][13] == 0
    assert stats["learned"][14] > 0  # the learned start was used
    # 16 whole rows against the quad two-body solution
    rows = [p for pr in _closest_pairs(st, 3) for p in pr]
    for name in ("guess", "plain"):
        for p in np.argsort(-off[name]):
            if len(rows) >= (9 if name == "guess" else 12):
                break
            if int(p) not in rows:
                rows.append(int(p))
    rng = np.random.default_rng(16)
    for p in [0, n - 1] + [int(x) for x in rng.integers(1, n - 1, 8)]:
        if len(rows) < 16 and p not in rows:
            rows.append(p)
    chi, clo, dq = O.rows_quad(st, rows, eps, dt)
    xq = st[:, rows] + dq
    xmax = np.abs(st[:3] + ref[:3]).max()
    vmax = np.abs(st[3:] + ref[3:]).max()
    corr = np.abs(chi).max(axis=0)  # largest component of each row's pair sum
    quad = {}
    for name in ("guess", "learned", "plain", "rkf"):
        x = st[:, rows] + out[name][:, rows]
        ep = np.abs(x[:3] - xq[:3]).max() / xmax
        ev = np.abs(x[3:] - xq[3:]).max() / vmax
        dd = np.abs(out[name][:, rows] - dq)
        ec = (dd[3:].max(axis=0) / np.abs(chi[3:]).max(axis=0)).max()
        ecp = (dd[:3].max(axis=0) / np.abs(chi[:3]).max(axis=0)).max()
        quad[name] = (ep, ev, ecp, ec)
        print(f"{name} vs quad on {len(rows)} rows: normwise positions {ep:.2e}, velocities {ev:.2e}; "
              f"relative to each row's Kepler correction: positions {ecp:.2e}, velocities {ec:.2e}")
    for name in ("guess", "learned"):
        ep, ev, ecp, ec = quad[name]
        assert ep < 1e-12 and ev < 1e-12, (name, ep, ev)
        assert ecp < 1e-11 and ec < 1e-11, (name, ecp, ec)


def _dim_to(ctx, st, eps_dim, t_end, chunks):
    """The DIM RKF45 (abserr 1e-16, the reference's controller) from 0 to t_end in `chunks` calls
    (the controller's 10000-step budget is per call; h carries over); returns (state, accepted)."""
    y = np.ascontiguousarray(st.T).copy()
    h, init, acc = 0.0, 1, 0
    for k in range(chunks):
        r = ctx.dim_rkf45(y, eps_dim, t_end * k / chunks, t_end * (k + 1) / chunks, h=h, init=init)
        assert r["ier"] == 0, (k, r["ier"])
        y, h, init, acc = r["y"].reshape(-1, 6), r["h"], 0, acc + r["naccept"]
    return np.ascontiguousarray(y.T), acc


def test_config1_uniform_ball_1000_steps_against_dim(ctx):
    """Config 1: N = 1332 uniform in the unit ball, Gaussian velocities rescaled to virial ratio
    0.5, G = 1, m = 1/N (eps = -2/N for BIA, -1/N for the DIM, as bia.f90:37 and dim_3D.f90:9 pair
    them), seed 1332; 1000 graph-replayed BIA-ALG steps of dt = 1e-3 against the DIM RKF45 to
    t = 1, the reference's own BIA-vs-DIM report (bia.f90:217-278).
    * First step: the reference solvers (plain mode) against the oracle and against the
      reference's CPU ALG loop (amdflang build of bia.f90's routines), within the per-particle
      condition bound.
    * Kepler drift: at t = 1 the error of the median particle against the DIM falls by about 4
      when dt is halved (2000 steps of 5e-4; the BIA splitting error is third order per step,
      second order over a fixed time): the method's error, not a solver's.
    * Reference solvers: the device trajectory follows the oracle's (the reference algorithm on
      the CPU) for 19 steps.  Within a few tens of steps the reference's ellipse Newton stops at
      its cap on a near-radial bound pair (on the oracle's trajectory the pair (401, 1278) at step
      20: L = 0.004, eccentricity near 1, whether it converges at its 19th iteration or returns
      a change 1e5 times the true one is decided by 1e-11 of the state) and the step goes wrong
      on its particles: found as the first step at which the plain and Kepler-drift steps from
      the same state disagree, checked against the quad two-body rows, with the pair in the
      failure log.  Such failures recur, so the plain trajectory ends far from the DIM (reported,
      not asserted).
    * |dE/E0| of BIA and of the DIM at t = 1 from the library's invariants (KE + PE = 2E/m)."""
    import torch
    import oracle as O
    from bia_3d_gpu_amd import MODE_GUESS, MODE_PLAIN, STATS_LEN, FailureLog, ic
    n = 1332
    st = ic.uniform_ball_virial(n, seed=1332)
    eps, eps_dim, dt = -2.0 / n, -1.0 / n, 1e-3
    # first step
    d1 = ctx.step_host(st, eps, dt, mode=MODE_PLAIN)
    b = O.sensitivity(st, eps, dt)
    ok, rp, rv = O.within_condition_bound(st, d1, O.bia_step(st, eps, dt), eps, bounds=b)
    print(f"first step vs oracle: error / allowed {rp:.2f}, {rv:.2f}")
    assert ok
    if O.ref_available():
        ok, rp, rv = O.within_condition_bound(st, d1, O.ref_cpu_rows(st, eps, dt), eps, bounds=b)
        print(f"first step vs the reference CPU ALG: error / allowed {rp:.2f}, {rv:.2f}")
        assert ok
    # the reference solvers for 19 steps: the device and the oracle
    s = torch.from_numpy(st).cuda()
    d = torch.empty_like(s)
    ctx.run_steps_(s, d, [n], 19, eps=eps, dt=dt, mode=MODE_PLAIN, use_graph=True)
    torch.cuda.synchronize()
    dev19 = s.cpu().numpy()
    orc19 = st.copy()
    for _ in range(19):
        orc19 = orc19 + O.bia_step(orc19, eps, dt)
    e19 = np.abs(dev19 - orc19).max() / np.abs(orc19).max()
    print(f"19 plain steps, device vs oracle: normwise {e19:.2e}")
    assert e19 < 2e-9  # about 1e-11 per step from the reference formulas' conditioning (acos near 1)
    # along the plain trajectory, the first step at which the reference solvers go wrong: the plain
    # and guess steps from the same state, until some particle's deltas differ by more than 1e-3
    # of the largest delta of their component
    s = torch.from_numpy(st).cuda()
    dg = torch.empty_like(s)
    c = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
    found = None
    for k in range(300):
        log = FailureLog(1 << 12)
        ctx.set_failure_log(log)
        try:
            ctx.step_batched_(s, d, [n], eps=eps, dt=dt, mode=MODE_PLAIN, stats=c)
        finally:
            ctx.set_failure_log(None)
        ctx.step_(s, dg, eps, dt, mode=MODE_GUESS)
        torch.cuda.synchronize()
        a_, b_ = d.cpu().numpy(), dg.cpu().numpy()
        off = (np.abs(a_ - b_) / np.abs(b_).max(axis=1, keepdims=True)).max(axis=0)
        if off.max() > 1e-3:
            found = (k + 1, s.cpu().numpy(), a_, b_, int(np.argmax(off)), log.records())
            break
        s.add_(d)
    assert found is not None, "the reference solvers stayed right for 300 steps"
    k, sk, dp_, dg_, p, rec = found
    _, _, dq = O.rows_quad(sk, [p], eps, dt)
    corr = np.abs(dq[:, 0] - np.r_[sk[3:, p] * dt, 0, 0, 0]).max()
    eg = np.abs(dg_[:, p] - dq[:, 0]).max() / corr
    ep = np.abs(dp_[:, p] - dq[:, 0]).max() / corr
    mine = rec[(rec["i"] == p) | (rec["j"] == p)]
    print(f"step {k}: the reference solvers go wrong on particle {p} (off by {ep:.1e} of its Kepler correction; "
          f"the Kepler drift {eg:.1e}); {len(rec)} capped pairs logged, on this particle: {mine}")
    assert eg < 1e-11 and ep > 1e-3
    assert len(mine) >= 1 and (mine["flags"] & 4).all() and (mine["branch"] == 2).any()
    if O.ref_available():
        r = mine[0]
        rel = sk[:, r["i"]] - sk[:, r["j"]]
        exi, eeta, yz = O.frame(rel[:3], rel[3:])
        yr, _, _, _ = O.ref_solve2d("ellipse", eps, dt, yz)
        hq, _ = O.pair_quad(eps, dt, rel)
        r_true = rel[:3] + rel[3:] * dt + 2 * hq[:3]
        r_ref = yr[0] * exi + yr[1] * eeta
        print(f"pair {r['i']}-{r['j']}: the reference's CPU find_ellipse puts the separation at {r_ref.round(5)}, "
              f"the two-body solution at {r_true.round(5)}")
    # t = 1 by BIA at dt and dt/2, and by the DIM
    final = {}
    for mode, name in ((MODE_PLAIN, "plain"), (MODE_GUESS, "guess")):
        for k, h in ((1000, dt), (2000, dt / 2)):
            s = torch.from_numpy(st).cuda()
            d = torch.empty_like(s)
            ctx.run_steps_(s, d, [n], k, eps=eps, dt=h, mode=mode, use_graph=True)
            torch.cuda.synchronize()
            final[(name, k)] = s.cpu().numpy()
    dim, nacc = _dim_to(ctx, st, eps_dim, 1.0, 20)
    e0 = ctx.invariants(st, eps)
    ed = ctx.invariants(dim, eps)
    E = lambda inv: inv["KE"] + inv["PE"]
    print(f"DIM to t = 1: {nacc} accepted steps, |dE/E0| {abs(E(ed) - E(e0)) / abs(E(e0)):.2e}")
    assert abs(E(ed) - E(e0)) < 1e-12 * abs(E(e0))
    med = {}
    for (name, k), x in final.items():
        dp = np.linalg.norm(x[:3] - dim[:3], axis=0)
        dv = np.linalg.norm(x[3:] - dim[3:], axis=0)
        eb = ctx.invariants(x, eps)
        med[(name, k)] = np.median(dp)
        print(f"{name} BIA, {k} steps: position error median {np.median(dp):.2e}, max {dp.max():.2e}; velocity "
              f"median {np.median(dv):.2e}, max {dv.max():.2e}; |dE/E0| {abs(E(eb) - E(e0)) / abs(E(e0)):.2e}")
        assert np.isfinite(x).all()
    ratio = med[("guess", 1000)] / med[("guess", 2000)]
    print(f"Kepler drift: median error ratio dt / (dt/2) = {ratio:.2f}")
    assert 2.5 < ratio < 8, ratio
    assert med[("guess", 1000)] < 1e-6


_ENSEMBLE = {}


def _mass_spectrum_ensemble():
    """64 Plummer systems, N log-uniform in [4096, 32768], masses log-uniform in [0.1, 10]
    normalised to 1 per system: sizes, SoA state [6, sum N], masses (built once per session)."""
    if not _ENSEMBLE:
        from bia_3d_gpu_amd import ic
        sizes = [int(n) for n in ic.ensemble_sizes(64, 4096, 32768, seed=4096)]
        rng = np.random.default_rng(4097)
        states, masses = [], []
        for k, n in enumerate(sizes):
            states.append(ic.plummer_system(n, seed=5000 + k))
            m = np.exp(rng.uniform(np.log(0.1), np.log(10), n))
            masses.append(m / m.sum())
        _ENSEMBLE.update(sizes=sizes, st=np.ascontiguousarray(np.concatenate(states, axis=1)),
                         m=np.concatenate(masses))
    return _ENSEMBLE["sizes"], _ENSEMBLE["st"].copy(), _ENSEMBLE["m"]


def test_mass_spectrum_ensemble_64_systems(ctx):
    """64 Plummer systems, N log-uniform in [4096, 32768] (9.6e9 pairs per step), masses
    log-uniform in [0.1, 10] normalised to 1 per system, gravity coupling with the centre-of-mass
    split, dt = 5e-4, one launch per step.  The first step against the per-pair RKF45 of the same
    pairs (no particle off by more than 1e-6 of the largest delta), then three Kepler-drift steps:
    the mass-weighted momentum of every system is kept to rounding and no solver flag is raised."""
    import torch
    from bia_3d_gpu_amd import COUPLING_GRAVITY, METHOD_ALG, METHOD_RKF, MODE_GUESS, MODE_PLAIN, STATS_LEN
    sizes, st, m_all = _mass_spectrum_ensemble()
    s = torch.from_numpy(st).cuda()
    mc = torch.from_numpy(m_all).cuda()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    out, stats = {}, {}
    for name, method, mode in (("kepler", METHOD_ALG, MODE_GUESS), ("rkf", METHOD_RKF, MODE_PLAIN)):
        d = torch.empty_like(s)
        c = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        ctx.step_batched_(s, d, sizes, dt=5e-4, method=method, mode=mode, coupling=COUPLING_GRAVITY, coup=1.0,
                          mass=mc, stats=c)
        torch.cuda.synchronize()
        out[name], stats[name] = d.cpu().numpy(), c.cpu().numpy()
    npairs = sum(n * (n - 1) // 2 for n in sizes)
    assert stats["kepler"][0:4].sum() == npairs and stats["rkf"][12] == 0
    off = (np.abs(out["kepler"] - out["rkf"]) / np.abs(out["rkf"]).max(axis=1, keepdims=True)).max(axis=0)
    print(f"{len(sizes)} systems, {sum(sizes)} particles, {npairs:.3e} pairs: Kepler drift vs RKF45 particles off by "
          f"more than 1e-6 / 1e-8 of the largest delta {(off > 1e-6).sum()} / {(off > 1e-8).sum()}")
    assert (off > 1e-6).sum() == 0
    d = torch.empty_like(s)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    p0 = [(m_all[a:b] * st[3:, a:b]).sum(axis=1) for a, b in zip(offs[:-1], offs[1:])]
    c = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
    for _ in range(3):
        ctx.step_batched_(s, d, sizes, dt=5e-4, mode=MODE_GUESS, coupling=COUPLING_GRAVITY, coup=1.0, mass=mc,
                          stats=c)
        s.add_(d)
    torch.cuda.synchronize()
    sn = s.cpu().numpy()
    worst = 0.0
    for k, (a, b) in enumerate(zip(offs[:-1], offs[1:])):
        p1 = (m_all[a:b] * sn[3:, a:b]).sum(axis=1)
        scale = (m_all[a:b] * np.abs(st[3:, a:b])).sum()
        worst = max(worst, np.abs(p1 - p0[k]).max() / scale)
    c = c.cpu().numpy()
    print(f"3 steps: momentum change at most {worst:.1e} of sum m|v| per system; flags {c[8:14]}; "
          f"{c[4:8].sum() / c[0:4].sum():.7f} evaluations per pair")
    assert worst < 1e-14 and c[10] == 0 and c[13] == 0


def test_mass_spectrum_ensemble_against_dim(ctx):
    """The same 64 systems, one analytic step (Kepler drift, gravity coupling, dt = 5e-4, one
    launch) against the direct-sum DIM RKF45 with the same masses (abserr 1e-16), system by
    system, with the energy of every system before and after from the mass-weighted invariants
    (the DIM keeps it to 4e-16, the analytic step to 1.4e-10 at most; bounds 1e-14 and 1e-8).
    The difference is the BIA splitting error, third order in dt in the velocities: in
    every system the median particle is within 1e-8 of the system's largest delta, and the worst
    particle of all 64 systems (a close three-body encounter) within 1e-3.  On the system holding
    that particle, halving dt must shrink its velocity error as a third-order error does (by more
    than 5; 8 in the limit, 7.1 for this particle in a 31-body neighbourhood integrated by an
    independent 8th-order integrator).  The relative deviation itself shrinks by 4 only, since
    the deltas it is measured against are proportional to dt.  An equal split of a pair's velocity change in place of the
    centre-of-mass split would move particle i's share by (1/2 - m_j/(m_i + m_j)) of the pair's
    change, up to half of it for these masses."""
    import torch
    from bia_3d_gpu_amd import COUPLING_GRAVITY, MODE_GUESS
    sizes, st, m_all = _mass_spectrum_ensemble()
    offs = np.concatenate([[0], np.cumsum(sizes)])

    def bia(state, szs, m, dt):
        s = torch.from_numpy(np.ascontiguousarray(state)).cuda()
        d = torch.empty_like(s)
        ctx.step_batched_(s, d, szs, dt=dt, mode=MODE_GUESS, coupling=COUPLING_GRAVITY, coup=1.0,
                          mass=torch.from_numpy(np.ascontiguousarray(m)).cuda())
        torch.cuda.synchronize()
        return d.cpu().numpy()

    def dim(state, m, dt):
        r = ctx.dim_rkf45(state.T, 0.0, 0.0, dt, coupling=COUPLING_GRAVITY, coup=1.0, mass=m)
        assert r["ier"] == 0
        return r["y"].reshape(-1, 6).T - state

    def deviation(db, dd):
        return (np.abs(db - dd) / np.abs(dd).max(axis=1, keepdims=True)).max(axis=0)

    db = bia(st, sizes, m_all, 5e-4)
    worst, med = [], []
    for k, (a, b) in enumerate(zip(offs[:-1], offs[1:])):
        rel = deviation(db[:, a:b], dim(st[:, a:b], m_all[a:b], 5e-4))
        worst.append((rel.max(), k, int(rel.argmax())))
        med.append(np.median(rel))
    # energy: the DIM conserves it to its tolerance, the analytic step to its splitting error
    sys_ = np.concatenate([dim(st[:, a:b], m_all[a:b], 5e-4) for a, b in zip(offs[:-1], offs[1:])], axis=1)
    e = []
    for state in (st, st + sys_, st + db):
        out = torch.zeros(len(sizes), 11, dtype=torch.float64, device="cuda")
        ctx.invariants_coupled_(torch.from_numpy(np.ascontiguousarray(state)).cuda(), sizes, out, COUPLING_GRAVITY,
                                1.0, mass=torch.from_numpy(m_all).cuda())
        e.append(out[:, 0].cpu().numpy() + out[:, 1].cpu().numpy())
    de_dim, de_bia = np.abs(e[1] / e[0] - 1), np.abs(e[2] / e[0] - 1)
    print(f"|dE/E0| over the step: DIM at most {de_dim.max():.1e}, analytic step median {np.median(de_bia):.1e}, "
          f"at most {de_bia.max():.1e}")
    assert de_dim.max() < 1e-14 and de_bia.max() < 1e-8
    w, k, i = max(worst)
    a, b = offs[k], offs[k + 1]
    ev = [np.abs(db[3:, a + i] - dim(st[:, a:b], m_all[a:b], 5e-4)[3:, i]).max()]
    ev.append(np.abs(bia(st[:, a:b], [sizes[k]], m_all[a:b], 2.5e-4)[3:, i]
                     - dim(st[:, a:b], m_all[a:b], 2.5e-4)[3:, i]).max())
    print(f"{len(sizes)} systems: BIA vs DIM, median particle at most {max(med):.2e} of the system's largest "
          f"delta; worst particle {w:.2e} (system {k}, N = {sizes[k]}, particle {i}); its velocity error "
          f"{ev[0]:.3e} at dt, {ev[1]:.3e} at dt / 2 (ratio {ev[0] / ev[1]:.2f})")
    assert max(med) < 1e-8 and w < 1e-3 and ev[0] / ev[1] > 5


def test_system_262144_with_64_bit_partial_offsets():
    """The largest single system tested: Plummer N = 262144 (4096 tiles a side), eps = -2/N,
    dt = 1e-3, one Kepler-drift step (3.4e10 pairs), unstaged (workspace limit off: the partials
    hold 1.6e9 doubles, 12.9 GB, so their byte offsets pass 2^32) and under the default 1 GiB
    limit (24 stages of 43 blocks of four row tiles, a 1.06 GB stage buffer): the two bitwise
    equal.  Checked
    further by (1) the same system in a shuffled order: every particle's delta agrees with the
    unshuffled one's up to summation order; (2) six particles,
    the first and last ones among them, against the oracle's reference-algorithm row for that
    particle (the particle moved to the front, so that its row covers all its pairs): velocities
    to 1e-9 of the largest velocity delta, positions normwise to 1e-11 — the reference
    algorithm's floor: it advances the whole relative state and subtracts the straight-line
    motion, leaving up to ~1e-14 of rounding in each weak pair's position departure, ~1e-12 summed
    over a particle's 262143 pairs; (3) the total velocity change is rounding."""
    import torch
    import oracle as O
    from bia_3d_gpu_amd import MODE_GUESS, Context, ic, partial_doubles, staging
    n, dt = 262144, 1e-3
    eps = -2.0 / n
    st = ic.plummer_system(n, seed=n)
    perm = np.random.default_rng(7).permutation(n)
    with Context(0) as c:
        c.set_stream(torch.cuda.current_stream().cuda_stream)
        out, plans = [], []
        for order, limit in ((None, 0), (None, 1 << 30), (perm, 1 << 30)):
            c.set_workspace_limit(limit)
            s = torch.from_numpy(np.ascontiguousarray(st if order is None else st[:, order])).cuda()
            d = torch.empty_like(s)
            c.step_(s, d, eps, dt, mode=MODE_GUESS)
            torch.cuda.synchronize()
            out.append(d.cpu().numpy())
            plans.append(c.plan_info())
            del s, d
    d_unstaged, d, dp = out
    print(f"plans (stages, workspace bytes): unstaged {plans[0]}, staged {plans[1]}")
    assert plans[0] == (1, 8 * partial_doubles(n)) and plans[1] == staging(n, 1 << 30) and plans[1][0] == 24
    assert plans[1][1] <= 1 << 30
    assert np.array_equal(d_unstaged, d)
    pair_part = d.copy()
    pair_part[:3] -= st[3:] * dt  # the position deltas without the drift v dt
    scale = np.abs(pair_part).max(axis=1, keepdims=True)
    shuffled = dp.copy()
    shuffled[:3] -= st[3:, perm] * dt
    err_shuffle = (np.abs(shuffled - pair_part[:, perm]) / scale).max()
    assert err_shuffle < 1e-10, err_shuffle
    mom = np.abs(d[3:].sum(axis=1)).max() / scale[3:].max()
    assert mom < 1e-9, mom
    worst_v, worst_x = 0.0, 0.0
    xmax = np.abs(st[:3]).max()
    for p in (0, 1, n // 2, n - 65, n - 2, n - 1):
        order = np.concatenate([[p], np.delete(np.arange(n), p)])
        ref = O.bia_step(st[:, order], eps, dt, method=1, rows=(0, 1))[:, 0]
        worst_v = max(worst_v, float((np.abs(d[3:, p] - ref[3:]) / scale[3:, 0]).max()))
        worst_x = max(worst_x, float(np.abs(d[:3, p] - ref[:3]).max() / xmax))
    print(f"N={n}: shuffled order {err_shuffle:.2e}, momentum {mom:.2e}; six particles against the oracle: "
          f"velocities {worst_v:.2e} of the largest velocity delta, positions {worst_x:.2e} normwise")
    assert worst_v < 1e-9 and worst_x < 1e-11, (worst_v, worst_x)
