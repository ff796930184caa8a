"""Batched (ensemble) steps, per-particle couplings, the hipGraph multi-step driver and the
autotuner, through the C ABI on the GPU."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _cat(states):
    return np.ascontiguousarray(np.concatenate(states, axis=1))


@pytest.mark.parametrize("mode", [0, 2])
def test_batched_equals_single_system_calls(ctx, mode):
    import torch
    from bia_3d_gpu_amd import ic
    sizes = [100, 333, 1000]
    states = [ic.plummer_system(n, seed=50 + k) for k, n in enumerate(sizes)]
    eps = [-2.0 / n for n in sizes]
    s = torch.from_numpy(_cat(states)).cuda()
    d = torch.empty_like(s)
    ctx.step_batched_(s, d, sizes, eps=eps, dt=1e-3, mode=mode)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    off = 0
    for st, n, e in zip(states, sizes, eps):
        single = ctx.step_host(st, e, 1e-3, mode=mode)
        assert np.array_equal(got[:, off:off + n], single)  # same tiles, same reduction order
        off += n


def test_batched_segment_lengths_equal_single_system_calls(ctx):
    """One list mixing the three segment lengths of the pair kernel's work items (four column tiles
    for N = 32768, two for N = 24576, one for the small systems): each system's deltas bitwise
    those of its own single-system step, and the small systems within the condition bound of the
    oracle's reference-algorithm step."""
    import torch
    from bia_3d_gpu_amd import MODE_GUESS, ic, segment_length
    sizes = [130, 32768, 777, 24576, 64]
    assert [segment_length(n) for n in sizes] == [1, 4, 1, 2, 1]
    states = [ic.plummer_system(n, seed=90 + k) for k, n in enumerate(sizes)]
    eps = [-2.0 / n for n in sizes]
    s = torch.from_numpy(_cat(states)).cuda()
    d = torch.empty_like(s)
    ctx.step_batched_(s, d, sizes, eps=eps, dt=1e-3, mode=MODE_GUESS)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    off = 0
    for st, n, e in zip(states, sizes, eps):
        si = torch.from_numpy(st).cuda()
        di = torch.empty_like(si)
        ctx.step_(si, di, e, 1e-3, mode=MODE_GUESS)
        torch.cuda.synchronize()
        assert np.array_equal(got[:, off:off + n], di.cpu().numpy()), n
        if n <= 777:  # against the oracle's reference-algorithm step, within its condition bound
            ok, rp, rv = O.within_condition_bound(st, got[:, off:off + n], O.bia_step(st, e, 1e-3), e, dt=1e-3)
            assert ok, (n, rp, rv)
        off += n


def test_gravity_coupling_equal_masses_is_the_reference_step(ctx):
    import torch
    from bia_3d_gpu_amd import ic, COUPLING_GRAVITY
    n = 700
    st = ic.plummer_system(n, seed=9)
    s = torch.from_numpy(st).cuda()
    m = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda")
    d = torch.empty_like(s)
    ctx.step_batched_(s, d, [n], dt=1e-3, coupling=COUPLING_GRAVITY, coup=1.0, mass=m)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), ctx.step_host(st, -2.0 / n, 1e-3))


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("kind", ["gravity", "coulomb"])
def test_coupled_step_vs_oracle(ctx, kind, mode):
    import torch
    from bia_3d_gpu_amd import ic, COUPLING_GRAVITY, COUPLING_COULOMB
    n = 512
    if kind == "gravity":
        st = ic.plummer_system(n, seed=12)
        m = np.random.default_rng(2).uniform(0.1, 10, n)
        m /= m.sum()
        q, coupling, coup, dt = None, COUPLING_GRAVITY, 1.0, 1e-3
    else:
        st, q = ic.coulomb_box(n, seed=13)
        m, coupling, coup, dt = None, COUPLING_COULOMB, 1.0, 5e-4
    s = torch.from_numpy(st).cuda()
    d = torch.empty_like(s)
    tm = None if m is None else torch.from_numpy(m).cuda()
    tq = None if q is None else torch.from_numpy(q).cuda()
    stats = torch.zeros(16, dtype=torch.int64, device="cuda")
    ctx.step_batched_(s, d, [n], dt=dt, coupling=coupling, coup=coup, mass=tm, charge=tq, stats=stats, mode=mode)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    ref = O.bia_step_coupled(st, coupling, coup, dt, mass=m, charge=q)
    b = O.sensitivity(st, coup, dt, coupling=coupling, mass=m, charge=q)
    ok, rp, rv = O.within_condition_bound(st, got, ref, coup, bounds=b)
    sc = stats.cpu().numpy()
    skipped = O.indeterminate(b)
    print(kind, "mode", mode, "branches", sc[:4], "flags", sc[8:14], f"error/allowed {rp:.2f} {rv:.2f}",
          f"particles excluded (a pair of indeterminate type or one the reference gets wrong): {skipped}")
    assert ok
    assert skipped <= 3 * n // 100  # pairs the reference gets wrong or cannot determine are rare
    if kind == "coulomb":
        assert sc[0] > 0 and sc[2] > 0  # attractive bound pairs and repulsive / unbound pairs


@pytest.mark.parametrize("mode", [0, 2])
def test_graph_driver_matches_eager_steps(ctx, mode):
    import torch
    from bia_3d_gpu_amd import ic
    sizes = [300, 129]
    st = _cat([ic.plummer_system(n, seed=20 + k) for k, n in enumerate(sizes)])
    eps = [-2.0 / n for n in sizes]
    a = torch.from_numpy(st).cuda()
    b = a.clone()
    da, db = torch.empty_like(a), torch.empty_like(b)
    for k in range(4):
        ctx.run_steps_(a, da, sizes, 1, eps=eps, dt=1e-3, mode=mode, use_graph=True)  # replay of the graph
        ctx.step_batched_(b, db, sizes, eps=eps, dt=1e-3, mode=mode)                  # eager work in between
        b.add_(db)
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"replay {k}"
    ctx.run_steps_(a, da, sizes, 3, eps=eps, dt=1e-3, mode=mode, use_graph=True)
    ctx.run_steps_(b, db, sizes, 3, eps=eps, dt=1e-3, mode=mode, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_graph_replay_after_plan_rebuild(ctx):
    """Graphs on size list A, then B (more systems, fewer pairs: the plan's descriptor and chunk
    buffers are reallocated), then A again on the same tensors: the graph captured on A reads
    the plan buffers through pointers, so it must be captured anew, not replayed.  Checked
    bitwise against eager steps on another context."""
    import torch
    from bia_3d_gpu_amd import Context, ic
    size_lists = {"A": [1500], "B": [64] * 40}
    runs = {}
    for key, sizes in size_lists.items():
        st = _cat([ic.plummer_system(n, seed=70 + k) for k, n in enumerate(sizes)])
        g = torch.from_numpy(st).cuda()
        runs[key] = (sizes, [-2.0 / n for n in sizes], g, g.clone(), torch.empty_like(g), torch.empty_like(g))
    with Context(0) as c:
        for key in ("A", "B", "A", "B"):
            sizes, eps, g, e, dg, de = runs[key]
            c.run_steps_(g, dg, sizes, 2, eps=eps, dt=1e-3, mode=2, use_graph=True)
            ctx.run_steps_(e, de, sizes, 2, eps=eps, dt=1e-3, mode=2, use_graph=False)
            torch.cuda.synchronize()
            assert torch.equal(g, e), key


@pytest.mark.parametrize("mode", [0, 2])
def test_autotuner_keeps_results(ctx, mode):
    import torch
    from bia_3d_gpu_amd import ic
    n = 2048
    st = ic.plummer_system(n, seed=31)
    s = torch.from_numpy(st).cuda()
    d0, d1 = torch.empty_like(s), torch.empty_like(s)
    ctx.step_batched_(s, d0, [n], eps=[-2.0 / n], dt=1e-3, mode=mode)
    ms, grid, occ = ctx.autotune_(s, d1, [n], eps=[-2.0 / n], dt=1e-3, mode=mode)
    ctx.step_batched_(s, d1, [n], eps=[-2.0 / n], dt=1e-3, mode=mode)
    torch.cuda.synchronize()
    ntiles = (n // 64) * (n // 64 + 1) // 2
    print(f"autotuner: grid {grid} of {ntiles} tiles, occupancy variant {occ}, {ms:.3f} ms")
    assert 0 < grid <= ntiles and ms > 0 and occ in (1, 4, 5)
    assert torch.equal(d0, d1)  # grid size and register allocation change scheduling only


def test_batched_rkf(ctx):
    import torch
    from bia_3d_gpu_amd import ic
    sizes = [64, 96]
    states = [ic.uniform_ball_virial(n, seed=40 + k) for k, n in enumerate(sizes)]
    eps = [-2.0 / n for n in sizes]
    s = torch.from_numpy(_cat(states)).cuda()
    d = torch.empty_like(s)
    ctx.step_batched_(s, d, sizes, eps=eps, dt=1e-2, method=0)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    off = 0
    for st, n, e in zip(states, sizes, eps):
        ref = O.bia_step(st, e, 1e-2, method=0)
        assert O.within_condition_bound(st, got[:, off:off + n], ref, e, 1e-2, method=0)[0]
        off += n


@pytest.mark.parametrize("mode", [0, 2])
def test_mass_spectrum_ensemble_vs_oracle(ctx, mode):
    """Config-5 style ensemble at test size: several Plummer systems with log-uniform masses in
    [0.1, 10] (normalised to total mass 1 per system), gravity coupling eps_ij = -(m_i + m_j),
    centre-of-mass split, one batched launch; every system against the oracle's coupled step."""
    import torch
    from bia_3d_gpu_amd import ic, COUPLING_GRAVITY
    sizes = [int(n) for n in ic.ensemble_sizes(5, 64, 700, seed=64)]
    rng = np.random.default_rng(65)
    states, masses = [], []
    for k, n in enumerate(sizes):
        states.append(ic.plummer_system(n, seed=70 + k))
        m = np.exp(rng.uniform(np.log(0.1), np.log(10), n))
        masses.append(m / m.sum())
    st = _cat(states)
    m_all = np.concatenate(masses)
    s = torch.from_numpy(st).cuda()
    d = torch.empty_like(s)
    ctx.step_batched_(s, d, sizes, dt=5e-4, coupling=COUPLING_GRAVITY, coup=1.0,
                      mass=torch.from_numpy(m_all).cuda(), mode=mode)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    off = 0
    for st_k, m_k, n in zip(states, masses, sizes):
        ref = O.bia_step_coupled(st_k, COUPLING_GRAVITY, 1.0, 5e-4, mass=m_k)
        b = O.sensitivity(st_k, 1.0, 5e-4, coupling=COUPLING_GRAVITY, mass=m_k)
        ok, rp, rv = O.within_condition_bound(st_k, got[:, off:off + n], ref, 1.0, bounds=b)
        print(f"system n={n}: error/allowed {rp:.3f} {rv:.3f}; excluded {O.indeterminate(b)}")
        assert ok and O.indeterminate(b) <= 3 * n // 100
        off += n


@pytest.mark.parametrize("mode", [0, 2])
def test_monitored_driver_matches_plain_steps(ctx, mode, tmp_path):
    """The monitored multi-step driver (graph replays in chunks, invariants and statistics
    recorded on the device, .bbb checkpoints written from pinned memory) against an eager loop
    of single steps with the same monitoring done by hand."""
    import torch
    from bia_3d_gpu_amd import STATS_LEN, driver, ic, io
    sizes = [200, 333, 64]
    st = _cat([ic.plummer_system(n, seed=70 + k) for k, n in enumerate(sizes)])
    eps = np.array([-2.0 / n for n in sizes])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    nsteps, dt = 7, 1e-3
    a = torch.from_numpy(st).cuda()
    lines = []
    rep = driver.run(ctx, a, sizes, nsteps, eps=eps, dt=dt, mode=mode, monitor_every=2, checkpoint_every=3,
                     checkpoint_path=str(tmp_path / "ck_{step:04d}.bbb"), log=lines.append)
    assert rep.steps == [0, 2, 4, 6, 7] and len(lines) == 5
    assert rep.checkpoints == [str(tmp_path / "ck_0003.bbb"), str(tmp_path / "ck_0006.bbb")]

    b = torch.from_numpy(st).cuda()
    d = torch.empty_like(b)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    acc = np.zeros(STATS_LEN, dtype=np.int64)
    out = torch.empty(11, dtype=torch.float64, device="cuda")
    for k in range(nsteps + 1):
        if k in rep.steps:
            i = rep.steps.index(k)
            assert np.array_equal(rep.stats[i], acc), k
            for s, n in enumerate(sizes):
                ctx.invariants_at_(b, int(offs[s]), n, float(eps[s]), out)
                assert np.array_equal(rep.invariants[i, s], out.cpu().numpy()), (k, s)
        if k in (3, 6):
            ck = io.read_bbb(str(tmp_path / f"ck_{k:04d}.bbb"))
            assert [c.shape[1] for c in ck] == sizes
            assert np.array_equal(np.concatenate(ck, axis=1), b.cpu().numpy()), k
        if k == nsteps:
            break
        s_k = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        ctx.step_batched_(b, d, sizes, eps=eps, dt=dt, mode=mode, stats=s_k)
        b.add_(d)
        acc += s_k.cpu().numpy()
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # resuming from the step-6 checkpoint gives the same step 7, also without the statistics
    # (a different kernel variant, the same rounding)
    c, csz = driver.load_checkpoint(str(tmp_path / "ck_0006.bbb"))
    assert csz == sizes
    r2 = driver.run(ctx, c, csz, 1, eps=eps, dt=dt, mode=mode, collect_stats=False)
    assert r2.stats is None
    torch.cuda.synchronize()
    assert torch.equal(a, c)
    h = rep.histogram()
    npairs = sum(n * (n - 1) // 2 for n in sizes)
    assert h["pairs"] == nsteps * npairs and h["evaluations_per_pair"] > 0
    assert np.isfinite(rep.energy_error()).all() and np.abs(rep.energy_error()).max() < 1e-3
    print(lines[-1])


def test_kernel_timing_ring(ctx):
    """With profiling on, every step's pair kernel is timed by its own pair of HIP events and
    bia_kernel_ms returns the last 256 durations, oldest first, after steps enqueued with no
    synchronisation in between; the newest equals bia_last_kernel_ms; the reduction kernel of the
    same steps has a ring of its own."""
    import torch
    from bia_3d_gpu_amd import MODE_GUESS, ic
    n = 700
    s = torch.from_numpy(ic.plummer_system(n, seed=31)).cuda()
    d = torch.empty_like(s)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.set_profiling(True)
    try:
        for _ in range(3):
            ctx.step_(s, d, -2.0 / n, 1e-3, mode=MODE_GUESS)
        three = ctx.kernel_ms()
        assert len(three) == 3 and all(t > 0 for t in three)
        assert three[-1] == ctx.last_kernel_ms()
        for _ in range(267):
            ctx.step_(s, d, -2.0 / n, 1e-3, mode=MODE_GUESS)
        ring = ctx.kernel_ms()
        assert len(ring) == 256 and all(t > 0 for t in ring) and ring[-1] == ctx.last_kernel_ms()
        assert len(ctx.kernel_ms(5)) == 5
        red = ctx.reduce_kernel_ms()
        assert len(red) == 256 and all(t > 0 for t in red)
    finally:
        ctx.set_profiling(False)
