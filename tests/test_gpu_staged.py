"""Staged steps (bia_set_workspace_limit): a single system whose tile partials exceed the
context's workspace limit runs in stages of row tiles, each followed by its reduction; every
particle's slots are still added in ascending order, so the deltas are bitwise those of the
unstaged step, in every mode, with statistics, under graph replay and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _steps(c, st, eps, dt, method, mode, nsteps=1, graph=False):
    import torch
    from bia_3d_gpu_amd import STATS_LEN
    s = torch.from_numpy(np.ascontiguousarray(st)).cuda()
    d = torch.empty_like(s)
    stats = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
    n = st.shape[1]
    if nsteps == 1:
        c.step_batched_(s, d, [n], eps=eps, dt=dt, method=method, mode=mode, stats=stats)
    else:
        c.run_steps_(s, d, [n], nsteps, eps=eps, dt=dt, method=method, mode=mode, use_graph=graph, stats=stats)
    torch.cuda.synchronize()
    return d.cpu().numpy(), s.cpu().numpy(), stats.cpu().numpy()


@pytest.mark.parametrize("method,mode", [(1, 0), (1, 2), (0, 0)])
def test_staged_step_is_bitwise_the_unstaged_one(method, mode):
    from bia_3d_gpu_amd import Context, ic
    n = 4000 if method == 1 else 1500
    st = ic.plummer_with_binaries(n, 77)
    eps, dt = -2.0 / n, 1e-3
    nt = (n + 63) // 64
    full = 6 * 8 * (64 * nt) ** 2 // 64
    with Context(0) as c:
        c.set_workspace_limit(0)
        want, _, wstats = _steps(c, st, eps, dt, method, mode)
        assert c.plan_info() == (1, full)
        # stage sizes: one row tile, odd sizes, the largest below the unstaged size (a stage of G
        # rows holds 12 G n_pad partials against 6 nt n_pad unstaged), and one above it (unstaged)
        for rows in (1, 5, 17, (nt - 1) // 2, nt // 2 + 1):
            limit = 12 * 8 * rows * 64 * nt
            c.set_workspace_limit(limit)
            got, _, gstats = _steps(c, st, eps, dt, method, mode)
            ns, wb = c.plan_info()
            print(f"method {method} mode {mode}: {rows} row tiles per stage -> {ns} stages, {wb / 1e6:.1f} MB "
                  f"(unstaged {full / 1e6:.1f} MB)")
            assert (ns, wb) == ((1, full) if limit >= full else (-(-nt // rows), limit))
            assert np.array_equal(got, want) and np.array_equal(gstats, wstats), rows


def test_staged_graph_replay_and_autotuner():
    """Three graph-replayed steps with stages equal three eager unstaged steps; the autotuner's
    grids (capped to each stage) leave the staged result unchanged."""
    from bia_3d_gpu_amd import Context, MODE_GUESS, ic
    import torch
    n = 6000
    st = ic.plummer_with_binaries(n, 78)
    eps, dt = -2.0 / n, 1e-3
    nt = (n + 63) // 64
    with Context(0) as c:
        c.set_workspace_limit(0)
        _, want, _ = _steps(c, st, eps, dt, 1, MODE_GUESS, nsteps=3, graph=False)
        c.set_workspace_limit(12 * 8 * 9 * 64 * nt)
        _, got, _ = _steps(c, st, eps, dt, 1, MODE_GUESS, nsteps=3, graph=True)
        assert c.plan_info()[0] == -(-nt // 9)
        assert np.array_equal(got, want)
        s = torch.from_numpy(st).cuda()
        d = torch.empty_like(s)
        c.autotune_(s, d, [n], eps=eps, dt=dt, mode=MODE_GUESS)
        c.step_(s, d, eps, dt, mode=MODE_GUESS)
        torch.cuda.synchronize()
        one, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        assert np.array_equal(d.cpu().numpy(), one)


def test_benchmark_system_staged_under_the_default_limit(ctx):
    """N = 65536 (3.2 GB of partials unstaged) under the default 1 GiB limit: staged, with
    the unstaged step's deltas bitwise; the staged workspace is reported."""
    import torch
    from bia_3d_gpu_amd import Context, MODE_GUESS, ic
    n = 65536
    st = ic.plummer_with_binaries(n, 65536)
    eps, dt = -2.0 / n, 1e-3
    with Context(0) as c:
        staged, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        ns, wb = c.plan_info()
        c.set_workspace_limit(0)
        full, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        ns0, wb0 = c.plan_info()
    print(f"N = 65536: {ns} stages, {wb / 1e9:.3f} GB of partials against {wb0 / 1e9:.3f} GB unstaged")
    assert ns > 1 and wb <= 1 << 30 and ns0 == 1
    assert np.array_equal(staged, full)
