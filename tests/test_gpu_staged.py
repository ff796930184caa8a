"""Staged steps (bia_set_workspace_limit): a single system whose partials exceed the context's
workspace limit runs in stages of blocks of row tiles, each followed by its reduction; every
particle's slots are still added in the same order, so the deltas are bitwise those of the
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
    from bia_3d_gpu_amd import Context, ic, partial_doubles, staging
    n = 4000 if method == 1 else 1500
    st = ic.plummer_with_binaries(n, 77)
    eps, dt = -2.0 / n, 1e-3
    nb = ((n + 63) // 64 + 3) // 4
    full = 8 * partial_doubles(n)
    with Context(0) as c:
        c.set_workspace_limit(0)
        want, _, wstats = _steps(c, st, eps, dt, method, mode)
        assert c.plan_info() == (1, full)
        # stage sizes: one block, odd sizes, the largest below the unstaged size (a stage of G
        # blocks from block 0 holds partial_doubles(n, G)), and the whole system (unstaged)
        for blocks in (1, 2, 3, (nb - 1) // 2, nb):
            limit = 8 * partial_doubles(n, blocks)
            c.set_workspace_limit(limit)
            got, _, gstats = _steps(c, st, eps, dt, method, mode)
            ns, wb = c.plan_info()
            print(f"method {method} mode {mode}: {blocks} blocks per stage -> {ns} stages, {wb / 1e6:.2f} MB "
                  f"(unstaged {full / 1e6:.2f} MB)")
            assert (ns, wb) == staging(n, limit) == ((1, full) if blocks == nb else (-(-nb // blocks), limit))
            assert np.array_equal(got, want) and np.array_equal(gstats, wstats), blocks


def test_staged_graph_replay_and_autotuner():
    """Three graph-replayed steps with stages equal three eager unstaged steps; the autotuner's
    grids (capped to each stage) leave the staged result unchanged."""
    from bia_3d_gpu_amd import Context, MODE_GUESS, ic, partial_doubles
    import torch
    n = 6000
    st = ic.plummer_with_binaries(n, 78)
    eps, dt = -2.0 / n, 1e-3
    nb = ((n + 63) // 64 + 3) // 4
    with Context(0) as c:
        c.set_workspace_limit(0)
        _, want, _ = _steps(c, st, eps, dt, 1, MODE_GUESS, nsteps=3, graph=False)
        c.set_workspace_limit(8 * partial_doubles(n, 3))
        _, got, _ = _steps(c, st, eps, dt, 1, MODE_GUESS, nsteps=3, graph=True)
        assert c.plan_info()[0] == -(-nb // 3)
        assert np.array_equal(got, want)
        s = torch.from_numpy(st).cuda()
        d = torch.empty_like(s)
        c.autotune_(s, d, [n], eps=eps, dt=dt, mode=MODE_GUESS)
        c.step_(s, d, eps, dt, mode=MODE_GUESS)
        torch.cuda.synchronize()
        one, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        assert np.array_equal(d.cpu().numpy(), one)


@pytest.mark.parametrize("n", [24576, 32768])
def test_staged_segments_of_two_and_four_tiles(n):
    """Staging with the longer work-item segments (two column tiles at N = 24576, four at
    N = 32768): three blocks per stage, the same deltas bitwise as the unstaged step."""
    from bia_3d_gpu_amd import Context, MODE_GUESS, ic, partial_doubles, segment_length, staging
    st = ic.plummer_with_binaries(n, 79)
    eps, dt = -2.0 / n, 1e-3
    nb = ((n + 63) // 64 + 3) // 4
    with Context(0) as c:
        c.set_workspace_limit(0)
        want, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        limit = 8 * partial_doubles(n, 3)
        c.set_workspace_limit(limit)
        got, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        ns, wb = c.plan_info()
    print(f"N = {n} (segments of {segment_length(n)}): {ns} stages of 3 blocks, {wb / 1e6:.1f} MB")
    assert (ns, wb) == staging(n, limit) == (-(-nb // 3), limit)
    assert np.array_equal(got, want)


def test_benchmark_system_staged(ctx):
    """N = 65536 (0.81 GB of partials unstaged, within the default 1 GiB limit) under a 256 MB
    limit: staged, with the unstaged step's deltas bitwise; the staged workspace is reported."""
    from bia_3d_gpu_amd import Context, MODE_GUESS, ic, staging
    n = 65536
    st = ic.plummer_with_binaries(n, 65536)
    eps, dt = -2.0 / n, 1e-3
    with Context(0) as c:
        full, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        ns0, wb0 = c.plan_info()
        c.set_workspace_limit(256 << 20)
        staged, _, _ = _steps(c, st, eps, dt, 1, MODE_GUESS)
        ns, wb = c.plan_info()
    print(f"N = 65536: {ns} stages, {wb / 1e9:.3f} GB of partials against {wb0 / 1e9:.3f} GB unstaged")
    assert ns0 == 1 and wb0 < 1 << 30 and (ns, wb) == staging(n, 256 << 20) and ns > 1 and wb <= 256 << 20
    assert np.array_equal(staged, full)
