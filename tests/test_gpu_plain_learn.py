"""The reference's own solvers started by the learned initialiser (BIA_MODE_PLAIN_LEARN, network 1)
on the GPU: trained on rows the reference solvers record, checked against the reference's fixture
and the oracle, against the same flow with an untrained (all-zero) table, and with a garbage
network, where the reference's guess is the fallback and the judge."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

POS_ABS, VEL_ABS = 3e-11, 1e-14  # the fixture tolerance of the plain mode (test_gpu_parity.py)


@pytest.fixture(scope="module")
def refctx():
    """A context of its own with network 1 trained (a small budget: 800k recorded rows, 800 Adam
    steps); network 0 stays untrained."""
    from bia_3d_gpu_amd import Context, learn
    with Context(0) as c:
        w, curve, rows = learn.train_reference(c, nsamples=400_000, steps=800, batch=8192, lr=3e-3, seed=31,
                                               per_system=100_000)
        yield c, w, curve, rows


def _evals(s):
    return int(s[4:8].sum()) / int(s[0:4].sum())


def test_training_on_reference_solver_rows(refctx):
    """Both branches are in the balanced rows, the loss falls by 4x or more, and selecting network
    1 for training left network 0 untouched (it has no weights, so the learned Kepler mode still
    refuses to run)."""
    from bia_3d_gpu_amd import MODE_LEARN
    c, w, curve, rows = refctx
    r = rows.cpu().numpy()
    ell, hyp = (r[:, 2] < 0).mean(), (r[:, 2] > 0).mean()
    first, last = float(curve[:5].mean()), float(curve[-20:].mean())
    print(f"rows: {len(r)} ({ell:.2f} elliptic, {hyp:.2f} hyperbolic); loss {first:.3e} -> {last:.3e}; "
          f"target variance {r[:, 3].var():.3e}")
    assert 0.3 < ell < 0.7 and 0.3 < hyp < 0.7
    assert last < 0.25 * first and last < 0.25 * r[:, 3].var()
    st = np.random.default_rng(1).normal(size=(6, 64))
    with pytest.raises(RuntimeError):
        c.step_host(st, -1e-3, 1e-3, mode=MODE_LEARN)


def test_fixture_with_learned_starts(refctx, fixture1332):
    """The reference's own N = 1332 fixture (eps = -6e-7, dt = 1, every pair hyperbolic) in the
    plain mode and with learned starts: both within the fixture tolerance of OUT/0001332_10^0.BIA
    and within the oracle's condition bound; every pair starts from the network and none falls
    back; the learned start's single set-up evaluation replaces the reference's two, and the
    network's start against the same flow from the reference's theta_gss (an all-zero network)
    is reported, not assumed."""
    from bia_3d_gpu_amd import MODE_PLAIN, MODE_PLAIN_LEARN
    c, w, _, _ = refctx
    st, eps, dt = fixture1332["state"], fixture1332["eps"], fixture1332["dt"]
    ref = O.bia_step(st, eps, dt)
    bounds = O.sensitivity(st, eps, dt)
    res = {}
    for name, mode in (("plain", MODE_PLAIN), ("learned", MODE_PLAIN_LEARN)):
        d, s = c.step_host(st, eps, dt, mode=mode, with_stats=True)
        p = np.abs(d[:3] - fixture1332["bia"][:3]).max()
        v = np.abs(d[3:] - fixture1332["bia"][3:]).max()
        ok, rp, rv = O.within_condition_bound(st, d, ref, eps, bounds=bounds)
        res[name] = s
        print(f"{name}: vs OUT pos {p:.2e} vel {v:.2e}; error/allowed {rp:.2f} {rv:.2f}; evaluations per pair "
              f"{_evals(s):.3f}; learned starts {s[14]}, fallbacks {s[15]}")
        assert p <= POS_ABS and v <= VEL_ABS and ok
    n = st.shape[1]
    assert res["learned"][14] == n * (n - 1) // 2 and res["learned"][15] == 0
    assert _evals(res["learned"]) < _evals(res["plain"]) - 0.9
    # the same flow from theta_gss: a network with all weights zero tabulates to zero
    c.mlp_select(1)
    try:
        c.mlp_set_weights(np.zeros_like(w))
        c.mlp_build_table(64)
        _, s0 = c.step_host(st, eps, dt, mode=MODE_PLAIN_LEARN, with_stats=True)
    finally:
        c.mlp_set_weights(w)
        c.mlp_build_table(64)
        c.mlp_select(0)
    print(f"the same flow from theta_gss (zero network): {_evals(s0):.3f} evaluations per pair; the trained "
          f"network's start: {_evals(res['learned']):.3f}")


def test_benchmark_binaries_with_learned_starts(refctx):
    """The benchmark system (N = 65536, 20 % in hard binaries, dt = 1e-3), one step: the reference
    solvers stop at their cap on ~1,600 elliptic pairs, and on a few of them (eccentricities above
    0.96) unconverged, which puts particles far off the per-pair RKF45; with the learned starts
    the Kepler equation of those pairs converges, and no particle is off RKF45 by more than 1e-6
    of the largest delta of its component (against 10 with the reference's starts)."""
    import torch
    from bia_3d_gpu_amd import METHOD_RKF, MODE_PLAIN, MODE_PLAIN_LEARN, STATS_LEN, ic
    c, _, _, _ = refctx
    n, dt = 65536, 1e-3
    eps = -2.0 / n
    st = ic.plummer_with_binaries(n, n)
    s = torch.from_numpy(st).cuda()
    c.set_stream(torch.cuda.current_stream().cuda_stream)
    out, stats = {}, {}
    for name, method, mode in (("plain", 1, MODE_PLAIN), ("learned", 1, MODE_PLAIN_LEARN), ("rkf", METHOD_RKF, 0)):
        d = torch.empty_like(s)
        k = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        c.step_batched_(s, d, [n], eps=[eps], dt=dt, method=method, mode=mode, stats=k)
        torch.cuda.synchronize()
        out[name], stats[name] = d.cpu().numpy(), k.cpu().numpy()
    r = out["rkf"]
    scale = np.abs(r).max(axis=1, keepdims=True)
    off = {}
    for name in ("plain", "learned"):
        off[name] = int((np.abs(out[name] - r) > 1e-6 * scale).any(axis=0).sum())
        s_ = stats[name]
        print(f"{name}: particles off RKF45 by > 1e-6 of the largest delta: {off[name]}; caps {s_[10]}, "
              f"evaluations per pair {_evals(s_):.3f}, learned starts {s_[14]}, fallbacks {s_[15]}")
    assert off["plain"] > 0 and off["learned"] == 0
    assert _evals(stats["learned"]) < _evals(stats["plain"])


def test_garbage_network_falls_back_to_the_reference_guess(refctx, fixture1332):
    """Network 1 with random weights of scale 3 (its table far from any root): the reference's
    guess takes over wherever the learned start's run does not converge, so the fixture step stays
    within the fixture tolerance and the oracle's condition bound, and on a strongly bound box
    (512 particles, eps = -2) within the bound too, finite wherever the reference solvers'
    step is (a near-parabolic pair whose branch the device's rounding puts on the hyperbola can
    break the reference's hyperbola there, with or without learned starts)."""
    from bia_3d_gpu_amd import MODE_PLAIN, MODE_PLAIN_LEARN, ic
    c, w, _, _ = refctx
    st, eps, dt = fixture1332["state"], fixture1332["eps"], fixture1332["dt"]
    ref = O.bia_step(st, eps, dt)
    bounds = O.sensitivity(st, eps, dt)
    box, q = ic.coulomb_box(512, seed=5)
    c.mlp_select(1)
    try:
        c.mlp_set_weights((np.random.default_rng(3).standard_normal(w.size) * 3).astype(np.float32))
        c.mlp_build_table(64)
        d, s = c.step_host(st, eps, dt, mode=MODE_PLAIN_LEARN, with_stats=True)
        db, sb = c.step_host(box, -2.0, 5e-4, mode=MODE_PLAIN_LEARN, with_stats=True)
    finally:
        c.mlp_set_weights(w)
        c.mlp_build_table(64)
        c.mlp_select(0)
    ok, rp, rv = O.within_condition_bound(st, d, ref, eps, bounds=bounds)
    p = np.abs(d[:3] - fixture1332["bia"][:3]).max()
    print(f"garbage network, fixture: vs OUT pos {p:.2e}; error/allowed {rp:.2f} {rv:.2f}; fallbacks {s[15]}; "
          f"attractive box: fallbacks {sb[15]} of {sb[0:4].sum()} pairs")
    assert ok and p <= POS_ABS and np.isfinite(d).all()
    dp = c.step_host(box, -2.0, 5e-4, mode=MODE_PLAIN)
    bad_l, bad_p = ~np.isfinite(db).all(axis=0), ~np.isfinite(dp).all(axis=0)
    print(f"attractive box: non-finite particles {bad_l.sum()} with learned starts, {bad_p.sum()} without")
    assert not (bad_l & ~bad_p).any()
    refb = O.bia_step(box, -2.0, 5e-4)
    # the particles of the reference solvers' own non-finite pairs compared as equal (they carry
    # a pair whose conic type rounding decides, which the bound excludes anyway)
    dbm = np.where(bad_p[None, :], refb, db)
    okb, rpb, rvb = O.within_condition_bound(box, dbm, refb, -2.0, 5e-4)
    print(f"attractive box vs oracle: error/allowed {rpb:.2f} {rvb:.2f}")
    assert okb


def test_config1_1000_steps_with_learned_starts(refctx):
    """Config 1 (N = 1332 uniform ball at virial ratio 0.5, seed 1332, eps = -2/N), 1000
    graph-replayed steps of dt = 1e-3.  The reference solvers go wrong on a bound pair at step 39
    and the plain trajectory ends 0.3 off the DIM for its worst particle; with the learned starts
    (the reference's guess as fallback and judge, a bisected start last) the trajectory ends where
    the Kepler drift's does (validated against the DIM RKF45 in test_gpu_configs.py): the median
    particle within 1e-8 of it and every particle within 1e-5 (two solvers' rounding-level
    differences, amplified by close encounters over 1000 steps: 3e-10 and 1.5e-6 measured; the
    drift's own worst particle is 1.5e-4 off the DIM), against 0.3 for the reference's starts, and
    the same |dE/E0|."""
    import torch
    from bia_3d_gpu_amd import MODE_GUESS, MODE_PLAIN, MODE_PLAIN_LEARN, STATS_LEN, ic
    c, _, _, _ = refctx
    n, dt = 1332, 1e-3
    eps = -2.0 / n
    st = ic.uniform_ball_virial(n, seed=1332)
    c.set_stream(torch.cuda.current_stream().cuda_stream)
    end, stats = {}, {}
    for name, mode in (("plain", MODE_PLAIN), ("learned", MODE_PLAIN_LEARN), ("guess", MODE_GUESS)):
        s = torch.from_numpy(st).cuda()
        d = torch.empty_like(s)
        k = torch.zeros(STATS_LEN, dtype=torch.int64, device="cuda")
        c.run_steps_(s, d, [n], 1000, eps=eps, dt=dt, mode=mode, stats=k)
        torch.cuda.synchronize()
        end[name], stats[name] = s.cpu().numpy(), k.cpu().numpy()
    inv = {k: c.invariants(v, eps) for k, v in end.items()}
    e0 = c.invariants(st, eps)
    dE = {k: abs((v["KE"] + v["PE"]) - (e0["KE"] + e0["PE"])) / abs(e0["KE"] + e0["PE"]) for k, v in inv.items()}
    dx = {k: np.linalg.norm(end[k][:3] - end["guess"][:3], axis=0) for k in ("plain", "learned")}
    for k in ("plain", "learned"):
        s_ = stats[k]
        print(f"{k}: position difference from the Kepler drift after 1000 steps: median {np.median(dx[k]):.2e}, "
              f"max {dx[k].max():.2e}; |dE/E0| {dE[k]:.2e} (Kepler drift {dE['guess']:.2e}); caps {s_[10]}, "
              f"fallbacks {s_[15]}, evaluations per pair {_evals(s_):.3f}")
    assert dx["plain"].max() > 1e-2
    assert np.median(dx["learned"]) < 1e-8 and dx["learned"].max() < 1e-5
    assert abs(dE["learned"] - dE["guess"]) < 0.05 * dE["guess"]
