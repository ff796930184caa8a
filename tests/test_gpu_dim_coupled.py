"""DIM RKF45 with mass and charge couplings (bia_dim_rkf45_coupled_host) against the analytic
pair step with the same couplings.

The reference's DIM has one eps for all pairs (dim/rkf_kernel.cu:29-53); the coupled variant
weights the direct sum by per-source and per-target factors.  What pins it:
  * with equal masses it is bitwise the reference-form DIM with eps = -G m;
  * for two bodies the analytic step is the exact two-body solution, so the two must agree to
    the DIM's own accuracy (unequal masses; attracting and repelling charges);
  * for a few bodies the analytic step's error against the direct sum is that of the BIA
    splitting: the pairwise Kepler drifts get the accelerations and jerks of the true motion
    right and miss the third-body terms from the snap on, so the local error is O(dt^4) in
    position and O(dt^3) in velocity.  Halving dt must divide the errors by 16 and 8.  A wrong
    mass split or coupling in either code would leave an O(dt) or O(dt^2) error (ratios 2 or 4).
"""
import numpy as np
import pytest

from bia_3d_gpu_amd._lib import (BIA_ERR_ARG, COUPLING_COULOMB, COUPLING_GRAVITY, COUPLING_SCALAR, METHOD_ALG,
                                 MODE_GUESS, MODE_PLAIN, BiaError)

pytestmark = pytest.mark.gpu


def _bia(ctx, st, dt, mode, coupling, coup, mass=None, charge=None):
    import torch
    s = torch.from_numpy(np.ascontiguousarray(st)).cuda()
    d = torch.zeros_like(s)
    cu = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    ctx.step_batched_(s, d, [st.shape[1]], dt=dt, method=METHOD_ALG, mode=mode, coupling=coupling, coup=coup,
                      mass=cu(mass), charge=cu(charge))
    torch.cuda.synchronize()
    return d.cpu().numpy()


def _dim(ctx, st, dt, coupling, coup, mass=None, charge=None):
    r = ctx.dim_rkf45(st.T, 0.0, 0.0, dt, coupling=coupling, coup=coup, mass=mass, charge=charge)
    assert r["ier"] == 0
    return r["y"].reshape(-1, 6).T - st, r


def test_equal_masses_reduce_to_the_reference_form(ctx):
    from bia_3d_gpu_amd import ic
    n = 300
    st = ic.plummer_system(n, seed=811)
    m = np.full(n, 1.0 / n)
    a = ctx.dim_rkf45(st.T, -1.0 / n, 0.0, 2e-3)
    b = ctx.dim_rkf45(st.T, 0.0, 0.0, 2e-3, coupling=COUPLING_GRAVITY, coup=1.0, mass=m)
    assert a["ier"] == b["ier"] == 0 and a["nde"] == b["nde"]
    assert np.array_equal(a["y"], b["y"]) and np.array_equal(a["ese"], b["ese"])


def _two_body(kind):
    # attracting: a bound pair; repelling: a hyperbola
    st = np.zeros((6, 2))
    st[0] = [0.6, -0.4]
    st[1] = [0.1, -0.05]
    st[3] = [0.05, -0.1]
    st[4] = [0.5, -0.3]
    st[5] = [0.1, 0.2]
    return st


@pytest.mark.parametrize("kind", ["gravity", "attract", "repel"])
def test_two_bodies_against_the_exact_solution(ctx, kind):
    """Unequal masses: the analytic step is exact for two bodies, the DIM within its tolerance."""
    st = _two_body(kind)
    m = np.array([0.7, 0.3]) if kind == "gravity" else np.array([1.0, 2.5])
    q = None if kind == "gravity" else (np.array([1.0, -1.0]) if kind == "attract" else np.array([1.0, 1.0]))
    cpl = COUPLING_GRAVITY if kind == "gravity" else COUPLING_COULOMB
    dt = 1.0
    dd, r = _dim(ctx, st, dt, cpl, 0.8, mass=m, charge=q)
    for mode in (MODE_PLAIN, MODE_GUESS):
        db = _bia(ctx, st, dt, mode, cpl, 0.8, mass=m, charge=q)
        err = np.abs(db - dd).max(axis=1) / np.abs(dd).max(axis=1)
        print(f"{kind} mode {mode}: {r['naccept']} DIM steps, worst relative deviation {err.max():.2e}")
        assert err.max() < 1e-12
    # momentum of the DIM solution
    p = (m * (st[3:] + dd[3:])).sum(axis=1) - (m * st[3:]).sum(axis=1)
    assert np.abs(p).max() < 1e-14 * (m * np.abs(st[3:])).sum()


def _few_body(seed, coulomb):
    rng = np.random.default_rng(seed)
    n = 6
    m = np.exp(rng.uniform(np.log(0.1), np.log(10), n))
    m /= m.sum()
    st = np.vstack([rng.normal(size=(3, n)), 0.3 * rng.normal(size=(3, n))])
    q = rng.permutation(np.repeat([1.0, -1.0], n // 2)) if coulomb else None
    return st, m, q


@pytest.mark.parametrize("coulomb", [False, True])
def test_splitting_error_order_against_the_direct_sum(ctx, coulomb):
    st, m, q = _few_body(7, coulomb)
    cpl, coup = (COUPLING_COULOMB, 0.05) if coulomb else (COUPLING_GRAVITY, 1.0)
    errs = []
    for dt in (4e-2, 2e-2, 1e-2, 5e-3):
        dd, _ = _dim(ctx, st, dt, cpl, coup, mass=m, charge=q)
        db = _bia(ctx, st, dt, MODE_GUESS, cpl, coup, mass=m, charge=q)
        e = np.abs(db - dd)
        errs.append((e[:3].max(), e[3:].max()))
        print(f"{'Coulomb' if coulomb else 'gravity'} dt {dt:.0e}: position error {errs[-1][0]:.3e}, "
              f"velocity error {errs[-1][1]:.3e}")
    for (p0, v0), (p1, v1) in zip(errs, errs[1:]):
        assert 12 < p0 / p1 < 20 and 6 < v0 / v1 < 10


def test_coupled_dim_argument_errors(ctx):
    st = _two_body("gravity")
    one = np.ones(2)
    calls = [lambda: _scalar_through_coupled(ctx, st)]
    for cpl, m, q in ((COUPLING_GRAVITY, None, None), (COUPLING_COULOMB, one, None), (7, one, one)):
        calls.append(lambda cpl=cpl, m=m, q=q: ctx.dim_rkf45(st.T, 0.0, 0.0, 0.1, coupling=cpl, coup=1.0, mass=m,
                                                             charge=q))
    for call in calls:
        with pytest.raises(BiaError) as e:
            call()
        assert e.value.status == BIA_ERR_ARG


def _scalar_through_coupled(ctx, st):
    """The coupled entry point itself rejects the scalar coupling (the Python wrapper sends it to
    the reference-form entry point instead)."""
    import ctypes as C
    from bia_3d_gpu_amd._lib import check, load
    y = np.ascontiguousarray(st.T).reshape(-1).copy()
    x, h, ini = C.c_double(0.0), C.c_double(0.0), C.c_int(1)
    i0, i1, i2, i3 = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    m = np.ones(2)
    s = load().bia_dim_rkf45_coupled_host(ctx._h, COUPLING_SCALAR, 1.0, 2, m.ctypes.data_as(C.c_void_p), None,
                                          C.byref(x), 0.1, y, C.byref(ini), 0.0, 1e-16, C.byref(h), None,
                                          C.byref(i0), C.byref(i1), C.byref(i2), C.byref(i3))
    check(s, "bia_dim_rkf45_coupled_host", ctx._h)


# ---- mass- and charge-weighted invariants (bia_invariants_coupled_batched_dev) ----------------

def _inv_coupled(ctx, st, sizes, coupling, coup, mass=None, charge=None):
    import torch
    cu = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    out = torch.zeros(len(sizes), 11, dtype=torch.float64, device="cuda")
    ctx.invariants_coupled_(cu(st), sizes, out, coupling, coup, mass=cu(mass), charge=cu(charge))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _inv_longdouble(st, w_k, w_p, f):
    """KE PE Lx Ly Lz Ua Va Wa Xa Ya Za by direct sums in 80-bit long double."""
    s = st.astype(np.longdouble)
    wk, wp = w_k.astype(np.longdouble), w_p.astype(np.longdouble)
    x, v = s[:3], s[3:]
    pe = np.longdouble(0)
    for i in range(s.shape[1] - 1):
        r = np.sqrt(((x[:, i + 1:] - x[:, i:i + 1]) ** 2).sum(axis=0))
        pe += (wp[i] * wp[i + 1:] / r).sum()
    L = (wk * np.cross(x.T, v.T).T).sum(axis=1)
    return np.concatenate([[(wk * (v ** 2).sum(axis=0)).sum(), np.longdouble(f) * pe], L, (wk * v).sum(axis=1),
                           (wk * x).sum(axis=1)])


def test_weighted_invariants_reduce_to_the_reference_form(ctx):
    """Equal masses m = 1/N: every weighted invariant is the reference-form one (eps = -2 G m)
    divided by N, to rounding."""
    import torch
    from bia_3d_gpu_amd import ic
    n = 1000
    st = ic.plummer_system(n, seed=812)
    got = _inv_coupled(ctx, st, [n], COUPLING_GRAVITY, 1.0, mass=np.full(n, 1.0 / n))[0]
    ref = torch.zeros(11, dtype=torch.float64, device="cuda")
    ctx.invariants_(torch.from_numpy(st).cuda(), ref, -2.0 / n)
    ref = ref.cpu().numpy() / n
    scale = np.array([abs(ref[0]), abs(ref[1])] + [np.abs(st).sum() / n] * 9)
    assert np.all(np.abs(got - ref) <= 1e-15 * scale)


@pytest.mark.parametrize("coulomb", [False, True])
def test_weighted_invariants_against_long_double(ctx, coulomb):
    """Unequal masses (and charges): the device sums against 80-bit direct sums, two systems in
    one call."""
    from bia_3d_gpu_amd import ic
    rng = np.random.default_rng(813)
    sizes = [700, 333]
    st = np.concatenate([ic.plummer_system(n, seed=814 + k) for k, n in enumerate(sizes)], axis=1)
    m = np.exp(rng.uniform(np.log(0.1), np.log(10), sum(sizes)))
    q = rng.choice([-1.0, 1.0], sum(sizes)) if coulomb else None
    cpl, coup = (COUPLING_COULOMB, 0.7) if coulomb else (COUPLING_GRAVITY, 1.3)
    got = _inv_coupled(ctx, st, sizes, cpl, coup, mass=m, charge=q)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for k, (a, b) in enumerate(zip(offs[:-1], offs[1:])):
        f = 2 * coup if coulomb else -2 * coup
        want = _inv_longdouble(st[:, a:b], m[a:b], (q if coulomb else m)[a:b], f)
        ab = np.abs(st[:, a:b])
        pw = np.abs(q[a:b]) if coulomb else m[a:b]
        scale = np.array([(m[a:b] * (st[3:, a:b] ** 2).sum(axis=0)).sum(),
                          abs(f) * float(_inv_longdouble(st[:, a:b], m[a:b], pw, 1.0)[1])]
                         + [(m[a:b] * ab[:3].sum(axis=0) * ab[3:].sum(axis=0)).sum()] * 3
                         + [(m[a:b] * ab[3:]).sum()] * 3 + [(m[a:b] * ab[:3]).sum()] * 3)
        err = np.abs(got[k] - want.astype(np.float64)) / scale
        print(f"{'Coulomb' if coulomb else 'gravity'} system {k}: worst relative error {err.max():.1e}")
        assert err.max() < 1e-16
