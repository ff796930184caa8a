"""The library's per-pair device code (bia_3d_gpu_amd/csrc/bia_pair.hpp) compiled for the host
(tests/hostsim) against the oracle: same control flow, same results up to the libm-level
differences of the half-angle residual form."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "hostsim", "libpairhost.so")


@pytest.fixture(scope="module")
def hs():
    if not os.path.exists(LIB):
        pytest.skip("host build of the device code missing (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    up = np.ctypeslib.ndpointer(np.uint32, flags="C")
    L.hs_pairs.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_int, dp, dp, ip, ip, up, C.c_int,
                           C.c_void_p, C.c_int]

    def run(method, rel, eps, dt, itr=18, mode=0, table=None):
        m = rel.shape[0]
        out = np.zeros((m, 6))
        br, it, fl = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.uint32)
        tp, tn = (None, 0) if table is None else (table.ctypes.data_as(C.c_void_p), table.shape[0])
        L.hs_pairs(method, m, eps, dt, itr, np.ascontiguousarray(rel), out, br, it, fl, mode, tp, tn)
        return out, br, it, fl
    return run


def _pairs(st):
    n = st.shape[1]
    I, J = np.triu_indices(n, 1)
    return np.ascontiguousarray(st.T[I] - st.T[J])


def _oracle_pairs(method, rel, eps, dt):
    out = np.zeros_like(rel)
    br = np.zeros(len(rel), int)
    fl = np.zeros(len(rel), int)
    for k in range(len(rel)):
        out[k], br[k], _, fl[k] = O.pair(method, eps, dt, rel[k])
    return out, br, fl


def test_fixture_pairs(hs, fixture1332):
    rel = _pairs(fixture1332["state"])[::53]
    a, br, it, fl = hs(1, rel, -6e-7, 1.0)
    b, bb, fb = _oracle_pairs(1, rel, -6e-7, 1.0)
    assert np.all(br == bb)
    assert np.abs(a[:, 3:] - b[:, 3:]).max() < 1e-20
    assert np.abs(a[:, :3] - b[:, :3]).max() < 1e-12


def test_strong_pairs_all_branches(hs):
    from bia_3d_gpu_amd import ic
    st = ic.uniform_ball_virial(90, seed=2)
    st[3:] *= 0.2
    rel = _pairs(st)
    eps, dt = -2 / 90, 4e-3
    a, br, it, fl = hs(1, rel, eps, dt)
    b, bb, fb = _oracle_pairs(1, rel, eps, dt)
    assert np.all(br == bb) and (br == 2).any() and (br == 0).any()
    good = (fb & (O.F_ITR_CAP | O.F_STOCK_DEFAULT)) == 0
    scale = np.abs(rel).max(0)
    err = (np.abs(a - b) / scale).max(1)
    assert np.quantile(err[good], 0.99) < 1e-12 and err[good].max() < 1e-9


def test_rkf_pairs(hs):
    from bia_3d_gpu_amd import ic
    st = ic.uniform_ball_virial(40, seed=3)
    rel = _pairs(st)
    a, br, it, fl = hs(0, rel, -2 / 40, 1e-2)
    b, bb, fb = _oracle_pairs(0, rel, -2 / 40, 1e-2)
    assert np.all(br == 3) and np.abs(a - b).max() < 1e-15


def test_kepler_drift_step_vs_oracle(hs):
    """BIA_MODE_GUESS (universal-variable Kepler drift from the series start): the whole step
    agrees with the oracle's reference step within the reference formulas' own rounding
    sensitivity, in about one evaluation of the time equation per pair."""
    from bia_3d_gpu_amd import ic
    n = 1024
    st = ic.plummer_with_binaries(n, seed=8)
    eps, dt = -2 / 65536, 1e-3
    I, J = np.triu_indices(n, 1)
    rel = _pairs(st)
    a, br, it, fl = hs(1, rel, eps, dt, mode=2)
    d = np.zeros((6, n))
    np.add.at(d.T, I, a)
    np.add.at(d.T, J, -a)
    d[:3] += st[3:] * dt
    ok, rp, rv = O.within_condition_bound(st, d, O.bia_step(st, eps, dt), eps, dt)
    assert ok, (rp, rv)
    assert it.mean() < 1.01 and (fl & O.F_ITR_CAP).sum() == 0
    assert (br == 0).mean() > 0.99 and (br == 2).any()  # the binaries are elliptic


def _rand_pairs(rng, m, eps, gmin, gmax):
    rels = []
    for _ in range(m):
        r = rng.uniform(0.5, 1.5)
        g = rng.uniform(gmin, gmax) * np.sqrt(2 * abs(eps) / r)  # in units of the escape speed
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        rels.append(np.concatenate([[r, 0, 0], g * u]))
    return np.array(rels)


@pytest.mark.parametrize("eps,gmin,gmax,dt", [(-1.0, 0.2, 0.95, 0.3),   # ellipses, some over whole periods
                                              (-1.0, 1.05, 3.0, 0.3),   # strongly deflected hyperbolae
                                              (+1.0, 0.1, 3.0, 0.3),    # repulsive
                                              (-1.0, 0.999, 1.001, 0.2)])  # near-parabolic
def test_kepler_drift_strong_pairs_vs_rkf45(hs, eps, gmin, gmax, dt):
    """Strong pairs against the oracle's independent adaptive RKF45 integration of the same pair
    (absolute tolerance 1e-16), including pairs where the reference's own Newton stops at its
    iteration cap away from the root."""
    rng = np.random.default_rng(int(10 * gmin + 100 * gmax))
    rels = _rand_pairs(rng, 150, eps, gmin, gmax)
    a, br, it, fl = hs(1, rels, eps, dt, mode=2)
    k, _, _, kf = hs(0, rels, eps, dt)
    good = kf == 0
    assert good.mean() > 0.9 and (fl & O.F_ITR_CAP).sum() == 0
    err = np.abs(a - k)[good].max(1) / np.abs(k[good]).max(1)
    assert np.quantile(err, 0.99) < 1e-11 and err.max() < 1e-9, err.max()


def test_kepler_drift_whole_periods_and_reversal(hs):
    """A bound pair over exactly m periods returns to its start (zero velocity change, position
    change -v0 dt / 2 after the drift is removed); stepping back by -dt undoes a step."""
    eps = -1.0
    mu = 1.0
    r0, v0 = 1.0, 1.1
    a = 1 / (2 / r0 - v0 * v0 / mu)
    P = 2 * np.pi * np.sqrt(a ** 3 / mu)
    rel = np.array([[r0, 0, 0, 0.3, v0 * np.sqrt(1 - 0.09 / v0 ** 2), 0]])
    rel[0, 3:] *= v0 / np.linalg.norm(rel[0, 3:])
    d, br, it, fl = hs(1, rel, eps, 3 * P, mode=2)
    assert br[0] == 2
    assert np.abs(d[0, 3:]).max() < 1e-12
    assert np.abs(d[0, :3] + rel[0, 3:] * 3 * P / 2).max() < 1e-12 * 3 * P
    # reversal: forward dt then back
    dt = 0.37
    f, _, _, _ = hs(1, rel, eps, dt, mode=2)
    moved = rel.copy()
    moved[0, :3] += 2 * f[0, :3] + rel[0, 3:] * dt
    moved[0, 3:] += 2 * f[0, 3:]
    b, _, _, _ = hs(1, moved, eps, -dt, mode=2)
    back = moved.copy()
    back[0, :3] += 2 * b[0, :3] - moved[0, 3:] * dt
    back[0, 3:] += 2 * b[0, 3:]
    assert np.abs(back - rel).max() < 1e-13


def test_kepler_drift_trivial_cases(hs):
    rel = np.array([[0.3, -0.2, 0.5, 0.1, 0.4, -0.3]])
    d, _, _, _ = hs(1, rel, -1.0, 0.0, mode=2)
    assert np.all(d == 0)  # dt = 0
    d, _, _, _ = hs(1, rel, 0.0, 0.7, mode=2)
    assert np.all(d == 0)  # no interaction: straight line exactly


def test_kepler_drift_where_the_reference_fails(hs):
    """Coulomb box (unit masses, charges +-1, 40 % of pairs bound): on a few bound pairs the
    reference's ellipse Newton ends at its iteration cap far from the root.  The Kepler drift
    agrees with the independent RKF45 integration on every capped pair, the reference algorithm
    (host build of the plain mode, which the oracle pins) does not on some."""
    from bia_3d_gpu_amd import ic
    st, q = ic.coulomb_box(512, seed=13)
    I, J = np.triu_indices(512, 1)
    rel = _pairs(st)
    eps = 2.0 * q[I] * q[J]
    sel = eps < 0
    rel = rel[sel]
    p, br, it, fl = hs(1, rel, -2.0, 5e-4)
    capped = np.where((br == 2) & ((fl & O.F_ITR_CAP) != 0))[0]
    assert len(capped) > 1000
    rel = rel[capped]
    p = p[capped]
    k, _, _, kf = hs(0, rel, -2.0, 5e-4)
    a, _, _, af = hs(1, rel, -2.0, 5e-4, mode=2)
    good = kf == 0
    scale = np.abs(k).max(1)
    err_kep = (np.abs(a - k).max(1) / scale)[good]
    err_ref = (np.abs(p - k).max(1) / scale)[good]
    assert (af & O.F_ITR_CAP).sum() == 0
    assert err_kep.max() < 1e-10
    assert (err_ref > 1e-3).sum() >= 3


def _universal_drift_40_digits(rel, eps, dt):
    """The pair's half deltas from the closed-form two-body solution in 40-digit arithmetic
    (mpmath): the universal anomaly s solves r0 s + eta G2(s) + zeta G3(s) = dt by bracketed
    root finding (the left side increases with s), G_k(s) = s^k c_k(beta s^2) summed to 80 terms,
    then the Lagrange coefficients f - 1 = -mu G2 / r0, g - dt = -mu G3, df = -mu G1 / (r0 r),
    dg - 1 = -mu G2 / r, halved.  Independent of the reference's solvers and of the library's
    Stumpff evaluation, Halley iteration and Taylor carry."""
    import mpmath as mp
    with mp.workdps(40):
        r0 = [mp.mpf(float(v)) for v in rel[:3]]
        v0 = [mp.mpf(float(v)) for v in rel[3:]]
        mu, tau = -mp.mpf(eps), mp.mpf(dt)
        R0 = mp.sqrt(sum(c * c for c in r0))
        eta = sum(a * b for a, b in zip(r0, v0))
        beta = 2 * mu / R0 - sum(c * c for c in v0)
        zeta = mu - beta * R0

        def G(s, k):
            z = beta * s * s
            return s ** k * mp.fsum((-z) ** j / mp.factorial(2 * j + k) for j in range(80))

        def f(s):
            return R0 * s + eta * G(s, 2) + zeta * G(s, 3) - tau
        lo, hi = mp.mpf(0), tau / R0
        while f(hi) < 0:
            lo, hi = hi, 2 * hi
        s = mp.findroot(f, (lo, hi), solver="anderson")
        rn = R0 + eta * G(s, 1) + zeta * G(s, 2)
        fm1, gmdt = -mu * G(s, 2) / R0, -mu * G(s, 3)
        fd, gdm1 = -mu * G(s, 1) / (R0 * rn), -mu * G(s, 2) / rn
        d = [(fm1 * r0[q] + gmdt * v0[q]) / 2 for q in range(3)] + [(fd * r0[q] + gdm1 * v0[q]) / 2 for q in range(3)]
        return np.array([float(c) for c in d])


def test_kepler_drift_against_40_digit_solution(hs):
    """Pairs of the benchmark's kind (Plummer with hard binaries, dt = 1e-3: the binaries and a
    sample of weak distant pairs) and strong pairs of every conic type, against the closed-form
    solution in 40 digits: within 1e-14 of each pair's largest delta component (the library forms
    the departure from straight-line motion directly, so weak pairs keep their full relative
    precision; an integrator that advances the whole state, like the per-pair RKF45, loses it to
    the rounding of |r| and stops near 1e-5 of these deltas)."""
    pytest.importorskip("mpmath")
    from bia_3d_gpu_amd import ic
    n = 4096
    st = ic.plummer_with_binaries(n, 4096)
    eps, dt = -2.0 / n, 1e-3
    x = st[:3].T
    rng = np.random.default_rng(40)
    i = rng.integers(0, n, 4000)
    j = rng.integers(0, n, 4000)
    keep = i != j
    close = [(a, np.argsort(np.linalg.norm(x - x[a], axis=1))[1]) for a in rng.integers(0, n, 400)]
    close = [(a, b) for a, b in close if np.linalg.norm(x[a] - x[b]) < 0.02][:20]
    pairs = list(zip(i[keep][:20], j[keep][:20])) + close
    rel = np.array([st[:, a] - st[:, b] for a, b in pairs])
    cases = [(rel, eps, dt)]
    rng2 = np.random.default_rng(41)
    for e, dt2 in ((-1.0, 0.3), (-1.0, 3.0), (1.0, 0.2)):  # bound (one over whole periods), repulsive
        r = rng2.normal(size=(6, 6))
        r[:, 3:] *= 0.7
        cases.append((r, e, dt2))
    worst = 0.0
    for rel_c, e, t in cases:
        got = hs(1, np.ascontiguousarray(rel_c), e, t, mode=2)[0]
        for k in range(len(rel_c)):
            ex = _universal_drift_40_digits(rel_c[k], e, t)
            err = np.abs(got[k] - ex).max() / np.abs(ex).max()
            worst = max(worst, err)
            assert err < 1e-14, (e, t, k, err)
    print(f"{sum(len(c[0]) for c in cases)} pairs: largest error {worst:.1e} of the pair's largest delta")


def _drift_closed_form(rel, eps, dt, dps=80, exact=False):
    """The pair's half deltas from the two-body solution with the Stumpff functions in closed
    form (cos/sin for bound, cosh/sinh for unbound orbits) and the time equation solved by
    bisection, in `dps` digits (mpmath): valid at any strength, where the 80-term series of
    _universal_drift_40_digits is not.  With exact, the mpmath values themselves."""
    import mpmath as mp
    with mp.workdps(dps):
        r0 = [mp.mpf(float(v)) for v in rel[:3]]
        v0 = [mp.mpf(float(v)) for v in rel[3:]]
        mu, tau = -mp.mpf(eps), mp.mpf(dt)
        R0 = mp.sqrt(sum(c * c for c in r0))
        eta = sum(a * b for a, b in zip(r0, v0))
        beta = 2 * mu / R0 - sum(c * c for c in v0)
        zeta = mu - beta * R0

        def gfun(s):  # G1, G2, G3 at s
            z = beta * s * s
            if z == 0:
                return s, s * s / 2, s ** 3 / 6
            q = mp.sqrt(abs(z))
            if z > 0:
                c1, c2, c3 = mp.sin(q) / q, (1 - mp.cos(q)) / z, (q - mp.sin(q)) / q ** 3
            else:
                c1, c2, c3 = mp.sinh(q) / q, (mp.cosh(q) - 1) / -z, (mp.sinh(q) - q) / q ** 3
            return s * c1, s * s * c2, s ** 3 * c3

        def t(s):
            _, g2, g3 = gfun(s)
            return R0 * s + eta * g2 + zeta * g3 - tau
        lo, hi = mp.mpf(0), tau / R0
        while t(hi) < 0:
            lo, hi = hi, 2 * hi
        for _ in range(dps * 4):
            mid = (lo + hi) / 2
            if t(mid) < 0:
                lo = mid
            else:
                hi = mid
        s = (lo + hi) / 2
        g1, g2, g3 = gfun(s)
        rn = R0 + eta * g1 + zeta * g2
        fm1, gmdt = -mu * g2 / R0, -mu * g3
        fd, gdm1 = -mu * g1 / (R0 * rn), -mu * g2 / rn
        d = [(fm1 * r0[q] + gmdt * v0[q]) / 2 for q in range(3)] + [(fd * r0[q] + gdm1 * v0[q]) / 2 for q in range(3)]
        return d if exact else np.array([float(c) for c in d])


def test_kepler_drift_converges_on_strong_pairs(hs):
    """Random pairs of every geometry up to max|a_k| = 1e3 (bound over many periods, fast
    flybys, repulsive, near-radial pairs that pass pericentre within dt): the safeguarded
    iteration converges on every one (no iteration cap, finite output, at most 30 evaluations),
    and a sample, with three near-radial incoming pairs that once stopped at the 60-evaluation cap
    (their Halley steps, from far up the exponential branch, crept down without halving; the
    iteration now bisects when the steps stop halving), agrees with the closed-form solution.
    The reference's own solvers (plain mode) stop at their cap on 17-40 % of such pairs."""
    pytest.importorskip("mpmath")
    rng = np.random.default_rng(77)
    sample = []
    for trial in range(60):
        m = 1000
        r = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-3, 1, (m, 1))
        v = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-3, 1, (m, 1))
        if trial % 3 == 2:  # fast pairs: max|a_k| up to 1e12, G functions that overflow beyond the root
            v *= 10.0 ** rng.uniform(0, 3, (m, 1))
        rel = np.ascontiguousarray(np.hstack([r, v]))
        eps = -10.0 ** rng.uniform(-4, 0.5) * (1 if trial % 4 else -1)
        dt = 10.0 ** rng.uniform(-3, 0.5)
        r0 = np.linalg.norm(r, axis=1)
        mu = -eps
        beta = 2 * mu / r0 - (v * v).sum(1)
        a = np.abs([(r * v).sum(1) * dt / r0 ** 2, (mu - beta * r0) * dt ** 2 / r0 ** 3, beta * dt ** 2 / r0 ** 2])
        keep = a.max(axis=0) <= (1e12 if trial % 3 == 2 else 1e3)
        out, br, it, fl = hs(1, np.ascontiguousarray(rel[keep]), eps, dt, mode=2)
        assert np.isfinite(out).all() and not (fl & 4).any() and it.max() <= 30, (trial, it.max())
        if trial % 3 == 2:
            continue
        k = int(np.argmax(it))  # the pair that took the most evaluations
        sample.append((rel[keep][k], eps, dt, out[k]))
    for c in ([0.0026037201609559564, -0.0070822235211471804, 0.024403191399523627, 1.6446235016284003,
               1.8344757887165375, -20.8830149885328, -3.9285356109812746, 0.0012847944556660402],
              [-0.2446222194714971, -0.05091575606808398, -0.48917818595901047, 9.262602111071313,
               4.075932064638315, 17.254861329258418, -7.267704183781688e-08, 0.04293330885457181],
              [0.0001719005647052639, -0.00304250995202669, -0.0009257702154520646, -0.028437539851761458,
               0.268365026562335, 0.09567271803022319, -7.267704183781688e-08, 0.04293330885457181]):
        rel = np.array([c[:6]])
        out, br, it, fl = hs(1, rel, c[6], c[7], mode=2)
        assert not (fl & 4).any() and it[0] <= 30, it[0]
        sample.append((rel[0], c[6], c[7], out[0]))
    worst = 0.0
    for rel, eps, dt, got in sample[::4] + sample[-3:]:
        ex = _drift_closed_form(rel, eps, dt)
        err = np.abs(got - ex).max() / np.abs(ex).max()
        worst = max(worst, err)
        assert err < 1e-10, (rel.tolist(), eps, dt, err)
    print(f"{len(sample[::4]) + 3} hardest pairs against the closed form: largest error {worst:.1e}")


def test_learned_start_cannot_derail_the_solver(hs):
    """The learned start with tables of random values (none trained): the proposal is held to a
    factor 2^6 of the series start and a residual that overflowed counts as beyond the root, so
    every pair up to max|a_k| = 1e3 still converges (before these rules, starts far up the
    exponential branch of strong hyperbolae gave -inf residuals taken as below the root)."""
    rng = np.random.default_rng(78)
    for scale in (5.0, 50.0):
        table = np.ascontiguousarray((rng.standard_normal((64, 64, 64)) * scale).astype(np.float32))
        for trial in range(8):
            m = 20000
            r = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-3, 1, (m, 1))
            v = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-3, 1, (m, 1))
            rel = np.hstack([r, v])
            eps = -10.0 ** rng.uniform(-4, 0.5) * (1 if trial % 4 else -1)
            dt = 10.0 ** rng.uniform(-3, 0.5)
            r0 = np.linalg.norm(r, axis=1)
            mu = -eps
            beta = 2 * mu / r0 - (v * v).sum(1)
            a = np.abs([(r * v).sum(1) * dt / r0 ** 2, (mu - beta * r0) * dt ** 2 / r0 ** 3, beta * dt ** 2 / r0 ** 2])
            keep = a.max(axis=0) <= 1e3
            out, br, it, fl = hs(1, np.ascontiguousarray(rel[keep]), eps, dt, mode=1, table=table)
            assert (fl & 64).any()  # learned starts taken
            assert np.isfinite(out).all() and not (fl & 4).any() and it.max() <= 30, (scale, trial, it.max())


def _ulp_error(y, ref):
    """|y - ref| in units of the last place of ref (ref in extended precision)."""
    ulp = np.spacing(np.abs(ref.astype(np.float64)))
    return (np.abs(y.astype(np.longdouble) - ref) / ulp).astype(np.float64)


def test_range_specific_sincos_and_log():
    """The plain mode's sin/cos of |x| <= pi and log of positive normal x (bia_pair.hpp
    sincos_pi, log_pos; host build of the same source) against 80-bit extended-precision values:
    within 0.7 ulp for sin and cos and 1.5 ulp for log, on random arguments, on the points where
    the table index switches ((k + 1/2) pi/64) or the reduction cancels (near k pi/64, pi/2, pi),
    and log near 1.  A wrong table entry or reduction constant, a dropped series term, the leading
    term's rounding left uncorrected (1.5 ulp) or an uncorrected quotient would show."""
    L = C.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.hs_sincos_pi.argtypes = [C.c_int64, dp, dp, dp]
    L.hs_log_pos.argtypes = [C.c_int64, dp, dp]
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("no 80-bit long double on this host")
    rng = np.random.default_rng(2016)
    pi = np.pi
    grid = np.concatenate([np.arange(1, 65), np.arange(0, 64) + 0.5]) * (pi / 64)
    edges = np.concatenate([(np.array([0.25, 0.5, 0.75, 1.0]) * pi * (1 + rng.uniform(-1e-6, 1e-6, (2000, 1)))).ravel(),
                            (grid * (1 + rng.uniform(-1e-9, 1e-9, (300, 1)))).ravel()])
    x = np.concatenate([rng.uniform(-pi, pi, 400000), edges, -edges, 10.0 ** rng.uniform(-300, 0, 20000),
                        np.array([0.0, pi, -pi, pi / 2, pi / 4, 3 * pi / 4])])
    x = np.ascontiguousarray(x[np.abs(x) <= pi])
    s, c = np.zeros_like(x), np.zeros_like(x)
    L.hs_sincos_pi(x.size, x, s, c)
    xl = x.astype(np.longdouble)
    es, ec = _ulp_error(s, np.sin(xl)), _ulp_error(c, np.cos(xl))
    print(f"sincos_pi on {x.size} arguments: max {es.max():.3f} ulp (sin), {ec.max():.3f} ulp (cos)")
    assert es.max() < 0.7 and ec.max() < 0.7
    y = np.concatenate([10.0 ** rng.uniform(-307, 307, 200000), 1 + rng.uniform(-0.3, 0.3, 200000),
                        1 + 10.0 ** rng.uniform(-15, -1, 20000), 1 - 10.0 ** rng.uniform(-15, -1, 20000),
                        np.array([1.0, 2.0, 0.5, np.sqrt(2.0), np.sqrt(0.5), 1e-300, 1e300])])
    out = np.zeros_like(y)
    L.hs_log_pos(y.size, np.ascontiguousarray(y), out)
    ref = np.log(y.astype(np.longdouble))
    nz = ref != 0
    el = _ulp_error(out[nz], ref[nz])
    print(f"log_pos on {y.size} arguments: max {el.max():.3f} ulp; log(1) = {out[~nz]}")
    assert el.max() < 1.5 and (out[~nz] == 0).all()


def test_range_specific_atan2_and_acos():
    """The plain mode's set-up inverses (bia_pair.hpp atan2_upper for y >= 0, acos_unit for
    |q| <= 1; host build of the same source) against 80-bit extended precision: within 2.5 ulp
    (atan2: the reduced argument is the quotient of a rounded difference and sum; 2.3 measured)
    and 1.1 ulp (acos),
    on random arguments over many binades and both signs of x, at the reduction's switch points
    (y/|x| = tan(pi/8), y = |x|; |q| = 1/2) and near the ends (q = +-1, y << |x|, y >> |x|).  A
    wrong polynomial coefficient, quadrant constant or reflection would show as many ulps."""
    L = C.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.hs_atan2_upper.argtypes = [C.c_int64, dp, dp, dp]
    L.hs_acos_unit.argtypes = [C.c_int64, dp, dp]
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("no 80-bit long double on this host")
    rng = np.random.default_rng(1974)
    n = 200000
    y = 10.0 ** rng.uniform(-8, 8, n)
    x = 10.0 ** rng.uniform(-8, 8, n) * rng.choice([-1.0, 1.0], n)
    t8 = np.tan(np.pi / 8)
    ye = np.abs(x[:40000]) * np.concatenate([t8 * (1 + rng.uniform(-1e-9, 1e-9, 20000)),
                                             1 + rng.uniform(-1e-9, 1e-9, 20000)])
    y = np.ascontiguousarray(np.concatenate([y, ye, np.abs(x[:2000]) * 1e-12, np.abs(x[:2000]) * 1e12]))
    x = np.ascontiguousarray(np.concatenate([x, x[:40000], x[:2000], x[:2000]]))
    out = np.zeros_like(y)
    L.hs_atan2_upper(y.size, y, x, out)
    ref = np.arctan2(y.astype(np.longdouble), x.astype(np.longdouble))
    ea = _ulp_error(out, ref)
    q = np.concatenate([rng.uniform(-1, 1, 200000), 0.5 * (1 + rng.uniform(-1e-9, 1e-9, 10000)),
                        -0.5 * (1 + rng.uniform(-1e-9, 1e-9, 10000)), 1 - 10.0 ** rng.uniform(-16, -1, 10000),
                        -1 + 10.0 ** rng.uniform(-16, -1, 10000), np.array([1.0, -1.0, 0.0, 0.5, -0.5])])
    q = np.ascontiguousarray(q)
    oq = np.zeros_like(q)
    L.hs_acos_unit(q.size, q, oq)
    refq = np.arccos(q.astype(np.longdouble))
    nz = refq != 0
    eq = _ulp_error(oq[nz], refq[nz])
    print(f"atan2_upper on {y.size} arguments: max {ea.max():.3f} ulp; acos_unit on {q.size}: max {eq.max():.3f} "
          f"ulp, acos(1) = {oq[~nz]}")
    assert ea.max() < 2.5 and eq.max() < 1.1 and (oq[~nz] == 0).all()


def test_quotient_from_a_shared_reciprocal():
    """The device's quotient a / b from a reciprocal r of b (bia_pair.hpp div_r: q0 = a r, then
    q0 + (a - q0 b) r with both steps fused multiply-adds), emulated here in exact rational
    arithmetic with one rounding per fused operation: it equals the correctly rounded quotient for
    r the correctly rounded 1/b and for r one ulp off either way (the device's fast_rcp is within an
    ulp), over operands spread across 120 binades, and on quotients that are exact or lie next to
    a rounding boundary.  Dropping the correction (q0 alone) fails on about a quarter of them.
    The same for the square root from a reciprocal square root (sqrt_r)."""
    import math
    from fractions import Fraction as F

    def fma(x, y, z):  # one rounding (float() of a Fraction rounds to nearest even)
        return float(F(x) * F(y) + F(z))

    rng = np.random.default_rng(8)
    n = 6000
    a = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-60, 60, n)
    b = rng.uniform(0.5, 1, n) * 2.0 ** rng.integers(-60, 60, n) * rng.choice([-1.0, 1.0], n)
    # exact quotients and quotients within a few ulps of a rounding boundary: a = b q (1 + k 2^-53)
    q = rng.uniform(1, 2, 400)
    a2 = np.array([float(F(bb) * F(qq)) for bb, qq in zip(b[:400], q)])
    a = np.concatenate([a, a2, b[400:600] * 3.0])
    b = np.concatenate([b, b[:400], b[400:600]])
    bad = plain = 0
    for x, y in zip(a.tolist(), b.tolist()):
        r0 = 1.0 / y
        q0 = x * r0
        plain += q0 != x / y
        for r in (r0, math.nextafter(r0, math.inf), math.nextafter(r0, -math.inf)):
            q0 = x * r
            bad += fma(fma(-q0, y, x), r, q0) != x / y
    print(f"{3 * a.size} quotients: {bad} not correctly rounded; without the correction {plain} of {a.size}")
    assert bad == 0 and plain > a.size // 10
    # the square root from a reciprocal square root y (sqrt_r: s0 = x y, then s0 + (x - s0^2) y / 2):
    # correctly rounded for y the rounded 1/sqrt(x) and one ulp off either way
    x = rng.uniform(0.5, 2, 6000) * 2.0 ** rng.integers(-80, 80, 6000)
    bad = plain = 0
    for v in x.tolist():
        y0 = 1.0 / math.sqrt(v)
        plain += v * y0 != math.sqrt(v)
        for y in (y0, math.nextafter(y0, math.inf), math.nextafter(y0, -math.inf)):
            s0 = v * y
            bad += fma(fma(-s0, s0, v), 0.5 * y, s0) != math.sqrt(v)
    print(f"{3 * x.size} square roots: {bad} not correctly rounded; without the correction {plain} of {x.size}")
    assert bad == 0 and plain > x.size // 10


def _binary_pairs(n=65536, seed=65536):
    """Each particle of the benchmark system with its nearest neighbour (its binary companion when
    it has one): 6526 elliptic pairs with eccentricities up to 0.9993, the rest hyperbolic."""
    from scipy.spatial import cKDTree
    from bia_3d_gpu_amd import ic
    st = ic.plummer_with_binaries(n, seed)
    _, idx = cKDTree(st[:3].T).query(st[:3].T, k=2)
    i = np.arange(n)
    keep = i < idx[:, 1]
    return np.ascontiguousarray(st[:, i[keep]].T - st[:, idx[keep, 1]].T), -2.0 / n, 1e-3


def _quad_error(rel, eps, dt, d):
    err = np.zeros(len(rel))
    for k in range(len(rel)):
        hi, _ = O.pair_quad(eps, dt, rel[k])
        err[k] = np.abs(d[k] - hi).max() / np.abs(hi).max()
    return err


def test_reference_solvers_with_learned_starts_mechanics(hs, fixture1332):
    """BIA_MODE_PLAIN_LEARN with an all-zero table, where the starts are the reference's own
    theta_gss (hyperbola) and E = M (ellipse): on the fixture's pairs the results agree with the
    reference solvers' to 1e-9 of each pair's delta and 1e-12 in position deltas, and the pairs
    take one evaluation fewer (the learned start's single set-up evaluation replaces the
    reference's two).  On the benchmark system's binaries, the reference ellipse Newton stops at
    its cap unconverged on 5 high-eccentricity pairs (e > 0.96) and returns wrong deltas; there
    the reference guess's run is kept only where its residual is smaller (the judge), and all 5
    come out within 1e-10 of the quad-precision two-body solution."""
    table = np.zeros((64, 64, 64), np.float32)
    rel = _pairs(fixture1332["state"])[::53]
    a, bra, ita, fla = hs(1, rel, -6e-7, 1.0)
    b, brb, itb, flb = hs(1, rel, -6e-7, 1.0, mode=3, table=table)
    assert np.all(bra == brb) and ((flb & 64) != 0).all()
    scale = np.abs(a).max(axis=1)
    assert (np.abs(a - b).max(axis=1) <= 1e-9 * scale).all()
    assert np.abs(a[:, :3] - b[:, :3]).max() < 1e-12
    print(f"fixture pairs: evaluations per pair {ita.mean():.3f} (reference) -> {itb.mean():.3f}")
    assert itb.mean() < ita.mean() - 0.9
    rel, eps, dt = _binary_pairs()
    a, bra, ita, fla = hs(1, rel, eps, dt)
    b, brb, itb, flb = hs(1, rel, eps, dt, mode=3, table=table)
    ell = bra == 2
    wrong = ell & (_quad_error(rel, eps, dt, a) > 1e-10) & ((fla & 4) != 0)
    print(f"binaries: {ell.sum()} ellipses, the reference wrong on {wrong.sum()} (capped unconverged); "
          f"{((flb & 128) != 0).sum()} fallbacks to the reference guess")
    assert wrong.sum() == 5
    assert (_quad_error(rel[wrong], eps, dt, b[wrong]) < 1e-10).all()
    # elsewhere the two agree to the reference formulas' own rounding: weak hyperbolic pairs carry
    # 1e-8 to 1e-7 of it in their position departure (both modes alike against the quad solution:
    # the largest difference, 1.1e-7, is a pair both of whose results are 4e-8 and 6e-8 off it)
    ok = ~wrong
    d = np.abs(a[ok] - b[ok]).max(axis=1) / np.abs(a[ok]).max(axis=1)
    assert (d <= 3e-7).all()
    far = d > 1e-9
    if far.any():
        qa, qb = _quad_error(rel[ok][far], eps, dt, a[ok][far]), _quad_error(rel[ok][far], eps, dt, b[ok][far])
        assert (qb <= 10 * qa + 1e-8).all()


def test_reference_solvers_with_learned_starts_survive_any_table(hs):
    """Tables of random values (none trained): the hyperbola's proposal is kept inside its branch
    and its factor exp(T / e) within e^4, the ellipse's within e of M, and a run that has not
    converged (cap, residual above 1e-12 of the time to reach, or — the reference loop's quirk —
    no iterate better than the start, which returns theta = 0) falls back to the reference's
    guess; so on strong pairs of every branch (attractive and repulsive, dt up to 3) every output
    is finite and agrees with the reference solvers' wherever neither stops at its cap.  Without
    the default-iterate rule, three such pairs came out 0.8 to 4474 times their delta off."""
    rng = np.random.default_rng(79)
    for scale in (5.0, 50.0):
        table = np.ascontiguousarray((rng.standard_normal((64, 64, 64)) * scale).astype(np.float32))
        for trial in range(6):
            m = 4000
            r = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-2, 0, (m, 1))
            v = rng.standard_normal((m, 3)) * 10.0 ** rng.uniform(-2, 0, (m, 1))
            rel = np.ascontiguousarray(np.hstack([r, v]))
            eps = -10.0 ** rng.uniform(-3, -0.5) * (1 if trial % 3 else -1)
            dt = 10.0 ** rng.uniform(-2, 0.5)
            a, bra, ita, fla = hs(1, rel, eps, dt)
            b, brb, itb, flb = hs(1, rel, eps, dt, mode=3, table=table)
            assert np.isfinite(b).all() and np.all(bra == brb)
            both = ((fla | flb) & 4) == 0  # neither run capped
            d = np.abs(a[both] - b[both]).max(axis=1) / np.abs(a[both]).max(axis=1)
            assert (d < 1e-6).all(), (scale, trial, d.max())


def test_reference_solvers_with_learned_starts_last_resort(hs):
    """A hyperbolic pair of the config-1 ball (N = 1332, eps = -2/N, dt = 1e-3) at step 457 of its
    trajectory, near-radial and close (theta_ini 0.14 inside -theta_0): the reference's Newton
    overshoots out of the branch, stops at its cap with no improved iterate and returns theta = 0,
    a velocity change of 0.81 where the two-body solution gives 0.0074.  With learned starts (an
    all-zero table, so the learned start is the reference's theta_gss) that run fails as well, and
    the bisected start that comes last converges to within 1e-10 of the quad-precision solution."""
    rel = np.array([float.fromhex(h) for h in ('-0x1.ef10c9e790b60p-8', '0x1.af59a8aa314e0p-8',
                                              '0x1.9392f96f9dd00p-9', '0x1.d2198a3cc9fc8p-1',
                                              '-0x1.3b5e176a59078p-1', '-0x1.a2f6248c2c580p-2')])[None, :]
    eps, dt = -2.0 / 1332, 1e-3
    table = np.zeros((64, 64, 64), np.float32)
    a, bra, ita, fla = hs(1, rel, eps, dt)
    b, brb, itb, flb = hs(1, rel, eps, dt, mode=3, table=table)
    ea, eb = _quad_error(rel, eps, dt, a)[0], _quad_error(rel, eps, dt, b)[0]
    print(f"reference: flags {fla[0]}, error {ea:.1e} of the pair's delta; learned starts: flags {flb[0]}, "
          f"{itb[0]} evaluations, error {eb:.1e}")
    assert bra[0] == 0 and (fla[0] & 4) and (fla[0] & 1) and ea > 1
    assert (flb[0] & 128) and eb < 1e-10


@pytest.mark.parametrize("case", ["fixture", "benchmark", "strong", "coulomb"])
def test_hyperbola_loop_repeat_exit_without_the_evaluation(case, fixture1332):
    """The library's hyperbola loop takes the reference's exact-repeat exit without evaluating an
    anomaly that repeats one of the last two it evaluated (the residual is a function of the
    anomaly, so it would repeat a history entry).  Against the reference loop with every
    iteration evaluated (bia/find_dXYUV_kernel.cu:186-217), from the same set-up and guess
    selection: best iterate, best residual, iteration count and flags bitwise equal on every
    hyperbolic pair, and the shortcut taken on most exits."""
    from bia_3d_gpu_amd import ic
    if case == "fixture":
        st, eps, dt, step = fixture1332["state"], fixture1332["eps"], fixture1332["dt"], 7
    elif case == "benchmark":
        st, eps, dt, step = ic.plummer_with_binaries(2048, 65536), -2 / 2048, 1e-3, 3
    elif case == "strong":
        st = ic.uniform_ball_virial(200, seed=5)
        st[3:] *= 0.3
        eps, dt, step = -2 / 200, 5e-3, 1
    else:
        st = ic.uniform_ball_virial(200, seed=6)
        eps, dt, step = 0.05, 1e-2, 1  # repulsive: every pair hyperbolic, theta_0 < pi/2
    rel = _pairs(st)[::step]
    L = C.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    L.hs_hyp_loops.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_int, dp, dp, dp, ip, ip]
    m = len(rel)
    lib, ref = np.zeros((m, 4)), np.zeros((m, 4))
    hyp, rep = np.zeros(m, np.int32), np.zeros(m, np.int32)
    L.hs_hyp_loops(m, eps, dt, 18, np.ascontiguousarray(rel), lib, ref, hyp, rep)
    h = hyp.astype(bool)
    cyc = (ref[h, 3].astype(int) & 2) != 0
    print(f"{case}: {h.sum()} hyperbolic pairs, {cyc.sum()} exact-repeat exits, {rep[h].sum()} on a repeated "
          f"anomaly; evaluations {ref[h, 2].sum():.0f}")
    assert h.sum() > 1000
    assert np.array_equal(lib[h].view(np.int64), ref[h].view(np.int64))
    assert rep[h].sum() > 0.5 * cyc.sum()
