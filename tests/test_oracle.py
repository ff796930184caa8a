"""Pins of the oracle (oracle/bia_oracle.c) before it is trusted as the checker.

* the reference's own outputs in OUT/ (N = 1332 BIA deltas, DIM final state, the DIM
  accepted-step trajectory of particle 1);
* the reference's own Fortran CPU routines compiled from the reference checkout
  (oracle/_ref, built by oracle/Makefile) where that build exists;
* what the mathematics fixes: conservation of the two-body energy and angular momentum,
  uniform rotation of a circular orbit, the straight-line limit, time reversal, the analytic
  solution against the independent numerical RKF45 on the same 2D problem, an exactly
  parabolic orbit, and the quad-precision potential sum against 50-digit arithmetic.
"""
import os

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference checkout absent)")


def W_L(y, eps):
    r = np.hypot(y[0], y[1])
    return y[2] ** 2 + y[3] ** 2 + 2 * eps / r, y[0] * y[3] - y[1] * y[2]


# --------------------------------------------------------------------------------------------
# reference fixtures
# --------------------------------------------------------------------------------------------
def test_step_matches_reference_BIA_fixture(fixture1332):
    st, eps, dt = fixture1332["state"], fixture1332["eps"], fixture1332["dt"]
    d = O.bia_step(st, eps, dt)
    ref = fixture1332["bia"]
    assert np.abs(d[3:] - ref[3:]).max() < 1e-14
    # position deltas: the reference's GPU (libdevice) and CPU (glibc) paths already differ by
    # 2.1e-11 on this fixture (acos near 1 for pairs starting near periapsis)
    assert np.abs(d[:3] - ref[:3]).max() < 3e-11
    # particle 1 (at rest at the origin) in the reference's text output
    line = O and np.array([float(x) for x in open(os.path.join(GOLDEN, "BIA_10^0.0001332")).read().split()])
    assert abs(line[0] - 1.0) == 0
    assert np.abs((st[:, 0] + d[:, 0]) - line[1:]).max() < 3e-11


@needs_ref
def test_step_matches_reference_cpu_routines(fixture1332):
    st, eps, dt = fixture1332["state"], fixture1332["eps"], fixture1332["dt"]
    a = O.bia_step(st, eps, dt)
    b = O.ref_cpu_rows(st, eps, dt)
    assert np.abs(a[3:] - b[3:]).max() < 1e-18
    assert np.abs(a[:3] - b[:3]).max() < 1e-12


def test_dim_matches_reference_fixture(fixture1332):
    st = fixture1332["state"]
    y = np.ascontiguousarray(st.T).reshape(-1)
    r = O.dim_rkf45(y, fixture1332["eps_dim"], 0.0, 1.0)
    assert r["ier"] == 0
    yn = r["y"].reshape(-1, 6).T
    ref = fixture1332["dim"]
    assert np.abs(yn[:3] - ref[:3]).max() < 1e-13 and np.abs(yn[3:] - ref[3:]).max() < 1e-14
    steps = np.loadtxt(os.path.join(GOLDEN, "0001332.DIM"), skiprows=1)
    assert r["naccept"] == len(steps) - 1 == 238  # the step-size controller's sequence
    assert np.abs(r["step_t"] - steps[1:, 0]).max() < 1e-7
    assert np.abs(r["step_y1"] - steps[1:, 1:]).max() < 1e-13


# --------------------------------------------------------------------------------------------
# against the reference's own CPU routines on random pairs
# --------------------------------------------------------------------------------------------
def _random_pairs(rng, count, eps):
    out = []
    for _ in range(count):
        r = 10 ** rng.uniform(-3, 0.5)
        g = np.sqrt(abs(eps) / r) * rng.uniform(0.05, 1.8)
        ang = rng.uniform(0.1, np.pi - 0.1)
        _, _, yz = O.frame(np.array([r, 0, 0.0]), g * np.array([np.cos(ang), np.sin(ang), 0]))
        out.append(yz)
    return out


@needs_ref
def test_solvers_match_reference_routines_where_both_converge():
    rng = np.random.default_rng(3)
    eps = -6e-7
    agree = {"ellipse": 0, "hyperbola": 0}
    total = {"ellipse": 0, "hyperbola": 0}
    for yz in _random_pairs(rng, 600, eps):
        Wr = np.hypot(yz[0], yz[1]) * yz[3] ** 2 + 2 * eps
        kind = "ellipse" if Wr < 0 else "hyperbola"
        dt = 10 ** rng.uniform(-3, -1)
        yn, it, fl = O.solve2d(kind, eps, dt, yz)
        yr, err, ir, ier = O.ref_solve2d(kind, eps, dt, yz)
        if fl & (O.F_ITR_CAP | O.F_STOCK_DEFAULT) or ir < 0:
            continue  # one of the two iteration schemes did not converge
        total[kind] += 1
        rel = max(np.abs(yn[:2] - yr[:2]).max() / np.hypot(yz[0], yz[1]), np.abs(yn[2:] - yr[2:]).max() / yz[3])
        agree[kind] += rel < 1e-10
    for k in total:
        assert total[k] > 50
        assert agree[k] >= 0.99 * total[k], (k, agree[k], total[k])


@needs_ref
def test_rkf_matches_reference_routine():
    rng = np.random.default_rng(4)
    eps = -6e-7
    for yz in _random_pairs(rng, 40, eps):
        dt = 10 ** rng.uniform(-3, -1)
        yn, steps, fl = O.solve2d("rkf", eps, dt, yz)
        yr, err, _, ier = O.ref_solve2d("rkf", eps, dt, yz)
        if fl or ier:
            continue
        assert np.abs(yn - yr).max() <= 1e-15 * max(1.0, np.abs(yr).max()), (yz, yn, yr)



def test_rkf_tolerance_variant():
    """orc_rkf_pair_tol (the per-pair RKF45 at a caller's relerr / abserr) is bitwise the
    reference-tolerance RKF45 at 0 / 1e-16; at relerr 1e-12 it takes fewer accepted steps on every
    pair that needs more than the controller's fewest, never more, and stays within 1e-10 of the conic solution where the reference's Newton converges (an error-norm or
    controller mistake in the variant would break the first check or the last)."""
    rng = np.random.default_rng(12)
    eps = -6e-7
    fewer, compared = 0, 0
    for yz in _random_pairs(rng, 40, eps):
        dt = 10 ** rng.uniform(-3, -1)
        ya, sa, fa = O.rkf2d_tol(eps, dt, yz, 0.0, 1e-16)
        yr, sr, fr = O.solve2d("rkf", eps, dt, yz)
        assert np.array_equal(ya, yr) and sa == sr and fa == fr
        yt, st_, ft = O.rkf2d_tol(eps, dt, yz, 1e-12, 1e-16)
        assert st_ <= sr
        if sr > 5:  # more than the fewest steps the controller takes over dt
            assert st_ < sr
            fewer += 1
        W = np.hypot(yz[0], yz[1]) * (yz[2] ** 2 + yz[3] ** 2) + 2 * eps
        yc, _, fc = O.solve2d("hyperbola" if W > 0 else "ellipse", eps, dt, yz)
        if fc or ft:
            continue
        compared += 1
        scale = max(np.abs(yc).max(), 1e-300)
        assert np.abs(yt - yc).max() <= 1e-10 * scale, (yz, dt, yt, yc)
    assert fewer >= 5 and compared >= 20, (fewer, compared)

# --------------------------------------------------------------------------------------------
# mathematics
# --------------------------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["ellipse", "hyperbola"])
def test_energy_and_angular_momentum_conserved(kind):
    rng = np.random.default_rng(5)
    eps = -1.0
    n = 0
    while n < 200:
        r = rng.uniform(0.5, 2)
        vc = np.sqrt(-eps / r)
        g = vc * (rng.uniform(0.3, 1.3) if kind == "ellipse" else rng.uniform(1.5, 4))
        ang = rng.uniform(0.2, np.pi - 0.2)
        _, _, yz = O.frame(np.array([r, 0, 0.0]), g * np.array([np.cos(ang), np.sin(ang), 0]))
        yn, it, fl = O.solve2d(kind, eps, rng.uniform(0.05, 2.0), yz)
        if fl & (O.F_ITR_CAP | O.F_STOCK_DEFAULT):
            continue
        w0, l0 = W_L(yz, eps)
        w1, l1 = W_L(yn, eps)
        assert abs(w1 - w0) < 1e-11 * abs(w0) + 1e-12
        assert abs(l1 - l0) < 1e-11 * abs(l0)
        n += 1


def test_circular_orbit_is_uniform_rotation():
    eps, r = -2.0, 1.5
    v = np.sqrt(-eps / r)
    for dt in (0.1, 0.7, 3.0):
        yn, it, fl = O.solve2d("ellipse", eps, dt, np.array([r, 0.0, 0.0, v]))
        ph = v / r * dt
        exact = np.array([r * np.cos(ph), r * np.sin(ph), -v * np.sin(ph), v * np.cos(ph)])
        assert np.abs(yn - exact).max() < 1e-12, (dt, yn, exact)


def test_straight_line_limit():
    """Weak coupling: the deviation from straight-line motion is linear in eps."""
    yz = np.array([1.3, -0.4, 0.0, 0.9])
    for dt in (0.1, 1.0):
        exact = np.array([yz[0], yz[1] + yz[3] * dt, 0, yz[3]])
        dev = {}
        for eps in (-1e-14, -1e-10, -1e-8):
            yn, it, fl = O.solve2d("hyperbola", eps, dt, yz)
            dev[eps] = np.abs(yn - exact).max()
        assert dev[-1e-14] < 1e-13
        assert 80 < dev[-1e-8] / dev[-1e-10] < 120


def test_reference_quirk_best_iterate_never_updated():
    """At eps = -1e-18 the dimensionless time is ~1e18 and no Newton residual improves on the
    initial one: the reference returns theta = 0 (find_dXYUV_kernel.cu:189-217), i.e. the
    periapsis state.  The oracle keeps that behaviour and reports it."""
    yz = np.array([1.3, -0.4, 0.0, 0.9])
    yn, it, fl = O.solve2d("hyperbola", -1e-18, 0.1, yz)
    assert fl & O.F_STOCK_DEFAULT
    assert abs(yn[1]) < 1e-12  # at periapsis instead of y = -0.31


def _co_failure(kind, eps, dt, yz):
    """True when the reference's own CPU routine also fails on this pair (IR = -1: the Newton
    iterate left (-theta_0, theta_0); or the iteration cap)."""
    if not O.ref_available():
        return True
    yf, err, ir, ier = O.ref_solve2d(kind, eps, dt, yz)
    return ir < 0 or ir >= 18


@pytest.mark.parametrize("kind,gscale", [("ellipse", 0.8), ("hyperbola", 2.5)])
def test_time_reversal(kind, gscale):
    eps = -1.0
    ok = 0
    for ang in (0.4, 1.0, 1.2, 2.3):
        r = 1.1
        g = gscale * np.sqrt(-eps / r)
        _, _, yz = O.frame(np.array([r, 0, 0.0]), g * np.array([np.cos(ang), np.sin(ang), 0]))
        yn, _, f1 = O.solve2d(kind, eps, 0.3, yz)
        # reverse the velocity and reflect y -> -y, which keeps L > 0 as the BIA frame always
        # does (e_zeta is taken along L); then undo both on the result
        yb = np.array([yn[0], -yn[1], -yn[2], yn[3]])
        back, _, f2 = O.solve2d(kind, eps, 0.3, yb)
        back = np.array([back[0], -back[1], -back[2], back[3]])
        good = np.abs(back - yz).max() < 1e-11
        if not good:  # only where the reference's iteration itself fails
            assert (f1 | f2) & O.F_ITR_CAP and (_co_failure(kind, eps, 0.3, yz) or _co_failure(kind, eps, 0.3, yb))
        ok += good
    assert ok >= 3


@pytest.mark.parametrize("kind,gscale", [("ellipse", 0.7), ("ellipse", 1.2), ("hyperbola", 1.6), ("hyperbola", 5.0)])
def test_analytic_matches_independent_rkf45(kind, gscale):
    """ALG (closed-form conic + Newton) against the adaptive numerical integration of the same
    2D equation of motion: independent formulations, agreement at the 1e-11 level wherever the
    reference's Newton iteration converges; every disagreement is a pair on which the
    reference's own CPU routine fails as well."""
    eps = -1.0
    agree = 0
    for ang in (0.3, 0.4, 1.0, 1.2, 1.9, 2.3, 2.7):
        r = 0.9
        g = gscale * np.sqrt(-eps / r)
        _, _, yz = O.frame(np.array([r, 0, 0.0]), g * np.array([np.cos(ang), np.sin(ang), 0]))
        dt = 0.25
        ya, _, fa = O.solve2d(kind, eps, dt, yz)
        yr, _, fr = O.solve2d("rkf", eps, dt, yz)
        assert fr == 0
        good = np.abs(ya - yr).max() < 1e-11 * max(1, np.abs(yr).max())
        if not good:
            assert fa & O.F_ITR_CAP and _co_failure(kind, eps, dt, yz), (kind, ang)
        agree += good
    assert agree >= 4


def test_exact_parabola():
    # r = 5 (3-4-5), g = 0.5, eps = -r g^2 / 2 = -0.625: r g^2 + 2 eps == 0 exactly
    yz = np.array([3.0, 4.0, 0.0, 0.5])
    eps = -0.625
    yn, it, fl = O.solve2d("parabola", eps, 0.8, yz)
    w, l = W_L(yn, eps)
    assert abs(w) < 1e-14 and abs(l - 1.5) < 1e-13
    yr, _, fr = O.solve2d("rkf", eps, 0.8, yz)
    assert np.abs(yn - yr).max() < 1e-11
    _, br, _, _ = O.pairs2d(1, eps, 0.8, yz[None])
    assert br[0] == O.PARABOLA


def test_half_split_and_frame_epilogue():
    """3D pair delta: velocity deltas of i and j are equal and opposite halves of the relative
    velocity change, and the position delta is half the deviation from straight-line motion."""
    rel = np.array([0.7, -0.2, 0.3, 0.1, 0.4, -0.5])
    eps, dt = -0.05, 0.4
    out, br, it, fl = O.pair(1, eps, dt, rel)
    exi, eeta, yz = O.frame(rel[:3], rel[3:])
    yn, _, _ = O.solve2d("hyperbola" if br == O.HYPERBOLA else "ellipse", eps, dt, yz)
    r_end = yn[0] * exi + yn[1] * eeta
    g_end = yn[2] * exi + yn[3] * eeta
    assert np.allclose(out[:3], (r_end - rel[:3] - rel[3:] * dt) / 2, atol=1e-15)
    assert np.allclose(out[3:], (g_end - rel[3:]) / 2, atol=1e-15)
    # relative 3D state after dt is consistent with the 3D two-body problem (energy, L vector)
    g0 = np.dot(rel[3:], rel[3:]) + 2 * eps / np.linalg.norm(rel[:3])
    g1 = np.dot(g_end, g_end) + 2 * eps / np.linalg.norm(r_end)
    assert abs(g1 - g0) < 1e-13
    assert np.allclose(np.cross(rel[:3], rel[3:]), np.cross(r_end, g_end), atol=1e-14)


# --------------------------------------------------------------------------------------------
# invariance
# --------------------------------------------------------------------------------------------
def test_quad_potential_against_50_digits():
    rng = np.random.default_rng(6)
    st = rng.normal(size=(6, 40))
    q = O.invariance(st, -1.0, "quad")
    mp.mp.dps = 50
    s = mp.mpf(0)
    for i in range(40):
        for j in range(i + 1, 40):
            s += 1 / mp.sqrt(sum((mp.mpf(st[c, i]) - mp.mpf(st[c, j])) ** 2 for c in range(3)))
    assert abs(q[1] + float(s)) <= 1e-16 * float(s)
    ke = sum(mp.mpf(st[c, i]) ** 2 for c in (3, 4, 5) for i in range(40))
    assert abs(q[0] - float(ke)) <= 1e-16 * float(ke)
    lz = sum(mp.mpf(st[0, i]) * mp.mpf(st[4, i]) - mp.mpf(st[1, i]) * mp.mpf(st[3, i]) for i in range(40))
    assert abs(q[4] - float(lz)) <= 1e-15 * max(1, abs(float(lz)))


@needs_ref
def test_invariance_matches_reference_routines(fixture1332):
    st, eps = fixture1332["state"], fixture1332["eps"]
    a = O.invariance(st, eps, "cpu")
    b = O.ref_invariance(st, eps, quad=False)
    assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(1, np.abs(b)))
    q = O.invariance(st, eps, "quad")
    qr = O.ref_invariance(st, eps, quad=True)
    assert np.all(np.abs(q - qr) <= 2e-16 * np.abs(qr) + 1e-300)
    g = O.invariance(st, eps, "gpu")
    assert abs(g[1] / q[1] - 1) < 1e-13 and abs(g[0] / q[0] - 1) < 1e-14


def test_fixture_energy_drift_of_reference_bia(fixture1332):
    """|dE/E0| of the reference's own BIA step on its fixture (method property at dt = 1)."""
    st, eps = fixture1332["state"], fixture1332["eps"]
    e0 = O.invariance(st, eps, "quad")
    e1 = O.invariance(st + fixture1332["bia"], eps, "quad")
    rel = abs((e1[0] + e1[1]) - (e0[0] + e0[1])) / abs(e0[0] + e0[1])
    assert 1e-13 < rel < 1e-12  # 4.4e-13


def test_sensitivity_bound_covers_the_reference_itself(fixture1332):
    """The parity bound used for the GPU (10x the formulas' sensitivity to one rounding unit)
    accepts the reference's own CUDA output against the glibc evaluation of the same formulas."""
    st, eps, dt = fixture1332["state"], fixture1332["eps"], fixture1332["dt"]
    ok, rp, rv = O.within_condition_bound(st, fixture1332["bia"], O.bia_step(st, eps, dt), eps, dt)
    assert ok and rp > 0.05  # the bound is tight enough to be exercised by the reference itself


def test_coupled_step_reduces_to_reference_for_equal_masses():
    from bia_3d_gpu_amd import ic
    st = ic.uniform_ball_virial(150, seed=3)
    a = O.bia_step(st, -2 / 150, 1e-3)
    b = O.bia_step_coupled(st, 1, 1.0, 1e-3, mass=np.full(150, 1 / 150))
    assert np.array_equal(a, b)


def test_coupled_step_conserves_momentum_and_centre_of_mass():
    from bia_3d_gpu_amd import ic
    n = 150
    st = ic.uniform_ball_virial(n, seed=4)
    m = np.random.default_rng(1).uniform(0.1, 10, n)
    m /= m.sum()
    d = O.bia_step_coupled(st, 1, 1.0, 1e-3, mass=m)
    assert np.abs((m * d[3:]).sum(1)).max() < 1e-15 * np.abs(m * d[3:]).sum()
    # sum m dx = sum m v dt (the pair terms cancel by the centre-of-mass split)
    drift = (m * st[3:]).sum(1) * 1e-3
    assert np.abs((m * d[:3]).sum(1) - drift).max() < 1e-15
    # Coulomb, unit masses: charges +-1, momentum conserved, like charges repel
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    dc = O.bia_step_coupled(st, 2, 1e-4, 1e-3, charge=q)
    assert np.abs(dc[3:].sum(1)).max() < 1e-15 * np.abs(dc[3:]).sum()


# --------------------------------------------------------------------------------------------
# two-body step in quadruple precision (orc_pair_quad / orc_rows_quad): the accuracy oracle of
# the pair propagators below 1e-12
# --------------------------------------------------------------------------------------------
def _quad_pairs(rng, m):
    """Pairs of every kind: weak distant ones, bound over whole periods, fast flybys, repulsive."""
    out = []
    for k in range(m):
        r = rng.standard_normal(3) * 10.0 ** rng.uniform(-2, 1)
        v = rng.standard_normal(3) * 10.0 ** rng.uniform(-2, 0.5)
        eps = -10.0 ** rng.uniform(-5, 0) * (1 if k % 5 else -1)
        dt = 10.0 ** rng.uniform(-3, 0.5)
        out.append((np.concatenate([r, v]), eps, dt))
    return out


def test_quad_pair_against_80_digits():
    """Every pair's quad half deltas (hi + lo) against the 80-digit closed-form solution to 1e-30
    of the pair's largest delta: a dropped term, a wrong sign or a lost digit of the Stumpff
    functions would show at 1e-16 or worse."""
    mp = pytest.importorskip("mpmath")
    from test_hostsim import _drift_closed_form
    rng = np.random.default_rng(1201)
    worst = 0.0
    for rel, eps, dt in _quad_pairs(rng, 40):
        hi, lo = O.pair_quad(eps, dt, rel)
        with mp.workdps(80):
            ex = _drift_closed_form(rel, eps, dt, exact=True)
            err = max(abs(mp.mpf(float(h)) + mp.mpf(float(l)) - e) for h, l, e in zip(hi, lo, ex))
            worst = max(worst, float(err / max(abs(e) for e in ex)))
    print(f"quad pair vs 80 digits: largest error {worst:.1e} of the pair's largest delta")
    assert worst < 1e-30


def test_quad_pair_conserves_energy_and_angular_momentum():
    """r1 = r0 + v0 dt + 2 d_r and v1 = v0 + 2 d_v from the quad deltas keep the relative energy
    v^2/2 - mu/r and the angular momentum r x v to 1e-28 (in 50 digits)."""
    mp = pytest.importorskip("mpmath")
    rng = np.random.default_rng(1202)
    with mp.workdps(50):
        for rel, eps, dt in _quad_pairs(rng, 30):
            hi, lo = O.pair_quad(eps, dt, rel)
            d = [mp.mpf(float(h)) + mp.mpf(float(l)) for h, l in zip(hi, lo)]
            r0 = [mp.mpf(float(c)) for c in rel[:3]]
            v0 = [mp.mpf(float(c)) for c in rel[3:]]
            r1 = [r0[q] + v0[q] * mp.mpf(dt) + 2 * d[q] for q in range(3)]
            v1 = [v0[q] + 2 * d[3 + q] for q in range(3)]
            mu = -mp.mpf(eps)
            e0 = sum(c * c for c in v0) / 2 - mu / mp.sqrt(sum(c * c for c in r0))
            e1 = sum(c * c for c in v1) / 2 - mu / mp.sqrt(sum(c * c for c in r1))
            l0 = [r0[1] * v0[2] - r0[2] * v0[1], r0[2] * v0[0] - r0[0] * v0[2], r0[0] * v0[1] - r0[1] * v0[0]]
            l1 = [r1[1] * v1[2] - r1[2] * v1[1], r1[2] * v1[0] - r1[0] * v1[2], r1[0] * v1[1] - r1[1] * v1[0]]
            escale = sum(c * c for c in v0) / 2 + abs(mu) / mp.sqrt(sum(c * c for c in r0))
            lscale = mp.sqrt(sum(c * c for c in r0) * sum(c * c for c in v0))
            assert abs(e1 - e0) < 1e-28 * escale, (rel, eps, dt)
            assert max(abs(a - b) for a, b in zip(l0, l1)) < 1e-28 * lscale, (rel, eps, dt)


def test_quad_rows_against_pairs_and_double_oracle():
    """orc_rows_quad on a 48-particle Plummer system with hard binaries: each row equals the
    signed sum of the quad pair deltas over its partners (each pair oriented i < j, drift added),
    and agrees with the double-precision oracle's whole step (the reference algorithm) to its
    rounding; a row with the partners' signs or orientation wrong would not."""
    from bia_3d_gpu_amd import ic
    mp = pytest.importorskip("mpmath")
    st = ic.plummer_with_binaries(48, seed=4801)
    n, eps, dt = st.shape[1], -2.0 / 48, 1e-3
    rows = [0, 1, 17, 46, 47]
    hi, lo, d = O.rows_quad(st, rows, eps, dt)
    with mp.workdps(40):
        for k, p in enumerate(rows):
            acc = [mp.mpf(0)] * 6
            for j in range(n):
                if j == p:
                    continue
                i0, j0 = min(p, j), max(p, j)
                ph, pl = O.pair_quad(eps, dt, st[:, i0] - st[:, j0])
                sg = 1 if p < j else -1
                acc = [a + sg * (mp.mpf(float(h)) + mp.mpf(float(l))) for a, h, l in zip(acc, ph, pl)]
            for c in range(6):
                got = mp.mpf(float(hi[c, k])) + mp.mpf(float(lo[c, k]))
                assert abs(got - acc[c]) <= 1e-30 * max(abs(x) for x in acc), (p, c)
                want = acc[c] + (mp.mpf(float(st[3 + c, p])) * mp.mpf(dt) if c < 3 else 0)
                assert d[c, k] == float(want)
    # every row against the double oracle's whole step: the reference algorithm is off the exact
    # solution by its own rounding sensitivity (up to 1e-10 of the largest delta near the binaries
    # here), so the comparison is the per-particle condition bound of the parity tests
    _, _, dall = O.rows_quad(st, list(range(n)), eps, dt)
    orc = O.bia_step(st, eps, dt)
    ok, rp, rv = O.within_condition_bound(st, orc, dall, eps, bounds=O.sensitivity(st, eps, dt))
    print(f"double oracle vs quad rows: error / allowed positions {rp:.2f}, velocities {rv:.2f}")
    assert ok


def test_config1_first_step_oracle_vs_reference_cpu():
    """Config 1 (seeded N = 1332 uniform ball at virial ratio 0.5, G = 1, m = 1/N, eps = -2/N,
    dt = 1e-3): the oracle's first step (GPU semantics) against the reference's own CPU ALG loop
    (bia.f90:142-196, amdflang build) within the per-particle condition bound."""
    if not O.ref_available():
        pytest.skip("reference CPU routines not built (oracle/_ref)")
    from bia_3d_gpu_amd import ic
    st = ic.uniform_ball_virial(1332, seed=1332)
    eps, dt = -2.0 / 1332, 1e-3
    orc = O.bia_step(st, eps, dt)
    ref = O.ref_cpu_rows(st, eps, dt)
    ok, rp, rv = O.within_condition_bound(st, ref, orc, eps, bounds=O.sensitivity(st, eps, dt))
    print(f"config 1 first step, reference CPU vs oracle: error / allowed {rp:.2f}, {rv:.2f}")
    assert ok
