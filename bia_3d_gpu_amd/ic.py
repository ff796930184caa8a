"""Seeded synthetic initial conditions for the benchmark configurations.

Units: G = 1 (gravity) or k = 1 (Coulomb).  States are structure-of-arrays ``[6, N]``.
For equal masses m the BIA pair coupling is eps = -2 G m (bia/bia.f90:37: twice the direct
N-body coupling, since the relative coordinate of a pair carries the reduced mass).
"""
from __future__ import annotations

import numpy as np


def _isotropic(rng, n):
    v = rng.normal(size=(3, n))
    return v / np.linalg.norm(v, axis=0)


def potential_energy(pos: np.ndarray, m: np.ndarray, G: float = 1.0, chunk: int = 2048) -> float:
    """W = -G sum_{i<j} m_i m_j / r_ij (blocked O(N^2) numpy, float64)."""
    n = pos.shape[1]
    w = 0.0
    for a in range(0, n, chunk):
        pa = pos[:, a:a + chunk]
        d = pa[:, :, None] - pos[:, None, :]
        r = np.sqrt((d * d).sum(0))
        idx = np.arange(a, min(a + chunk, n))
        r[np.arange(len(idx)), idx] = np.inf
        w += (m[a:a + chunk, None] * m[None, :] / r).sum()
    return -0.5 * G * w


def uniform_ball_virial(n: int, seed: int, q: float = 0.5, G: float = 1.0) -> np.ndarray:
    """Config 1: positions uniform in the unit ball, isotropic Gaussian velocities rescaled to
    virial ratio T/|W| = q, equal masses 1/n."""
    rng = np.random.default_rng(seed)
    r = rng.random(n) ** (1.0 / 3.0)
    pos = _isotropic(rng, n) * r
    vel = rng.normal(size=(3, n))
    m = np.full(n, 1.0 / n)
    vel -= (vel * m).sum(1, keepdims=True) / m.sum()
    T = 0.5 * (m * (vel ** 2).sum(0)).sum()
    W = potential_energy(pos, m, G)
    vel *= np.sqrt(q * abs(W) / T)
    return np.ascontiguousarray(np.vstack([pos, vel]))


def plummer(n: int, rng, a: float = 1.0, M: float = 1.0, G: float = 1.0, rmax: float = 10.0) -> np.ndarray:
    """Plummer sphere (scale radius a, mass M) by Aarseth-Henon-Wielen sampling, truncated at
    r < rmax * a.  Returns [6, n]."""
    out = np.zeros((6, n))
    k = 0
    while k < n:
        need = n - k
        x1 = rng.random(need)
        r = a / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
        ok = np.isfinite(r) & (r < rmax * a)
        r = r[ok]
        cnt = r.size
        pos = _isotropic(rng, cnt) * r
        # speed from the distribution function by rejection: g(q) = q^2 (1-q^2)^3.5
        qs = np.empty(cnt)
        filled = 0
        while filled < cnt:
            qq = rng.random(2 * (cnt - filled))
            gg = rng.random(2 * (cnt - filled)) * 0.1
            acc = qq[gg < qq ** 2 * (1 - qq ** 2) ** 3.5]
            take = min(acc.size, cnt - filled)
            qs[filled:filled + take] = acc[:take]
            filled += take
        vesc = np.sqrt(2 * G * M) * (r * r + a * a) ** -0.25
        vel = _isotropic(rng, cnt) * qs * vesc
        out[:3, k:k + cnt] = pos
        out[3:, k:k + cnt] = vel
        k += cnt
    out[:3] -= out[:3].mean(1, keepdims=True)
    out[3:] -= out[3:].mean(1, keepdims=True)
    return out


def plummer_with_binaries(n: int, seed: int, frac: float = 0.2, amin: float = 1e-3, amax: float = 1e-2,
                          G: float = 1.0) -> np.ndarray:
    """Config 4: Plummer sphere (a = 1, M = 1) of n equal-mass particles (m = 1/n) in which a
    fraction ``frac`` of the particles are paired into hard binaries: semi-major axis
    log-uniform in [amin, amax], thermal eccentricities f(e) = 2e, random orientation and
    phase, centred on a Plummer-sampled centre of mass."""
    rng = np.random.default_rng(seed)
    m = 1.0 / n
    nb = int(round(frac * n)) // 2
    ncm = n - nb  # single stars + binary centres of mass
    cm = plummer(ncm, rng, G=G)
    st = np.zeros((6, n))
    st[:, :ncm - nb] = cm[:, :ncm - nb]
    mu = G * 2 * m
    sma = np.exp(rng.uniform(np.log(amin), np.log(amax), nb))
    ecc = np.sqrt(rng.random(nb))
    M_anom = rng.uniform(0, 2 * np.pi, nb)
    E = M_anom.copy()
    for _ in range(50):  # Kepler's equation by Newton (e < 1)
        E -= (E - ecc * np.sin(E) - M_anom) / (1 - ecc * np.cos(E))
    rx = sma * (np.cos(E) - ecc)
    ry = sma * np.sqrt(1 - ecc ** 2) * np.sin(E)
    rr = sma * (1 - ecc * np.cos(E))
    nmean = np.sqrt(mu / sma ** 3)
    vx = -sma * nmean * np.sin(E) * sma / rr
    vy = sma * nmean * np.sqrt(1 - ecc ** 2) * np.cos(E) * sma / rr
    # random orientation: rotate the orbital plane by a random rotation matrix
    rel = np.zeros((6, nb))
    for b in range(nb):
        qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        rel[:3, b] = qm @ np.array([rx[b], ry[b], 0.0])
        rel[3:, b] = qm @ np.array([vx[b], vy[b], 0.0])
    c = cm[:, ncm - nb:]
    s0 = ncm - nb
    st[:, s0:s0 + nb] = c + 0.5 * rel
    st[:, s0 + nb:s0 + 2 * nb] = c - 0.5 * rel
    perm = rng.permutation(n)
    st = st[:, perm]
    st[:3] -= st[:3].mean(1, keepdims=True)
    st[3:] -= st[3:].mean(1, keepdims=True)
    return np.ascontiguousarray(st)


def plummer_system(n: int, seed: int, G: float = 1.0) -> np.ndarray:
    """Config 3: one Plummer sphere (a = 1, M = 1, m = 1/n) from its own seed."""
    return np.ascontiguousarray(plummer(n, np.random.default_rng(seed), G=G))


def ensemble_sizes(count: int, nmin: int, nmax: int, seed: int) -> np.ndarray:
    """N per system, log-uniform in [nmin, nmax] (configs 3 and 5)."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(np.int64)


def coulomb_box(n: int, seed: int, bound_frac: float = 0.4, k: float = 1.0, sample: int = 200_000):
    """Config 2: n unit masses with charges +-1 (n/2 each, shuffled) uniform in the unit cube;
    Gaussian velocities whose sigma is bisected so that ``bound_frac`` of sampled pairs are
    bound (g^2 + 2 eps_ij / r < 0, eps_ij = k q_i q_j (1/m_i + 1/m_j)).
    Returns (state [6, n], charges [n])."""
    rng = np.random.default_rng(seed)
    pos = rng.random((3, n))
    q = np.ones(n)
    q[: n // 2] = -1
    rng.shuffle(q)
    base = rng.normal(size=(3, n))
    i = rng.integers(0, n, sample)
    j = rng.integers(0, n, sample)
    ok = i != j
    i, j = i[ok], j[ok]
    r = np.linalg.norm(pos[:, i] - pos[:, j], axis=0)
    eps = k * q[i] * q[j] * 2.0
    g2u = ((base[:, i] - base[:, j]) ** 2).sum(0)

    def frac(sig):
        return np.mean(sig * sig * g2u + 2 * eps / r < 0)

    lo, hi = 1e-6, 1e3
    for _ in range(200):
        mid = np.sqrt(lo * hi)
        if frac(mid) > bound_frac:
            lo = mid
        else:
            hi = mid
    sig = np.sqrt(lo * hi)
    vel = base * sig
    vel -= vel.mean(1, keepdims=True)
    return np.ascontiguousarray(np.vstack([pos, vel])), q
