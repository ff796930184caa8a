"""Coefficients of the plain-mode kernel's range-specific sin, cos and log (bia_pair.hpp,
sincos_small and log_pos), computed for this library by the Remez exchange algorithm in 60-digit
arithmetic (mpmath).  Run it to print the C constants the header holds:

    python3 bia_3d_gpu_amd/csrc/minimax_coefficients.py

Each polynomial p of degree n in z minimises max |w(z) (p(z) - f(z))| on [0, zmax]:

* sin b = b + b^3 p_s(b^2), |b| <= pi/4:  f_s(z) = (sin(sqrt z) - sqrt z) / z^(3/2), w = z
  (the error of sin b relative to b);
* cos b = 1 - b^2/2 + b^4 p_c(b^2), |b| <= pi/4:  f_c(z) = (cos(sqrt z) - 1 + z/2) / z^2,
  w = z^2;
* ln m = 2 atanh(u) = 2u + 2u v p_l(v), v = u^2, |u| <= (sqrt 2 - 1)/(sqrt 2 + 1):
  f_l(v) = (atanh(sqrt v) - sqrt v) / v^(3/2), w = v.

The exchange: start from Chebyshev points, solve p(x_i) - f(x_i) = (-1)^i E / w(x_i) at n + 2
points, move the points to the alternating extrema of the weighted error (dense sampling, then
golden-section refinement), repeat until E settles.  The weighted error is then checked in 60
digits for the coefficients rounded to double.
"""
from __future__ import annotations

import mpmath as mp

mp.mp.dps = 60


def remez(f, w, a, b, n, iters=30):
    """minimax p (coefficients low to high) of degree n for w (p - f) on [a, b]; returns (p, E)."""
    m = n + 2
    # Chebyshev extrema as the first reference (the end a excluded: the weight vanishes there)
    xs = [a + (b - a) * (1 - mp.cos(mp.pi * (k + 1) / (m + 0.5))) / 2 for k in range(m)]
    E = mp.mpf(0)
    coef = None
    for _ in range(iters):
        A = mp.matrix(m, m)
        rhs = mp.matrix(m, 1)
        for i, x in enumerate(xs):
            for j in range(n + 1):
                A[i, j] = x ** j
            A[i, n + 1] = (-1) ** i / w(x)
            rhs[i] = f(x)
        sol = mp.lu_solve(A, rhs)
        coef = [sol[j] for j in range(n + 1)]
        E = sol[n + 1]

        def err(x):
            return w(x) * (sum(c * x ** j for j, c in enumerate(coef)) - f(x))

        # alternating extrema of the weighted error on a dense grid, refined locally
        grid = [a + (b - a) * k / 4000 for k in range(1, 4001)]
        vals = [err(x) for x in grid]
        ext = []
        for k in range(len(grid)):
            left = vals[k - 1] if k > 0 else None
            right = vals[k + 1] if k + 1 < len(grid) else None
            v = vals[k]
            if (left is None or abs(v) >= abs(left)) and (right is None or abs(v) >= abs(right)):
                lo = grid[k - 1] if k > 0 else grid[k]
                hi = grid[k + 1] if k + 1 < len(grid) else grid[k]
                ext.append(golden_max(lambda t: abs(err(t)), lo, hi) if hi > lo else grid[k])
        # keep the largest alternating set of m points
        pts = []
        for x in ext:
            if pts and mp.sign(err(x)) == mp.sign(err(pts[-1])):
                if abs(err(x)) > abs(err(pts[-1])):
                    pts[-1] = x
            else:
                pts.append(x)
        while len(pts) > m:  # drop the smaller end
            if abs(err(pts[0])) < abs(err(pts[-1])):
                pts.pop(0)
            else:
                pts.pop()
        if len(pts) < m:
            break
        xs = pts
    return coef, E


def golden_max(g, lo, hi, steps=80):
    """the maximiser of a unimodal g on [lo, hi] by golden-section search"""
    r = (mp.sqrt(5) - 1) / 2
    x1, x2 = hi - r * (hi - lo), lo + r * (hi - lo)
    g1, g2 = g(x1), g(x2)
    for _ in range(steps):
        if g1 < g2:
            lo, x1, g1 = x1, x2, g2
            x2 = lo + r * (hi - lo)
            g2 = g(x2)
        else:
            hi, x2, g2 = x2, x1, g1
            x1 = hi - r * (hi - lo)
            g1 = g(x1)
    return (lo + hi) / 2


def f_sin(z):
    if z == 0:
        return mp.mpf(-1) / 6
    r = mp.sqrt(z)
    return (mp.sin(r) - r) / (z * r)


def f_cos(z):
    if z == 0:
        return mp.mpf(1) / 24
    return (mp.cos(mp.sqrt(z)) - 1 + z / 2) / (z * z)


def f_log(v):
    if v == 0:
        return mp.mpf(1) / 3
    r = mp.sqrt(v)
    return (mp.atanh(r) - r) / (v * r)


def check(f, w, a, b, coef):
    """max weighted error of the double-rounded coefficients (60 digits, dense grid)"""
    cd = [mp.mpf(float(c)) for c in coef]
    worst = mp.mpf(0)
    for k in range(1, 20001):
        x = a + (b - a) * k / 20000
        e = abs(w(x) * (sum(c * x ** j for j, c in enumerate(cd)) - f(x)))
        worst = max(worst, e)
    return worst


def main():
    zs = (mp.pi / 4) ** 2
    us = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)
    for name, f, w, b, n in (("sin", f_sin, lambda z: z, zs, 5), ("cos", f_cos, lambda z: z * z, zs, 5),
                             ("log", f_log, lambda v: v, us * us, 6)):
        coef, E = remez(f, w, mp.mpf(0), b, n)
        worst = check(f, w, mp.mpf(0), b, coef)
        print(f"// {name}: degree {n}, weighted minimax error {mp.nstr(abs(E), 3)}, "
              f"with double coefficients {mp.nstr(worst, 3)} (2^{mp.nstr(mp.log(worst, 2), 4)})")
        for j, c in enumerate(coef):
            print(f"constexpr double k{name.capitalize()}{j} = {float(c)!r};")


if __name__ == "__main__":
    main()
