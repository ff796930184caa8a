"""Constants of the plain-mode kernel's range-specific sin, cos and log (bia_pair.hpp): the
coefficients of log_pos's polynomial, computed for this library by the Remez exchange algorithm in
60-digit arithmetic (mpmath), the same for the arctangent and arcsine polynomials of atan_ratio and
acos_ref, and sincos_pi's table of sin and cos at k pi / 64 with its reduction constants.  Run it to print the C constants the header holds:

    python3 bia_3d_gpu_amd/csrc/minimax_coefficients.py

sincos_pi: a = |x| <= pi is reduced to r = a - k pi/64, |r| <= pi/128, k = rint(a 64/pi), with
pi/64 = P1 + P2 + P3 (P1 to 46 bits, so k P1 is exact for k <= 64); sin a and cos a follow from the
table's sin and cos of k pi/64, each as a pair of doubles (hi + lo to 2^-106), and the Taylor
polynomials of sin r and cos r - 1 (through r^7 and r^6: the first terms left out are below 4e-19
and 4e-18 of the result).

Each polynomial p of degree n in z minimises max |w(z) (p(z) - f(z))| on [0, zmax]:

* ln m = 2 atanh(u) = 2u + 2u v p_l(v), v = u^2, |u| <= (sqrt 2 - 1)/(sqrt 2 + 1):
  f_l(v) = (atanh(sqrt v) - sqrt v) / v^(3/2), w = v (the error relative to 2u);
* atan v = v + v z p_t(z), z = v^2, |v| <= tan(pi/8): f_t(z) = (atan(sqrt z) - sqrt z) / z^(3/2),
  w = z (relative to v);
* asin w = w + w z p_a(z), z = w^2, 0 <= w <= 1/2: f_a(z) = (asin(sqrt z) - sqrt z) / z^(3/2),
  w = z (relative to w).

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


def f_atan(z):
    if z == 0:
        return mp.mpf(-1) / 3
    r = mp.sqrt(z)
    return (mp.atan(r) - r) / (z * r)


def f_asin(z):
    if z == 0:
        return mp.mpf(1) / 6
    r = mp.sqrt(z)
    return (mp.asin(r) - r) / (z * r)


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


def sincos_table():
    """sin and cos of k pi / 64, k = 0..64, each as hi + lo (hi the double nearest, lo the double
    nearest the remainder), and the reduction constants"""
    pi64 = mp.pi / 64
    e = int(mp.floor(mp.log(pi64, 2)))
    scale = mp.mpf(2) ** (45 - e)
    p1 = mp.floor(pi64 * scale) / scale  # 46 significant bits
    assert mp.mpf(float(p1)) == p1
    p2 = float(pi64 - p1)
    p3 = float(pi64 - p1 - p2)
    print(f"constexpr double kPio64Hi = {float(p1)!r}, kPio64Mid = {p2!r}, kPio64Lo = {p3!r},")
    print(f"                 k64oPi = {float(64 / mp.pi)!r};")
    print("constexpr SinCosEntry kSinCosTab[65] = {")
    for k in range(65):
        s, c = mp.sinpi(mp.mpf(k) / 64), mp.cospi(mp.mpf(k) / 64)
        sh, ch = float(s), float(c)
        print(f"    {{{sh!r}, {float(s - sh)!r}, {ch!r}, {float(c - ch)!r}}},")
    print("};")


def main():
    sincos_table()
    us = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)
    t8 = mp.tan(mp.pi / 8)
    for name, f, w, b, n in (("log", f_log, lambda v: v, us * us, 6), ("atan", f_atan, lambda z: z, t8 * t8, 10),
                             ("asin", f_asin, lambda z: z, mp.mpf(1) / 4, 12)):
        coef, E = remez(f, w, mp.mpf(0), b, n)
        worst = check(f, w, mp.mpf(0), b, coef)
        print(f"// {name}: degree {n}, weighted minimax error {mp.nstr(abs(E), 3)}, "
              f"with double coefficients {mp.nstr(worst, 3)} (2^{mp.nstr(mp.log(worst, 2), 4)})")
        for j, c in enumerate(coef):
            print(f"constexpr double k{name.capitalize()}{j} = {float(c)!r};")


if __name__ == "__main__":
    main()
