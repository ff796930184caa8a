"""ctypes front end of the checker (test infrastructure only).

Loads ``oracle/liboracle.so`` (the C restatement of the reference's pairwise BIA
path, ``bia_oracle.c``) and, when it has been built, ``oracle/_ref/libbiaref.so``
(the reference's own Fortran CPU routines compiled by ``oracle/Makefile``).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module.  The product library never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

HYPERBOLA, PARABOLA, ELLIPSE, RKF = 0, 1, 2, 3
F_STOCK_DEFAULT, F_CYCLE, F_ITR_CAP, F_ACOS_CLAMP, F_RKF_FAIL, F_NONFINITE = 1, 2, 4, 8, 16, 32

_lib = None
_ref = None


def lib():
    """The C restatement; raises if it has not been built (``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_frame.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_hyperbola.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_ellipse.argtypes = L.orc_hyperbola.argtypes
        L.orc_parabola.argtypes = [C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_rkf_pair.argtypes = L.orc_parabola.argtypes
        L.orc_pair.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_pairs2d.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int64, _dp, _dp, _ip, _ip, _ip]
        L.orc_bia_rows.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_int, C.c_int,
                                   _dp, C.c_void_p]
        for f in ("orc_invariance_gpu", "orc_invariance_cpu", "orc_invariance_quad"):
            getattr(L, f).argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_dim_rkf45.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.c_double, _dp, C.c_double,
                                    C.c_double, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_int]
        L.orc_dim_rkf45.restype = C.c_int
        L.orc_dim_func.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_bia_rows_coupled.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_int, _dp]
        L.orc_conic_indeterminate.argtypes = [C.c_double, _dp]
        L.orc_sensitivity.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_void_p,
                                      C.c_void_p, C.c_int, _dp, _dp]
        L.orc_rkf_pair_tol.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_pair_quad.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_rows_quad.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_int, _ip, _dp, _dp, _dp]
        L.orc_rows_rkf_tol.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       _dp, C.c_void_p, C.c_void_p, C.c_int, _ip, _dp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def rkf2d_tol(eps: float, dt: float, yz, relerr: float, abserr: float):
    """The per-pair RKF45 of one 2D pair at the given tolerances; returns (yn, steps, flags)."""
    yz = np.ascontiguousarray(yz, dtype=np.float64)
    yn = np.zeros(4)
    it, fl = C.c_int(0), C.c_int(0)
    lib().orc_rkf_pair_tol(eps, dt, relerr, abserr, yz, yn, C.byref(it), C.byref(fl))
    return yn, it.value, fl.value


def rows_rkf_tol(state_soa, rows, dt: float, relerr: float, abserr: float, eps: float = 0.0, coupling: int = 0,
                 coup: float = 1.0, mass=None, charge=None):
    """Deltas [6, len(rows)] of whole rows of the given particles with the per-pair RKF45 at the
    given tolerances (couplings as bia_step_coupled); returns (delta, pairs that failed)."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((6, r.size))
    m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    q = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
    nf = C.c_int(0)
    lib().orc_rows_rkf_tol(st.shape[1], coupling, coup, eps, dt, relerr, abserr, st,
                           None if m is None else m.ctypes.data_as(C.c_void_p),
                           None if q is None else q.ctypes.data_as(C.c_void_p), r.size, r, out, C.byref(nf))
    return out, nf.value


def pair_quad(eps: float, dt: float, rel):
    """One pair's half deltas from the two-body solution in quadruple precision, as (hi, lo)
    doubles with hi + lo the quad result."""
    hi, lo = np.zeros(6), np.zeros(6)
    lib().orc_pair_quad(eps, dt, np.ascontiguousarray(rel, dtype=np.float64), hi, lo)
    return hi, lo


def rows_quad(state_soa, rows, eps: float, dt: float):
    """Whole rows of the given particles (scalar coupling) in quadruple precision: returns
    (corr_hi, corr_lo, delta), each [6, len(rows)]; corr = the sum of the particle's pair half
    deltas as hi + lo, delta = corr + v dt rounded once."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    hi, lo, d = (np.zeros((6, r.size)) for _ in range(3))
    lib().orc_rows_quad(st.shape[1], eps, dt, st, r.size, r, hi, lo, d)
    return hi, lo, d


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libbiaref.so"))


def ref():
    """The reference's own Fortran CPU routines (``make -C oracle ref``)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libbiaref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle ref` where the reference is mounted")
        L = C.CDLL(path)
        ip = C.POINTER(C.c_int)
        L.ref_hyperbola.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp, C.POINTER(C.c_double), ip, ip]
        L.ref_ellipse.argtypes = L.ref_hyperbola.argtypes
        L.ref_parabola.argtypes = [C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_double)]
        L.ref_rkf.argtypes = [C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_double), ip]
        L.ref_invariance.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_cpu_rows.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_int, C.c_int, _dp]
        _ref = L
    return _ref


# ------------------------------------------------------------------------------------------
# oracle wrappers
# ------------------------------------------------------------------------------------------
def frame(r, g):
    exi, eeta, yz = np.zeros(3), np.zeros(3), np.zeros(4)
    lib().orc_frame(np.ascontiguousarray(r, float), np.ascontiguousarray(g, float), exi, eeta, yz)
    return exi, eeta, yz


def solve2d(kind: str, eps: float, dt: float, yz, itr: int = 18):
    """One 2D pair with a named solver; returns (yn, iters, flags)."""
    yz = np.ascontiguousarray(yz, dtype=np.float64)
    yn = np.zeros(4)
    it, fl = C.c_int(0), C.c_int(0)
    L = lib()
    if kind == "hyperbola":
        L.orc_hyperbola(eps, dt, itr, yz, yn, C.byref(it), C.byref(fl))
    elif kind == "ellipse":
        L.orc_ellipse(eps, dt, itr, yz, yn, C.byref(it), C.byref(fl))
    elif kind == "parabola":
        L.orc_parabola(eps, dt, yz, yn, C.byref(it), C.byref(fl))
    elif kind == "rkf":
        L.orc_rkf_pair(eps, dt, yz, yn, C.byref(it), C.byref(fl))
    else:
        raise ValueError(kind)
    return yn, it.value, fl.value


def pairs2d(method: int, eps: float, dt: float, yz: np.ndarray, itr: int = 18):
    """Batch of 2D pairs (yz: [n,4]) through the classifier; returns (yn, branch, iters, flags)."""
    yz = np.ascontiguousarray(yz, dtype=np.float64).reshape(-1, 4)
    n = yz.shape[0]
    yn = np.zeros_like(yz)
    br, it, fl = (np.zeros(n, np.int32) for _ in range(3))
    lib().orc_pairs2d(method, eps, dt, itr, n, yz, yn, br, it, fl)
    return yn, br, it, fl


def pair(method: int, eps: float, dt: float, rel, itr: int = 18):
    out = np.zeros(6)
    br, it, fl = C.c_int(0), C.c_int(0), C.c_int(0)
    lib().orc_pair(method, eps, dt, itr, np.ascontiguousarray(rel, float), out, C.byref(br), C.byref(it),
                   C.byref(fl))
    return out, br.value, it.value, fl.value


def bia_step(state_soa: np.ndarray, eps: float, dt: float, method: int = 1, itr: int = 18,
             rows=None, hist: bool = False):
    """Whole BIA step (or rows [i0, i1)); returns deltas SoA [6, n] (drift included)."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    n = st.shape[1]
    i0, i1 = (0, n) if rows is None else rows
    d = np.zeros((6, n))
    h = np.zeros((4, 64), np.int64) if hist else None
    lib().orc_bia_rows(method, n, eps, dt, itr, st, i0, i1, d,
                       h.ctypes.data_as(C.c_void_p) if hist else None)
    return (d, h) if hist else d


def bia_step_coupled(state_soa: np.ndarray, coupling: int, coup: float, dt: float, mass=None, charge=None,
                     method: int = 1, itr: int = 18) -> np.ndarray:
    """Step with per-pair couplings (1 gravity with masses, 2 Coulomb); deltas SoA [6, n]."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    n = st.shape[1]
    d = np.zeros((6, n))
    m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    q = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
    lib().orc_bia_rows_coupled(method, n, coupling, coup, dt, itr, st,
                               None if m is None else m.ctypes.data_as(C.c_void_p),
                               None if q is None else q.ctypes.data_as(C.c_void_p), 0, n, d)
    return d


def invariance(state_soa: np.ndarray, eps: float, kind: str = "gpu") -> np.ndarray:
    """KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za (kind: gpu order, cpu order, quad)."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    out = np.zeros(11)
    getattr(lib(), f"orc_invariance_{kind}")(st.shape[1], eps, st, out)
    return out


def dim_rkf45(y_aos: np.ndarray, eps: float, x0: float, xe: float, relerr: float = 0.0,
              abserr: float = 1e-16, cap: int = 20000):
    """DIM direct-sum RKF45 on AoS state [n*6]; returns dict."""
    y = np.ascontiguousarray(y_aos, dtype=np.float64).copy().reshape(-1)
    np_ = y.size // 6
    ese = np.zeros_like(y)
    st_t = np.zeros(cap)
    st_y = np.zeros((cap, 6))
    x = C.c_double(x0)
    nde, nacc, nrej = C.c_int(0), C.c_int(0), C.c_int(0)
    ier = lib().orc_dim_rkf45(np_, eps, C.byref(x), xe, y, relerr, abserr, ese.ctypes.data_as(C.c_void_p),
                              C.byref(nde), C.byref(nacc), C.byref(nrej), st_t.ctypes.data_as(C.c_void_p),
                              st_y.ctypes.data_as(C.c_void_p), cap)
    k = min(nacc.value, cap)
    return dict(y=y, ier=ier, x=x.value, nde=nde.value, naccept=nacc.value, nreject=nrej.value,
                ese=ese, step_t=st_t[:k].copy(), step_y1=st_y[:k].copy())


# ------------------------------------------------------------------------------------------
# reference-routine wrappers (amdflang build of the reference's CPU Fortran)
# ------------------------------------------------------------------------------------------
def ref_solve2d(kind: str, eps: float, dt: float, yz, itr: int = 18):
    yz = np.ascontiguousarray(yz, dtype=np.float64)
    yn = np.zeros(4)
    err = C.c_double(0)
    ir, ier = C.c_int(0), C.c_int(0)
    R = ref()
    if kind == "hyperbola":
        R.ref_hyperbola(eps, dt, itr, yz, yn, C.byref(err), C.byref(ir), C.byref(ier))
    elif kind == "ellipse":
        R.ref_ellipse(eps, dt, itr, yz, yn, C.byref(err), C.byref(ir), C.byref(ier))
    elif kind == "parabola":
        R.ref_parabola(eps, dt, yz, yn, C.byref(err))
    elif kind == "rkf":
        R.ref_rkf(eps, dt, yz, yn, C.byref(err), C.byref(ier))
    else:
        raise ValueError(kind)
    return yn, err.value, ir.value, ier.value


def ref_invariance(state_soa: np.ndarray, eps: float, quad: bool) -> np.ndarray:
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    out = np.zeros(11)
    cols = [np.ascontiguousarray(st[k]) for k in range(6)]
    ref().ref_invariance(st.shape[1], eps, int(quad), *cols, out)
    return out


def ref_cpu_rows(state_soa: np.ndarray, eps: float, dt: float, method: int = 1, itr: int = 18, rows=None):
    """The reference CPU pair loop (bia.f90:142-196) over rows [i0, i1); deltas SoA [6, n]."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    n = st.shape[1]
    i0, i1 = (0, n) if rows is None else rows
    d = np.zeros((6, n))
    ref().ref_cpu_rows(method, n, eps, dt, itr, st, i0, i1, d)
    return d


# ------------------------------------------------------------------------------------------
# parity bound: the reference formulas' own sensitivity to one rounding unit
# ------------------------------------------------------------------------------------------
def sensitivity(state_soa: np.ndarray, eps: float, dt: float, method: int = 1, itr: int = 18, samples: int = 6,
                coupling: int = 0, mass=None, charge=None):
    """Per-particle sums over its pairs of the largest change of the pair's (position, velocity)
    half-deltas under random 1-ulp perturbations of the pair's six relative-state inputs.

    Two correct evaluations of the reference formulas whose libm results differ by one ulp land
    about this far apart.  Two kinds of pair have no answer to match and give their two
    particles an infinite bound (``indeterminate`` counts such particles): a pair whose conic
    type is decided by rounding (classifier value W below 64 rounding units of its terms,
    orc_conic_indeterminate), and a pair on which the reference Newton stopped unconverged and
    the independent RKF45 of the same pair disagrees with it by more than 1e-6.  It is large where the reference's formulation is ill-conditioned:
    acos(q) with q near 1 (pairs starting near periapsis, bia/find_dXYUV_kernel.cu:124) and
    near-parabolic hyperbolae (W = g^2 + 2 eps/r cancels, :98-107)."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    n = st.shape[1]
    bp, bv = np.zeros(n), np.zeros(n)
    m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    q = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
    lib().orc_sensitivity(method, n, coupling, eps, dt, itr, st, None if m is None else m.ctypes.data_as(C.c_void_p),
                          None if q is None else q.ctypes.data_as(C.c_void_p), samples, bp, bv)
    return bp, bv


def indeterminate(bounds) -> int:
    """Number of particles excluded by sensitivity() (a pair of indeterminate type, or one the
    reference algorithm gets wrong)."""
    return int(np.count_nonzero(~np.isfinite(bounds[0])))


def conic_indeterminate(eps: float, rel) -> bool:
    return bool(lib().orc_conic_indeterminate(eps, np.ascontiguousarray(rel, dtype=np.float64)))


def within_condition_bound(state_soa, d, ref, eps, dt: float = None, factor: float = 10.0, floor: float = 1e-13,
                           method: int = 1, bounds=None):
    """Per-particle check of deltas d against ref: |d - ref| <= factor * sensitivity + floor *
    max|x + ref| (positions and velocities separately).  Returns (ok, worst ratio pos, vel).

    factor = 10: on the reference's own N = 1332 fixture its CUDA output (OUT/*.BIA) and the
    glibc evaluation of the same formulas differ by up to 7x this sensitivity estimate."""
    st = np.asarray(state_soa, dtype=np.float64).reshape(6, -1)
    if bounds is None:
        if dt is None:
            raise ValueError("dt is needed to evaluate the sensitivity")
        bounds = sensitivity(st, eps, dt, method)
    bp, bv = bounds
    ep = np.abs(d[:3] - ref[:3]).max(0)
    ev = np.abs(d[3:] - ref[3:]).max(0)
    ap = factor * bp + floor * np.abs(st[:3] + ref[:3]).max()
    av = factor * bv + floor * np.abs(st[3:] + ref[3:]).max()
    rp, rv = (ep / ap).max(), (ev / av).max()
    return bool(rp <= 1 and rv <= 1), float(rp), float(rv)
