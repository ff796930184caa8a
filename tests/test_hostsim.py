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
                           C.c_void_p, C.c_int, C.c_int]

    def run(method, rel, eps, dt, itr=18, mode=0, table=None):
        m = rel.shape[0]
        out = np.zeros((m, 6))
        br, it, fl = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.uint32)
        tp, nv, nx = (None, 0, 0) if table is None else (table.ctypes.data_as(C.c_void_p), table.shape[1],
                                                         table.shape[2])
        L.hs_pairs(method, m, eps, dt, itr, np.ascontiguousarray(rel), out, br, it, fl, mode, tp, nx, nv)
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


def test_guess_mode_safeguarded_newton(hs):
    """BIA_MODE_GUESS (the reference guess + bracketed Newton with a rounding-aware stop): same
    results as the reference iteration wherever that converges, and converged results on the
    strongly deflected pairs where the reference's Newton leaves (-theta_0, theta_0)."""
    from bia_3d_gpu_amd import ic
    st = ic.plummer_with_binaries(1024, seed=8)
    rel = _pairs(st)[::7]
    eps, dt = -2 / 65536, 1e-3
    a, br, it, fl = hs(1, rel, eps, dt, mode=2)
    b, bb, fb = _oracle_pairs(1, rel, eps, dt)
    hyp = br == 0
    assert hyp.mean() > 0.99
    assert (fl[hyp] & 128).sum() == 0  # no fallback needed
    assert it[hyp].mean() < 3.0        # evaluations per pair (the reference spends ~5.5)
    good = (fb & (O.F_ITR_CAP | O.F_STOCK_DEFAULT)) == 0
    err = (np.abs(a - b) / np.abs(rel).max(0)).max(1)
    assert np.quantile(err[good & hyp], 0.999) < 1e-11
    # strongly deflected pairs (e ~ 1-3): the reference's Newton fails on some; the bracketed
    # one converges (checked against the independent RKF45 integration of the same pair)
    eps2 = -1.0
    rng = np.random.default_rng(1)
    rels = []
    for _ in range(200):
        r = 0.9
        g = rng.uniform(1.5, 2.5) * np.sqrt(-eps2 / r)
        ang = rng.uniform(0.1, np.pi - 0.1)
        rels.append(np.concatenate([[r, 0, 0], g * np.array([np.cos(ang), np.sin(ang), 0])]))
    rels = np.array(rels)
    a, br, it, fl = hs(1, rels, eps2, 0.25, mode=2)
    k, _, _, _ = hs(0, rels, eps2, 0.25)
    assert np.all(br == 0) and (fl & 128).sum() == 0
    assert np.abs(a - k).max() < 1e-10
    b, bb, fb = _oracle_pairs(1, rels, eps2, 0.25)
    assert ((fb & O.F_ITR_CAP) != 0).sum() > 0  # the reference iteration does fail on some of these
