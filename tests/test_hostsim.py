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
    L.hs_pairs.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_int, dp, dp, ip, ip, up]

    def run(method, rel, eps, dt, itr=18):
        m = rel.shape[0]
        out = np.zeros((m, 6))
        br, it, fl = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.uint32)
        L.hs_pairs(method, m, eps, dt, itr, np.ascontiguousarray(rel), out, br, it, fl)
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
