"""File formats (round trips, the reference's fixtures) and the synthetic initial conditions."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_read_reference_fixtures(fixture1332):
    st = fixture1332["state"]
    assert st.shape == (6, 1332)
    assert np.all(st[:, 0] == 0)  # particle 1 at rest at the origin (IC/bbb_maker_REV.f90)
    assert fixture1332["dt"] == 1.0 and fixture1332["t"] == 1.0


def test_round_trips(tmp_path, fixture1332):
    from bia_3d_gpu_amd import io
    st = fixture1332["state"]
    io.write_bbb(str(tmp_path / "a.bbb"), [st, st * 2])
    back = io.read_bbb(str(tmp_path / "a.bbb"))
    assert len(back) == 2 and np.array_equal(back[1], st * 2)
    raw = open(os.path.join(GOLDEN, "0001332_10^0.bbb"), "rb").read()
    assert open(tmp_path / "a.bbb", "rb").read()[: len(raw)] == raw  # byte-identical record
    io.write_timed(str(tmp_path / "a.BIA"), [(1.0, fixture1332["bia"])])
    assert open(tmp_path / "a.BIA", "rb").read() == open(os.path.join(GOLDEN, "0001332_10^0.BIA"), "rb").read()


def test_text_line_format():
    from bia_3d_gpu_amd import io
    line = open(os.path.join(GOLDEN, "BIA_10^0.0001332")).read().rstrip("\n")
    v = io.read_particle_line(os.path.join(GOLDEN, "BIA_10^0.0001332"))
    assert io.format_particle_line(v[0], v[1:]) == line


def test_plummer_properties():
    from bia_3d_gpu_amd import ic
    st = ic.plummer_system(20000, seed=9)
    r = np.linalg.norm(st[:3], axis=0)
    assert abs(np.median(r) - 1.305) < 0.05  # half-mass radius of a = 1
    m = np.full(st.shape[1], 1 / st.shape[1])
    T = 0.5 * (m * (st[3:] ** 2).sum(0)).sum()
    W = ic.potential_energy(st[:3], m)
    assert abs(2 * T / abs(W) - 1) < 0.05  # virial equilibrium


def test_binaries_config():
    from bia_3d_gpu_amd import ic
    n = 4096
    st = ic.plummer_with_binaries(n, seed=4)
    assert st.shape == (6, n)
    # count tightly bound close pairs: separation below the widest apocentre a(1+e) < 0.02
    from scipy.spatial import cKDTree
    t = cKDTree(st[:3].T)
    pairs = np.array(sorted(t.query_pairs(0.021)))
    eps = -2.0 / n
    d = st[:, pairs[:, 0]] - st[:, pairs[:, 1]]
    E = (d[3:] ** 2).sum(0) + 2 * eps / np.linalg.norm(d[:3], axis=0)
    nb = int(round(0.2 * n)) // 2
    # every planted binary is a bound close pair; a few chance close encounters of single
    # stars in the dense core are bound as well
    assert nb <= (E < 0).sum() <= 1.2 * nb


def test_uniform_ball_virial_ratio():
    from bia_3d_gpu_amd import ic
    st = ic.uniform_ball_virial(1332, seed=1332)
    m = np.full(1332, 1 / 1332)
    T = 0.5 * (m * (st[3:] ** 2).sum(0)).sum()
    assert abs(T / abs(ic.potential_energy(st[:3], m)) - 0.5) < 1e-12
    assert np.linalg.norm(st[:3], axis=0).max() <= 1


def test_coulomb_bound_fraction():
    from bia_3d_gpu_amd import ic
    st, q = ic.coulomb_box(2000, seed=16384)
    assert q.sum() == 0
    rng = np.random.default_rng(0)
    i, j = rng.integers(0, 2000, 50000), rng.integers(0, 2000, 50000)
    ok = i != j
    i, j = i[ok], j[ok]
    d = st[:, i] - st[:, j]
    E = (d[3:] ** 2).sum(0) + 2 * 2 * q[i] * q[j] / np.linalg.norm(d[:3], axis=0)
    assert abs((E < 0).mean() - 0.4) < 0.02


def test_malformed_record_files(tmp_path):
    """Truncated files, negative or oversized record markers and mismatched trailing markers
    raise ValueError (a negative marker once made the reader loop for ever)."""
    import struct
    from bia_3d_gpu_amd import io
    good = tmp_path / "good.bbb"
    io.write_bbb(str(good), [np.arange(12, dtype=float).reshape(6, 2)])
    data = good.read_bytes()
    cases = {"truncated": data[:-3], "short marker": data + b"\x01\x00",
             "negative": struct.pack("<i", -8) + data, "oversized": struct.pack("<i", 1 << 30) + data[4:],
             "mismatch": data[:-4] + struct.pack("<i", 7)}
    for name, blob in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.bbb"
        p.write_bytes(blob)
        with pytest.raises(ValueError):
            io.read_records(str(p))
    assert len(io.read_records(str(good))) == 1
