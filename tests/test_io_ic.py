"""File formats (round trips, the reference's fixtures) and the synthetic initial conditions."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


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


def test_native_record_files(tmp_path, fixture1332):
    """The library's record functions (bia_records_scan / bia_record_read / bia_record_write, the
    C side of io.read_bbb etc.): the reference's fixtures read as the raw records say; a
    multi-record file of uneven N (an N_calc ensemble) round-trips with its heads; size queries,
    records past the last and short rows are argument errors; malformed files (bad markers, a
    payload that is not whole particles) are BIA_ERR_IO and raise ValueError through io."""
    import ctypes as C
    import struct
    from bia_3d_gpu_amd import _lib, io
    L = _lib.load()
    fx = os.path.join(GOLDEN, "0001332_10^0.DIM")
    nrec, nmax = C.c_longlong(), C.c_longlong()
    assert L.bia_records_scan(fx.encode(), 1, C.byref(nrec), C.byref(nmax)) == _lib.BIA_OK
    assert (nrec.value, nmax.value) == (1, 1332)
    raw = np.frombuffer(io.read_records(fx)[0], dtype="<f8")
    (t, dim), = io.read_timed(fx)
    assert t == raw[0] and np.array_equal(dim, raw[1:].reshape(-1, 6).T)
    # an ensemble of uneven N with heads, written by appending
    rng = np.random.default_rng(62)
    recs = [(float(k) * 0.5, rng.normal(size=(6, n))) for k, n in enumerate([5, 1, 300, 0, 17])]
    p = str(tmp_path / "ens.DIM")
    io.write_timed(p, recs)
    back = io.read_timed(p)
    assert len(back) == 5 and all(a[0] == b[0] and np.array_equal(a[1], b[1]) for a, b in zip(recs, back))
    assert L.bia_records_scan(p.encode(), 1, C.byref(nrec), C.byref(nmax)) == _lib.BIA_OK
    assert (nrec.value, nmax.value) == (5, 300)
    n = C.c_longlong()
    assert L.bia_record_read(p.encode(), 1, 2, None, None, 0, C.byref(n)) == _lib.BIA_OK and n.value == 300
    head, soa = np.zeros(1), np.zeros((6, 299))
    assert L.bia_record_read(p.encode(), 1, 2, head.ctypes.data, soa.ctypes.data, 299, C.byref(n)) == _lib.BIA_ERR_ARG
    assert L.bia_record_read(p.encode(), 1, 5, None, None, 0, C.byref(n)) == _lib.BIA_ERR_ARG
    # rows of stride ld > n
    wide = np.zeros((6, 310))
    assert L.bia_record_read(p.encode(), 1, 2, head.ctypes.data, wide.ctypes.data, 310, C.byref(n)) == _lib.BIA_OK
    assert head[0] == 1.0 and np.array_equal(wide[:, :300], recs[2][1]) and not wide[:, 300:].any()
    # malformed: 7 doubles are not whole particles; markers as in the raw reader's test
    bad = tmp_path / "bad.bbb"
    bad.write_bytes(struct.pack("<i", 56) + np.arange(7.0).tobytes() + struct.pack("<i", 56))
    assert L.bia_records_scan(str(bad).encode(), 0, C.byref(nrec), C.byref(nmax)) == _lib.BIA_ERR_IO
    good = tmp_path / "good.bbb"
    io.write_bbb(str(good), [np.arange(12, dtype=float).reshape(6, 2)])
    data = good.read_bytes()
    for name, blob in {"truncated": data[:-3], "negative": struct.pack("<i", -8) + data,
                       "oversized": struct.pack("<i", 1 << 30) + data[4:],
                       "mismatch": data[:-4] + struct.pack("<i", 7), "shape": bad.read_bytes()}.items():
        q = tmp_path / f"m_{name}.bbb"
        q.write_bytes(blob)
        with pytest.raises(ValueError):
            io.read_bbb(str(q))
    with pytest.raises(FileNotFoundError):
        io.read_bbb(str(tmp_path / "missing.bbb"))


def test_fortran_module_reads_records(tmp_path, fixture1332):
    """The bind(C) module's record interfaces from Fortran (amdflang): a program counts and reads
    the reference's .bbb fixture through bia_records_scan / bia_record_read into soa(ld, 6) and
    prints particle 2; the values equal the Python reader's."""
    import shutil
    import subprocess
    if shutil.which("amdflang") is None:
        pytest.skip("amdflang absent")
    root = ROOT
    fdir = os.path.join(root, "bia_3d_gpu_amd", "fortran")
    prog = tmp_path / "rd.f90"
    prog.write_text(
        "program rd\n"
        "  use bia_hip\n"
        "  use iso_c_binding\n"
        "  implicit none\n"
        "  integer(c_long_long) :: nrec, nmax, n\n"
        "  real(c_double), allocatable :: soa(:, :)\n"
        "  real(c_double) :: head(1)\n"
        "  character(len=512) :: path\n"
        "  integer(c_int) :: s\n"
        "  call get_command_argument(1, path)\n"
        "  s = bia_records_scan(trim(path)//c_null_char, 0, nrec, nmax)\n"
        "  if (s /= BIA_OK) stop 1\n"
        "  allocate(soa(nmax, 6))\n"
        "  s = bia_record_read(trim(path)//c_null_char, 0, 0_c_long_long, head, soa, nmax, n)\n"
        "  if (s /= BIA_OK) stop 2\n"
        "  print '(i0,1x,i0,6es25.16e3)', nrec, n, soa(2, :)\n"
        "end program rd\n")
    exe = tmp_path / "rd"
    subprocess.run(["amdflang", "-O1", "-module-dir", str(tmp_path), "-o", str(exe),
                    os.path.join(fdir, "bia_hip.f90"), str(prog), "-L", os.path.join(root, "bia_3d_gpu_amd"),
                    "-lbia_hip", "-Wl,-rpath," + os.path.join(root, "bia_3d_gpu_amd")], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe), os.path.join(GOLDEN, "0001332_10^0.bbb")], check=True, capture_output=True,
                         text=True).stdout.split()
    assert int(out[0]) == 1 and int(out[1]) == 1332
    assert np.array_equal(np.array([float(v) for v in out[2:]]), fixture1332["state"][:, 1])
