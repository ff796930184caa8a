"""The reference's unmodified Fortran host programs (bia/bia.f90, dim/dim_3D.f90), compiled by
oracle/Makefile and linked against libbia_hip.so, run on the reference's N=1332 fixture in the
directory layout they expect (../OUT relative to the working directory)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def _layout(tmp_path, with_dim=True):
    out = tmp_path / "OUT"
    run = tmp_path / "run"
    out.mkdir()
    run.mkdir()
    names = ["0001332_10^0.bbb"] + (["0001332_10^0.DIM"] if with_dim else [])
    for nm in names:
        shutil.copy(os.path.join(GOLDEN, nm), out / nm)
    return out, run


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "bia_gpu_alg")), reason="reference build absent")
def test_reference_bia_f90_on_library(tmp_path):
    from bia_3d_gpu_amd import io
    out, run = _layout(tmp_path)
    p = subprocess.run([os.path.join(REF_BIN, "bia_gpu_alg")], cwd=run, capture_output=True, text=True, timeout=600)
    print(p.stdout[-3000:])
    assert p.returncode == 0, p.stderr[-2000:]
    # the program's own solver timing (bia/bia.f90:212; 0.193 s on the reference's TITAN V)
    timing = [ln for ln in p.stdout.splitlines() if "for BIA-ALG" in ln]
    print("solver timing of the unmodified host program on libbia_hip:", timing)
    assert timing
    (dt, got), = io.read_timed(str(out / "0001332_10^0.BIA"))
    (_, ref), = io.read_timed(os.path.join(GOLDEN, "0001332_10^0.BIA"))
    dp, dv = np.abs(got[:3] - ref[:3]).max(), np.abs(got[3:] - ref[3:]).max()
    print(f"reference bia.f90 + libbia_hip vs OUT/0001332_10^0.BIA: pos {dp:.3e} vel {dv:.3e}")
    assert dt == 1.0 and dp < 3e-11 and dv < 1e-14


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "dim_gpu")), reason="reference build absent")
def test_reference_dim_f90_on_library(tmp_path):
    from bia_3d_gpu_amd import io
    out, run = _layout(tmp_path, with_dim=False)
    trace = tmp_path / "trace.DIM"
    env = dict(os.environ, BIA_DIM_TRACE=str(trace))
    p = subprocess.run([os.path.join(REF_BIN, "dim_gpu")], cwd=run, capture_output=True, text=True, timeout=600,
                       env=env)
    print(p.stdout[-3000:])
    assert p.returncode == 0, p.stderr[-2000:]
    (t, got), = io.read_timed(str(out / "0001332_10^0.DIM"))
    (_, ref), = io.read_timed(os.path.join(GOLDEN, "0001332_10^0.DIM"))
    dp, dv = np.abs(got[:3] - ref[:3]).max(), np.abs(got[3:] - ref[3:]).max()
    print(f"reference dim_3D.f90 + libbia_hip vs OUT/0001332_10^0.DIM: pos {dp:.3e} vel {dv:.3e}")
    print("timing of the unmodified dim_3D.f90 on libbia_hip (its own cpu_time, particle-1 trace on):",
          [ln.strip() for ln in p.stdout.splitlines() if ln.strip().startswith("CPU =")])
    assert t == 1.0 and dp < 1e-13 and dv < 1e-13
    mine = np.loadtxt(trace, skiprows=1)
    theirs = np.loadtxt(os.path.join(GOLDEN, "0001332.DIM"), skiprows=1)
    assert mine.shape == theirs.shape
    # accepted step times: the controller works on error ratios at the rounding floor
    # (ABSERR 1e-16), so step sizes agree to ~1e-9 relative, not to the last bit
    assert np.abs(mine[:, 0] - theirs[:, 0]).max() < 1e-7


FORTRAN_DRIVER = os.path.join(ROOT, "bia_3d_gpu_amd", "fortran", "bia_run")


@pytest.mark.skipif(not os.path.exists(FORTRAN_DRIVER), reason="Fortran driver not built")
@pytest.mark.parametrize("mode", [0, 2])
def test_fortran_driver_over_the_bind_c_module(ctx, tmp_path, mode):
    """bia_3d_gpu_amd/fortran/bia_run (the bind(C) module's host driver: read .bbb, invariants,
    one whole-step call, write .BIA) on the reference's N=1332 fixture: the reference solvers'
    deltas within the fixture tolerances of the reference's own .BIA, and both modes bitwise
    equal to the same step through the Python binding of the same entry point."""
    from bia_3d_gpu_amd import io
    prefix = tmp_path / "fx"
    shutil.copy(os.path.join(GOLDEN, "0001332_10^0.bbb"), str(prefix) + ".bbb")
    shutil.copy(os.path.join(GOLDEN, "0001332_10^0.DIM"), str(prefix) + ".DIM")
    p = subprocess.run([FORTRAN_DRIVER, str(prefix), "-6e-7", "1.0", "1", str(mode)], capture_output=True,
                       text=True, timeout=300)
    print(p.stdout[-2000:])
    assert p.returncode == 0, p.stderr[-2000:]
    (dt, got), = io.read_timed(str(prefix) + ".run.BIA")
    st = io.read_bbb(str(prefix) + ".bbb")[0]
    assert dt == 1.0
    assert np.array_equal(got, ctx.step_host(st, -6e-7, 1.0, mode=mode))
    if mode == 0:
        (_, ref), = io.read_timed(os.path.join(GOLDEN, "0001332_10^0.BIA"))
        assert np.abs(got[:3] - ref[:3]).max() < 3e-11 and np.abs(got[3:] - ref[3:]).max() < 1e-14
