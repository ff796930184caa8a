"""Argument checks of the Python device-tensor entry points (bia_3d_gpu_amd/api.py).

A float32, strided, short or host tensor handed to a device entry point would become an
out-of-bounds device access or a silently wrong result, so the wrappers reject such arguments
with ValueError before the library is called.  Runs on the CPU: the checks come before any call
into libbia_hip.so (the context below has no device handle at all, so a check that let an
argument through would fail with another error than the one asserted).
"""
import numpy as np
import pytest
import torch

from bia_3d_gpu_amd import Context
from bia_3d_gpu_amd._lib import COUPLING_COULOMB, COUPLING_GRAVITY, STATS_LEN


@pytest.fixture()
def ctx():
    c = Context.__new__(Context)  # no bia_create: nothing here may reach the library
    c._h = None
    c.device = 0
    return c


def _state(n=8, dtype=torch.float64):
    return torch.zeros((6, n), dtype=dtype)


def _raises(fn, match):
    with pytest.raises(ValueError, match=match):
        fn()


def test_state_and_delta_checks(ctx):
    s, d = _state(), _state()
    _raises(lambda: ctx.step_(s.float(), d, -1.0, 1e-3), "dtype float64")
    _raises(lambda: ctx.step_(s, d.float(), -1.0, 1e-3), "dtype float64")
    _raises(lambda: ctx.step_(torch.zeros((8, 6), dtype=torch.float64).t(), d, -1.0, 1e-3), "contiguous")
    _raises(lambda: ctx.step_(torch.zeros((5, 8), dtype=torch.float64), d, -1.0, 1e-3), r"\[6, ld\]")
    _raises(lambda: ctx.step_(s, _state(7), -1.0, 1e-3), "shape")
    _raises(lambda: ctx.step_(np.zeros((6, 8)), d, -1.0, 1e-3), "torch tensor")
    # well-formed but on the host
    _raises(lambda: ctx.step_(s, d, -1.0, 1e-3), "device")


def test_stats_checks(ctx):
    s, d = _state(), _state()
    _raises(lambda: ctx.step_(s, d, -1.0, 1e-3, stats=torch.zeros(STATS_LEN, dtype=torch.int32)), "int64")
    _raises(lambda: ctx.step_(s, d, -1.0, 1e-3, stats=torch.zeros(STATS_LEN - 1, dtype=torch.int64)),
            "at least 16")


def test_batched_checks(ctx):
    s, d = _state(8), _state(8)
    m = torch.ones(8, dtype=torch.float64)
    _raises(lambda: ctx.step_batched_(s, d, [4, 5], eps=-1.0), "only 8 columns")
    _raises(lambda: ctx.step_batched_(s, d, [0, 8], eps=-1.0), ">= 1")
    _raises(lambda: ctx.step_batched_(s, d, [], eps=-1.0), "at least one system")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4]), "needs eps")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4], coupling=COUPLING_GRAVITY), "masses")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4], coupling=COUPLING_COULOMB, mass=m), "charges")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4], coupling=COUPLING_GRAVITY, mass=m[:7]), "at least 8")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4], coupling=COUPLING_GRAVITY, mass=m.float()), "float64")
    _raises(lambda: ctx.step_batched_(s, d, [4, 4], coupling=9), "unknown")
    _raises(lambda: ctx.run_steps_(s, d, [4, 4], -1, eps=-1.0), "nsteps")
    _raises(lambda: ctx.autotune_(s, d, [4, 5], eps=-1.0), "only 8 columns")
    # a short mass array is fine when it covers the particles of the listed systems
    _raises(lambda: ctx.step_batched_(s, d, [3, 3], coupling=COUPLING_GRAVITY, mass=m[:6]), "device")


def test_invariant_and_mlp_checks(ctx):
    s = _state(8)
    out = torch.zeros(11, dtype=torch.float64)
    _raises(lambda: ctx.invariants_(s, out[:10], -1.0), "at least 11")
    _raises(lambda: ctx.invariants_batched_(s, [4, 4], -1.0, out), "at least 22")
    _raises(lambda: ctx.invariants_at_(s, 4, 5, -1.0, out), "inside the state")
    _raises(lambda: ctx.invariants_coupled_(s, [8], out, 0), "gravity or Coulomb")
    rows = torch.zeros((16, 4), dtype=torch.float32)
    _raises(lambda: ctx.mlp_forward_(rows.double(), torch.zeros(16)), "float32")
    _raises(lambda: ctx.mlp_forward_(torch.zeros((16, 3)), torch.zeros(16)), r"\[B, 4\]")
    _raises(lambda: ctx.mlp_forward_(rows, torch.zeros(15)), "at least 16")
    _raises(lambda: ctx.record_samples_(s, -1.0, 1e-3, 8, 1, torch.zeros((7, 4))), "at least 32")
