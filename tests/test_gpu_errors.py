"""Argument and call-sequence errors of the C ABI (include/bia_hip.h status codes).

Every invalid call below is rejected on the host before anything is launched, with the status the
header documents; the context stays usable, and the next valid call on it gives bitwise the
result of the same call on a fresh context.  The reference's own entry points return no status
(find_dXYUV_kernel.cu:526, invariance.cu:50); this library's versions of them print the failing
status and abort the program, which is not exercised here.
"""
import numpy as np
import pytest

from bia_3d_gpu_amd._lib import (BIA_ERR_ARG, BIA_ERR_STATE, COUPLING_COULOMB, COUPLING_GRAVITY, METHOD_RKF, MODE_GUESS,
                                 MODE_LEARN, MODE_PLAIN, BiaError)

pytestmark = pytest.mark.gpu

SIZES = [64, 66]
EPS = [-2.0 / 64, -2.0 / 66]


def _status(fn) -> int:
    with pytest.raises(BiaError) as e:
        fn()
    return e.value.status


def _raw(name, *args) -> int:
    """Status of a direct call of a C entry point (the Python wrappers reject most of these
    arguments themselves before calling the library, tests/test_api_checks.py)."""
    from bia_3d_gpu_amd._lib import load
    return getattr(load(), name)(*args)


def _raw_step(ctx, s, d, sizes, eps, method=1, mode=MODE_GUESS, coupling=0, mass=None, charge=None, itr=18):
    sz = np.ascontiguousarray(sizes, dtype=np.int32)
    e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps if eps is not None else 0.0, dtype=np.float64),
                                             (len(sz),)))
    ptr = lambda t: None if t is None else t.data_ptr()
    return _raw("bia_step_batched_dev", ctx.handle, method, mode, coupling, len(sz), sz, e, 1.0, 1e-3, itr,
                s.data_ptr(), s.shape[1], ptr(mass), ptr(charge), d.data_ptr(), None)


@pytest.fixture(scope="module")
def arrays():
    import torch
    from bia_3d_gpu_amd import ic
    st = np.concatenate([ic.plummer_system(n, seed=300 + k) for k, n in enumerate(SIZES)], axis=1)
    s = torch.from_numpy(np.ascontiguousarray(st)).cuda()
    m = torch.full((s.shape[1],), 1.0 / 64, dtype=torch.float64, device="cuda")
    return s, torch.zeros_like(s), m


def _valid_step(c, s, mode=MODE_GUESS):
    import torch
    d = torch.zeros_like(s)
    c.step_batched_(s, d, SIZES, eps=EPS, dt=1e-3, mode=mode)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def test_step_argument_errors(ctx, arrays):
    s, d, m = arrays
    n = s.shape[1]
    raw = lambda *a, **k: _raw_step(ctx, s, d, *a, **k)
    assert raw([0, n], [-1.0, -1.0]) == BIA_ERR_ARG            # empty system
    assert raw([-1, n + 1], [-1.0, -1.0]) == BIA_ERR_ARG       # negative size
    assert raw([64, 67], EPS) == BIA_ERR_ARG                   # sizes exceed ld
    assert raw([], []) == BIA_ERR_ARG                          # no system
    assert raw(SIZES, EPS, method=7) == BIA_ERR_ARG
    assert raw(SIZES, EPS, mode=7) == BIA_ERR_ARG
    assert raw(SIZES, EPS, method=METHOD_RKF, mode=MODE_GUESS) == BIA_ERR_ARG
    assert raw(SIZES, EPS, coupling=3) == BIA_ERR_ARG
    assert raw(SIZES, None, coupling=COUPLING_GRAVITY) == BIA_ERR_ARG     # no masses
    assert raw(SIZES, None, coupling=COUPLING_COULOMB, mass=m) == BIA_ERR_ARG  # no charges
    assert raw(SIZES, EPS, mode=MODE_PLAIN, itr=-1) == BIA_ERR_ARG
    assert raw(SIZES, EPS, mode=MODE_PLAIN, itr=21) == BIA_ERR_ARG  # history holds 20
    sz = np.ascontiguousarray(SIZES, dtype=np.int32)
    e = np.ascontiguousarray(EPS, dtype=np.float64)
    assert _raw("bia_run_steps_dev", ctx.handle, 1, MODE_GUESS, 0, 2, sz, e, 1.0, 1e-3, 18, s.data_ptr(),
                s.shape[1], None, None, d.data_ptr(), -1, 1, None) == BIA_ERR_ARG
    bad = np.ascontiguousarray([64, 67], dtype=np.int32)
    assert _raw("bia_autotune_dev", ctx.handle, 1, MODE_GUESS, 0, 2, bad, e, 1.0, 1e-3, 18, s.data_ptr(),
                s.shape[1], None, None, d.data_ptr(), None, None) == BIA_ERR_ARG
    assert _status(lambda: ctx.step_host(np.zeros((6, 0)), -1.0, 1e-3)) == BIA_ERR_ARG
    # the same arguments through the Python wrapper are rejected before the library is called
    with pytest.raises(ValueError):
        ctx.step_batched_(s, d, [64, 67], eps=EPS, dt=1e-3, mode=MODE_GUESS)


def test_invariant_argument_errors(ctx, arrays):
    import torch
    s, _, _ = arrays
    out = torch.zeros(22, dtype=torch.float64, device="cuda")
    e = np.ascontiguousarray(EPS, dtype=np.float64)
    for bad in ([0, 130], [64, 67]):
        sz = np.ascontiguousarray(bad, dtype=np.int32)
        assert _raw("bia_invariants_batched_dev", ctx.handle, 2, sz, e, s.data_ptr(), s.shape[1],
                    out.data_ptr()) == BIA_ERR_ARG
    assert _status(lambda: ctx.invariants(np.zeros((6, 0)), -1.0)) == BIA_ERR_ARG


def test_context_errors():
    from bia_3d_gpu_amd import Context
    assert _status(lambda: Context(1 << 20)) == BIA_ERR_ARG  # no such device


def test_learned_mode_needs_a_network(arrays):
    """Learned mode before a table is built is a call-sequence error, with the reason in
    bia_last_error."""
    from bia_3d_gpu_amd import Context
    s, d, _ = arrays
    with Context(0) as c:
        with pytest.raises(BiaError) as e:
            c.step_batched_(s, d, SIZES, eps=EPS, mode=MODE_LEARN)
        assert e.value.status == BIA_ERR_STATE
        assert "bia_mlp_build_table" in str(e.value)


def test_errors_leave_the_context_usable(ctx, arrays):
    """After the rejected calls, the same context steps bitwise like a fresh one, and a plan
    for other sizes built in between does not leak into the next call."""
    from bia_3d_gpu_amd import Context
    s, d, _ = arrays
    for bad in ([0, 130], [64, 67], [-1, 131]):
        assert _raw_step(ctx, s, d, bad, [-1.0, -1.0]) == BIA_ERR_ARG
    got = _valid_step(ctx, s)
    with Context(0) as fresh:
        want = _valid_step(fresh, s)
    assert np.array_equal(got, want)
    assert np.all(np.isfinite(got)) and np.any(got[3:] != 0)
