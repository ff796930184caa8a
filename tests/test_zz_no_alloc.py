"""The no-allocation step (bia_step_noalloc_dev) against bia_step_dev, through the C ABI: two calls
in a row, the second on a smaller system, in the reference solvers' mode and the Kepler drift;
bitwise the same deltas (the same kernels in the same order, the plan and partials in the
library's static device array instead of the context's buffers).  Sizes beyond the array are
rejected before anything runs."""
import numpy as np
import pytest

from bia_3d_gpu_amd import MODE_GUESS, MODE_PLAIN, ic
from bia_3d_gpu_amd._lib import BIA_ERR_ARG, BIA_OK, load

pytestmark = pytest.mark.gpu


def _step(ctx, name, st, eps, dt, mode):
    import torch
    s = torch.from_numpy(np.ascontiguousarray(st)).cuda()
    d = torch.full_like(s, float("nan"))
    n = s.shape[1]
    if name == "bia_step_dev":
        status = load().bia_step_dev(ctx.handle, 1, mode, n, eps, dt, 18, s.data_ptr(), n, d.data_ptr(), None)
    else:
        status = load().bia_step_noalloc_dev(ctx.handle, 1, mode, n, eps, dt, 18, s.data_ptr(), n, d.data_ptr())
    torch.cuda.synchronize()
    return status, d.cpu().numpy()


@pytest.mark.parametrize("mode", [MODE_PLAIN, MODE_GUESS])
def test_noalloc_step_matches_step_dev(ctx, mode):
    for n, seed in ((3000, 31), (700, 32)):  # the second call smaller than the first
        st = ic.plummer_with_binaries(n, seed)
        eps, dt = -2.0 / n, 1e-3
        s1, want = _step(ctx, "bia_step_dev", st, eps, dt, mode)
        s2, got = _step(ctx, "bia_step_noalloc_dev", st, eps, dt, mode)
        assert s1 == BIA_OK and s2 == BIA_OK
        print(f"mode {mode}, n = {n}: max |difference| {np.abs(got - want).max():.1e}")
        assert np.isfinite(got).all() and np.array_equal(got, want)


def test_noalloc_step_rejects_sizes_beyond_its_array(ctx):
    st = ic.plummer_system(8193, 33)
    status, d = _step(ctx, "bia_step_noalloc_dev", st, -2.0 / 8193, 1e-3, MODE_GUESS)
    assert status == BIA_ERR_ARG
    assert np.isnan(d).all()  # nothing ran
    assert b"8192" in load().bia_last_error(ctx.handle)
