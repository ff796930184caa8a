"""Python host interface of libbia_hip.so.

Two layers, mirroring include/bia_hip.h:

* reference-named wrappers (``find_dxyuv_kernel``, ``invariance_gpu``, ``rkf_c``) with the
  argument meaning of the reference's Fortran call sites (bia/bia.f90:67,135-139;
  dim/dim_3D.f90:30), operating in place on numpy arrays like the Fortran does;
* :class:`Context` for the device-resident whole-step, invariant and DIM entry points.

State layout for the new entry points: structure of arrays ``[6, n]`` (x, y, z, u, v, w).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import METHOD_ALG, METHOD_RKF, MODE_GUESS, MODE_LEARN, MODE_PLAIN, MODE_PLAIN_LEARN, STATS_LEN, check, load

__all__ = ["find_dxyuv_kernel", "invariance_gpu", "rkf_c", "reference_step", "Context", "FailureLog",
           "PAIR_RECORD_DTYPE",
           "METHOD_ALG", "METHOD_RKF", "MODE_PLAIN", "MODE_LEARN", "MODE_GUESS", "MODE_PLAIN_LEARN", "STATS_LEN",
           "flops_per_step"]

INVARIANT_NAMES = ("KE", "PE", "Lx", "Ly", "Lz", "Ua", "Va", "Wa", "Xa", "Ya", "Za")


def _f64(a) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array (updated in place)")
    return a


# ------------------------------------------------------------------------------------------------
# argument checks of the device-tensor entry points: a wrong dtype, a strided or short tensor or a
# host tensor would become an out-of-bounds device access or silently wrong results, so every
# entry point checks its tensors here, before any library call (dtype, layout and size first, the
# device last)
# ------------------------------------------------------------------------------------------------
def _dev(t, name: str, dtype: str, min_numel: int = 0, shape=None):
    """Check a device tensor argument; returns it.  ``dtype`` is a torch dtype name."""
    if t is None or not hasattr(t, "data_ptr") or not hasattr(t, "is_cuda"):
        raise ValueError(f"{name}: expected a torch tensor")
    if str(t.dtype) != "torch." + dtype:
        raise ValueError(f"{name}: expected dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.numel() < min_numel:
        raise ValueError(f"{name}: needs at least {min_numel} elements, has {t.numel()}")
    return t


def _on_device(**tensors):
    """The last check of an entry point: every tensor argument lives on the device."""
    for name, t in tensors.items():
        if t is not None and not t.is_cuda:
            raise ValueError(f"{name}: expected a device (cuda) tensor")


def _state_pair(state, delta=None):
    """[6, ld] float64 device state (and a delta of the same shape); returns ld."""
    if state is None or not hasattr(state, "shape") or len(state.shape) != 2 or state.shape[0] != 6:
        raise ValueError("state: expected a [6, ld] tensor")
    _dev(state, "state", "float64")
    if delta is not None:
        _dev(delta, "delta", "float64", shape=state.shape)
    return int(state.shape[1])


def _sizes_fit(sizes, ld: int) -> np.ndarray:
    sz = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
    if sz.size < 1 or (sz < 1).any():
        raise ValueError("sizes: need at least one system and every size >= 1")
    if int(sz.astype(np.int64).sum()) > ld:
        raise ValueError(f"sizes: the systems hold {int(sz.sum())} particles, the state only {ld} columns")
    return sz


def _stats_arg(stats):
    return None if stats is None else _dev(stats, "stats", "int64", STATS_LEN)


def _couplings(coupling: int, mass, charge, ntot: int):
    if coupling not in (_lib.COUPLING_SCALAR, _lib.COUPLING_GRAVITY, _lib.COUPLING_COULOMB):
        raise ValueError(f"coupling: unknown value {coupling}")
    if coupling == _lib.COUPLING_GRAVITY and mass is None:
        raise ValueError("mass: the gravity coupling needs per-particle masses")
    if coupling == _lib.COUPLING_COULOMB and charge is None:
        raise ValueError("charge: the Coulomb coupling needs per-particle charges")
    if mass is not None:
        _dev(mass, "mass", "float64", ntot)
    if charge is not None:
        _dev(charge, "charge", "float64", ntot)


def _ptr(t):
    return None if t is None else t.data_ptr()


def _rows(rows):
    """float32 sample rows [B, 4] (u1, u2, u3, target) on the device."""
    if rows is None or not hasattr(rows, "shape") or len(rows.shape) != 2 or rows.shape[1] != 4:
        raise ValueError("rows: expected a [B, 4] tensor")
    return _dev(rows, "rows", "float32", 4)


# ------------------------------------------------------------------------------------------------
# reference-named wrappers
# ------------------------------------------------------------------------------------------------
def find_dxyuv_kernel(METHOD: int, N: int, I: int, EPS: float, DT: float, ITR: int,
                      dX, dY, dZ, dU, dV, dW, RXYZUVW, dXYZUVW) -> None:
    """Per-particle pair kernel, bia/find_dXYUV_kernel.cu:526-561 semantics.

    ``I`` is 1-based; ``RXYZUVW[6j:6j+6] = state_I - state_j``.  Accumulates the pair deltas into
    ``dX..dW`` in place (+ at I-1, - at j > I-1) and writes ``dXYZUVW[6j:6j+6]`` for j > I-1.
    """
    arrs = [_f64(a) for a in (dX, dY, dZ, dU, dV, dW, RXYZUVW, dXYZUVW)]
    if any(a.size < N for a in arrs[:6]) or arrs[6].size < 6 * N or arrs[7].size < 6 * N:
        raise ValueError("array shorter than N")
    load().find_dxyuv_kernel_(int(METHOD), int(N), int(I), float(EPS), float(DT), int(ITR), *arrs)


def invariance_gpu(N: int, eps_th: float, X, Y, Z, U, V, W) -> dict:
    """Invariants, bia/invariance.cu:50-110; returns KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za."""
    cols = [np.ascontiguousarray(a, dtype=np.float64) for a in (X, Y, Z, U, V, W)]
    if any(c.size < N for c in cols):
        raise ValueError("array shorter than N")
    outs = [C.c_double() for _ in range(11)]
    load().invariance_gpu_(int(N), float(eps_th), *cols, *[C.byref(o) for o in outs])
    return {k: o.value for k, o in zip(INVARIANT_NAMES, outs)}


def rkf_c(neq: int, eps_th: float, X0: float, XE: float, YH0, INIT: int = 1, RELERR: float = 0.0,
          ABSERR: float = 1e-16, H: float = 0.0) -> dict:
    """DIM direct-sum RKF45, dim/rkf_kernel.cu:194-495 semantics (dim/dim_3D.f90:30 call).

    Note: as in the reference, an integration failure prints the reason and exits the process
    with status 0; use :meth:`Context.dim_rkf45` for an error return instead.
    """
    yh0 = np.ascontiguousarray(YH0, dtype=np.float64).copy()
    yhn = np.zeros_like(yh0)
    ese = np.zeros_like(yh0)
    wl0, wln, werr = np.zeros_like(yh0), np.zeros_like(yh0), np.zeros_like(yh0)
    wrk = np.zeros(11 * neq)
    x0, xe, h = C.c_double(X0), C.c_double(XE), C.c_double(H)
    init, nde, ier = C.c_int(INIT), C.c_int(0), C.c_int(0)
    rel, ab = C.c_double(RELERR), C.c_double(ABSERR)
    load().rkf_c_(int(neq), float(eps_th), C.byref(x0), C.byref(xe), yh0, C.byref(init), C.byref(rel),
                  C.byref(ab), yhn, ese, C.byref(nde), C.byref(ier), C.byref(h), wl0, wln, werr, wrk)
    return dict(YHN=yhn, YH0=yh0, ESE=ese, X0=x0.value, NDE=nde.value, IER=ier.value, H=h.value, INIT=init.value)


def reference_step(state_soa: np.ndarray, eps: float, dt: float, method: int = METHOD_ALG,
                   itr: int = 18) -> np.ndarray:
    """One BIA step driven exactly like the reference host loop (bia/bia.f90:126-196): for each
    particle i, build the relative AoS row, call ``find_dxyuv_kernel``, then add the drift.
    Returns the deltas ``[6, n]``."""
    st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
    n = st.shape[1]
    d = np.zeros((6, n))
    dX, dY, dZ, dU, dV, dW = (np.ascontiguousarray(d[c]) for c in range(6))
    rel = np.zeros(6 * n)
    out = np.zeros(6 * n)
    for i in range(n):
        rel.reshape(n, 6)[:] = (st[:, i:i + 1] - st).T
        find_dxyuv_kernel(method, n, i + 1, eps, dt, itr, dX, dY, dZ, dU, dV, dW, rel, out)
        dX[i] += st[3, i] * dt
        dY[i] += st[4, i] * dt
        dZ[i] += st[5, i] * dt
    return np.stack([dX, dY, dZ, dU, dV, dW])


# ------------------------------------------------------------------------------------------------
# algorithmic work per step (flop convention: + - x / sqrt = 1, transcendental = 20)
# ------------------------------------------------------------------------------------------------
def _items_before(nt: int, seg: int, r: int) -> int:
    """Work items of the first r blocks of four row tiles (csrc/bia_step.hip, Layout): block R has
    ceil(nt / seg) - (4 / seg) R segments of seg column tiles."""
    return r * (-(-nt // seg)) - (4 // seg) * r * (r - 1) // 2


def segment_length(n: int) -> int:
    """Column tiles per work item of an n-particle system (csrc/bia_step.hip, choose_seg): the
    longest of 4, 2, 1 that still gives the system 8192 items or more, else 1."""
    nt = (n + 63) // 64
    nb = (nt + 3) // 4
    for seg in (4, 2):
        if _items_before(nt, seg, nb) >= 8192:
            return seg
    return 1


def partial_doubles(n: int, blocks: int | None = None) -> int:
    """Doubles of the pair kernel's partial sums for one n-particle system (the first ``blocks``
    blocks of four row tiles; all of them by default), as the library lays them out
    (csrc/bia_step.hip, Layout): block R holds its column partials for the columns from 256 R to
    the padded end, 6 (npad - 256 R), then one row partial set of 6 x 256 per work item."""
    nt = (n + 63) // 64
    npad, nb = 64 * nt, (nt + 3) // 4
    r = nb if blocks is None else blocks
    return 6 * (r * npad - 128 * r * (r - 1)) + 6 * 256 * _items_before(nt, segment_length(n), r)


def staging(n: int, limit_bytes: int) -> tuple[int, int]:
    """(stages, workspace bytes) of one n-particle system under a workspace limit (0 = none): the
    largest G blocks per stage whose first stage fits the limit (at least one block)."""
    nt = (n + 63) // 64
    nb = (nt + 3) // 4
    full = 8 * partial_doubles(n)
    if limit_bytes <= 0 or full <= limit_bytes:
        return 1, full
    g = 1
    while g < nb and 8 * partial_doubles(n, g + 1) <= limit_bytes:
        g += 1
    if g >= nb:
        return 1, full
    return -(-nb // g), 8 * partial_doubles(n, g)


def flops_per_step(stats, mode: int = MODE_PLAIN) -> float:
    """Algorithmic FP64 flops of one step from its BIA_STATS counters (+ - x / sqrt = 1 flop,
    each libm transcendental = 20).

    Plain mode (the reference algorithm), per pair, with k = Newton corrections after the two
    set-up evaluations (hyperbola) or Newton corrections (ellipse, parabola) and s accepted RKF
    steps, counted from the reference GPU source: hyperbola 635 + 122 k, ellipse 524 + 47 k,
    parabola 472 + 76 k, RKF 80 + 790 s.

    Guess and learned modes (universal-variable Kepler drift, csrc/bia_pair.hpp kepler_pair),
    counted from that source with a reciprocal or reciprocal square root as one division (and
    one square root), comparisons and min/max not counted: 109 per pair (relative state 6,
    kep_setup 23, features 7, series start 9, kep_output 52 with its Taylor carry of the G
    functions, accumulation 12) + 58 per evaluation of the time equation (kep_g 22, kep_halley
    32, step and convergence test 4), + 120 per learned start (three log2 features, trilinear lookup,
    exp, scale).  The bound-orbit period reduction, taken by few pairs, is not counted.

    Reference solvers with learned starts: the plain-mode counts with the evaluations the pairs
    made (the learned start's one evaluation takes the place of the reference's two, which the
    hyperbola's 635 include), + 80 per learned start (two log2 features, a reciprocal, the
    bilinear lookup, expm1 or the ellipse's reduction).
    """
    s = np.asarray(stats, dtype=np.float64)
    if mode not in (MODE_PLAIN, MODE_PLAIN_LEARN):
        return 109 * s[0:4].sum() + 58 * s[4:8].sum() + 120 * s[14]
    n_h, n_p, n_e, n_r = s[0:4]
    it_h, it_p, it_e, it_r = s[4:8]
    k_h = it_h - 2 * n_h
    extra = 80 * s[14] if mode == MODE_PLAIN_LEARN else 0.0
    return (635 * n_h + 122 * k_h) + (472 * n_p + 76 * it_p) + (524 * n_e + 47 * it_e) + (80 * n_r + 790 * it_r) + extra


# ------------------------------------------------------------------------------------------------
# per-pair failure records (bia_set_failure_log)
# ------------------------------------------------------------------------------------------------
PAIR_RECORD_DTYPE = np.dtype([("sys", "<i4"), ("i", "<i4"), ("j", "<i4"), ("branch", "<i4"), ("iters", "<i4"),
                              ("flags", "<u4"), ("error", "<f8")])
assert PAIR_RECORD_DTYPE.itemsize == 32


class FailureLog:
    """Device buffers of bia_set_failure_log: room for ``capacity`` records of the pairs a solver
    could not finish (iteration cap, RKF45 failure, non-finite output) and their count, which
    includes the pairs that did not fit.  The reference's CPU loop writes the same pairs to its
    units 88 / 99 (bia/bia.f90:151-154,167-170)."""

    def __init__(self, capacity: int = 4096, device: str = "cuda"):
        import torch
        if capacity < 0:
            raise ValueError("capacity must be >= 0")
        self.capacity = int(capacity)
        self.buf = torch.zeros(max(1, self.capacity) * PAIR_RECORD_DTYPE.itemsize, dtype=torch.uint8, device=device)
        self.count = torch.zeros(1, dtype=torch.int64, device=device)

    def reset(self):
        self.count.zero_()

    def total(self) -> int:
        """Pairs logged since the last reset (written or not)."""
        return int(self.count.item())

    def records(self) -> np.ndarray:
        """The written records, sorted by (sys, i, j) (the device writes them in no particular order)."""
        n = min(self.total(), self.capacity)
        raw = self.buf[: n * PAIR_RECORD_DTYPE.itemsize].cpu().numpy()
        return np.sort(raw.view(PAIR_RECORD_DTYPE), order=["sys", "i", "j"])


# ------------------------------------------------------------------------------------------------
# context API
# ------------------------------------------------------------------------------------------------
class Context:
    """Owns a bia_ctx (device workspace, stream).  One per host thread."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        check(L.bia_create(C.byref(h), int(device)), "bia_create")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            load().bia_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_handle: int):
        check(load().bia_set_stream(self._h, C.c_void_p(int(stream_handle))), "bia_set_stream", self._h)

    def set_profiling(self, on: bool = True):
        check(load().bia_set_profiling(self._h, int(on)), "bia_set_profiling", self._h)

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(load().bia_last_kernel_ms(self._h, C.byref(ms)), "bia_last_kernel_ms", self._h)
        return ms.value

    def kernel_ms(self, max_launches: int = 256) -> list[float]:
        """Durations of the pair-kernel launches recorded since profiling was switched on (the
        last 256 at most), oldest first; waits for them."""
        buf = (C.c_float * max_launches)()
        n = C.c_int()
        check(load().bia_kernel_ms(self._h, buf, int(max_launches), C.byref(n)), "bia_kernel_ms", self._h)
        return [buf[i] for i in range(n.value)]

    def reduce_kernel_ms(self, max_launches: int = 256) -> list[float]:
        """Durations of the tile-partial reduction kernel of the same steps, oldest first."""
        buf = (C.c_float * max_launches)()
        n = C.c_int()
        check(load().bia_reduce_kernel_ms(self._h, buf, int(max_launches), C.byref(n)), "bia_reduce_kernel_ms",
              self._h)
        return [buf[i] for i in range(n.value)]

    def synchronize(self):
        check(load().bia_synchronize(self._h), "bia_synchronize", self._h)

    def set_workspace_limit(self, nbytes: int = 1 << 30):
        """Largest tile-partial workspace of a step (bytes, 0 = no limit); a single system beyond it
        runs in stages of row tiles with the same deltas bitwise."""
        check(load().bia_set_workspace_limit(self._h, int(nbytes)), "bia_set_workspace_limit", self._h)

    def plan_info(self) -> tuple[int, int]:
        """(stages, tile-partial workspace bytes) of the last step's plan."""
        ns, wb = C.c_int(), C.c_ulonglong()
        check(load().bia_plan_info(self._h, C.byref(ns), C.byref(wb)), "bia_plan_info", self._h)
        return ns.value, wb.value

    def set_rkf_tolerance(self, relerr: float = 0.0, abserr: float = 1e-16):
        """Per-pair RKF45 tolerances of the following steps (the reference's 0 / 1e-16 by
        default, bia/bia-rkf.cu:5)."""
        check(load().bia_set_rkf_tolerance(self._h, float(relerr), float(abserr)), "bia_set_rkf_tolerance", self._h)

    def set_failure_log(self, log: "FailureLog | None"):
        """Record the pairs the solvers could not finish in ``log`` during the following steps
        (None switches it off).  Steps with a log run the pair kernel's statistics variant."""
        if log is None:
            check(load().bia_set_failure_log(self._h, None, 0, None), "bia_set_failure_log", self._h)
            self._log = None
            return
        _dev(log.buf, "log.buf", "uint8", log.capacity * PAIR_RECORD_DTYPE.itemsize)
        _dev(log.count, "log.count", "int64", 1)
        _on_device(buf=log.buf, count=log.count)
        check(load().bia_set_failure_log(self._h, log.buf.data_ptr(), log.capacity, log.count.data_ptr()),
              "bia_set_failure_log", self._h)
        self._log = log  # the buffers must outlive the steps that write them

    # -- whole step on device memory (torch tensors) -------------------------------------------
    def step_(self, state, delta, eps: float, dt: float, method: int = METHOD_ALG, mode: int = MODE_PLAIN,
              itr: int = 18, stats=None):
        """Device step: ``state`` and ``delta`` are CUDA/HIP float64 tensors ``[6, ld]``."""
        n = _state_pair(state, delta)
        _stats_arg(stats)
        _on_device(state=state, delta=delta, stats=stats)
        check(load().bia_step_dev(self._h, method, mode, n, eps, dt, itr, state.data_ptr(), n, delta.data_ptr(),
                                  _ptr(stats)), "bia_step_dev", self._h)
        return delta

    # -- whole step on host arrays --------------------------------------------------------------
    def step_host(self, state_soa: np.ndarray, eps: float, dt: float, method: int = METHOD_ALG,
                  mode: int = MODE_PLAIN, itr: int = 18, with_stats: bool = False):
        st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
        n = st.shape[1]
        cols = [np.ascontiguousarray(st[c]) for c in range(6)]
        outs = [np.zeros(n) for _ in range(6)]
        stats = np.zeros(STATS_LEN, dtype=np.uint64)
        check(load().bia_step_host(self._h, method, mode, n, eps, dt, itr, *cols, *outs,
                                   stats.ctypes.data_as(C.c_void_p)), "bia_step_host", self._h)
        d = np.stack(outs)
        return (d, stats) if with_stats else d

    def invariants(self, state_soa: np.ndarray, eps: float) -> dict:
        st = np.ascontiguousarray(state_soa, dtype=np.float64).reshape(6, -1)
        out = np.zeros(11)
        cols = [np.ascontiguousarray(st[c]) for c in range(6)]
        check(load().bia_invariants_host(self._h, st.shape[1], eps, *cols, out), "bia_invariants_host", self._h)
        return dict(zip(INVARIANT_NAMES, out.tolist()))

    def invariants_(self, state, out, eps: float):
        """Device invariants: ``state`` [6, n] float64 tensor, ``out`` 11-element float64 tensor."""
        _state_pair(state)
        _dev(out, "out", "float64", 11)
        _on_device(state=state, out=out)
        check(load().bia_invariants_dev(self._h, state.shape[1], eps, state.data_ptr(), state.shape[1],
                                        out.data_ptr()), "bia_invariants_dev", self._h)
        return out

    # -- batched / coupled steps, autotuner, multi-step driver ---------------------------------
    def _batched_args(self, sizes, eps, coupling):
        sizes = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
        if coupling == _lib.COUPLING_SCALAR and eps is None:
            raise ValueError("eps: the scalar coupling needs eps (one value or one per system)")
        e = np.ascontiguousarray(np.zeros(len(sizes)) if eps is None else np.broadcast_to(eps, len(sizes)),
                                 dtype=np.float64)
        return sizes, e

    def _step_args(self, state, delta, sizes, eps, coupling, mass, charge, stats=None):
        """Checks of the batched entry points' tensors; returns (sizes, eps) for the call."""
        ld = _state_pair(state, delta)
        sz = _sizes_fit(sizes, ld)
        _couplings(coupling, mass, charge, int(sz.astype(np.int64).sum()))
        _stats_arg(stats)
        out = self._batched_args(sz, eps, coupling)
        _on_device(state=state, delta=delta, mass=mass, charge=charge, stats=stats)
        return out

    def step_batched_(self, state, delta, sizes, eps=None, dt: float = 1e-3, method: int = METHOD_ALG,
                      mode: int = MODE_PLAIN, coupling: int = 0, coup: float = 1.0, mass=None, charge=None,
                      itr: int = 18, stats=None):
        """Systems of the given sizes concatenated column-wise in state/delta ([6, ld] CUDA f64)."""
        sz, e = self._step_args(state, delta, sizes, eps, coupling, mass, charge, stats)
        ptr = _ptr
        check(load().bia_step_batched_dev(self._h, method, mode, coupling, len(sz), sz, e, coup, dt, itr,
                                          state.data_ptr(), state.shape[1], ptr(mass), ptr(charge), delta.data_ptr(),
                                          ptr(stats)), "bia_step_batched_dev", self._h)
        return delta

    def autotune_(self, state, delta, sizes, eps=None, dt: float = 1e-3, method: int = METHOD_ALG,
                  mode: int = MODE_PLAIN, coupling: int = 0, coup: float = 1.0, mass=None, charge=None, itr: int = 18):
        sz, e = self._step_args(state, delta, sizes, eps, coupling, mass, charge)
        ptr = _ptr
        ms, g = C.c_float(), C.c_int()
        check(load().bia_autotune_dev(self._h, method, mode, coupling, len(sz), sz, e, coup, dt, itr,
                                      state.data_ptr(), state.shape[1], ptr(mass), ptr(charge), delta.data_ptr(),
                                      C.byref(ms), C.byref(g)), "bia_autotune_dev", self._h)
        return ms.value, g.value // 10, g.value % 10

    def run_steps_(self, state, delta, sizes, nsteps: int, eps=None, dt: float = 1e-3, method: int = METHOD_ALG,
                   mode: int = MODE_PLAIN, coupling: int = 0, coup: float = 1.0, mass=None, charge=None,
                   itr: int = 18, use_graph: bool = True, stats=None):
        """nsteps of (step, state += delta) on the device; ``stats`` (int64 CUDA tensor of
        STATS_LEN, optional) accumulates the per-pair statistics of every step."""
        if int(nsteps) < 0:
            raise ValueError("nsteps must be >= 0")
        sz, e = self._step_args(state, delta, sizes, eps, coupling, mass, charge, stats)
        ptr = _ptr
        check(load().bia_run_steps_dev(self._h, method, mode, coupling, len(sz), sz, e, coup, dt, itr,
                                       state.data_ptr(), state.shape[1], ptr(mass), ptr(charge), delta.data_ptr(),
                                       nsteps, int(use_graph), ptr(stats)), "bia_run_steps_dev", self._h)
        return state

    def invariants_at_(self, state, offset: int, n: int, eps: float, out):
        """Device invariants of the n particles at column ``offset`` of a [6, ld] state (one system
        of an ensemble); ``out`` is an 11-element float64 CUDA tensor (any offset into a larger one)."""
        ld = _state_pair(state)
        _dev(out, "out", "float64", 11)
        if int(offset) < 0 or int(n) < 1 or int(offset) + int(n) > ld:
            raise ValueError("offset / n: the system must lie inside the state's columns")
        _on_device(state=state, out=out)
        ptr = state.data_ptr() + 8 * int(offset)
        check(load().bia_invariants_dev(self._h, int(n), float(eps), ptr, ld, out.data_ptr()),
              "bia_invariants_dev", self._h)
        return out

    def invariants_batched_(self, state, sizes, eps, out):
        """Device invariants of every system of a [6, ld] state (systems of ``sizes`` side by
        side, ``eps`` per system or one for all) into ``out``, a float64 CUDA tensor of
        ``len(sizes) * 11`` elements (e.g. [nsys, 11]); one launch per kernel for the list."""
        sz, e = self._batched_args(_sizes_fit(sizes, _state_pair(state)), eps, 0)
        _dev(out, "out", "float64", 11 * len(sz))
        _on_device(state=state, out=out)
        check(load().bia_invariants_batched_dev(self._h, len(sz), sz, e, state.data_ptr(), state.shape[1],
                                                out.data_ptr()), "bia_invariants_batched_dev", self._h)
        return out

    def invariants_coupled_(self, state, sizes, out, coupling: int, coup: float = 1.0, mass=None, charge=None):
        """Mass- and charge-weighted invariants of every system (``mass`` / ``charge`` float64
        CUDA tensors with the state's columns): KE = sum m|v|^2, PE = -2G sum m_i m_j / r or
        2k sum q_i q_j / r, so KE + PE = 2E; L and the sums weighted by the masses."""
        if coupling not in (_lib.COUPLING_GRAVITY, _lib.COUPLING_COULOMB):
            raise ValueError("coupling: the weighted invariants take the gravity or Coulomb coupling")
        sz = _sizes_fit(sizes, _state_pair(state))
        _couplings(coupling, mass, charge, int(sz.astype(np.int64).sum()))
        _dev(out, "out", "float64", 11 * len(sz))
        _on_device(state=state, out=out, mass=mass, charge=charge)
        ptr = _ptr
        check(load().bia_invariants_coupled_batched_dev(self._h, coupling, coup, len(sz), sz, state.data_ptr(),
                                                        state.shape[1], ptr(mass), ptr(charge), out.data_ptr()),
              "bia_invariants_coupled_batched_dev", self._h)
        return out

    # -- learned initialiser ---------------------------------------------------------------------
    def mlp_select(self, net: int):
        """Make network ``net`` (0: the Kepler drift's start, 1: the reference solvers' starts) the
        one the following mlp_* and record_samples_ calls act on."""
        check(load().bia_mlp_select(self._h, int(net)), "bia_mlp_select", self._h)

    def mlp_init(self, seed: int = 1):
        check(load().bia_mlp_init(self._h, int(seed)), "bia_mlp_init", self._h)

    def mlp_set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(load().bia_mlp_set_weights(self._h, w, w.size), "bia_mlp_set_weights", self._h)

    def mlp_get_weights(self) -> np.ndarray:
        n = load().bia_mlp_nparam()
        w = np.zeros(n, dtype=np.float32)
        check(load().bia_mlp_get_weights(self._h, w, n), "bia_mlp_get_weights", self._h)
        return w

    def mlp_forward_(self, rows, y):
        """rows: float32 CUDA tensor [B, 4]; y: float32 CUDA tensor [B]."""
        _rows(rows)
        _dev(y, "y", "float32", rows.shape[0])
        _on_device(rows=rows, y=y)
        check(load().bia_mlp_forward_dev(self._h, rows.data_ptr(), rows.shape[0], y.data_ptr()),
              "bia_mlp_forward_dev", self._h)
        return y

    def mlp_grad_(self, rows, grad):
        """Full-batch gradient of mean((y - t)^2): grad is a float32 CUDA tensor of nparam + 1."""
        _rows(rows)
        _dev(grad, "grad", "float32", 4482)  # bia_mlp_nparam() + the loss
        _on_device(rows=rows, grad=grad)
        check(load().bia_mlp_grad_dev(self._h, rows.data_ptr(), rows.shape[0], grad.data_ptr()),
              "bia_mlp_grad_dev", self._h)
        return grad

    def mlp_train_(self, rows, steps: int, batch: int = 8192, lr: float = 3e-3, seed: int = 7) -> np.ndarray:
        _rows(rows)
        _on_device(rows=rows)
        curve = np.zeros(steps, dtype=np.float32)
        check(load().bia_mlp_train_dev(self._h, rows.data_ptr(), rows.shape[0], steps, batch, lr, int(seed),
                                       curve.ctypes.data_as(C.c_void_p)), "bia_mlp_train_dev", self._h)
        return curve

    def mlp_build_table(self, n: int = 64):
        check(load().bia_mlp_build_table(self._h, n), "bia_mlp_build_table", self._h)

    def record_samples_(self, state, eps: float, dt: float, nsamples: int, seed: int, out):
        """out: float32 CUDA tensor [nsamples, 4]."""
        _state_pair(state)
        _dev(out, "out", "float32", 4 * int(nsamples))
        _on_device(state=state, out=out)
        check(load().bia_record_samples_dev(self._h, state.shape[1], eps, dt, state.data_ptr(), state.shape[1],
                                            nsamples, int(seed), out.data_ptr()), "bia_record_samples_dev", self._h)
        return out

    def dim_rkf45(self, y_aos: np.ndarray, eps: float, x0: float, xe: float, relerr: float = 0.0,
                  abserr: float = 1e-16, h: float = 0.0, init: int = 1, coupling: int = _lib.COUPLING_SCALAR,
                  coup: float = 1.0, mass=None, charge=None) -> dict:
        """DIM RKF45 of the AoS state y_aos [N, 6] from x0 to xe.  With the scalar coupling every
        pair has the reference's eps; with COUPLING_GRAVITY / COUPLING_COULOMB (coup = G or k) the
        force is weighted by the host arrays mass / charge and eps is unused."""
        y = np.ascontiguousarray(y_aos, dtype=np.float64).reshape(-1).copy()
        ese = np.zeros_like(y)
        x = C.c_double(x0)
        hh = C.c_double(h)
        ini = C.c_int(init)
        nde, nacc, nrej, ier = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        tail = (C.byref(hh), ese.ctypes.data_as(C.c_void_p), C.byref(nde), C.byref(nacc), C.byref(nrej), C.byref(ier))
        if coupling == _lib.COUPLING_SCALAR:
            name = "bia_dim_rkf45_host"
            s = load().bia_dim_rkf45_host(self._h, y.size // 6, eps, C.byref(x), xe, y, C.byref(ini), relerr, abserr,
                                          *tail)
        else:
            def host(a):
                return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            m, q = host(mass), host(charge)
            for a in (m, q):
                if a is not None and a.size != y.size // 6:
                    raise ValueError("mass / charge need one value per particle")
            name = "bia_dim_rkf45_coupled_host"
            s = load().bia_dim_rkf45_coupled_host(self._h, coupling, coup, y.size // 6,
                                                  None if m is None else m.ctypes.data_as(C.c_void_p),
                                                  None if q is None else q.ctypes.data_as(C.c_void_p),
                                                  C.byref(x), xe, y, C.byref(ini), relerr, abserr, *tail)
        if s not in (_lib.BIA_OK, _lib.BIA_ERR_RKF):
            check(s, name, self._h)
        return dict(y=y, ese=ese, x=x.value, h=hh.value, nde=nde.value, naccept=nacc.value,
                    nreject=nrej.value, ier=ier.value)
