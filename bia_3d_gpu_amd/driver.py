"""Device-resident multi-step driver with in-run monitoring.

The reference advances one realisation by one BIA step per ``i_calc`` and reports progress with
``cpu_time`` prints (bia/bia.f90:212-213), writes the deltas to ``.BIA`` and particle 1 to a text
line (bia/bia.f90:206-207), and evaluates the invariants once, on the host, after the loop
(bia/bia.f90:236-241).  This driver runs many steps of a whole ensemble without leaving the
device:

* steps run as replays of one captured hipGraph (``bia_run_steps_dev``), in chunks between the
  monitoring and checkpoint steps;
* the per-pair statistics (branch counts, solver evaluations, status bits) accumulate in one
  device counter array during every step; at each monitoring step the running totals are copied,
  on the device, into a per-interval history;
* the invariants (KE, PE, L, momentum and centre sums) of all systems are evaluated on the device
  at each monitoring step (``bia_invariants_batched_dev``: one launch per kernel for the whole
  ensemble) into a ``[n_monitor, nsys, 11]`` array;
* checkpoints are ``.bbb`` files (one record per system, as the reference's multi-realisation
  ``.bbb`` inputs, IC/bbb_maker_REV.f90:71): the state is copied into pinned host memory on the
  stream and a writer thread waits for that copy and writes the file while the next steps run.

Nothing is copied to the host between steps unless a ``log`` callback asks for a line at each
monitoring step.  Monitoring and checkpoints read the state without changing it, so the final
state is bitwise the one of the same number of plain steps.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import io
from .api import INVARIANT_NAMES, METHOD_ALG, MODE_PLAIN, STATS_LEN, Context
from ._lib import COUPLING_SCALAR

BRANCHES = ("hyperbola", "parabola", "ellipse", "rkf")
STATUS_BITS = ("best-iterate never updated", "exact-repeat exit", "iteration cap", "acos clamp",
               "RKF failure", "non-finite", "learned start", "unused")


def schedule(nsteps: int, monitor_every: int = 0, checkpoint_every: int = 0) -> list[tuple[int, bool, bool]]:
    """The steps at which the driver stops: (step, monitor, checkpoint), ascending, ending at
    ``nsteps`` (always monitored).  Step 0 (the initial state) is monitored separately."""
    if nsteps < 0:
        raise ValueError("nsteps must be >= 0")
    stops = {}
    for every, what in ((monitor_every, 0), (checkpoint_every, 1)):
        if every and every > 0:
            for s in range(every, nsteps + 1, every):
                stops.setdefault(s, [False, False])[what] = True
    if nsteps > 0:
        stops.setdefault(nsteps, [False, False])[0] = True
    return [(s, m, c) for s, (m, c) in sorted(stops.items())]


@dataclass
class Report:
    """What a run leaves behind (host copies, made once at the end of the run)."""

    sizes: list[int]
    steps: list[int]                 # monitored steps, 0 first
    invariants: np.ndarray           # [len(steps), nsys, 11] in the order of INVARIANT_NAMES
    stats: np.ndarray | None         # [len(steps), STATS_LEN] running totals at each monitored step
                                     # (None when the run did not collect them)
    checkpoints: list[str] = field(default_factory=list)
    seconds: float = 0.0             # wall time of the run (steps, monitoring, checkpoint copies)

    def invariant(self, name: str) -> np.ndarray:
        """[len(steps), nsys] values of one invariant."""
        return self.invariants[:, :, INVARIANT_NAMES.index(name)]

    def energy(self) -> np.ndarray:
        return self.invariant("KE") + self.invariant("PE")

    def energy_error(self) -> np.ndarray:
        """(E(step) - E(0)) / |E(0)| per monitored step and system."""
        e = self.energy()
        return (e - e[0]) / np.abs(e[0])

    def momentum_error(self) -> np.ndarray:
        """[len(steps), nsys, 3] change of the velocity sums (Ua, Va, Wa) since step 0."""
        m = self.invariants[:, :, 5:8]
        return m - m[0]

    def interval_stats(self) -> np.ndarray:
        """[len(steps) - 1, STATS_LEN] counters of the steps between consecutive monitored steps."""
        if self.stats is None:
            raise ValueError("this run did not collect solver statistics (collect_stats=False)")
        return np.diff(self.stats.astype(np.int64), axis=0)

    def histogram(self, row: np.ndarray | None = None) -> dict:
        """Branch counts, solver evaluations per pair and status-bit counts of a counter row
        (default: the whole run)."""
        if row is None and self.stats is None:
            raise ValueError("this run did not collect solver statistics (collect_stats=False)")
        r = self.stats[-1] if row is None else row
        pairs = int(r[0:4].sum())
        out = {"pairs": pairs}
        for k, b in enumerate(BRANCHES):
            out[b] = int(r[k])
        out["evaluations_per_pair"] = float(r[4:8].sum()) / pairs if pairs else 0.0
        for k, b in enumerate(STATUS_BITS[:7]):
            out[b] = int(r[8 + k])
        return out


class _CheckpointWriter:
    """Writes one ``.bbb`` at a time from a pinned host buffer, after the device copy into it
    (recorded on the stream) has completed."""

    def __init__(self):
        self._thread = None
        self._error = None

    def submit(self, path: str, host, sizes, event):
        self.join()

        def work():
            try:
                event.synchronize()
                st = host.numpy()
                offs = np.concatenate([[0], np.cumsum(sizes)])
                io.write_bbb(path, [st[:, offs[k]:offs[k + 1]] for k in range(len(sizes))])
            except BaseException as e:  # re-raised on the caller's thread by join()
                self._error = e

        self._thread = threading.Thread(target=work, name="bia-checkpoint", daemon=True)
        self._thread.start()

    def join(self):
        if self._thread is not None:
            self._thread.join()
            self._thread = None
        if self._error is not None:
            e, self._error = self._error, None
            raise e


def run(ctx: Context, state, sizes, nsteps: int, *, eps, dt: float, method: int = METHOD_ALG,
        mode: int = MODE_PLAIN, coupling: int = COUPLING_SCALAR, coup: float = 1.0, mass=None, charge=None,
        itr: int = 18, monitor_every: int = 0, checkpoint_every: int = 0, checkpoint_path: str | None = None,
        use_graph: bool = True, collect_stats: bool = True, log=None) -> Report:
    """Advance ``state`` ([6, ld] float64 CUDA tensor, systems of ``sizes`` side by side) by
    ``nsteps`` BIA steps in place.

    ``eps``: one value per system (or one for all).  Every ``monitor_every`` steps (and at the
    last step) the invariants of each system and the running statistics are recorded on the
    device; every ``checkpoint_every`` steps the state is written to
    ``checkpoint_path.format(step=k)`` as a ``.bbb`` file with one record per system.
    ``log(line)``, if given, receives one line per monitoring step (this copies the monitored
    values to the host at that step).  ``collect_stats=False`` runs the steps without the per-pair
    statistics (the pair kernel's statistics variant costs 3-10 % more time per step).

    The invariants are those of the reference's equal-mass scalar coupling (its
    ``Invariance_GPU``); with the gravity or Coulomb couplings they are left NaN and only the
    statistics are monitored.
    """
    import torch

    sizes = [int(n) for n in sizes]
    nsys = len(sizes)
    if nsys < 1 or min(sizes) < 1:
        raise ValueError("sizes must list at least one system of at least one particle")
    ld = int(state.shape[1])
    if state.dtype != torch.float64 or state.dim() != 2 or state.shape[0] != 6 or not state.is_contiguous():
        raise ValueError("state must be a contiguous [6, ld] float64 tensor")
    if sum(sizes) > ld:
        raise ValueError(f"sizes sum to {sum(sizes)} > ld {ld}")
    eps_arr = np.broadcast_to(np.asarray(eps, dtype=np.float64), (nsys,)).copy()
    if checkpoint_every > 0 and not checkpoint_path:
        raise ValueError("checkpoint_every needs checkpoint_path (a format string with {step})")
    plan = schedule(nsteps, monitor_every, checkpoint_every)
    nmon = 1 + sum(1 for _, m, _ in plan if m)

    dev = state.device
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)  # library work and the torch copies below share one stream
    delta = torch.empty_like(state)
    acc = torch.zeros(STATS_LEN, dtype=torch.int64, device=dev) if collect_stats else None
    hist = torch.zeros((nmon, STATS_LEN), dtype=torch.int64, device=dev)
    inv = torch.full((nmon, nsys, 11), float("nan"), dtype=torch.float64, device=dev)
    host = torch.empty((6, ld), dtype=torch.float64, pin_memory=True) if checkpoint_every > 0 else None
    writer = _CheckpointWriter()
    written: list[str] = []
    steps_mon = [0]

    def monitor(k: int, step: int):
        if coupling == COUPLING_SCALAR:
            ctx.invariants_batched_(state, sizes, eps_arr, inv[k])
        if acc is not None:
            hist[k].copy_(acc)
        if log is not None:
            e = inv[:k + 1, :, 0] + inv[:k + 1, :, 1]
            de = ((e[k] - e[0]) / e[0].abs()).abs().max().item()
            line = f"step {step:8d}: max |dE/E0| {de:.3e}"
            if acc is not None:
                row = hist[k].cpu().numpy()
                pairs = int(row[0:4].sum())
                ev = float(row[4:8].sum()) / pairs if pairs else 0.0
                line += (f"; {pairs} pair solves so far, {ev:.4f} evaluations per pair, "
                         f"{int(row[8 + 2])} at the iteration cap")
            log(line)

    t0 = time.perf_counter()
    monitor(0, 0)
    done, k = 0, 1
    for step, mon, ckpt in plan:
        if step > done:
            ctx.run_steps_(state, delta, sizes, step - done, eps=eps_arr, dt=dt, method=method, mode=mode,
                           coupling=coupling, coup=coup, mass=mass, charge=charge, itr=itr, use_graph=use_graph,
                           stats=acc)
            done = step
        if mon:
            monitor(k, step)
            steps_mon.append(step)
            k += 1
        if ckpt:
            writer.join()  # the pinned buffer is free again
            host.copy_(state, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            path = checkpoint_path.format(step=step)
            writer.submit(path, host, sizes, ev)
            written.append(path)
    writer.join()
    stream.synchronize()
    seconds = time.perf_counter() - t0
    return Report(sizes=sizes, steps=steps_mon, invariants=inv.cpu().numpy(),
                  stats=hist.cpu().numpy() if collect_stats else None, checkpoints=written, seconds=seconds)


def load_checkpoint(path: str, device="cuda"):
    """A ``.bbb`` checkpoint (or any multi-realisation ``.bbb``) as a [6, N] CUDA tensor of the
    systems side by side, and the list of their sizes."""
    import torch
    states = io.read_bbb(path)
    sizes = [s.shape[1] for s in states]
    st = np.ascontiguousarray(np.concatenate(states, axis=1))
    return torch.from_numpy(st).to(device), sizes
