"""Training driver of the learned initialiser (GPU kernels in csrc/bia_learn.hip).

1. record: the solver records rows [u1, u2, u3, target] from pairs of given systems that are not
   weak over their dt (the only pairs whose start the network sets);
2. balance: rows are resampled to a flat distribution of the largest |u_k| (pair strength);
3. train: Adam on mini-batches (BF16 MFMA forward/backward, FP32 master weights);
4. tabulate: the network is evaluated on an n^3 grid that the pair kernel interpolates.

Network 0 starts the Kepler drift (BIA_MODE_LEARN), network 1 the reference's own hyperbola and
ellipse Newton iterations (BIA_MODE_PLAIN_LEARN); ``train_reference`` trains network 1 the same way
on rows the reference solvers record.

No reference counterpart.  Row format and features: include/bia_hip.h (learned initialiser).
"""
from __future__ import annotations

import numpy as np

from . import ic

TABLE_N = 64


def record(ctx, systems, per_system: int, seed: int, device="cuda:0"):
    """systems: list of (state [6, n] numpy, eps, dt).  Returns a CUDA float32 tensor of valid rows."""
    import torch

    out = []
    for k, (st, eps, dt) in enumerate(systems):
        s = torch.from_numpy(np.ascontiguousarray(st)).to(device)
        rows = torch.empty((per_system, 4), dtype=torch.float32, device=device)
        ctx.record_samples_(s, eps, dt, per_system, seed * 1000 + k + 1, rows)
        out.append(rows[torch.isfinite(rows[:, 0])])
    return torch.cat(out)


def balance(rows, count: int, seed: int, bins: int = 32):
    """Resample rows to a flat distribution of max_k |u_k| (strength of the pair over dt)."""
    import torch

    m = rows[:, :3].abs().amax(dim=1)
    lo, hi = float(m.min()), float(m.max())
    idx = torch.clamp(((m - lo) / max(hi - lo, 1e-9) * bins).long(), 0, bins - 1)
    counts = torch.bincount(idx, minlength=bins).float()
    wts = 1.0 / counts[idx]
    g = torch.Generator(device=rows.device)
    g.manual_seed(seed)
    pick = torch.multinomial(wts, count, replacement=True, generator=g)
    return rows[pick].contiguous()


def training_systems(seed: int = 2024):
    """Source runs of the recorder: config-1 balls (also at the reference fixture's coupling and
    step, eps = -6e-7, dt = 1), a config-2 Coulomb box (its attractive and repulsive couplings)
    and a Plummer sphere with hard binaries (another seed than the benchmark's), each at step
    lengths where a good share of pairs is strong over dt."""
    s1 = ic.uniform_ball_virial(1332, seed=seed)
    s2, _ = ic.coulomb_box(4096, seed=seed + 1)
    s4 = ic.plummer_with_binaries(65536, seed=seed + 2)
    s5 = ic.uniform_ball_virial(1332, seed=seed + 3)
    return [
        (s5, -6e-7, 1.0),
        (s1, -2.0 / 1332, 1e-2),
        (s1, -2.0 / 1332, 3e-2),
        (s2, -2.0, 5e-3),
        (s2, +2.0, 5e-3),
        (s4, -2.0 / 65536, 1e-2),
        (s4, -2.0 / 65536, 1e-1),
    ]


def train_default(ctx, nsamples: int = 10_000_000, steps: int = 3000, batch: int = 16384, lr: float = 3e-3,
                  seed: int = 2024, log=print):
    """Record, balance, train and tabulate; returns (weights, loss curve, rows used)."""
    import torch

    systems = training_systems(seed)
    raw = record(ctx, systems, per_system=max(nsamples // len(systems), 1), seed=seed)
    rows = balance(raw, nsamples, seed)
    del raw
    ctx.mlp_init(seed)
    curve = ctx.mlp_train_(rows, steps=steps, batch=batch, lr=lr, seed=seed)
    ctx.mlp_build_table(TABLE_N)
    torch.cuda.synchronize()
    log(f"learned initialiser: {rows.shape[0]} rows, loss {curve[0]:.3e} -> {curve[-10:].mean():.3e}")
    return ctx.mlp_get_weights(), curve, rows


REFERENCE_NET = 1


def training_systems_reference(seed: int = 2025):
    """Source runs of the reference solvers' recorder: config-1 balls (N = 1332, virial 0.5, at
    their step and longer ones, and at the reference fixture's coupling), config-2 Coulomb boxes
    (N = 4096, their attractive and repulsive couplings: bound pairs of every eccentricity and
    strongly deflected hyperbolae) and a Plummer sphere with hard binaries (another seed than the
    benchmark's: weak hyperbolae and high-eccentricity binaries)."""
    s1 = ic.uniform_ball_virial(1332, seed=seed)
    s5 = ic.uniform_ball_virial(1332, seed=seed + 3)
    s2, _ = ic.coulomb_box(4096, seed=seed + 1)
    s4 = ic.plummer_with_binaries(65536, seed=seed + 2)
    return [
        (s1, -2.0 / 1332, 1e-3),
        (s1, -2.0 / 1332, 3e-2),
        (s5, -6e-7, 1.0),
        (s2, -2.0, 5e-4),
        (s2, -2.0, 5e-3),
        (s2, +2.0, 5e-4),
        (s4, -2.0 / 65536, 1e-3),
        (s4, -2.0 / 65536, 1e-2),
    ]


def balance_reference(rows, count: int, seed: int, bins: int = 32):
    """Resample reference-solver rows to equal counts per branch (u3 = -1 ellipse, +1 hyperbola),
    within a branch each row weighted by 1 / sqrt(the count of its bin of the eccentricity feature
    u1: log10(e - 1) / 7 for the hyperbola, 1 - 2 sqrt(1 - e) for the ellipse), so rare
    eccentricities are raised without letting them crowd out the common ones."""
    import torch

    br = (rows[:, 2] > 0).long()
    u1 = rows[:, 0]
    lo, hi = float(u1.min()), float(u1.max())
    idx = torch.clamp(((u1 - lo) / max(hi - lo, 1e-9) * bins).long(), 0, bins - 1) + bins * br
    counts = torch.bincount(idx, minlength=2 * bins).float()
    root = counts.sqrt()
    mass = torch.stack([root[:bins].sum(), root[bins:].sum()])  # sum of the weights 1/sqrt(count) per branch
    wts = 1.0 / (root[idx] * mass[br])  # each branch half of the draws
    g = torch.Generator(device=rows.device)
    g.manual_seed(seed)
    pick = torch.multinomial(wts, count, replacement=True, generator=g)
    return rows[pick].contiguous()


def train_reference(ctx, nsamples: int = 10_000_000, steps: int = 3000, batch: int = 16384, lr: float = 3e-3,
                    seed: int = 2025, per_system: int | None = None, log=print):
    """Record, balance, train and tabulate network 1 (the reference solvers' starts); network 0
    is selected again afterwards.  Returns (weights, loss curve, rows used)."""
    import torch

    systems = training_systems_reference(seed)
    ctx.mlp_select(REFERENCE_NET)
    try:
        raw = record(ctx, systems, per_system=per_system or max(nsamples // len(systems), 1), seed=seed)
        rows = balance_reference(raw, nsamples, seed)
        nell = int((raw[:, 2] < 0).sum())
        del raw
        ctx.mlp_init(seed)
        curve = ctx.mlp_train_(rows, steps=steps, batch=batch, lr=lr, seed=seed)
        ctx.mlp_build_table(TABLE_N)
        torch.cuda.synchronize()
        w = ctx.mlp_get_weights()
    finally:
        ctx.mlp_select(0)
    log(f"reference-solver starts: {rows.shape[0]} rows ({nell} recorded elliptic), loss {curve[0]:.3e} -> "
        f"{curve[-10:].mean():.3e}")
    return w, curve, rows
