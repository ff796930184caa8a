"""Training driver of the learned initialiser (GPU kernels in csrc/bia_learn.hip).

1. record: the solver records (x1, v, b, t) rows from random hyperbolic pairs of given systems;
2. balance: rows are resampled to a flat distribution in log e (e = v^-4);
3. train: Adam on mini-batches (BF16 MFMA forward/backward, FP32 master weights);
4. tabulate: the network is evaluated on a grid that the pair kernel interpolates.

No reference counterpart.  Row format and features: include/bia_hip.h (learned initialiser).
"""
from __future__ import annotations

import numpy as np

from . import ic

NX, NV = 1024, 256


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


def balance_log_e(rows, count: int, seed: int, bins: int = 64):
    """Resample rows to a flat distribution in log10(e) (e = v^-4)."""
    import torch

    loge = -4.0 * torch.log10(rows[:, 1].clamp_min(1e-12))
    lo, hi = float(loge.min()), float(loge.max())
    idx = torch.clamp(((loge - lo) / max(hi - lo, 1e-9) * bins).long(), 0, bins - 1)
    counts = torch.bincount(idx, minlength=bins).float()
    wts = 1.0 / counts[idx]
    g = torch.Generator(device=rows.device)
    g.manual_seed(seed)
    pick = torch.multinomial(wts, count, replacement=True, generator=g)
    return rows[pick].contiguous()


def training_systems(seed: int = 2024):
    """Config-1 and config-2 style systems (the recorder's source runs) plus a Plummer-with-binaries
    system of another seed than the benchmark's, so the samples span e from ~1 to ~1e6."""
    s1 = ic.uniform_ball_virial(1332, seed=seed)
    s2, q = ic.coulomb_box(4096, seed=seed + 1)
    s4 = ic.plummer_with_binaries(65536, seed=seed + 2)
    return [
        (s1, -2.0 / 1332, 1e-3),          # config 1: gravity, eps = -2Gm
        (s2, -2.0, 5e-4),                 # config 2: attractive Coulomb pairs (eps = -2 k)
        (s2, +2.0, 5e-4),                 # config 2: repulsive Coulomb pairs
        (s4, -2.0 / 65536, 1e-3),         # Plummer + binaries
    ]


def train_default(ctx, nsamples: int = 10_000_000, steps: int = 3000, batch: int = 16384, lr: float = 3e-3,
                  seed: int = 2024, log=print):
    """Record, balance, train and tabulate; returns (weights, loss curve, rows used)."""
    import torch

    raw = record(ctx, training_systems(seed), per_system=max(nsamples // 2, 1), seed=seed)
    rows = balance_log_e(raw, nsamples, seed)
    del raw
    ctx.mlp_init(seed)
    curve = ctx.mlp_train_(rows, steps=steps, batch=batch, lr=lr, seed=seed)
    ctx.mlp_build_table(NX, NV)
    torch.cuda.synchronize()
    log(f"learned initialiser: {rows.shape[0]} rows, loss {curve[0]:.3e} -> {curve[-10:].mean():.3e}")
    return ctx.mlp_get_weights(), curve, rows
