"""The learned membrane substepper end to end: training rows from plain
substepping, feature/target normalisation, training on the GPU with the
library's BF16-MFMA kernels, residual threshold from the plain method, and
the surrogate-vs-plain comparison (deformation index, cell volume, wall
time per simulated time).

Every computation runs in libgpucell_hip.so (engine, MLP); torch holds the
device buffers and does the column statistics of the training rows.
"""
from __future__ import annotations

import time

import numpy as np

from . import analysis
from .engine import Engine
from .surrogate import Mlp

NF, NO = 36, 24


def collect(cases, steps: int, every: int = 1, warmup: int = 5, max_rows: int = 10 ** 6,
            k: int = 8):
    """Training rows from plain substepping (mode 1, k = 8): each case runs
    `warmup` steps, then `steps` steps of which every `every`-th is recorded.
    Returns CUDA tensors X [n][36], Y [n][24] (raw)."""
    import torch
    per_case = []
    for c in cases:
        n_rec = len(range(0, steps, every))
        per_case.append(min(n_rec * c.nCellPar, max_rows))
    total = min(sum(per_case), max_rows)
    X = torch.zeros((total, NF), dtype=torch.float32, device="cuda")
    Y = torch.zeros((total, NO), dtype=torch.float32, device="cuda")
    row = 0
    for c in cases:
        e = Engine(c)
        e.set_membrane_mode(1, k=k)
        e.step(warmup)
        for s in range(steps):
            if s % every == 0 and row + c.nCellPar <= total:
                e.record_samples(X[row:], Y[row:])
                e.step(1)
                row += c.nCellPar
                e.record_samples(None, None)
            else:
                e.step(1)
        e.close()
    return X[:row], Y[:row]


def fit_norm(X, Y) -> np.ndarray:
    """mean[36] std[36] mean[24] std[24] of the raw rows (std floored)."""
    import torch
    def ms(t):
        t64 = t.double()
        m = t64.mean(0)
        s = t64.std(0).clamp_min(1e-6)
        return m.cpu().numpy(), s.cpu().numpy()
    mx, sx = ms(X)
    my, sy = ms(Y)
    # constant features (the reserved column, the six-neighbour pad of a
    # mesh with no five-neighbour vertex) keep unit scale
    sx = np.where(sx < 1e-5, 1.0, sx)
    sy = np.where(sy < 1e-12, 1.0, sy)
    return np.concatenate([mx, sx, my, sy]).astype(np.float32)


def normalised(X, Y, norm):
    import torch
    n = torch.from_numpy(norm).to(X.device)
    Xn = (X - n[:NF]) / n[NF:2 * NF]
    Yn = (Y - n[2 * NF:2 * NF + NO]) / n[2 * NF + NO:]
    return Xn.contiguous(), Yn.contiguous()


def train(X, Y, epochs: int = 8, batch: int = 4096, lr: float = 2e-3, seed: int = 1):
    """Returns (Mlp, norm, loss per epoch, seconds)."""
    import torch
    norm = fit_norm(X, Y)
    Xn, Yn = normalised(X, Y, norm)
    net = Mlp(seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = net.train(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    torch.cuda.synchronize()
    return net, norm, loss, time.perf_counter() - t0


def plain_residual_scale(case, steps: int = 6, warmup: int = 2) -> float:
    """Largest residual R of the exact (plain, 8-substep) trajectories."""
    e = Engine(case)
    e.set_membrane_mode(1, k=8)
    e.step(warmup)
    r = 0.0
    for _ in range(steps):
        e.step(1)
        r = max(r, e.membrane_stats()["plain_residual"])
    e.close()
    return r


def _cell_metrics(e, case):
    S = e.state()
    cells = analysis.unwrap_cells(S, case)
    di = np.array([analysis.deformation_index(cells[c], analysis.cell_tri0(case, c))
                   for c in range(case.nCell)])
    return di, S["vol"].astype(np.float64).copy()


def compare(case, net, norm, threshold: float, steps: int, warmup: int = 2, graph: bool = True,
            relax: int = 600, skin: float = 0.0):
    """Plain (mode 1, k = 8) and surrogate (mode 2) runs of `steps` steps
    from the same state (engines with Verlet skin `skin`, 0 = per-step
    neighbour build); returns accuracy and wall time per step.  All start
    from the case relaxed for `relax` reference steps (mode 0): a freshly
    placed suspension goes through a start-up transient (pressure waves from
    the lattice start and the clearance around the cells) in which cell
    volumes swing by ~1 % over the first few hundred steps.  The volume
    change over the measured steps is reported for both modes and for the
    same steps without substepping (mode 0): the fluid's pressure
    fluctuations change cell volumes alike in all three, so the difference
    to mode 0 is what the substepping method adds.  With `graph` the modes
    replay their steps from a HIP graph (the surrogate's fallback is decided
    on the device, so neither mode reads back per step)."""
    import torch
    out = {}
    runs = {}
    for mode in (0, 1, 2):
        e = Engine(case, skin=skin)
        if relax:
            e.step(relax, graph=graph)
        if mode == 1:
            e.set_membrane_mode(1, k=8)
        elif mode == 2:
            e.set_membrane_mode(2, k=8, mlp=net, norm=norm, threshold=threshold)
        e.step(warmup, graph=graph)
        di0, v0 = _cell_metrics(e, case)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.step(steps, graph=graph)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        di, v = _cell_metrics(e, case)
        runs[mode] = {"di": di, "vol": v, "vol0": v0, "ms_per_step": 1e3 * dt,
                      "stats": e.membrane_stats() if mode else None}
        e.close()
    p, s = runs[1], runs[2]
    out["ms_per_step_plain"] = p["ms_per_step"]
    out["ms_per_step_surrogate"] = s["ms_per_step"]
    out["speedup"] = p["ms_per_step"] / s["ms_per_step"]
    out["di_rel_err_max"] = float(np.max(np.abs(s["di"] - p["di"]) / np.maximum(np.abs(p["di"]), 1e-12)))
    out["vol_rel_diff_max"] = float(np.max(np.abs(s["vol"] - p["vol"]) / p["vol"]))
    out["vol_drift_surrogate_max"] = float(np.max(np.abs(s["vol"] - s["vol0"]) / s["vol0"]))
    out["vol_drift_plain_max"] = float(np.max(np.abs(p["vol"] - p["vol0"]) / p["vol0"]))
    z = runs[0]
    out["vol_change_no_substeps_max"] = float(np.max(np.abs(z["vol"] - z["vol0"]) / z["vol0"]))
    out["ms_per_step_no_substeps"] = z["ms_per_step"]
    st = s["stats"]
    out["fallback_rate"] = st["fallbacks"] / max(1, st["steps"])
    out["steps"] = steps
    out["relax_steps"] = relax
    out["graph"] = bool(graph)
    return out
