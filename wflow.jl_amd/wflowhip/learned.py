"""Training the learned Newton warm start from the plain solver's own solves.

`harvest` steps a parameter ensemble with the plain solver and collects (features, target) rows
of sampled kinematic-wave solves on the device; `fit` trains the network (wfh_mlp_*) on them with
the library's matrix-core kernels. The network then proposes the start of the ensemble's solves
(Ensemble.set_learned); the iteration and its stopping rule stay the reference's
(surface_process.jl:24-70).
"""
import numpy as np

from .model import Ensemble, Mlp


def harvest(catchment, factors, steps, forcing, every=1, cap=2_000_000, **cfg):
    """Rows [n, 10] (torch, device): features [8], target log10 q - 5 log10 u_ref, Newton updates,
    from `steps` plain steps of the ensemble `factors` (kv_0, f, n multipliers) on `catchment`.
    forcing(k) returns the step's (P, E, T) (numpy, land order, or device tensors)."""
    import torch

    rows = torch.empty((cap, 10), dtype=torch.float32, device="cuda")
    ens = Ensemble(catchment, *factors, **cfg)
    ens.set_harvest(rows, every=every)
    for k in range(steps):
        ens.run_timestep(*forcing(k))
    n = ens.harvest_count()
    ens.set_harvest(None)
    ens.close()
    return rows[:n].contiguous()


def fit(rows, epochs=6, batch=8192, lr=2e-3, seed=0, device=0, log=None):
    """A network trained on harvested rows; returns (network, per-epoch losses)."""
    import torch

    x = rows[:, :8].contiguous()
    t_raw = rows[:, 8].contiguous()
    xm = x.mean(0).cpu().numpy().astype(np.float32)
    xs = np.maximum(x.std(0).cpu().numpy().astype(np.float32), 1e-3)
    ym = float(t_raw.mean())
    ys = max(float(t_raw.std()), 1e-6)
    t = ((t_raw - ym) / ys).contiguous()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    net = Mlp(device=device, seed=seed)
    net.set_normalization(xm, xs, ym, ys)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(x.shape[0], device="cuda", generator=g)
        losses.append(net.train_epoch(x[perm].contiguous(), t[perm].contiguous(), batch=batch, lr=lr))
        if log:
            log(f"epoch {ep}: mse {losses[-1]:.4f} (normalised target)")
    return net, losses


def mean_updates(hist):
    """Mean Newton updates per solve from a kw_histogram row (bins of KW_BIN_EDGES)."""
    mids = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 12.5, 24.5, 48.5, 96.5, 192.5, 1628, 3000], np.float64)
    h = np.asarray(hist, np.float64)
    return float((h * mids).sum() / max(h.sum(), 1.0))
