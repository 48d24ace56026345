"""Lateral snow transport (snow_gravitational_transport__flag) on the GPU against the oracle:
snow storage and snow water passed downhill by accucapacityflux! in the reference's topological
order, snow_in / snow_out in the land balance.

Bar: 1e-10 relative per cell per step, as for the rest of the step.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from test_gpu_reservoir import RES_VARS
from wflowhip import Model, Var, synth

pytestmark = pytest.mark.gpu

SNOW_VARS = [Var.SNOW_IN, Var.SNOW_OUT, Var.SNOW_STORAGE, Var.SNOW_WATER]


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def cold(c, k):
    P, E, T = synth.forcing_numpy(c, k)
    return P, E, T - (16.0 if k < 4 else 4.0)  # snow first, then some melt


def run_snow(c, steps, extra=(), **kw):
    gpu = Model(c, snow_transport=True, **kw)
    ref = O.OracleModel(c, snow_transport=True, **kw)
    moved = 0.0
    for k in range(steps):
        P, E, T = cold(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        compare(gpu, ref, STATE_VARS + FLUX_VARS + SNOW_VARS + list(extra), f"step {k}")
        moved += float(ref.get(Var.SNOW_OUT).sum())
    gpu.close()
    ref.close()
    assert moved > 0.0


def test_snow_transport_daily():
    _say("snow transport daily 64x64")
    c = synth.make_catchment(64, cell=500.0, dt=86400.0, river_threshold=60, seed=71)
    run_snow(c, 8)


def test_snow_transport_hourly_with_reservoirs():
    _say("snow transport + reservoirs hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=72)
    synth.add_reservoirs(c, 4, seed=6)
    run_snow(c, 10, extra=RES_VARS)


def test_snow_transport_adaptive():
    _say("snow transport adaptive 64x64")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=73)
    run_snow(c, 6, adaptive=True)


def test_snow_transport_chain_chunks():
    """A catchment whose main stem gives chain chunks of the transport's lists: both snow
    families run there as lane-by-lane chains."""
    from test_gpu_edges import _land_chain_chunks
    _say("snow transport with chain chunks 512x512")
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=500, seed=36)
    probe = Model(c, snow_transport=True)
    n_chain = _land_chain_chunks(probe)  # no reservoirs: the transport's lists are the same cells
    probe.close()
    _say(f"  {n_chain} chain chunks")
    assert n_chain > 0
    run_snow(c, 4)


def test_snow_transport_4096_hourly_on_basins():
    """The full 4096^2 hourly grid (250 m, river threshold 2000 cells) with lateral snow
    transport and cold forcing (snow first, then some melt), on the GPU against the oracle run on
    complete basins of it (snow moves only down the flow graph, so a set of whole basins is an
    exact sub-problem): random small basins over 8 steps, checked every 4th step, and the largest
    basin (about 6 M cells) over 2 steps. Every state, flux and snow variable must match within
    RTOL."""
    import time
    import torch
    from test_gpu_parity import RTOL, rel_err
    from wflowhip import _abi

    t0 = time.time()
    c = synth.make_catchment(4096, cell=250.0, dt=3600.0, river_threshold=2000, seed=1)
    gpu = Model(c, snow_transport=True)
    node = gpu.node_order()
    inv = np.empty_like(node)
    inv[node] = np.arange(node.size)
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[torch.as_tensor(node, device="cuda")].contiguous()
    lab = synth.basin_ids(c.arrays["ldd"], c.nx, c.ny, c.arrays["cell_index"])
    ids, counts = np.unique(lab, return_counts=True)
    rng = np.random.default_rng(26)
    small, total = [], 0
    for k in rng.permutation(len(ids)):
        if counts[k] < 20000 and total < 150000:
            small.append(ids[k])
            total += counts[k]
    riv = c.arrays["river_location"].astype(bool)
    parts = []
    for name, basin_set, steps in (("small", np.array(small), 8), ("largest", ids[np.argmax(counts)][None], 2)):
        mask = np.isin(lab, basin_set)
        sel = torch.as_tensor(inv[np.nonzero(mask)[0]], device="cuda")
        parts.append((name, mask, sel, mask[riv], O.OracleModel(c.subset(mask), snow_transport=True), steps))
    _say(f"snow transport 4096 basins set up in {time.time() - t0:.0f}s: {[(p[0], int(p[1].sum())) for p in parts]}")
    moved = 0.0
    for k in range(8):
        P, E, T = synth.forcing_torch(c, k, elev)
        T = (T - (16.0 if k < 4 else 4.0)).contiguous()
        gpu.set_forcing(P, E, T, memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        gpu.update_model()
        for name, mask, sel, rmask, ref, steps in parts:
            if k >= steps:
                continue
            ref.step(*(f[sel].cpu().numpy() for f in (P, E, T)))
            moved += float(ref.get(Var.SNOW_OUT).sum())
            if (k + 1) % 4 != 0 and k != steps - 1:
                continue
            for v in STATE_VARS + FLUX_VARS + SNOW_VARS:
                g = gpu.get(v)
                if v in _abi.RIVER_VARS:
                    g = g[rmask]
                elif v in _abi.LAYERED_VARS:
                    g = g.reshape(-1, c.n)[:, mask].ravel()
                else:
                    g = g[mask]
                e = rel_err(g, ref.get(v))
                assert e <= RTOL, (name, k, v.name, e)
        _say(f"snow transport 4096 step {k} done ({time.time() - t0:.0f}s)")
    assert moved > 0.0
    gpu.close()
    for p in parts:
        p[4].close()

