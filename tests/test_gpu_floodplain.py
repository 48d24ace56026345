"""The 1-D floodplain of the kinematic-wave river (floodplain_1d__flag) on the GPU against the
oracle: channel-floodplain exchange, the floodplain's Manning capacity and accumulation down the
river graph, flux_in, the combined river averages and the river balance with floodplain storage;
fixed and adaptive sub-steps, with reservoirs, and restart.

On the GPU a floodplain model runs one launch per river sub-step (exchange, river sub-step,
floodplain accumulation, inflow). Bar: 1e-10 relative per cell per step, as for the river.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, RTOL, STATE_VARS, compare, rel_err
from test_gpu_reservoir import RES_VARS
from wflowhip import Model, Var, synth
from wflowhip import _abi
from wflowhip._abi import FLOODPLAIN_VARS

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def run_fp(c, steps, res=False, **kw):
    gpu = Model(c, **kw)
    ref = O.OracleModel(c, **kw)
    extra = RES_VARS if res else []
    for k in range(steps):
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        assert gpu.last_substeps() == ref.last_substeps(), k
        compare(gpu, ref, STATE_VARS + FLUX_VARS + FLOODPLAIN_VARS + extra, f"step {k}")
    fs = ref.get(Var.FLOODPLAIN_STORAGE)
    gpu.close()
    ref.close()
    return fs


def test_floodplain_hourly():
    _say("floodplain hourly 96x96 (4 river sub-steps)")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=91)
    synth.add_floodplain(c, seed=1)
    fs = run_fp(c, 8)
    assert 0 < int((fs > 0).sum()) < fs.size


def test_floodplain_daily():
    _say("floodplain daily 64x64 (96 river sub-steps)")
    c = synth.make_catchment(64, cell=1000.0, dt=86400.0, river_threshold=60, seed=92)
    synth.add_floodplain(c, seed=2, bankfull_scale=0.1)
    fs = run_fp(c, 5)
    assert int((fs > 0).sum()) > 0


def test_floodplain_adaptive():
    _say("floodplain adaptive 64x64")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=93)
    synth.add_floodplain(c, seed=3)
    run_fp(c, 5, adaptive=True)


def test_floodplain_with_reservoirs():
    _say("floodplain + reservoirs hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=94)
    synth.add_reservoirs(c, 6, seed=4, curves=True)
    synth.add_floodplain(c, seed=4)
    run_fp(c, 6, res=True)


def test_floodplain_state_roundtrip():
    _say("floodplain restart 48x48")
    c = synth.make_catchment(48, cell=500.0, dt=3600.0, river_threshold=40, seed=95)
    synth.add_floodplain(c, seed=5)
    a = Model(c)
    for k in range(3):
        a.run_timestep(*synth.forcing_numpy(c, k))
    state = a.save_state()
    b = Model(c)
    b.load_state(state)
    for k in range(3, 6):
        a.run_timestep(*synth.forcing_numpy(c, k))
        b.run_timestep(*synth.forcing_numpy(c, k))
    for v in STATE_VARS + [Var.FLOODPLAIN_STORAGE]:
        assert np.array_equal(a.get_value(v), b.get_value(v), equal_nan=True), v
    a.close()
    b.close()


def test_floodplain_with_linked_free_weirs_and_snow_transport():
    _say("floodplain + linked free weirs + snow transport hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 12, seed=2, curves=True, linked=True)
    synth.add_floodplain(c, seed=6)
    assert int((c.arrays["reservoir_lower"] > 0).sum()) == 2
    run_fp(c, 4, res=True, snow_transport=True)


def test_floodplain_chain_chunks():
    """A catchment whose main stem gives chain chunks: river and floodplain families run as
    lane-by-lane chains there."""
    _say("floodplain with chain chunks 512x512")
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=500, seed=36)
    synth.add_floodplain(c, seed=7)
    fs = run_fp(c, 3)
    assert int((fs > 0).sum()) > 0


def test_floodplain_4096_hourly_on_basins():
    """The full 4096^2 hourly grid (250 m, river threshold 2000 cells) with a floodplain on every
    river cell, on the GPU against the oracle run on complete basins of it (basins exchange no
    water, so a set of whole basins is an exact sub-problem): random small basins over 12 steps,
    checked every 4th step, and the largest basin (about 6 M cells, holding the longest flow
    path) over 4 steps. Every state, flux and floodplain variable of every oracle cell must match
    within RTOL, and the largest basin's floodplain must hold water at its checked steps."""
    import time
    import torch

    t0 = time.time()
    c = synth.make_catchment(4096, cell=250.0, dt=3600.0, river_threshold=2000, seed=1)
    synth.add_floodplain(c, seed=8)
    gpu = Model(c)
    node = gpu.node_order()
    inv = np.empty_like(node)
    inv[node] = np.arange(node.size)
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[torch.as_tensor(node, device="cuda")].contiguous()
    lab = synth.basin_ids(c.arrays["ldd"], c.nx, c.ny, c.arrays["cell_index"])
    ids, counts = np.unique(lab, return_counts=True)
    rng = np.random.default_rng(25)
    small, total = [], 0
    for k in rng.permutation(len(ids)):
        if counts[k] < 20000 and total < 150000:
            small.append(ids[k])
            total += counts[k]
    riv = c.arrays["river_location"].astype(bool)
    parts = []
    for name, basin_set, steps in (("small", np.array(small), 12), ("largest", ids[np.argmax(counts)][None], 4)):
        mask = np.isin(lab, basin_set)
        sel = torch.as_tensor(inv[np.nonzero(mask)[0]], device="cuda")
        parts.append((name, mask, sel, mask[riv], O.OracleModel(c.subset(mask)), steps))
    _say(f"floodplain 4096 basins set up in {time.time() - t0:.0f}s: {[(p[0], int(p[1].sum())) for p in parts]}")
    flooded = {}
    for k in range(12):
        P, E, T = synth.forcing_torch(c, k, elev)
        gpu.set_forcing(P, E, T, memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        gpu.update_model()
        for name, mask, sel, rmask, ref, steps in parts:
            if k >= steps:
                continue
            ref.step(*(f[sel].cpu().numpy() for f in (P, E, T)))
            if (k + 1) % 4 != 0 and k != steps - 1:
                continue
            for v in STATE_VARS + FLUX_VARS + FLOODPLAIN_VARS:
                g = gpu.get(v)
                if v in _abi.RIVER_VARS:
                    g = g[rmask]
                elif v in _abi.LAYERED_VARS:
                    g = g.reshape(-1, c.n)[:, mask].ravel()
                else:
                    g = g[mask]
                e = rel_err(g, ref.get(v))
                assert e <= RTOL, (name, k, v.name, e)
            flooded[name] = int((ref.get(Var.FLOODPLAIN_STORAGE) > 0).sum())
            if name == "largest":
                assert flooded[name] > 0, (k, "the largest basin's floodplain holds no water")
        _say(f"floodplain 4096 step {k} done ({time.time() - t0:.0f}s)")
    _say(f"flooded cells at the last checked step: {flooded}")
    assert flooded["largest"] > 0
    gpu.close()
    for p in parts:
        p[4].close()

