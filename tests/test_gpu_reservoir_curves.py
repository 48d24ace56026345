"""Reservoirs with rating curves (outflow type 1, an H-Q table per julian day), free weirs
(type 2) and interpolated storage curves (S-H tables) on the GPU against the oracle, with the
julian day moving over the year's end (update_index_hq!, reservoir.jl:379-386) and one
reservoir that overflows its rating curve's maximum storage.

The GPU interpolates by bisection with the table extremes precomputed at create; the oracle
scans the table as the reference's interpolate_linear does. Bar: 1e-10 relative per cell and
per reservoir per step, as for the other reservoir types.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from test_gpu_reservoir import RES_VARS
from wflowhip import Model, Var, _abi, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _day(k, start):
    return (start - 1 + k) % 365 + 1


def run_curves(c, steps, start_day, **kw):
    gpu = Model(c, **kw)
    ref = O.OracleModel(c, **kw)
    compare(gpu, ref, RES_VARS[:2], "init")
    for k in range(steps):
        gpu.set_julian_day(_day(k, start_day))
        ref.set_julian_day(_day(k, start_day))
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        compare(gpu, ref, STATE_VARS + FLUX_VARS + RES_VARS, f"step {k} day {_day(k, start_day)}")
    out = ref.get(Var.RESERVOIR_OUTFLOW)
    gpu.close()
    ref.close()
    return out


def test_curves_hourly():
    _say("reservoir curves hourly 96x96, 8 reservoirs")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=61)
    synth.add_reservoirs(c, 8, seed=6, curves=True)
    out = run_curves(c, 10, start_day=362)
    assert np.all(np.isfinite(out))


def test_curves_daily():
    _say("reservoir curves daily 64x64 (96 river sub-steps)")
    c = synth.make_catchment(64, cell=1000.0, dt=86400.0, river_threshold=60, seed=62)
    synth.add_reservoirs(c, 8, seed=7, radius=2, curves=True)
    run_curves(c, 6, start_day=58)


def test_curves_adaptive():
    _say("reservoir curves adaptive 64x64")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=63)
    synth.add_reservoirs(c, 6, seed=8, curves=True)
    run_curves(c, 6, start_day=200, adaptive=True)


def test_julian_day_argument_checked():
    c = synth.make_catchment(32, cell=500.0, dt=3600.0, river_threshold=20, seed=64)
    synth.add_reservoirs(c, 4, seed=9, curves=True)
    gpu = Model(c)
    for bad in (0, 366, -1):
        with pytest.raises(RuntimeError, match="julian day"):
            gpu.set_julian_day(bad)
    gpu.close()


def test_linked_free_weirs_hourly():
    """Free weirs linked to a lower reservoir downstream, with reverse flow: the weir hands its
    reverse flow to the lower reservoir's lane within each river sub-step launch."""
    _say("linked free weirs hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 12, seed=2, curves=True, linked=True)
    assert int((c.arrays["reservoir_lower"] > 0).sum()) == 2
    out = run_curves(c, 6, start_day=100)
    assert np.all(out[c.arrays["reservoir_lower"] > 0] < 0.0)
    # an observed outflow replaces a linked weir's update: its lower reservoir gets no transfer
    up = int(np.nonzero(c.arrays["reservoir_lower"])[0][0])
    c.arrays["reservoir_outflow_obs"][up] = 0.5
    run_curves(c, 2, start_day=100)


def test_linked_free_weirs_daily_and_adaptive():
    _say("linked free weirs daily and adaptive 96x96")
    c = synth.make_catchment(96, cell=1000.0, dt=86400.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 12, seed=2, curves=True, linked=True)
    run_curves(c, 3, start_day=300)
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 12, seed=2, curves=True, linked=True)
    run_curves(c, 3, start_day=1, adaptive=True)


@pytest.mark.parametrize("linked", [False, True])
def test_curves_4096_hourly_on_basins(linked):
    """The full 4096^2 hourly grid (250 m, river threshold 2000 cells) with 64 reservoirs of
    outflow types 1-4 (rating curves, storage curves, one overflowing) on the GPU against the
    oracle run on complete basins of it (no reservoir area crosses a basin divide, which the test
    asserts, so a set of whole basins with their reservoirs is an exact sub-problem). The basins
    of fewer than 1 M cells that hold reservoirs run 8 steps over the year's end, the 2 M-cell
    basin with 13 reservoirs 2 steps. With linked=True, free weirs are linked to a lower
    reservoir downstream (two such pairs lie in the 2 M-cell basin). Every state, flux and
    reservoir variable must match within RTOL."""
    import time
    import torch
    from test_gpu_parity import RTOL, rel_err

    t0 = time.time()
    c = synth.make_catchment(4096, cell=250.0, dt=3600.0, river_threshold=2000, seed=1)
    synth.add_reservoirs(c, 64, seed=1, radius=2, curves=True, linked=linked)
    A = c.arrays
    lab = synth.basin_ids(A["ldd"], c.nx, c.ny, A["cell_index"])
    outlet, cov = A["reservoir_outlet"], A["reservoir_coverage"]
    covered = np.nonzero(cov)[0]
    assert np.array_equal(lab[covered], lab[outlet[cov[covered] - 1]])
    ids, counts = np.unique(lab, return_counts=True)
    size = dict(zip(ids.tolist(), counts.tolist()))
    held = np.unique(lab[outlet])
    small = np.array([b for b in held if size[int(b)] < 1000000])
    mid = np.array([b for b in held if 1000000 <= size[int(b)] < 2500000])
    assert small.size > 0 and mid.size == 1
    gpu = Model(c)
    node = gpu.node_order()
    inv = np.empty_like(node)
    inv[node] = np.arange(node.size)
    elev = torch.as_tensor(A["elevation"], device="cuda")[torch.as_tensor(node, device="cuda")].contiguous()
    riv = A["river_location"].astype(bool)
    parts = []
    for name, basin_set, steps in (("small", small, 8), ("mid", mid, 2)):
        mask = np.isin(lab, basin_set)
        keep = np.nonzero(mask[outlet])[0]
        sel = torch.as_tensor(inv[np.nonzero(mask)[0]], device="cuda")
        parts.append((name, mask, sel, mask[riv], keep, O.OracleModel(c.subset(mask)), steps))
    if linked:
        assert int((A["reservoir_lower"][parts[1][4]] > 0).sum()) == 2
    _say(f"reservoirs 4096 (linked={linked}) basins set up in {time.time() - t0:.0f}s: "
         f"{[(p[0], int(p[1].sum()), int(p[4].size)) for p in parts]}")
    for k in range(8):
        day = _day(k // 3, 364)
        gpu.set_julian_day(day)
        P, E, T = synth.forcing_torch(c, k, elev)
        gpu.set_forcing(P, E, T, memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        gpu.update_model()
        for name, mask, sel, rmask, keep, ref, steps in parts:
            if k >= steps:
                continue
            ref.set_julian_day(day)
            ref.step(*(f[sel].cpu().numpy() for f in (P, E, T)))
            if (k + 1) % 4 != 0 and k != steps - 1:
                continue
            for v in STATE_VARS + FLUX_VARS + RES_VARS:
                g = gpu.get(v)
                if v in _abi.RESERVOIR_VARS:
                    g = g[keep]
                elif v in _abi.RIVER_VARS:
                    g = g[rmask]
                elif v in _abi.LAYERED_VARS:
                    g = g.reshape(-1, c.n)[:, mask].ravel()
                else:
                    g = g[mask]
                e = rel_err(g, ref.get(v))
                assert e <= RTOL, (name, k, day, v.name, e)
        _say(f"reservoirs 4096 step {k} day {day} done ({time.time() - t0:.0f}s)")
    gpu.close()
    for p in parts:
        p[5].close()

