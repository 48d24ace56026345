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
from wflowhip import Model, Var, synth

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
