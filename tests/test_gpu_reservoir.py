"""Reservoirs on the GPU against the oracle: the reservoir update inside the river sweep
(simple, modified Puls, observed outflow, abstraction), the outlet's overland and subsurface flow
going to the reservoir, and precipitation / evaporation over each reservoir's area.

Bar: 1e-10 relative per cell (and per reservoir) per step, as for the model without reservoirs.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from wflowhip import Model, Var, synth

pytestmark = pytest.mark.gpu

RES_VARS = [Var.RESERVOIR_STORAGE, Var.RESERVOIR_WATERLEVEL, Var.RESERVOIR_OUTFLOW, Var.RESERVOIR_OUTFLOW_AVERAGE,
            Var.RESERVOIR_INFLOW_AVERAGE, Var.RESERVOIR_ACTEVAP_AVERAGE, Var.RESERVOIR_ABSTRACTION_AVERAGE,
            Var.RESERVOIR_PRECIPITATION, Var.RESERVOIR_EVAPORATION, Var.BALANCE_RESERVOIR_ERROR]


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def run_res(c, steps, **kw):
    gpu = Model(c, **kw)
    ref = O.OracleModel(c, **kw)
    compare(gpu, ref, RES_VARS[:2], "init")
    for k in range(steps):
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        compare(gpu, ref, STATE_VARS + FLUX_VARS + RES_VARS, f"step {k}")
    out = ref.get(Var.RESERVOIR_OUTFLOW)
    gpu.close()
    ref.close()
    return out


def test_reservoirs_hourly():
    _say("reservoirs hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=51)
    synth.add_reservoirs(c, 6, seed=1)
    out = run_res(c, 12)
    assert np.all(np.isfinite(out))


def test_reservoirs_daily():
    _say("reservoirs daily 64x64 (96 river sub-steps)")
    c = synth.make_catchment(64, cell=1000.0, dt=86400.0, river_threshold=60, seed=52)
    synth.add_reservoirs(c, 5, seed=2, radius=2)
    run_res(c, 6)


def test_reservoirs_adaptive():
    _say("reservoirs adaptive 64x64")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 4, seed=3)
    run_res(c, 6, adaptive=True)


def test_reservoirs_no_balance():
    _say("reservoirs without balance 48x48")
    c = synth.make_catchment(48, cell=500.0, dt=3600.0, river_threshold=40, seed=54)
    synth.add_reservoirs(c, 3, seed=4)
    gpu = Model(c, water_mass_balance=False)
    ref = O.OracleModel(c, water_mass_balance=False)
    for k in range(4):
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
    compare(gpu, ref, STATE_VARS + RES_VARS[:7], "no balance")
    gpu.close()
    ref.close()


def test_reservoir_state_roundtrip():
    _say("reservoir restart 48x48")
    c = synth.make_catchment(48, cell=500.0, dt=3600.0, river_threshold=40, seed=55)
    synth.add_reservoirs(c, 3, seed=5)
    a = Model(c)
    for k in range(3):
        a.run_timestep(*synth.forcing_numpy(c, k))
    state = a.save_state()
    b = Model(c)
    b.load_state(state)
    for k in range(3, 6):
        a.run_timestep(*synth.forcing_numpy(c, k))
        b.run_timestep(*synth.forcing_numpy(c, k))
    for v in STATE_VARS + RES_VARS[:3]:
        assert np.array_equal(a.get_value(v), b.get_value(v), equal_nan=True), v
    a.close()
    b.close()
