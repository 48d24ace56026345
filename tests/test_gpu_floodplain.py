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
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from test_gpu_reservoir import RES_VARS
from wflowhip import Model, Var, synth
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
