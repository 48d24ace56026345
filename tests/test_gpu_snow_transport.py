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
