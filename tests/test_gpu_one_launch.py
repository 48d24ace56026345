"""Which models take one launch per step (each chunk's vertical update inside the routing kernel)
and which keep the separate vertical kernel, as the library reports it: the vertical kernel's
byte count is 0 and its event time about 0 for a one-launch model. The results of both kinds are
compared with the oracle in the parity, edge, reservoir and snow-transport tests; here a
one-launch model and a two-kernel model of the same catchment (lateral snow transport switched
on, no snow falling, so it moves nothing) must agree to the parity bar (1e-10 relative per cell
and step)."""
import sys

import numpy as np
import pytest

from test_gpu_edges import with_layers
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from wflowhip import Model, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _vertical_bytes(c, **kw):
    m = Model(c, **kw)
    try:
        return m.kernel_bytes()[0]
    finally:
        m.close()


def test_which_models_take_one_launch():
    _say("one launch: plain, reservoirs, six layers, snow transport, adaptive")
    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=30, seed=911)
    assert _vertical_bytes(c) == 0
    r = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=30, seed=912)
    synth.add_reservoirs(r, 2, seed=3)
    assert _vertical_bytes(r) == 0
    six = with_layers(synth.make_catchment(48, cell=1000.0, dt=86400.0, river_threshold=40, seed=913),
                      (50.0, 100.0, 200.0, 400.0, 600.0))
    assert _vertical_bytes(six) > 0
    assert _vertical_bytes(c, snow_transport=True) > 0
    assert _vertical_bytes(c, adaptive=True) > 0


def test_one_launch_equals_two_kernels():
    _say("one launch against two kernels: 64x64 hourly, 6 steps, warm forcing")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=40, seed=914)
    one = Model(c)
    two = Model(c, snow_transport=True)  # two kernels; with no snow on the ground nothing moves
    assert one.kernel_bytes()[0] == 0 and two.kernel_bytes()[0] > 0
    one.set_timing(True)
    for k in range(6):
        P, E, T = synth.forcing_numpy(c, k)
        T = np.maximum(T, 278.15)  # 5 degC and above: no snowfall, so no transport
        one.run_timestep(P, E, T)
        two.run_timestep(P, E, T)
        compare(one, two, STATE_VARS + FLUX_VARS, f"step {k}")
    vert_ms, route_ms = one.timing()
    assert vert_ms < 0.05 * route_ms
    one.close()
    two.close()
