"""wfh_step_noalloc (the step whose routing scratch lives in the library's fixed per-device array)
against wfh_step through the C ABI: two models stepped side by side on the same forcing, every
state and flux compared bitwise after each step. The larger model runs first and a smaller one
after it, so the second call finds the scratch as the first left it (counters, slots).
Also: the size limit and the unsupported modes are rejected with an error, before any launch."""
import sys

import numpy as np
import pytest

from test_gpu_parity import FLUX_VARS, STATE_VARS
from wflowhip import Model, Var, synth
from wflowhip.model import WflowError

pytestmark = pytest.mark.gpu

EXTRA = [Var.SNOW_IN, Var.SNOW_OUT]


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _bitwise(a, b):
    a, b = np.asarray(a), np.asarray(b)
    same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
    return bool(same.all())


def _side_by_side(c, steps, **kw):
    ref = Model(c, **kw)
    na = Model(c, **kw)
    for k in range(steps):
        P, E, T = synth.forcing_numpy(c, k)
        ref.run_timestep(P, E, T)
        na.set_forcing(P, E, T)
        na.update_model_noalloc()
        na.sync()
        ref.sync()
        for v in STATE_VARS + FLUX_VARS + (EXTRA if kw.get("snow_transport") else []):
            assert _bitwise(na.get(v), ref.get(v)), f"step {k}: {v.name} differs from wfh_step"
    assert na.last_substeps() == ref.last_substeps()
    ref.close()
    na.close()


def test_noalloc_matches_step_larger_then_smaller():
    _say("no-alloc: 128x128 hourly with lateral snow transport, then 48x48 daily with reservoirs")
    big = synth.make_catchment(128, cell=250.0, dt=3600.0, river_threshold=150, seed=211)
    _side_by_side(big, 4, snow_transport=True)
    small = synth.make_catchment(48, cell=1000.0, dt=86400.0, river_threshold=50, seed=212)
    synth.add_reservoirs(small, 3, seed=5)
    _side_by_side(small, 3)


def test_noalloc_rejects_what_it_does_not_cover():
    _say("no-alloc: rejections (adaptive, scratch above the limit)")
    c = synth.make_catchment(32, cell=1000.0, dt=86400.0, river_threshold=30, seed=213)
    m = Model(c, adaptive=True)
    P, E, T = synth.forcing_numpy(c, 0)
    m.set_forcing(P, E, T)
    with pytest.raises(WflowError) as e:
        m.update_model_noalloc()
    assert e.value.code == -1 and "fixed sub-steps" in str(e.value)
    m.close()
    # daily with 900 s land sub-steps: (1 + 96) slots per cell exceed 64 MiB at 300^2 cells
    big = synth.make_catchment(300, cell=1000.0, dt=86400.0, river_threshold=400, seed=214)
    m = Model(big, land_dt=900.0)
    P, E, T = synth.forcing_numpy(big, 0)
    m.set_forcing(P, E, T)
    with pytest.raises(WflowError) as e:
        m.update_model_noalloc()
    assert e.value.code == -1 and "fixed limit" in str(e.value)
    # the handle stays usable: the rejected call launched nothing
    m.update_model()
    m.sync()
    m.close()
