"""(Skipped while wfh_run is switched off.) The multi-step run (wfh_run: many steps in one launch, a chunk's next step overlapping the
previous step's long chains) against the same steps taken one at a time with wfh_step: every
state and flux, and every step's river discharge average, bitwise equal; on small catchments,
one with chain chunks, a daily model with six layers, and the full 4096^2 hourly grid. Plus a
non-finite frame and the models the run does not cover."""
import sys

import numpy as np
import pytest

from test_gpu_parity import FLUX_VARS, STATE_VARS
from wflowhip import Model, Var, synth
from wflowhip import _abi
from wflowhip.model import WflowError

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="wfh_run is switched off: its first device run (test_run_small_hourly) did "
                                      "not finish and the cause is not established")]


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _frames(c, m, k0, steps):
    import torch

    node = torch.as_tensor(m.node_order(), device="cuda")
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[node].contiguous()
    fr = [synth.forcing_torch(c, k0 + k, elev) for k in range(steps)]
    return tuple(torch.stack([f[i] for f in fr]).contiguous() for i in range(3))


def _run_against_steps(c, steps, warm=2, **kw):
    import torch

    a, b = Model(c, **kw), Model(c, **kw)
    P, E, T = _frames(c, a, 0, warm + steps)
    for m in (a, b):
        for k in range(warm):
            m.set_forcing(P[k], E[k], T[k], memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
            m.update_model()
    q = torch.full((steps, c.nriv), float("nan"), dtype=torch.float64, device="cuda")
    a.run(P[warm:].contiguous(), E[warm:].contiguous(), T[warm:].contiguous(), river_q_avg=q)
    qb = []
    for k in range(warm, warm + steps):
        b.set_forcing(P[k], E[k], T[k], memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        b.update_model()
        qb.append(b.get(Var.RIVER_Q_AVERAGE))
    a.sync()
    b.sync()
    for v in STATE_VARS + FLUX_VARS:
        x, y = a.get(v), b.get(v)
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), v.name
    qa = q.cpu().numpy()
    assert np.array_equal(qa.view(np.uint64), np.stack(qb).view(np.uint64))
    a.close()
    b.close()


def test_run_small_hourly():
    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=30, seed=901)
    _run_against_steps(c, 12)


def test_run_daily_six_layers_no_snow():
    from test_gpu_edges import with_layers

    c = with_layers(synth.make_catchment(64, cell=1000.0, dt=86400.0, river_threshold=40, seed=902, snow=False),
                    (50.0, 100.0, 200.0, 400.0, 600.0))
    _run_against_steps(c, 8)


def test_run_with_chain_chunks():
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=500, seed=36)
    _run_against_steps(c, 16)


def test_run_4096_hourly_24_steps():
    import time

    t0 = time.time()
    c = synth.make_catchment(4096, cell=250.0, dt=3600.0, river_threshold=2000, seed=1)
    _say(f"4096^2 catchment built ({time.time() - t0:.0f}s)")
    _run_against_steps(c, 24)
    _say(f"4096^2: 24-step run equal to 24 steps ({time.time() - t0:.0f}s)")


def test_run_nonfinite_frame_and_rejections():
    import torch

    c = synth.make_catchment(32, cell=250.0, dt=3600.0, river_threshold=30, seed=903)
    m = Model(c)
    P, E, T = _frames(c, m, 0, 4)
    P[2, c.n // 3] = float("nan")
    m.run(P, E, T)
    with pytest.raises(WflowError) as e:
        m.sync()
    assert e.value.code == -4
    m.close()
    r = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=30, seed=904)
    synth.add_reservoirs(r, 2, seed=1)
    m = Model(r)
    P, E, T = _frames(r, m, 0, 2)
    with pytest.raises(WflowError) as e:
        m.run(P, E, T)
    assert e.value.code == -1
    m.close()
    with pytest.raises(ValueError):
        m2 = Model(c)
        try:
            m2.run(torch.zeros((2, c.n + 1), dtype=torch.float64, device="cuda"), E[:2], T[:2])
        finally:
            m2.close()
