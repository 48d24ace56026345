"""Graph mode (the step's launches replayed from a HIP graph) and the routing-grid autotuner of
the single-model step: results must be bitwise those of the plain wfh_step on the same inputs,
for a batch of catchments (the batched configuration) and for a grid with reservoirs and lateral
snow transport (two dataflow launches per step)."""
import sys

import numpy as np
import pytest

from test_gpu_parity import FLUX_VARS, STATE_VARS
from wflowhip import Model, Var, _abi, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return bool(((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))).all())


def _compare(ref, other, where):
    for v in STATE_VARS + FLUX_VARS:
        assert _same(other.get(v), ref.get(v)), f"{where}: {v.name}"


def test_graph_and_autotune_on_a_batch():
    _say("graph mode + autotuner: 6-catchment batch, 12 hourly steps")
    ms = [synth.make_catchment(nx, cell=1000.0, dt=3600.0, river_threshold=40, seed=600 + k)
          for k, nx in enumerate((40, 64, 24, 96, 48, 72))]
    b = synth.make_batch(ms)
    ref = Model(b)
    gm = Model(b)
    gm.set_graph(True)
    tm = Model(b)
    tm.autotune(1)
    for k in range(12):
        f = synth.forcing_batch_numpy(ms, k)
        for m in (ref, gm, tm):
            m.run_timestep(*f)
        _compare(ref, gm, f"graph step {k}")
        _compare(ref, tm, f"autotune step {k}")
    s = tm.schedule()
    _say(f"  autotuned schedule {s}")
    assert s["tuning"] == 0 and 1 <= s["route_blocks"] <= s["max_blocks"]
    tm.set_graph(True)  # the tuned grid, replayed
    for k in range(12, 16):
        f = synth.forcing_batch_numpy(ms, k)
        for m in (ref, tm):
            m.run_timestep(*f)
        _compare(ref, tm, f"tuned graph step {k}")
    for m in (ref, gm, tm):
        m.close()


def test_graph_with_reservoirs_snow_transport_and_device_forcing():
    import torch

    _say("graph mode: 96x96 hourly with reservoirs and snow transport, node-order device forcing")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=100, seed=610)
    synth.add_reservoirs(c, 3, seed=2, curves=True)
    ref = Model(c, snow_transport=True)
    gm = Model(c, snow_transport=True)
    gm.set_graph(True)
    node = torch.as_tensor(ref.node_order(), device="cuda")
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[node].contiguous()
    for k in range(8):
        fr = synth.forcing_torch(c, k, elev)
        torch.cuda.synchronize()
        for m in (ref, gm):
            m.set_julian_day(1 + k // 4)
            m.set_forcing(*fr, memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
            m.update_model()
        ref.sync()
        gm.sync()
        _compare(ref, gm, f"step {k}")
        for v in (Var.RESERVOIR_STORAGE, Var.RESERVOIR_OUTFLOW, Var.SNOW_OUT):
            assert _same(gm.get(v), ref.get(v)), (k, v.name)
    for m in (ref, gm):
        m.close()


def test_enabling_graph_mode_unbinds_node_order_forcing():
    """Forcing bound in node order before graph mode is enabled is not what the graph reads: the
    next step without new forcing is refused, and setting the forcing again continues the run,
    equal to a run without graph mode."""
    import torch

    from wflowhip.model import WflowError

    c = synth.make_catchment(40, cell=250.0, dt=3600.0, river_threshold=30, seed=606)
    a, b = Model(c), Model(c)
    node = torch.as_tensor(a.node_order(), device="cuda")
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[node].contiguous()
    fr = [synth.forcing_torch(c, k, elev) for k in range(3)]
    NODE = _abi.WFH_MEM_DEVICE_NODE_ORDER
    for m in (a, b):
        m.set_forcing(*fr[0], memkind=NODE)
        m.update_model()
    a.set_forcing(*fr[1], memkind=NODE)
    a.set_graph(True)
    with pytest.raises(WflowError) as e:
        a.update_model()
    assert e.value.code == -6
    for k in (1, 2):
        a.set_forcing(*fr[k], memkind=NODE)
        a.update_model()
        b.set_forcing(*fr[k], memkind=NODE)
        b.update_model()
    for v in STATE_VARS:
        assert np.array_equal(a.get(v).view(np.uint64), b.get(v).view(np.uint64)), v.name
    a.close()
    b.close()
