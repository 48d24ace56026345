"""BMI get_value / set_value with device memory (WFH_MEM_DEVICE in land-index order,
WFH_MEM_DEVICE_NODE_ORDER in the library's node order) against the host-memory variant."""
import numpy as np
import pytest

from wflowhip import Model, Var, synth

pytestmark = pytest.mark.gpu

GET_VARS = [Var.USTORELAYERDEPTH, Var.SATWATERDEPTH, Var.LAND_Q, Var.RIVER_Q, Var.RIVER_H, Var.RESERVOIR_STORAGE,
            Var.RESERVOIR_OUTFLOW, Var.LAND_ALPHA, Var.FLOW_FRACTION_TO_RIVER, Var.VWC, Var.BALANCE_LAND_ERROR]


def _model():
    c = synth.make_catchment(48, cell=500.0, dt=3600.0, river_threshold=40, seed=81)
    synth.add_reservoirs(c, 3, seed=8)
    m = Model(c)
    for k in range(2):
        m.run_timestep(*synth.forcing_numpy(c, k))
    return c, m


def test_device_get_equals_host_get():
    import torch
    c, m = _model()
    node = m.node_order()
    for v in GET_VARS:
        host = m.get_value(v)
        dev = torch.empty(host.size, dtype=torch.float64, device="cuda")
        m.get_value_device(v, dev)
        assert np.array_equal(dev.cpu().numpy(), host, equal_nan=True), v
        m.get_value_device(v, dev, node_order=True)
        got = dev.cpu().numpy()
        if host.size % c.n == 0 and host.size >= c.n:  # land variable (possibly layered): node order
            layers = host.size // c.n
            want = host.reshape(layers, c.n)[:, node].ravel()
        else:
            want = host
        assert np.array_equal(got, want, equal_nan=True), v
    m.close()


def test_device_set_equals_host_set():
    import torch
    c, a = _model()
    _, b = _model()
    rng = np.random.default_rng(3)
    for v in (Var.SATWATERDEPTH, Var.USTORELAYERDEPTH, Var.RIVER_H, Var.RESERVOIR_STORAGE, Var.LAND_Q):
        base = a.get_value(v)
        new = base * rng.uniform(0.9, 1.1, base.size)
        a.set_value(v, new)
        t = torch.as_tensor(new, device="cuda")
        torch.cuda.synchronize()
        b.set_value_device(v, t)
        assert np.array_equal(a.get_value(v), b.get_value(v)), v
    assert np.array_equal(a.get_value(Var.RIVER_STORAGE), b.get_value(Var.RIVER_STORAGE))
    # node-order set then one more step: both models stay identical
    node = a.node_order()
    sat = a.get_value(Var.SATWATERDEPTH)
    a.set_value(Var.SATWATERDEPTH, sat)
    t = torch.as_tensor(sat[node], device="cuda")
    torch.cuda.synchronize()
    b.set_value_device(Var.SATWATERDEPTH, t, node_order=True)
    P, E, T = synth.forcing_numpy(c, 2)
    a.run_timestep(P, E, T)
    b.run_timestep(P, E, T)
    for v in (Var.SATWATERDEPTH, Var.RIVER_Q, Var.RESERVOIR_STORAGE):
        assert np.array_equal(a.get_value(v), b.get_value(v)), v
    a.close()
    b.close()
