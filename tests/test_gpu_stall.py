"""Bounded waits of the routing kernel, forced from the host (wfh_test_route_stall): one cell's
wait for its upstream outflow is redirected to a slot that nothing fills. Its wave idles up to
the spin limit and raises the abort flag, every other wave leaves the kernel, the step ends and
wfh_sync reports WFH_ERR_DEVICE; the handle then refuses further steps. The limit is lowered so
that a stall lasts milliseconds; the time per idle poll measured from two limits, times the
default limit (2e6), is the longest a stall of the product kernel lasts before it is reported.
The allocation-free step stalls and reports the same way, and its fixed scratch is clean again
for the next call."""
import sys
import time

import numpy as np
import pytest

from wflowhip import Model, synth
from wflowhip.model import WflowError

pytestmark = pytest.mark.gpu

DEFAULT_SPIN_LIMIT = 2_000_000


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _nodes_with_upstream(m, c):
    land = m.node_order()
    down = synth.land_downstream(c)
    has_up = np.zeros(c.n, bool)
    has_up[down[down >= 0]] = True
    return np.nonzero(has_up[land])[0]


def _stalled(c, limit, pick, noalloc=False):
    """Seconds of a step whose routing stalls at the picked node; asserts how it ends."""
    m = Model(c)
    f = synth.forcing_numpy(c, 0)
    m.run_timestep(*f)
    m.sync()
    nodes = _nodes_with_upstream(m, c)
    node = int(nodes[pick(len(nodes))])
    m.test_route_stall(limit, node)
    m.set_forcing(*f)
    t0 = time.perf_counter()
    if noalloc:
        m.update_model_noalloc()
    else:
        m.update_model()
    with pytest.raises(WflowError) as e:
        m.sync()
    dt = time.perf_counter() - t0
    assert e.value.code == -5, e.value
    with pytest.raises(WflowError) as e:
        m.run_timestep(*f)
    assert e.value.code == -6
    m.close()
    return dt


def test_a_wait_that_is_never_met_ends_the_step_with_device_error():
    c = synth.make_catchment(256, cell=250.0, dt=3600.0, river_threshold=60, seed=801)
    for where, pick in (("last", lambda k: k - 1), ("middle", lambda k: k // 2), ("first", lambda k: 0)):
        dt = _stalled(c, 20_000, pick)
        _say(f"256^2, stall at the {where} cell with an upstream cell, spin limit 2e4: step ended in {dt * 1e3:.1f} ms")
        assert dt < 30.0


def test_stall_time_scales_with_the_limit_and_the_default_bound():
    c = synth.make_catchment(1024, cell=250.0, dt=3600.0, river_threshold=200, seed=802)
    # a clean step of this model, for scale
    m = Model(c)
    f = synth.forcing_numpy(c, 0)
    for _ in range(2):
        m.run_timestep(*f)
    m.sync()
    t0 = time.perf_counter()
    m.run_timestep(*f)
    m.sync()
    clean = time.perf_counter() - t0
    m.close()
    lo, hi = 20_000, 100_000
    t_lo = _stalled(c, lo, lambda k: k // 2)
    t_hi = _stalled(c, hi, lambda k: k // 2)
    per_poll = max(t_hi - t_lo, 0.0) / (hi - lo)
    bound = per_poll * DEFAULT_SPIN_LIMIT
    _say(f"1024^2: clean step {clean * 1e3:.2f} ms; stalled step {t_lo * 1e3:.1f} ms at limit {lo}, "
         f"{t_hi * 1e3:.1f} ms at limit {hi}; {per_poll * 1e9:.0f} ns per idle poll; "
         f"default limit {DEFAULT_SPIN_LIMIT}: a stall is reported after about {bound:.2f} s")
    assert t_hi > t_lo
    assert bound < 60.0


def test_noalloc_step_stalls_and_its_scratch_is_clean_again():
    c = synth.make_catchment(128, cell=250.0, dt=3600.0, river_threshold=40, seed=803)
    _stalled(c, 20_000, lambda k: k // 2, noalloc=True)
    a, b = Model(c), Model(c)
    for k in range(3):
        f = synth.forcing_numpy(c, k)
        a.set_forcing(*f)
        a.update_model_noalloc()
        b.run_timestep(*f)
    from test_gpu_parity import STATE_VARS

    for v in STATE_VARS:
        assert np.array_equal(a.get(v).view(np.uint64), b.get(v).view(np.uint64)), v.name
    a.close()
    b.close()
