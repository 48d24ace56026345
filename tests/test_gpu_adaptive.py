"""Adaptive kinematic-wave sub-steps (kinematic_wave__adaptive_time_step_flag) on the GPU against
the oracle: every sub-step length (subsurface minimum, overland 2 % and river 5 % quantiles of the
Courant time steps) and every state and flux after each step.

Bar: as for the fixed sub-steps, 1e-10 relative per cell per step; the sub-step counts must be
equal (a different count means a different sub-step length somewhere).
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare
from wflowhip import Model, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def run_adaptive(c, steps, **kw):
    gpu = Model(c, adaptive=True, **kw)
    ref = O.OracleModel(c, adaptive=True, **kw)
    counts = []
    for k in range(steps):
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        assert gpu.last_substeps() == ref.last_substeps(), (k, gpu.last_substeps(), ref.last_substeps())
        counts.append(ref.last_substeps())
        compare(gpu, ref, STATE_VARS + FLUX_VARS, f"step {k}")
    gpu.close()
    ref.close()
    return counts


def test_adaptive_hourly():
    _say("adaptive hourly 96x96")
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=200, seed=31)
    counts = run_adaptive(c, 12)
    assert max(s[2] for s in counts) > 1 and max(s[1] for s in counts) > 1


def test_adaptive_daily():
    _say("adaptive daily 64x64")
    c = synth.make_catchment(64, cell=1000.0, dt=86400.0, river_threshold=100, seed=32)
    counts = run_adaptive(c, 8)
    assert max(s[2] for s in counts) > 10


def test_adaptive_subsurface_substeps():
    """A small alpha coefficient splits the subsurface step into several sub-steps."""
    _say("adaptive subsurface alpha 1e-3 48x48")
    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=60, seed=33)
    counts = run_adaptive(c, 6, subsurface_alpha=1e-3)
    assert max(s[0] for s in counts) > 1


def test_adaptive_no_balance_no_snow():
    _say("adaptive no snow / no balance 48x48")
    c = synth.make_catchment(48, cell=1000.0, dt=86400.0, river_threshold=60, seed=34, snow=False,
                             infiltration_reduction=False)
    run_adaptive(c, 4, water_mass_balance=False)


def test_adaptive_512_hourly():
    """A larger selection (262 144 overland keys) over a few steps."""
    _say("adaptive hourly 512x512")
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=2000, seed=35)
    counts = run_adaptive(c, 4)
    assert max(s[2] for s in counts) > 1


def test_adaptive_then_state_roundtrip():
    """Restart equals the continuous run in adaptive mode (run.jl:70-72 style)."""
    _say("adaptive restart 48x48")
    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=60, seed=36)
    a = Model(c, adaptive=True)
    for k in range(3):
        a.run_timestep(*synth.forcing_numpy(c, k))
    state = a.save_state()
    b = Model(c, adaptive=True)
    b.load_state(state)
    for k in range(3, 6):
        a.run_timestep(*synth.forcing_numpy(c, k))
        b.run_timestep(*synth.forcing_numpy(c, k))
    for v in STATE_VARS:
        assert np.array_equal(a.get_value(v), b.get_value(v), equal_nan=True), v
    a.close()
    b.close()


def _chain_chunks(gpu):
    """Chunks of river positions that the river launch runs as chains: river cells sorted by the
    library's node order; full chunks of 64 with at most 4 upstream cells per lane, where lane
    l's only upstream cell inside the chunk is lane l - 1."""
    lo, ld, ro, rd, rl, _ = gpu.topology()
    node = gpu.node_order()
    pos_of_land = np.empty_like(node)
    pos_of_land[node] = np.arange(node.size)
    rp_r = np.argsort(pos_of_land[rl], kind="stable")
    rpos = np.empty_like(rp_r)
    rpos[rp_r] = np.arange(rp_r.size)
    ups = [[] for _ in range(rp_r.size)]
    for r in range(rd.size):
        if rd[r] >= 0:
            ups[rd[r]].append(r)
    count = 0
    for c in range(rp_r.size // 64):
        ok = True
        for l in range(64):
            p = c * 64 + l
            up = [int(rpos[u]) for u in ups[rp_r[p]]]
            inside = [u for u in up if u // 64 == c]
            ok = ok and len(up) <= 4 and inside == ([p - 1] if l > 0 else [])
        count += ok
    return count


def test_adaptive_river_chain_chunks():
    """A catchment whose main stem gives chain chunks (run as a lane-by-lane chain): the same
    results as the oracle."""
    _say("adaptive with chain chunks 512x512")
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=500, seed=36)
    probe = Model(c, adaptive=True)
    n_chain = _chain_chunks(probe)
    probe.close()
    _say(f"  {n_chain} chain chunks")
    assert n_chain > 0
    run_adaptive(c, 3)
