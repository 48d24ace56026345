"""Reservoirs on the CPU: the oracle's reservoir update pinned by the reference's reservoir unit
tests (tests/golden/kat_reservoir.json), closed forms of the simple and modified Puls updates,
and the oracle model with reservoirs checked for water-balance closure.

Reference: Wflow/src/routing/surface/reservoir.jl, surface_kinwave.jl:441-489 (coupling into the
river sweep), network.jl:296-360 and domain.jl:95-121 (reservoir outlets removed from the
upstream lists), io.jl:125-143 (forcing over the reservoir area).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from wflowhip import Var, synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "kat_reservoir.json")))
RTOL = GOLD["rtol"]


def approx(a, b):
    return a == b or abs(a - b) <= RTOL * max(abs(a), abs(b))


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["ref"].split(" ")[0] + c["call"])
def test_reservoir_known_answers(case):
    p = dict(case["res"])
    from_wl = p.pop("storage_from_waterlevel", False)
    r = O.Reservoir.make(**p)
    if from_wl:  # initialize_storage, linear profile  reservoir.jl:343-362
        r.storage = r.area * r.waterlevel
    e = case["expect"]
    dt = case["dt"]
    if case["call"] == "update":
        O.update_reservoir(r, case["inflow"], dt)
    elif case["call"] == "at_river":
        qin = O.update_reservoir_at_river(r, case["q_v"], dt)
        assert approx(qin, e["downstream_qin"])
    else:  # two-node river: node 1's kinematic wave, then its reservoir
        g = json.load(open(os.path.join(HERE, "golden", "kat_reference.json")))["kinwave_river_update_2node"]
        q1, _ = O.kinematic_wave(0.0, g["q_prev"][0], 0.0, g["alpha"][0], dt, g["flow_length"][0])
        out = O.update_reservoir_at_river(r, q1, dt)
        assert approx(out, g["reservoir_outflow"])
    for k, v in e.items():
        if k == "downstream_qin":
            continue
        got = r.storage / r.area if k == "waterlevel_of_storage" else getattr(r, k)
        assert approx(got, v), (k, got, v)


def test_simple_reservoir_closed_form():
    """Below the full target and with storage left after the demand, the simple reservoir
    releases exactly scurve(fill) x demand and keeps the rest (reservoir.jl:389-421)."""
    r = O.Reservoir.make(outflow_type=4, area=1e6, demand=2.0, maximum_release=10.0, maximum_storage=1e7,
                         target_full_fraction=0.9, target_minimum_fraction=0.3, precipitation=0.0, evaporation=0.0,
                         storage=4e6, waterlevel=4.0)
    dt = 3600.0
    O.update_reservoir(r, 5.0, dt)
    s0 = 4e6 + 5.0 * dt
    fac = 1.0 / (1.0 + math.exp(-30.0 * (s0 / 1e7 - 0.3)))
    assert r.outflow == pytest.approx(fac * 2.0, rel=1e-14)
    assert r.storage == pytest.approx(s0 - fac * 2.0 * dt, rel=1e-14)
    assert r.waterlevel == pytest.approx(4.0 + (r.storage - 4e6) / 1e6, rel=1e-14)


def test_modified_puls_closed_form():
    """Modified Puls: with Q = c (H - H0)^2 and S = A H, the implicit step solves
    S/dt + I = Q + S_new/dt, so outflow = c (S_new/A - H0)^2 (reservoir.jl:427-456)."""
    area, c, h0, dt = 2e6, 0.5, 1.0, 3600.0
    r = O.Reservoir.make(outflow_type=3, area=area, threshold=h0, rating_curve_coefficient=c, precipitation=0.0,
                         evaporation=0.0, storage=area * 3.0, waterlevel=3.0)
    O.update_reservoir(r, 20.0, dt)
    h = r.storage / area
    assert r.outflow == pytest.approx(c * (h - h0) ** 2, rel=1e-9)
    assert r.storage / dt + r.outflow == pytest.approx(area * 3.0 / dt + 20.0, rel=1e-13)


def _balance_max(m, vars_):
    return max(float(np.abs(m.get(v)).max()) for v in vars_)


@pytest.mark.parametrize("adaptive", [False, True])
def test_oracle_with_reservoirs_closes_balance(adaptive):
    c = synth.make_catchment(48, cell=500.0, dt=3600.0, river_threshold=40, seed=41)
    synth.add_reservoirs(c, 4, seed=3)
    assert c.nres == 4
    m = O.OracleModel(c, adaptive=adaptive)
    for k in range(6):
        m.step(*synth.forcing_numpy(c, k))
        assert _balance_max(m, (Var.BALANCE_LAND_ERROR, Var.BALANCE_SSF_ERROR, Var.BALANCE_OVERLAND_ERROR,
                                Var.BALANCE_RIVER_ERROR)) < 1e-9
        st = m.get(Var.RESERVOIR_STORAGE)
        # reservoir balance relative to its throughput (storage up to 1e7 m3)
        assert np.all(np.abs(m.get(Var.BALANCE_RESERVOIR_ERROR)) <= 1e-12 * np.maximum(1.0, st / 3600.0))
    out = m.get(Var.RESERVOIR_OUTFLOW)
    assert np.all(np.isfinite(out)) and np.all(out >= 0.0)
    # the reservoir area gets no precipitation in the land model, the reservoir gets its mean
    P, _, _ = synth.forcing_numpy(c, 5)
    cov = c.arrays["reservoir_coverage"]
    for k in range(c.nres):
        assert m.get(Var.RESERVOIR_PRECIPITATION)[k] == pytest.approx(P[cov == k + 1].mean(), rel=1e-12)
    m.close()


def test_reservoir_outlet_is_removed_from_land_inflow():
    """The outlet's overland and subsurface flow go to the reservoir, not to the land cell
    downstream (filter_upstream_nodes on the land graph, domain.jl:117-121)."""
    c = synth.make_catchment(32, cell=500.0, dt=3600.0, river_threshold=20, seed=42)
    synth.add_reservoirs(c, 2, seed=1)
    m = O.OracleModel(c)
    for k in range(3):
        m.step(*synth.forcing_numpy(c, k))
    outlets = c.arrays["reservoir_outlet"]
    np.testing.assert_array_equal(m.get(Var.RESERVOIR_INFLOW_AVERAGE) >= 0.0, True)
    lo, ld, ro, rd, rl = m.topology()
    # the downstream land cell's overland inflow sums only its other upstream cells
    ol_q = m.get(Var.LAND_Q)
    qin = m.get(Var.LAND_QIN)
    fftr = m.get(Var.FLOW_FRACTION_TO_RIVER)
    for o in outlets:
        d = ld[o]
        ups = [u for u in np.nonzero(ld == d)[0] if u not in set(outlets)]
        want = 0.0
        for u in sorted(ups):
            want = want + ol_q[u] * (1.0 - fftr[u])
        if m.get(Var.SURFACE_FLOW_WIDTH)[d] > 0.0:
            assert qin[d] == want
    m.close()
