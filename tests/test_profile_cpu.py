"""Exponential-constant Ksat profile (KvExponentialConstant / KhExponentialConstant): below the
depth z_exp the vertical and horizontal conductivity stay at their value at z_exp
(Wflow/src/utils.jl:740-762, subsurface_process.jl:27-41, utils.jl:940-975). The SSF functions
are pinned by the reference's known answers (tests/golden/kat_reference.json, z_exp = 0.2,
routing_process.jl:22-37, 199-253); here the whole oracle step is checked against the
exponential profile where both must agree."""
import numpy as np

import oracle as O
from wflowhip import Var, synth

VARS = [Var.SATWATERDEPTH, Var.USTORELAYERDEPTH, Var.SSF_Q, Var.LAND_Q, Var.RIVER_Q, Var.TRANSFER, Var.RECHARGE,
        Var.ACTUAL_LEAKAGE, Var.SSF_QMAX, Var.BALANCE_LAND_ERROR]


def test_z_exp_at_soil_thickness_equals_exponential():
    """With z_exp at the soil thickness no depth in the column exceeds it, so every result is
    the exponential profile's, to the bit (q_constant is 0, and x + 0 = x)."""
    c = synth.make_catchment(32, cell=500.0, dt=86400.0, river_threshold=30, seed=91)
    a = O.OracleModel(c)
    c.arrays["z_exp"] = np.ascontiguousarray(c.arrays["soil_thickness"].copy())
    b = O.OracleModel(c)
    for k in range(4):
        F = synth.forcing_numpy(c, k)
        a.step(*F)
        b.step(*F)
        for v in VARS:
            assert np.array_equal(a.get(v), b.get(v), equal_nan=True), (k, v)


def test_shallow_z_exp_changes_flow_and_closes_balance():
    c = synth.make_catchment(32, cell=500.0, dt=86400.0, river_threshold=30, seed=92)
    a = O.OracleModel(c)
    c.arrays["z_exp"] = np.full(c.n, 0.3)
    b = O.OracleModel(c)
    # more conductive at depth than the exponential profile: larger q_max everywhere
    assert np.all(b.get(Var.SSF_QMAX) > a.get(Var.SSF_QMAX))
    for k in range(4):
        F = synth.forcing_numpy(c, k)
        a.step(*F)
        b.step(*F)
        assert float(np.abs(b.get(Var.BALANCE_LAND_ERROR)).max()) < 1e-9
        # the subsurface balance closes except where the water table reached the soil bottom: there
        # kinematic_wave_ssf lowers q by q_excess after net_flux was formed
        # (subsurface_process.jl:118-123), an imbalance of the reference's own algorithm
        above = b.get(Var.WATER_TABLE_DEPTH) < c.arrays["soil_thickness"]
        assert float(np.abs(b.get(Var.BALANCE_SSF_ERROR))[above].max()) < 1e-9
    assert not np.array_equal(a.get(Var.SSF_Q), b.get(Var.SSF_Q))
