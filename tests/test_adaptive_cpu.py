"""Adaptive kinematic-wave sub-steps on the CPU: the oracle's stable time steps and quantile
pinned by the reference's known answer, by numpy and by the textbook definition in exact
arithmetic, and the oracle's adaptive step checked for water-balance closure.

Reference: stable_timestep  Wflow/src/routing/surface/surface_kinwave.jl:674-704 (quantile! of
Statistics.jl 1.11.1, default type 7) and lateral_subsurface_flow.jl:314-344; the sub-step loops
surface_kinwave.jl:347-385, :613-662 and lateral_subsurface_flow.jl:279-304.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from wflowhip import Var, _abi, synth
from wflowhip.model import load_library


def test_stable_timestep_known_answer():
    """Wflow/test/routing_process.jl:290-383: two river cells, p = 0.05 -> 994.6119029285007
    (the reference asserts it with rtol 1.5e-8; the restatement gives the same double)."""
    q = [0.5499295110293246, 3.0005238507869465]
    alpha = [2.544585458995107, 2.5507721996678145]
    flow_length = [1059.8125, 951.96875]
    assert O.kw_stable_timestep(q, alpha, flow_length, 0.05) == 994.6119029285007


def test_stable_timestep_special_cases():
    # one cell with flow: its own Courant step; none: the 600 s default (surface_kinwave.jl:695-702)
    q, a, fl = [2.0, 0.0, 1e-31], [1.5, 1.5, 1.5], [1000.0, 900.0, 800.0]
    cel = 1.0 / (1.5 * 0.6 * 2.0 ** -0.4)
    assert O.kw_stable_timestep(q, a, fl, 0.05) == pytest.approx(1000.0 / cel, rel=1e-15)
    assert O.kw_stable_timestep([0.0, 1e-30], a[:2], fl[:2], 0.05) == 600.0


def _type7_exact(v, p):
    """Hyndman and Fan type 7 in exact rational arithmetic: h = (n - 1) p + 1 (1-based),
    x[floor h] + (h - floor h) (x[floor h + 1] - x[floor h])."""
    x = sorted(Fraction(t) for t in v)
    n = len(x)
    h = (n - 1) * Fraction(p) + 1
    j = min(max(int(h), 1), n - 1)
    g = h - j
    return float(x[j - 1] + g * (x[j] - x[j - 1]))


@pytest.mark.parametrize("seed", range(6))
def test_quantile_matches_definition(seed):
    rng = np.random.default_rng(seed)
    for _ in range(150):
        n = int(rng.integers(2, 60))
        v = rng.lognormal(4.0, 2.0, n)
        if rng.random() < 0.3:  # ties
            v[rng.integers(0, n, n // 2)] = v[0]
        p = float(rng.choice([0.02, 0.05, 0.5, 0.95, 0.0, 1.0, rng.random()]))
        got = O.quantile(v, p)
        exact = _type7_exact(v, p)
        # the reference forms the fractional index in floating point (aleph = fma(n, p, 1 - p),
        # gamma = aleph - j), so gamma carries up to half an ulp of aleph <= n
        tol = 4.0 * 2.0 ** -52 * n * float(v.max() - v.min())
        assert got == pytest.approx(exact, rel=4e-15, abs=tol), (n, p)
        assert got == pytest.approx(float(np.quantile(v, p)), rel=4e-15, abs=tol)
        assert v.min() <= got <= v.max()


def test_quantile_ends_and_order():
    v = np.array([5.0, 1.0, 3.0, 2.0, 4.0])
    assert O.quantile(v, 0.0) == 1.0
    assert O.quantile(v, 1.0) == 5.0
    assert O.quantile(v, 0.5) == 3.0
    assert O.quantile(v, 0.25) == 2.0
    # two values: a + gamma (b - a) with gamma = fma(2, p, 1 - p) - 1
    assert O.quantile([10.0, 20.0], 0.05) == pytest.approx(10.5, rel=1e-15)
    # nearly equal values take the a + gamma (b - a) branch and stay within [a, b]
    a = 1.0
    b = np.nextafter(np.nextafter(a, 2.0), 2.0)
    assert a <= O.quantile([b, a], 0.3) <= b


def _balance_max(m):
    return max(float(np.abs(m.get(v)).max()) for v in (Var.BALANCE_LAND_ERROR, Var.BALANCE_SSF_ERROR,
                                                       Var.BALANCE_OVERLAND_ERROR, Var.BALANCE_RIVER_ERROR))


@pytest.mark.parametrize("dt,cell,alpha", [(3600.0, 250.0, 1.0), (86400.0, 1000.0, 1.0), (3600.0, 250.0, 1e-3)])
def test_oracle_adaptive_closes_balance(dt, cell, alpha):
    """Adaptive sub-steps of unequal length keep every balance closed (run_sbm.jl:1238-1248 asserts
    |error| < 1e-9); the river takes several sub-steps, and a small subsurface alpha coefficient
    splits the subsurface step."""
    c = synth.make_catchment(48, cell=cell, dt=dt, river_threshold=60, seed=21)
    m = O.OracleModel(c, adaptive=True, subsurface_alpha=alpha)
    subs = []
    for k in range(5):
        m.step(*synth.forcing_numpy(c, k))
        subs.append(m.last_substeps())
        assert _balance_max(m) < 1e-9
    assert max(s[2] for s in subs) > 1
    if alpha < 1.0:
        assert max(s[0] for s in subs) > 1
    m.close()


def test_oracle_fixed_substep_counts():
    c = synth.make_catchment(24, cell=1000.0, dt=86400.0, river_threshold=20, seed=22)
    m = O.OracleModel(c)
    m.step(*synth.forcing_numpy(c, 0))
    assert m.last_substeps() == (1, 24, 96)
    m.close()


def test_create_rejects_bad_adaptive_coefficient():
    c = synth.make_catchment(16, cell=1000.0, dt=86400.0, river_threshold=10, seed=2)
    lib = load_library()
    for alpha in (0.0, -1.0, float("nan")):
        cfg = c.config(adaptive=True, subsurface_alpha=alpha)
        st = c.static()
        h = ctypes.c_void_p()
        assert lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h)) == _abi.WFH_ERR_ARG
        assert "alpha" in lib.wfh_last_error(None).decode()
