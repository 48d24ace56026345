"""Lateral snow transport (snow_gravitational_transport__flag) on the CPU: accucapacityflux pinned
by the reference's known answers (PCRaster accucapacity examples, Wflow/test/routing_process.jl:
255-288), and the oracle's model with transport checked for snow-mass conservation and balance.

Reference: lateral_snow_transport!  Wflow/src/routing/surface/surface_process.jl:9-19,
accucapacityflux!  routing/utils.jl:82-108, flux_in!  routing/utils.jl:162-167.
"""
import numpy as np
import pytest

import oracle as O
from wflowhip import Var, synth

# graph of the reference test: 1 -> 4, 2 -> 5, 3 -> 5, 4 -> 6, 5 -> 6 (0-based below)
DOWN = [3, 4, 4, 5, 5, -1]
ORDER = [0, 1, 2, 3, 4, 5]


def test_accucapacityflux_example1():
    dt = 86400.0
    material = dt * np.array([0.5, 2.0, 2.0, 0.5, 2.0, 0.5])
    flux, new = O.accucapacityflux(ORDER, DOWN, material, np.full(6, 1.5), dt)
    assert list(new) == list(dt * np.array([0.0, 0.5, 0.5, 0.0, 3.5, 1.5]))
    assert list(flux) == [0.5, 1.5, 1.5, 1.0, 1.5, 1.5]


def test_accucapacityflux_example2():
    dt = 86400.0
    material = dt * np.full(6, 10.0)
    flux, new = O.accucapacityflux(ORDER, DOWN, material, np.array([2.0, 30, 30, 2, 30, 2]), dt)
    assert list(new) == list(dt * np.array([8.0, 0.0, 0.0, 10.0, 0.0, 40.0]))
    assert list(flux) == [2.0, 10.0, 10.0, 2.0, 30.0, 2.0]


def test_oracle_snow_transport_conserves_snow():
    """Transport moves snow between cells: with forcing that only freezes (no melt, no rain) the
    catchment's snow total changes only by what leaves through the pits; every cell's balance
    closes with snow_in / snow_out as boundary fluxes (mass_balance.jl:215-253)."""
    c = synth.make_catchment(40, cell=250.0, dt=86400.0, river_threshold=30, seed=61)
    m = O.OracleModel(c, snow_transport=True)
    moved = 0.0
    for k in range(6):
        P, E, T = synth.forcing_numpy(c, k)
        T = np.full_like(T, 263.15)  # -10 C: precipitation falls as snow, nothing melts
        m.step(P, E, T)
        assert float(np.abs(m.get(Var.BALANCE_LAND_ERROR)).max()) < 1e-9
        sin, sout = m.get(Var.SNOW_IN), m.get(Var.SNOW_OUT)
        assert np.all(sout >= 0.0) and np.all(sin >= 0.0)
        moved += float(sout.sum())
    assert moved > 0.0
    m.close()


def test_oracle_without_transport_has_no_snow_fluxes():
    c = synth.make_catchment(24, cell=250.0, dt=86400.0, river_threshold=20, seed=62)
    m = O.OracleModel(c)
    P, E, T = synth.forcing_numpy(c, 0)
    m.step(P, E, np.full_like(T, 263.15))
    assert not m.get(Var.SNOW_OUT).any() and not m.get(Var.SNOW_IN).any()
    m.close()
