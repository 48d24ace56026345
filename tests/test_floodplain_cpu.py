"""The 1-D floodplain of the kinematic-wave river on the CPU: the oracle's floodplain profile and
flood depth pinned by closed forms (a floodplain of constant width is a rectangle), the
width rule for a narrowing volume table, and the oracle model with a floodplain checked for
water-balance closure (channel plus floodplain storage, mass_balance.jl:296-313) with cells
that flood and cells that do not.

The reference's floodplain tests (Wflow/test/run_sbm.jl:920-990) run the Moselle model, whose
netCDF data is not part of the reference tree; they are therefore not used here.

Reference: Wflow/src/routing/surface/floodplain.jl:29-147, 264-341, 380-402 and
surface_kinwave.jl:387-434, 505-537, 572-662.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from wflowhip import Var, _abi, synth
from wflowhip.model import load_library


def test_profile_of_a_rectangle():
    """Constant floodplain width W (volumes W d L): width W at every depth, flow area W d,
    wetted perimeter (W - w) + 2 d, storage W d L."""
    d = np.array([0.5, 1.0, 2.0, 4.0])
    w, fl, W = 30.0, 800.0, 450.0
    d0, s, width, a, p = O.floodplain_profile(d, W * d * fl, w, fl)
    assert np.array_equal(d0, [0.0, 0.5, 1.0, 2.0, 4.0])
    assert width[0] == w
    np.testing.assert_allclose(width[1:], W, rtol=1e-14)
    np.testing.assert_allclose(a, W * d0, rtol=1e-14)
    np.testing.assert_allclose(s, W * d0 * fl, rtol=1e-14)
    np.testing.assert_allclose(p, (W - w) + 2.0 * d0, rtol=1e-14)


def test_profile_narrowing_segment_is_widened():
    """A volume table whose width would shrink keeps the previous width (floodplain.jl:89-100)."""
    d = np.array([1.0, 2.0, 3.0])
    fl, w = 100.0, 10.0
    vol = np.cumsum(np.array([200.0, 150.0, 300.0]) * 1.0 * fl)  # widths 200, 150 (-> 200), 300
    _, s, width, a, p = O.floodplain_profile(d, vol, w, fl)
    np.testing.assert_allclose(width, [10.0, 200.0, 200.0, 300.0], rtol=1e-15)
    np.testing.assert_allclose(a, [0.0, 200.0, 400.0, 700.0], rtol=1e-15)
    np.testing.assert_allclose(s, [0.0, 2e4, 4e4, 7e4], rtol=1e-15)
    # wetted perimeter: p1 = (200 - 10) + 2, then +(0 + 2), +(100 + 2); p0 = p1 - 2
    np.testing.assert_allclose(p, [190.0, 192.0, 194.0, 296.0], rtol=1e-15)


def test_flood_depth_inverts_the_storage():
    d = np.array([0.5, 1.0, 2.0, 4.0])
    fl = 800.0
    W = np.array([100.0, 300.0, 600.0, 1000.0])
    vol = np.cumsum(W * np.diff(np.concatenate([[0.0], d])) * fl)
    for h in (0.0, 0.2, 0.5, 0.75, 1.9, 3.0, 4.0):
        # storage at depth h by the rectangle segments
        seg = np.clip(h - np.concatenate([[0.0], d[:-1]]), 0.0, np.diff(np.concatenate([[0.0], d])))
        st = float(np.sum(W * seg * fl))
        assert O.flood_depth(d, vol, 20.0, fl, st) == pytest.approx(h, rel=1e-13, abs=1e-14)
    # above the table: extrapolated with the last segment's width
    top = float(vol[-1])
    assert O.flood_depth(d, vol, 20.0, fl, top + 1000.0 * 2.0 * fl) == pytest.approx(6.0, rel=1e-13)


def test_exchange_of_a_rectangle():
    """Above bankfull the excess storage F fills a rectangle of width W: flood depth F / (W L),
    the channel holds (bankfull depth + h) w L and the floodplain (W - w) h L; below bankfull
    the floodplain returns all its water to the channel (surface_kinwave.jl:585-599)."""
    d = np.array([0.5, 1.0, 2.0, 4.0])
    W, w, L, bfd, dt = 400.0, 40.0, 1000.0, 2.0, 900.0
    vol = W * d * L
    fs, fh, ex = O.floodplain_exchange(d, vol, w, L, bfd, 100000.0, 30000.0, dt)
    h = 50000.0 / (W * L)
    assert fh == pytest.approx(h, rel=1e-15)
    assert fs == pytest.approx((W - w) * h * L, rel=1e-14)
    assert ex == pytest.approx(((bfd + h) * w * L - 100000.0) / dt, rel=1e-14)
    assert fs + (bfd + h) * w * L == pytest.approx(130000.0, rel=1e-15)  # water is conserved
    # deeper than the first segment: h = 1.5 m
    F = W * 1.5 * L
    fs, fh, ex = O.floodplain_exchange(d, vol, w, L, bfd, bfd * w * L + F, 0.0, dt)
    assert fh == pytest.approx(1.5, rel=1e-14)
    assert fs == pytest.approx((W - w) * 1.5 * L, rel=1e-14)
    # below bankfull
    fs, fh, ex = O.floodplain_exchange(d, vol, w, L, bfd, 50000.0, 10000.0, dt)
    assert (fs, fh) == (0.0, 0.0)
    assert ex == 10000.0 / dt


@pytest.mark.parametrize("h", [0.125, 0.5, 1.5, 3.0])
def test_capacity_of_a_rectangle(h):
    """Manning flow of the floodplain part of a rectangle: area (W - w) h, wetted perimeter
    (W - w) + 2 h, Q = R^(2/3) S^(1/2) A / n (surface_process.jl:166-169)."""
    d = np.array([0.5, 1.0, 2.0, 4.0])
    W, w, L, n, S = 400.0, 40.0, 1000.0, 0.07, 2e-4
    vol = W * d * L
    A = (W - w) * h
    P = (W - w) + 2.0 * h
    want = (A / P) ** (2.0 / 3.0) * math.sqrt(S) * A / n
    got = O.floodplain_capacity(d, vol, w, L, h, n, S, downstream=(vol, w, L))
    assert got == pytest.approx(want, rel=1e-14)
    assert O.floodplain_capacity(d, vol, w, L, h, n, S) == got  # a pit uses its own area


def test_capacity_zero_cases():
    d = np.array([0.5, 1.0, 2.0, 4.0])
    W, w, L = 400.0, 40.0, 1000.0
    vol = W * d * L
    assert O.floodplain_capacity(d, vol, w, L, 0.0, 0.07, 1e-3) == 0.0
    # a downstream cell without floodplain beyond its channel (width w everywhere): its area is 0
    assert O.floodplain_capacity(d, vol, w, L, 0.8, 0.07, 1e-3, downstream=(w * d * L, w, L)) == 0.0


def _balance(m):
    return float(np.abs(m.get(Var.BALANCE_RIVER_ERROR)).max())


@pytest.mark.parametrize("dt,cell,adaptive,scale", [(3600.0, 250.0, False, 0.02), (86400.0, 1000.0, False, 0.1),
                                                    (3600.0, 250.0, True, 0.02)])
def test_oracle_floodplain_closes_balance(dt, cell, adaptive, scale):
    c = synth.make_catchment(56, cell=cell, dt=dt, river_threshold=50, seed=81)
    synth.add_floodplain(c, seed=2, bankfull_scale=scale)
    m = O.OracleModel(c, adaptive=adaptive)
    for k in range(5):
        m.step(*synth.forcing_numpy(c, k))
        assert _balance(m) < 1e-9
        for v in (Var.BALANCE_LAND_ERROR, Var.BALANCE_SSF_ERROR, Var.BALANCE_OVERLAND_ERROR):
            assert float(np.abs(m.get(v)).max()) < 1e-9
    fs = m.get(Var.FLOODPLAIN_STORAGE)
    assert 0 < int((fs > 0).sum()) < fs.size  # some cells flood, some do not
    assert np.all(fs >= 0.0)
    # river averages are channel plus floodplain (surface_kinwave.jl:652-659)
    np.testing.assert_array_equal(m.get(Var.RIVER_Q_AVERAGE),
                                  m.get(Var.RIVER_Q_CHANNEL_AVERAGE) + m.get(Var.FLOODPLAIN_Q_AVERAGE))
    m.close()


def test_floodplain_inputs_checked():
    c = synth.make_catchment(24, cell=1000.0, dt=86400.0, river_threshold=20, seed=83)
    synth.add_floodplain(c, seed=4)
    c.arrays["floodplain_depth"] = np.array([0.5, 0.4, 2.0, 3.5, 5.0])  # not ascending
    with pytest.raises(RuntimeError):
        O.OracleModel(c)
    lib = load_library()
    cfg, st = c.config(), c.static()
    h = ctypes.c_void_p()
    assert lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h)) == _abi.WFH_ERR_ARG
    assert "floodplain" in lib.wfh_last_error(None).decode()
