"""Rating curves, free weirs and interpolated storage curves of reservoirs on the CPU: the
oracle's interpolate_linear, update_reservoir_hq, update_reservoir_free_weir, get_maximum_storage
and initialize_storage pinned by numpy's interp, closed forms and the special cases the
reference's code fixes; the oracle model with every outflow type checked for water-balance
closure; both libraries' checks of the tables.

The reference's own known answers for these types (Wflow/test/reservoir.jl:224-379) read the
tables reservoir_sh_1.csv, reservoir_sh_2.csv and reservoir_hq_2.csv, which are not part of the
reference tree (its test data is downloaded); those answers are therefore not used here.

Reference: Wflow/src/routing/surface/reservoir.jl:305-386 (storage curves, interpolate_linear,
update_index_hq!), :457-553 (rating curve, free weir), :583-635 (update_reservoir_model!).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from wflowhip import Var, _abi, synth
from wflowhip.model import load_library

DT = 86400.0


# ---- interpolate_linear ---------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(5))
def test_interpolate_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    for _ in range(200):
        n = int(rng.integers(2, 40))
        xp = np.sort(rng.uniform(-50.0, 400.0, n))
        fp = rng.uniform(0.0, 1e6, n) if rng.random() < 0.5 else np.sort(rng.uniform(0.0, 1e9, n))
        x = float(rng.uniform(xp[0] - 10.0, xp[-1] + 10.0))
        got = O.interpolate_linear(x, xp, fp)
        if x <= xp[0]:
            assert got == fp.min()
        elif x >= xp[-1]:
            assert got == fp.max()
        else:
            assert got == pytest.approx(float(np.interp(x, xp, fp)), rel=1e-12, abs=1e-9 * np.abs(fp).max())


def test_interpolate_knots_and_linear_tables():
    xp = np.array([0.0, 1.5, 2.0, 7.25, 10.0])
    fp = np.array([3.0, -1.0, 8.0, 2.5, 4.0])
    for i in range(1, 4):  # at a knot: exactly fp there (x - xp[i1] = 0)
        assert O.interpolate_linear(xp[i], xp, fp) == fp[i]
    # a linear table reproduces the line
    a, b = 3.7, -12.5
    for x in np.linspace(0.01, 9.99, 57):
        assert O.interpolate_linear(x, xp, a * xp + b) == pytest.approx(a * x + b, rel=1e-14, abs=1e-13)


def test_interpolate_ends_take_minimum_and_maximum():
    """Outside the table the reference returns minimum(fp) / maximum(fp), not the end values
    (a rating curve's discharge column need not be monotone)."""
    xp = np.array([1.0, 2.0, 3.0])
    fp = np.array([5.0, 1.0, 7.0])
    assert O.interpolate_linear(0.5, xp, fp) == 1.0
    assert O.interpolate_linear(1.0, xp, fp) == 1.0
    assert O.interpolate_linear(3.0, xp, fp) == 7.0
    fp2 = np.array([5.0, 9.0, 7.0])
    assert O.interpolate_linear(4.0, xp, fp2) == 9.0
    assert O.interpolate_linear(1.5, xp, fp2) == 7.0


def test_interpolate_repeated_knot_takes_last():
    """findall(xp .<= x) then last: at a repeated knot the later row wins (a step in the table)."""
    xp = np.array([0.0, 1.0, 1.0, 2.0])
    fp = np.array([0.0, 10.0, 20.0, 30.0])
    assert O.interpolate_linear(1.0, xp, fp) == 20.0
    assert O.interpolate_linear(1.5, xp, fp) == 25.0
    assert O.interpolate_linear(0.5, xp, fp) == 5.0


def test_interpolate_one_row_and_nan():
    assert O.interpolate_linear(3.0, [2.0], [9.0]) == 9.0
    assert O.interpolate_linear(1.0, [2.0], [9.0]) == 9.0
    assert math.isnan(O.interpolate_linear(float("nan"), [0.0, 1.0], [0.0, 1.0]))


# ---- storage curves ---------------------------------------------------------------------------

def test_initial_and_maximum_storage():
    area = 2e6
    h = np.array([0.0, 2.0, 5.0, 10.0])
    s = np.array([0.0, 3e6, 1.2e7, 3.5e7])
    hq_h = np.array([0.0, 4.0, 8.0])
    q = np.zeros((3, 365))
    lin = O.Reservoir.make(outflow_type=1, area=area, waterlevel=3.0, hq=(hq_h, q))
    assert O.reservoir_initial_storage(lin) == area * 3.0
    assert O.reservoir_maximum_storage(lin) == area * 8.0  # area x maximum(H of the H-Q table)
    ip = O.Reservoir.make(outflow_type=1, area=area, waterlevel=3.0, sh=(h, s), hq=(hq_h, q))
    assert O.reservoir_initial_storage(ip) == pytest.approx(3e6 + (1.2e7 - 3e6) / 3.0, rel=1e-15)
    assert O.reservoir_maximum_storage(ip) == pytest.approx(1.2e7 + (3.5e7 - 1.2e7) * 3.0 / 5.0, rel=1e-15)


def test_linear_table_equals_linear_curve():
    """An S-H table S = A H gives the water level of the linear curve S / A (up to rounding)."""
    area = 1.5e6
    h = np.linspace(0.0, 20.0, 9)
    kw = dict(outflow_type=3, area=area, threshold=1.0, rating_curve_coefficient=0.3, precipitation=1e-8,
              evaporation=2e-8, waterlevel=4.0)
    a = O.Reservoir.make(**kw)
    b = O.Reservoir.make(sh=(h, area * h), **kw)
    a.storage = O.reservoir_initial_storage(a)
    b.storage = O.reservoir_initial_storage(b)
    assert a.storage == pytest.approx(b.storage, rel=1e-15)
    for inflow in (5.0, 0.0, 40.0):
        O.update_reservoir(a, inflow, 3600.0)
        O.update_reservoir(b, inflow, 3600.0)
        assert b.outflow == pytest.approx(a.outflow, rel=1e-12)
        assert b.storage == pytest.approx(a.storage, rel=1e-12)
        assert b.waterlevel == pytest.approx(b.storage / area, rel=1e-14)
        assert b.waterlevel == pytest.approx(a.waterlevel, rel=1e-12)


# ---- rating curve (outflow type 1) ---------------------------------------------------------

def _hq_reservoir(day, **kw):
    """H-Q table with Q linear in H and a day-dependent slope: Q = (c + 1) H in column c."""
    h = np.array([0.0, 10.0])
    q = np.outer(h, np.arange(1, 366, dtype=float))
    r = O.Reservoir.make(outflow_type=1, hq=(h, q), julian_day=day, precipitation=0.0, evaporation=0.0, **kw)
    r.maximum_storage = O.reservoir_maximum_storage(r)
    return r


def test_rating_curve_column_of_the_day():
    area = 1e8  # enough stored water for the largest column (365 x 2 m3/s)
    for day in (1, 2, 59, 60, 365):
        r = _hq_reservoir(day, area=area, waterlevel=2.0, storage=2.0 * area)
        O.update_reservoir(r, 100.0, DT)
        assert r.outflow == pytest.approx(day * 2.0, rel=1e-15)  # column day - 1 holds day x H
        storage_input = 2.0 * area / DT + 100.0
        assert r.storage == pytest.approx((storage_input - day * 2.0) * DT, rel=1e-15)
        assert r.waterlevel == pytest.approx(2.0 + (r.storage - 2.0 * area) / area, rel=1e-15)


def test_rating_curve_limited_by_available_water():
    area = 1e4
    r = _hq_reservoir(200, area=area, waterlevel=5.0, storage=10.0)  # 1000 m3/s wanted, little stored
    O.update_reservoir(r, 0.5, DT)
    assert r.outflow == pytest.approx(10.0 / DT + 0.5, rel=1e-15)
    assert r.storage == pytest.approx(0.0, abs=1e-9)


def test_rating_curve_overflow():
    """Above the table's top the storage is held at the maximum storage and the excess leaves."""
    area = 1e6
    r = _hq_reservoir(1, area=area, waterlevel=9.0, storage=9.0 * area)
    maxsto = r.maximum_storage
    assert maxsto == 10.0 * area
    inflow = 500.0
    O.update_reservoir(r, inflow, DT)
    storage_input = 9.0 * area / DT + inflow
    assert r.storage == maxsto
    assert r.outflow == pytest.approx(storage_input - maxsto / DT, rel=1e-14)
    assert r.outflow > 9.0  # the rating curve's 9 m3/s plus the overflow


# ---- free weir (outflow type 2) -------------------------------------------------------------

def test_free_weir_closed_form():
    area, thr, c, e = 5e7, 2.0, 140.0, 1.5
    for wl in (2.5, 4.0, 7.0):
        r = O.Reservoir.make(outflow_type=2, area=area, threshold=thr, rating_curve_coefficient=c,
                             rating_curve_exponent=e, precipitation=0.0, evaporation=0.0, waterlevel=wl,
                             storage=wl * area)
        O.update_reservoir(r, 50.0, DT)
        dh = wl - thr
        want = min(c * dh ** e, dh * area / DT)
        assert r.outflow == pytest.approx(want, rel=1e-14)
        assert r.storage == pytest.approx((wl * area / DT + 50.0 - want) * DT, rel=1e-14)


def test_free_weir_limits():
    # below the threshold: no outflow
    r = O.Reservoir.make(outflow_type=2, area=1e6, threshold=3.0, rating_curve_coefficient=10.0,
                         rating_curve_exponent=1.5, precipitation=0.0, evaporation=0.0, waterlevel=2.9,
                         storage=2.9e6)
    O.update_reservoir(r, 1.0, DT)
    assert r.outflow == 0.0
    assert r.storage == pytest.approx(2.9e6 + DT, rel=1e-15)
    # a small reservoir: the outflow is capped at dh x area / dt
    r = O.Reservoir.make(outflow_type=2, area=1e3, threshold=1.0, rating_curve_coefficient=100.0,
                         rating_curve_exponent=1.5, precipitation=0.0, evaporation=0.0, waterlevel=3.0,
                         storage=3e3)
    O.update_reservoir(r, 0.0, DT)
    assert r.outflow == pytest.approx(2.0 * 1e3 / DT, rel=1e-15)
    # storage_input is clamped at zero before the weir flow is taken
    r = O.Reservoir.make(outflow_type=2, area=1e6, threshold=0.0, rating_curve_coefficient=0.001,
                         rating_curve_exponent=1.0, precipitation=0.0, evaporation=0.0, waterlevel=1.0,
                         storage=1e6)
    O.update_reservoir(r, -1e3, DT)
    assert r.storage == pytest.approx(-0.001 * DT, rel=1e-12)


# ---- the oracle model with every type --------------------------------------------------------

@pytest.mark.parametrize("dt,cell,adaptive", [(3600.0, 250.0, False), (86400.0, 1000.0, False),
                                              (3600.0, 250.0, True)])
def test_oracle_with_curves_closes_balance(dt, cell, adaptive):
    c = synth.make_catchment(56, cell=cell, dt=dt, river_threshold=50, seed=71)
    synth.add_reservoirs(c, 8, seed=5, curves=True)
    assert list(c.arrays["reservoir_outflow_type"]) == [1, 2, 3, 4, 1, 2, 3, 4]
    m = O.OracleModel(c, adaptive=adaptive)
    for k in range(5):
        m.set_julian_day((362 + k) % 365 + 1)
        m.step(*synth.forcing_numpy(c, k))
        for v in (Var.BALANCE_LAND_ERROR, Var.BALANCE_SSF_ERROR, Var.BALANCE_OVERLAND_ERROR, Var.BALANCE_RIVER_ERROR):
            assert float(np.abs(m.get(v)).max()) < 1e-9
        st = m.get(Var.RESERVOIR_STORAGE)
        assert np.all(np.abs(m.get(Var.BALANCE_RESERVOIR_ERROR)) <= 1e-12 * np.maximum(1.0, st / 3600.0))
        sct = c.arrays["reservoir_storage_curve_type"]
        wl = m.get(Var.RESERVOIR_WATERLEVEL)
        ptr, sh_h, sh_s = (c.arrays[k] for k in ("reservoir_sh_ptr", "reservoir_sh_h", "reservoir_sh_s"))
        for r in np.nonzero(sct == 2)[0]:  # the water level sits on the S-H table
            a, b = ptr[r], ptr[r + 1]
            assert wl[r] == pytest.approx(float(np.interp(st[r], sh_s[a:b], sh_h[a:b])), rel=1e-12)
    out = m.get(Var.RESERVOIR_OUTFLOW)
    assert np.all(np.isfinite(out))
    m.close()


def test_oracle_rating_curve_overflow_in_model():
    """Reservoir 4 starts above its H-Q table's top (5 m against 4.5 m): the first sub-step
    cuts its storage to the maximum storage (S at 4.5 m), and it stays at or below it."""
    c = synth.make_catchment(56, cell=250.0, dt=3600.0, river_threshold=50, seed=71)
    synth.add_reservoirs(c, 8, seed=5, curves=True)
    m = O.OracleModel(c)
    ptr, sh_h, sh_s = (c.arrays[k] for k in ("reservoir_sh_ptr", "reservoir_sh_h", "reservoir_sh_s"))
    a, b = ptr[4], ptr[5]
    maxsto = O.interpolate_linear(4.5, sh_h[a:b], sh_s[a:b])
    assert m.get(Var.RESERVOIR_STORAGE)[4] == O.interpolate_linear(5.0, sh_h[a:b], sh_s[a:b]) > maxsto
    for k in range(3):
        m.step(*synth.forcing_numpy(c, k))
        assert m.get(Var.RESERVOIR_STORAGE)[4] <= maxsto
        assert m.get(Var.RESERVOIR_WATERLEVEL)[4] <= 4.5
    with pytest.raises(ValueError):
        m.set_julian_day(366)
    m.close()


# ---- table checks (both libraries) -----------------------------------------------------------

def _bad_tables():
    def cat():
        c = synth.make_catchment(32, cell=500.0, dt=3600.0, river_threshold=20, seed=42)
        return synth.add_reservoirs(c, 4, seed=1, curves=True)

    c = cat()
    c.arrays["reservoir_sh_h"] = c.arrays["reservoir_sh_h"][::-1].copy()
    yield "decreasing S-H", c
    c = cat()
    c.arrays["reservoir_hq_q"][7] = np.nan
    yield "NaN in H-Q", c
    c = cat()
    p = c.arrays["reservoir_hq_ptr"]
    p[1:] = 0  # reservoir 0 (rating curve) without rows
    yield "empty H-Q", c
    c = cat()
    c.arrays.pop("reservoir_sh_ptr")
    yield "missing S-H", c
    c = cat()
    c.arrays["reservoir_outflow_type"][1] = 5
    yield "outflow type 5", c
    c = cat()
    c.arrays["reservoir_storage_curve_type"][0] = 3
    yield "storage curve type 3", c


@pytest.mark.parametrize("what,c", list(_bad_tables()), ids=lambda v: v if isinstance(v, str) else "")
def test_bad_tables_rejected(what, c):
    with pytest.raises(RuntimeError):
        O.OracleModel(c)
    lib = load_library()
    cfg, st = c.config(), c.static()
    h = ctypes.c_void_p()
    assert lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h)) == _abi.WFH_ERR_ARG, what
    assert "reservoir" in lib.wfh_last_error(None).decode()


def test_subset_carries_reservoirs_and_floodplain():
    """A basin subset (the CPU baseline's sample) keeps the reservoirs whose outlets it holds,
    with their tables, and the floodplain profiles of its river cells: basins exchange no water,
    so the subset model reproduces the full model on those basins exactly."""
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=51)
    synth.add_reservoirs(c, 8, seed=5, curves=True)
    synth.add_floodplain(c, seed=1)
    lab = synth.basin_ids(c.arrays["ldd"], c.nx, c.ny, c.arrays["cell_index"])
    ids, cnt = np.unique(lab, return_counts=True)
    mk = np.isin(lab, ids[np.argsort(-cnt)[:3]])
    sub = c.subset(mk)
    assert sub.nres == 8
    full, part = O.OracleModel(c), O.OracleModel(sub)
    riv = c.arrays["river_location"].astype(bool)
    for k in range(3):
        P, E, T = synth.forcing_numpy(c, k)
        full.step(P, E, T)
        part.step(P[mk], E[mk], T[mk])
        np.testing.assert_array_equal(full.get(Var.RIVER_Q)[mk[riv]], part.get(Var.RIVER_Q))
        np.testing.assert_array_equal(full.get(Var.FLOODPLAIN_STORAGE)[mk[riv]], part.get(Var.FLOODPLAIN_STORAGE))
        np.testing.assert_array_equal(full.get(Var.RESERVOIR_STORAGE), part.get(Var.RESERVOIR_STORAGE))
    full.close()
    part.close()


def _pair(wl_up, wl_lo, area_lo=8e6, lower_sh=None):
    up = O.Reservoir.make(outflow_type=2, area=5e7, threshold=2.0, rating_curve_coefficient=140.0,
                          rating_curve_exponent=1.5, precipitation=0.0, evaporation=0.0, waterlevel=wl_up,
                          storage=wl_up * 5e7)
    lo = O.Reservoir.make(outflow_type=3, area=area_lo, threshold=1.0, rating_curve_coefficient=0.3,
                          precipitation=0.0, evaporation=0.0, waterlevel=wl_lo, storage=wl_lo * area_lo, sh=lower_sh)
    up.lower = ctypes.addressof(lo)
    return up, lo


def test_linked_free_weir_reverse_flow():
    """The lower reservoir above the upper one: the weir flow reverses, -c (H_lo - thr)^e capped
    at (H_lo - thr) A_lo / dt, and the water leaves the lower reservoir (reservoir.jl:506-551)."""
    for area_lo in (8e6, 1e4):  # uncapped, then capped by the lower reservoir's volume above thr
        up, lo = _pair(3.0, 4.0, area_lo)
        s_lo0, cout0 = lo.storage, lo.outflow_cumulative
        O.update_reservoir(up, 10.0, DT)
        dh = 4.0 - 2.0
        want = max(-140.0 * dh ** 1.5, -dh * area_lo / DT)
        assert up.outflow == pytest.approx(want, rel=1e-14)
        assert up.storage == pytest.approx((3.0 * 5e7 / DT + 10.0 - want) * DT, rel=1e-14)
        assert lo.storage == pytest.approx(s_lo0 + want * DT, rel=1e-14)
        assert lo.waterlevel == pytest.approx(4.0 + want * DT / area_lo, rel=1e-13)
        assert lo.outflow == -up.outflow
        assert lo.outflow_cumulative == cout0 - want * DT
        # the two storages together changed by the inflow alone
        assert up.storage + lo.storage == pytest.approx(3.0 * 5e7 + s_lo0 + 10.0 * DT, rel=1e-14)


def test_linked_free_weir_forward_and_interpolated_lower():
    up, lo = _pair(4.0, 3.0)  # upper above: ordinary free weir, the lower reservoir untouched
    s_lo0 = lo.storage
    O.update_reservoir(up, 0.0, DT)
    assert up.outflow == pytest.approx(min(140.0 * 2.0 ** 1.5, 2.0 * 5e7 / DT), rel=1e-14)
    assert lo.storage == s_lo0
    h = np.linspace(0.0, 10.0, 11)
    up, lo = _pair(3.0, 4.0, 8e6, lower_sh=(h, 8e6 * h * (1.0 + 0.1 * h)))
    lo.storage = O.reservoir_initial_storage(lo)
    s0 = lo.storage
    O.update_reservoir(up, 0.0, DT)
    assert lo.storage == pytest.approx(s0 + up.outflow * DT, rel=1e-14)
    assert lo.waterlevel == pytest.approx(float(np.interp(lo.storage, 8e6 * h * (1.0 + 0.1 * h), h)), rel=1e-12)


def test_linked_reservoirs_in_model():
    """A model with linked free weirs (synthetic: each free weir linked to the next reservoir
    downstream, which starts higher): the flow reverses and every reservoir's balance closes."""
    c = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(c, 12, seed=2, curves=True, linked=True)
    lw = c.arrays["reservoir_lower"]
    assert int((lw > 0).sum()) == 2
    m = O.OracleModel(c)
    for k in range(4):
        m.step(*synth.forcing_numpy(c, k))
        st = m.get(Var.RESERVOIR_STORAGE)
        assert np.all(np.abs(m.get(Var.BALANCE_RESERVOIR_ERROR)) <= 1e-12 * np.maximum(1.0, st / 3600.0))
    assert np.all(m.get(Var.RESERVOIR_OUTFLOW)[lw > 0] < 0.0)
    m.close()
    # links must point downstream, from free weirs, one per lower reservoir
    bad = synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=60, seed=53)
    synth.add_reservoirs(bad, 12, seed=2, curves=True, linked=True)
    up = int(np.nonzero(lw)[0][0])
    bad.arrays["reservoir_lower"][:] = 0
    bad.arrays["reservoir_lower"][lw[up] - 1] = up + 1  # the lower one linked back upstream
    with pytest.raises(RuntimeError):
        O.OracleModel(bad)
    lib = load_library()
    cfg, st = bad.config(), bad.static()
    h = ctypes.c_void_p()
    assert lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h)) == _abi.WFH_ERR_ARG
