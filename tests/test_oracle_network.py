"""Pins of the oracle where the reference holds no known answer: the network sweeps of the step
(kinwave_subsurface_update!, kinwave_land_update!, kinwave_river_update!) by network
conservation, and single soil and Newton functions by closed forms.

Network conservation. Each cell's upstream fold adds the discharge its upstream cells computed
in the same sub-step, weighted by the flow fraction to the river: summed over the network the
inflows equal the discharge of every cell that has a downstream cell, sub-step by sub-step, so
over a step

    sum_v qin_avg(v) + sum_v to_river_avg(v) = sum_{u with downstream} q_avg(u)   (subsurface)
    sum_v qin_avg(v) + sum_v to_river_avg(v) = sum_u q_avg(u) w_u                   (overland)
    sum_r qin_avg(r)                           = sum_{r with downstream} q_avg(r)   (river)

and the storage of the overland and river networks changes by the lateral inflow less what
leaves the network. These sums tie the fold to the topology and the sub-step: a fold of the
discharge of the sub-step before, of the wrong cell or with a dropped weight breaks them, while
each cell's own balance still closes (it uses the inflow the cell was given). The tests plant
the stale fold in the oracle (wfo_set_mutation) and show that the checks catch it.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from wflowhip import Var, synth

REL = 1e-12


def _fsum(x):
    return math.fsum(np.asarray(x, dtype=np.float64).tolist())


def _rel(a, b, scale):
    # the scale has a floor of 1 (m3 or m3 s-1): the KIN_WAVE_MIN_FLOW floor of the Newton
    # solve (surface_process.jl:61) adds about alpha 1e-18 dx of storage per cell and solve,
    # which the relative error of a near-empty network (the first hours) would magnify
    return abs(a - b) / max(scale, 1.0)


class Net:
    """The oracle model with the identities of its last step."""

    def __init__(self, c, mutation=0, **kw):
        self.c = c
        self.m = O.OracleModel(c, **kw)
        if mutation:
            self.m.set_mutation(mutation)
        lo, ld, ro, rd, rl = self.m.topology()
        self.down = ld
        self.rdown = rd
        self.f = self.m.get(Var.FLOW_FRACTION_TO_RIVER)
        self.sfw = self.m.get(Var.SURFACE_FLOW_WIDTH)
        has = ld >= 0
        self.has = has
        # overland: the share of u's discharge that enters its downstream cell's overland store
        # or the river (kinwave_land_update! adds the (1 - f) share only where the downstream
        # surface flow width is positive)
        wq = np.zeros(c.n)
        wq[has] = (1.0 - self.f[has]) * (self.sfw[ld[has]] > 0.0)
        self.w_qin = wq
        self.dt = c.dt

    def step(self, k):
        m = self.m
        ol0 = m.get(Var.LAND_STORAGE)
        r0 = m.get(Var.RIVER_STORAGE)
        P, E, T = synth.forcing_numpy(self.c, k)
        m.step(P, E, T)
        out = {}
        # subsurface
        qin, tor, q = m.get(Var.SSF_QIN_AVERAGE), m.get(Var.SSF_TO_RIVER_AVERAGE), m.get(Var.SSF_Q_AVERAGE)
        lhs, rhs = _fsum(qin) + _fsum(tor), _fsum(q[self.has])
        out["ssf"] = _rel(lhs, rhs, _fsum(np.abs(q)))
        # overland fold
        qin, tor, q = m.get(Var.LAND_QIN_AVERAGE), m.get(Var.LAND_TO_RIVER_AVERAGE), m.get(Var.LAND_Q_AVERAGE)
        lhs = _fsum(qin) + _fsum(tor)
        rhs = _fsum(q[self.has] * self.w_qin[self.has]) + _fsum(q[self.has] * self.f[self.has])
        out["land"] = _rel(lhs, rhs, _fsum(np.abs(q)))
        # overland storage: lateral inflow in, everything not entering a downstream overland store out
        ol1, inw = m.get(Var.LAND_STORAGE), m.get(Var.LAND_INWATER)
        leaves = _fsum(q) - _fsum(q * self.w_qin)
        d_st = _fsum(ol1) - _fsum(ol0)
        out["land_storage"] = _rel(d_st, self.dt * (_fsum(inw) - leaves),
                                   _fsum(np.abs(ol1)) + self.dt * (_fsum(np.abs(inw)) + _fsum(np.abs(q))))
        # river fold and storage
        if self.c.nriv:
            qin, q = m.get(Var.RIVER_QIN_AVERAGE), m.get(Var.RIVER_Q_AVERAGE)
            rhas = self.rdown >= 0
            out["river"] = _rel(_fsum(qin), _fsum(q[rhas]), _fsum(np.abs(q)))
            r1, inw = m.get(Var.RIVER_STORAGE), m.get(Var.RIVER_INWATER)
            d_st = _fsum(r1) - _fsum(r0)
            out["river_storage"] = _rel(d_st, self.dt * (_fsum(inw) - _fsum(q[~rhas])),
                                        _fsum(np.abs(r1)) + self.dt * (_fsum(np.abs(inw)) + _fsum(np.abs(q))))
        return out

    def close(self):
        self.m.close()


# storage sums carry the Newton stopping residual (|f| <= 1e-12 per solve, larger where the
# rounding noise of f exceeds it), so they are held to a looser bound than the folds
BOUND = {"ssf": REL, "land": REL, "river": REL, "land_storage": 1e-9, "river_storage": 1e-9}


@pytest.mark.parametrize("case", ["hourly_one_substep", "hourly", "daily"])
def test_sweeps_conserve_water_over_the_network(case):
    if case == "hourly_one_substep":  # one sub-step per family: the step's sums are per sub-step
        c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=301)
        kw = dict(river_dt=3600.0)
        steps = 8
    elif case == "hourly":  # 4 river sub-steps
        c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=302)
        kw = {}
        steps = 8
    else:  # 24 overland and 96 river sub-steps
        c = synth.make_catchment(48, cell=1000.0, dt=86400.0, river_threshold=40, seed=303)
        kw = {}
        steps = 6
    net = Net(c, **kw)
    worst = {}
    for k in range(steps):
        for name, e in net.step(k).items():
            worst[name] = max(worst.get(name, 0.0), e)
            assert e <= BOUND[name], f"step {k}: {name} off by {e:.3e}"
    net.close()
    assert set(worst) == set(BOUND)


@pytest.mark.parametrize("mutation,family", [(1, "ssf"), (2, "land"), (3, "river")])
def test_stale_fold_is_caught(mutation, family):
    """The same model with one family folding the discharge of the sub-step before: its
    network sum breaks (the other families' sums, whose inputs the mutation also changes, may
    still hold); every cell's own balance still closes, so the per-cell checks alone would
    miss it."""
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=304)
    net = Net(c, mutation=mutation)
    worst = 0.0
    for k in range(8):
        e = net.step(k)
        worst = max(worst, e[family])
        for v in (Var.BALANCE_SSF_ERROR, Var.BALANCE_OVERLAND_ERROR, Var.BALANCE_RIVER_ERROR):
            b = net.m.get(v)
            assert np.all(np.abs(b) < 1e-9), f"{v.name} should close per cell even with the mutant"
    net.close()
    assert worst > 1e-6, f"the stale {family} fold went unnoticed (worst {worst:.3e})"


def test_chain_reaches_the_steady_state_closed_form():
    """A chain of cells with the same lateral inflow: swept in topological order with the
    oracle's kinematic_wave, the steady state is q_k = (k + 1) q_lat dx (storage no longer
    changes, so each cell passes on its inflow plus its lateral inflow)."""
    L, qlat, alpha, dx, dt = 40, 2e-4, 3.0, 250.0, 900.0
    q = np.zeros(L)
    for it in range(4000):
        prev = q.copy()
        qin = 0.0
        for k in range(L):
            q[k], _ = O.kinematic_wave(qin, q[k], qlat, alpha, dt, dx)
            qin = q[k]
        if np.max(np.abs(q - prev)) == 0.0:
            break
    want = (np.arange(L) + 1) * qlat * dx
    assert np.max(np.abs(q - want) / want) < 1e-9, np.max(np.abs(q - want) / want)


def _soil(nl, thick, theta_s=0.45, theta_r=0.05, theta_fc=0.25, **kw):
    s = O.SoilCell()
    cum = [0.0]
    for t in thick:
        cum.append(cum[-1] + t)
    s.set(N=nl, nl=nl, n_unsat=nl, theta_s=theta_s, theta_r=theta_r, theta_fc=theta_fc,
          soil_thickness=cum[-1], act_thick=list(thick), cum_depth=cum, ustorethick=list(thick),
          kvfrac=[1.0] * nl, **kw)
    return s


def test_actual_infiltration_single_layer_closed_form():
    """One unsaturated layer holding x above its capacity t (theta_s - theta_r): the excess x
    leaves, actual infiltration = infiltration - x / dt; below capacity nothing changes."""
    dt = 3600.0
    for over in (0.0, 0.0125, 0.25):
        t = 0.5
        cap = t * (0.45 - 0.05)
        s = _soil(1, [t], infiltration=3e-6, ustore=[cap + over])
        O.soil_call("soil_actual_infiltration", s, dt)
        assert s.ustore[0] == pytest.approx(cap, rel=1e-15) if over else s.ustore[0] == cap + over
        assert s.actual_infiltration == pytest.approx(3e-6 - over / dt, rel=1e-13, abs=1e-22)
    s = _soil(1, [0.5], infiltration=3e-6, ustore=[0.01])
    O.soil_call("soil_actual_infiltration", s, dt)
    assert s.ustore[0] == 0.01 and s.actual_infiltration == 3e-6


def test_actual_infiltration_cascade_conserves_water():
    """Several layers: water above a layer's capacity moves one layer up, and what the top layer
    cannot hold leaves as the reduction of the infiltration. Total water is conserved, every
    layer ends at or below capacity, and the amount that leaves is the closed form of the
    bottom-up cascade, max(0, u0 - c0 + max(0, u1 - c1 + max(0, u2 - c2)))."""
    rng = np.random.default_rng(5)
    dt = 86400.0
    for _ in range(200):
        nl = int(rng.integers(2, 5))
        thick = list(rng.uniform(0.05, 0.8, nl))
        s = _soil(nl, thick, infiltration=1e-6)
        caps = [t * (0.45 - 0.05) for t in thick]
        u = [c * rng.uniform(0.2, 1.6) for c in caps]
        s.set(ustore=u)
        O.soil_call("soil_actual_infiltration", s, dt)
        after = s.arr("ustore", nl)
        out = 0.0
        for k in range(nl - 1, -1, -1):
            out = max(0.0, u[k] + out - caps[k])
        assert np.all(after <= np.array(caps) * (1 + 1e-15))
        assert math.fsum(after) + out == pytest.approx(math.fsum(u), rel=1e-14)
        assert s.actual_infiltration == pytest.approx(1e-6 - out / dt, rel=1e-12, abs=1e-20)


def test_leakage_closed_forms():
    """leakage! = max(0, min(max_leakage, min(drainable / dt, Ksat at the soil bottom))), with
    the bottom layer's kvfrac and the depth of the soil bottom (or z_exp above it). Each regime
    is reached with values whose result is exact."""
    dt = 3600.0
    kv0 = 2.0 ** -20
    base = dict(kv_0=kv0, f=0.0, max_leakage=1.0, drainable=1.0, zi=0.3)
    # Ksat-limited: f = 0, so Ksat = kvfrac[nl-1] kv_0 exactly (and not layer 0's kvfrac)
    s = _soil(3, [0.1, 0.3, 0.8], **base)
    s.set(kvfrac=[1.0, 0.75, 0.5])
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == 0.5 * kv0
    # max_leakage-limited
    s = _soil(3, [0.1, 0.3, 0.8], **dict(base, max_leakage=2.0 ** -24))
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == 2.0 ** -24
    # drainable-limited
    s = _soil(3, [0.1, 0.3, 0.8], **dict(base, drainable=2.0 ** -40 * dt))
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == 2.0 ** -40
    # zero maximum leakage (the synthetic workloads' default)
    s = _soil(3, [0.1, 0.3, 0.8], **dict(base, max_leakage=0.0))
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == 0.0
    # exponential decay with the soil-bottom depth 1.2 m, not the water table 0.3 m
    f = 1.7
    s = _soil(3, [0.1, 0.3, 0.8], **dict(base, f=f))
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == pytest.approx(kv0 * math.exp(-f * 1.2), rel=1e-15)
    # exponential-constant profile: Ksat frozen at z_exp above the soil bottom
    s = _soil(3, [0.1, 0.3, 0.8], **dict(base, f=f))
    s.z_exp = 0.4
    O.soil_call("soil_leakage", s, dt)
    assert s.leakage == pytest.approx(kv0 * math.exp(-f * 0.4), rel=1e-15)


def test_update_ustorelayerdepth_closed_forms():
    """update_ustorelayerdepth!: when the water table rises, each unsaturated layer keeps its
    water content (u / thickness) and layers that become saturated lose their store; when it
    falls, each layer gains the drained thickness times (theta_fc - theta_r)."""
    thick = [0.1, 0.3, 0.8, 1.0]
    cum = np.cumsum([0.0] + thick)
    for zi_prev, zi in [(1.7, 0.9), (1.7, 0.25), (0.9, 0.35), (0.35, 1.7), (0.25, 2.1), (0.9, 1.3)]:
        s = _soil(4, thick)
        # layer thicknesses above the previous water table (set_layerthickness)
        t0 = [min(max(zi_prev - cum[k], 0.0), thick[k]) for k in range(4)]
        n0 = sum(1 for t in t0 if t > 0)
        u0 = [0.02 * (k + 1) if k < n0 else 0.0 for k in range(4)]
        s.set(ustorethick=[t if t > 0 else float("nan") for t in t0], n_unsat=n0, ustore=u0, zi=zi_prev)
        O.lib().wfo_update_ustorelayerdepth(O.ctypes.byref(s), O._c(zi_prev), O._c(zi))
        t1 = [min(max(zi - cum[k], 0.0), thick[k]) for k in range(4)]
        u1 = s.arr("ustore", 4)
        assert s.zi == zi and s.n_unsat == sum(1 for t in t1 if t > 0)
        for k in range(4):
            if zi < zi_prev:  # rise
                if t1[k] > 0:
                    assert u1[k] * t0[k] == pytest.approx(u0[k] * t1[k], rel=1e-14), (zi_prev, zi, k)
                elif k < n0:
                    assert u1[k] == 0.0
            else:  # fall
                assert u1[k] - u0[k] == pytest.approx((t1[k] - t0[k]) * (0.25 - 0.05), rel=1e-13, abs=1e-17)


def test_kw_ssf_newton_raphson_linear_root():
    """kw_ssf_newton_raphson solves q dt/dx + q / c = C, which is linear in q: the root is
    C / (dt/dx + 1/c) exactly; the oracle must agree to the rounding of its two evaluations."""
    rng = np.random.default_rng(7)
    for _ in range(500):
        dt = float(rng.choice([900.0, 3600.0, 86400.0]))
        dx = float(rng.uniform(100.0, 1500.0))
        cel = float(10.0 ** rng.uniform(-9, -2))
        C = float(10.0 ** rng.uniform(-8, 3))
        q0 = float(10.0 ** rng.uniform(-9, 1))
        got = O.kw_ssf_newton_raphson(q0, C, cel, dt, dx)
        root = Fraction(C) / (Fraction(dt) / Fraction(dx) + 1 / Fraction(cel))
        assert abs(Fraction(got) - root) <= root * Fraction(1, 10 ** 12), (q0, C, cel, dt, dx)


def _kw_root_bracket(dtdx, alpha, C, bits=200):
    """Bisection in exact rationals on f(u) = dtdx u^5 + alpha u^3 - C (increasing for u > 0)."""
    dtdx, alpha, C = Fraction(dtdx), Fraction(alpha), Fraction(C)
    lo, hi = Fraction(0), Fraction(1)
    while dtdx * hi ** 5 + alpha * hi ** 3 < C:
        hi *= 2
    for _ in range(bits):
        mid = (lo + hi) / 2
        if dtdx * mid ** 5 + alpha * mid ** 3 < C:
            lo = mid
        else:
            hi = mid
    return lo, hi


def test_kinematic_wave_root_bracket():
    """kinematic_wave stops where |f(u)| <= 1e-12 in floating point: its q = u^5 lies within
    the root's bracket widened by what that residual (plus the rounding of f) allows,
    |u - u*| <= (1e-12 + 8 eps C) / f'(u*). The root is bracketed by bisection in exact
    rationals. A dropped term, a wrong exponent or a wrong coefficient moves q far outside."""
    rng = np.random.default_rng(11)
    eps = 2.0 ** -52
    checked = 0
    for _ in range(120):
        dt = float(rng.choice([900.0, 3600.0]))
        dx = float(rng.uniform(200.0, 1400.0))
        alpha = float(rng.uniform(0.5, 8.0))
        q_in = float(10.0 ** rng.uniform(-6, 1.5))
        q_prev = float(10.0 ** rng.uniform(-6, 1.5))
        q_lat = float(10.0 ** rng.uniform(-8, -3))
        q, A = O.kinematic_wave(q_in, q_prev, q_lat, alpha, dt, dx)
        dtdx = dt / dx
        u_prev = float(O.math("pow", np.array([q_prev]), np.array([0.2]))[0])  # the reference's pow
        C = dtdx * q_in + alpha * u_prev * u_prev * u_prev + dt * q_lat
        lo, hi = _kw_root_bracket(dtdx, alpha, C)
        ustar = float(lo)
        fprime = 5 * dtdx * ustar ** 4 + 3 * alpha * ustar ** 2
        du = Fraction((1e-12 + 8 * eps * C) / fprime)
        qlo, qhi = (lo - du) ** 5, (hi + du) ** 5
        qf = Fraction(q)
        tol = Fraction(4 * eps) * qhi  # rounding of q = u^3 u u
        assert qlo - tol <= qf <= qhi + tol, (q_in, q_prev, q_lat, alpha, dt, dx)
        checked += 1
    assert checked == 120
