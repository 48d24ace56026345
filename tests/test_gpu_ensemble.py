"""Parameter ensemble (wfh_ens_*) on the GPU against the oracle.

Each member m is the catchment with kv_0, f and Manning n scaled by its factors; the oracle runs
that member's catchment as a model of its own, and every state of the member must match it to
1e-10 relative per cell and step (bitwise in practice), with the same largest land-balance error.
The schedule (stages, strands, block size) must not change any bit, and the autotuner's steps are
real steps of the model."""
import sys

import numpy as np
import pytest

import oracle as O
from wflowhip import Ensemble, Var, member_catchment, synth

pytestmark = pytest.mark.gpu

RTOL = 1e-10
ENS_STATES = [Var.CANOPY_STORAGE, Var.SNOW_STORAGE, Var.SNOW_WATER, Var.SOIL_TEMPERATURE, Var.SATWATERDEPTH,
              Var.USTORELAYERDEPTH, Var.SSF_Q, Var.LAND_Q, Var.LAND_H, Var.RIVER_Q, Var.RIVER_H, Var.RIVER_STORAGE,
              Var.RIVER_Q_AVERAGE]
PARAMS = [Var.LAND_ALPHA, Var.RIVER_ALPHA, Var.SSF_QMAX]


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    d[same] = 0.0
    return float(d.max()) if d.size else 0.0


def factors(members, seed, sigma=0.3):
    rng = np.random.default_rng(seed)
    f = np.exp(sigma * rng.standard_normal((3, members)))
    f[:, 0] = 1.0  # member 0 is the unperturbed catchment
    return f


def compare_member(ens, m, ref, where, vars_=ENS_STATES):
    bad = {}
    for v in vars_:
        e = rel_err(ens.get(v, m), ref.get(v))
        if not e <= RTOL:
            bad[v.name] = e
    assert not bad, f"{where} member {m}: {bad}"


def run_against_oracle(c, members, steps, check_members, seed=1, **kw):
    fa, fb, fc = factors(members, seed)
    ens = Ensemble(c, fa, fb, fc, **kw)
    refs = {m: O.OracleModel(member_catchment(c, fa[m], fb[m], fc[m]), **kw) for m in check_members}
    for m, ref in refs.items():
        compare_member(ens, m, ref, "init", ENS_STATES + PARAMS)
    for k in range(steps):
        P, E, T = synth.forcing_numpy(c, k)
        ens.run_timestep(P, E, T)
        bal = ens.balance()
        for m, ref in refs.items():
            ref.step(P, E, T)
            compare_member(ens, m, ref, f"step {k}")
            assert bal[m] == np.max(np.abs(ref.get(Var.BALANCE_LAND_ERROR))), (k, m)
        assert np.all(bal < 1e-9)
    for ref in refs.values():
        ref.close()
    return ens, (fa, fb, fc)


def test_ensemble_hourly_matches_oracle_per_member():
    _say("ensemble 48x48 hourly, 64 members, 8 steps, members 0 1 31 63 against the oracle")
    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=40, seed=401)
    ens, _ = run_against_oracle(c, 64, 8, [0, 1, 31, 63])
    ens.close()


def test_ensemble_six_river_substeps():
    """Six river sub-steps per step (river_dt 600 s): the chain stages read the inflow of the first
    four sub-steps ahead and of the others in the sub-step loop."""
    _say("ensemble 40x40 hourly, river_dt 600 s (6 river sub-steps), 64 members, 4 steps")
    c = synth.make_catchment(40, cell=250.0, dt=3600.0, river_threshold=30, seed=403)
    ens, _ = run_against_oracle(c, 64, 4, [0, 45], river_dt=600.0)
    ens.close()


def test_ensemble_daily_two_groups():
    _say("ensemble 40x40 daily (land_dt = dt), 128 members, 5 steps")
    c = synth.make_catchment(40, cell=1000.0, dt=86400.0, river_threshold=30, seed=402)
    ens, _ = run_against_oracle(c, 128, 5, [0, 64, 127], land_dt=86400.0, river_dt=21600.0)
    ens.close()


def test_ensemble_schedules_are_bitwise_equal():
    """Every schedule (wide-level threshold, cells per strand, block size) and the autotuner
    give the same bits for every member."""
    _say("ensemble schedules 64x64 hourly, 64 members")
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=403)
    fa, fb, fc = factors(64, 3)
    runs = []
    for sched in [None, (64, 1, 1), (256, 10 ** 9, 1), (128, 16, 16), "autotune"]:
        ens = Ensemble(c, fa, fb, fc)
        if sched == "autotune":
            ens.set_autotune(1)
        elif sched is not None:
            ens.set_schedule(*sched)
        for k in range(6):
            P, E, T = synth.forcing_numpy(c, k)
            ens.run_timestep(P, E, T)
        s = ens.schedule()
        _say(f"  schedule {sched}: {s}")
        runs.append({(v, m): ens.get(v, m) for v in ENS_STATES for m in (0, 5, 63)})
        ens.close()
    for r in runs[1:]:
        for key, a in r.items():
            b = runs[0][key]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), key


def test_ensemble_rejects_unsupported():
    c = synth.make_catchment(16, cell=1000.0, dt=86400.0, river_threshold=10, seed=404)
    f = np.ones(64)
    with pytest.raises(Exception):
        Ensemble(c, f[:32], f[:32], f[:32])  # not a multiple of 64
    with pytest.raises(Exception):
        Ensemble(c, f, f, f, adaptive=True)
    bad = f.copy()
    bad[3] = np.nan
    with pytest.raises(Exception):
        Ensemble(c, bad, f, f)
