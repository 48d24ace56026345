"""The learned Newton warm start in the ensemble (configs: learned kinematic-wave start, batched).

The network is trained on the GPU on rows the plain solver harvests from another catchment, then
proposes the starts of an ensemble's overland and river solves. Reported: the Newton-update
histograms with and without it, how many solves took its start, and the per-member error
against the FP64 oracle (the plain method). Only the start changes; the iteration and its 1e-12
residual stop are the reference's, so results stay within that stop's band of the plain method,
and every member's water balance still closes."""
import sys

import numpy as np
import pytest

import oracle as O
from wflowhip import Ensemble, Var, member_catchment, synth
from wflowhip import learned as LW

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _factors(M, seed):
    f = np.exp(0.3 * np.random.default_rng(seed).standard_normal((3, M)))
    f[:, 0] = 1.0
    return f


def test_learned_warm_start_trains_and_keeps_results_within_the_stop_band():
    train_c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=501)
    rows = LW.harvest(train_c, _factors(64, 11), 12, lambda k: synth.forcing_numpy(train_c, k), every=8,
                      cap=400_000)
    _say(f"harvested {rows.shape[0]} rows")
    assert rows.shape[0] > 50_000
    net, losses = LW.fit(rows, epochs=6, log=_say)
    assert losses[-1] < losses[0]

    c = synth.make_catchment(48, cell=250.0, dt=3600.0, river_threshold=40, seed=502)
    fa, fb, fc = _factors(64, 12)
    plain = Ensemble(c, fa, fb, fc)
    learn = Ensemble(c, fa, fb, fc)
    learn.set_learned(net)
    plain.kw_histogram(True)
    learn.kw_histogram(True)
    refs = {m: O.OracleModel(member_catchment(c, fa[m], fb[m], fc[m])) for m in (0, 9)}
    worst = {}
    for k in range(10):
        P, E, T = synth.forcing_numpy(c, 100 + k)
        plain.run_timestep(P, E, T)
        learn.run_timestep(P, E, T)
        for ref in refs.values():
            ref.step(P, E, T)
        assert np.all(learn.balance() < 1e-9)
        for m, ref in refs.items():
            for v in (Var.LAND_Q, Var.RIVER_Q, Var.SATWATERDEPTH, Var.SSF_Q):
                a, b = learn.get(v, m), ref.get(v)
                big = np.abs(b) > 1e-3
                if big.any():
                    worst[v.name] = max(worst.get(v.name, 0.0), float(np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))))
                # the plain ensemble is bitwise the oracle
                assert np.array_equal(plain.get(v, m), b)
    hp, _ = plain.kw_histogram(False)
    hl, used = learn.kw_histogram(False)
    for fam, name in ((0, "overland"), (1, "river")):
        _say(f"{name}: plain mean updates {LW.mean_updates(hp[fam]):.3f}, learned {LW.mean_updates(hl[fam]):.3f}; "
             f"learned starts taken {int(used[fam])} of {int(hl[fam].sum())}")
        _say(f"  histogram plain   {hp[fam].tolist()}")
        _say(f"  histogram learned {hl[fam].tolist()}")
    _say(f"max relative error against the oracle (cells with |value| > 1e-3): {worst}")
    assert int(used.sum()) > 0
    assert all(v < 1e-6 for v in worst.values()), worst
    for e in (plain, learn):
        e.close()
    net.close()
