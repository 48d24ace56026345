"""Batched catchments: several members stacked into one model (synth.make_batch).

A member's results inside a batch must equal the member run on its own, bit for bit: basins
exchange no water and the step does not depend on the order in which basins are visited. The
CPU test checks the batch construction on the oracle; the GPU test checks the HIP path on a
batch of members of uneven size against per-member oracle runs."""
import numpy as np
import pytest

import oracle as O
from wflowhip import Var, synth, _abi

CELL_VARS = [Var.SATWATERDEPTH, Var.USTORELAYERDEPTH, Var.SSF_Q, Var.LAND_Q, Var.LAND_H, Var.TOTAL_STORAGE,
             Var.BALANCE_LAND_ERROR, Var.SSF_Q_AVERAGE, Var.CANOPY_STORAGE]
RIVER_VARS = [Var.RIVER_Q, Var.RIVER_H, Var.RIVER_Q_AVERAGE, Var.BALANCE_RIVER_ERROR]


def members_small():
    return [synth.make_catchment(40, 24, cell=1000.0, dt=3600.0, river_threshold=40, seed=31),
            synth.make_catchment(24, 36, cell=500.0, dt=3600.0, river_threshold=30, seed=32),
            synth.make_catchment(56, 20, cell=250.0, dt=3600.0, river_threshold=50, seed=33)]


def member_slices(b, m, var, nlayers):
    if var in _abi.RIVER_VARS:
        return slice(b.river_offsets[m], b.river_offsets[m + 1]), None
    if var in _abi.LAYERED_VARS:
        return slice(b.cell_offsets[m], b.cell_offsets[m + 1]), nlayers
    return slice(b.cell_offsets[m], b.cell_offsets[m + 1]), None


def pick(b, m, var, values):
    sl, nl = member_slices(b, m, var, b.nlayers)
    if nl is None:
        return values[sl]
    return values.reshape(nl, -1)[:, sl].ravel()


def check_members(batch_model, b, ms, refs):
    for m, ref in enumerate(refs):
        for v in CELL_VARS + RIVER_VARS:
            got = pick(b, m, v, batch_model.get(v))
            exp = ref.get(v)
            assert got.shape == exp.shape, (m, v)
            same = (got == exp) | (np.isnan(got) & np.isnan(exp))
            assert same.all(), (m, v.name)


def test_batch_layout():
    ms = members_small()
    b = synth.make_batch(ms)
    assert b.n == sum(c.n for c in ms) and b.nriv == sum(c.nriv for c in ms)
    assert np.all(np.diff(b.arrays["cell_index"]) > 0)
    assert list(b.cell_offsets) == list(np.cumsum([0] + [c.n for c in ms]))
    # brooks_corey_c is layer-major: member 1's layer-2 values sit in batch layer 2
    n1 = ms[1].n
    got = b.arrays["brooks_corey_c"].reshape(4, -1)[2, b.cell_offsets[1]:b.cell_offsets[2]]
    assert np.array_equal(got, ms[1].arrays["brooks_corey_c"].reshape(4, n1)[2])


def test_batch_equals_members_oracle():
    ms = members_small()
    b = synth.make_batch(ms)
    bm = O.OracleModel(b)
    refs = [O.OracleModel(c) for c in ms]
    for k in range(4):
        P, E, T = synth.forcing_batch_numpy(ms, k)
        bm.step(P, E, T)
        for m, c in enumerate(ms):
            refs[m].step(*synth.forcing_numpy(c, k, seed=7 * m))
    check_members(bm, b, ms, refs)


@pytest.mark.gpu
def test_batch_gpu_equals_member_oracles():
    from wflowhip import Model
    ms = members_small() + [synth.make_catchment(96, cell=250.0, dt=3600.0, river_threshold=200, seed=34)]
    b = synth.make_batch(ms)
    g = Model(b)
    refs = [O.OracleModel(c) for c in ms]
    for k in range(6):
        g.run_timestep(*synth.forcing_batch_numpy(ms, k))
        for m, c in enumerate(ms):
            refs[m].step(*synth.forcing_numpy(c, k, seed=7 * m))
    check_members(g, b, ms, refs)
    g.close()


@pytest.mark.gpu
def test_batch_64_catchments_real_sizes():
    """The 64-catchment batch at its real sizes (log-uniform 1e4 .. 2e6 cells, seeded as in
    bench.py --workload batch, 21.9 M cells, dx 1 km, hourly) in one model on the GPU, each
    member against its own oracle run: the smallest, median and largest members over 24 steps,
    every other member over the first 6 (members run in parallel threads)."""
    import math
    import sys
    import time
    from concurrent.futures import ThreadPoolExecutor

    from wflowhip import Model

    def say(msg):
        sys.__stderr__.write(f"[batch 64] {msg}\n")
        sys.__stderr__.flush()

    t0 = time.time()
    rng = np.random.default_rng(2)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e6), 64))
    ms = [synth.make_catchment(int(round(math.sqrt(s))), cell=1000.0, dt=3600.0, river_threshold=100, seed=1000 + m)
          for m, s in enumerate(sizes)]
    b = synth.make_batch(ms)
    g = Model(b)
    order = np.argsort([c.n for c in ms])
    long_run = {int(order[0]), int(order[len(order) // 2]), int(order[-1])}
    refs = [O.OracleModel(c) for c in ms]
    say(f"{b.n} cells in 64 members ({ms[order[0]].n} .. {ms[order[-1]].n}); set up in {time.time() - t0:.0f}s")
    pool = ThreadPoolExecutor(16)
    for k in range(24):
        active = [m for m in range(64) if k < 6 or m in long_run]
        fs = list(pool.map(lambda m: synth.forcing_numpy(ms[m], k, seed=7 * m), range(64)))
        g.run_timestep(*(np.ascontiguousarray(np.concatenate([f[i] for f in fs])) for i in range(3)))
        list(pool.map(lambda m: refs[m].step(*fs[m]), active))
        if k == 5 or k == 23:
            checked = active
            vals = {v: g.get(v) for v in CELL_VARS + RIVER_VARS}
            for m in checked:
                for v in CELL_VARS + RIVER_VARS:
                    got = pick(b, m, v, vals[v])
                    exp = refs[m].get(v)
                    same = (got == exp) | (np.isnan(got) & np.isnan(exp))
                    assert same.all(), (k, m, v.name)
            say(f"step {k + 1}: members {sorted(checked) if len(checked) < 64 else 'all 64'} equal to their oracle "
                f"runs ({time.time() - t0:.0f}s)")
    pool.shutdown()
    g.close()
