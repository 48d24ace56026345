"""Edge cases of the step on the GPU against the oracle: a catchment without river cells, soil
layer counts from 1 to 5 configured layers (N = 2 and N = 6, the lowest layer filling the rest of
the soil), a ragged catchment (inactive raster cells, cell_index a strict subset), and a
catchment of two cells."""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare, run_pair
from wflowhip import Model, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def with_layers(c, layers_mm):
    """The same catchment with soil_layer__thickness = layers_mm (N = len + 1 layers)."""
    n = c.n
    N = len(layers_mm) + 1
    c.nlayers_cfg = len(layers_mm)
    c.layers_mm = tuple(float(x) for x in layers_mm)
    c.arrays["brooks_corey_c"] = np.ascontiguousarray(np.repeat(np.linspace(9.4, 10.3, N), n))
    c.arrays["kvfrac"] = np.ascontiguousarray(np.repeat(np.linspace(1.0, 0.6, N), n))
    c.arrays.pop("rootfraction", None)
    return c


def test_no_river_cells():
    _say("no river cells 40x40 (fixed and adaptive)")
    c = synth.make_catchment(40, cell=250.0, dt=3600.0, river_threshold=10 ** 9, seed=101)
    assert c.nriv == 0
    for kw in ({}, {"adaptive": True}):
        gpu, ref = run_pair(c, 4, **kw)
        gpu.close()
        ref.close()


@pytest.mark.parametrize("layers", [(300.0,), (50.0, 100.0, 200.0, 400.0, 600.0)])
def test_layer_counts(layers):
    _say(f"soil layers {layers} 48x48 daily")
    c = with_layers(synth.make_catchment(48, cell=1000.0, dt=86400.0, river_threshold=60, seed=102), layers)
    gpu, ref = run_pair(c, 5)
    gpu.close()
    ref.close()


def test_ragged_catchment():
    """Only the cells of some basins are active: cell_index skips raster cells (io.jl:1346-1351
    enumerates active cells only); no D8 direction points to an inactive cell."""
    _say("ragged catchment 64x64 basins subset")
    c = synth.make_catchment(64, cell=500.0, dt=3600.0, river_threshold=60, seed=103)
    lab = synth.basin_ids(c.arrays["ldd"], c.nx, c.ny)
    ids, counts = np.unique(lab, return_counts=True)
    keep = ids[np.argsort(-counts)[: max(2, ids.size // 3)]]
    sub = c.subset(np.isin(lab, keep))
    assert sub.n < c.n and sub.arrays["cell_index"][-1] >= sub.n
    gpu, ref = run_pair(sub, 6)
    gpu.close()
    ref.close()


def test_tiny_catchments():
    """A 2 x 2 raster (n = 4) and a single pit cell."""
    _say("tiny catchments: 2x2 and one cell")
    c = synth.make_catchment(2, cell=1000.0, dt=86400.0, river_threshold=10 ** 9, seed=104)
    gpu, ref = run_pair(c, 4)
    gpu.close()
    ref.close()
    pit = np.zeros(c.n, bool)
    pit[int(np.nonzero(c.arrays["ldd"] == 5)[0][0])] = True
    one = c.subset(pit)
    assert one.n == 1
    gpu, ref = run_pair(one, 4)
    gpu.close()
    ref.close()


def _land_chain_chunks(gpu):
    """Chunks of the node order that the routing kernels run as chains: full chunks of 64 with
    at most 4 upstream cells per lane, where lane l's only upstream cell inside the chunk is
    lane l - 1 (no reservoirs here, so the upstream lists are the flow graph's)."""
    lo, ld, _, _, _, _ = gpu.topology()
    node = gpu.node_order()
    pos = np.empty_like(node)
    pos[node] = np.arange(node.size)
    ups = [[] for _ in range(node.size)]
    for i in range(ld.size):
        if ld[i] >= 0:
            ups[pos[ld[i]]].append(int(pos[i]))
    count = 0
    for c in range(node.size // 64):
        ok = True
        for l in range(64):
            q = c * 64 + l
            inside = [u for u in ups[q] if u // 64 == c]
            ok = ok and len(ups[q]) <= 4 and inside == ([q - 1] if l > 0 else [])
        count += ok
    return count


def _river_chain_chunks(gpu):
    """Chunks of the node order whose 64 cells are all river cells forming a chain of the river
    lists (the persistent kernel runs their river family as a diagonal over a chain)."""
    lo, ld, ro, rd, rl, _ = gpu.topology()
    node = gpu.node_order()
    pos = np.empty_like(node)
    pos[node] = np.arange(node.size)
    river_at = np.full(node.size, -1)
    river_at[pos[rl]] = np.arange(rl.size)
    ups = [[] for _ in range(rl.size)]
    for r in range(rd.size):
        if rd[r] >= 0:
            ups[rd[r]].append(int(pos[rl[r]]))
    count = 0
    for c in range(node.size // 64):
        ok = True
        for l in range(64):
            q = c * 64 + l
            r = river_at[q]
            if r < 0:
                ok = False
                break
            inside = [u for u in ups[r] if u // 64 == c]
            ok = ok and len(ups[r]) <= 4 and inside == ([q - 1] if l > 0 else [])
        count += ok
    return count


def test_fixed_substeps_with_chain_chunks():
    """Hourly fixed sub-steps (1 subsurface, 1 overland, 4 river): the main stem's chunks run
    their subsurface and overland families as lane-by-lane chains and their river family as a
    diagonal over the chain in the persistent kernel."""
    _say("fixed sub-steps with chain chunks 512x512")
    c = synth.make_catchment(512, cell=250.0, dt=3600.0, river_threshold=500, seed=36)
    gpu = Model(c)
    n_chain, n_rchain = _land_chain_chunks(gpu), _river_chain_chunks(gpu)
    _say(f"  {n_chain} land chain chunks, {n_rchain} river chain chunks")
    assert n_chain > 0 and n_rchain > 0
    ref = O.OracleModel(c)
    for k in range(3):
        P, E, T = synth.forcing_numpy(c, k)
        gpu.run_timestep(P, E, T)
        ref.step(P, E, T)
        compare(gpu, ref, STATE_VARS + FLUX_VARS, f"step {k}")
    gpu.close()
    ref.close()
