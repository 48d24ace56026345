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
