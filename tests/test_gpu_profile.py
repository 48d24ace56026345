"""Exponential-constant Ksat profile on the GPU: against the oracle, and the routing kernel's
extended instantiation against the plain one where both must agree bitwise."""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import FLUX_VARS, STATE_VARS, compare, run_pair
from wflowhip import Model, synth

pytestmark = pytest.mark.gpu


def test_shallow_z_exp_daily():
    print("z_exp 0.3 m daily 48x48", file=sys.__stderr__, flush=True)
    c = synth.make_catchment(48, cell=500.0, dt=86400.0, river_threshold=40, seed=93)
    c.arrays["z_exp"] = np.random.default_rng(1).uniform(0.2, 1.5, c.n)
    gpu, ref = run_pair(c, 8)
    gpu.close()
    ref.close()


def test_z_exp_hourly_adaptive():
    print("z_exp hourly adaptive 64x64", file=sys.__stderr__, flush=True)
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=94)
    c.arrays["z_exp"] = np.full(c.n, 0.4)
    gpu, ref = run_pair(c, 6, adaptive=True)
    gpu.close()
    ref.close()


def test_extended_kernel_equals_plain_kernel():
    """z_exp at the soil thickness runs the extended routing kernel on a model whose results
    equal the plain exponential model's, so the two instantiations must agree to the bit."""
    print("extended vs plain kernel 64x64", file=sys.__stderr__, flush=True)
    c = synth.make_catchment(64, cell=250.0, dt=3600.0, river_threshold=60, seed=95)
    a = Model(c)
    c.arrays["z_exp"] = np.ascontiguousarray(c.arrays["soil_thickness"].copy())
    b = Model(c)
    for k in range(6):
        F = synth.forcing_numpy(c, k)
        a.run_timestep(*F)
        b.run_timestep(*F)
    for v in STATE_VARS + FLUX_VARS:
        assert np.array_equal(a.get_value(v), b.get_value(v), equal_nan=True), v
    a.close()
    b.close()
