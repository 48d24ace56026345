"""The kernels' kinematic_wave (wflow.jl_amd/csrc/wfh_physics.h, compiled for the host) against
the oracle's literal restatement of the reference loop (Wflow surface_process.jl:24-70).

The kernels shortcut the reference's 3000-update limit once the Newton iterate repeats (a
fixed point or a two-cycle, which is what happens when |C| is too large for the absolute
1e-12 residual test to ever pass). The result must be bit-identical to running all updates.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def host_physics(tmp_path_factory):
    out = tmp_path_factory.mktemp("phys") / "libphys.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-mfma",
                    "-D__host__=", "-D__device__=", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "wflow.jl_amd", "csrc"), "-o", str(out),
                    os.path.join(HERE, "physics_host.cpp")], check=True)
    L = ctypes.CDLL(str(out))
    L.wfh_host_kinematic_wave.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return L


def oracle_iterations(row):
    L = O.lib()
    o = (ctypes.c_double * 2)()
    it = ctypes.c_int64(0)
    L.wfo_kinematic_wave_it.argtypes = [ctypes.c_double] * 6 + [ctypes.c_void_p, ctypes.c_void_p]
    L.wfo_kinematic_wave_it(*[float(x) for x in row], o, ctypes.addressof(it))
    return o[0], o[1], it.value


def test_kinematic_wave_host_build_equals_oracle(host_physics):
    rng = np.random.default_rng(7)
    n = 4000
    q_in = 10.0 ** rng.uniform(-8, 5.5, n) * (rng.random(n) > 0.05)
    q_prev = np.where(rng.random(n) < 0.15, 1e-30, 10.0 ** rng.uniform(-8, 5.5, n))
    q_lat = 10.0 ** rng.uniform(-9, -1, n) * (rng.random(n) > 0.1)
    alpha = rng.uniform(0.5, 20.0, n)
    dt = rng.choice([900.0, 3600.0, 994.6119029285007], n)
    dx = rng.uniform(200.0, 1500.0, n)
    rows = np.ascontiguousarray(np.stack([q_in, q_prev, q_lat, alpha, dt, dx], axis=1))
    got = np.empty((n, 2))
    host_physics.wfh_host_kinematic_wave(rows.ctypes.data, got.ctypes.data, n)
    ref = np.array([oracle_iterations(r)[:2] for r in rows])
    iters = np.array([oracle_iterations(r)[2] for r in rows])
    assert (iters >= 3000).sum() >= 50, "the sample must exercise the iteration limit"
    assert np.array_equal(got, ref)


def test_kinematic_wave_limit_cases_two_cycle_parity(host_physics):
    """Large constant terms that end on a two-cycle: both parities of the cycle entry."""
    # q_prev = 0.98 q_in enters its two-cycles after an odd number of updates from the end,
    # q_prev = 0.1 q_in after an even number (checked with a plain restatement of the loop)
    rows = []
    for fac in (0.98, 0.1):
        for C_q in np.linspace(8e3, 9e4, 400):
            rows.append((C_q, C_q * fac, 1e-4, 3.0, 900.0, 250.0))
    rows = np.ascontiguousarray(np.array(rows))
    got = np.empty((len(rows), 2))
    host_physics.wfh_host_kinematic_wave(rows.ctypes.data, got.ctypes.data, len(rows))
    ref = np.array([oracle_iterations(r)[:2] for r in rows])
    assert np.array_equal(got, ref)
