"""The non-finite forcing test of the kernels (include/wfh_math.h wfh_nonfinite3, compiled for
the host): NaN and infinities of any sign in any of the three forcing values are caught, and no
finite value is (largest, smallest normal and subnormal, zeros)."""
import ctypes
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("nf") / "libnf.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-mfma",
                    "-D__host__=", "-D__device__=", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "wflow.jl_amd", "csrc"), "-o", str(out),
                    os.path.join(HERE, "physics_host.cpp")], check=True)
    L = ctypes.CDLL(str(out))
    L.wfh_host_nonfinite3.argtypes = [ctypes.c_double] * 3
    L.wfh_host_nonfinite3.restype = ctypes.c_int
    return L


def test_nonfinite_values_are_caught(lib):
    bad = [math.nan, -math.nan, math.inf, -math.inf]
    for v in bad:
        for pos in range(3):
            args = [1e-7, 2e-8, 290.0]
            args[pos] = v
            assert lib.wfh_host_nonfinite3(*args) == 1, (v, pos)


def test_finite_values_pass(lib):
    for v in (0.0, -0.0, 1.7976931348623157e308, -1.7976931348623157e308, 2.2250738585072014e-308, 5e-324, 1e-30):
        assert lib.wfh_host_nonfinite3(v, v, v) == 0, v
