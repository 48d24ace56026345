"""The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle's CPU
tests run again in a child process that loads the sanitizer build of the same sources
(oracle/Makefile, target `sanitize`) with the sanitizer runtime preloaded. An out-of-bounds
access, a use after free or undefined behaviour (signed overflow, a bad shift, a misaligned or
null access) aborts the child with the sanitizer's report."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_LIB = os.path.join(ORACLE, "build", "libwflow_oracle_san.so")
TESTS = ["tests/test_oracle_kat.py", "tests/test_oracle_network.py", "tests/test_reservoir_cpu.py",
         "tests/test_reservoir_curves_cpu.py", "tests/test_floodplain_cpu.py", "tests/test_snow_transport_cpu.py",
         "tests/test_adaptive_cpu.py", "tests/test_kinwave_shortcut.py"]


def _env():
    rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(rt) or not os.path.exists(rt):
        pytest.skip("gcc has no AddressSanitizer runtime")
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["WFO_LIB"] = SAN_LIB
    # Python itself is not instrumented and keeps memory until exit: leak reports are off
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def test_oracle_tests_pass_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = _env()
    probe = ("import sys; sys.path.insert(0, 'oracle'); import oracle; oracle.lib(); "
             "print(open('/proc/self/maps').read())")
    maps = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=300)
    assert maps.returncode == 0, maps.stderr[-2000:]
    assert "libwflow_oracle_san.so" in maps.stdout and "libasan" in maps.stdout
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        *TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in r.stdout
