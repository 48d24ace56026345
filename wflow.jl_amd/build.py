"""Build recipe of the package's native libraries (in-tree, into wflow.jl_amd/lib/).

  libwflow_hip.so   the HIP implementation behind include/wflow_hip.h (gfx950)
  libwfh_synth.so   host-only synthetic catchment generator (depression filling, D8)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",                      # keep the reference's operation order (no FMA fusion)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=HERE)


def build(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc")]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", os.path.join(LIB, "libwfh_synth.so"),
          os.path.join(HERE, "csrc", "synth.cpp")])
    _run(["hipcc", *HIP_FLAGS, *inc, "-o", os.path.join(LIB, "libwflow_hip.so"),
          os.path.join(HERE, "csrc", "wflow_hip.hip")])


if __name__ == "__main__":
    build()
    sys.exit(0)
