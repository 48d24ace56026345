"""C ABI checks that need no GPU: the library loads, exports every entry point declared in
include/wflow_hip.h, and rejects bad arguments before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from wflowhip import _abi, synth
from wflowhip.model import load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "include", "wflow_hip.h")).read()


def header_functions():
    return sorted(set(re.findall(r"^(?:int32_t|int64_t|void|const char\*)\s+(wfh_\w+)\(", HEADER, re.M)))


def test_header_functions_are_exported():
    lib = load_library()
    names = header_functions()
    assert set(names) == set(_abi.EXPORTS), names
    for name in names:
        assert hasattr(lib, name), name


def test_var_enum_matches_header():
    body = HEADER.split("typedef enum wfh_var {")[1].split("} wfh_var;")[0]
    names = re.findall(r"^\s*WFH_(\w+)\s*(?:=\s*0)?,?", body, re.M)
    assert names == [v.name for v in _abi.Var], names


def test_struct_fields_match_header():
    body = HEADER.split("typedef struct wfh_static {")[1].split("} wfh_static;")[0]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            fields.append(re.findall(r"\w+", part)[-1])
    assert fields == [f for f, _ in _abi.STATIC_FIELDS]


def _create(c, **kw):
    lib = load_library()
    cfg = c.config(**kw)
    st = c.static()
    h = ctypes.c_void_p()
    rc = lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h))
    return rc, lib.wfh_last_error(None).decode(), h


def test_create_rejects_bad_arguments():
    c = synth.make_catchment(16, cell=1000.0, dt=86400.0, river_threshold=10, seed=2)
    lib = load_library()
    h = ctypes.c_void_p()
    assert lib.wfh_create(None, None, ctypes.byref(h)) == _abi.WFH_ERR_ARG
    cfg = c.config()
    cfg.abi_version = 99
    st = c.static()
    assert lib.wfh_create(ctypes.byref(cfg), ctypes.byref(st), ctypes.byref(h)) == _abi.WFH_ERR_ARG
    c2 = synth.make_catchment(16, cell=1000.0, dt=86400.0, river_threshold=10, seed=2)
    del c2.arrays["theta_s"]
    rc, msg, _ = _create(c2)
    assert rc == _abi.WFH_ERR_ARG and "theta_s" in msg


def test_create_detects_cycle():
    """flowgraph raises on a cyclic LDD (Wflow/src/routing/utils.jl:26-31)."""
    c = synth.make_catchment(16, cell=1000.0, dt=86400.0, river_threshold=10, seed=2)
    ldd = c.arrays["ldd"]
    ldd[0], ldd[1] = 6, 4  # cells (0,0) -> (1,0) -> (0,0)
    rc, msg, _ = _create(c)
    assert rc == _abi.WFH_ERR_CYCLE, msg


def test_struct_layout_matches_header(tmp_path):
    """Offsets and sizes of wfh_config and wfh_static as a C compiler lays them out, against the
    ctypes mirror (a field of the wrong width would shift every field after it)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = (("wfh_config", _abi.WfhConfig), ("wfh_static", _abi.WfhStatic))
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "wflow_hip.h"', "int main(void) {"]
    for name, py in structs:
        lines.append(f'printf("%zu\\n", sizeof({name}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cc, "-I", include, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = []
    for _, py in structs:
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want


def test_run_rejects_a_null_handle():
    """wfh_run's first argument check runs without a device: a null handle is WFH_ERR_ARG."""
    from wflowhip.model import load_library

    lib = load_library()
    assert lib.wfh_run(None, 1, None, None, None, None) == -1
