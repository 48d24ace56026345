"""The Julia module julia/WflowHIP.jl (not runnable here: no Julia) against the library's ABI: its
wfh_config / wfh_static mirrors have the header's fields in the header's order with the same
sizes (the ctypes mirror it is compared with is itself checked against the compiled library's
offsets), its wfh_var list numbers the variables as the header does, and every entry point it
calls is exported."""
import ctypes
import os
import re

from wflowhip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "julia", "WflowHIP.jl")).read()

JL_SIZE = {"Int32": 4, "Int64": 8, "Float64": 8, "PD": 8}


def _fields(struct):
    body = re.search(r"struct %s\n(.*?)\nend" % struct, SRC, re.S).group(1)
    out = []
    for line in body.splitlines():
        m = re.match(r"\s+(\w+)::([\w{}, ]+?)(\s*=.*)?$", line)
        if m:
            out.append((m.group(1), m.group(2).strip()))
    return out


def _size(t):
    if t.startswith("Ptr{"):
        return 8
    m = re.match(r"NTuple\{(\w+), Float64\}", t)
    if m:
        return 8 * (_abi.WFH_MAX_LAYERS if m.group(1) == "WFH_MAX_LAYERS" else int(m.group(1)))
    return JL_SIZE[t]


def test_config_and_static_mirror_the_header():
    for jl, ct in (("WfhConfig", _abi.WfhConfig._fields_), ("WfhStatic", _abi.STATIC_FIELDS)):
        f = _fields(jl)
        assert [n for n, _ in f] == [n for n, _ in ct], jl
        for (n, t), (_, c) in zip(f, ct):
            assert _size(t) == ctypes.sizeof(c), (jl, n, t)


def test_var_numbering_and_entry_points():
    names = re.findall(r":(\w+)", re.search(r"const WFH_VARS = \((.*?)\)\n", SRC, re.S).group(1))
    assert names == [v.name for v in _abi.Var][: len(names)]
    for acc in re.findall(r"^\s+:(\w+) =>", SRC, re.M):
        assert acc in names, acc
    for sym in set(re.findall(r"ccall\(\(:(\w+), LIB\)", SRC)):
        assert sym in _abi.EXPORTS, sym
    assert re.search(r"WFH_ABI_VERSION = Int32\((\d+)\)", SRC).group(1) == str(_abi.WFH_ABI_VERSION)
