"""Python binding of the CPU oracle (oracle/wflow_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.

Parity unpinned (no reference vector and no closed form): the functions listed as such in the
header of oracle/wflow_oracle.c: the soil-flow composition and the step end to end beyond the
parts pinned by tests/test_oracle_network.py, reservoir area means, the quantile's fma branch, and
elementary functions at the ulp level against Julia's Base.Math.
"""

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "wflow.jl_amd"))

from wflowhip import _abi  # noqa: E402

# WFO_LIB names another build of the same sources (the sanitizer build, oracle/Makefile)
LIB_PATH = os.environ.get("WFO_LIB") or os.path.join(HERE, "build", "libwflow_oracle.so")
_lib = None

_c = ctypes.c_double
_D = ctypes.POINTER(ctypes.c_double)


class SoilCell(ctypes.Structure):
    """ctypes mirror of wfo_soil (oracle/wflow_oracle.h)."""
    N = 8
    _fields_ = (
        [("N", ctypes.c_int32), ("nl", ctypes.c_int32), ("n_unsat", ctypes.c_int32), ("pad_", ctypes.c_int32)]
        + [(k, _c) for k in ("theta_s theta_r theta_fc soil_water_capacity soil_thickness kv_0 f "
                             "air_entry_pressure h1 h2 h3_high h3_low h4 alpha_h1 "
                             "infiltcap_soil infiltcap_compacted max_leakage compacted_fraction rootdistpar "
                             "cap_hmax cap_n w_soil cf_soil rooting_depth soil_fraction").split()]
        + [(k, _c * 8) for k in ("c", "kvfrac", "rootfraction", "act_thick")]
        + [("cum_depth", _c * 9)]
        + [(k, _c * 8) for k in ("ustore", "ustorethick")]
        + [(k, _c) for k in ("satwaterdepth zi drainable ustorecapacity ustoredepth total_soil_water_storage "
                             "tsoil f_infred h3 water_flux_surface pot_transpiration pot_soilevaporation "
                             "infiltration infiltration_excess transfer soil_evaporation soil_evap_sat "
                             "ae_ustore ae_sat transpiration actual_infiltration saturation_excess "
                             "act_inf_soil act_inf_compacted excess_soil excess_compacted "
                             "capflux leakage recharge aet runoff net_runoff exfiltration").split()]
        + [(k, _c * 8) for k in ("vwc", "rvwc")]
        + [(k, _c) for k in ("root_zone_storage", "vwc_root", "rvwc_root", "z_exp")]
    )

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.z_exp = float("inf")  # exponential profile

    def set(self, **kw):
        for k, v in kw.items():
            cur = getattr(self, k)
            if isinstance(cur, ctypes.Array):
                vals = list(v) + [float("nan")] * (len(cur) - len(v))
                for i, x in enumerate(vals):
                    cur[i] = x
            else:
                setattr(self, k, v)
        return self

    def arr(self, k, m=None):
        a = np.array(getattr(self, k)[:])
        return a if m is None else a[:m]


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.wfo_create.restype = ctypes.c_void_p
        L.wfo_create.argtypes = [ctypes.POINTER(_abi.WfhConfig), ctypes.POINTER(_abi.WfhStatic), ctypes.c_char_p,
                                 ctypes.c_int64]
        L.wfo_destroy.argtypes = [ctypes.c_void_p]
        L.wfo_step.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.wfo_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, _D, ctypes.c_int64]
        L.wfo_set.argtypes = [ctypes.c_void_p, ctypes.c_int32, _D, ctypes.c_int64]
        for f in ("wfo_n", "wfo_nriv", "wfo_kinwave_iterations"):
            getattr(L, f).restype = ctypes.c_int64
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.wfo_nlayers.argtypes = [ctypes.c_void_p]
        L.wfo_set_mutation.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.wfo_set_mutation.restype = ctypes.c_int32
        L.wfo_topology.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5
        L.wfo_toposort_dfs.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        dbl_fns = {
            "wfo_scurve": 4, "wfo_vwc_brooks_corey": 5, "wfo_head_brooks_corey": 5, "wfo_feddes_h3": 3,
            "wfo_rwu_reduction_feddes": 6, "wfo_soil_temperature": 3, "wfo_kw_ssf_newton_raphson": 5,
            "wfo_ssf_celerity": 6,
        }
        for name, k in dbl_fns.items():
            getattr(L, name).restype = _c
            getattr(L, name).argtypes = [_c] * k
        L.wfo_infiltration_reduction_factor.restype = _c
        L.wfo_infiltration_reduction_factor.argtypes = [_c, _c, ctypes.c_int, ctypes.c_int]
        L.wfo_soil_evaporation_unsaturated_store.restype = _c
        L.wfo_soil_evaporation_unsaturated_store.argtypes = [_c, _c, _c, ctypes.c_int, _c, _c]
        L.wfo_soil_evaporation_saturated_store.restype = _c
        L.wfo_soil_evaporation_saturated_store.argtypes = [_c, ctypes.c_int, _c, _c, _c, _c]
        L.wfo_math.argtypes = [ctypes.c_int32, _D, _D, _D, ctypes.c_int64]
        L.wfo_quantile.restype = _c
        L.wfo_quantile.argtypes = [_D, ctypes.c_int64, _c]
        L.wfo_kw_stable_timestep.restype = _c
        L.wfo_kw_stable_timestep.argtypes = [_D, _D, _D, ctypes.c_int64, _c, _D]
        L.wfo_last_substeps.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.wfo_update_reservoir.argtypes = [ctypes.c_void_p, _c, _c]
        L.wfo_accucapacityflux.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, _D, _D, _D, _c]
        L.wfo_update_reservoir_at_river.restype = _c
        L.wfo_update_reservoir_at_river.argtypes = [ctypes.c_void_p, _c, _c]
        L.wfo_interpolate_linear.restype = _c
        L.wfo_interpolate_linear.argtypes = [_c, _D, _D, ctypes.c_int64, ctypes.c_int64]
        for f in ("wfo_reservoir_maximum_storage", "wfo_reservoir_initial_storage"):
            getattr(L, f).restype = _c
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.wfo_floodplain_profile.argtypes = [ctypes.c_int64, _D, _D, ctypes.c_int64, _c, _c, _D]
        L.wfo_flood_depth.restype = _c
        L.wfo_flood_depth.argtypes = [_D, ctypes.c_int64, _D, _c, _c]
        L.wfo_floodplain_exchange.argtypes = [_D, ctypes.c_int64, _D] + [_c] * 6 + [_D]
        L.wfo_floodplain_capacity.restype = _c
        L.wfo_floodplain_capacity.argtypes = [_D, _D, ctypes.c_int64, _D, _c, _c, _c]
        L.wfo_set_julian_day.restype = ctypes.c_int32
        L.wfo_set_julian_day.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        _lib = L
    return _lib


MATH_OPS = {"exp": 0, "log": 1, "pow": 2, "cbrt": 3, "cld": 4, "round_sig12": 5, "min": 6, "max": 7, "pow10": 8}


def math(op, x, y=None):
    """Element-wise shared elementary function (include/wfh_math.h) over numpy arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    rc = lib().wfo_math(MATH_OPS[op], x.ctypes.data_as(_D), y.ctypes.data_as(_D), out.ctypes.data_as(_D), x.size)
    assert rc == 0
    return out


class Reservoir(ctypes.Structure):
    """wfo_reservoir (oracle/wflow_oracle.h): one reservoir of the reference's ReservoirModel."""
    _fields_ = [("outflow_type", ctypes.c_int32)] + [(f, ctypes.c_double) for f in (
        "area", "demand", "maximum_release", "maximum_storage", "target_full_fraction", "target_minimum_fraction",
        "threshold", "rating_curve_coefficient", "outflow_obs", "external_inflow", "precipitation", "evaporation",
        "inflow_overland", "inflow_subsurface", "storage", "waterlevel", "outflow", "inflow_cumulative",
        "outflow_cumulative", "actevap_cumulative", "abstraction_cumulative")] + [
        ("storage_curve_type", ctypes.c_int32), ("rating_curve_exponent", ctypes.c_double),
        ("sh_n", ctypes.c_int64), ("sh_h", _D), ("sh_s", _D),
        ("hq_n", ctypes.c_int64), ("hq_h", _D), ("hq_q", _D), ("hq_col", ctypes.c_int32),
        ("lower", ctypes.c_void_p)]

    @classmethod
    def make(cls, sh=None, hq=None, julian_day=1, **kw):
        """sh = (H, S) storage curve table (sets storage curve type 2); hq = (H, Q[rows][365])
        rating curve table. The tables are kept alive by the returned object."""
        r = cls()
        for f, t in cls._fields_:
            if t is ctypes.c_double:
                setattr(r, f, float("nan"))
        for f in ("external_inflow", "inflow_overland", "inflow_subsurface", "inflow_cumulative",
                  "outflow_cumulative", "actevap_cumulative", "abstraction_cumulative"):
            setattr(r, f, 0.0)
        r.storage_curve_type = 1
        r._keep = []
        if sh is not None:
            h, s_ = (np.ascontiguousarray(a, np.float64) for a in sh)
            r._keep += [h, s_]
            r.sh_n, r.sh_h, r.sh_s = h.size, h.ctypes.data_as(_D), s_.ctypes.data_as(_D)
            r.storage_curve_type = 2
        if hq is not None:
            h = np.ascontiguousarray(hq[0], np.float64)
            q = np.ascontiguousarray(hq[1], np.float64).reshape(h.size, 365)
            r._keep += [h, q]
            r.hq_n, r.hq_h, r.hq_q = h.size, h.ctypes.data_as(_D), q.ctypes.data_as(_D)
        r.hq_col = julian_day - 1
        for k, v in kw.items():
            setattr(r, k, float("nan") if v is None else v)
        return r


def interpolate_linear(x, xp, fp):
    """interpolate_linear(x, xp, fp)  reservoir.jl:364-377."""
    xp = np.ascontiguousarray(xp, np.float64)
    fp = np.ascontiguousarray(fp, np.float64)
    return lib().wfo_interpolate_linear(float(x), xp.ctypes.data_as(_D), fp.ctypes.data_as(_D), xp.size, 1)


def floodplain_profile(depth, volume, flow_width, flow_length):
    """FloodPlainProfile (floodplain.jl:48-147) of one river cell: returns (depth with 0,
    storage, width, flow area, wetted perimeter)."""
    d = np.ascontiguousarray(depth, np.float64)
    v = np.ascontiguousarray(volume, np.float64)
    out = np.zeros(4 * (d.size + 1))
    lib().wfo_floodplain_profile(d.size, d.ctypes.data_as(_D), v.ctypes.data_as(_D), 1, flow_width, flow_length,
                                 out.ctypes.data_as(_D))
    s, w, a, p = out.reshape(4, d.size + 1)
    return np.concatenate([[0.0], d]), s, w, a, p


def flood_depth(depth, volume, flow_width, flow_length, flood_storage):
    """compute_flood_depth (floodplain.jl:316-327) on the profile of floodplain_profile."""
    d = np.ascontiguousarray(depth, np.float64)
    v = np.ascontiguousarray(volume, np.float64)
    prof = np.zeros(4 * (d.size + 1))
    lib().wfo_floodplain_profile(d.size, d.ctypes.data_as(_D), v.ctypes.data_as(_D), 1, flow_width, flow_length,
                                 prof.ctypes.data_as(_D))
    d0 = np.concatenate([[0.0], d])
    return lib().wfo_flood_depth(prof.ctypes.data_as(_D), d.size + 1, d0.ctypes.data_as(_D), flood_storage,
                                 flow_length)


def _profile_buf(depth, volume, flow_width, flow_length):
    d = np.ascontiguousarray(depth, np.float64)
    v = np.ascontiguousarray(volume, np.float64)
    prof = np.zeros(4 * (d.size + 1))
    lib().wfo_floodplain_profile(d.size, d.ctypes.data_as(_D), v.ctypes.data_as(_D), 1, flow_width, flow_length,
                                 prof.ctypes.data_as(_D))
    return prof, np.concatenate([[0.0], d])


def floodplain_exchange(depth, volume, flow_width, flow_length, bankfull_depth, river_storage, fp_storage, dt):
    """river_channel_floodplain_exchange! (surface_kinwave.jl:572-601) for one river cell:
    returns (floodplain storage, flood depth, exchange flux)."""
    prof, d0 = _profile_buf(depth, volume, flow_width, flow_length)
    out = np.zeros(3)
    lib().wfo_floodplain_exchange(prof.ctypes.data_as(_D), d0.size, d0.ctypes.data_as(_D), river_storage, fp_storage,
                                  bankfull_depth, flow_width, flow_length, dt, out.ctypes.data_as(_D))
    return tuple(out)


def floodplain_capacity(depth, volume, flow_width, flow_length, h, mannings_n, slope, downstream=None):
    """Manning capacity of update_floodplain_model! (surface_kinwave.jl:395-425); downstream =
    (volume, flow_width, flow_length) of the downstream cell, None at a pit."""
    prof, d0 = _profile_buf(depth, volume, flow_width, flow_length)
    pd = None
    if downstream is not None:
        pd, _ = _profile_buf(depth, *downstream)
    return lib().wfo_floodplain_capacity(prof.ctypes.data_as(_D), None if pd is None else pd.ctypes.data_as(_D),
                                         d0.size, d0.ctypes.data_as(_D), h, mannings_n, slope)


def reservoir_maximum_storage(res):
    """get_maximum_storage  reservoir.jl:324-340."""
    return lib().wfo_reservoir_maximum_storage(ctypes.byref(res))


def reservoir_initial_storage(res):
    """initialize_storage  reservoir.jl:342-362."""
    return lib().wfo_reservoir_initial_storage(ctypes.byref(res))


def update_reservoir(res, inflow, dt):
    """update_reservoir_model!(reservoir_model, i, inflow, dt)  reservoir.jl:583-635."""
    lib().wfo_update_reservoir(ctypes.byref(res), inflow, dt)


def update_reservoir_at_river(res, q_v, dt):
    """update_reservoir_model! at a river cell (surface_kinwave.jl:441-489); returns the outflow."""
    return lib().wfo_update_reservoir_at_river(ctypes.byref(res), q_v, dt)


def accucapacityflux(order, down, material, capacity, dt):
    """accucapacityflux! (routing/utils.jl:82-108): returns (flux, new material)."""
    order = np.ascontiguousarray(order, np.int64)
    down = np.ascontiguousarray(down, np.int64)
    mat = np.array(material, dtype=np.float64)
    cap = np.ascontiguousarray(capacity, dtype=np.float64)
    flux = np.zeros_like(mat)
    lib().wfo_accucapacityflux(mat.size, order.ctypes.data, down.ctypes.data, mat.ctypes.data_as(_D),
                               cap.ctypes.data_as(_D), flux.ctypes.data_as(_D), dt)
    return flux, mat


def quantile(v, p):
    """Statistics.quantile!(v, p) as the reference's adaptive time step evaluates it."""
    v = np.array(v, dtype=np.float64)
    assert v.size >= 2
    return lib().wfo_quantile(v.ctypes.data_as(_D), v.size, p)


def kw_stable_timestep(q, alpha, flow_length, p):
    """stable_timestep of a kinematic-wave river / overland model (surface_kinwave.jl:674-704)."""
    q, alpha, fl = (np.ascontiguousarray(x, dtype=np.float64) for x in (q, alpha, flow_length))
    buf = np.empty(max(q.size, 1), np.float64)
    return lib().wfo_kw_stable_timestep(q.ctypes.data_as(_D), alpha.ctypes.data_as(_D), fl.ctypes.data_as(_D),
                                        q.size, p, buf.ctypes.data_as(_D))


def _out(k):
    return (ctypes.c_double * k)()


# ---- process functions (one call = one reference function call) --------------------------

def kinematic_wave(q_in, q_prev, q_lat, alpha, dt, dx):
    o = _out(2)
    lib().wfo_kinematic_wave(_c(q_in), _c(q_prev), _c(q_lat), _c(alpha), _c(dt), _c(dx), o)
    return o[0], o[1]


def ssf_celerity(zi, slope, sy, kh0, f, z_exp=float("nan")):
    return lib().wfo_ssf_celerity(zi, slope, sy, kh0, f, z_exp)


def kw_ssf_newton_raphson(q, C, cel, dt, dx):
    return lib().wfo_kw_ssf_newton_raphson(q, C, cel, dt, dx)


def kinematic_wave_ssf(q_in, q_prev, zi_prev, q_net, slope, sy, d, dt, dx, dw, qmax, kh0, f, soil, z_exp=float("nan")):
    o = _out(4)
    lib().wfo_kinematic_wave_ssf(*(_c(x) for x in (q_in, q_prev, zi_prev, q_net, slope, sy, d, dt, dx, dw, qmax,
                                                   kh0, f, z_exp)), ctypes.byref(soil), o)
    return tuple(o)


def rainfall_interception_gash(smax, er, gap, precip, storage, max_evap, dt):
    o = _out(4)
    lib().wfo_rainfall_interception_gash(*(_c(x) for x in (smax, er, gap, precip, storage, max_evap, dt)), o)
    return tuple(o)


def rainfall_interception_modrut(precip, pot_evap, storage, gap, smax, dt):
    o = _out(4)
    lib().wfo_rainfall_interception_modrut(*(_c(x) for x in (precip, pot_evap, storage, gap, smax, dt)), o)
    return tuple(o)


def precipitation_hbv(precip, temp, tti, tt):
    o = _out(2)
    lib().wfo_precipitation_hbv(*(_c(x) for x in (precip, temp, tti, tt)), o)
    return tuple(o)


def snowpack_hbv(ss, sw, sp, lp, temp, ttm, cfmax, whc, dt):
    o = _out(5)
    lib().wfo_snowpack_hbv(*(_c(x) for x in (ss, sw, sp, lp, temp, ttm, cfmax, whc, dt)), o)
    return tuple(o)


def infiltration(pot, compacted, cap_soil, cap_compacted, ustorecap, fred, dt):
    o = _out(2)
    lib().wfo_infiltration(*(_c(x) for x in (pot, compacted, cap_soil, cap_compacted, ustorecap, fred, dt)), o)
    return tuple(o)


def unsatzone_flow_layer(usd, kv_z, l_sat, c, dt):
    o = _out(2)
    lib().wfo_unsatzone_flow_layer(*(_c(x) for x in (usd, kv_z, l_sat, c, dt)), o)
    return tuple(o)


def actual_infiltration_soil_path(pot, act, compacted, cap_soil, cap_compacted, fred):
    o = _out(2)
    lib().wfo_actual_infiltration_soil_path(*(_c(x) for x in (pot, act, compacted, cap_soil, cap_compacted, fred)), o)
    return tuple(o)


def water_table_change(soil, net_flux, sy, dt):
    o = _out(2)
    lib().wfo_water_table_change(ctypes.byref(soil), _c(net_flux), _c(sy), _c(dt), o)
    return tuple(o)


def compute_mass_balance_error(tin, tout, rate):
    o = _out(2)
    lib().wfo_compute_mass_balance_error(_c(tin), _c(tout), _c(rate), o)
    return tuple(o)


def soil_call(name, soil, *args):
    """Run one soil-model update (e.g. 'soil_transpiration') on a single SoilCell."""
    getattr(lib(), "wfo_" + name)(ctypes.byref(soil), *(_c(a) for a in args))
    return soil


def toposort_dfs(n, out_lists):
    """Graphs.topological_sort_by_dfs on out-neighbour lists (0-based ids)."""
    ptr = np.zeros(n + 1, np.int64)
    idx = []
    for v in range(n):
        lst = sorted(out_lists[v])
        idx.extend(lst)
        ptr[v + 1] = len(idx)
    idx = np.array(idx if idx else [0], np.int64)
    order = np.zeros(n, np.int64)
    rc = lib().wfo_toposort_dfs(n, ptr.ctypes.data, idx.ctypes.data, order.ctypes.data)
    if rc != 0:
        raise ValueError("cycle")
    return order


# ---- whole model ---------------------------------------------------------------------------

class OracleModel:
    """update_model! on the CPU for a synthetic Catchment (wflowhip.synth)."""

    def __init__(self, catchment, water_mass_balance=True, **cfg_kw):
        self.c = catchment
        self.cfg = catchment.config(water_mass_balance=water_mass_balance, **cfg_kw)
        self._static = catchment.static()
        err = ctypes.create_string_buffer(256)
        self.h = lib().wfo_create(ctypes.byref(self.cfg), ctypes.byref(self._static), err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.n = lib().wfo_n(self.h)
        self.nriv = lib().wfo_nriv(self.h)
        self.nres = getattr(catchment, "nres", 0)
        self.nlayers = lib().wfo_nlayers(self.h)

    def close(self):
        if self.h:
            lib().wfo_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, precip, pet, temp):
        P = [np.ascontiguousarray(x, dtype=np.float64) for x in (precip, pet, temp)]
        rc = lib().wfo_step(self.h, *(x.ctypes.data_as(_D) for x in P))
        if rc != 0:
            raise RuntimeError(f"wfo_step failed {rc}")

    def get(self, var):
        cnt = _abi.var_count(var, self.n, self.nriv, self.nlayers, self.nres)
        out = np.empty(cnt, np.float64)
        rc = lib().wfo_get(self.h, int(var), out.ctypes.data_as(_D), cnt)
        if rc != 0:
            raise KeyError(var)
        return out

    def set(self, var, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        rc = lib().wfo_set(self.h, int(var), v.ctypes.data_as(_D), v.size)
        if rc != 0:
            raise KeyError(var)

    def topology(self):
        lo = np.empty(self.n, np.int64)
        ld = np.empty(self.n, np.int64)
        ro = np.empty(max(self.nriv, 1), np.int64)
        rd = np.empty(max(self.nriv, 1), np.int64)
        rl = np.empty(max(self.nriv, 1), np.int64)
        lib().wfo_topology(self.h, lo.ctypes.data, ld.ctypes.data, ro.ctypes.data, rd.ctypes.data, rl.ctypes.data)
        return lo, ld, ro[: self.nriv], rd[: self.nriv], rl[: self.nriv]

    def kinwave_iterations(self):
        return lib().wfo_kinwave_iterations(self.h)

    def set_julian_day(self, day):
        """update_index_hq!  reservoir.jl:379-386."""
        if lib().wfo_set_julian_day(self.h, int(day)) != 0:
            raise ValueError(f"julian day {day} outside 1..365")

    def set_mutation(self, mutation):
        """Plant a known mistake in the sweeps (tests of the network checks only):
        1 subsurface, 2 overland, 3 river fold the upstream discharge of the sub-step before."""
        if lib().wfo_set_mutation(self.h, int(mutation)) != 0:
            raise ValueError(f"unknown mutation {mutation}")

    def last_substeps(self):
        """(subsurface, overland, river) sub-step counts of the last step."""
        out = (ctypes.c_int32 * 3)()
        lib().wfo_last_substeps(self.h, out)
        return tuple(out)
