"""Synthetic catchments and forcing that stand in for the netCDF inputs of a Wflow `sbm` model.

Workload specification (all values converted to the reference's internal SI units):

* DEM: white noise shaped to a |k|^-2 power spectrum (amplitude |k|^-1, zero mean), scaled
  affinely to 0-1500 m, depression-filled by Priority-Flood+epsilon, D8 steepest descent
  (PCRaster LDD codes, first maximum in code order 1,2,3,4,6,7,8,9 wins, no positive drop ->
  pit). All raster cells are active; cells are enumerated with x fastest.
* River cells: upstream area (cells, itself included) > threshold. River width 2-50 m, linear
  in log(upstream area); bankfull depth 0.27 * width^0.39 m; river length = D8 flow length;
  river slope = land slope. Manning n 0.036 (river), 0.07 (land).
* Land slope: gradient magnitude of the unfilled DEM (central differences).
* Soil: KsatVer lognormal (median 250 mm/d, sigma 0.8), theta_s U[0.35, 0.5],
  theta_r U[0.03, 0.1], f U[1e-4, 1e-3] mm^-1, soil thickness 2200 mm with layers
  100/300/800 mm (+ remainder), Brooks-Corey c 9.43/9.82/10.24/10.24, KsatHorFrac 100,
  reference defaults elsewhere (air entry -10 cm, Feddes h1..h4 = 0/-100/-400,-1000/-16000 cm,
  alpha_h1 1, cap_hmax 2000 mm, cap_n 2, rootdistpar -500 mm^-1, w_soil 0.1125,
  cf_soil 0.038, infiltration capacity of compacted soil 10 mm/d, no leakage),
  compacted fraction 0.01, field capacity derived (Brooks-Corey at h = -1 m).
* Vegetation: canopy gap fraction U[0.2, 0.8], canopy storage U[0.5, 2.5] mm, rooting depth
  U[300, 1500] mm, crop factor 1, Gash E/R ratio 0.11.
* Snow (HBV): degree-day 3.75 mm/degC/d, thresholds 0 degC, interval 1 degC, whc 0.1.
* Forcing per step: precipitation gamma(shape 0.4, mean 3 mm/d) with a 40 % dry mask;
  temperature 10 + 12 sin(2 pi (day - 105) / 365) + N(0, 3) - 6.5 degC/km * elevation,
  hourly steps add 4 sin(2 pi (hour - 9) / 24); PET = clip(3 + 3 sin(seasonal) +
  0.1 (T - 10), 0, 6) mm/d, hourly steps multiply by a half-sine daylight shape of mean 1.
"""

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi

_LIB = None
MM = 1e-3
MM_PER_DAY = (1.0 * (1.0 / 86400.0)) * 1e-3


def _synth_lib():
    global _LIB
    if _LIB is None:
        here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        path = os.path.join(here, "lib", "libwfh_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.synth_fill_depressions.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        lib.synth_d8.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p]
        lib.synth_accumulate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p]
        lib.synth_accumulate.restype = ctypes.c_int64
        _LIB = lib
    return _LIB


def fractal_dem(nx, ny, seed, zmax=1500.0):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny, nx))
    F = np.fft.rfft2(w)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    F = F / k
    F[0, 0] = 0.0
    z = np.fft.irfft2(F, s=(ny, nx))
    z = (z - z.min()) / (z.max() - z.min()) * zmax
    return np.ascontiguousarray(z)  # z[y, x]


@dataclass
class Catchment:
    nx: int
    ny: int
    cell: float
    dt: float
    arrays: dict = field(default_factory=dict)  # static maps in SI, land order
    nlayers_cfg: int = 3
    layers_mm: tuple = (100.0, 300.0, 800.0)
    snow: bool = True
    infiltration_reduction: bool = True
    max_depth: int = 0

    @property
    def n(self):
        return int(self.arrays["cell_index"].shape[0])

    @property
    def nriv(self):
        return int(self.arrays["river_slope"].shape[0])

    @property
    def nlayers(self):
        return self.nlayers_cfg + 1

    @property
    def nres(self):
        a = self.arrays.get("reservoir_outlet")
        return 0 if a is None else int(a.shape[0])

    def config(self, water_mass_balance=True, device=0, land_dt=3600.0, river_dt=900.0, subsurface_dt=86400.0,
               adaptive=False, subsurface_alpha=1.0, snow_transport=False, floodplain=None):
        """floodplain=None: on when the catchment has floodplain profiles (add_floodplain)."""
        c = _abi.WfhConfig()
        c.abi_version = _abi.WFH_ABI_VERSION
        c.device = device
        c.dt = self.dt
        c.land_dt = land_dt
        c.river_dt = river_dt
        c.subsurface_dt = subsurface_dt
        c.n_config_layers = self.nlayers_cfg
        c.snow = int(self.snow)
        for k, v in enumerate(self.layers_mm):
            c.soil_layer_thickness_mm[k] = v
        c.soil_infiltration_reduction = int(self.infiltration_reduction)
        c.water_mass_balance = int(water_mass_balance)
        c.kinematic_wave_adaptive = int(adaptive)
        c.subsurface_alpha_coefficient = subsurface_alpha
        c.snow_gravitational_transport = int(snow_transport)
        c.floodplain_1d = int("floodplain_volume" in self.arrays if floodplain is None else floodplain)
        return c

    def static(self):
        """wfh_static view of the arrays (the arrays stay owned by this object)."""
        s = _abi.WfhStatic()
        s.nx, s.ny, s.n, s.nriv = self.nx, self.ny, self.n, self.nriv
        s.nres = self.nres
        fd = self.arrays.get("floodplain_depth")
        s.floodplain_ndepth = 0 if fd is None else int(fd.size)
        for name, ctype in _abi.STATIC_FIELDS[4:]:
            if name in ("nres", "floodplain_ndepth"):
                continue
            a = self.arrays.get(name)
            if a is None:
                setattr(s, name, ctype())
            else:
                assert a.flags["C_CONTIGUOUS"], name
                setattr(s, name, a.ctypes.data_as(ctype))
        return s

    def subset(self, mask):
        """The same maps restricted to the cells where mask (land order) is true; used to run
        complete basins on their own (basins exchange no water). Reservoirs whose outlet is kept
        come along (their area restricted to the kept cells), and so do the floodplain profiles
        of the kept river cells."""
        mask = np.asarray(mask, bool)
        out = Catchment(self.nx, self.ny, self.cell, self.dt, {}, self.nlayers_cfg, self.layers_mm, self.snow,
                        self.infiltration_reduction)
        riv = self.arrays["river_location"].astype(bool)
        rmask = mask[riv]
        n = self.n
        A = self.arrays
        if "floodplain_volume" in A:
            nd = A["floodplain_depth"].size
            out.arrays["floodplain_depth"] = A["floodplain_depth"].copy()
            out.arrays["floodplain_volume"] = np.ascontiguousarray(A["floodplain_volume"].reshape(nd, -1)[:, rmask])
            for k in ("floodplain_manning_n", "floodplain_slope"):
                out.arrays[k] = np.ascontiguousarray(A[k][rmask])
        if "reservoir_outlet" in A:
            self._subset_reservoirs(out, mask)
        for k, a in self.arrays.items():
            if k.startswith("reservoir_") or k.startswith("floodplain_"):
                continue
            if k.startswith("river_") and k not in ("river_location", "river_width_land", "river_length_land"):
                out.arrays[k] = np.ascontiguousarray(a[rmask])
            elif a.ndim == 1 and a.shape[0] == n:
                out.arrays[k] = np.ascontiguousarray(a[mask])
            elif a.ndim == 1 and a.shape[0] % n == 0:  # layer-major [N][n]
                out.arrays[k] = np.ascontiguousarray(a.reshape(-1, n)[:, mask].ravel())
            else:
                raise ValueError(k)
        return out

    def _subset_reservoirs(self, out, mask):
        A, B = self.arrays, out.arrays
        outlet = A["reservoir_outlet"]
        nres = outlet.size
        keep = np.nonzero(mask[outlet])[0]
        new_index = np.cumsum(mask) - 1
        B["reservoir_outlet"] = np.ascontiguousarray(new_index[outlet[keep]].astype(np.int64))
        remap = np.zeros(nres + 1, np.int32)
        remap[keep + 1] = np.arange(1, keep.size + 1, dtype=np.int32)
        B["reservoir_coverage"] = np.ascontiguousarray(remap[A["reservoir_coverage"][mask]])
        if "reservoir_lower" in A:  # 1-based links; a link to a dropped reservoir goes
            B["reservoir_lower"] = np.ascontiguousarray(remap[A["reservoir_lower"][keep]])
        for k, a in A.items():
            if k.startswith("reservoir_") and k not in ("reservoir_outlet", "reservoir_coverage", "reservoir_lower") and \
                    not k.startswith(("reservoir_sh_", "reservoir_hq_")):
                assert a.shape[0] == nres, k
                B[k] = np.ascontiguousarray(a[keep])
        for tab, cols in (("sh", ("h", "s")), ("hq", ("h", "q"))):
            ptr = A.get(f"reservoir_{tab}_ptr")
            if ptr is None:
                continue
            width = {"h": 1, "s": 1, "q": 365}
            new_ptr, parts = [0], {c: [] for c in cols}
            for r in keep:
                a, b = int(ptr[r]), int(ptr[r + 1])
                for c in cols:
                    w = width[c]
                    parts[c].append(A[f"reservoir_{tab}_{c}"][a * w:b * w])
                new_ptr.append(new_ptr[-1] + b - a)
            B[f"reservoir_{tab}_ptr"] = np.array(new_ptr, np.int64)
            for c in cols:
                v = np.concatenate(parts[c]) if parts[c] else np.zeros(0)
                B[f"reservoir_{tab}_{c}"] = np.ascontiguousarray(v if v.size else np.zeros(width[c]))


def make_catchment(nx, ny=None, cell=1000.0, dt=86400.0, river_threshold=100, seed=0, snow=True,
                   infiltration_reduction=True):
    ny = nx if ny is None else ny
    lib = _synth_lib()
    n = nx * ny
    z = fractal_dem(nx, ny, seed)
    zf = np.empty_like(z)
    lib.synth_fill_depressions(z.ctypes.data, nx, ny, zf.ctypes.data)
    ldd = np.empty(n, np.uint8)
    lib.synth_d8(zf.ctypes.data, nx, ny, cell, ldd.ctypes.data)
    area = np.empty(n, np.int64)
    depth = np.empty(n, np.int32)
    maxdepth = lib.synth_accumulate(ldd.ctypes.data, nx, ny, area.ctypes.data, depth.ctypes.data)
    if maxdepth < 0:
        raise RuntimeError("invalid synthetic LDD")
    gy, gx = np.gradient(z, cell)
    slope = np.sqrt(gx * gx + gy * gy).ravel()
    rng = np.random.default_rng(seed + 7919)
    A = {}
    A["cell_index"] = np.arange(n, dtype=np.int64)
    A["ldd"] = ldd
    A["x_length"] = np.full(n, float(cell))
    A["y_length"] = np.full(n, float(cell))
    A["land_slope"] = np.ascontiguousarray(slope)
    A["land_manning_n"] = np.full(n, 0.07)
    riv = area > river_threshold
    A["river_location"] = riv.astype(np.uint8)
    lo, hi = math.log(river_threshold), math.log(max(area.max(), river_threshold + 1))
    width = np.clip(2.0 + 48.0 * (np.log(np.maximum(area, 1)) - lo) / (hi - lo), 2.0, 50.0)
    diag = np.isin(ldd, (1, 3, 7, 9))
    flen = np.where(diag, cell * math.sqrt(2.0), float(cell))
    A["river_width_land"] = np.where(riv, width, 0.0)
    A["river_length_land"] = np.where(riv, flen, 0.0)
    A["water_fraction"] = np.zeros(n)
    A["elevation"] = np.ascontiguousarray(zf.ravel())
    A["canopy_gap_fraction"] = rng.uniform(0.2, 0.8, n)
    A["max_canopy_storage"] = rng.uniform(0.5, 2.5, n) * MM
    A["rooting_depth"] = rng.uniform(300.0, 1500.0, n) * MM
    A["crop_coefficient"] = np.ones(n)
    A["evap_precip_ratio"] = np.full(n, 0.11)
    A["cfmax"] = np.full(n, 3.75 * MM_PER_DAY)
    A["tt"] = np.full(n, 273.15)
    A["tti"] = np.full(n, 1.0)
    A["ttm"] = np.full(n, 273.15)
    A["whc"] = np.full(n, 0.1)
    A["theta_s"] = rng.uniform(0.35, 0.5, n)
    A["theta_r"] = rng.uniform(0.03, 0.1, n)
    A["kv_0"] = 250.0 * np.exp(0.8 * rng.standard_normal(n)) * MM_PER_DAY
    A["f"] = rng.uniform(1e-4, 1e-3, n) / MM
    A["air_entry_pressure"] = np.full(n, -10.0 * 1e-2)
    A["h1"] = np.full(n, 0.0)
    A["h2"] = np.full(n, -100.0 * 1e-2)
    A["h3_high"] = np.full(n, -400.0 * 1e-2)
    A["h3_low"] = np.full(n, -1000.0 * 1e-2)
    A["h4"] = np.full(n, -16000.0 * 1e-2)
    A["alpha_h1"] = np.full(n, 1.0)
    A["soil_thickness"] = np.full(n, 2200.0 * MM)
    A["infiltcap_compacted"] = np.full(n, 10.0 * MM_PER_DAY)
    A["max_leakage"] = np.zeros(n)
    A["compacted_fraction"] = np.full(n, 0.01)
    A["rootdistpar"] = np.full(n, -500.0 / MM)
    A["cap_hmax"] = np.full(n, 2000.0 * MM)
    A["cap_n"] = np.full(n, 2.0)
    A["w_soil"] = np.full(n, 0.1125)
    A["cf_soil"] = np.full(n, 0.038)
    A["brooks_corey_c"] = np.ascontiguousarray(np.repeat([9.43, 9.82, 10.24, 10.24], n).astype(np.float64))
    A["kvfrac"] = np.ones(4 * n)
    A["khfrac"] = np.full(n, 100.0)
    A["river_slope"] = np.ascontiguousarray(slope[riv])
    A["river_manning_n"] = np.full(int(riv.sum()), 0.036)
    A["river_bankfull_depth"] = np.ascontiguousarray(0.27 * width[riv] ** 0.39)
    A["river_length"] = np.ascontiguousarray(flen[riv])
    A["river_width"] = np.ascontiguousarray(width[riv])
    c = Catchment(nx, ny, float(cell), float(dt), A, snow=snow, infiltration_reduction=infiltration_reduction)
    c.max_depth = int(maxdepth)
    c.upstream_area = area
    c.depth = depth  # topological level of every cell (longest upstream path)
    return c


def make_batch(members):
    """Several catchments as one model: the members are stacked in a virtual raster (width =
    the widest member, one empty row between members) so that no D8 direction can reach
    another member; every member keeps its own cells, parameters and local flow network.
    Cells stay member by member in each member's own order, so member m occupies cells
    [cell_offsets[m], cell_offsets[m + 1]) and river cells [river_offsets[m],
    river_offsets[m + 1]) of every per-cell or per-river-cell array. A member's results are
    those of the member run on its own: basins exchange no water and the step's result does
    not depend on the order in which basins are visited.

    All members must share the time step, the soil-layer configuration and the snow and
    infiltration-reduction switches (one model configuration)."""
    m0 = members[0]
    for c in members[1:]:
        if (c.dt, c.nlayers_cfg, tuple(c.layers_mm), c.snow, c.infiltration_reduction) != \
           (m0.dt, m0.nlayers_cfg, tuple(m0.layers_mm), m0.snow, m0.infiltration_reduction):
            raise ValueError("batch members must share one model configuration")
    width = max(c.nx for c in members)
    names = list(m0.arrays.keys())
    parts = {k: [] for k in names}
    cell_off, riv_off, y0 = [0], [0], 0
    for c in members:
        n = c.n
        if set(c.arrays.keys()) != set(names):
            raise ValueError("batch members must provide the same maps")
        for k in names:
            a = c.arrays[k]
            if k == "cell_index":
                x = a % c.nx
                y = a // c.nx
                parts[k].append(((y + y0) * width + x).astype(np.int64))
            elif k.startswith("river_") and k not in ("river_location", "river_width_land", "river_length_land"):
                parts[k].append(a)
            elif a.ndim == 1 and a.shape[0] == n:
                parts[k].append(a)
            elif a.ndim == 1 and a.shape[0] % n == 0:  # layer-major [N][n]
                parts[k].append(a.reshape(-1, n))
            else:
                raise ValueError(k)
        y0 += c.ny + 1
        cell_off.append(cell_off[-1] + n)
        riv_off.append(riv_off[-1] + c.nriv)
    A = {}
    for k in names:
        if parts[k][0].ndim == 2:
            A[k] = np.ascontiguousarray(np.concatenate(parts[k], axis=1).ravel())
        else:
            A[k] = np.ascontiguousarray(np.concatenate(parts[k]))
    b = Catchment(width, y0 - 1, m0.cell, m0.dt, A, m0.nlayers_cfg, m0.layers_mm, m0.snow, m0.infiltration_reduction)
    b.max_depth = max(c.max_depth for c in members)
    b.cell_offsets = np.array(cell_off, np.int64)
    b.river_offsets = np.array(riv_off, np.int64)
    if all(hasattr(c, "depth") for c in members):
        b.depth = np.concatenate([c.depth for c in members])
    return b


def forcing_batch_numpy(members, step, seed=0):
    """Forcing of a batch: each member's own forcing_numpy, concatenated in member order."""
    fs = [forcing_numpy(c, step, seed=seed + 7 * m) for m, c in enumerate(members)]
    return tuple(np.ascontiguousarray(np.concatenate([f[i] for f in fs])) for i in range(3))


def basin_ids(ldd, nx, ny, cell_index=None):
    """Outlet (pit) id of every cell: cells with the same id form one closed basin. Cells sit at
    raster positions cell_index (default: every cell of the nx x ny raster, in order); the
    generator's directions never leave the active cells."""
    n = ldd.shape[0]
    dxs = np.array([0, -1, 0, 1, -1, 0, 1, -1, 0, 1])
    dys = np.array([0, -1, -1, -1, 0, 0, 0, 1, 1, 1])
    idx = np.arange(n)
    pos = idx if cell_index is None else np.asarray(cell_index, np.int64)
    tgt = (pos % nx + dxs[ldd]) + (pos // nx + dys[ldd]) * nx
    if cell_index is not None:
        tgt = np.searchsorted(pos, tgt).clip(0, n - 1)
    down = np.where(ldd == 5, idx, tgt)
    lab = down.copy()
    for _ in range(64):  # pointer jumping: log2(path length) rounds
        nxt = lab[lab]
        if np.array_equal(nxt, lab):
            break
        lab = nxt
    return lab


def land_downstream(catchment):
    """Land index of each cell's D8 downstream cell (-1 for pits)."""
    c = catchment
    ldd = c.arrays["ldd"]
    n = c.n
    pos = c.arrays["cell_index"]
    dxs = np.array([0, -1, 0, 1, -1, 0, 1, -1, 0, 1])
    dys = np.array([0, -1, -1, -1, 0, 0, 0, 1, 1, 1])
    tgt = (pos % c.nx + dxs[ldd]) + (pos // c.nx + dys[ldd]) * c.nx
    k = np.searchsorted(pos, tgt).clip(0, n - 1)
    ok = (ldd != 5) & (pos[k] == tgt)
    return np.where(ok, k, -1)


def add_reservoirs(catchment, count, seed=0, radius=1, curves=False, linked=False):
    """Place `count` reservoirs on river cells (in place): outlets on river cells whose river
    continues downstream, at most one reservoir feeding any river cell, each with the cells
    within `radius` raster steps of its outlet as its area (cells of an earlier reservoir
    excluded). Types alternate simple (4) and modified Puls (3); some abstract water (negative
    external inflow) and one uses observed outflow. Sizes are scaled to the cell size.

    curves=True cycles the outflow types 1 (rating curve), 2 (free weir), 3 and 4, gives every
    third reservoir (k % 3 == 1) an interpolated storage curve (S-H table), and gives the
    rating-curve reservoirs H-Q tables with a seasonal cycle over the 365 day columns; reservoir 4
    starts above the top of its H-Q table, so it overflows. linked=True (with curves) links every
    free weir to the first unlinked reservoir downstream of it on the river, if any, and starts
    that lower reservoir above the free weir's level, so the flow reverses."""
    c = catchment
    rng = np.random.default_rng(seed + 4242)
    down = land_downstream(c)
    riv = c.arrays["river_location"].astype(bool)
    cand = [i for i in np.nonzero(riv)[0] if down[i] >= 0 and riv[down[i]]]
    rng.shuffle(cand)
    outlets, fed = [], set()
    for i in cand:
        if len(outlets) == count:
            break
        if down[i] in fed or i in fed or any(down[o] == i for o in outlets):
            continue
        outlets.append(int(i))
        fed.add(int(down[i]))
    if len(outlets) < count:
        raise ValueError("not enough river cells for the reservoirs")
    outlets = np.array(sorted(outlets), np.int64)
    n = c.n
    pos = c.arrays["cell_index"]
    x, y = pos % c.nx, pos // c.nx
    cov = np.zeros(n, np.int32)
    for k, o in enumerate(outlets):
        near = (np.abs(x - x[o]) <= radius) & (np.abs(y - y[o]) <= radius) & (cov == 0)
        near[o] = True
        cov[near] = k + 1
    m = outlets.size
    cell_area = c.cell * c.cell
    A = c.arrays
    A["reservoir_outlet"] = outlets
    A["reservoir_coverage"] = cov
    A["reservoir_outflow_type"] = np.where(np.arange(m) % 2 == 0, 4, 3).astype(np.int32)
    ncov = np.bincount(cov, minlength=m + 1)[1:].astype(float)
    area = ncov * cell_area * rng.uniform(0.6, 1.0, m)
    A["reservoir_area"] = area
    A["reservoir_waterlevel"] = rng.uniform(2.0, 6.0, m)
    maxsto = area * rng.uniform(6.0, 9.0, m)
    A["reservoir_maximum_storage"] = maxsto
    A["reservoir_demand"] = maxsto * rng.uniform(2e-8, 8e-8, m)          # m3/s
    A["reservoir_maximum_release"] = maxsto * rng.uniform(2e-7, 6e-7, m)
    A["reservoir_target_full_fraction"] = rng.uniform(0.6, 0.9, m)
    A["reservoir_target_minimum_fraction"] = rng.uniform(0.2, 0.4, m)
    A["reservoir_threshold"] = rng.uniform(0.0, 3.0, m)
    A["reservoir_rating_curve_coefficient"] = rng.uniform(0.05, 0.5, m)
    ext = np.zeros(m)
    ext[1::3] = -maxsto[1::3] * 1e-8
    A["reservoir_external_inflow"] = ext
    obs = np.full(m, np.nan)
    if m >= 3:
        obs[2] = maxsto[2] * 3e-8
    A["reservoir_outflow_obs"] = obs
    if curves:
        _add_reservoir_curves(c, rng)
        if linked:
            _link_free_weirs(c)
    return c


def _link_free_weirs(c):
    A = c.arrays
    down = land_downstream(c)
    outlets = A["reservoir_outlet"]
    at = {int(o): k for k, o in enumerate(outlets)}
    lower = np.zeros(outlets.size, np.int32)
    taken = set()
    wl, thr = A["reservoir_waterlevel"], A["reservoir_threshold"]
    for k in np.nonzero(A["reservoir_outflow_type"] == 2)[0]:
        i = int(down[outlets[k]])
        while i >= 0:
            j = at.get(i)
            if j is not None and j not in taken and j != k:
                lower[k] = j + 1
                taken.add(j)
                wl[j] = max(wl[j], wl[k] + 0.5, thr[k] + 1.0)
                break
            i = int(down[i])
    A["reservoir_lower"] = lower


def add_floodplain(catchment, seed=0, depths=(0.5, 1.0, 2.0, 3.5, 5.0), bankfull_scale=0.02):
    """Floodplain profiles for every river cell (in place): cumulative flood volumes at `depths`
    of a floodplain widening with depth (5 to 60 times the river width), with a narrowing step
    in every seventh cell (which the profile construction widens back); floodplain Manning n and
    slope. The bankfull depths are scaled by `bankfull_scale` so that the synthetic rivers leave
    their channels."""
    c = catchment
    A = c.arrays
    rng = np.random.default_rng(seed + 777)
    nriv = int(A["river_width"].size)
    d = np.asarray(depths, np.float64)
    h = np.diff(np.concatenate([[0.0], d]))
    fw = A["river_width"]
    fl = A["river_length"]
    base = fw[None, :] * rng.uniform(5.0, 20.0, nriv)[None, :]
    grow = np.cumprod(rng.uniform(1.0, 1.6, (d.size, nriv)), axis=0)
    width = base * grow
    width[2, ::7] *= 0.7  # a narrowing segment
    vol = np.cumsum(width * h[:, None] * fl[None, :], axis=0)
    A["floodplain_depth"] = d
    A["floodplain_volume"] = np.ascontiguousarray(vol)
    A["floodplain_manning_n"] = rng.uniform(0.05, 0.1, nriv)
    slope = A["river_slope"].copy()
    slope[::11] = 0.0  # clamped to 1e-5
    A["floodplain_slope"] = slope
    A["river_bankfull_depth"] = np.ascontiguousarray(A["river_bankfull_depth"] * bankfull_scale)
    return c


def _add_reservoir_curves(c, rng):
    A = c.arrays
    m = A["reservoir_outlet"].size
    k = np.arange(m)
    types = np.array([1, 2, 3, 4], np.int32)[k % 4]
    sct = np.where(k % 3 == 1, 2, 1).astype(np.int32)
    area = A["reservoir_area"]
    wl = A["reservoir_waterlevel"]
    A["reservoir_outflow_type"] = types
    A["reservoir_storage_curve_type"] = sct
    A["reservoir_rating_curve_exponent"] = rng.uniform(1.4, 1.6, m)
    sh_ptr, sh_h, sh_s = [0], [], []
    hq_ptr, hq_h, hq_q = [0], [], []
    day = np.arange(1, 366)
    for r in range(m):
        if sct[r] == 2:  # convex storage curve on a ragged water-level grid, 0 .. 12 m
            rows = 10 + r % 7
            h = np.sort(np.concatenate([[0.0, 12.0], rng.uniform(0.2, 11.8, rows - 2)]))
            sh_h += list(h)
            sh_s += list(area[r] * h * (1.0 + 0.05 * h))
        sh_ptr.append(len(sh_h))
        if types[r] == 1:  # Q = a (H - H_spill)^1.5, seasonal factor per day column
            hmax = 4.5 if r == 4 else 12.0
            if r == 4:
                wl[r] = 5.0
            rows = 8 + r % 5
            h = np.linspace(0.0, hmax, rows)
            spill = rng.uniform(0.5, 1.5)
            a = area[r] * rng.uniform(2e-7, 6e-7)
            q = a * np.maximum(h - spill, 0.0)[:, None] ** 1.5 * (1.0 + 0.3 * np.sin(2 * np.pi * day / 365.0))[None, :]
            hq_h += list(h)
            hq_q += list(q.ravel())
        hq_ptr.append(len(hq_h))
    A["reservoir_sh_ptr"] = np.array(sh_ptr, np.int64)
    A["reservoir_sh_h"] = np.array(sh_h if sh_h else [0.0])
    A["reservoir_sh_s"] = np.array(sh_s if sh_s else [0.0])
    A["reservoir_hq_ptr"] = np.array(hq_ptr, np.int64)
    A["reservoir_hq_h"] = np.array(hq_h if hq_h else [0.0])
    A["reservoir_hq_q"] = np.array(hq_q if hq_q else [0.0] * 365)


# ------------------------------------------------------------------------------------------
# forcing
# ------------------------------------------------------------------------------------------

def _daylight_shape():
    h = np.arange(24) + 0.5
    s = np.maximum(0.0, np.sin(np.pi * (h - 6.0) / 12.0))
    return s / s.mean()


_DAYLIGHT = _daylight_shape()


def forcing_numpy(catchment, step, seed=0):
    """(precipitation [m s-1], PET [m s-1], temperature [K]) for model step `step` (0-based)."""
    n = catchment.n
    dt = catchment.dt
    rng = np.random.default_rng([seed, step, 12345])
    t_days = step * dt / 86400.0
    doy = t_days % 365.0
    hour = (step * dt / 3600.0) % 24.0
    season = math.sin(2.0 * math.pi * (doy - 105.0) / 365.0)
    wet = rng.random(n) >= 0.4
    p = rng.gamma(0.4, 3.0 / 0.4, n) * wet
    elev = catchment.arrays["elevation"]
    temp_c = 10.0 + 12.0 * season + rng.normal(0.0, 3.0, n) - 6.5e-3 * elev
    if dt < 86400.0:
        temp_c = temp_c + 4.0 * math.sin(2.0 * math.pi * (hour - 9.0) / 24.0)
    pet = np.clip(3.0 + 3.0 * season + 0.1 * (temp_c - 10.0), 0.0, 6.0)
    if dt < 86400.0:
        pet = pet * _DAYLIGHT[int(hour) % 24]
    return (np.ascontiguousarray(p * MM_PER_DAY), np.ascontiguousarray(pet * MM_PER_DAY),
            np.ascontiguousarray(temp_c + 273.15))


def forcing_torch(catchment, step, elevation_dev, seed=0, generator=None):
    """Device-side variant of forcing_numpy (same distributions, torch RNG on the GPU)."""
    import torch
    n = catchment.n
    dt = catchment.dt
    dev = elevation_dev.device
    g = generator
    if g is None:
        g = torch.Generator(device=dev)
        g.manual_seed((seed * 1000003 + step) & 0x7FFFFFFF)
    t_days = step * dt / 86400.0
    doy = t_days % 365.0
    hour = (step * dt / 3600.0) % 24.0
    season = math.sin(2.0 * math.pi * (doy - 105.0) / 365.0)
    wet = (torch.rand(n, generator=g, device=dev, dtype=torch.float64) >= 0.4).to(torch.float64)
    conc = torch.full((n,), 0.4, device=dev, dtype=torch.float64)
    p = torch._standard_gamma(conc, generator=g) * (3.0 / 0.4) * wet
    noise = torch.randn(n, generator=g, device=dev, dtype=torch.float64) * 3.0
    temp_c = 10.0 + 12.0 * season + noise - 6.5e-3 * elevation_dev
    if dt < 86400.0:
        temp_c = temp_c + 4.0 * math.sin(2.0 * math.pi * (hour - 9.0) / 24.0)
    pet = torch.clamp(3.0 + 3.0 * season + 0.1 * (temp_c - 10.0), 0.0, 6.0)
    if dt < 86400.0:
        pet = pet * float(_DAYLIGHT[int(hour) % 24])
    return (p * MM_PER_DAY).contiguous(), (pet * MM_PER_DAY).contiguous(), (temp_c + 273.15).contiguous()
