"""ctypes binding of libwflow_hip.so and the Model class that mirrors the reference step API."""

import ctypes
import os

import numpy as np

from . import _abi
from ._abi import Var

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path():
    return os.path.join(os.path.dirname(_HERE), "lib", "libwflow_hip.so")


def load_library():
    """Load the HIP library; raises when it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    try:
        # torch ships its own HIP runtime under the same soname; loading it first makes this
        # library bind to that single copy, so device buffers can be shared with torch.
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.wfh_create.argtypes = [ctypes.POINTER(_abi.WfhConfig), ctypes.POINTER(_abi.WfhStatic), ctypes.POINTER(P)]
    L.wfh_create.restype = i32
    L.wfh_set_forcing.argtypes = [P, P, P, P, i32]
    L.wfh_set_forcing.restype = i32
    L.wfh_step.argtypes = [P]
    L.wfh_step.restype = i32
    L.wfh_step_noalloc.argtypes = [P]
    L.wfh_step_noalloc.restype = i32
    L.wfh_set_graph.argtypes = [P, i32]
    L.wfh_set_graph.restype = i32
    L.wfh_autotune.argtypes = [P, i32]
    L.wfh_autotune.restype = i32
    L.wfh_get_schedule.argtypes = [P, P, P, P, P]
    L.wfh_get_schedule.restype = i32
    L.wfh_test_route_stall.argtypes = [P, i64, i64]
    L.wfh_test_route_stall.restype = i32
    L.wfh_run.argtypes = [P, i32, P, P, P, P]
    L.wfh_run.restype = i32
    L.wfh_get.argtypes = [P, i32, P, i64, i32]
    L.wfh_get.restype = i32
    L.wfh_set.argtypes = [P, i32, P, i64, i32]
    L.wfh_set.restype = i32
    L.wfh_get_topology.argtypes = [P, P, P, P, P, P, P]
    L.wfh_get_topology.restype = i32
    L.wfh_get_node_order.argtypes = [P, P]
    L.wfh_get_node_order.restype = i32
    L.wfh_state_size.argtypes = [P]
    L.wfh_state_size.restype = i64
    L.wfh_save_state.argtypes = [P, P, i64]
    L.wfh_save_state.restype = i32
    L.wfh_load_state.argtypes = [P, P, i64]
    L.wfh_load_state.restype = i32
    L.wfh_sync.argtypes = [P]
    L.wfh_sync.restype = i32
    L.wfh_set_timing.argtypes = [P, i32]
    L.wfh_set_timing.restype = i32
    L.wfh_get_timing.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.wfh_get_timing.restype = i32
    L.wfh_last_error.argtypes = [P]
    L.wfh_last_error.restype = ctypes.c_char_p
    L.wfh_destroy.argtypes = [P]
    L.wfh_destroy.restype = None
    L.wfh_kernel_bytes.argtypes = [P, P, P]
    L.wfh_kernel_bytes.restype = i32
    L.wfh_last_substeps.argtypes = [P, P, P, P]
    L.wfh_last_substeps.restype = i32
    L.wfh_set_julian_day.argtypes = [P, i32]
    L.wfh_set_julian_day.restype = i32
    _LIB = L
    return L


class WflowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wflow_hip error {code}: {msg}")
        self.code = code


class Model:
    """One Wflow `sbm` model on one GPU (Model(config) for the sbm concept, sbm_model.jl:7-57)."""

    def __init__(self, catchment, water_mass_balance=True, device=0, **cfg_kw):
        self.lib = load_library()
        self.catchment = catchment
        self.cfg = catchment.config(water_mass_balance=water_mass_balance, device=device, **cfg_kw)
        self._static = catchment.static()
        h = ctypes.c_void_p()
        rc = self.lib.wfh_create(ctypes.byref(self.cfg), ctypes.byref(self._static), ctypes.byref(h))
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_last_error(None).decode())
        self.h = h
        self.n = catchment.n
        self.nriv = catchment.nriv
        self.nres = catchment.nres
        self.nlayers = catchment.nlayers
        self._forcing_keep = None

    # ---- lifetime -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.wfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_last_error(self.h).decode())

    # ---- step ------------------------------------------------------------------------------
    def set_forcing(self, precipitation, potential_evaporation, temperature, memkind=None):
        """Forcing for the next step. numpy arrays -> host copy; torch CUDA tensors -> device
        pointers (land order, or node order with memkind=WFH_MEM_DEVICE_NODE_ORDER)."""
        arrs = (precipitation, potential_evaporation, temperature)
        if isinstance(precipitation, np.ndarray):
            arrs = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in arrs)
            ptrs = [a.ctypes.data for a in arrs]
            kind = _abi.WFH_MEM_HOST if memkind is None else memkind
        else:  # torch tensors on the GPU
            for a in arrs:
                assert a.is_cuda and a.dtype.itemsize == 8 and a.is_contiguous() and a.numel() == self.n
            ptrs = [a.data_ptr() for a in arrs]
            kind = _abi.WFH_MEM_DEVICE if memkind is None else memkind
        self._forcing_keep = arrs
        self._check(self.lib.wfh_set_forcing(self.h, *ptrs, kind))

    def update_model(self):
        """update_model! (asynchronous on the model's stream)."""
        self._check(self.lib.wfh_step(self.h))

    def update_model_noalloc(self):
        """update_model! through wfh_step_noalloc: the routing scratch comes from the library's
        fixed per-device array, so the step allocates nothing (asynchronous like update_model)."""
        self._check(self.lib.wfh_step_noalloc(self.h))

    def set_graph(self, enable=True):
        """Replay the step's launches from a HIP graph (fixed sub-steps)."""
        self._check(self.lib.wfh_set_graph(self.h, int(enable)))

    def autotune(self, trials=1):
        """Time the candidate routing grids on the next trials x 4 steps and keep the fastest."""
        self._check(self.lib.wfh_autotune(self.h, int(trials)))

    def schedule(self):
        v = [ctypes.c_int32() for _ in range(4)]
        self._check(self.lib.wfh_get_schedule(self.h, *(ctypes.addressof(x) for x in v)))
        return dict(zip(("route_blocks", "max_blocks", "tuning", "graph_mode"), (x.value for x in v)))

    def run(self, precipitation, potential_evaporation, temperature, river_q_avg=None):
        """`nsteps` steps in one launch (wfh_run): forcing frames as device tensors [nsteps, n] in
        node order (node_order()); river_q_avg, a device tensor [nsteps, nriv], receives each
        step's river discharge average. Equal to nsteps run_timestep calls bit for bit."""
        frames = [precipitation, potential_evaporation, temperature]
        nsteps = int(frames[0].shape[0])
        for f in frames:
            if not (f.is_cuda and f.is_contiguous() and tuple(f.shape) == (nsteps, self.n)):
                raise ValueError("forcing frames must be contiguous device tensors [nsteps, n] in node order")
        out = 0
        if river_q_avg is not None:
            if not (river_q_avg.is_cuda and river_q_avg.is_contiguous() and
                    tuple(river_q_avg.shape) == (nsteps, self.nriv)):
                raise ValueError("river_q_avg must be a contiguous device tensor [nsteps, nriv]")
            out = river_q_avg.data_ptr()
        self._keep_run = frames + ([river_q_avg] if river_q_avg is not None else [])
        self._check(self.lib.wfh_run(self.h, nsteps, *(f.data_ptr() for f in frames), out))

    def test_route_stall(self, spin_limit, node=-1):
        """Testing only: the routing kernel's spin limit and, for node >= 0, a wait of that cell
        (node order) that can never be met (wfh_test_route_stall)."""
        self._check(self.lib.wfh_test_route_stall(self.h, int(spin_limit), int(node)))

    def run_timestep(self, precipitation, potential_evaporation, temperature):
        """run_timestep! with the step's forcing supplied by the caller."""
        self.set_forcing(precipitation, potential_evaporation, temperature)
        self.update_model()

    def sync(self):
        self._check(self.lib.wfh_sync(self.h))

    # ---- BMI-style access --------------------------------------------------------------------
    def get_value(self, var):
        cnt = _abi.var_count(var, self.n, self.nriv, self.nlayers, self.nres)
        out = np.empty(cnt, np.float64)
        self._check(self.lib.wfh_get(self.h, int(var), out.ctypes.data, cnt, _abi.WFH_MEM_HOST))
        return out

    get = get_value

    def set_value(self, var, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.wfh_set(self.h, int(var), v.ctypes.data, v.size, _abi.WFH_MEM_HOST))

    set = set_value

    def get_value_device(self, var, out, node_order=False):
        """BMI get_value into a device tensor (land-index order, or node order), no host copy."""
        kind = _abi.WFH_MEM_DEVICE_NODE_ORDER if node_order else _abi.WFH_MEM_DEVICE
        cnt = _abi.var_count(var, self.n, self.nriv, self.nlayers, self.nres)
        assert out.is_cuda and out.dtype.itemsize == 8 and out.is_contiguous() and out.numel() == cnt
        self._check(self.lib.wfh_get(self.h, int(var), out.data_ptr(), cnt, kind))
        return out

    def set_value_device(self, var, src, node_order=False):
        """BMI set_value from a device tensor (the tensor must be ready: synchronise its stream)."""
        kind = _abi.WFH_MEM_DEVICE_NODE_ORDER if node_order else _abi.WFH_MEM_DEVICE
        assert src.is_cuda and src.dtype.itemsize == 8 and src.is_contiguous()
        self._check(self.lib.wfh_set(self.h, int(var), src.data_ptr(), src.numel(), kind))

    def topology(self):
        n, nr = self.n, self.nriv
        lo, ld = np.empty(n, np.int64), np.empty(n, np.int64)
        ro, rd, rl = (np.empty(max(nr, 1), np.int64) for _ in range(3))
        ml = ctypes.c_int32()
        self._check(self.lib.wfh_get_topology(self.h, lo.ctypes.data, ld.ctypes.data, ro.ctypes.data,
                                              rd.ctypes.data, rl.ctypes.data, ctypes.addressof(ml)))
        return lo, ld, ro[:nr], rd[:nr], rl[:nr], ml.value

    def node_order(self):
        out = np.empty(self.n, np.int64)
        self._check(self.lib.wfh_get_node_order(self.h, out.ctypes.data))
        return out

    def save_state(self):
        size = self.lib.wfh_state_size(self.h)
        buf = np.empty(size, np.uint8)
        self._check(self.lib.wfh_save_state(self.h, buf.ctypes.data, size))
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._check(self.lib.wfh_load_state(self.h, buf.ctypes.data, buf.size))

    def set_timing(self, enable=True):
        self._check(self.lib.wfh_set_timing(self.h, int(enable)))

    def kernel_bytes(self):
        """Algorithmic bytes of one launch of (vertical kernel, routing kernel)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.wfh_kernel_bytes(self.h, ctypes.addressof(a), ctypes.addressof(b)))
        return a.value, b.value

    def set_julian_day(self, day):
        """Rating-curve column (julian day 1..365) of the following steps (update_index_hq!
        Wflow/src/routing/surface/reservoir.jl:379-386)."""
        self._check(self.lib.wfh_set_julian_day(self.h, int(day)))

    def last_substeps(self):
        """(subsurface, overland, river) sub-step counts of the last step."""
        v = [ctypes.c_int32() for _ in range(3)]
        self._check(self.lib.wfh_last_substeps(self.h, *(ctypes.addressof(x) for x in v)))
        return tuple(x.value for x in v)

    def timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.wfh_get_timing(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def _bind_ensemble(L):
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.wfh_ens_create.argtypes = [ctypes.POINTER(_abi.WfhConfig), ctypes.POINTER(_abi.WfhStatic), i32, P, P, P,
                                 ctypes.POINTER(P)]
    L.wfh_ens_create.restype = i32
    for name, args in (("wfh_ens_set_forcing", [P, P, P, P, i32]), ("wfh_ens_step", [P]),
                       ("wfh_ens_get", [P, i32, i32, P, i64]), ("wfh_ens_balance", [P, P]),
                       ("wfh_ens_set_schedule", [P, i32, i32, i32]), ("wfh_ens_set_autotune", [P, i32]),
                       ("wfh_ens_get_schedule", [P, P, P, P, P, P]), ("wfh_ens_sync", [P]),
                       ("wfh_ens_set_timing", [P, i32]), ("wfh_ens_get_timing", [P, P]),
                       ("wfh_ens_step_bytes", [P, P])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = i32
    L.wfh_ens_last_error.argtypes = [P]
    L.wfh_ens_last_error.restype = ctypes.c_char_p
    L.wfh_ens_destroy.argtypes = [P]
    L.wfh_ens_destroy.restype = None


class Ensemble:
    """M members of one catchment whose kv_0, f and Manning n maps are scaled by a factor per
    member, stepped together (member-innermost layout, graph-replayed stages). A member's
    results are those of Model on the catchment with its scaled maps (member_catchment)."""

    def __init__(self, catchment, kv0_factor, f_factor, n_factor, water_mass_balance=True, device=0, **cfg_kw):
        self.lib = load_library()
        if not hasattr(self.lib, "_ens_bound"):
            _bind_ensemble(self.lib)
            self.lib._ens_bound = True
        self.catchment = catchment
        self.factors = [np.ascontiguousarray(a, dtype=np.float64) for a in (kv0_factor, f_factor, n_factor)]
        self.members = int(self.factors[0].size)
        self.cfg = catchment.config(water_mass_balance=water_mass_balance, device=device, **cfg_kw)
        self._static = catchment.static()
        h = ctypes.c_void_p()
        rc = self.lib.wfh_ens_create(ctypes.byref(self.cfg), ctypes.byref(self._static), self.members,
                                     *(a.ctypes.data for a in self.factors), ctypes.byref(h))
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_ens_last_error(None).decode())
        self.h = h
        self.n, self.nriv, self.nlayers = catchment.n, catchment.nriv, catchment.nlayers

    def close(self):
        if getattr(self, "h", None):
            self.lib.wfh_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_ens_last_error(self.h).decode())

    def set_forcing(self, precipitation, potential_evaporation, temperature, memkind=None):
        arrs = (precipitation, potential_evaporation, temperature)
        if isinstance(precipitation, np.ndarray):
            arrs = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in arrs)
            ptrs = [a.ctypes.data for a in arrs]
            kind = _abi.WFH_MEM_HOST if memkind is None else memkind
        else:
            ptrs = [a.data_ptr() for a in arrs]
            kind = _abi.WFH_MEM_DEVICE if memkind is None else memkind
        self._check(self.lib.wfh_ens_set_forcing(self.h, *ptrs, kind))

    def update_model(self):
        self._check(self.lib.wfh_ens_step(self.h))

    def run_timestep(self, precipitation, potential_evaporation, temperature):
        self.set_forcing(precipitation, potential_evaporation, temperature)
        self.update_model()

    def sync(self):
        self._check(self.lib.wfh_ens_sync(self.h))

    def get(self, var, member):
        cnt = _abi.var_count(var, self.n, self.nriv, self.nlayers)
        out = np.empty(cnt, np.float64)
        self._check(self.lib.wfh_ens_get(self.h, int(var), int(member), out.ctypes.data, cnt))
        return out

    def balance(self):
        out = np.empty(self.members, np.float64)
        self._check(self.lib.wfh_ens_balance(self.h, out.ctypes.data))
        return out

    def set_schedule(self, block, wide, strand_cells):
        self._check(self.lib.wfh_ens_set_schedule(self.h, int(block), int(wide), int(strand_cells)))

    def set_autotune(self, trials=1):
        self._check(self.lib.wfh_ens_set_autotune(self.h, int(trials)))

    def schedule(self):
        v = [ctypes.c_int32() for _ in range(5)]
        self._check(self.lib.wfh_ens_get_schedule(self.h, *(ctypes.addressof(x) for x in v)))
        return dict(zip(("block", "wide", "strand_cells", "stages", "tuning"), (x.value for x in v)))

    def set_timing(self, enable=True):
        self._check(self.lib.wfh_ens_set_timing(self.h, int(enable)))

    def timing(self):
        a = ctypes.c_double()
        self._check(self.lib.wfh_ens_get_timing(self.h, ctypes.byref(a)))
        return a.value

    def step_bytes(self):
        a = ctypes.c_int64()
        self._check(self.lib.wfh_ens_step_bytes(self.h, ctypes.addressof(a)))
        return a.value


def member_catchment(catchment, kv0_factor, f_factor, n_factor):
    """The catchment of one ensemble member: kv_0, f and Manning n (land and river) scaled."""
    import copy

    c = copy.copy(catchment)
    c.arrays = dict(catchment.arrays)
    c.arrays["kv_0"] = catchment.arrays["kv_0"] * float(kv0_factor)
    c.arrays["f"] = catchment.arrays["f"] * float(f_factor)
    c.arrays["land_manning_n"] = catchment.arrays["land_manning_n"] * float(n_factor)
    c.arrays["river_manning_n"] = catchment.arrays["river_manning_n"] * float(n_factor)
    return c


MLP_NPARAM = 64 * 8 + 64 + 64 * 64 + 64 + 64 + 1


def _bind_mlp(L):
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.wfh_mlp_create.argtypes = [i32, ctypes.c_uint64, ctypes.POINTER(P)]
    L.wfh_mlp_create.restype = i32
    L.wfh_mlp_set_normalization.argtypes = [P, P, P, ctypes.c_float, ctypes.c_float]
    L.wfh_mlp_set_normalization.restype = i32
    for name in ("wfh_mlp_set_params", "wfh_mlp_get_params"):
        getattr(L, name).argtypes = [P, P]
        getattr(L, name).restype = i32
    L.wfh_mlp_train.argtypes = [P, P, P, i64, i64, ctypes.c_float, P, P]
    L.wfh_mlp_train.restype = i32
    L.wfh_mlp_predict.argtypes = [P, P, i64, P]
    L.wfh_mlp_predict.restype = i32
    L.wfh_mlp_last_error.argtypes = [P]
    L.wfh_mlp_last_error.restype = ctypes.c_char_p
    L.wfh_mlp_destroy.argtypes = [P]
    L.wfh_mlp_destroy.restype = None
    L.wfh_ens_set_learned.argtypes = [P, P, ctypes.c_double]
    L.wfh_ens_set_learned.restype = i32
    L.wfh_ens_set_harvest.argtypes = [P, P, i64, i32]
    L.wfh_ens_set_harvest.restype = i32
    L.wfh_ens_harvest_count.argtypes = [P, P]
    L.wfh_ens_harvest_count.restype = i32
    L.wfh_ens_kw_histogram.argtypes = [P, i32, P]
    L.wfh_ens_kw_histogram.restype = i32


def _lib_with_mlp():
    L = load_library()
    if not hasattr(L, "_mlp_bound"):
        _bind_mlp(L)
        L._mlp_bound = True
    return L


class Mlp:
    """The learned Newton warm start network (8-64-64-1, BF16 on the matrix cores): features of a
    kinematic-wave solve -> log10 q - 5 log10 (reference start u). Trained on the GPU with the
    library's own forward and backward kernels. Data are torch tensors on the device."""

    def __init__(self, device=0, seed=0):
        self.lib = _lib_with_mlp()
        h = ctypes.c_void_p()
        rc = self.lib.wfh_mlp_create(int(device), int(seed), ctypes.byref(h))
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_mlp_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.wfh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise WflowError(rc, self.lib.wfh_mlp_last_error(self.h).decode())

    def set_normalization(self, x_mean, x_std, y_mean, y_std):
        xm = np.ascontiguousarray(x_mean, np.float32)
        xs = np.ascontiguousarray(x_std, np.float32)
        self._check(self.lib.wfh_mlp_set_normalization(self.h, xm.ctypes.data, xs.ctypes.data, float(y_mean),
                                                        float(y_std)))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_NPARAM
        self._check(self.lib.wfh_mlp_set_params(self.h, p.ctypes.data))

    def params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        self._check(self.lib.wfh_mlp_get_params(self.h, p.ctypes.data))
        return p

    def train_epoch(self, x, t, batch=8192, lr=1e-3):
        """One pass of Adam steps over device tensors x [n, 8] (float32), t [n] (normalised);
        returns the pass's mean squared error."""
        assert x.is_cuda and x.dtype.itemsize == 4 and x.is_contiguous() and t.is_contiguous()
        loss = ctypes.c_double()
        self._check(self.lib.wfh_mlp_train(self.h, x.data_ptr(), t.data_ptr(), int(t.numel()), int(batch), float(lr),
                                           ctypes.byref(loss), None))
        return loss.value

    def gradient(self, x, t):
        """Gradient of the mean squared error over all samples (no update), and the loss."""
        g = np.empty(MLP_NPARAM, np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.wfh_mlp_train(self.h, x.data_ptr(), t.data_ptr(), int(t.numel()), 64, 0.0,
                                           ctypes.byref(loss), g.ctypes.data))
        return g, loss.value

    def predict(self, x, out):
        self._check(self.lib.wfh_mlp_predict(self.h, x.data_ptr(), int(out.numel()), out.data_ptr()))
        return out


def _ens_learned_methods():
    def set_learned(self, mlp, gate=float("inf")):
        L = _lib_with_mlp()
        self._check(L.wfh_ens_set_learned(self.h, mlp.h if mlp is not None else None, float(gate)))
        self._mlp = mlp

    def set_harvest(self, rows, every=1):
        L = _lib_with_mlp()
        if rows is None:
            self._check(L.wfh_ens_set_harvest(self.h, None, 0, 1))
        else:
            assert rows.is_cuda and rows.dtype.itemsize == 4 and rows.dim() == 2 and rows.shape[1] == 10
            self._check(L.wfh_ens_set_harvest(self.h, rows.data_ptr(), int(rows.shape[0]), int(every)))
        self._harvest = rows

    def harvest_count(self):
        L = _lib_with_mlp()
        v = ctypes.c_int64()
        self._check(L.wfh_ens_harvest_count(self.h, ctypes.addressof(v)))
        return v.value

    def kw_histogram(self, enable=None):
        """(counts [2][16], learned starts [2]); enable True restarts counting, False stops."""
        L = _lib_with_mlp()
        out = np.zeros(34, np.uint64)
        flag = 2 if enable is None else int(bool(enable))
        self._check(L.wfh_ens_kw_histogram(self.h, flag, out.ctypes.data))
        return out[:32].reshape(2, 16), out[32:]

    Ensemble.set_learned = set_learned
    Ensemble.set_harvest = set_harvest
    Ensemble.harvest_count = harvest_count
    Ensemble.kw_histogram = kw_histogram


_ens_learned_methods()

# Upper edges of the Newton-update bins of Ensemble.kw_histogram
KW_BIN_EDGES = [0, 1, 2, 3, 4, 5, 6, 7, 8, 16, 32, 64, 128, 256, 2999, 3000]
