"""ctypes mirror of include/wflow_hip.h (the C ABI of the HIP library).

The structures and the variable enumeration must match the header field for field; the
test-suite checks the sizes against the compiled library's view of them.
"""

import ctypes
import enum

WFH_ABI_VERSION = 3
WFH_MAX_LAYERS = 8

WFH_OK = 0
WFH_ERR_ARG = -1
WFH_ERR_HIP = -2
WFH_ERR_CYCLE = -3
WFH_ERR_NONFINITE = -4
WFH_ERR_DEVICE = -5
WFH_ERR_STATE = -6

WFH_MEM_HOST = 0
WFH_MEM_DEVICE = 1
WFH_MEM_DEVICE_NODE_ORDER = 2

_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)


class WfhConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("land_dt", ctypes.c_double),
        ("river_dt", ctypes.c_double),
        ("subsurface_dt", ctypes.c_double),
        ("n_config_layers", ctypes.c_int32),
        ("snow", ctypes.c_int32),
        ("soil_layer_thickness_mm", ctypes.c_double * WFH_MAX_LAYERS),
        ("soil_infiltration_reduction", ctypes.c_int32),
        ("water_mass_balance", ctypes.c_int32),
        ("kinematic_wave_adaptive", ctypes.c_int32),
        ("subsurface_alpha_coefficient", ctypes.c_double),
        ("snow_gravitational_transport", ctypes.c_int32),
        ("floodplain_1d", ctypes.c_int32),
    ]


# (name, ctype) in header order; arrays are host pointers
STATIC_FIELDS = [
    ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("n", ctypes.c_int64), ("nriv", ctypes.c_int64),
    ("cell_index", _I64), ("ldd", _U8), ("x_length", _D), ("y_length", _D), ("land_slope", _D),
    ("land_manning_n", _D), ("river_location", _U8), ("river_width_land", _D), ("river_length_land", _D),
    ("water_fraction", _D), ("elevation", _D),
    ("canopy_gap_fraction", _D), ("max_canopy_storage", _D), ("rooting_depth", _D), ("crop_coefficient", _D),
    ("evap_precip_ratio", _D),
    ("cfmax", _D), ("tt", _D), ("tti", _D), ("ttm", _D), ("whc", _D),
    ("theta_s", _D), ("theta_r", _D), ("theta_fc", _D), ("kv_0", _D), ("f", _D), ("air_entry_pressure", _D),
    ("h1", _D), ("h2", _D), ("h3_high", _D), ("h3_low", _D), ("h4", _D), ("alpha_h1", _D),
    ("soil_thickness", _D), ("infiltcap_compacted", _D), ("max_leakage", _D), ("compacted_fraction", _D),
    ("rootdistpar", _D), ("cap_hmax", _D), ("cap_n", _D), ("w_soil", _D), ("cf_soil", _D),
    ("brooks_corey_c", _D), ("kvfrac", _D), ("rootfraction", _D),
    ("khfrac", _D), ("z_exp", _D),
    ("river_slope", _D), ("river_manning_n", _D), ("river_bankfull_depth", _D), ("river_length", _D),
    ("river_width", _D),
    ("nres", ctypes.c_int64), ("reservoir_outlet", _I64), ("reservoir_coverage", _I32),
    ("reservoir_outflow_type", _I32), ("reservoir_area", _D), ("reservoir_waterlevel", _D),
    ("reservoir_demand", _D), ("reservoir_maximum_release", _D), ("reservoir_maximum_storage", _D),
    ("reservoir_target_full_fraction", _D), ("reservoir_target_minimum_fraction", _D), ("reservoir_threshold", _D),
    ("reservoir_rating_curve_coefficient", _D), ("reservoir_external_inflow", _D), ("reservoir_outflow_obs", _D),
    ("reservoir_storage_curve_type", _I32), ("reservoir_rating_curve_exponent", _D),
    ("reservoir_sh_ptr", _I64), ("reservoir_sh_h", _D), ("reservoir_sh_s", _D),
    ("reservoir_hq_ptr", _I64), ("reservoir_hq_h", _D), ("reservoir_hq_q", _D),
    ("floodplain_ndepth", ctypes.c_int32), ("floodplain_depth", _D), ("floodplain_volume", _D),
    ("floodplain_manning_n", _D), ("floodplain_slope", _D),
    ("reservoir_lower", _I32),
]


class WfhStatic(ctypes.Structure):
    _fields_ = STATIC_FIELDS


class Var(enum.IntEnum):
    CANOPY_STORAGE = 0
    SNOW_STORAGE = enum.auto()
    SNOW_WATER = enum.auto()
    SOIL_TEMPERATURE = enum.auto()
    SATWATERDEPTH = enum.auto()
    USTORELAYERDEPTH = enum.auto()
    WATER_TABLE_DEPTH = enum.auto()
    USTOREDEPTH = enum.auto()
    TRANSFER = enum.auto()
    RECHARGE = enum.auto()
    ACTUAL_INFILTRATION = enum.auto()
    INFILTRATION_EXCESS = enum.auto()
    SATURATION_EXCESS = enum.auto()
    TRANSPIRATION = enum.auto()
    SOIL_EVAPORATION = enum.auto()
    ACTUAL_CAPILLARY_FLUX = enum.auto()
    ACTUAL_LEAKAGE = enum.auto()
    ACTUAL_EVAPOTRANSPIRATION = enum.auto()
    INTERCEPTION_RATE = enum.auto()
    NET_RUNOFF = enum.auto()
    RUNOFF = enum.auto()
    NET_RUNOFF_RIVER = enum.auto()
    EXFILTRATION = enum.auto()
    TOTAL_STORAGE = enum.auto()
    SNOW_WATER_EQUIVALENT = enum.auto()
    VWC = enum.auto()
    VWC_ROOT_ZONE = enum.auto()
    ROOT_ZONE_STORAGE = enum.auto()
    SSF_Q = enum.auto()
    SSF_QIN_AVERAGE = enum.auto()
    SSF_Q_AVERAGE = enum.auto()
    SSF_TO_RIVER_AVERAGE = enum.auto()
    SSF_STORAGE = enum.auto()
    SSF_QMAX = enum.auto()
    LAND_Q = enum.auto()
    LAND_QIN = enum.auto()
    LAND_H = enum.auto()
    LAND_STORAGE = enum.auto()
    LAND_Q_AVERAGE = enum.auto()
    LAND_QIN_AVERAGE = enum.auto()
    LAND_TO_RIVER_AVERAGE = enum.auto()
    LAND_INWATER = enum.auto()
    RIVER_Q = enum.auto()
    RIVER_QIN = enum.auto()
    RIVER_H = enum.auto()
    RIVER_STORAGE = enum.auto()
    RIVER_Q_AVERAGE = enum.auto()
    RIVER_QIN_AVERAGE = enum.auto()
    RIVER_INWATER = enum.auto()
    BALANCE_LAND_ERROR = enum.auto()
    BALANCE_LAND_RELERROR = enum.auto()
    BALANCE_SSF_ERROR = enum.auto()
    BALANCE_OVERLAND_ERROR = enum.auto()
    BALANCE_RIVER_ERROR = enum.auto()
    LAND_ALPHA = enum.auto()
    RIVER_ALPHA = enum.auto()
    THETA_FC = enum.auto()
    SURFACE_FLOW_WIDTH = enum.auto()
    FLOW_FRACTION_TO_RIVER = enum.auto()
    RESERVOIR_STORAGE = enum.auto()
    RESERVOIR_WATERLEVEL = enum.auto()
    RESERVOIR_OUTFLOW = enum.auto()
    RESERVOIR_OUTFLOW_AVERAGE = enum.auto()
    RESERVOIR_INFLOW_AVERAGE = enum.auto()
    RESERVOIR_ACTEVAP_AVERAGE = enum.auto()
    RESERVOIR_ABSTRACTION_AVERAGE = enum.auto()
    RESERVOIR_PRECIPITATION = enum.auto()
    RESERVOIR_EVAPORATION = enum.auto()
    BALANCE_RESERVOIR_ERROR = enum.auto()
    SNOW_IN = enum.auto()
    SNOW_OUT = enum.auto()
    FLOODPLAIN_STORAGE = enum.auto()
    FLOODPLAIN_H = enum.auto()
    FLOODPLAIN_Q = enum.auto()
    FLOODPLAIN_Q_AVERAGE = enum.auto()
    FLOODPLAIN_QIN_AVERAGE = enum.auto()
    RIVER_Q_CHANNEL_AVERAGE = enum.auto()
    VAR_COUNT = enum.auto()


LAYERED_VARS = {Var.USTORELAYERDEPTH, Var.VWC}
RIVER_VARS = {Var.RIVER_Q, Var.RIVER_QIN, Var.RIVER_H, Var.RIVER_STORAGE, Var.RIVER_Q_AVERAGE,
              Var.RIVER_QIN_AVERAGE, Var.RIVER_INWATER, Var.BALANCE_RIVER_ERROR, Var.RIVER_ALPHA,
              Var.FLOODPLAIN_STORAGE, Var.FLOODPLAIN_H, Var.FLOODPLAIN_Q, Var.FLOODPLAIN_Q_AVERAGE,
              Var.FLOODPLAIN_QIN_AVERAGE, Var.RIVER_Q_CHANNEL_AVERAGE}
FLOODPLAIN_VARS = [Var.FLOODPLAIN_STORAGE, Var.FLOODPLAIN_H, Var.FLOODPLAIN_Q, Var.FLOODPLAIN_Q_AVERAGE,
                   Var.FLOODPLAIN_QIN_AVERAGE, Var.RIVER_Q_CHANNEL_AVERAGE]
RESERVOIR_VARS = {Var.RESERVOIR_STORAGE, Var.RESERVOIR_WATERLEVEL, Var.RESERVOIR_OUTFLOW,
                  Var.RESERVOIR_OUTFLOW_AVERAGE, Var.RESERVOIR_INFLOW_AVERAGE, Var.RESERVOIR_ACTEVAP_AVERAGE,
                  Var.RESERVOIR_ABSTRACTION_AVERAGE, Var.RESERVOIR_PRECIPITATION, Var.RESERVOIR_EVAPORATION,
                  Var.BALANCE_RESERVOIR_ERROR}
SETTABLE_VARS = {Var.CANOPY_STORAGE, Var.SNOW_STORAGE, Var.SNOW_WATER, Var.SOIL_TEMPERATURE, Var.SATWATERDEPTH,
                 Var.USTORELAYERDEPTH, Var.SSF_Q, Var.LAND_Q, Var.LAND_H, Var.RIVER_Q, Var.RIVER_H,
                 Var.RESERVOIR_STORAGE, Var.RESERVOIR_WATERLEVEL, Var.FLOODPLAIN_STORAGE}

# entry points declared in include/wflow_hip.h (checked against the built library)
EXPORTS = ["wfh_create", "wfh_set_forcing", "wfh_step", "wfh_get", "wfh_set", "wfh_get_topology",
           "wfh_state_size", "wfh_save_state", "wfh_load_state", "wfh_sync", "wfh_set_timing",
           "wfh_get_timing", "wfh_last_error", "wfh_destroy", "wfh_get_node_order",
           "wfh_kernel_bytes", "wfh_last_substeps", "wfh_set_julian_day", "wfh_step_noalloc", "wfh_set_graph", "wfh_autotune",
           "wfh_get_schedule", "wfh_test_route_stall", "wfh_run",
           "wfh_ens_create", "wfh_ens_set_forcing", "wfh_ens_step", "wfh_ens_get", "wfh_ens_balance",
           "wfh_ens_set_schedule", "wfh_ens_set_autotune", "wfh_ens_get_schedule", "wfh_ens_sync",
           "wfh_ens_set_timing", "wfh_ens_get_timing", "wfh_ens_step_bytes", "wfh_ens_last_error",
           "wfh_ens_destroy", "wfh_mlp_create", "wfh_mlp_set_normalization", "wfh_mlp_set_params",
           "wfh_mlp_get_params", "wfh_mlp_train", "wfh_mlp_predict", "wfh_mlp_last_error", "wfh_mlp_destroy",
           "wfh_ens_set_learned", "wfh_ens_set_harvest", "wfh_ens_harvest_count", "wfh_ens_kw_histogram"]


def var_count(var, n, nriv, nlayers, nres=0):
    var = Var(var)
    if var in RESERVOIR_VARS:
        return nres
    if var in LAYERED_VARS:
        return nlayers * n
    if var in RIVER_VARS:
        return nriv
    return n
