"""The reference's TOML model configuration (Config, Wflow/src/config_structure.jl:28-110) mapped
onto wfh_config for what the GPU step implements.

Only the [time] time step and the [model] section are read. Input and output sections (netCDF
paths, variable names, states) are the caller's business and are not interpreted here. A [model]
setting the GPU step does not implement raises ValueError naming it, so a configuration is
never silently run as something else.
"""
import tomli

from . import _abi

# ModelSection defaults (config_structure.jl:62-110) for the keys the GPU step uses
_DEFAULTS = {
    "type": None,
    "cold_start__flag": True,
    "cell_length_in_meter__flag": False,
    "reservoir__flag": False,
    "snow__flag": False,
    "snow_gravitational_transport__flag": False,
    "glacier__flag": False,
    "soil_infiltration_reduction__flag": False,
    "soil_layer__thickness": [100, 300, 800],
    "saturated_hydraulic_conductivity_profile": "exponential",
    "water_mass_balance__flag": False,
    "land_routing": "kinematic_wave",
    "river_routing": "kinematic_wave",
    "pit__flag": False,
    "river_streamorder__min_count": 6,
    "land_streamorder__min_count": 5,
    "kinematic_wave__adaptive_time_step_flag": False,
    "river_kinematic_wave__time_step": 900.0,
    "land_kinematic_wave__time_step": 3600.0,
    "subsurface_kinematic_wave__time_step": 86400.0,
    "subsurface_kinematic_wave__alpha_coefficient": 1.0,
    "floodplain_1d__flag": False,
}
# settings that only change the reference's CPU thread schedule, not the results
_SCHEDULE_ONLY = ("river_streamorder__min_count", "land_streamorder__min_count")
# ModelSection settings of other routing methods and model types: no effect on the kinematic-wave
# sbm step, accepted and ignored
_OTHER_METHODS = (
    "river_local_inertial_flow__alpha_coefficient", "land_local_inertial_flow__alpha_coefficient",
    "land_local_inertial_flow__theta_coefficient", "land_surface_water_flow_threshold__depth",
    "river_water_flow__froude_limit_flag", "land_surface_water_flow__froude_limit_flag",
    "river_water_flow_threshold__depth", "river_staggered_manning_flow__alpha_coefficient",
    "conductivity_profile", "drain__flag", "constanthead__flag", "subsurface_water_flow__alpha_coefficient",
    "river_subsurface_exchange_head_based__flag", "rainfall_erosion", "overland_flow_erosion",
    "run_river_model__flag", "land_transport", "river_transport")
_REQUIRE = {  # key: the only value the GPU step implements
    "type": "sbm",
    "glacier__flag": False,
    "land_routing": "kinematic_wave",
    "river_routing": "kinematic_wave",
    "pit__flag": False,
}
_PROFILES = ("exponential", "exponential_constant")


def load(path=None, text=None):
    """Parse a Wflow TOML file (or its text) into (wfh_config, info). info holds the settings the
    caller has to act on: 'reservoirs' (fill wfh_static's reservoir arrays and tables, and call
    wfh_set_julian_day for rating curves), 'floodplain' (fill wfh_static's floodplain_* fields;
    wfh_config.floodplain_1d is set), 'exponential_constant'
    (fill wfh_static.z_exp), 'cold_start', 'cell_length_in_meter' and 'schedule_only' (keys that
    do not change results)."""
    if text is None:
        with open(path, "rb") as fh:
            doc = tomli.load(fh)
    else:
        doc = tomli.loads(text)
    model = dict(_DEFAULTS)
    unknown = []
    for k, v in doc.get("model", {}).items():
        if k in model:
            model[k] = v
        elif k not in _OTHER_METHODS:
            unknown.append(k)  # incl. water_demand: allocation and demand are not implemented
    if unknown:
        raise ValueError("[model] settings the GPU step does not implement: " + ", ".join(sorted(unknown)))
    for k, want in _REQUIRE.items():
        if model[k] != want:
            raise ValueError(f"[model] {k} = {model[k]!r}: the GPU step implements {want!r} only")
    if model["saturated_hydraulic_conductivity_profile"] not in _PROFILES:
        raise ValueError("[model] saturated_hydraulic_conductivity_profile = "
                         f"{model['saturated_hydraulic_conductivity_profile']!r}: exponential or "
                         "exponential_constant only")
    if model["snow_gravitational_transport__flag"] and not model["snow__flag"]:
        model["snow_gravitational_transport__flag"] = False  # no-op without a snow model (sbm.jl:98-100)
    dt = doc.get("time", {}).get("timestepsecs")
    if dt is None:
        raise ValueError("[time] timestepsecs is required (the reference falls back to the forcing file)")
    layers = [float(x) for x in model["soil_layer__thickness"]]
    if not 1 <= len(layers) <= _abi.WFH_MAX_LAYERS - 1:
        raise ValueError("soil_layer__thickness: 1 to 7 layers")
    c = _abi.WfhConfig()
    c.abi_version = _abi.WFH_ABI_VERSION
    c.device = 0
    c.dt = float(dt)
    c.land_dt = float(model["land_kinematic_wave__time_step"])
    c.river_dt = float(model["river_kinematic_wave__time_step"])
    c.subsurface_dt = float(model["subsurface_kinematic_wave__time_step"])
    c.n_config_layers = len(layers)
    c.snow = int(bool(model["snow__flag"]))
    for k, v in enumerate(layers):
        c.soil_layer_thickness_mm[k] = v
    c.soil_infiltration_reduction = int(bool(model["soil_infiltration_reduction__flag"]))
    c.water_mass_balance = int(bool(model["water_mass_balance__flag"]))
    c.kinematic_wave_adaptive = int(bool(model["kinematic_wave__adaptive_time_step_flag"]))
    c.subsurface_alpha_coefficient = float(model["subsurface_kinematic_wave__alpha_coefficient"])
    c.snow_gravitational_transport = int(bool(model["snow_gravitational_transport__flag"]))
    c.floodplain_1d = int(bool(model["floodplain_1d__flag"]))
    info = {
        "reservoirs": bool(model["reservoir__flag"]),
        "floodplain": bool(model["floodplain_1d__flag"]),
        "exponential_constant": model["saturated_hydraulic_conductivity_profile"] == "exponential_constant",
        "cold_start": bool(model["cold_start__flag"]),
        "cell_length_in_meter": bool(model["cell_length_in_meter__flag"]),
        "schedule_only": {k: model[k] for k in _SCHEDULE_ONLY},
    }
    return c, info
