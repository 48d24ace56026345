# WflowHIP: Wflow.jl's `update_model!` for an `sbm` model with kinematic-wave routing, run on one
# MI355X through lib/libwflow_hip.so (include/wflow_hip.h).
#
#   g = WflowHIP.GpuStep(model, dataset)       # dataset: the model's static maps (NCDataset)
#   for t in times
#       Wflow.run_timestep!(model; update_func = m -> WflowHIP.hip_update!(g, m))
#   end
#
# `run_timestep!(model; update_func)` (Wflow/src/Wflow.jl:311-322) reads the forcing into
# model.land.atmospheric_forcing, calls update_func(model), then writes the outputs. hip_update!
# hands the forcing to the GPU, runs one step there, and copies back every array that the
# model's configured outputs and states read ([output.netcdf_grid], [output.netcdf_scalar],
# [output.csv], [state], [API]), so that Wflow's writers see the step's values.
#
# Covered: what wfh_create accepts (include/wflow_hip.h): kinematic-wave land, river and
# subsurface routing, fixed or adaptive sub-steps, exponential or exponential-constant Ksat
# profiles, optional snow, frozen-soil reduction, lateral snow transport and water balance.
# Reservoirs and the 1-D floodplain have fields in wfh_static too; this module does not fill them
# and refuses models that set reservoir__flag or floodplain_1d__flag.
#
# The struct layouts below mirror include/wflow_hip.h field by field (tests/test_julia_shim_cpu.py
# checks names and order against the library's ctypes mirror, whose offsets are checked against
# the compiled library).
module WflowHIP

using Wflow
using StaticArrays

const LIB = normpath(joinpath(@__DIR__, "..", "wflow.jl_amd", "lib", "libwflow_hip.so"))
const WFH_ABI_VERSION = Int32(3)
const WFH_MAX_LAYERS = 8
const WFH_MEM_HOST = Int32(0)

struct WfhConfig
    abi_version::Int32
    device::Int32
    dt::Float64
    land_dt::Float64
    river_dt::Float64
    subsurface_dt::Float64
    n_config_layers::Int32
    snow::Int32
    soil_layer_thickness_mm::NTuple{WFH_MAX_LAYERS, Float64}
    soil_infiltration_reduction::Int32
    water_mass_balance::Int32
    kinematic_wave_adaptive::Int32
    subsurface_alpha_coefficient::Float64
    snow_gravitational_transport::Int32
    floodplain_1d::Int32
end

const PD = Ptr{Float64}
Base.@kwdef struct WfhStatic
    nx::Int64 = 0
    ny::Int64 = 0
    n::Int64 = 0
    nriv::Int64 = 0
    cell_index::Ptr{Int64} = C_NULL
    ldd::Ptr{UInt8} = C_NULL
    x_length::PD = C_NULL
    y_length::PD = C_NULL
    land_slope::PD = C_NULL
    land_manning_n::PD = C_NULL
    river_location::Ptr{UInt8} = C_NULL
    river_width_land::PD = C_NULL
    river_length_land::PD = C_NULL
    water_fraction::PD = C_NULL
    elevation::PD = C_NULL
    canopy_gap_fraction::PD = C_NULL
    max_canopy_storage::PD = C_NULL
    rooting_depth::PD = C_NULL
    crop_coefficient::PD = C_NULL
    evap_precip_ratio::PD = C_NULL
    cfmax::PD = C_NULL
    tt::PD = C_NULL
    tti::PD = C_NULL
    ttm::PD = C_NULL
    whc::PD = C_NULL
    theta_s::PD = C_NULL
    theta_r::PD = C_NULL
    theta_fc::PD = C_NULL
    kv_0::PD = C_NULL
    f::PD = C_NULL
    air_entry_pressure::PD = C_NULL
    h1::PD = C_NULL
    h2::PD = C_NULL
    h3_high::PD = C_NULL
    h3_low::PD = C_NULL
    h4::PD = C_NULL
    alpha_h1::PD = C_NULL
    soil_thickness::PD = C_NULL
    infiltcap_compacted::PD = C_NULL
    max_leakage::PD = C_NULL
    compacted_fraction::PD = C_NULL
    rootdistpar::PD = C_NULL
    cap_hmax::PD = C_NULL
    cap_n::PD = C_NULL
    w_soil::PD = C_NULL
    cf_soil::PD = C_NULL
    brooks_corey_c::PD = C_NULL
    kvfrac::PD = C_NULL
    rootfraction::PD = C_NULL
    khfrac::PD = C_NULL
    z_exp::PD = C_NULL
    river_slope::PD = C_NULL
    river_manning_n::PD = C_NULL
    river_bankfull_depth::PD = C_NULL
    river_length::PD = C_NULL
    river_width::PD = C_NULL
    nres::Int64 = 0
    reservoir_outlet::Ptr{Int64} = C_NULL
    reservoir_coverage::Ptr{Int32} = C_NULL
    reservoir_outflow_type::Ptr{Int32} = C_NULL
    reservoir_area::PD = C_NULL
    reservoir_waterlevel::PD = C_NULL
    reservoir_demand::PD = C_NULL
    reservoir_maximum_release::PD = C_NULL
    reservoir_maximum_storage::PD = C_NULL
    reservoir_target_full_fraction::PD = C_NULL
    reservoir_target_minimum_fraction::PD = C_NULL
    reservoir_threshold::PD = C_NULL
    reservoir_rating_curve_coefficient::PD = C_NULL
    reservoir_external_inflow::PD = C_NULL
    reservoir_outflow_obs::PD = C_NULL
    reservoir_storage_curve_type::Ptr{Int32} = C_NULL
    reservoir_rating_curve_exponent::PD = C_NULL
    reservoir_sh_ptr::Ptr{Int64} = C_NULL
    reservoir_sh_h::PD = C_NULL
    reservoir_sh_s::PD = C_NULL
    reservoir_hq_ptr::Ptr{Int64} = C_NULL
    reservoir_hq_h::PD = C_NULL
    reservoir_hq_q::PD = C_NULL
    floodplain_ndepth::Int32 = 0
    floodplain_depth::PD = C_NULL
    floodplain_volume::PD = C_NULL
    floodplain_manning_n::PD = C_NULL
    floodplain_slope::PD = C_NULL
    reservoir_lower::Ptr{Int32} = C_NULL
end

# wfh_var, in header order (0-based)
const WFH_VARS = (:CANOPY_STORAGE, :SNOW_STORAGE, :SNOW_WATER, :SOIL_TEMPERATURE, :SATWATERDEPTH,
    :USTORELAYERDEPTH, :WATER_TABLE_DEPTH, :USTOREDEPTH, :TRANSFER, :RECHARGE, :ACTUAL_INFILTRATION,
    :INFILTRATION_EXCESS, :SATURATION_EXCESS, :TRANSPIRATION, :SOIL_EVAPORATION, :ACTUAL_CAPILLARY_FLUX,
    :ACTUAL_LEAKAGE, :ACTUAL_EVAPOTRANSPIRATION, :INTERCEPTION_RATE, :NET_RUNOFF, :RUNOFF,
    :NET_RUNOFF_RIVER, :EXFILTRATION, :TOTAL_STORAGE, :SNOW_WATER_EQUIVALENT, :VWC, :VWC_ROOT_ZONE,
    :ROOT_ZONE_STORAGE, :SSF_Q, :SSF_QIN_AVERAGE, :SSF_Q_AVERAGE, :SSF_TO_RIVER_AVERAGE, :SSF_STORAGE,
    :SSF_QMAX, :LAND_Q, :LAND_QIN, :LAND_H, :LAND_STORAGE, :LAND_Q_AVERAGE, :LAND_QIN_AVERAGE,
    :LAND_TO_RIVER_AVERAGE, :LAND_INWATER, :RIVER_Q, :RIVER_QIN, :RIVER_H, :RIVER_STORAGE,
    :RIVER_Q_AVERAGE, :RIVER_QIN_AVERAGE, :RIVER_INWATER)
wfh(var::Symbol) = Int32(findfirst(==(var), WFH_VARS) - 1)

# Where each variable the GPU computes lives in the Wflow model (the lenses of Wflow's standard
# name maps, Wflow/src/standard_name/standard_name_sbm.jl and standard_name_routing.jl).
const ACCESSORS = (
    :CANOPY_STORAGE => m -> m.land.interception.variables.canopy_storage,
    :INTERCEPTION_RATE => m -> m.land.interception.variables.interception_rate,
    :SNOW_STORAGE => m -> m.land.snow.variables.snow_storage,
    :SNOW_WATER => m -> m.land.snow.variables.snow_water,
    :SNOW_WATER_EQUIVALENT => m -> m.land.snow.variables.snow_water_equivalent,
    :SOIL_TEMPERATURE => m -> m.land.soil.variables.soil_surface_temperature,
    :SATWATERDEPTH => m -> m.land.soil.variables.saturated_water_depth,
    :USTORELAYERDEPTH => m -> m.land.soil.variables.unsaturated_layer_depth,
    :WATER_TABLE_DEPTH => m -> m.land.soil.variables.water_table_depth,
    :USTOREDEPTH => m -> m.land.soil.variables.unsaturated_store_depth,
    :TRANSFER => m -> m.land.soil.variables.transfer,
    :RECHARGE => m -> m.land.soil.variables.recharge,
    :ACTUAL_INFILTRATION => m -> m.land.soil.variables.actual_infiltration,
    :TRANSPIRATION => m -> m.land.soil.variables.transpiration,
    :ACTUAL_CAPILLARY_FLUX => m -> m.land.soil.variables.actual_capillary_flux,
    :ACTUAL_LEAKAGE => m -> m.land.soil.variables.actual_leakage,
    :ACTUAL_EVAPOTRANSPIRATION => m -> m.land.soil.variables.actual_evapotranspiration,
    :NET_RUNOFF => m -> m.land.soil.variables.net_runoff,
    :RUNOFF => m -> m.land.soil.variables.runoff,
    :EXFILTRATION => m -> m.land.soil.variables.exfiltration_saturated_water,
    :TOTAL_STORAGE => m -> m.land.soil.variables.total_storage,
    :VWC => m -> m.land.soil.variables.volumetric_water_content,
    :VWC_ROOT_ZONE => m -> m.land.soil.variables.volumetric_water_content_root_zone,
    :ROOT_ZONE_STORAGE => m -> m.land.soil.variables.root_zone_storage,
    :SSF_Q => m -> m.routing.subsurface_flow.variables.q,
    :SSF_Q_AVERAGE => m -> m.routing.subsurface_flow.variables.q_average,
    :SSF_QIN_AVERAGE => m -> m.routing.subsurface_flow.variables.q_in_average,
    :SSF_TO_RIVER_AVERAGE => m -> m.routing.subsurface_flow.variables.to_river_average,
    :SSF_STORAGE => m -> m.routing.subsurface_flow.variables.storage,
    :SSF_QMAX => m -> m.routing.subsurface_flow.variables.q_max,
    :LAND_Q => m -> m.routing.overland_flow.variables.q,
    :LAND_H => m -> m.routing.overland_flow.variables.h,
    :LAND_STORAGE => m -> m.routing.overland_flow.variables.storage,
    :LAND_Q_AVERAGE => m -> m.routing.overland_flow.variables.q_average,
    :LAND_TO_RIVER_AVERAGE => m -> m.routing.overland_flow.variables.to_river_average,
    :RIVER_Q => m -> m.routing.river_flow.variables.q,
    :RIVER_H => m -> m.routing.river_flow.variables.h,
    :RIVER_STORAGE => m -> m.routing.river_flow.variables.storage,
    :RIVER_Q_AVERAGE => m -> m.routing.river_flow.variables.q_average,
    :RIVER_INWATER => m -> m.routing.river_flow.boundary_conditions.inwater,
)
# the subsurface model's own copy of the water table depth is the soil's after a step
const ALIASES = (:WATER_TABLE_DEPTH => m -> m.routing.subsurface_flow.variables.water_table_depth,)
const STATE_VARS = (:CANOPY_STORAGE, :SNOW_STORAGE, :SNOW_WATER, :SOIL_TEMPERATURE, :SATWATERDEPTH,
    :USTORELAYERDEPTH, :SSF_Q, :LAND_Q, :LAND_H, :RIVER_Q, :RIVER_H)

mutable struct GpuStep
    handle::Ptr{Cvoid}
    keep::Vector{Any}                     # host arrays the static struct points into
    copyback::Vector{Tuple{Int32, Any}}   # (wfh_var, destination array) after every step
    layer_buf::Matrix{Float64}            # [n][N] staging for layered variables
end

last_error(h) = unsafe_string(ccall((:wfh_last_error, LIB), Cstring, (Ptr{Cvoid},), h))
check(g::GpuStep, rc) = rc == 0 || error("wflow_hip error $rc: ", last_error(g.handle))

# SVector AoS (one vector per cell) -> layer-major [N][n] array, as wfh_static's layered fields
layer_major(v::Vector{<:SVector{N}}) where {N} = [v[i][k] for i in eachindex(v), k in 1:N]
layer_major(v::Vector{Float64}) = reshape(copy(v), :, 1)

"""
    build_config_and_static(model, dataset; device = 0) -> (cfg, static, keep)

wfh_config and wfh_static for `model`, taking the derived parameters from the model (the values
Wflow computed from its static maps) and the raw maps that wfh_create derives itself (land and
river Manning n, river width and length on the land grid, the open-water fraction, elevation)
from `dataset` through Wflow's own reader `ncread`. `keep` holds every array the pointers refer to
and must outlive wfh_create.
"""
function build_config_and_static(model, dataset; device = 0)
    config = model.config
    ms = config.model
    ms.reservoir__flag && error("WflowHIP: reservoirs are not mapped by this module")
    ms.floodplain_1d__flag && error("WflowHIP: the 1-D floodplain is not mapped by this module")
    ms.glacier__flag && error("wflow_hip: glaciers are not supported")
    (ms.land_routing == Wflow.RoutingType.kinematic_wave &&
     ms.river_routing == Wflow.RoutingType.kinematic_wave) ||
        error("wflow_hip: kinematic-wave land and river routing only")
    net = model.domain.land.network
    lp = model.domain.land.parameters
    rp = model.domain.river.parameters
    sp = model.land.soil.parameters
    veg = model.land.vegetation_parameters
    keep = Any[]
    P(a::AbstractVector{Float64}) = (b = collect(a); push!(keep, b); pointer(b))
    P(a::AbstractMatrix{Float64}) = (b = collect(a); push!(keep, b); pointer(b))
    nx, ny = net.modelsize
    n = length(net.indices)
    cell_index = Int64[(I[1] - 1) + (I[2] - 1) * nx for I in net.indices]
    issorted(cell_index; lt = <=) || error("WflowHIP: land indices are not strictly ascending")
    push!(keep, cell_index)
    ldd = collect(UInt8, net.local_drain_direction)
    push!(keep, ldd)
    river_location = UInt8.(lp.river_location)
    push!(keep, river_location)
    rd(name, type; default = nothing) = begin
        v = Wflow.ncread(dataset, config, name, type; sel = net.indices, logging = false)
        isnothing(default) ? Float64.(v) : Float64.(coalesce.(v, default))
    end
    kvp = sp.kv_profile
    kvexp = kvp isa Wflow.KvExponentialConstant ? kvp.exponential : kvp
    kvexp isa Wflow.KvExponential ||
        error("wflow_hip: exponential or exponential-constant Ksat profiles only")
    z_exp = kvp isa Wflow.KvExponentialConstant ? P(kvp.z_exp) : C_NULL
    ssf = model.routing.subsurface_flow.parameters
    gash = model.land.interception isa Wflow.GashInterceptionModel
    snow = model.land.snow isa Wflow.SnowHbvModel
    snp = snow ? model.land.snow.parameters : nothing
    zeros_n = zeros(n)
    push!(keep, zeros_n)
    thick = ms.soil_layer__thickness
    length(thick) < WFH_MAX_LAYERS || error("wflow_hip: at most $(WFH_MAX_LAYERS - 1) configured soil layers")
    mm = ntuple(k -> k <= length(thick) ? Float64(thick[k]) : 0.0, WFH_MAX_LAYERS)
    cfg = WfhConfig(WFH_ABI_VERSION, Int32(device), Float64(config.time.timestepsecs),
        ms.land_kinematic_wave__time_step, ms.river_kinematic_wave__time_step,
        ms.subsurface_kinematic_wave__time_step, Int32(length(thick)), Int32(snow), mm,
        Int32(ms.soil_infiltration_reduction__flag), Int32(ms.water_mass_balance__flag),
        Int32(ms.kinematic_wave__adaptive_time_step_flag), ms.subsurface_kinematic_wave__alpha_coefficient,
        Int32(ms.snow_gravitational_transport__flag), Int32(0))
    st = WfhStatic(;
        nx, ny, n, nriv = length(rp.flow_length),
        cell_index = pointer(cell_index), ldd = pointer(ldd),
        x_length = P(lp.x_length), y_length = P(lp.y_length), land_slope = P(lp.slope),
        land_manning_n = P(rd("land_surface_water_flow__manning_n_parameter", Wflow.Routing)),
        river_location = pointer(river_location),
        river_width_land = P(rd("river__width", Wflow.Routing; default = 0.0)),
        river_length_land = P(rd("river__length", Wflow.Routing; default = 0.0)),
        water_fraction = P(rd("land_water_covered__area_fraction", Wflow.LandHydrologySBM)),
        elevation = ms.snow_gravitational_transport__flag ?
                    P(rd("land_surface__elevation", Wflow.LandHydrologySBM)) : pointer(zeros_n),
        canopy_gap_fraction = P(veg.canopy_gap_fraction),
        max_canopy_storage = P(veg.maximum_canopy_storage),
        rooting_depth = P(veg.rooting_depth), crop_coefficient = P(veg.crop_coefficient),
        evap_precip_ratio = gash ?
                            P(model.land.interception.parameters.evaporation_to_precipitation_ratio) :
                            pointer(zeros_n),
        cfmax = snow ? P(snp.degree_day_factor) : pointer(zeros_n),
        tt = snow ? P(snp.temperature_threshold_snowfall) : pointer(zeros_n),
        tti = snow ? P(snp.temperature_interval_snowfall) : pointer(zeros_n),
        ttm = snow ? P(snp.temperature_threshold_melt) : pointer(zeros_n),
        whc = snow ? P(snp.water_holding_capacity) : pointer(zeros_n),
        theta_s = P(sp.theta_s), theta_r = P(sp.theta_r), theta_fc = P(sp.theta_fc),
        kv_0 = P(kvexp.kv_0), f = P(kvexp.hydraulic_conductivity_scale_parameter),
        air_entry_pressure = P(sp.air_entry_pressure),
        h1 = P(sp.h1), h2 = P(sp.h2), h3_high = P(sp.h3_high), h3_low = P(sp.h3_low), h4 = P(sp.h4),
        alpha_h1 = P(sp.alpha_h1), soil_thickness = P(sp.soil_thickness),
        infiltcap_compacted = P(sp.infiltration_capacity_compacted_soil),
        max_leakage = P(sp.maximum_leakage), compacted_fraction = P(sp.compacted_soil_area_fraction),
        rootdistpar = P(sp.wet_root_distribution_parameter), cap_hmax = P(sp.cap_hmax),
        cap_n = P(sp.cap_n), w_soil = P(sp.w_soil), cf_soil = P(sp.cf_soil),
        brooks_corey_c = P(layer_major(sp.brooks_corey_exponent)),
        kvfrac = P(layer_major(sp.vertical_hydraulic_conductivity_factor)),
        rootfraction = P(layer_major(sp.rootfraction)),
        khfrac = P(ssf.horizontal_to_vertical_hydraulic_conductivity_ratio), z_exp,
        river_slope = P(rp.slope),
        river_manning_n = P(model.routing.river_flow.parameters.flow.mannings_n),
        river_bankfull_depth = P(model.routing.river_flow.parameters.bankfull_depth),
        river_length = P(rp.flow_length), river_width = P(rp.flow_width))
    return cfg, st, keep
end

"Names of every model array that the configured outputs, states and API variables read."
function configured_names(config)
    names = String[]
    out = config.output
    append!(names, String.(keys(out.netcdf_grid.variables)))
    append!(names, [v.parameter for v in out.netcdf_scalar.variable])
    append!(names, [c.parameter for c in out.csv.column])
    append!(names, String.(keys(config.state.variables)))
    append!(names, config.API.variables)
    return unique(names)
end

function GpuStep(model, dataset; device = 0)
    cfg, st, keep = build_config_and_static(model, dataset; device)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve keep ccall((:wfh_create, LIB), Int32, (Ref{WfhConfig}, Ref{WfhStatic}, Ref{Ptr{Cvoid}}),
        cfg, st, h)
    rc == 0 || error("wfh_create error $rc: ", last_error(h[]))
    n = Int(st.n)
    g = GpuStep(h[], Any[], Tuple{Int32, Any}[], Matrix{Float64}(undef, n, cfg.n_config_layers + 1))
    finalizer(x -> ccall((:wfh_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.handle), g)
    # every configured name resolves (Wflow's own lookup) to a model array; each must be one the
    # GPU step computes, and is then copied back after every step
    by_array = IdDict{Any, Symbol}()
    for (var, acc) in (ACCESSORS..., ALIASES...)
        a = try
            acc(model)
        catch
            nothing
        end
        isnothing(a) || (by_array[a] = var)
    end
    wanted = IdDict{Any, Symbol}()
    for var in STATE_VARS
        a = try
            Dict(ACCESSORS)[var](model)
        catch
            nothing
        end
        isnothing(a) || (wanted[a] = var)
    end
    for name in configured_names(model.config)
        a, _ = Wflow.get_field_in_model(model, name)
        haskey(by_array, a) || error("WflowHIP: output `$name` is not computed by the GPU step")
        wanted[a] = by_array[a]
    end
    for (a, var) in wanted
        push!(g.copyback, (wfh(var), a))
    end
    return g
end

function copy_back!(g::GpuStep, var::Int32, dst::Vector{Float64})
    check(g, ccall((:wfh_get, LIB), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Int32),
        g.handle, var, dst, length(dst), WFH_MEM_HOST))
end
function copy_back!(g::GpuStep, var::Int32, dst::Vector{<:SVector{N}}) where {N}
    buf = g.layer_buf   # wfh_get returns [N][n] layer-major: column k of buf is layer k
    size(buf, 2) == N || error("WflowHIP: layer count mismatch")
    check(g, ccall((:wfh_get, LIB), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Int32),
        g.handle, var, buf, length(buf), WFH_MEM_HOST))
    @inbounds for i in eachindex(dst)
        dst[i] = SVector{N}(ntuple(k -> buf[i, k], N))
    end
end

"""
    hip_update!(g, model)

`update_model!(model)` (Wflow/src/sbm_model.jl:60-92) on the GPU: the step's forcing from
model.land.atmospheric_forcing, the rating-curve day from the clock, one step, then every
configured output and state array copied back into the model.
"""
function hip_update!(g::GpuStep, model)
    f = model.land.atmospheric_forcing
    check(g, ccall((:wfh_set_forcing, LIB), Int32,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32),
        g.handle, f.precipitation, f.potential_evaporation, f.temperature, WFH_MEM_HOST))
    clock = model.clock
    check(g, ccall((:wfh_set_julian_day, LIB), Int32, (Ptr{Cvoid}, Int32), g.handle,
        Int32(Wflow.julian_day(clock.time - clock.dt))))
    check(g, ccall((:wfh_step, LIB), Int32, (Ptr{Cvoid},), g.handle))
    for (var, dst) in g.copyback
        copy_back!(g, var, dst)
    end
    return nothing
end

end # module
