/*
 * wflow_hip.h - C ABI of the MI355X (gfx950) implementation of the Wflow.jl `sbm` time step.
 *
 * The replaceable unit in the reference is
 *     update_model!(model::AbstractModel{<:SbmModel})        Wflow/src/sbm_model.jl:60-92
 * called through the step hook
 *     run_timestep!(model; update_func = update_model!)      Wflow/src/Wflow.jl:311-322
 * One wfh_step() call performs exactly one update_model!: land hydrology (interception, snow,
 * open-water runoff, SBM soil column), lateral subsurface flow (kinematic wave), the soil
 * storage update, overland and river kinematic-wave routing and the total-storage update.
 * storage_prev!/compute_mass_balance! (Wflow.jl:315-317) sit outside update_model! in the
 * reference; here they are optional and fused into the same step (wfh_config.water_mass_balance).
 *
 * Conventions
 *  - All quantities are in the reference's internal SI units (m, s, K, m3 s-1; units.jl:98-121).
 *  - Land arrays have n entries in land-index order (active cells in column-major raster order,
 *    Wflow/src/utils.jl:85-99); river arrays have nriv entries ordered by ascending land index
 *    of the river cells (Wflow/src/network.jl:214-258).
 *  - Layered arrays are layer-major: a[k * n + i] is layer k of cell i (the reference stores
 *    Vector{SVector{N}}; the host shim converts).
 *  - Every function returns WFH_OK (0) or a negative status; no exception crosses the ABI.
 *    wfh_last_error() gives the message of the last failure on a handle.
 *  - A handle owns one HIP stream and its device memory. Calls on one handle must come from
 *    one host thread at a time. wfh_step() is asynchronous with fixed sub-steps; wfh_get()/wfh_sync()
 *    synchronise.
 */
#ifndef WFLOW_HIP_H
#define WFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFH_ABI_VERSION 3
#define WFH_MAX_LAYERS 8

enum wfh_status {
    WFH_OK = 0,
    WFH_ERR_ARG = -1,       /* invalid argument (null pointer, size, enum)            */
    WFH_ERR_HIP = -2,       /* a HIP runtime call failed                               */
    WFH_ERR_CYCLE = -3,     /* the drainage network has a cycle (routing/utils.jl:26) */
    WFH_ERR_NONFINITE = -4, /* non-finite forcing in a step (reported by wfh_sync)       */
    WFH_ERR_DEVICE = -5,    /* a kernel reported a failure (dependency wait timed out) */
    WFH_ERR_STATE = -6      /* call out of order (e.g. wfh_step before wfh_set_forcing) */
};

/* WFH_MEM_HOST / WFH_MEM_DEVICE: arrays in land-index order in host / device memory.
 * WFH_MEM_DEVICE_NODE_ORDER: device arrays already in the library's node order
 * (wfh_get_node_order), bound without a copy. */
enum wfh_memkind { WFH_MEM_HOST = 0, WFH_MEM_DEVICE = 1, WFH_MEM_DEVICE_NODE_ORDER = 2 };

/* Model settings: the fields of the reference's `ModelSection` (config_structure.jl:62-115)
 * that change the per-step arithmetic, plus the model time step. */
typedef struct wfh_config {
    int32_t abi_version;                 /* = WFH_ABI_VERSION                                  */
    int32_t device;                      /* HIP device ordinal                                 */
    double dt;                           /* model time step [s] (Clock.dt)                     */
    double land_dt;                      /* land_kinematic_wave__time_step [s], default 3600    */
    double river_dt;                     /* river_kinematic_wave__time_step [s], default 900    */
    double subsurface_dt;                /* subsurface_kinematic_wave__time_step, default 86400 */
    int32_t n_config_layers;             /* length of soil_layer__thickness (N - 1)            */
    int32_t snow;                        /* snow__flag                                         */
    double soil_layer_thickness_mm[WFH_MAX_LAYERS]; /* soil_layer__thickness [mm]            */
    int32_t soil_infiltration_reduction; /* soil_infiltration_reduction__flag                  */
    int32_t water_mass_balance;          /* water_mass_balance__flag                           */
    /* kinematic_wave__adaptive_time_step_flag (config_structure.jl:84): 0 = fixed sub-steps of
     * land_dt / river_dt / subsurface_dt; 1 = adaptive sub-steps (surface_kinwave.jl:674-704,
     * lateral_subsurface_flow.jl:314-344), the *_dt fields are then ignored.            */
    int32_t kinematic_wave_adaptive;
    double subsurface_alpha_coefficient; /* subsurface_kinematic_wave__alpha_coefficient (:88), 1.0 */
    int32_t snow_gravitational_transport; /* snow_gravitational_transport__flag (:70): lateral snow
                                             transport downhill (surface_process.jl:9-19), needs snow */
    int32_t floodplain_1d;               /* floodplain_1d__flag: a 1-D floodplain beside every river
                                            cell (surface_kinwave.jl:387-434, 572-601); needs the
                                            floodplain_* fields of wfh_static                   */
} wfh_config;

/* Static maps as the reference reads them (standard names in brackets), already in SI.
 * Pointers are host pointers and are only read during wfh_create(). */
typedef struct wfh_static {
    int64_t nx, ny;               /* raster size                                            */
    int64_t n;                    /* number of active land cells                            */
    int64_t nriv;                 /* number of river cells (= count of river_location)     */
    const int64_t* cell_index;    /* [n] linear raster index x + y*nx, strictly ascending   */
    const uint8_t* ldd;           /* [n] basin__local_drain_direction (PCRaster codes 1-9)  */
    const double* x_length;       /* [n] cell length in x [m]                               */
    const double* y_length;       /* [n] cell length in y [m]                               */
    const double* land_slope;     /* [n] land_surface__slope (clamped to >= 1e-5)           */
    const double* land_manning_n; /* [n] land_surface_water_flow__manning_n_parameter       */
    const uint8_t* river_location;/* [n] river_location__mask                               */
    const double* river_width_land;  /* [n] river__width on the land grid (river fraction)  */
    const double* river_length_land; /* [n] river__length on the land grid                  */
    const double* water_fraction; /* [n] land_water_covered__area_fraction                  */
    const double* elevation;      /* [n] land_surface__elevation                            */
    /* vegetation */
    const double* canopy_gap_fraction;   /* [n] */
    const double* max_canopy_storage;    /* [n] vegetation_water__storage_capacity [m]      */
    const double* rooting_depth;         /* [n] vegetation_root__depth [m]                  */
    const double* crop_coefficient;      /* [n] vegetation__crop_factor                     */
    const double* evap_precip_ratio;     /* [n] Gash E/R ratio (used when dt >= 23 h)       */
    /* snow (HBV) */
    const double* cfmax;          /* [n] snowpack__degree_day_coefficient [m K-1 s-1]       */
    const double* tt;             /* [n] snowfall temperature threshold [K]                 */
    const double* tti;            /* [n] snowfall temperature interval [K]                  */
    const double* ttm;            /* [n] melting temperature threshold [K]                  */
    const double* whc;            /* [n] liquid water holding capacity [-]                  */
    /* SBM soil */
    const double* theta_s, *theta_r;
    const double* theta_fc;       /* [n] or NULL: Brooks-Corey field capacity (soil_process.jl:138-160) */
    const double* kv_0;           /* [n] surface vertical Ksat [m s-1]                      */
    const double* f;              /* [n] Ksat scale parameter [m-1]                         */
    const double* air_entry_pressure; /* [n] [m]                                            */
    const double* h1, *h2, *h3_high, *h3_low, *h4, *alpha_h1; /* Feddes [m], [-]            */
    const double* soil_thickness; /* [n] [m]                                                */
    const double* infiltcap_compacted; /* [n] [m s-1]                                       */
    const double* max_leakage;    /* [n] [m s-1]                                            */
    const double* compacted_fraction; /* [n]                                                */
    const double* rootdistpar;    /* [n] wet-root sigmoid parameter [m-1]                   */
    const double* cap_hmax;       /* [n] [m]                                                */
    const double* cap_n;          /* [n]                                                    */
    const double* w_soil;         /* [n]                                                    */
    const double* cf_soil;        /* [n]                                                    */
    const double* brooks_corey_c; /* [N][n] layer-major, N = n_config_layers + 1            */
    const double* kvfrac;         /* [N][n]                                                 */
    const double* rootfraction;   /* [N][n] or NULL: derived from rooting_depth (soil.jl:526-548) */
    /* lateral subsurface flow */
    const double* khfrac;         /* [n] horizontal/vertical Ksat ratio                     */
    const double* z_exp;          /* [n] or NULL: depth [m] below which the vertical and horizontal
                                     Ksat stay constant (KvExponentialConstant / KhExponentialConstant,
                                     soil.jl:222-271, lateral_subsurface_flow.jl:73-116); NULL = the
                                     exponential profile                                       */
    /* river (nriv entries) */
    const double* river_slope;    /* river__slope (clamped to >= 1e-5)                     */
    const double* river_manning_n;
    const double* river_bankfull_depth;
    const double* river_length;
    const double* river_width;
    /* reservoirs (reservoir__flag; Wflow/src/routing/surface/reservoir.jl, network.jl:296-360),
     * nres = 0 for none. Outflow types 1 (rating curve, H-Q table per julian day), 2 (free weir,
     * optionally linked to a lower reservoir), 3 (modified Puls) and 4 (simple), or the observed outflow
     * where given; linear or interpolated (S-H table) storage curve (fields at the end). */
    int64_t nres;
    const int64_t* reservoir_outlet;       /* [nres] land index of each outlet, ascending; a river
                                              cell whose river flows on to a river cell   */
    const int32_t* reservoir_coverage;     /* [n] 1-based reservoir whose area holds the cell
                                              (reservoir_area__count), 0 = none; NULL = outlets */
    const int32_t* reservoir_outflow_type; /* [nres] ReservoirOutflowType: 1, 2, 3 or 4      */
    const double* reservoir_area;          /* [nres] [m2]                                      */
    const double* reservoir_waterlevel;    /* [nres] initial water level [m]                   */
    const double* reservoir_demand;        /* [nres] simple: [m3 s-1]                          */
    const double* reservoir_maximum_release;         /* [nres] simple: [m3 s-1]               */
    const double* reservoir_maximum_storage;         /* [nres] simple (and observed): [m3];
                                                        rating curve: derived from the tables */
    const double* reservoir_target_full_fraction;    /* [nres] simple [-]                     */
    const double* reservoir_target_minimum_fraction; /* [nres] simple [-]                     */
    const double* reservoir_threshold;               /* [nres] modified Puls, free weir [m]   */
    const double* reservoir_rating_curve_coefficient;/* [nres] modified Puls, free weir [-]   */
    const double* reservoir_external_inflow; /* [nres] or NULL (0) [m3 s-1]; negative = abstraction */
    const double* reservoir_outflow_obs;     /* [nres] or NULL (NaN = not observed) [m3 s-1]       */
    /* Storage and rating curves (ReservoirProfileType, SH and HQ tables of reservoir.jl:5-45 and
     * io.jl:1129-1151). Tables are in CSR form: reservoir k owns rows ptr[k] .. ptr[k+1]-1. H and S
     * must be finite and non-decreasing within a reservoir (interpolate_linear assumes it). */
    const int32_t* reservoir_storage_curve_type;    /* [nres] or NULL (all 1): 1 linear S = A H,
                                                       2 interpolation of the S-H table         */
    const double* reservoir_rating_curve_exponent;  /* [nres] free weir [-]                     */
    const int64_t* reservoir_sh_ptr;  /* [nres+1] or NULL: S-H rows (>= 1 for storage curve 2) */
    const double* reservoir_sh_h;     /* [rows] water level H [m]                                */
    const double* reservoir_sh_s;     /* [rows] storage S [m3]                                   */
    const int64_t* reservoir_hq_ptr;  /* [nres+1] or NULL: H-Q rows (>= 1 for outflow type 1)  */
    const double* reservoir_hq_h;     /* [rows] water level H [m]                                */
    const double* reservoir_hq_q;     /* [rows][365] outflow Q [m3 s-1], one column per julian
                                         day (read_hq_csv's Q matrix, row-major)                */
    /* 1-D floodplain (FloodPlainProfile / FloodPlainParameters, floodplain.jl:29-147, 374-402),
     * read when wfh_config.floodplain_1d is set */
    int32_t floodplain_ndepth;        /* number of flood depths (flood_depth, without 0 m)       */
    const double* floodplain_depth;   /* [ndepth] flood depths [m], ascending                    */
    const double* floodplain_volume;  /* [ndepth][nriv] cumulative floodplain volume at each depth
                                         (floodplain_water__sum_of_volume_per_depth) [m3]       */
    const double* floodplain_manning_n; /* [nriv] (floodplain_water_flow__manning_n_parameter)   */
    const double* floodplain_slope;     /* [nriv] (floodplain__slope), clamped at 1e-5           */
    /* [nres] or NULL: 1-based lower reservoir linked to a free weir (lower_reservoir_ind,
     * reservoir.jl:152-159), 0 = none. The lower reservoir's outlet must lie downstream of the
     * free weir's on the river, and at most one free weir may link to a reservoir. */
    const int32_t* reservoir_lower;
} wfh_static;

/* Readable / writable variables (BMI get_value / set_value equivalents, bmi.jl:203-262). */
typedef enum wfh_var {
    /* land states [n] */
    WFH_CANOPY_STORAGE = 0,
    WFH_SNOW_STORAGE,
    WFH_SNOW_WATER,
    WFH_SOIL_TEMPERATURE,
    WFH_SATWATERDEPTH,
    WFH_USTORELAYERDEPTH,        /* [N][n] */
    /* land fluxes / diagnostics [n] */
    WFH_WATER_TABLE_DEPTH,
    WFH_USTOREDEPTH,
    WFH_TRANSFER,
    WFH_RECHARGE,
    WFH_ACTUAL_INFILTRATION,
    WFH_INFILTRATION_EXCESS,
    WFH_SATURATION_EXCESS,
    WFH_TRANSPIRATION,
    WFH_SOIL_EVAPORATION,
    WFH_ACTUAL_CAPILLARY_FLUX,
    WFH_ACTUAL_LEAKAGE,
    WFH_ACTUAL_EVAPOTRANSPIRATION,
    WFH_INTERCEPTION_RATE,
    WFH_NET_RUNOFF,
    WFH_RUNOFF,
    WFH_NET_RUNOFF_RIVER,
    WFH_EXFILTRATION,
    WFH_TOTAL_STORAGE,
    WFH_SNOW_WATER_EQUIVALENT,
    WFH_VWC,                     /* [N][n] volumetric water content               */
    WFH_VWC_ROOT_ZONE,
    WFH_ROOT_ZONE_STORAGE,
    /* lateral subsurface flow [n] */
    WFH_SSF_Q,
    WFH_SSF_QIN_AVERAGE,
    WFH_SSF_Q_AVERAGE,
    WFH_SSF_TO_RIVER_AVERAGE,
    WFH_SSF_STORAGE,
    WFH_SSF_QMAX,
    /* overland flow [n] */
    WFH_LAND_Q,
    WFH_LAND_QIN,
    WFH_LAND_H,
    WFH_LAND_STORAGE,
    WFH_LAND_Q_AVERAGE,
    WFH_LAND_QIN_AVERAGE,
    WFH_LAND_TO_RIVER_AVERAGE,
    WFH_LAND_INWATER,
    /* river flow [nriv] */
    WFH_RIVER_Q,
    WFH_RIVER_QIN,
    WFH_RIVER_H,
    WFH_RIVER_STORAGE,
    WFH_RIVER_Q_AVERAGE,
    WFH_RIVER_QIN_AVERAGE,
    WFH_RIVER_INWATER,
    /* water balance (water_mass_balance = 1) */
    WFH_BALANCE_LAND_ERROR,      /* [n]    */
    WFH_BALANCE_LAND_RELERROR,   /* [n]    */
    WFH_BALANCE_SSF_ERROR,       /* [n]    */
    WFH_BALANCE_OVERLAND_ERROR,  /* [n]    */
    WFH_BALANCE_RIVER_ERROR,     /* [nriv] */
    /* derived parameters (read only) */
    WFH_LAND_ALPHA,              /* [n]    */
    WFH_RIVER_ALPHA,             /* [nriv] */
    WFH_THETA_FC,                /* [n]    */
    WFH_SURFACE_FLOW_WIDTH,      /* [n]    */
    WFH_FLOW_FRACTION_TO_RIVER,  /* [n]    */
    /* reservoirs [nres] (reservoir.jl ReservoirVariables / ReservoirBC) */
    WFH_RESERVOIR_STORAGE,
    WFH_RESERVOIR_WATERLEVEL,
    WFH_RESERVOIR_OUTFLOW,
    WFH_RESERVOIR_OUTFLOW_AVERAGE,
    WFH_RESERVOIR_INFLOW_AVERAGE,
    WFH_RESERVOIR_ACTEVAP_AVERAGE,       /* [m s-1] */
    WFH_RESERVOIR_ABSTRACTION_AVERAGE,   /* actual abstraction from negative external inflow */
    WFH_RESERVOIR_PRECIPITATION,         /* mean over the reservoir area [m s-1] */
    WFH_RESERVOIR_EVAPORATION,           /* mean potential evaporation [m s-1] */
    WFH_BALANCE_RESERVOIR_ERROR,
    /* lateral snow transport [n] (snow_in / snow_out of the snow model) [m s-1] */
    WFH_SNOW_IN,
    WFH_SNOW_OUT,
    /* 1-D floodplain [nriv] (FloodPlainVariables, floodplain.jl:319-341); with a floodplain,
     * WFH_RIVER_Q_AVERAGE and WFH_RIVER_QIN_AVERAGE are channel plus floodplain */
    WFH_FLOODPLAIN_STORAGE,              /* [m3] (a state) */
    WFH_FLOODPLAIN_H,                    /* flood depth [m] */
    WFH_FLOODPLAIN_Q,                    /* [m3 s-1] of the last sub-step */
    WFH_FLOODPLAIN_Q_AVERAGE,
    WFH_FLOODPLAIN_QIN_AVERAGE,
    WFH_RIVER_Q_CHANNEL_AVERAGE,         /* channel discharge alone */
    WFH_VAR_COUNT
} wfh_var;

typedef struct wfh_handle wfh_handle;

/* Model(config) for the sbm concept: derives topology (flowgraph, topological order, upstream
 * lists) and parameters, sets the cold-start states (sbm_model.jl:7-57, soil.jl:150-201,
 * utils.jl:916-937) and uploads everything to the device.
 * Replaces: Model(config, ::SbmModel)  Wflow/src/sbm_model.jl:7  */
int32_t wfh_create(const wfh_config* cfg, const wfh_static* st, wfh_handle** out);

/* Forcing for the next step (precipitation [m s-1], potential evaporation [m s-1],
 * temperature [K]). memkind WFH_MEM_HOST: host arrays in land-index order, uploaded (on a copy
 * stream, overlapping the previous step) and permuted on the device; the call returns once the
 * caller's arrays may be reused. WFH_MEM_DEVICE: device arrays in land-index order, gathered into
 * the library's buffers on the handle's stream (they must be ready, and stay valid until the next
 * step has started). WFH_MEM_DEVICE_NODE_ORDER: device arrays in the library's node order, bound
 * without a copy: they are read by every following wfh_step until the next wfh_set_forcing, so
 * they must stay alive and unmodified until then (in graph mode they are copied instead).
 * A non-finite value (NaN or infinity) makes the step's results unusable: the vertical kernel
 * flags it and wfh_sync (or wfh_get) returns WFH_ERR_NONFINITE; the handle then accepts no further
 * step until wfh_load_state restores a saved state.
 * Replaces: update_forcing!/load_dynamic_input!  Wflow/src/io.jl:108-178  */
int32_t wfh_set_forcing(wfh_handle* h, const double* precipitation, const double* potential_evaporation,
                        const double* temperature, int32_t memkind);

/* One update_model! step.  Replaces: update_model!  Wflow/src/sbm_model.jl:60-92
 * With kinematic_wave_adaptive set, every routing sub-step length depends on the state the
 * previous sub-step left (a global quantile or minimum over the cells), so the host reads each
 * one back before it launches the sub-step: the call returns when the step has completed. */
int32_t wfh_step(wfh_handle* h);

/* wfh_step without any device allocation by the library, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It has no
 * counterpart in the reference. The routing scratch of the step (wave tickets, abort flag and
 * one outflow slot per cell and sub-step) comes from a fixed device array of 64 MiB that every GPU
 * the library runs on holds until the process ends, instead of the handle's own buffers; the
 * results are those of wfh_step. Limit: (S_s + S_l) * n + S_r * max(nriv, 1) + 32 words of 8 bytes
 * must fit in 64 MiB (S_* = sub-steps per step), else WFH_ERR_ARG. Fixed sub-steps only: adaptive
 * sub-steps, a floodplain or linked free weirs return WFH_ERR_ARG. Every call on one GPU shares
 * that GPU's array and clears it first, so calls that overlap on one GPU (two handles stepped on
 * concurrently) race; serialise them. Asynchronous like wfh_step; wfh_sync reports an abort. */
int32_t wfh_step_noalloc(wfh_handle* h);

/* Graph mode (enable = 1): the fixed-sub-step step's kernel launches are recorded into a HIP
 * graph once (per starting ticket parity) and every following wfh_step replays it. No
 * counterpart in the reference. Forcing bound in node order (WFH_MEM_DEVICE_NODE_ORDER) is then
 * copied into the library's buffers; forcing bound that way before graph mode was enabled must
 * be set again (wfh_step returns WFH_ERR_STATE until it is). Results are wfh_step's. Adaptive,
 * floodplain and linked-weir models (one launch per sub-step) return WFH_ERR_ARG. */
int32_t wfh_set_graph(wfh_handle* h, int32_t enable);
/* Autotuner of the routing schedule (no counterpart in the reference): the next trials x 4 steps
 * run the candidate persistent routing grids (all resident blocks, 3/4, 1/2, 1/4 of them) in
 * turn, each timed with HIP events (those steps synchronise); the fastest is then kept. The steps
 * are real steps of the model and every grid gives the same results. trials = 0 stops. */
int32_t wfh_autotune(wfh_handle* h, int32_t trials);
/* The routing grid in use and its maximum (blocks of 256 threads), whether the autotuner is
 * still running and whether graph mode is on; any pointer may be NULL. */
int32_t wfh_get_schedule(wfh_handle* h, int32_t* route_blocks, int32_t* max_blocks, int32_t* tuning,
                         int32_t* graph_mode);
/* nsteps steps of update_model! in one launch: the time loop of run! (Wflow/src/Wflow.jl:324-372)
 * over the forcing frames P, E, T ([nsteps][n] device arrays in the library's node order,
 * wfh_get_node_order; frame k is step k's forcing). The results equal nsteps calls of
 * wfh_set_forcing + wfh_step bit for bit; on the device a chunk of cells starts its next step as
 * soon as it and the chunks downstream of it have finished the previous one, so the steps
 * overlap. river_q_avg, if not NULL, receives every step's river discharge average
 * (WFH_RIVER_Q_AVERAGE, [nsteps][nriv] in river order). Asynchronous like wfh_step (errors,
 * non-finite forcing included, at wfh_sync); the frames must stay valid until it has completed.
 * Fixed sub-steps only, without reservoirs, lateral snow transport, the exponential-constant
 * Ksat profile or a floodplain (else WFH_ERR_ARG).
 * Currently switched off: returns WFH_ERR_STATE (its first device run did not finish and the
 * cause is not established); call wfh_step per step instead.
 * Replaces: the time loop of run!  Wflow/src/Wflow.jl:324-372 (run_timestep! per step) */
int32_t wfh_run(wfh_handle* h, int32_t nsteps, const double* P, const double* E, const double* T,
                double* river_q_avg);

/* Testing only (no counterpart in the reference): sets the routing kernel's spin limit (idle polls
 * before a waiting wave raises the abort flag; 2000000 after wfh_create, at least 256) and, for
 * node >= 0, redirects the first upstream edge of that cell (node order, wfh_get_node_order) to
 * the outflow slot of a cell of another 64-cell chunk that publishes nothing, so that the cell's
 * wait can never be met. The next step's routing kernel then ends by the abort and wfh_sync
 * returns WFH_ERR_DEVICE; the handle is for destroying afterwards. For testing the bounded waits
 * from the host; node < 0 only sets the limit. */
int32_t wfh_test_route_stall(wfh_handle* h, int64_t spin_limit, int64_t node);

/* Julian day (1..365; February 29 counts as March 1) that selects the rating-curve column of
 * reservoirs with outflow type 1 in the following steps; 1 after wfh_create. The reference sets
 * it from the step's start time before the river update: julian_day(clock.time - clock.dt).
 * Replaces: update_index_hq!  Wflow/src/routing/surface/reservoir.jl:379-386 (called from
 * surface_kinwave.jl:629), julian_day  Wflow/src/utils.jl:674-678 */
int32_t wfh_set_julian_day(wfh_handle* h, int32_t julian_day);

/* Number of subsurface, overland and river sub-steps the last wfh_step() took (fixed or
 * adaptive). No counterpart in the reference (its loop counters are local). */
int32_t wfh_last_substeps(wfh_handle* h, int32_t* subsurface, int32_t* land, int32_t* river);

/* BMI.get_value / set_value (Wflow/src/bmi.jl:203-262), land-index / river-index order.
 * count = number of doubles the caller provides (n, nriv, nres or N*n). memkind WFH_MEM_HOST:
 * host memory; WFH_MEM_DEVICE: device memory in land-index order (permuted by a kernel);
 * WFH_MEM_DEVICE_NODE_ORDER: device memory in the library's node order (copied). A device source
 * must be ready when the call is made (synchronise the stream that wrote it); every variant has
 * completed when it returns. */
int32_t wfh_get(wfh_handle* h, int32_t var, double* dst, int64_t count, int32_t memkind);
int32_t wfh_set(wfh_handle* h, int32_t var, const double* src, int64_t count, int32_t memkind);

/* Topology as the reference builds it: topological order of the land and river graphs
 * (Graphs.topological_sort_by_dfs, network.jl:99/244) as 0-based node ids, and the flow-graph
 * edges (downstream node or -1 for pits).  Any pointer may be NULL.  Replaces: NetworkLand /
 * NetworkRiver fields order, graph (network.jl:48-258). */
int32_t wfh_get_topology(wfh_handle* h, int64_t* land_order, int64_t* land_down, int64_t* river_order,
                         int64_t* river_down, int64_t* river_land_index, int32_t* max_level);

/* The library's internal cell order (topological level, then land index): land_index_of_node[q]
 * is the land index stored at node q. No counterpart in the reference; it lets a caller produce
 * forcing directly in node order (WFH_MEM_DEVICE_NODE_ORDER). */
int32_t wfh_get_node_order(wfh_handle* h, int64_t* land_index_of_node);

/* Raw state checkpoint of all prognostic variables (states.jl:73-170 set).
 * wfh_state_size() gives the byte count. */
int64_t wfh_state_size(const wfh_handle* h);
int32_t wfh_save_state(wfh_handle* h, void* buf, int64_t bytes);
int32_t wfh_load_state(wfh_handle* h, const void* buf, int64_t bytes);

int32_t wfh_sync(wfh_handle* h);
/* Kernel timing: with wfh_set_timing(h, 1) every following step records HIP events on the
 * handle's stream around each kernel; wfh_get_timing() returns the mean duration (ms) of the
 * vertical and the routing kernel over the steps recorded since timing was enabled. A model
 * with fixed sub-steps, at most four layers and no lateral snow transport, floodplain or linked
 * free weirs does the vertical update inside the routing launch: its vertical time is then
 * about zero and the routing time covers the whole step. */
int32_t wfh_set_timing(wfh_handle* h, int32_t enable);
int32_t wfh_get_timing(wfh_handle* h, double* vertical_ms, double* routing_ms);
/* Algorithmic HBM bytes of one launch of each kernel for this model: every input array read
 * once and every output array written once (parameter maps stored once when uniform count
 * 8 bytes), upstream discharge read once per edge and sub-step. Counts the fixed sub-steps'
 * single routing launch (with adaptive sub-steps the routing is one launch per sub-step and
 * this count does not apply). When the routing launch does the vertical update (see
 * wfh_get_timing), vertical_bytes is 0 and routing_bytes counts both halves without any
 * read-back of what the vertical half wrote (the launch does re-read the vertical fluxes, which
 * shows as traffic above this count). No counterpart in the reference; used for
 * the roofline of the measurement harness. */
int32_t wfh_kernel_bytes(wfh_handle* h, int64_t* vertical_bytes, int64_t* routing_bytes);
const char* wfh_last_error(const wfh_handle* h);
void wfh_destroy(wfh_handle* h);

/* ---- Parameter ensembles -------------------------------------------------------------------
 * M members of one catchment that differ in three parameter maps, each scaled by a factor per
 * member: kv_0 (KsatVer), f and Manning n (land and river). No counterpart in the reference,
 * which runs one parameter set per model: a member's step is update_model! (sbm_model.jl:60-92)
 * of the model built from the scaled maps (Model(config), sbm_model.jl:7-57), and its results are
 * those of wfh_step on that model. All members are stepped together: states are held
 * member-innermost (cell v, member m at v * M + m) and the step's launches run from a HIP graph.
 * Fixed sub-steps with one overland and one subsurface sub-step per step (land_dt and
 * subsurface_dt >= dt) and at most 8 river sub-steps; no reservoirs, floodplain, snow transport
 * or exponential-constant profile. members: a multiple of 64, at most 4096; factors: [members],
 * finite and positive. */
typedef struct wfh_ensemble wfh_ensemble;
int32_t wfh_ens_create(const wfh_config* cfg, const wfh_static* st, int32_t members, const double* kv0_factor,
                       const double* f_factor, const double* n_factor, wfh_ensemble** out);
/* Forcing of the next step, shared by all members (as wfh_set_forcing; always copied). */
int32_t wfh_ens_set_forcing(wfh_ensemble* x, const double* precipitation, const double* potential_evaporation,
                            const double* temperature, int32_t memkind);
/* One update_model! of every member (asynchronous; the graph of the current schedule). */
int32_t wfh_ens_step(wfh_ensemble* x);
/* Member m's field (land-index or river-index order) of a state or river output: canopy, snow,
 * soil temperature, satwaterdepth, ustorelayerdepth, subsurface q and q_max, overland q, h and
 * alpha, river q, h, storage, q average and alpha. Synchronises. */
int32_t wfh_ens_get(wfh_ensemble* x, int32_t var, int32_t member, double* dst, int64_t count);
/* Largest |land balance error| of the last step per member ([members]; water_mass_balance). */
int32_t wfh_ens_balance(wfh_ensemble* x, double* max_abs_error);
/* Schedule of the step: threads per block (64, 128, 256), the level width from which a level is
 * a stage of its own, and cells per strand in such stages. Results do not depend on it. */
int32_t wfh_ens_set_schedule(wfh_ensemble* x, int32_t block, int32_t wide, int32_t strand_cells);
/* Autotuner: the next steps run the candidate schedules in turn, `trials` steps each, timed; the
 * fastest is then kept (trials = 0 stops tuning). The steps are real steps of the model. */
int32_t wfh_ens_set_autotune(wfh_ensemble* x, int32_t trials);
int32_t wfh_ens_get_schedule(wfh_ensemble* x, int32_t* block, int32_t* wide, int32_t* strand_cells, int32_t* stages,
                             int32_t* tuning);
int32_t wfh_ens_sync(wfh_ensemble* x);
/* Step timing (HIP events around the graph launch; synchronises every step while enabled). */
int32_t wfh_ens_set_timing(wfh_ensemble* x, int32_t enable);
int32_t wfh_ens_get_timing(wfh_ensemble* x, double* step_ms);
/* Algorithmic HBM bytes of one ensemble step (the roofline of the measurement harness). */
int32_t wfh_ens_step_bytes(wfh_ensemble* x, int64_t* bytes);
const char* wfh_ens_last_error(const wfh_ensemble* x);
void wfh_ens_destroy(wfh_ensemble* x);

/* ---- Learned Newton warm start ---------------------------------------------------------------
 * No counterpart in the reference: kinematic_wave (surface_process.jl:24-70) starts its Newton
 * iteration from u_prev = q_prev^0.2 (or cbrt(C / alpha)). These entry points train and use a
 * network that proposes a start instead: 8 features (log q_in, log q_prev, signed log of the
 * lateral inflow as discharge, log alpha, beta, log dx, log dt, log slope) -> log10 q - 5 log10
 * (reference start u), an 8-64-64-1 perceptron with BF16 weights and activations and FP32
 * accumulation on the matrix cores. The iteration and its stopping rule stay the reference's;
 * the proposed start is taken only where its residual is below the reference start's, so the
 * result is the reference's Newton from a different start (within the same 1e-12 residual). */
typedef struct wfh_mlp wfh_mlp;
int32_t wfh_mlp_create(int32_t device, uint64_t seed, wfh_mlp** out);
/* feature mean / standard deviation [8] and target mean / standard deviation */
int32_t wfh_mlp_set_normalization(wfh_mlp* m, const float* x_mean, const float* x_std, float y_mean, float y_std);
/* flat parameters [4737]: W1 [64][8], b1 [64], W2 [64][64], b2 [64], w3 [64], b3 (host floats) */
int32_t wfh_mlp_set_params(wfh_mlp* m, const float* params);
int32_t wfh_mlp_get_params(wfh_mlp* m, float* params);
/* One pass over n device samples (x [n][8], t [n] normalised targets) in minibatches of `batch`
 * (multiple of 64), an Adam step each; lr = 0: the gradient of the mean squared error over all n
 * (no update). loss: mean squared error of the pass; grad [4737] (host, may be NULL): the last
 * minibatch's gradient. */
int32_t wfh_mlp_train(wfh_mlp* m, const float* x, const float* t, int64_t n, int64_t batch, float lr, double* loss,
                      float* grad);
/* network outputs (normalised target space) for device x [n][8] into device y [n] */
int32_t wfh_mlp_predict(wfh_mlp* m, const float* x, int64_t n, float* y);
const char* wfh_mlp_last_error(const wfh_mlp* m);
void wfh_mlp_destroy(wfh_mlp* m);
/* The ensemble's kinematic-wave solves with the network's start (mlp NULL: off); a wave consults
 * the network where some member's previous discharge is below gate (INFINITY: always). */
int32_t wfh_ens_set_learned(wfh_ensemble* x, wfh_mlp* mlp, double gate);
/* Harvest training rows (features [8], target, Newton updates: 10 floats) of one plain solve in
 * `every` into the device buffer rows [cap][10] during the following steps (rows NULL: stop). */
int32_t wfh_ens_set_harvest(wfh_ensemble* x, float* rows, int64_t cap, int32_t every);
int32_t wfh_ens_harvest_count(wfh_ensemble* x, int64_t* rows);
/* Newton updates per solve since enabled: [2][16] bins (overland, river) + learned starts taken
 * [2]; enable 1 restarts, 0 stops, other values keep; out (34 values) may be NULL. */
int32_t wfh_ens_kw_histogram(wfh_ensemble* x, int32_t enable, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* WFLOW_HIP_H */
