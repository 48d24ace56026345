/*
 * wflow_oracle.h - CPU FP64 oracle of the Wflow.jl `sbm` time step.
 *
 * TEST INFRASTRUCTURE ONLY. This restatement of the reference algorithm is the checker for
 * the HIP implementation: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it. The product never links or calls it.
 */
#ifndef WFLOW_ORACLE_H
#define WFLOW_ORACLE_H

#include <stdint.h>

#include "wflow_hip.h"

#define WFO_NMAX WFH_MAX_LAYERS

/* One SBM soil column: the fields of SbmSoilParameters / SbmSoilVariables / SbmSoilBC
 * (Wflow/src/soil/soil.jl:4-224) for a single cell; N-vectors hold the SVector{N} entries. */
typedef struct wfo_soil {
    int32_t N;       /* maximum_number_of_layers */
    int32_t nl;      /* number_of_layers         */
    int32_t n_unsat; /* n_unsatlayers            */
    int32_t pad_;
    /* parameters */
    double theta_s, theta_r, theta_fc, soil_water_capacity, soil_thickness, kv_0, f;
    double air_entry_pressure, h1, h2, h3_high, h3_low, h4, alpha_h1;
    double infiltcap_soil, infiltcap_compacted, max_leakage, compacted_fraction, rootdistpar;
    double cap_hmax, cap_n, w_soil, cf_soil, rooting_depth, soil_fraction;
    double c[WFO_NMAX], kvfrac[WFO_NMAX], rootfraction[WFO_NMAX], act_thick[WFO_NMAX];
    double cum_depth[WFO_NMAX + 1];
    /* variables */
    double ustore[WFO_NMAX], ustorethick[WFO_NMAX];
    double satwaterdepth, zi, drainable, ustorecapacity, ustoredepth, total_soil_water_storage;
    double tsoil, f_infred, h3;
    /* boundary conditions */
    double water_flux_surface, pot_transpiration, pot_soilevaporation;
    /* fluxes */
    double infiltration, infiltration_excess, transfer, soil_evaporation, soil_evap_sat;
    double ae_ustore, ae_sat, transpiration, actual_infiltration, saturation_excess;
    double act_inf_soil, act_inf_compacted, excess_soil, excess_compacted;
    double capflux, leakage, recharge, aet, runoff, net_runoff, exfiltration;
    double vwc[WFO_NMAX], rvwc[WFO_NMAX], root_zone_storage, vwc_root, rvwc_root;
    double z_exp; /* KvExponentialConstant depth (INFINITY: exponential profile) */
} wfo_soil;

/* One reservoir (ReservoirParameters / ReservoirVariables / ReservoirBC of
 * Wflow/src/routing/surface/reservoir.jl). */
typedef struct wfo_reservoir {
    int32_t outflow_type; /* 1 rating curve, 2 free weir, 3 modified Puls, 4 simple */
    double area, demand, maximum_release, maximum_storage, target_full_fraction, target_minimum_fraction;
    double threshold, rating_curve_coefficient, outflow_obs, external_inflow;
    double precipitation, evaporation;           /* [m s-1] over the reservoir area */
    double inflow_overland, inflow_subsurface;   /* [m3 s-1] */
    double storage, waterlevel, outflow;
    double inflow_cumulative, outflow_cumulative, actevap_cumulative, abstraction_cumulative;
    int32_t storage_curve_type;     /* 1 linear S = A H, 2 interpolation of the S-H table */
    double rating_curve_exponent;   /* free weir */
    int64_t sh_n;                   /* S-H table rows (storage curve 2) */
    const double *sh_h, *sh_s;
    int64_t hq_n;                   /* H-Q table rows (outflow type 1) */
    const double *hq_h, *hq_q;      /* hq_q: [hq_n][365], one column per julian day */
    int32_t hq_col;                 /* col_index_hq - 1 */
    struct wfo_reservoir* lower;    /* linked lower reservoir of a free weir (lower_reservoir_ind), or NULL */
} wfo_reservoir;

/* interpolate_linear(x, xp, fp)  reservoir.jl:364-377; fp read with stride fs */
double wfo_interpolate_linear(double x, const double* xp, const double* fp, int64_t n, int64_t fs);
/* get_maximum_storage  reservoir.jl:324-340 (a reservoir with a rating curve) */
double wfo_reservoir_maximum_storage(const wfo_reservoir* r);
/* initialize_storage  reservoir.jl:342-362 for r->waterlevel */
double wfo_reservoir_initial_storage(const wfo_reservoir* r);

/* update_reservoir_model!(reservoir_model, i, inflow, dt)  reservoir.jl:583-635 */
void wfo_update_reservoir(wfo_reservoir* r, double inflow, double dt);
/* update_reservoir_model!(reservoir_model, river_flow_vars, network, v, dt)  surface_kinwave.jl:441-489
 * for river discharge q_v at the outlet; returns the outflow that becomes the downstream qin. */
double wfo_update_reservoir_at_river(wfo_reservoir* r, double q_v, double dt);

typedef struct wfo_model wfo_model;

#ifdef __cplusplus
extern "C" {
#endif

wfo_model* wfo_create(const wfh_config* cfg, const wfh_static* st, char* err, int64_t errlen);
void wfo_destroy(wfo_model* m);
int32_t wfo_step(wfo_model* m, const double* precip, const double* pet, const double* temp);
int32_t wfo_get(const wfo_model* m, int32_t var, double* dst, int64_t count);
int32_t wfo_set(wfo_model* m, int32_t var, const double* src, int64_t count);
int64_t wfo_n(const wfo_model* m);
int64_t wfo_nriv(const wfo_model* m);
int32_t wfo_nlayers(const wfo_model* m);
int32_t wfo_topology(const wfo_model* m, int64_t* land_order, int64_t* land_down, int64_t* river_order,
                     int64_t* river_down, int64_t* river_land_index);
int64_t wfo_kinwave_iterations(const wfo_model* m);
/* A planted mistake, for tests that show their checks catch it (never set otherwise): the
 * subsurface / overland / river sweep folds the upstream discharge of the sub-step before
 * (the discharge at the start of the sweep) instead of the one its upstream cells have just
 * computed. Returns 0, or -1 for an unknown code. */
enum { WFO_MUT_NONE = 0, WFO_MUT_STALE_SSF = 1, WFO_MUT_STALE_LAND = 2, WFO_MUT_STALE_RIVER = 3 };
int32_t wfo_set_mutation(wfo_model* m, int32_t mutation);
/* subsurface, overland and river sub-step counts of the last wfo_step */
void wfo_last_substeps(const wfo_model* m, int32_t* out3);
/* FloodPlainProfile  floodplain.jl:48-147 for one river cell: nd0 flood depths (ascending, without
 * 0), their cumulative volumes vol[j * stride]; out = [4][nd0 + 1] storage, width, flow area,
 * wetted perimeter, depth 0 first */
void wfo_floodplain_profile(int64_t nd0, const double* depth, const double* vol, int64_t stride, double fw, double fl,
                            double* out);
/* compute_flood_depth  floodplain.jl:316-327 on such a profile; depth0 = [0, depth...] */
double wfo_flood_depth(const double* prof, int64_t nd, const double* depth0, double flood_storage, double fl);
/* river_channel_floodplain_exchange!  surface_kinwave.jl:572-601 for one river cell with profile
 * prof ([4][nd], depth0 = [0, depths...]): out3 = floodplain storage, flood depth, exchange flux */
void wfo_floodplain_exchange(const double* prof, int64_t nd, const double* depth0, double river_storage0,
                             double fp_storage0, double bankfull_depth, double flow_width, double flow_length, double dt,
                             double* out3);
/* Manning flow capacity of update_floodplain_model!  surface_kinwave.jl:395-425 at flood depth h;
 * prof_ds = the downstream cell's profile, NULL at a pit */
double wfo_floodplain_capacity(const double* prof, const double* prof_ds, int64_t nd, const double* depth0, double h,
                               double mannings_n, double slope);
/* update_index_hq!  reservoir.jl:379-386: the rating-curve column (julian day 1..365); 0 or -1 */
int32_t wfo_set_julian_day(wfo_model* m, int32_t julian_day);
/* Statistics.quantile!(v, p) of k >= 2 values (sorts v); the kinematic-wave stable time step
 * (surface_kinwave.jl:674-704) over n cells, buf holding n values. */
double wfo_quantile(double* v, int64_t k, double p);
/* accucapacityflux! (routing/utils.jl:82-108) over n nodes in topological order `order` with
 * downstream `down` (-1 = pit); mutates material, writes flux. */
void wfo_accucapacityflux(int64_t n, const int64_t* order, const int64_t* down, double* material,
                          const double* capacity, double* flux, double dt);
double wfo_kw_stable_timestep(const double* q, const double* alpha, const double* flow_length, int64_t n,
                              double p, double* buf);

/* Element-wise evaluation of the shared elementary functions (include/wfh_math.h) for the
 * accuracy tests: op 0 exp, 1 log, 2 pow(x, y), 3 cbrt, 4 cld(x, y), 5 round(x; sigdigits=12),
 * 6 Julia min(x, y), 7 Julia max(x, y), 8 10.0^x for integral x >= 0. y is ignored by the
 * one-argument functions. */
int32_t wfo_math(int32_t op, const double* x, const double* y, double* out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
