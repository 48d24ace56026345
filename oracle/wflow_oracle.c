/*
 * wflow_oracle.c - CPU FP64 oracle of the Wflow.jl `sbm` time step (update_model!).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product never does.
 *
 * A line-by-line restatement of the Julia reference (paths relative to the reference root,
 * Wflow.jl 1.1.0-dev). Every function names the Julia it follows. It keeps the reference's
 * operation order (no FMA contraction; build with -ffp-contract=off) and Julia's semantics
 * for min/max/clamp, cld and round(; sigdigits) (include/wfh_math.h). exp/log come from
 * include/wfh_math.h, the same elementary functions the HIP kernels use, so that ulp-level
 * differences cannot change Newton or Brooks-Corey iteration counts between the two.
 *
 * Configuration covered: sbm model, kinematic wave land/river/subsurface routing with fixed or
 * adaptive sub-time steps, exponential and exponential-constant Ksat profiles, reservoirs of
 * outflow types 1-4 (linear or interpolated storage curves, linked free weirs), the 1-D river
 * floodplain, lateral snow transport, optional HBV snow and frozen-soil infiltration
 * reduction, Gash (dt >= 23 h) or modified Rutter interception with a static canopy. No
 * glacier, paddy or water demand.
 *
 * Pinned without a reference vector (tests/test_oracle_network.py):
 *   - the network sweeps in wfo_step (kinwave_subsurface_update!, kinwave_land_update!,
 *     kinwave_river_update!) by network conservation: the folded inflows equal the discharge of
 *     the upstream cells of the same sub-step summed over the network, and the network storage
 *     changes by the lateral inflow less the outflow; a planted stale fold (wfo_set_mutation)
 *     fails them while every per-cell balance still closes; a chain reaches the closed-form
 *     steady state q_k = (k + 1) q_lat dx;
 *   - wfo_soil_actual_infiltration (one layer: closed form; several: conservation and the
 *     closed form of the bottom-up cascade), wfo_soil_leakage (each limiting regime exactly),
 *     wfo_update_ustorelayerdepth (water content kept on a rise, field-capacity water added on
 *     a fall), wfo_kw_ssf_newton_raphson (the linear root in exact rationals) and
 *     wfo_kinematic_wave (inside the root's exact-rational bracket widened by its stopping
 *     residual).
 *
 * Parity unpinned (no reference vector and no closed form; checked only by line-by-line review
 * and by invariants: per-cell balance closure, restart equal to a continuous run, and results
 * that do not depend on the sweep order):
 *   - the composition wfo_soil_water_flow and wfo_step end to end beyond the pinned parts above
 *     (the land balance closes per cell, the reference's own bound);
 *   - reservoir area means (reservoir forcing in wfo_step): a left fold here, Julia's pairwise
 *     `mean` in the reference, so the last bits are unpinned for areas of more than a few cells;
 *   - in wfo_quantile, the fma and the isapprox(a, b) branch of Statistics.jl's interpolation;
 *   - exp/log/pow/cbrt (include/wfh_math.h) at the ulp level against Julia's Base.Math, which
 *     is not vendored (they are checked against libm and exact rationals instead).
 */
#include "wflow_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wfh_math.h"

#define NMAX WFO_NMAX
#define NaN (wfh_from_bits(0x7ff8000000000000ULL))

/* ------------------------------------------------------------------------------------ */
/* unit factors (Wflow/src/units.jl:55-67, 239-310)                                      */
/* ------------------------------------------------------------------------------------ */
/* from_SI(x, MM_PER_DAY): factor built in field order d^-1 then mm^1 */
static double mm_per_day_factor(void) { return (1.0 * (1.0 / 86400.0)) * 1e-3; }

/* ------------------------------------------------------------------------------------ */
/* process functions                                                                     */
/* ------------------------------------------------------------------------------------ */

/* scurve  Wflow/src/utils.jl:27-30 */
double wfo_scurve(double x, double a, double b, double c) { return 1.0 / (b + wfh_exp(-c * (x - a))); }

/* bounded_power  Wflow/src/utils.jl:1070-1076 */
static double bounded_power(double base, double power) { return base > 1.0 ? 1.0 : wfh_pow(base, power); }

/* rainfall_interception_gash  Wflow/src/vegetation/rainfall_interception.jl:9-70 */
void wfo_rainfall_interception_gash(double smax, double er, double gap, double precip, double storage,
                                    double max_evap, double dt, double* out4) {
    double throughfall, interception, stem_flow;
    if (smax > 0.0) {
        double fs, fi, psat;
        if (gap < 1.0 / 1.1) {
            fs = 0.1 * gap;
            fi = 1.0 - 1.1 * gap;
            if (er > fi)
                psat = 0.0;
            else
                psat = -smax / (er * dt) * wfh_log(1.0 - er / fi);
        } else {
            fs = 1.0 - gap;
            fi = 0.0;
            psat = 0.0;
        }
        if (precip > psat) {
            double iwet = fi * psat - smax / dt;
            double isat = er * (precip - psat);
            double idry = smax / dt;
            interception = iwet + isat + idry;
        } else {
            interception = fi * precip;
        }
        stem_flow = fs * precip;
        throughfall = precip - interception - stem_flow;
        if (interception > max_evap) {
            double drainage = interception - max_evap;
            interception = max_evap;
            throughfall += drainage;
        }
    } else {
        throughfall = precip;
        interception = 0.0;
        stem_flow = 0.0;
    }
    out4[0] = throughfall;
    out4[1] = interception;
    out4[2] = stem_flow;
    out4[3] = storage;
}

/* rainfall_interception_modrut  Wflow/src/vegetation/rainfall_interception.jl:78-130 */
void wfo_rainfall_interception_modrut(double precip, double pot_evap, double storage, double gap, double smax,
                                      double dt, double* out4) {
    double fs, pcan;
    if (gap < 1.0 / 1.1) {
        fs = 0.1 * gap;
        pcan = (1.0 - gap - fs) * precip;
    } else {
        fs = 1.0 - gap;
        pcan = 0.0;
    }
    double stemflow = fs * precip;
    double throughfall = gap * precip;
    if (storage > smax) {
        double drainage = storage - smax;
        storage = smax;
        throughfall += drainage / dt;
    }
    storage += pcan * dt;
    double max_evap = storage / dt, evap;
    if (pot_evap > max_evap) {
        evap = max_evap;
        storage = 0.0;
    } else {
        evap = pot_evap;
        storage -= evap * dt;
    }
    if (storage > smax) {
        double drainage = storage - smax;
        storage = smax;
        throughfall += drainage / dt;
    }
    out4[0] = throughfall;
    out4[1] = evap;
    out4[2] = stemflow;
    out4[3] = storage;
}

/* precipitation_hbv  Wflow/src/snow/snow_process.jl:90-116 (rfcf = sfcf = 1) */
void wfo_precipitation_hbv(double precip, double temp, double tti, double tt, double* out2) {
    double rainfrac;
    if (tti == 0.0) {
        rainfrac = temp > tt ? 1.0 : 0.0;
    } else {
        double frac = (temp - (tt - tti / 2.0)) / tti;
        rainfrac = wfh_clamp(frac, 0.0, 1.0);
    }
    double snowfrac = 1.0 - rainfrac;
    out2[0] = snowfrac * 1.0 * precip;
    out2[1] = rainfrac * 1.0 * precip;
}

/* snowpack_hbv  Wflow/src/snow/snow_process.jl:26-70 (cfr = 0.05) */
void wfo_snowpack_hbv(double ss, double sw, double sp, double lp, double temp, double ttm, double cfmax,
                      double whc, double dt, double* out5) {
    double melt;
    if (temp > ttm) {
        double pot = cfmax * (temp - ttm);
        melt = wfh_min(pot, ss / dt);
        ss -= melt * dt;
        sw += melt * dt;
    } else {
        melt = 0.0;
        double pot_refr = cfmax * 0.05 * (ttm - temp);
        double refr = wfh_min(pot_refr * dt, sw);
        ss += refr;
        sw -= refr;
    }
    ss = wfh_max(ss, 0.0);
    sw = wfh_max(sw, 0.0);
    ss += sp * dt;
    sw += lp * dt;
    double maxw = ss * whc, runoff;
    if (sw > maxw) {
        runoff = (sw - maxw) / dt;
        sw = maxw;
    } else {
        runoff = 0.0;
    }
    out5[0] = ss;
    out5[1] = sw;
    out5[2] = sw + ss;
    out5[3] = melt;
    out5[4] = runoff;
}

/* infiltration  Wflow/src/soil/soil_process.jl:16-41 */
void wfo_infiltration(double pot, double compacted, double cap_soil, double cap_compacted, double ustorecap,
                      double fred, double dt, double* out2) {
    double soilinf = pot * (1.0 - compacted);
    double pathinf = pot * compacted;
    double max_soil = wfh_min(cap_soil * fred, soilinf);
    double max_path = wfh_min(cap_compacted * fred, pathinf);
    out2[0] = wfh_min(max_path + max_soil, wfh_max(0.0, ustorecap / dt));
    out2[1] = (soilinf - max_soil) + (pathinf - max_path);
}

/* unsatzone_flow_layer  Wflow/src/soil/soil_process.jl:51-92 */
void wfo_unsatzone_flow_layer(double usd, double kv_z, double l_sat, double c, double dt, double* out2) {
    if (usd <= 0.0) {
        out2[0] = 0.0;
        out2[1] = 0.0;
        return;
    }
    double st_sat = wfh_max(0.0, usd - l_sat);
    double st = kv_z * bounded_power(usd / l_sat, c);
    double sum_ast = wfh_min(st, st_sat / dt);
    usd -= sum_ast * dt;
    double remainder = wfh_min((st - sum_ast) * dt, usd);
    int64_t its = (int64_t)wfh_cld(remainder, 2e-4);
    for (int64_t i = 0; i < its; ++i) {
        st = (kv_z / (double)its) * bounded_power(usd / l_sat, c);
        double st_max = usd / dt;
        if (st < st_max) {
            usd -= st * dt;
            sum_ast += st;
        } else {
            usd = 0.0;
            sum_ast += st_max;
            break;
        }
    }
    out2[0] = usd;
    out2[1] = sum_ast;
}

/* vwc_brooks_corey  Wflow/src/soil/soil_process.jl:99-106 */
double wfo_vwc_brooks_corey(double h, double hb, double ts, double tr, double c) {
    if (h < hb) {
        double lam = 2.0 / (c - 3.0);
        return (ts - tr) * wfh_pow(hb / h, lam) + tr;
    }
    return ts;
}

/* head_brooks_corey  Wflow/src/soil/soil_process.jl:113-130 */
double wfo_head_brooks_corey(double vwc, double ts, double tr, double c, double hb) {
    double lam = 2.0 / (c - 3.0);
    if (lam > 0.0) return hb / wfh_pow(vwc / (ts - tr), 1.0 / lam);
    return hb;
}

/* field_capacity  Wflow/src/soil/soil_process.jl:138-160 (h = -1 m) */
static double field_capacity(const double* thick, int nl, double ts, double tr, const double* c, double hb) {
    double fc = 0.0, total = 0.0;
    for (int i = 0; i < nl; ++i) {
        fc += wfo_vwc_brooks_corey(-1.0, hb, ts, tr, c[i]) * thick[i];
        total += thick[i];
    }
    return fc / total;
}

/* feddes_h3  Wflow/src/soil/soil_process.jl:166-176 */
double wfo_feddes_h3(double h3_high, double h3_low, double tpot) {
    double td = tpot / mm_per_day_factor();
    if (td <= 1.0) return h3_low;
    if (td < 5.0) return h3_low + (h3_high - h3_low) * (td - 1.0) / (5.0 - 1.0);
    return h3_high;
}

/* rwu_reduction_feddes  Wflow/src/soil/soil_process.jl:183-200 */
double wfo_rwu_reduction_feddes(double h, double h1, double h2, double h3, double h4, double alpha_h1) {
    if (h < h4) return 0.0;
    if (h < h3) return (h - h4) / (h3 - h4);
    if (alpha_h1 == 0.0) {
        if (h < h2) return 1.0;
        if (h < h1) return (h1 - h) / (h1 - h2);
        return 0.0;
    }
    return 1.0;
}

/* soil_temperature  Wflow/src/soil/soil_process.jl:210-213 */
double wfo_soil_temperature(double tprev, double w, double temp) { return tprev + w * (temp - tprev); }

/* infiltration_reduction_factor  Wflow/src/soil/soil_process.jl:229-244 */
double wfo_infiltration_reduction_factor(double tsoil, double cf, int modelsnow, int reduction) {
    if (modelsnow && reduction) {
        double bb = 1.0 / (1.0 - cf);
        return wfo_scurve(tsoil, 0.0 + 273.15, bb, 8.0) + cf;
    }
    return 1.0;
}

/* soil_evaporation_unsaturated_store  Wflow/src/soil/soil_process.jl:247-271 */
double wfo_soil_evaporation_unsaturated_store(double pot, double usd, double ust, int n_unsat, double zi,
                                              double theta_e) {
    if (n_unsat == 0) return 0.0;
    if (n_unsat == 1) return pot * wfh_min(1.0, usd / (zi * theta_e));
    return pot * wfh_min(1.0, usd / (ust * theta_e));
}

/* soil_evaporation_saturated_store  Wflow/src/soil/soil_process.jl:274-294 */
double wfo_soil_evaporation_saturated_store(double pot, int n_unsat, double thick, double zi,
                                            double theta_d, double dt) {
    if (n_unsat == 0 || n_unsat == 1) {
        double e = pot * wfh_min(1.0, (thick - zi) / thick);
        return wfh_min(e, (thick - zi) * theta_d / dt);
    }
    return 0.0;
}

/* actual_infiltration_soil_path  Wflow/src/soil/soil_process.jl:297-323 */
void wfo_actual_infiltration_soil_path(double pot, double act, double compacted, double cap_soil,
                                       double cap_compacted, double fred, double* out2) {
    double soilinf = pot * (1.0 - compacted);
    double pathinf = pot * compacted;
    if (act > 0.0) {
        double ms = wfh_min(cap_soil * fred, soilinf);
        double mp = wfh_min(cap_compacted * fred, pathinf);
        out2[0] = act * ms / (mp + ms);
        out2[1] = act * mp / (mp + ms);
    } else {
        out2[0] = 0.0;
        out2[1] = 0.0;
    }
}

/* kinematic_wave  Wflow/src/routing/surface/surface_process.jl:24-70; *iters (optional)
 * receives the number of Newton updates. */
void wfo_kinematic_wave_it(double q_in, double q_prev, double q_lat, double alpha, double dt, double dx,
                           double* out2, int64_t* iters) {
    if (q_in + q_prev + q_lat == 0.0) {
        out2[0] = 0.0;
        out2[1] = 0.0;
        return;
    }
    double dt_dx = dt / dx;
    double u_prev = q_prev >= 0.0 ? wfh_pow(q_prev, 0.2) : 0.0;
    double C = dt_dx * q_in + alpha * u_prev * u_prev * u_prev + dt * q_lat;
    double u = u_prev > 0.0 ? u_prev : wfh_cbrt(C / alpha);
    double c1 = 5.0 * dt_dx, c2 = 3.0 * alpha;
    int64_t it = 0;
    for (int k = 0; k < 3000; ++k) {
        double u2 = u * u;
        double u3 = u2 * u;
        double f_u = u3 * (dt_dx * u2 + alpha) - C;
        if (wfh_fabs(f_u) <= 1.0e-12) break;
        double df = u2 * (c1 * u2 + c2);
        u -= f_u / df;
        ++it;
        if (wfh_isnan(u) || u <= 0.0) u = WFH_KIN_WAVE_MIN_FLOW_QROOT;
    }
    if (iters) *iters += it;
    u = wfh_max(u, WFH_KIN_WAVE_MIN_FLOW_QROOT);
    double u3 = u * u * u;
    out2[0] = u3 * u * u;
    out2[1] = alpha * u3;
}
void wfo_kinematic_wave(double q_in, double q_prev, double q_lat, double alpha, double dt, double dx,
                        double* out2) {
    wfo_kinematic_wave_it(q_in, q_prev, q_lat, alpha, dt, dx, out2, NULL);
}

/* ssf_celerity  Wflow/src/routing/subsurface/subsurface_process.jl:6-21 (KhExponential);
 * :27-41 (KhExponentialConstant) when z_exp is finite. */
double wfo_ssf_celerity(double zi, double slope, double sy, double kh0, double f, double z_exp) {
    double z = zi;
    if (!wfh_isnan(z_exp)) z = zi < z_exp ? zi : z_exp;
    return (kh0 * wfh_exp(-f * z) * slope) / sy;
}

/* kw_ssf_newton_raphson  Wflow/src/routing/subsurface/subsurface_process.jl:57-78 */
double wfo_kw_ssf_newton_raphson(double q, double C, double cel, double dt, double dx) {
    int count = 0;
    double dt_dx = dt / dx;
    double cinv = 1.0 / cel;
    double df = dt_dx + cinv;
    for (;;) {
        double r = dt_dx * q + cinv * q - C;
        q -= (r / df);
        if (wfh_isnan(q)) q = 0.0;
        q = wfh_max(q, WFH_KIN_WAVE_MIN_FLOW);
        if (wfh_fabs(r) <= 1.0e-12 || count >= 3000) break;
        count += 1;
    }
    return q;
}

/* set_layerthickness  Wflow/src/utils.jl:390-404 */
static void set_layerthickness(double ref, const double* cum, const double* thick, int N, double* out) {
    for (int i = 0; i < N; ++i) out[i] = thick[i] * NaN;
    for (int i = 0; i < N; ++i) {
        if (ref > cum[i + 1])
            out[i] = thick[i];
        else if (ref - cum[i] > 0.0)
            out[i] = ref - cum[i];
    }
}

/* number_of_active_layers  Wflow/src/utils.jl:406-409 */
static int number_of_active_layers(const double* t, int N) {
    int k = N;
    for (int i = 0; i < N; ++i)
        if (wfh_isnan(t[i])) --k;
    return k;
}

/* lower_bound_drainable_porosity  Wflow/src/utils.jl:1134-1136 */
static double lower_bound_drainable_porosity(double ts, double tfc) { return wfh_max(ts - tfc, 0.02); }

/* sum(@view v[1:k]): sequential left fold (Base.mapreduce_impl for short arrays) */
static double sum_first(const double* v, int k) {
    if (k <= 0) return 0.0;
    double s = v[0];
    for (int i = 1; i < k; ++i) s = s + v[i];
    return s;
}

/* hydraulic_conductivity_at_depth(::KvExponential / ::KvExponentialConstant)  Wflow/src/utils.jl:
 * 727-762 (m is 0-based; z_exp = INFINITY for the exponential profile) */
static double kv_at_depth(const wfo_soil* s, double z, int m) {
    return s->kvfrac[m] * s->kv_0 * wfh_exp(-s->f * (z < s->z_exp ? z : s->z_exp));
}

/* water_table_change  Wflow/src/utils.jl:1090-1131 */
void wfo_water_table_change(const wfo_soil* s, double net_flux, double sy, double dt, double* out2) {
    double theta_e = s->theta_s - s->theta_r;
    double dh;
    if (net_flux <= 0.0) {
        dh = net_flux * dt / sy;
    } else {
        dh = 0.0;
        for (int k = s->n_unsat - 1; k >= 0; --k) {
            double cap = wfh_max(s->ustorethick[k] * theta_e - s->ustore[k], 0.0) / dt;
            double fl = wfh_min(net_flux, cap);
            if (cap <= net_flux) {
                dh += s->ustorethick[k];
            } else {
                double syd = theta_e - (s->ustore[k] / s->ustorethick[k]);
                dh += fl * dt / syd;
            }
            net_flux -= fl;
            if (net_flux == 0.0) break;
        }
    }
    out2[0] = dh;
    out2[1] = wfh_max(net_flux, 0.0);
}

/* update_ustorelayerdepth!(soil, zi_prev, water_table_depth, i)  Wflow/src/soil/soil.jl:1213-1259 */
void wfo_update_ustorelayerdepth(wfo_soil* s, double zi_prev, double zi) {
    const int N = s->N;
    int nprev = s->n_unsat;
    double tprev[NMAX];
    memcpy(tprev, s->ustorethick, sizeof tprev);
    double t[NMAX];
    set_layerthickness(zi, s->cum_depth, s->act_thick, N, t);
    int nnew = number_of_active_layers(t, N);
    double u[NMAX];
    memcpy(u, s->ustore, sizeof u);
    if (zi < zi_prev) {
        for (int k = nnew; k <= nprev; ++k) { /* 1-based k as in the reference */
            if (k == 0) continue;
            if (wfh_isnan(t[k - 1]))
                u[k - 1] = 0.0;
            else
                u[k - 1] = (t[k - 1] / tprev[k - 1]) * u[k - 1];
        }
    } else {
        for (int k = nprev; k <= nnew; ++k) {
            if (k == 0) continue;
            double th_prev = wfh_isnan(tprev[k - 1]) ? 0.0 : tprev[k - 1];
            double dthick = t[k - 1] - th_prev;
            u[k - 1] = u[k - 1] + dthick * (s->theta_fc - s->theta_r);
        }
    }
    s->n_unsat = nnew;
    memcpy(s->ustore, u, sizeof u);
    memcpy(s->ustorethick, t, sizeof t);
    s->zi = zi;
}

/* kinematic_wave_ssf (KhExponential / KhExponentialConstant)
 * Wflow/src/routing/subsurface/subsurface_process.jl:89-172.  out4 = (q, zi, exfilt, net_flux) */
void wfo_kinematic_wave_ssf(double q_in, double q_prev, double zi_prev, double q_net_bnds, double slope,
                            double sy, double d, double dt, double dx, double dw, double q_max, double kh0,
                            double f, double z_exp, wfo_soil* soil, double* out4) {
    if (q_in + q_prev == 0.0 && q_net_bnds <= 0.0) {
        out4[0] = 0.0;
        out4[1] = d;
        out4[2] = 0.0;
        out4[3] = 0.0;
        return;
    }
    double q = (q_prev + q_in) / 2.0;
    double cel = wfo_ssf_celerity(zi_prev, slope, sy, kh0, f, z_exp);
    double C = (dt / dx) * (q_in + q_net_bnds) + q_prev / cel;
    q = wfo_kw_ssf_newton_raphson(q, C, cel, dt, dx);
    q = wfh_min(q, (q_max * dw));
    double net_flux = (q_in + q_net_bnds - q) / (dw * dx);
    double w[2];
    wfo_water_table_change(soil, net_flux, sy, dt, w);
    double exfilt = w[1];
    double zi = zi_prev - w[0];
    if (zi > d) {
        double q_excess = (dw * dx) * sy * (zi - d) / dt;
        q = wfh_max(q - q_excess, WFH_KIN_WAVE_MIN_FLOW);
    }
    zi = wfh_clamp(zi, 0.0, d);
    double its_f = WFH_CEIL(wfh_round_sig12(wfh_fabs(zi - zi_prev) / 0.1));
    int64_t its = (int64_t)its_f;
    if (its > 1) {
        double dt_s = dt / (double)its;
        double q_sum = 0.0, exfilt_sum = 0.0, net_flux_sum = 0.0;
        for (int64_t i = 0; i < its; ++i) {
            cel = wfo_ssf_celerity(zi_prev, slope, sy, kh0, f, z_exp);
            C = (dt_s / dx) * q_in + q_prev / cel + q_net_bnds * (dt_s / dx);
            q = wfo_kw_ssf_newton_raphson(q_prev, C, cel, dt_s, dx);
            q = wfh_min(q, (q_max * dw));
            net_flux = (q_in + q_net_bnds - q) / (dw * dx);
            wfo_water_table_change(soil, net_flux, sy, dt_s, w);
            exfilt = w[1];
            zi = zi_prev - w[0];
            if (zi > d) {
                double q_excess = (dw * dx) * sy * (zi - d) / dt_s;
                q = wfh_max(q - q_excess, WFH_KIN_WAVE_MIN_FLOW);
            }
            zi = wfh_clamp(zi, 0.0, d);
            wfo_update_ustorelayerdepth(soil, zi_prev, zi);
            exfilt_sum += exfilt;
            net_flux_sum += net_flux;
            q_sum += q;
            q_prev = q;
            zi_prev = zi;
        }
        q = q_sum / (double)its;
        exfilt = exfilt_sum / (double)its;
        net_flux = net_flux_sum / (double)its;
    } else {
        wfo_update_ustorelayerdepth(soil, zi_prev, zi);
    }
    out4[0] = q;
    out4[1] = zi;
    out4[2] = exfilt;
    out4[3] = net_flux;
}

/* compute_mass_balance_error  Wflow/src/mass_balance.jl:84-93 */
void wfo_compute_mass_balance_error(double tin, double tout, double rate, double* out2) {
    double err = tin - tout - rate;
    double avg = (tin + tout) / 2.0;
    out2[0] = err;
    out2[1] = avg == 0.0 ? 0.0 : err / avg;
}

/* ------------------------------------------------------------------------------------ */
/* SBM soil model, one column (Wflow/src/soil/soil.jl)                                   */
/* ------------------------------------------------------------------------------------ */

/* update_diagnostic_vars!  soil.jl:1400-1436 */
/* ---- 1-D floodplain (Wflow/src/routing/surface/floodplain.jl) ---- */

/* interpolation_indices  floodplain.jl:272-284: i1 = last index with v[i] <= x (first index when
 * none), i2 = i1 + 1 except at the end; 0-based */
static void fp_interpolation_indices(double x, const double* v, int64_t n, int64_t* i1, int64_t* i2) {
    int64_t a = 0;
    for (int64_t i = 0; i < n; ++i)
        if (v[i] <= x) a = i;
    *i1 = a;
    *i2 = a == n - 1 ? a : a + 1;
}

void wfo_floodplain_profile(int64_t nd0, const double* depth, const double* vol, int64_t stride, double fw, double fl,
                            double* out) {
    const int64_t nd = nd0 + 1;
    double *S = out, *W = S + nd, *A = W + nd, *P = A + nd;
    double* seg = (double*)calloc((size_t)nd, sizeof(double));
    for (int64_t j = 0; j < nd; ++j) S[j] = W[j] = A[j] = P[j] = 0.0;
    W[0] = fw;
    for (int64_t j = 0; j < nd - 1; ++j) {
        const double h = depth[j] - (j ? depth[j - 1] : 0.0);
        const double diff_storage = vol[j * stride] - (j ? vol[(j - 1) * stride] : 0.0);
        W[j + 1] = diff_storage / (h * fl);
        if (W[j + 1] < W[j]) W[j + 1] = W[j]; /* width constant or increasing (the reference warns) */
        A[j + 1] = W[j + 1] * h;
        P[j + 1] = (W[j + 1] - W[j]) + 2.0 * h;
        seg[j + 1] = A[j + 1] * fl;
        if (j == 0) P[0] = P[1] - 2.0 * h;
    }
    for (int64_t j = 2; j < nd; ++j) P[j] = P[j - 1] + P[j];   /* cumsum of rows 2:end */
    for (int64_t j = 1; j < nd; ++j) A[j] = A[j - 1] + A[j];   /* cumsum */
    S[0] = seg[0];
    for (int64_t j = 1; j < nd; ++j) S[j] = S[j - 1] + seg[j]; /* cumsum of the segments */
    free(seg);
}

double wfo_flood_depth(const double* prof, int64_t nd, const double* depth0, double flood_storage, double fl) {
    const double *S = prof, *W = prof + nd;
    int64_t i1, i2;
    fp_interpolation_indices(flood_storage, S, nd, &i1, &i2);
    const double delta_A = (flood_storage - S[i1]) / fl;
    const double delta_h = delta_A / W[i2];
    return depth0[i1] + delta_h;
}

/* compute_floodplain_flow_area  floodplain.jl:330-341 with compute_flood_flow_area :286-300 */
static double fp_flow_area(const double* W, const double* A, const double* depth, double h, int64_t i1, int64_t i2) {
    const double channel_area = W[0] * h;
    const double delta_h = h - depth[i1];
    const double flood_flow_area = A[i1] + (W[i2] * delta_h);
    return wfh_max(flood_flow_area - channel_area, 0.0);
}

/* ---- reservoirs (Wflow/src/routing/surface/reservoir.jl), linear storage curve ---- */

/* update_reservoir_simple  reservoir.jl:389-421; inflow, precipitation, evaporation in m3 s-1 */
static void reservoir_simple(const wfo_reservoir* r, double inflow, double precipitation, double evaporation,
                             double dt, double* outflow, double* storage_out) {
    double storage = r->storage + (inflow + precipitation - evaporation) * dt;
    storage = wfh_max(storage, 0.0);
    double fill_fraction = storage / r->maximum_storage;
    double fac = wfo_scurve(fill_fraction, r->target_minimum_fraction, 1.0, 30.0);
    double demand_release = wfh_min(fac * r->demand, storage / dt);
    storage -= demand_release * dt;
    double release_wanted = wfh_max(0.0, (storage - r->maximum_storage * r->target_full_fraction) / dt);
    double overflow_q = wfh_max(0.0, (storage - r->maximum_storage) / dt);
    double release_realized = wfh_min(release_wanted, overflow_q + r->maximum_release - demand_release);
    storage -= release_realized * dt;
    *outflow = release_realized + demand_release;
    *storage_out = storage;
}

/* update_reservoir_modified_puls  reservoir.jl:427-456 */
static void reservoir_modified_puls(const wfo_reservoir* r, double inflow, double precipitation, double evaporation,
                                    double dt, double* outflow_out, double* storage_out) {
    double res_factor = r->area / (dt * sqrt(r->rating_curve_coefficient));
    double si_factor = r->storage / dt + precipitation - evaporation + inflow;
    double si_factor_adj = si_factor - r->area * r->threshold / dt;
    double outflow;
    if (si_factor_adj > 0.0) {
        double quadratic_sol_term = -res_factor + sqrt((res_factor * res_factor + 4.0 * si_factor_adj));
        outflow = quadratic_sol_term > 0.0 ? 0.25 * (quadratic_sol_term * quadratic_sol_term) : 0.0;
    } else {
        outflow = 0.0;
    }
    outflow = wfh_min(outflow, si_factor);
    *outflow_out = outflow;
    *storage_out = (si_factor - outflow) * dt;
}

/* interpolate_linear  reservoir.jl:364-377, literally: the ends of the table give minimum(fp) /
 * maximum(fp); otherwise i1 is the last index with xp[i1] <= x (findall), i2 = i1 + 1. A NaN x
 * (no such index: the reference raises) gives NaN. */
double wfo_interpolate_linear(double x, const double* xp, const double* fp, int64_t n, int64_t fs) {
    double xmin = xp[0], xmax = xp[0], fmin = fp[0], fmax = fp[0];
    for (int64_t i = 1; i < n; ++i) {
        xmin = wfh_min(xmin, xp[i]);
        xmax = wfh_max(xmax, xp[i]);
        fmin = wfh_min(fmin, fp[i * fs]);
        fmax = wfh_max(fmax, fp[i * fs]);
    }
    if (x <= xmin) return fmin;
    if (x >= xmax) return fmax;
    int64_t i1 = -1;
    for (int64_t i = 0; i < n; ++i)
        if (xp[i] <= x) i1 = i;
    if (i1 < 0 || i1 + 1 >= n) return NaN;
    const int64_t i2 = i1 + 1;
    const double f1 = fp[i1 * fs], f2 = fp[i2 * fs];
    return f1 * (1.0 - (x - xp[i1]) / (xp[i2] - xp[i1])) + f2 * (x - xp[i1]) / (xp[i2] - xp[i1]);
}

double wfo_reservoir_maximum_storage(const wfo_reservoir* r) {
    double hmax = r->hq_h[0];
    for (int64_t i = 1; i < r->hq_n; ++i) hmax = wfh_max(hmax, r->hq_h[i]);
    if (r->storage_curve_type == 2) return wfo_interpolate_linear(hmax, r->sh_h, r->sh_s, r->sh_n, 1);
    return r->area * hmax;
}

double wfo_reservoir_initial_storage(const wfo_reservoir* r) {
    if (r->storage_curve_type == 2) return wfo_interpolate_linear(r->waterlevel, r->sh_h, r->sh_s, r->sh_n, 1);
    return r->area * r->waterlevel;
}

/* update_reservoir_hq  reservoir.jl:457-483: the outflow of the rating curve column of the day at
 * the current water level, then the overflow above the maximum storage */
static void reservoir_hq(const wfo_reservoir* r, double inflow, double precipitation, double evaporation,
                         double dt, double* outflow_out, double* storage_out) {
    const double storage_input = r->storage / dt + precipitation - evaporation + inflow;
    double outflow = wfo_interpolate_linear(r->waterlevel, r->hq_h, r->hq_q + r->hq_col, r->hq_n, 365);
    outflow = wfh_min(outflow, storage_input);
    double storage = (storage_input - outflow) * dt;
    const double overflow = wfh_max(0.0, (storage - r->maximum_storage) / dt);
    storage = wfh_min(storage, r->maximum_storage);
    outflow += overflow;
    *outflow_out = outflow;
    *storage_out = storage;
}

/* update_reservoir_free_weir  reservoir.jl:485-553. With a linked lower reservoir whose water
 * level is above this one's, the flow reverses (negative outflow, from the lower reservoir) and
 * the lower reservoir's storage, water level, outflow and cumulative outflow are updated in
 * place. */
static void reservoir_free_weir(const wfo_reservoir* r, double inflow, double precipitation, double evaporation,
                                double dt, double* outflow_out, double* storage_out) {
    wfo_reservoir* lo = r->lower;
    const double diff_wl = lo ? r->waterlevel - lo->waterlevel : 0.0;
    const double storage_input = wfh_max(r->storage / dt + precipitation - evaporation + inflow, 0.0);
    double outflow;
    if (diff_wl >= 0.0) {
        if (r->waterlevel > r->threshold) {
            const double dh = r->waterlevel - r->threshold;
            outflow = r->rating_curve_coefficient * wfh_pow(dh, r->rating_curve_exponent);
            const double maxflow = dh * r->area / dt;
            outflow = wfh_min(outflow, maxflow);
        } else {
            outflow = 0.0;
        }
    } else {
        if (lo->waterlevel > r->threshold) {
            const double dh = lo->waterlevel - r->threshold;
            outflow = -r->rating_curve_coefficient * wfh_pow(dh, r->rating_curve_exponent);
            const double maxflow = dh * lo->area / dt;
            outflow = wfh_max(outflow, -maxflow);
        } else {
            outflow = 0.0;
        }
    }
    const double storage = (storage_input - outflow) * dt;
    if (diff_wl < 0.0) {
        const double lower_res_storage = lo->storage + outflow * dt;
        const double lower_res_waterlevel =
            lo->storage_curve_type == 2 ? wfo_interpolate_linear(lower_res_storage, lo->sh_s, lo->sh_h, lo->sh_n, 1)
                                        : lo->waterlevel + (lower_res_storage - lo->storage) / lo->area;
        lo->outflow = -outflow;
        lo->outflow_cumulative -= outflow * dt;
        lo->storage = lower_res_storage;
        lo->waterlevel = lower_res_waterlevel;
    }
    *outflow_out = outflow;
    *storage_out = storage;
}

/* update_reservoir_outflow_obs  reservoir.jl:555-575 */
static void reservoir_outflow_obs(const wfo_reservoir* r, double inflow, double precipitation, double evaporation,
                                  double dt, double* outflow_out, double* storage_out) {
    double storage_input = wfh_max(r->storage / dt + precipitation - evaporation + inflow, 0.0);
    double outflow = wfh_min(r->outflow_obs, storage_input);
    double storage = (storage_input - outflow) * dt;
    if (!wfh_isnan(r->maximum_storage)) {
        double overflow = wfh_max(0.0, (storage - r->maximum_storage) / dt);
        storage -= overflow * dt;
        outflow += overflow;
    }
    *outflow_out = outflow;
    *storage_out = storage;
}

void wfo_update_reservoir(wfo_reservoir* r, double inflow, double dt) {
    double precipitation = r->precipitation * r->area;
    double available_storage = r->storage + (inflow + precipitation) * dt;
    double potential_evaporation = r->evaporation * r->area;
    double evaporation = wfh_min(available_storage / dt, potential_evaporation);
    double outflow, storage;
    if (!wfh_isnan(r->outflow_obs))
        reservoir_outflow_obs(r, inflow, precipitation, evaporation, dt, &outflow, &storage);
    else if (r->outflow_type == 1)
        reservoir_hq(r, inflow, precipitation, evaporation, dt, &outflow, &storage);
    else if (r->outflow_type == 2)
        reservoir_free_weir(r, inflow, precipitation, evaporation, dt, &outflow, &storage);
    else if (r->outflow_type == 3)
        reservoir_modified_puls(r, inflow, precipitation, evaporation, dt, &outflow, &storage);
    else
        reservoir_simple(r, inflow, precipitation, evaporation, dt, &outflow, &storage);
    double waterlevel = r->storage_curve_type == 2
                            ? wfo_interpolate_linear(storage, r->sh_s, r->sh_h, r->sh_n, 1)
                            : r->waterlevel + (storage - r->storage) / r->area;
    r->storage = storage;
    r->waterlevel = waterlevel;
    r->outflow = outflow;
    r->inflow_cumulative += inflow * dt;
    r->outflow_cumulative += outflow * dt;
    r->actevap_cumulative += evaporation / r->area * dt;
}

double wfo_update_reservoir_at_river(wfo_reservoir* r, double q_v, double dt) {
    double inflow_ext = r->external_inflow, inflow;
    if (inflow_ext < 0.0) {
        double abstraction = wfh_min(-inflow_ext, (r->storage / dt) * 0.98);
        r->abstraction_cumulative += abstraction * dt;
        inflow = -abstraction;
    } else {
        inflow = inflow_ext;
    }
    double net_inflow = q_v + r->inflow_overland + r->inflow_subsurface + inflow;
    wfo_update_reservoir(r, net_inflow, dt);
    return r->outflow;
}

void wfo_soil_update_diagnostic_vars(wfo_soil* s) {
    s->ustoredepth = sum_first(s->ustore, s->nl);
    s->zi = wfh_max(0.0, s->soil_thickness - s->satwaterdepth / (s->theta_s - s->theta_r));
    s->drainable = (s->soil_thickness - s->zi) * lower_bound_drainable_porosity(s->theta_s, s->theta_fc);
    s->ustorecapacity = s->soil_water_capacity - s->satwaterdepth - s->ustoredepth;
    set_layerthickness(s->zi, s->cum_depth, s->act_thick, s->N, s->ustorethick);
    s->n_unsat = number_of_active_layers(s->ustorethick, s->N);
    s->total_soil_water_storage = s->satwaterdepth + s->ustoredepth;
}

/* update_bc_soil_model!  soil.jl:658-682 (no paddy, no irrigation). wfs_runoff, runoff_river
 * and runoff_land come from the open-water runoff model. */
void wfo_soil_update_bc(wfo_soil* s, double canopy_potevap, double interception_rate, double pet,
                        double wfs_runoff, double runoff_river, double runoff_land, double dt) {
    (void)dt;
    s->pot_transpiration = wfh_max(0.0, canopy_potevap - interception_rate);
    s->pot_soilevaporation = s->soil_fraction * pet;
    s->pot_soilevaporation = s->pot_soilevaporation - 0.0;
    s->water_flux_surface = wfh_max(((wfs_runoff + 0.0) - runoff_river) - runoff_land + 0.0 / dt, 0.0);
}

/* unsaturated_zone_flow!  soil.jl:764-804 */
void wfo_soil_unsaturated_zone_flow(wfo_soil* s, double dt) {
    if (s->n_unsat > 0) {
        double z[NMAX];
        z[0] = s->ustorethick[0];
        for (int k = 1; k < s->N; ++k) z[k] = z[k - 1] + s->ustorethick[k];
        double flow_rate = 0.0;
        for (int m = 0; m < s->n_unsat; ++m) {
            double l_sat = s->ustorethick[m] * (s->theta_s - s->theta_r);
            double kv_z = kv_at_depth(s, z[m], m);
            double usd = (m == 0) ? s->ustore[m] + s->infiltration * dt : s->ustore[m] + flow_rate * dt;
            double o[2];
            wfo_unsatzone_flow_layer(usd, kv_z, l_sat, s->c[m], dt, o);
            s->ustore[m] = o[0];
            flow_rate = o[1];
        }
        s->transfer = flow_rate;
    } else {
        s->transfer = 0.0;
    }
}

/* soil_evaporation!  soil.jl:814-856 */
void wfo_soil_evaporation(wfo_soil* s, double dt) {
    double pot = s->pot_soilevaporation;
    double seu = wfo_soil_evaporation_unsaturated_store(pot, s->ustore[0], s->ustorethick[0], s->n_unsat, s->zi,
                                                        s->theta_s - s->theta_r);
    seu = wfh_min(seu, s->ustore[0] / dt);
    pot -= seu;
    s->ustore[0] = s->ustore[0] - seu * dt;
    double thd = lower_bound_drainable_porosity(s->theta_s, s->theta_fc);
    double ses = wfo_soil_evaporation_saturated_store(pot, s->n_unsat, s->act_thick[0], s->zi, thd, dt);
    s->soil_evap_sat = ses;
    s->soil_evaporation = seu + ses;
    s->drainable -= ses * dt;
}

/* transpiration!  soil.jl:865-975 */
void wfo_soil_transpiration(wfo_soil* s, double dt) {
    const double rd = s->rooting_depth;
    const double pt = s->pot_transpiration;
    s->h3 = wfo_feddes_h3(s->h3_high, s->h3_low, pt);
    double sum_rf = 0.0, rf_lowest = 0.0;
    for (int k = 0; k < s->n_unsat; ++k) {
        double rfu;
        if (k == s->n_unsat - 1 && s->zi < rd) {
            double rootlength = wfh_min(s->act_thick[k], rd - s->cum_depth[k]);
            rfu = s->rootfraction[k] * (s->ustorethick[k] / rootlength);
        } else {
            rfu = s->rootfraction[k];
        }
        sum_rf += rfu;
        rf_lowest = rfu;
    }
    double actevap = 0.0;
    for (int k = 0; k < s->n_unsat; ++k) {
        double rfu = (k < s->n_unsat - 1) ? s->rootfraction[k] : rf_lowest;
        double rfs = rd > 0.0 ? wfh_max((1.0 / sum_rf), 1.0) * rfu : 0.0;
        double vwc = wfh_max(s->ustore[k] / s->ustorethick[k], 1e-7);
        double head = wfo_head_brooks_corey(vwc, s->theta_s, s->theta_r, s->c[k], s->air_entry_pressure);
        double alpha = wfo_rwu_reduction_feddes(head, s->h1, s->h2, s->h3, s->h4, s->alpha_h1);
        double availcap = wfh_min(1.0, wfh_max(0.0, (rd - s->cum_depth[k]) / s->ustorethick[k]));
        double maxextr = s->ustore[k] * availcap / dt;
        double layer = wfh_min(alpha * rfs * pt, maxextr);
        s->ustore[k] = s->ustore[k] - layer * dt;
        actevap += layer;
    }
    double wetroots = wfo_scurve(s->zi, rd, 1.0, s->rootdistpar);
    double alpha = wfo_rwu_reduction_feddes(0.0, s->h1, s->h2, s->h3, s->h4, s->alpha_h1);
    double rest = pt - actevap;
    double aes = wfh_min(rest * wetroots * alpha, s->drainable / dt);
    s->ae_ustore = actevap;
    s->ae_sat = aes;
    s->drainable -= aes * dt;
    s->transpiration = actevap + aes;
}

/* actual_infiltration!  soil.jl:987-1017 */
void wfo_soil_actual_infiltration(wfo_soil* s, double dt) {
    double excess = 0.0;
    for (int k = s->n_unsat - 1; k >= 0; --k) {
        excess = wfh_max(0.0, s->ustore[k] - s->ustorethick[k] * (s->theta_s - s->theta_r));
        s->ustore[k] = s->ustore[k] - excess;
        if (k > 0) s->ustore[k - 1] = s->ustore[k - 1] + excess;
    }
    s->actual_infiltration = s->infiltration - excess / dt;
}

/* capillary_flux!  soil.jl:1050-1111 */
void wfo_soil_capillary_flux(wfo_soil* s, double dt) {
    if (s->n_unsat > 0) {
        double ksat = kv_at_depth(s, s->zi, s->n_unsat - 1);
        double mx = wfh_max(0.0, wfh_min(wfh_min(wfh_min(ksat, s->ae_ustore), s->ustorecapacity / dt),
                                         s->drainable / dt));
        double capflux;
        if (s->zi > s->rooting_depth)
            capflux = mx * wfh_pow(1.0 - wfh_min(s->zi, s->cap_hmax) / (s->cap_hmax), s->cap_n);
        else
            capflux = 0.0;
        double net = capflux, act = 0.0;
        for (int k = s->n_unsat - 1; k >= 0; --k) {
            double toadd = wfh_min(
                net, wfh_max((s->ustorethick[k] * (s->theta_s - s->theta_r) - s->ustore[k]) / dt, 0.0));
            s->ustore[k] = s->ustore[k] + toadd * dt;
            net -= toadd;
            act += toadd;
        }
        s->capflux = act;
    } else {
        s->capflux = 0.0;
    }
}

/* leakage!  soil.jl:1118-1136 */
void wfo_soil_leakage(wfo_soil* s, double dt) {
    double deepksat = kv_at_depth(s, s->soil_thickness, s->nl - 1);
    double deeptransfer = wfh_min(s->drainable / dt, deepksat);
    s->leakage = wfh_max(0.0, wfh_min(s->max_leakage, deeptransfer));
}

/* update_soil_water_flow!  soil.jl:1150-1211 (AET without interception; the caller adds it,
 * sbm.jl:130). temperature is the air temperature forcing. */
void wfo_soil_water_flow(wfo_soil* s, double temperature, int modelsnow, int reduction, double ae_openw_r,
                         double ae_openw_l, double dt) {
    wfo_soil_update_diagnostic_vars(s);
    if (modelsnow) s->tsoil = wfo_soil_temperature(s->tsoil, s->w_soil, temperature);
    s->f_infred = wfo_infiltration_reduction_factor(s->tsoil, s->cf_soil, modelsnow, reduction);
    double o[2];
    wfo_infiltration(s->water_flux_surface, s->compacted_fraction, s->infiltcap_soil, s->infiltcap_compacted,
                     s->ustorecapacity, s->f_infred, dt, o);
    s->infiltration = o[0];
    s->infiltration_excess = o[1];
    wfo_soil_unsaturated_zone_flow(s, dt);
    wfo_soil_evaporation(s, dt);
    wfo_soil_transpiration(s, dt);
    wfo_soil_actual_infiltration(s, dt);
    s->saturation_excess = (s->water_flux_surface - s->actual_infiltration) - s->infiltration_excess;
    wfo_actual_infiltration_soil_path(s->water_flux_surface, s->actual_infiltration, s->compacted_fraction,
                                      s->infiltcap_soil, s->infiltcap_compacted, s->f_infred, o);
    s->act_inf_soil = o[0];
    s->act_inf_compacted = o[1];
    s->excess_soil = wfh_max(s->water_flux_surface * (1.0 - s->compacted_fraction) - s->act_inf_soil, 0.0);
    s->excess_compacted = wfh_max(s->water_flux_surface * s->compacted_fraction - s->act_inf_compacted, 0.0);
    s->ustoredepth = sum_first(s->ustore, s->nl);
    s->ustorecapacity = s->soil_water_capacity - s->satwaterdepth - s->ustoredepth;
    wfo_soil_capillary_flux(s, dt);
    wfo_soil_leakage(s, dt);
    s->recharge = (s->transfer - s->capflux - s->leakage - s->ae_sat - s->soil_evap_sat);
    s->aet = s->soil_evaporation + s->transpiration + ae_openw_r + ae_openw_l + 0.0;
}

/* update_soil_water_storage! (per cell body)  soil.jl:1294-1392; exfilt = exfiltwater_average */
void wfo_soil_water_storage(wfo_soil* s, double exfilt, double runoff_land, double ae_openw_l) {
    const double ts = s->theta_s, tr = s->theta_r;
    double usd = sum_first(s->ustore, s->n_unsat);
    double sbm_runoff = wfh_max(0.0, exfilt + s->saturation_excess + runoff_land + s->infiltration_excess);
    for (int k = 0; k < s->nl; ++k) {
        if (k < s->n_unsat)
            s->vwc[k] = (s->ustore[k] + (s->act_thick[k] - s->ustorethick[k]) * (ts - tr)) / s->act_thick[k] + tr;
        else
            s->vwc[k] = ts;
        s->rvwc[k] = (s->vwc[k] / ts) / 1e-2;
    }
    double rootstore_unsat = 0.0;
    for (int k = 0; k < s->n_unsat; ++k)
        rootstore_unsat += wfh_min(1.0, (wfh_max(0.0, s->rooting_depth - s->cum_depth[k]) / s->ustorethick[k])) *
                           s->ustore[k];
    double rootstore_sat = wfh_max(0.0, s->rooting_depth - s->zi) * (ts - tr);
    double rzs = rootstore_sat + rootstore_unsat;
    double vwc_root = rzs / s->rooting_depth + tr;
    double swd = (s->soil_thickness - s->zi) * (ts - tr);
    double drainable = (s->soil_thickness - s->zi) * lower_bound_drainable_porosity(ts, s->theta_fc);
    double cap = s->soil_water_capacity - swd - usd;
    s->ustorecapacity = cap;
    s->ustoredepth = usd;
    s->satwaterdepth = swd;
    s->drainable = drainable;
    s->exfiltration = exfilt;
    s->runoff = sbm_runoff;
    s->root_zone_storage = rzs;
    s->vwc_root = vwc_root;
    s->rvwc_root = (vwc_root / ts) / 1e-2;
    s->total_soil_water_storage = swd + usd;
    s->net_runoff = s->runoff - ae_openw_l;
}

/* ------------------------------------------------------------------------------------ */
/* Graph utilities                                                                       */
/* ------------------------------------------------------------------------------------ */

/* Graphs.topological_sort_by_dfs (Graphs.jl 1.14, iterative colouring DFS: vertices in order
 * 1..n, follow the first unvisited out-neighbour, append a vertex to the post-order when it
 * has none left, reverse at the end). Out-neighbour lists must be sorted ascending.
 * Returns 0, or -1 if the graph has a cycle. */
int32_t wfo_toposort_dfs(int64_t n, const int64_t* out_ptr, const int64_t* out_idx, int64_t* order) {
    uint8_t* color = (uint8_t*)calloc((size_t)n, 1);
    int64_t* stack = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* post = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t np = 0;
    int32_t rc = 0;
    for (int64_t v = 0; v < n && rc == 0; ++v) {
        if (color[v]) continue;
        int64_t sp = 0;
        stack[sp++] = v;
        color[v] = 1;
        while (sp > 0) {
            int64_t u = stack[sp - 1];
            int64_t w = -1;
            for (int64_t e = out_ptr[u]; e < out_ptr[u + 1]; ++e) {
                int64_t x = out_idx[e];
                if (color[x] == 1) {
                    rc = -1;
                    break;
                } else if (color[x] == 0) {
                    w = x;
                    break;
                }
            }
            if (rc) break;
            if (w >= 0) {
                color[w] = 1;
                stack[sp++] = w;
            } else {
                color[u] = 2;
                post[np++] = u;
                --sp;
            }
        }
    }
    if (rc == 0)
        for (int64_t i = 0; i < n; ++i) order[i] = post[n - 1 - i];
    free(color);
    free(stack);
    free(post);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* model                                                                                 */
/* ------------------------------------------------------------------------------------ */

struct wfo_model {
    wfh_config cfg;
    int64_t n, nriv;
    int N;
    int gash;
    int64_t kw_iters;
    /* land topology */
    int64_t *order, *down, *up_ptr, *up_idx;
    /* river topology */
    int64_t *rorder, *rdown, *rup_ptr, *rup_idx, *river_land, *land_river;
    /* land parameters (domain.jl:218-261) */
    double *area, *flow_width, *flow_length, *slope, *river_fraction, *water_fraction, *sfw, *fftr;
    double *alpha_land, *elevation;
    /* vegetation / interception */
    double *gap, *smax, *kc, *er;
    double *canopy, *canopy_potevap, *interception_rate, *throughfall, *stemflow;
    /* snow */
    double *cfmax, *tt, *tti, *ttm, *whc;
    double *snow_s, *snow_w, *swe, *snow_runoff;
    /* open water runoff */
    double *runoff_river, *runoff_land, *aeow_r, *aeow_l, *net_runoff_river, *wfs;
    /* soil */
    wfo_soil* soil;
    /* lateral subsurface flow */
    double *kh0, *sy, *ssf_d, *qmax, *ssf_q, *ssf_zi, *ssf_qin, *ssf_head, *ssf_storage;
    double *ssf_qin_cum, *ssf_q_cum, *ssf_exf_cum, *ssf_qnet_cum, *ssf_tor_cum, *qnet_bnds;
    double *ssf_qin_avg, *ssf_q_avg, *ssf_exf_avg, *ssf_qnet_avg, *ssf_tor_avg;
    double *rech_rate, *rech_flux_cum, *rech_flux_avg;
    /* overland */
    double *ol_q, *ol_qin, *ol_h, *ol_storage, *ol_qlat, *ol_inwater, *ol_q_cum, *ol_qin_cum, *ol_tor_cum;
    double *ol_q_avg, *ol_qin_avg, *ol_tor_avg;
    /* river */
    double *r_alpha, *r_fw, *r_fl, *r_q, *r_qin, *r_h, *r_storage, *r_qlat, *r_inwater, *r_q_cum, *r_qin_cum;
    double *r_q_avg, *r_qin_avg, *r_area;
    /* forcing of the current step */
    double *precip, *pet, *temp;
    /* total storage and balance */
    double *total_storage, *land_store_prev, *ol_store_prev, *r_store_prev;
    double *bal_land, *bal_land_rel, *bal_ssf, *bal_ol, *bal_r;
    /* adaptive sub-steps: stable time step buffer, sub-step counts of the last step */
    double* stbuf;
    int32_t nsub[3];
    /* planted mistake for the tests of the network checks (wfo_set_mutation), and the copy of the
     * upstream discharge at the start of the sub-step that the mutated folds read */
    int32_t mutation;
    double* stale;
    /* lateral snow transport */
    double *snow_in, *snow_out, *snow_flux_w, *snow_cap_s, *snow_cap_w;
    /* upstream lists without reservoir outlets (filter_upstream_nodes, utils.jl:61-71) */
    int64_t *fup_ptr, *fup_idx, *rfup_ptr, *rfup_idx;
    /* reservoirs */
    int64_t nres;
    wfo_reservoir* res;
    int64_t *res_land, *res_river, *land_res, *river_res, *cov_ptr, *cov_idx;
    uint8_t* land_cov;
    double *res_store_prev, *res_outflow_avg, *res_inflow_avg, *res_actevap_avg, *res_abs_avg, *bal_res;
    double* res_tab; /* copies of the S-H and H-Q tables the reservoirs point into */
    /* 1-D floodplain (floodplain.jl); fp_prof: [nriv][4][fp_nd] storage, width, flow area,
     * wetted perimeter of FloodPlainProfile, depth 0 included */
    int fp;
    int64_t fp_nd;
    double *fp_depth, *fp_prof, *fp_n, *fp_slope, *r_bfs, *r_bfd;
    double *fp_storage, *fp_h, *fp_q, *fp_qin, *fp_exch, *fp_cap, *fp_q_cum, *fp_qin_cum, *fp_q_avg, *fp_qin_avg;
    double* r_qch_avg;
};

static double* dalloc(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
static int64_t* ialloc(int64_t n) { return (int64_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t)); }

static const int PCR_DX[10] = {0, -1, 0, 1, -1, 0, 1, -1, 0, 1};
static const int PCR_DY[10] = {0, -1, -1, -1, 0, 0, 0, 1, 1, 1};

/* flowgraph  Wflow/src/routing/utils.jl:6-33 on a set of active raster indices (sorted):
 * down[i] = downstream node or -1 (pit or invalid direction -> pit). */
static void flowgraph(const int64_t* idx, int64_t n, const uint8_t* ldd, int64_t nx, int64_t ny, int64_t* down,
                      uint8_t* ldd_out) {
    for (int64_t i = 0; i < n; ++i) {
        int l = ldd[i];
        ldd_out[i] = (uint8_t)l;
        down[i] = -1;
        if (l == 5 || l < 1 || l > 9) {
            if (l < 1 || l > 9) ldd_out[i] = 5;
            continue;
        }
        int64_t x = idx[i] % nx + PCR_DX[l], y = idx[i] / nx + PCR_DY[l];
        if (x < 0 || y < 0 || x >= nx || y >= ny) {
            ldd_out[i] = 5;
            continue;
        }
        int64_t t = x + y * nx;
        int64_t lo = 0, hi = n; /* searchsortedfirst */
        while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (idx[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= n || idx[lo] != t) {
            ldd_out[i] = 5;
            continue;
        }
        down[i] = lo;
    }
}

/* upstream lists in ascending id order (Graphs inneighbors) as CSR */
static void build_upstream(int64_t n, const int64_t* down, int64_t* ptr, int64_t** idx_out) {
    for (int64_t i = 0; i <= n; ++i) ptr[i] = 0;
    for (int64_t i = 0; i < n; ++i)
        if (down[i] >= 0) ptr[down[i] + 1]++;
    for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    int64_t* idx = ialloc(ptr[n]);
    int64_t* fill = ialloc(n);
    for (int64_t i = 0; i < n; ++i) /* ascending i -> ascending lists */
        if (down[i] >= 0) idx[ptr[down[i]] + fill[down[i]]++] = i;
    free(fill);
    *idx_out = idx;
}

static int32_t toposort_down(int64_t n, const int64_t* down, int64_t* order) {
    int64_t* op = ialloc(n + 1);
    int64_t* oi = ialloc(n);
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        op[i] = k;
        if (down[i] >= 0) oi[k++] = down[i];
    }
    op[n] = k;
    int32_t rc = wfo_toposort_dfs(n, op, oi, order);
    free(op);
    free(oi);
    return rc;
}

/* get_flow_length / get_flow_width  Wflow/src/utils.jl:418-450 */
static double flow_length_of(int ldd, double xl, double yl) {
    if (ldd == 2 || ldd == 8) return yl;
    if (ldd == 4 || ldd == 6) return xl;
    return hypot(xl, yl);
}
static double flow_width_of(int ldd, double xl, double yl) {
    if (ldd == 2 || ldd == 8) return xl;
    if (ldd == 4 || ldd == 6) return yl;
    return (xl * yl) / hypot(xl, yl);
}

#define REQ(p)                                                                     \
    do {                                                                           \
        if (!(st->p)) {                                                            \
            snprintf(err, (size_t)errlen, "wfh_static.%s is NULL", #p);            \
            goto fail;                                                             \
        }                                                                          \
    } while (0)

wfo_model* wfo_create(const wfh_config* cfg, const wfh_static* st, char* err, int64_t errlen) {
    wfo_model* m = (wfo_model*)calloc(1, sizeof(wfo_model));
    const int64_t n = st->n;
    int64_t nriv = 0;
    uint8_t* ldd2 = NULL;
    m->cfg = *cfg;
    m->n = n;
    if (n < 1 || cfg->n_config_layers < 1 || cfg->n_config_layers + 1 > NMAX) {
        snprintf(err, (size_t)errlen, "bad sizes");
        goto fail;
    }
    REQ(cell_index); REQ(ldd); REQ(x_length); REQ(y_length); REQ(land_slope); REQ(land_manning_n);
    REQ(river_location); REQ(river_width_land); REQ(river_length_land); REQ(water_fraction);
    REQ(canopy_gap_fraction); REQ(max_canopy_storage); REQ(rooting_depth); REQ(crop_coefficient);
    REQ(theta_s); REQ(theta_r); REQ(kv_0); REQ(f); REQ(air_entry_pressure); REQ(h1); REQ(h2);
    REQ(h3_high); REQ(h3_low); REQ(h4); REQ(alpha_h1); REQ(soil_thickness); REQ(infiltcap_compacted);
    REQ(max_leakage); REQ(compacted_fraction); REQ(rootdistpar); REQ(cap_hmax); REQ(cap_n); REQ(w_soil);
    REQ(cf_soil); REQ(brooks_corey_c); REQ(kvfrac); REQ(khfrac); REQ(elevation);
    m->N = cfg->n_config_layers + 1;
    m->gash = cfg->dt >= 23.0 * 3600.0;
    for (int64_t i = 0; i < n; ++i) nriv += st->river_location[i] != 0;
    if (nriv != st->nriv) {
        snprintf(err, (size_t)errlen, "nriv %lld != count(river_location) %lld", (long long)st->nriv,
                 (long long)nriv);
        goto fail;
    }
    m->nriv = nriv;
    /* ---- land topology (network.jl:87-112) ---- */
    m->down = ialloc(n);
    ldd2 = (uint8_t*)malloc((size_t)n);
    flowgraph(st->cell_index, n, st->ldd, st->nx, st->ny, m->down, ldd2);
    m->order = ialloc(n);
    if (toposort_down(n, m->down, m->order)) {
        snprintf(err, (size_t)errlen, "cycle in land flow graph");
        goto fail;
    }
    m->up_ptr = ialloc(n + 1);
    build_upstream(n, m->down, m->up_ptr, &m->up_idx);
    /* ---- river topology (network.jl:175-258) ---- */
    m->river_land = ialloc(nriv);
    m->land_river = ialloc(n);
    {
        int64_t r = 0;
        for (int64_t i = 0; i < n; ++i) {
            m->land_river[i] = -1;
            if (st->river_location[i]) {
                m->river_land[r] = i;
                m->land_river[i] = r++;
            }
        }
        int64_t* ridx = ialloc(nriv);
        uint8_t* rldd = (uint8_t*)malloc((size_t)(nriv > 0 ? nriv : 1));
        uint8_t* rldd2 = (uint8_t*)malloc((size_t)(nriv > 0 ? nriv : 1));
        for (int64_t k = 0; k < nriv; ++k) {
            ridx[k] = st->cell_index[m->river_land[k]];
            rldd[k] = ldd2[m->river_land[k]];
        }
        m->rdown = ialloc(nriv);
        flowgraph(ridx, nriv, rldd, st->nx, st->ny, m->rdown, rldd2);
        free(ridx);
        free(rldd);
        free(rldd2);
        m->rorder = ialloc(nriv);
        if (toposort_down(nriv, m->rdown, m->rorder)) {
            snprintf(err, (size_t)errlen, "cycle in river flow graph");
            goto fail;
        }
        m->rup_ptr = ialloc(nriv + 1);
        build_upstream(nriv, m->rdown, m->rup_ptr, &m->rup_idx);
    }
    /* ---- land parameters (domain.jl:218-261, 284-366) ---- */
    m->area = dalloc(n); m->flow_width = dalloc(n); m->flow_length = dalloc(n); m->slope = dalloc(n);
    m->river_fraction = dalloc(n); m->water_fraction = dalloc(n); m->sfw = dalloc(n); m->fftr = dalloc(n);
    m->alpha_land = dalloc(n); m->elevation = dalloc(n);
    for (int64_t i = 0; i < n; ++i) {
        double xl = st->x_length[i], yl = st->y_length[i];
        m->area[i] = xl * yl;
        m->flow_width[i] = flow_width_of(ldd2[i], xl, yl);
        m->flow_length[i] = flow_length_of(ldd2[i], xl, yl);
        double s = st->land_slope[i];
        m->slope[i] = wfh_clamp(s, 0.00001, INFINITY);
        m->river_fraction[i] =
            st->river_location[i] ? wfh_min((st->river_length_land[i] * st->river_width_land[i]) / (m->area[i]), 1.0)
                                  : 0.0;
        m->water_fraction[i] = wfh_max(st->water_fraction[i] - m->river_fraction[i], 0.0);
        double land_area = (1.0 - m->river_fraction[i]) * m->area[i];
        m->sfw[i] = st->river_location[i] ? land_area / m->flow_length[i] : m->flow_width[i];
        m->elevation[i] = st->elevation[i];
    }
    /* get_flow_fraction_to_river  utils.jl:493-510 */
    for (int64_t r = 0; r < nriv; ++r) {
        int64_t i = m->river_land[r];
        for (int64_t e = m->up_ptr[i]; e < m->up_ptr[i + 1]; ++e) {
            int64_t j = m->up_idx[e];
            if (ldd2[j] != ldd2[i]) m->fftr[j] = m->slope[j] / (m->slope[i] + m->slope[j]);
        }
    }
    /* ManningFlowParameters  surface_kinwave.jl:32-44 (land: wetted perimeter = surface width) */
    {
        const double hrp = 2.0 / 3.0;
        const double apow = hrp * 0.6;
        for (int64_t i = 0; i < n; ++i) {
            double at = wfh_pow(st->land_manning_n[i] / sqrt(m->slope[i]), 0.6);
            m->alpha_land[i] = at * wfh_pow(m->sfw[i], apow);
        }
        m->r_alpha = dalloc(nriv); m->r_fw = dalloc(nriv); m->r_fl = dalloc(nriv); m->r_area = dalloc(nriv);
        for (int64_t r = 0; r < nriv; ++r) {
            double s = wfh_clamp(st->river_slope[r], 0.00001, INFINITY);
            m->r_fw[r] = st->river_width[r];
            m->r_fl[r] = st->river_length[r];
            double wp = m->r_fw[r] + st->river_bankfull_depth[r];
            double at = wfh_pow(st->river_manning_n[r] / sqrt(s), 0.6);
            m->r_alpha[r] = at * wfh_pow(wp, apow);
            m->r_area[r] = m->area[m->river_land[r]];
        }
    }
    /* ---- 1-D floodplain: FloodPlainProfile  floodplain.jl:48-147, FloodPlainModel :380-402,
     * bankfull_storage  surface_kinwave.jl:83 ---- */
    m->fp = m->cfg.floodplain_1d != 0;
    if (m->fp) {
        const int64_t nd = (int64_t)st->floodplain_ndepth + 1;
        int ok = st->floodplain_ndepth >= 1 && st->floodplain_depth && st->floodplain_volume &&
                 st->floodplain_manning_n && st->floodplain_slope && st->river_bankfull_depth;
        for (int64_t j = 0; ok && j < nd - 1; ++j)
            ok = isfinite(st->floodplain_depth[j]) && st->floodplain_depth[j] > (j ? st->floodplain_depth[j - 1] : 0.0);
        for (int64_t j = 0; ok && j < (nd - 1) * nriv; ++j) ok = isfinite(st->floodplain_volume[j]);
        if (!ok) {
            snprintf(err, (size_t)errlen, "floodplain: flood depths (ascending, > 0), volumes, manning_n, slope "
                     "and river_bankfull_depth needed");
            goto fail;
        }
        m->fp_nd = nd;
        m->fp_depth = dalloc(nd);
        for (int64_t j = 1; j < nd; ++j) m->fp_depth[j] = st->floodplain_depth[j - 1];
        m->fp_prof = dalloc(nriv * 4 * nd);
        for (int64_t r = 0; r < nriv; ++r)
            wfo_floodplain_profile(nd - 1, st->floodplain_depth, st->floodplain_volume + r, nriv, m->r_fw[r], m->r_fl[r],
                                   m->fp_prof + r * 4 * nd);
        m->fp_n = dalloc(nriv); m->fp_slope = dalloc(nriv); m->r_bfs = dalloc(nriv); m->r_bfd = dalloc(nriv);
        m->fp_storage = dalloc(nriv); m->fp_h = dalloc(nriv); m->fp_q = dalloc(nriv); m->fp_qin = dalloc(nriv);
        m->fp_exch = dalloc(nriv); m->fp_cap = dalloc(nriv); m->fp_q_cum = dalloc(nriv); m->fp_qin_cum = dalloc(nriv);
        m->fp_q_avg = dalloc(nriv); m->fp_qin_avg = dalloc(nriv); m->r_qch_avg = dalloc(nriv);
        for (int64_t r = 0; r < nriv; ++r) {
            m->fp_n[r] = st->floodplain_manning_n[r];
            m->fp_slope[r] = wfh_clamp(st->floodplain_slope[r], 0.00001, INFINITY);
            m->r_bfd[r] = st->river_bankfull_depth[r];
            m->r_bfs[r] = st->river_bankfull_depth[r] * m->r_fw[r] * m->r_fl[r];
        }
    }
    /* ---- reservoirs (network.jl:296-360, reservoir.jl:47-300) and the upstream lists without
     * reservoir outlets (domain.jl:95-121, filter_upstream_nodes utils.jl:61-71) ---- */
    m->nres = st->nres;
    m->land_res = ialloc(n);
    m->river_res = ialloc(nriv);
    for (int64_t i = 0; i < n; ++i) m->land_res[i] = -1;
    for (int64_t r = 0; r < nriv; ++r) m->river_res[r] = -1;
    m->land_cov = (uint8_t*)calloc((size_t)n, 1);
    if (m->nres < 0) {
        snprintf(err, (size_t)errlen, "nres < 0");
        goto fail;
    }
    if (m->nres > 0) {
        const int64_t nr = m->nres;
        if (!st->reservoir_outlet || !st->reservoir_outflow_type || !st->reservoir_area || !st->reservoir_waterlevel) {
            snprintf(err, (size_t)errlen, "reservoir_outlet / outflow_type / area / waterlevel missing");
            goto fail;
        }
        m->res = (wfo_reservoir*)calloc((size_t)nr, sizeof(wfo_reservoir));
        m->res_land = ialloc(nr);
        m->res_river = ialloc(nr);
        int64_t* taken = ialloc(nriv > 0 ? nriv : 1); /* downstream river cells fed by a reservoir */
        for (int64_t k = 0; k < nr; ++k) {
            const int64_t i = st->reservoir_outlet[k];
            if (i < 0 || i >= n || (k > 0 && i <= st->reservoir_outlet[k - 1]) || m->land_river[i] < 0 ||
                m->rdown[m->land_river[i]] < 0 || taken[m->rdown[m->land_river[i]]]) {
                snprintf(err, (size_t)errlen, "reservoir %lld: outlet must be a river cell with a river cell "
                         "downstream, outlets ascending, one reservoir per downstream cell", (long long)k);
                free(taken);
                goto fail;
            }
            taken[m->rdown[m->land_river[i]]] = 1;
            m->res_land[k] = i;
            m->res_river[k] = m->land_river[i];
            m->land_res[i] = k;
            m->river_res[m->land_river[i]] = k;
            const int32_t t = st->reservoir_outflow_type[k];
            const int32_t sct = st->reservoir_storage_curve_type ? st->reservoir_storage_curve_type[k] : 1;
            if (t < 1 || t > 4 || (sct != 1 && sct != 2)) {
                snprintf(err, (size_t)errlen, "reservoir %lld: outflow type %d / storage curve type %d not supported",
                         (long long)k, t, sct);
                free(taken);
                goto fail;
            }
            wfo_reservoir* R = &m->res[k];
            R->outflow_type = t;
            R->storage_curve_type = sct;
            R->area = st->reservoir_area[k];
            /* S-H and H-Q tables (read_sh_csv / read_hq_csv  io.jl:1129-1147) */
            if (sct == 2 || t == 1) {
                const int64_t* ptr = sct == 2 ? st->reservoir_sh_ptr : st->reservoir_hq_ptr;
                int ok = ptr != NULL;
                if (sct == 2) {
                    ok = ok && st->reservoir_sh_h && st->reservoir_sh_s && ptr[k + 1] > ptr[k];
                    if (ok) {
                        R->sh_n = ptr[k + 1] - ptr[k];
                        R->sh_h = st->reservoir_sh_h + ptr[k];
                        R->sh_s = st->reservoir_sh_s + ptr[k];
                        for (int64_t j = 0; j < R->sh_n; ++j)
                            ok = ok && isfinite(R->sh_h[j]) && isfinite(R->sh_s[j]) &&
                                 (j == 0 || (R->sh_h[j] >= R->sh_h[j - 1] && R->sh_s[j] >= R->sh_s[j - 1]));
                    }
                }
                if (t == 1) {
                    ptr = st->reservoir_hq_ptr;
                    ok = ok && ptr && st->reservoir_hq_h && st->reservoir_hq_q && ptr[k + 1] > ptr[k];
                    if (ok) {
                        R->hq_n = ptr[k + 1] - ptr[k];
                        R->hq_h = st->reservoir_hq_h + ptr[k];
                        R->hq_q = st->reservoir_hq_q + ptr[k] * 365;
                        for (int64_t j = 0; j < R->hq_n; ++j) {
                            ok = ok && isfinite(R->hq_h[j]) && (j == 0 || R->hq_h[j] >= R->hq_h[j - 1]);
                            for (int c = 0; c < 365; ++c) ok = ok && isfinite(R->hq_q[j * 365 + c]);
                        }
                    }
                }
                if (!ok) {
                    snprintf(err, (size_t)errlen, "reservoir %lld: missing, empty, non-finite or decreasing S-H / "
                             "H-Q table", (long long)k);
                    free(taken);
                    goto fail;
                }
            }
#define RES_PRM(field, name, dflt) R->field = st->name ? st->name[k] : (dflt)
            RES_PRM(demand, reservoir_demand, NaN);
            RES_PRM(maximum_release, reservoir_maximum_release, NaN);
            RES_PRM(maximum_storage, reservoir_maximum_storage, NaN);
            RES_PRM(target_full_fraction, reservoir_target_full_fraction, NaN);
            RES_PRM(target_minimum_fraction, reservoir_target_minimum_fraction, NaN);
            RES_PRM(threshold, reservoir_threshold, NaN);
            RES_PRM(rating_curve_coefficient, reservoir_rating_curve_coefficient, NaN);
            RES_PRM(external_inflow, reservoir_external_inflow, 0.0);
            RES_PRM(outflow_obs, reservoir_outflow_obs, NaN);
            RES_PRM(rating_curve_exponent, reservoir_rating_curve_exponent, NaN);
#undef RES_PRM
            R->waterlevel = st->reservoir_waterlevel[k];
            R->storage = wfo_reservoir_initial_storage(R);
            if (t == 1) R->maximum_storage = wfo_reservoir_maximum_storage(R); /* reservoir.jl:170-175 */
            R->sh_h = R->sh_s = R->hq_h = R->hq_q = NULL; /* re-pointed into res_tab below */
            R->hq_col = 0;
            R->outflow = NaN;
            R->lower = NULL;
        }
        free(taken);
        /* linked free weirs (lower_reservoir_ind, reservoir.jl:152-159): the lower reservoir's
         * outlet must lie downstream of the upper one's on the river, and at most one upper
         * reservoir may link to a lower one (the GPU hands the transfer down the river) */
        if (st->reservoir_lower) {
            int64_t* has_upper = ialloc(nr);
            for (int64_t k = 0; k < nr; ++k) {
                const int32_t lw = st->reservoir_lower[k];
                if (lw == 0) continue;
                int ok = lw >= 1 && lw <= nr && lw - 1 != k && m->res[k].outflow_type == 2 && !has_upper[lw - 1];
                if (ok) { /* walk down the river from the upper outlet */
                    ok = 0;
                    for (int64_t rr = m->rdown[m->res_river[k]]; rr >= 0; rr = m->rdown[rr])
                        if (rr == m->res_river[lw - 1]) {
                            ok = 1;
                            break;
                        }
                }
                if (!ok) {
                    snprintf(err, (size_t)errlen, "reservoir %lld: linked lower reservoir %d must be another reservoir "
                             "whose outlet lies downstream on the river, linked from one free weir only",
                             (long long)k, lw);
                    free(has_upper);
                    goto fail;
                }
                has_upper[lw - 1] = 1;
                m->res[k].lower = &m->res[lw - 1];
            }
            free(has_upper);
        }
        {
            int64_t tot = 0;
            for (int64_t k = 0; k < nr; ++k) tot += 2 * m->res[k].sh_n + 366 * m->res[k].hq_n;
            m->res_tab = dalloc(tot > 0 ? tot : 1);
            double* w = m->res_tab;
            for (int64_t k = 0; k < nr; ++k) {
                wfo_reservoir* R = &m->res[k];
                const int64_t s0 = R->sh_n ? st->reservoir_sh_ptr[k] : 0, h0 = R->hq_n ? st->reservoir_hq_ptr[k] : 0;
                if (R->sh_n) {
                    memcpy(w, st->reservoir_sh_h + s0, sizeof(double) * (size_t)R->sh_n);
                    memcpy(w + R->sh_n, st->reservoir_sh_s + s0, sizeof(double) * (size_t)R->sh_n);
                    R->sh_h = w;
                    R->sh_s = w + R->sh_n;
                    w += 2 * R->sh_n;
                }
                if (R->hq_n) {
                    memcpy(w, st->reservoir_hq_h + h0, sizeof(double) * (size_t)R->hq_n);
                    memcpy(w + R->hq_n, st->reservoir_hq_q + h0 * 365, sizeof(double) * (size_t)(365 * R->hq_n));
                    R->hq_h = w;
                    R->hq_q = w + R->hq_n;
                    w += 366 * R->hq_n;
                }
            }
        }
        /* coverage lists in ascending land index (findall order of reservoir_area__count) */
        m->cov_ptr = ialloc(nr + 1);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t id = st->reservoir_coverage ? st->reservoir_coverage[i] : m->land_res[i] + 1;
            if (id < 0 || id > nr) {
                snprintf(err, (size_t)errlen, "reservoir_coverage out of range");
                goto fail;
            }
            if (id > 0) m->cov_ptr[id]++;
        }
        for (int64_t k = 0; k < nr; ++k) {
            if (m->cov_ptr[k + 1] == 0) {
                snprintf(err, (size_t)errlen, "reservoir %lld has an empty area", (long long)k);
                goto fail;
            }
            m->cov_ptr[k + 1] += m->cov_ptr[k];
        }
        m->cov_idx = ialloc(m->cov_ptr[nr]);
        int64_t* fill = ialloc(nr);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t id = st->reservoir_coverage ? st->reservoir_coverage[i] : m->land_res[i] + 1;
            if (id > 0) {
                m->cov_idx[m->cov_ptr[id - 1] + fill[id - 1]++] = i;
                m->land_cov[i] = 1;
            }
        }
        free(fill);
        m->res_store_prev = dalloc(nr); m->res_outflow_avg = dalloc(nr); m->res_inflow_avg = dalloc(nr);
        m->res_actevap_avg = dalloc(nr); m->res_abs_avg = dalloc(nr); m->bal_res = dalloc(nr);
    }
    {
        int64_t* df = ialloc(n);
        for (int64_t i = 0; i < n; ++i) df[i] = m->land_res[i] >= 0 ? -1 : m->down[i];
        m->fup_ptr = ialloc(n + 1);
        build_upstream(n, df, m->fup_ptr, &m->fup_idx);
        free(df);
        int64_t* rf = ialloc(nriv > 0 ? nriv : 1);
        for (int64_t r = 0; r < nriv; ++r) rf[r] = m->river_res[r] >= 0 ? -1 : m->rdown[r];
        m->rfup_ptr = ialloc(nriv + 1);
        build_upstream(nriv, rf, m->rfup_ptr, &m->rfup_idx);
        free(rf);
    }
    /* ---- vegetation, snow ---- */
    m->gap = dalloc(n); m->smax = dalloc(n); m->kc = dalloc(n); m->er = dalloc(n);
    m->canopy = dalloc(n); m->canopy_potevap = dalloc(n); m->interception_rate = dalloc(n);
    m->throughfall = dalloc(n); m->stemflow = dalloc(n);
    m->cfmax = dalloc(n); m->tt = dalloc(n); m->tti = dalloc(n); m->ttm = dalloc(n); m->whc = dalloc(n);
    m->snow_s = dalloc(n); m->snow_w = dalloc(n); m->swe = dalloc(n); m->snow_runoff = dalloc(n);
    for (int64_t i = 0; i < n; ++i) {
        m->gap[i] = st->canopy_gap_fraction[i];
        m->smax[i] = st->max_canopy_storage[i];
        m->kc[i] = st->crop_coefficient[i];
        m->er[i] = st->evap_precip_ratio ? st->evap_precip_ratio[i] : NaN;
        if (cfg->snow) {
            m->cfmax[i] = st->cfmax[i]; m->tt[i] = st->tt[i]; m->tti[i] = st->tti[i];
            m->ttm[i] = st->ttm[i]; m->whc[i] = st->whc[i];
        }
    }
    m->runoff_river = dalloc(n); m->runoff_land = dalloc(n); m->aeow_r = dalloc(n); m->aeow_l = dalloc(n);
    m->net_runoff_river = dalloc(n); m->wfs = dalloc(n);
    /* ---- soil (soil.jl:333-615, sbm.jl:57-58) ---- */
    m->soil = (wfo_soil*)calloc((size_t)n, sizeof(wfo_soil));
    {
        const int N = m->N;
        double thick_cfg[NMAX], cum_cfg[NMAX + 1];
        for (int k = 0; k < N - 1; ++k) thick_cfg[k] = cfg->soil_layer_thickness_mm[k] * 1e-3;
        thick_cfg[N - 1] = NaN;
        cum_cfg[0] = 0.0;
        cum_cfg[1] = thick_cfg[0];
        for (int k = 1; k < N; ++k) cum_cfg[k + 1] = cum_cfg[k] + thick_cfg[k];
        for (int64_t i = 0; i < n; ++i) {
            wfo_soil* s = &m->soil[i];
            s->N = N;
            s->theta_s = st->theta_s[i]; s->theta_r = st->theta_r[i];
            s->kv_0 = st->kv_0[i]; s->f = st->f[i]; s->z_exp = st->z_exp ? st->z_exp[i] : INFINITY; s->air_entry_pressure = st->air_entry_pressure[i];
            s->h1 = st->h1[i]; s->h2 = st->h2[i]; s->h3_high = st->h3_high[i]; s->h3_low = st->h3_low[i];
            s->h4 = st->h4[i]; s->alpha_h1 = st->alpha_h1[i]; s->soil_thickness = st->soil_thickness[i];
            s->infiltcap_compacted = st->infiltcap_compacted[i]; s->max_leakage = st->max_leakage[i];
            s->compacted_fraction = st->compacted_fraction[i]; s->rootdistpar = st->rootdistpar[i];
            s->cap_hmax = st->cap_hmax[i]; s->cap_n = st->cap_n[i]; s->w_soil = st->w_soil[i];
            s->cf_soil = st->cf_soil[i];
            for (int k = 0; k < N; ++k) {
                s->c[k] = st->brooks_corey_c[(int64_t)k * n + i];
                s->kvfrac[k] = st->kvfrac[(int64_t)k * n + i];
            }
            s->infiltcap_soil = s->kv_0 * s->kvfrac[0];
            set_layerthickness(s->soil_thickness, cum_cfg, thick_cfg, N, s->act_thick);
            s->cum_depth[0] = 0.0;
            s->cum_depth[1] = s->act_thick[0];
            for (int k = 1; k < N; ++k) s->cum_depth[k + 1] = s->cum_depth[k] + s->act_thick[k];
            s->nl = number_of_active_layers(s->act_thick, N);
            if (st->theta_fc)
                s->theta_fc = st->theta_fc[i];
            else
                s->theta_fc = field_capacity(s->act_thick, s->nl, s->theta_s, s->theta_r, s->c, s->air_entry_pressure);
            /* default root fraction with the unclipped rooting depth (soil.jl:526-548) */
            double rdu = st->rooting_depth[i];
            for (int k = 0; k < N; ++k) {
                if (st->rootfraction) {
                    s->rootfraction[k] = st->rootfraction[(int64_t)k * n + i];
                } else if (rdu > 0.0) {
                    if ((rdu - s->cum_depth[k]) >= s->act_thick[k])
                        s->rootfraction[k] = s->act_thick[k] / rdu;
                    else
                        s->rootfraction[k] = wfh_max((rdu - s->cum_depth[k]) / rdu, 0.0);
                } else {
                    s->rootfraction[k] = 0.0;
                }
            }
            s->soil_water_capacity = s->soil_thickness * (s->theta_s - s->theta_r);
            s->rooting_depth = wfh_min(s->soil_thickness * 0.99, rdu);
            /* SbmSoilVariables cold state (soil.jl:150-201) */
            s->satwaterdepth = 0.85 * s->soil_water_capacity;
            s->zi = wfh_max(0.0, s->soil_thickness - s->satwaterdepth / (s->theta_s - s->theta_r));
            s->drainable = (s->soil_thickness - s->zi) * lower_bound_drainable_porosity(s->theta_s, s->theta_fc);
            set_layerthickness(s->zi, s->cum_depth, s->act_thick, N, s->ustorethick);
            s->n_unsat = number_of_active_layers(s->ustorethick, N);
            for (int k = 0; k < NMAX; ++k) s->ustore[k] = 0.0;
            s->ustoredepth = 0.0;
            s->ustorecapacity = s->soil_water_capacity - s->satwaterdepth;
            s->total_soil_water_storage = s->satwaterdepth + s->ustoredepth;
            s->tsoil = 10.0 + 273.15;
            s->f_infred = 1.0;
            /* soil_fraction!  soil.jl:643-655 (static canopy, no glacier) */
            s->soil_fraction = wfh_max(m->gap[i] - m->water_fraction[i] - m->river_fraction[i] - 0.0, 0.0);
        }
    }
    /* ---- lateral subsurface flow (lateral_subsurface_flow.jl:81-171, utils.jl:916-937) ---- */
    m->kh0 = dalloc(n); m->sy = dalloc(n); m->ssf_d = dalloc(n); m->qmax = dalloc(n); m->ssf_q = dalloc(n);
    m->ssf_zi = dalloc(n); m->ssf_qin = dalloc(n); m->ssf_head = dalloc(n); m->ssf_storage = dalloc(n);
    m->ssf_qin_cum = dalloc(n); m->ssf_q_cum = dalloc(n); m->ssf_exf_cum = dalloc(n); m->ssf_qnet_cum = dalloc(n);
    m->ssf_tor_cum = dalloc(n); m->qnet_bnds = dalloc(n); m->ssf_qin_avg = dalloc(n); m->ssf_q_avg = dalloc(n);
    m->ssf_exf_avg = dalloc(n); m->ssf_qnet_avg = dalloc(n); m->ssf_tor_avg = dalloc(n);
    m->rech_rate = dalloc(n); m->rech_flux_cum = dalloc(n); m->rech_flux_avg = dalloc(n);
    for (int64_t i = 0; i < n; ++i) {
        wfo_soil* s = &m->soil[i];
        m->kh0[i] = st->khfrac[i] * s->kv_0;
        m->sy[i] = lower_bound_drainable_porosity(s->theta_s, s->theta_fc);
        m->ssf_d[i] = s->soil_thickness;
        m->ssf_zi[i] = s->zi;
        m->ssf_storage[i] = m->sy[i] * (m->ssf_d[i] - m->ssf_zi[i]) * m->area[i];
        m->ssf_head[i] = m->elevation[i] - m->ssf_zi[i];
        double f = s->f;
        const double ze = s->z_exp;
        if (ze < INFINITY) { /* initialize_lateral_ssf_model!(::KhExponentialConstant)  utils.jl:940-975 */
            const double q_constant = m->kh0[i] * wfh_exp(-f * ze) * m->slope[i] * (m->ssf_d[i] - ze);
            m->qmax[i] = ((m->kh0[i] * m->slope[i]) / f) * (1.0 - wfh_exp(-f * ze)) + q_constant;
            if (m->ssf_zi[i] < ze)
                m->ssf_q[i] = (((m->kh0[i] * m->slope[i]) / f) * (wfh_exp(-f * m->ssf_zi[i]) - wfh_exp(-f * ze)) +
                               q_constant) * m->flow_width[i];
            else
                m->ssf_q[i] = m->kh0[i] * wfh_exp(-f * m->ssf_zi[i]) * m->slope[i] * (m->ssf_d[i] - m->ssf_zi[i]) *
                              m->flow_width[i];
        } else { /* initialize_lateral_ssf_model!(::KhExponential)  utils.jl:916-937 */
            m->qmax[i] = ((m->kh0[i] * m->slope[i]) / f) * (1.0 - wfh_exp(-f * m->ssf_d[i]));
            m->ssf_q[i] = ((m->kh0[i] * m->slope[i]) / f) *
                          (wfh_exp(-f * m->ssf_zi[i]) - wfh_exp(-f * m->ssf_d[i])) * m->flow_width[i];
        }
    }
    /* ---- overland / river state ---- */
    m->ol_q = dalloc(n); m->ol_qin = dalloc(n); m->ol_h = dalloc(n); m->ol_storage = dalloc(n);
    m->ol_qlat = dalloc(n); m->ol_inwater = dalloc(n); m->ol_q_cum = dalloc(n); m->ol_qin_cum = dalloc(n);
    m->ol_tor_cum = dalloc(n); m->ol_q_avg = dalloc(n); m->ol_qin_avg = dalloc(n); m->ol_tor_avg = dalloc(n);
    m->r_q = dalloc(nriv); m->r_qin = dalloc(nriv); m->r_h = dalloc(nriv); m->r_storage = dalloc(nriv);
    m->r_qlat = dalloc(nriv); m->r_inwater = dalloc(nriv); m->r_q_cum = dalloc(nriv); m->r_qin_cum = dalloc(nriv);
    m->r_q_avg = dalloc(nriv); m->r_qin_avg = dalloc(nriv);
    m->precip = dalloc(n); m->pet = dalloc(n); m->temp = dalloc(n);
    m->total_storage = dalloc(n); m->land_store_prev = dalloc(n); m->ol_store_prev = dalloc(n);
    m->r_store_prev = dalloc(nriv); m->bal_land = dalloc(n); m->bal_land_rel = dalloc(n); m->bal_ssf = dalloc(n);
    m->bal_ol = dalloc(n); m->bal_r = dalloc(nriv);
    m->stbuf = dalloc(n > nriv ? n : nriv);
    m->stale = dalloc(n > nriv ? n : nriv);
    m->snow_in = dalloc(n); m->snow_out = dalloc(n); m->snow_flux_w = dalloc(n);
    m->snow_cap_s = dalloc(n); m->snow_cap_w = dalloc(n);
    free(ldd2);
    return m;
fail:
    free(ldd2);
    wfo_destroy(m);
    return NULL;
}

void wfo_destroy(wfo_model* m) {
    if (!m) return;
    void** p[] = {
        (void**)&m->order, (void**)&m->down, (void**)&m->up_ptr, (void**)&m->up_idx, (void**)&m->rorder,
        (void**)&m->rdown, (void**)&m->rup_ptr, (void**)&m->rup_idx, (void**)&m->river_land, (void**)&m->land_river,
        (void**)&m->area, (void**)&m->flow_width, (void**)&m->flow_length, (void**)&m->slope,
        (void**)&m->river_fraction, (void**)&m->water_fraction, (void**)&m->sfw, (void**)&m->fftr,
        (void**)&m->alpha_land, (void**)&m->elevation, (void**)&m->gap, (void**)&m->smax, (void**)&m->kc,
        (void**)&m->er, (void**)&m->canopy, (void**)&m->canopy_potevap, (void**)&m->interception_rate,
        (void**)&m->throughfall, (void**)&m->stemflow, (void**)&m->cfmax, (void**)&m->tt, (void**)&m->tti,
        (void**)&m->ttm, (void**)&m->whc, (void**)&m->snow_s, (void**)&m->snow_w, (void**)&m->swe,
        (void**)&m->snow_runoff, (void**)&m->runoff_river, (void**)&m->runoff_land, (void**)&m->aeow_r,
        (void**)&m->aeow_l, (void**)&m->net_runoff_river, (void**)&m->wfs, (void**)&m->soil, (void**)&m->kh0,
        (void**)&m->sy, (void**)&m->ssf_d, (void**)&m->qmax, (void**)&m->ssf_q, (void**)&m->ssf_zi,
        (void**)&m->ssf_qin, (void**)&m->ssf_head, (void**)&m->ssf_storage, (void**)&m->ssf_qin_cum,
        (void**)&m->ssf_q_cum, (void**)&m->ssf_exf_cum, (void**)&m->ssf_qnet_cum, (void**)&m->ssf_tor_cum,
        (void**)&m->qnet_bnds, (void**)&m->ssf_qin_avg, (void**)&m->ssf_q_avg, (void**)&m->ssf_exf_avg,
        (void**)&m->ssf_qnet_avg, (void**)&m->ssf_tor_avg, (void**)&m->rech_rate, (void**)&m->rech_flux_cum,
        (void**)&m->rech_flux_avg, (void**)&m->ol_q, (void**)&m->ol_qin, (void**)&m->ol_h, (void**)&m->ol_storage,
        (void**)&m->ol_qlat, (void**)&m->ol_inwater, (void**)&m->ol_q_cum, (void**)&m->ol_qin_cum,
        (void**)&m->ol_tor_cum, (void**)&m->ol_q_avg, (void**)&m->ol_qin_avg, (void**)&m->ol_tor_avg,
        (void**)&m->r_alpha, (void**)&m->r_fw, (void**)&m->r_fl, (void**)&m->r_q, (void**)&m->r_qin,
        (void**)&m->r_h, (void**)&m->r_storage, (void**)&m->r_qlat, (void**)&m->r_inwater, (void**)&m->r_q_cum,
        (void**)&m->r_qin_cum, (void**)&m->r_q_avg, (void**)&m->r_qin_avg, (void**)&m->r_area,
        (void**)&m->precip, (void**)&m->pet, (void**)&m->temp, (void**)&m->total_storage,
        (void**)&m->land_store_prev, (void**)&m->ol_store_prev, (void**)&m->r_store_prev, (void**)&m->bal_land,
        (void**)&m->bal_land_rel, (void**)&m->bal_ssf, (void**)&m->bal_ol, (void**)&m->bal_r, (void**)&m->stbuf, (void**)&m->stale,
        (void**)&m->fup_ptr, (void**)&m->fup_idx, (void**)&m->rfup_ptr, (void**)&m->rfup_idx, (void**)&m->res,
        (void**)&m->res_land, (void**)&m->res_river, (void**)&m->land_res, (void**)&m->river_res,
        (void**)&m->cov_ptr, (void**)&m->cov_idx, (void**)&m->land_cov, (void**)&m->res_store_prev,
        (void**)&m->res_outflow_avg, (void**)&m->res_inflow_avg, (void**)&m->res_actevap_avg, (void**)&m->res_abs_avg,
        (void**)&m->bal_res, (void**)&m->res_tab, (void**)&m->fp_depth, (void**)&m->fp_prof, (void**)&m->fp_n,
        (void**)&m->fp_slope, (void**)&m->r_bfs, (void**)&m->r_bfd, (void**)&m->fp_storage, (void**)&m->fp_h,
        (void**)&m->fp_q, (void**)&m->fp_qin, (void**)&m->fp_exch, (void**)&m->fp_cap, (void**)&m->fp_q_cum,
        (void**)&m->fp_qin_cum, (void**)&m->fp_q_avg, (void**)&m->fp_qin_avg, (void**)&m->r_qch_avg, (void**)&m->snow_in, (void**)&m->snow_out, (void**)&m->snow_flux_w,
        (void**)&m->snow_cap_s, (void**)&m->snow_cap_w};
    for (size_t k = 0; k < sizeof p / sizeof p[0]; ++k) free(*p[k]);
    free(m);
}

/* check_timestepsize  Wflow/src/routing/timestepping.jl:11-17 */
static double check_timestepsize(double ts, double t, double end) { return t + ts > end ? end - t : ts; }

static int cmp_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* quantile!(v, p) of k >= 2 values (Statistics.jl 1.11.1, default type 7; the order statistics
 * come from a full sort, the interpolation is wfh_quantile_index/_interp). Sorts v. */
double wfo_quantile(double* v, int64_t k, double p) {
    qsort(v, (size_t)k, sizeof(double), cmp_double);
    double g;
    const int64_t j = wfh_quantile_index(k, p, &g);
    return wfh_quantile_interp(v[j - 1], v[j], g);
}

/* stable_timestep(flow_model, flow_length, p) of the kinematic-wave river and overland models
 * Wflow/src/routing/surface/surface_kinwave.jl:674-704: the p-quantile of the Courant time steps
 * flow_length / celerity over the cells with q > KIN_WAVE_MIN_FLOW, celerity =
 * inv(alpha BETA q^(BETA - 1)) with the reference's pow (utils.jl:470); 600 s without flow.
 * buf holds n values. */
double wfo_kw_stable_timestep(const double* q, const double* alpha, const double* flow_length, int64_t n,
                              double p, double* buf) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i)
        if (q[i] > WFH_KIN_WAVE_MIN_FLOW) {
            const double c = 1.0 / (alpha[i] * 0.6 * wfh_pow(q[i], (0.6 - 1.0)));
            buf[k++] = flow_length[i] / c;
        }
    if (k == 1) return buf[0];
    if (k > 0) return wfo_quantile(buf, k, p);
    return 600.0;
}

/* stable_timestep(::LateralSSFModel, domain)  lateral_subsurface_flow.jl:314-344: the minimum
 * of flow_length / ssf_celerity over the cells with a positive water table depth (0.5 s when
 * there is none), times the alpha coefficient. */
static double ssf_stable_timestep(const wfo_model* m) {
    double mn = 0.0;
    int found = 0;
    for (int64_t i = 0; i < m->n; ++i)
        if (m->ssf_zi[i] > 0.0) {
            const double c = wfo_ssf_celerity(m->ssf_zi[i], m->slope[i], m->sy[i], m->kh0[i], m->soil[i].f,
                                              m->soil[i].z_exp);
            const double v = m->flow_length[i] / c;
            mn = found ? wfh_min(mn, v) : v;
            found = 1;
        }
    return (found ? mn : 0.5) * m->cfg.subsurface_alpha_coefficient;
}

/* accucapacityflux!(flux, material, network, capacity, dt)  Wflow/src/routing/utils.jl:82-108:
 * in the reference's topological order each cell passes min(material / dt, capacity) on to its
 * downstream cell (the unfiltered flow graph), which adds it to its own material. */
void wfo_accucapacityflux(int64_t n, const int64_t* order, const int64_t* down, double* material,
                          const double* capacity, double* flux, double dt) {
    for (int64_t k = 0; k < n; ++k) {
        const int64_t v = order[k];
        const double flux_val = wfh_min(material[v] / dt, capacity[v]);
        const double material_update = flux_val * dt;
        material[v] -= material_update;
        flux[v] = flux_val;
        if (down[v] >= 0) material[down[v]] += material_update;
    }
}
static void accucapacityflux(const wfo_model* m, double* material, const double* capacity, double* flux, double dt) {
    wfo_accucapacityflux(m->n, m->order, m->down, material, capacity, flux, dt);
}

/* lateral_snow_transport!  Wflow/src/routing/surface/surface_process.jl:9-19 (TAN80 = 5.67,
 * snow_storage_max = 10 m); flux_in! (routing/utils.jl) sums snow_out over the upstream lists
 * (without reservoir outlets) in ascending id. */
static void lateral_snow_transport(wfo_model* m, double dt) {
    const int64_t n = m->n;
    for (int64_t i = 0; i < n; ++i) {
        const double frac = wfh_min(0.5, m->slope[i] / 5.67) * wfh_min(1.0, m->snow_s[i] / 10.0);
        m->snow_cap_s[i] = frac * m->snow_s[i] / dt;
        m->snow_cap_w[i] = m->snow_w[i] * frac / dt;
    }
    accucapacityflux(m, m->snow_s, m->snow_cap_s, m->snow_out, dt);
    accucapacityflux(m, m->snow_w, m->snow_cap_w, m->snow_flux_w, dt);
    for (int64_t i = 0; i < n; ++i) m->snow_out[i] += m->snow_flux_w[i];
    for (int64_t v = 0; v < n; ++v) {
        double s = 0.0;
        for (int64_t e = m->fup_ptr[v]; e < m->fup_ptr[v + 1]; ++e) s = s + m->snow_out[m->fup_idx[e]];
        m->snow_in[v] = s;
    }
}

/* compute_total_storage  mass_balance.jl:96-115 */
static double land_total_storage(const wfo_model* m, int64_t i) {
    double snow = m->cfg.snow ? m->snow_s[i] + m->snow_w[i] : 0.0 + 0.0;
    return m->soil[i].total_soil_water_storage + m->canopy[i] + snow + 0.0 * 0.0 + 0.0;
}

/* update_model!  Wflow/src/sbm_model.jl:60-92, preceded by storage_prev! and followed by
 * compute_mass_balance! (Wflow.jl:315-317) when water_mass_balance is set. */
/* river_channel_floodplain_exchange!  surface_kinwave.jl:572-601: the water above bankfull is
 * shared instantaneously between channel and floodplain; the change of the channel's storage
 * enters the river's kinematic wave as a lateral inflow */
void wfo_floodplain_exchange(const double* prof, int64_t nd, const double* depth0, double river_storage0,
                             double fp_storage0, double bankfull_depth, double flow_width, double flow_length, double dt,
                             double* out3) {
    const double bankfull_storage = bankfull_depth * flow_width * flow_length;
    const double storage_total = river_storage0 + fp_storage0;
    double delta_river_storage, fs, fh;
    if (storage_total > bankfull_storage) {
        const double flood_storage = storage_total - bankfull_storage;
        const double h = wfo_flood_depth(prof, nd, depth0, flood_storage, flow_length);
        const double river_storage = (bankfull_depth + h) * flow_width * flow_length;
        delta_river_storage = river_storage - river_storage0;
        fs = wfh_max(storage_total - river_storage, 0.0);
        fh = fs > 0.0 ? h : 0.0;
    } else {
        delta_river_storage = wfh_max(storage_total - river_storage0, 0.0);
        fh = 0.0;
        fs = 0.0;
    }
    out3[0] = fs;
    out3[1] = fh;
    out3[2] = delta_river_storage / dt;
}

static void river_floodplain_exchange(wfo_model* m, double dt) {
    const int64_t nd = m->fp_nd;
    for (int64_t i = 0; i < m->nriv; ++i) {
        double o[3];
        wfo_floodplain_exchange(m->fp_prof + i * 4 * nd, nd, m->fp_depth, m->r_storage[i], m->fp_storage[i], m->r_bfd[i],
                                m->r_fw[i], m->r_fl[i], dt, o);
        m->fp_storage[i] = o[0];
        m->fp_h[i] = o[1];
        m->fp_exch[i] = o[2];
    }
}

double wfo_floodplain_capacity(const double* prof, const double* prof_ds, int64_t nd, const double* depth0, double h,
                               double mannings_n, double slope) {
    double cap = 0.0;
    if (h > 0.0) {
        int64_t i1, i2;
        fp_interpolation_indices(h, depth0, nd, &i1, &i2);
        const double* Wv = prof + nd;
        const double flow_area = fp_flow_area(Wv, Wv + nd, depth0, h, i1, i2);
        double flow_area_ds = flow_area;
        if (prof_ds) flow_area_ds = fp_flow_area(prof_ds + nd, prof_ds + 2 * nd, depth0, h, i1, i2);
        if (flow_area > 1.0e-05 && flow_area_ds > 1.0e-05) {
            const double wetted_perimeter = Wv[2 * nd + i1] + 2.0 * (h - depth0[i1]);
            const double hydraulic_radius = flow_area / wetted_perimeter;
            /* manning_flow  surface_process.jl:166-169 */
            cap = wfh_cbrt(hydraulic_radius * hydraulic_radius) * sqrt(slope) * flow_area / mannings_n;
        }
    }
    return cap;
}

/* update_floodplain_model!  surface_kinwave.jl:387-430: Manning flow capacity from the flood
 * depth (zero unless both this cell's and the downstream cell's floodplain flow area exceed
 * 1e-5 m2), accucapacityflux! of the floodplain storage down the river graph in its
 * topological order (routing/utils.jl:82-108, the storage is updated in place), flux_in!
 * (:162-167) over the upstream lists without reservoir outlets */
static void update_floodplain(wfo_model* m, double dt) {
    const int64_t nd = m->fp_nd, nriv = m->nriv;
    for (int64_t k = 0; k < nriv; ++k) {
        const int64_t v = m->rorder[k], ds = m->rdown[v];
        m->fp_cap[v] = wfo_floodplain_capacity(m->fp_prof + v * 4 * nd, ds >= 0 ? m->fp_prof + ds * 4 * nd : NULL, nd,
                                               m->fp_depth, m->fp_h[v], m->fp_n[v], m->fp_slope[v]);
    }
    /* the same accucapacityflux! as the snow transport's, over the river graph and its order */
    wfo_accucapacityflux(nriv, m->rorder, m->rdown, m->fp_storage, m->fp_cap, m->fp_q, dt);
    for (int64_t r = 0; r < nriv; ++r) m->fp_q_cum[r] += m->fp_q[r] * dt;
    for (int64_t v = 0; v < nriv; ++v) {
        double s = 0.0;
        for (int64_t e = m->rfup_ptr[v]; e < m->rfup_ptr[v + 1]; ++e) s = s + m->fp_q[m->rfup_idx[e]];
        m->fp_qin[v] = s;
    }
    for (int64_t r = 0; r < nriv; ++r) m->fp_qin_cum[r] += m->fp_qin[r] * dt;
}

int32_t wfo_step(wfo_model* m, const double* precip, const double* pet, const double* temp) {
    const int64_t n = m->n, nriv = m->nriv;
    const double dt = m->cfg.dt;
    const int snow = m->cfg.snow;
    memcpy(m->precip, precip, sizeof(double) * (size_t)n);
    memcpy(m->pet, pet, sizeof(double) * (size_t)n);
    memcpy(m->temp, temp, sizeof(double) * (size_t)n);
    /* precipitation and potential evaporation over each reservoir area: their mean goes to the
     * reservoir and the land model gets zero there (update_forcing!  io.jl:125-143; the mean is a
     * left fold over the area's cells in ascending index, then / count) */
    for (int64_t k = 0; k < m->nres; ++k) {
        double sp = 0.0, se = 0.0;
        for (int64_t e = m->cov_ptr[k]; e < m->cov_ptr[k + 1]; ++e) {
            sp += m->precip[m->cov_idx[e]];
            se += m->pet[m->cov_idx[e]];
        }
        const double cnt = (double)(m->cov_ptr[k + 1] - m->cov_ptr[k]);
        m->res[k].precipitation = sp / cnt;
        m->res[k].evaporation = se / cnt;
    }
    for (int64_t i = 0; i < n && m->nres > 0; ++i)
        if (m->land_cov[i]) {
            m->precip[i] = 0.0;
            m->pet[i] = 0.0;
        }
    if (m->cfg.water_mass_balance) { /* storage_prev!  mass_balance.jl:156-183 */
        for (int64_t i = 0; i < n; ++i) {
            m->land_store_prev[i] = land_total_storage(m, i);
            m->ol_store_prev[i] = m->ol_storage[i];
        }
        for (int64_t r = 0; r < nriv; ++r) /* get_storage  mass_balance.jl:132-140 */
            m->r_store_prev[r] = m->fp ? m->r_storage[r] + m->fp_storage[r] : m->r_storage[r];
        for (int64_t k = 0; k < m->nres; ++k) m->res_store_prev[k] = m->res[k].storage;
    }
    /* ---- update_land_hydrology_model!  sbm.jl:82-132 ---- */
    for (int64_t i = 0; i < n; ++i) { /* interception  canopy.jl:54-98 / 115-143 */
        double o[4];
        m->canopy_potevap[i] = m->kc[i] * m->pet[i] * (1.0 - m->gap[i]);
        if (m->gash)
            wfo_rainfall_interception_gash(m->smax[i], m->er[i], m->gap[i], m->precip[i], m->canopy[i],
                                           m->canopy_potevap[i], dt, o);
        else
            wfo_rainfall_interception_modrut(m->precip[i], m->canopy_potevap[i], m->canopy[i], m->gap[i],
                                             m->smax[i], dt, o);
        m->throughfall[i] = o[0];
        m->interception_rate[i] = o[1];
        m->stemflow[i] = o[2];
        m->canopy[i] = o[3];
    }
    if (snow) { /* update_bc_snow_model! / update_snow_model!  snow.jl:123-177 */
        for (int64_t i = 0; i < n; ++i) {
            double eff = m->throughfall[i] + m->stemflow[i];
            double p2[2], o[5];
            wfo_precipitation_hbv(eff, m->temp[i], m->tti[i], m->tt[i], p2);
            wfo_snowpack_hbv(m->snow_s[i], m->snow_w[i], p2[0], p2[1], m->temp[i], m->ttm[i], m->cfmax[i], m->whc[i],
                             dt, o);
            m->snow_s[i] = o[0];
            m->snow_w[i] = o[1];
            m->swe[i] = o[2];
            m->snow_runoff[i] = o[4];
        }
        if (m->cfg.snow_gravitational_transport) lateral_snow_transport(m, dt);
    }
    /* update_bc_open_water_runoff_model! / update_open_water_runoff_model!  runoff.jl:37-111 */
    for (int64_t i = 0; i < n; ++i) {
        m->wfs[i] = snow ? m->snow_runoff[i] + 0.0 * 0.0 : m->throughfall[i] + m->stemflow[i];
        double wd_land = m->ol_h[i];
        double wd_river = m->land_river[i] >= 0 ? m->r_h[m->land_river[i]] : 0.0;
        m->runoff_river[i] = wfh_min(1.0, m->river_fraction[i]) * m->wfs[i];
        m->runoff_land[i] = wfh_min(1.0, m->water_fraction[i]) * m->wfs[i];
        m->aeow_r[i] = m->river_fraction[i] * wfh_min(wd_river / dt, m->pet[i]);
        m->aeow_l[i] = m->water_fraction[i] * wfh_min(wd_land / dt, m->pet[i]);
        m->net_runoff_river[i] = m->runoff_river[i] - m->aeow_r[i];
    }
    for (int64_t i = 0; i < n; ++i) { /* soil_fraction!, update_bc_soil_model!, update_soil_water_flow! */
        wfo_soil* s = &m->soil[i];
        wfo_soil_update_bc(s, m->canopy_potevap[i], m->interception_rate[i], m->pet[i], m->wfs[i],
                           m->runoff_river[i], m->runoff_land[i], dt);
        wfo_soil_water_flow(s, m->temp[i], snow, m->cfg.soil_infiltration_reduction, m->aeow_r[i], m->aeow_l[i], dt);
        s->aet = s->aet + m->interception_rate[i];
    }
    /* ---- recharge and water-table hand-off (sbm_model.jl:74-81) ---- */
    for (int64_t i = 0; i < n; ++i) {
        m->rech_rate[i] = m->soil[i].recharge;
        m->ssf_zi[i] = m->soil[i].zi;
    }
    m->nsub[0] = m->nsub[1] = m->nsub[2] = 0;
    /* ---- update_subsurface_flow_model!  lateral_subsurface_flow.jl:279-304 ---- */
    for (int64_t i = 0; i < n; ++i) {
        m->ssf_tor_cum[i] = 0.0;
        m->rech_flux_cum[i] = 0.0;
        m->ssf_exf_cum[i] = 0.0;
        m->ssf_qin_cum[i] = 0.0;
        m->ssf_q_cum[i] = 0.0;
        m->ssf_qnet_cum[i] = 0.0;
    }
    {
        double t = 0.0;
        while (t < dt) {
            double dts = m->cfg.kinematic_wave_adaptive ? ssf_stable_timestep(m) : m->cfg.subsurface_dt;
            dts = check_timestepsize(dts, t, dt);
            ++m->nsub[0];
            for (int64_t i = 0; i < n; ++i) { /* flux!(::RechargeModel)  boundary_conditions.jl:219-236 */
                m->qnet_bnds[i] = 0.0;
                double fl = m->rech_rate[i] * m->area[i];
                if (m->ssf_zi[i] >= m->ssf_d[i]) fl = wfh_max(0.0, fl);
                m->rech_flux_cum[i] += fl * dts;
                m->qnet_bnds[i] += fl;
            }
            const double* sq_up = m->ssf_q;
            if (m->mutation == WFO_MUT_STALE_SSF) {
                memcpy(m->stale, m->ssf_q, (size_t)n * sizeof(double));
                sq_up = m->stale;
            }
            for (int64_t k = 0; k < n; ++k) { /* kinwave_subsurface_update!  :198-273 */
                int64_t v = m->order[k];
                double qin = 0.0, tor = 0.0;
                for (int64_t e = m->fup_ptr[v]; e < m->fup_ptr[v + 1]; ++e) {
                    int64_t u = m->fup_idx[e];
                    qin = qin + sq_up[u] * (1.0 - m->fftr[u]);
                }
                for (int64_t e = m->fup_ptr[v]; e < m->fup_ptr[v + 1]; ++e) {
                    int64_t u = m->fup_idx[e];
                    tor = tor + sq_up[u] * m->fftr[u];
                }
                m->ssf_qin[v] = qin;
                m->ssf_tor_cum[v] += tor * dts;
                double o[4];
                wfo_kinematic_wave_ssf(qin, m->ssf_q[v], m->ssf_zi[v], m->qnet_bnds[v], m->slope[v], m->sy[v],
                                       m->ssf_d[v], dts, m->flow_length[v], m->flow_width[v], m->qmax[v], m->kh0[v],
                                       m->soil[v].f, m->soil[v].z_exp, &m->soil[v], o);
                m->ssf_q[v] = o[0];
                m->ssf_zi[v] = o[1];
                m->ssf_qin_cum[v] += m->ssf_qin[v] * dts;
                m->ssf_q_cum[v] += m->ssf_q[v] * dts;
                m->ssf_exf_cum[v] += o[2] * dts;
                m->ssf_qnet_cum[v] += o[3] * m->area[v] * dts;
                m->ssf_head[v] = m->elevation[v] - m->ssf_zi[v];
                m->ssf_storage[v] = m->sy[v] * (m->ssf_d[v] - m->ssf_zi[v]) * m->area[v];
            }
            t += dts;
        }
        for (int64_t i = 0; i < n; ++i) { /* average_flux_vars!  groundwater.jl:630-636; flux_to_river! */
            m->rech_flux_avg[i] = m->rech_flux_cum[i] / dt;
            m->ssf_qin_avg[i] = m->ssf_qin_cum[i] / dt;
            m->ssf_q_avg[i] = m->ssf_q_cum[i] / dt;
            m->ssf_qnet_avg[i] = m->ssf_qnet_cum[i] / dt;
            m->ssf_exf_avg[i] = m->ssf_exf_cum[i] / dt;
            m->ssf_tor_avg[i] = m->ssf_tor_cum[i] / dt;
        }
    }
    /* ---- update_soil_water_storage!  soil.jl:1294-1392 ---- */
    for (int64_t i = 0; i < n; ++i)
        wfo_soil_water_storage(&m->soil[i], m->ssf_exf_avg[i], m->runoff_land[i], m->aeow_l[i]);
    /* ---- surface_routing!  surface_routing.jl:7-46 ---- */
    for (int64_t i = 0; i < n; ++i) { /* update_lateral_inflow! (overland)  surface_kinwave.jl:740-766 */
        m->ol_inwater[i] = (m->soil[i].net_runoff + 0.0) * m->area[i] + 0.0;
        m->ol_qlat[i] = m->ol_inwater[i] / m->flow_length[i];
        m->ol_q_cum[i] = 0.0;
        m->ol_qin_cum[i] = 0.0;
        m->ol_tor_cum[i] = 0.0;
    }
    {
        double t = 0.0;
        while (t < dt) { /* update_overland_flow_model!  :347-385 */
            double dts = m->cfg.kinematic_wave_adaptive
                             ? wfo_kw_stable_timestep(m->ol_q, m->alpha_land, m->flow_length, n, 0.02, m->stbuf)
                             : m->cfg.land_dt;
            dts = check_timestepsize(dts, t, dt);
            ++m->nsub[1];
            for (int64_t i = 0; i < n; ++i) m->ol_qin[i] = 0.0;
            const double* oq_up = m->ol_q;
            if (m->mutation == WFO_MUT_STALE_LAND) {
                memcpy(m->stale, m->ol_q, (size_t)n * sizeof(double));
                oq_up = m->stale;
            }
            for (int64_t k = 0; k < n; ++k) { /* kinwave_land_update!  :293-341 */
                int64_t v = m->order[k];
                double tor = 0.0;
                for (int64_t e = m->fup_ptr[v]; e < m->fup_ptr[v + 1]; ++e) {
                    int64_t u = m->fup_idx[e];
                    tor = tor + oq_up[u] * m->fftr[u];
                }
                m->ol_tor_cum[v] += tor * dts;
                if (m->sfw[v] > 0.0) {
                    double qin = 0.0;
                    for (int64_t e = m->fup_ptr[v]; e < m->fup_ptr[v + 1]; ++e) {
                        int64_t u = m->fup_idx[e];
                        qin = qin + oq_up[u] * (1.0 - m->fftr[u]);
                    }
                    m->ol_qin[v] = qin;
                }
                double o[2];
                wfo_kinematic_wave_it(m->ol_qin[v], m->ol_q[v], m->ol_qlat[v], m->alpha_land[v], dts,
                                      m->flow_length[v], o, &m->kw_iters);
                m->ol_q[v] = o[0];
                if (m->sfw[v] > 0.0) m->ol_h[v] = o[1] / m->sfw[v];
                m->ol_storage[v] = m->flow_length[v] * m->sfw[v] * m->ol_h[v];
                m->ol_q_cum[v] += m->ol_q[v] * dts;
                m->ol_qin_cum[v] += m->ol_qin[v] * dts;
            }
            t += dts;
        }
        for (int64_t i = 0; i < n; ++i) {
            m->ol_q_avg[i] = m->ol_q_cum[i] / dt;
            m->ol_tor_avg[i] = m->ol_tor_cum[i] / dt;
            m->ol_qin_avg[i] = m->ol_qin_cum[i] / dt;
        }
    }
    for (int64_t k = 0; k < m->nres; ++k) {
        /* update_inflow!  surface_kinwave.jl:772-788; set_reservoir_vars!  :227-237 */
        wfo_reservoir* R = &m->res[k];
        R->inflow_overland = m->ol_q_avg[m->res_land[k]];
        R->inflow_subsurface = m->ssf_q_avg[m->res_land[k]];
        R->inflow_cumulative = 0.0;
        R->abstraction_cumulative = 0.0;
        R->outflow_cumulative = 0.0;
        R->actevap_cumulative = 0.0;
    }
    for (int64_t r = 0; r < nriv; ++r) { /* update_lateral_inflow! (river)  :710-734 */
        int64_t i = m->river_land[r];
        m->r_inwater[r] = (m->ssf_tor_avg[i] + m->ol_tor_avg[i] + m->net_runoff_river[i] * m->area[i]) +
                          0.0 * m->r_area[r];
        m->r_qlat[r] = m->r_inwater[r] / m->r_fl[r];
        m->r_q_cum[r] = 0.0;
        m->r_qin_cum[r] = 0.0;
    }
    if (m->fp)
        for (int64_t r = 0; r < nriv; ++r) {
            m->fp_q_cum[r] = 0.0;
            m->fp_qin_cum[r] = 0.0;
        }
    {
        double t = 0.0;
        while (t < dt) { /* update_river_flow_model!  :613-662 */
            double dts = m->cfg.kinematic_wave_adaptive
                             ? wfo_kw_stable_timestep(m->r_q, m->r_alpha, m->r_fl, nriv, 0.05, m->stbuf)
                             : m->cfg.river_dt;
            dts = check_timestepsize(dts, t, dt);
            ++m->nsub[2];
            if (m->fp) river_floodplain_exchange(m, dts);
            for (int64_t r = 0; r < nriv; ++r) m->r_qin[r] = 0.0;
            const double* rq_up = m->r_q;
            if (m->mutation == WFO_MUT_STALE_RIVER) {
                memcpy(m->stale, m->r_q, (size_t)nriv * sizeof(double));
                rq_up = m->stale;
            }
            for (int64_t k = 0; k < nriv; ++k) { /* kinwave_river_update!  :492-566 */
                int64_t v = m->rorder[k];
                double s = 0.0;
                for (int64_t e = m->rfup_ptr[v]; e < m->rfup_ptr[v + 1]; ++e) s = s + rq_up[m->rfup_idx[e]];
                m->r_qin[v] += s;
                double inflow = 0.0 / m->r_fl[v];
                inflow -= 0.0 / m->r_fl[v];
                if (m->fp) inflow += m->fp_exch[v] / m->r_fl[v];
                double o[2];
                wfo_kinematic_wave_it(m->r_qin[v], m->r_q[v], m->r_qlat[v] + inflow, m->r_alpha[v], dts, m->r_fl[v],
                                      o, &m->kw_iters);
                m->r_q[v] = o[0];
                if (m->river_res[v] >= 0) /* the reservoir's outflow is the downstream inflow */
                    m->r_qin[m->rdown[v]] = wfo_update_reservoir_at_river(&m->res[m->river_res[v]], m->r_q[v], dts);
                m->r_h[v] = o[1] / m->r_fw[v];
                m->r_storage[v] = m->r_fl[v] * o[1];
                m->r_q_cum[v] += m->r_q[v] * dts;
                m->r_qin_cum[v] += m->r_qin[v] * dts;
            }
            if (m->fp) update_floodplain(m, dts);
            t += dts;
        }
        for (int64_t r = 0; r < nriv; ++r) {
            m->r_q_avg[r] = m->r_q_cum[r] / dt;
            m->r_qin_avg[r] = m->r_qin_cum[r] / dt;
        }
        if (m->fp) /* surface_kinwave.jl:652-659 */
            for (int64_t r = 0; r < nriv; ++r) {
                m->fp_q_avg[r] = m->fp_q_cum[r] / dt;
                m->r_qch_avg[r] = m->r_q_avg[r];
                m->r_q_avg[r] = m->r_qch_avg[r] + m->fp_q_avg[r];
                m->fp_qin_avg[r] = m->fp_qin_cum[r] / dt;
                m->r_qin_avg[r] = m->r_qin_avg[r] + m->fp_qin_avg[r];
            }
        for (int64_t k = 0; k < m->nres; ++k) { /* average_reservoir_vars!  :244-258 */
            m->res_outflow_avg[k] = m->res[k].outflow_cumulative / dt;
            m->res_inflow_avg[k] = m->res[k].inflow_cumulative / dt;
            m->res_abs_avg[k] = m->res[k].abstraction_cumulative / dt;
        }
    }
    /* ---- update_total_water_storage!  sbm.jl:143-182 ---- */
    for (int64_t i = 0; i < n; ++i) m->total_storage[i] = 0.0;
    for (int64_t r = 0; r < nriv; ++r) {
        int64_t i = m->river_land[r];
        m->total_storage[i] = (m->r_h[r] * m->r_fw[r] * m->r_fl[r]) / (m->area[i]);
    }
    for (int64_t i = 0; i < n; ++i) {
        double ss = snow ? m->snow_s[i] : 0.0, sw = snow ? m->snow_w[i] : 0.0;
        m->total_storage[i] = m->total_storage[i] + ((((ss + sw) + 0.0 * 0.0) + m->canopy[i]) + 0.0);
        double sub = m->soil[i].ustoredepth + m->soil[i].satwaterdepth;
        double lat = m->ol_h[i] * (1.0 - m->river_fraction[i]);
        m->total_storage[i] += sub + lat;
    }
    if (m->cfg.water_mass_balance) { /* compute_mass_balance!  mass_balance.jl:215-549 */
        for (int64_t i = 0; i < n; ++i) {
            double a = m->area[i];
            double tin = m->ssf_qin_avg[i] / a + (m->precip[i] + 0.0) + m->snow_in[i] + 0.0 / m->area[i];
            double vout = m->soil[i].net_runoff + m->soil[i].aet + m->net_runoff_river[i] + m->soil[i].leakage + 0.0;
            double tout = m->ssf_q_avg[i] / a + vout + m->snow_out[i] + 0.0 / m->area[i];
            double rate = (land_total_storage(m, i) - m->land_store_prev[i]) / dt;
            double o[2];
            wfo_compute_mass_balance_error(tin, tout, rate, o);
            m->bal_land[i] = o[0];
            m->bal_land_rel[i] = o[1];
            /* subsurface (recharge is a boundary flux here) */
            double fin = 0.0, fout = 0.0, fa = m->rech_flux_avg[i];
            if (fa > 0.0) fin += fa;
            else fout -= fa;
            wfo_compute_mass_balance_error(m->ssf_qin_avg[i] + fin, m->ssf_q_avg[i] + fout, m->ssf_qnet_avg[i], o);
            m->bal_ssf[i] = o[0];
            /* overland */
            wfo_compute_mass_balance_error(m->ol_inwater[i] + m->ol_qin_avg[i], m->ol_q_avg[i],
                                           (m->ol_storage[i] - m->ol_store_prev[i]) / dt, o);
            m->bal_ol[i] = o[0];
        }
        for (int64_t r = 0; r < nriv; ++r) {
            double o[2];
            const double storage = m->fp ? m->r_storage[r] + m->fp_storage[r] : m->r_storage[r];
            wfo_compute_mass_balance_error(m->r_inwater[r] + m->r_qin_avg[r] + wfh_max(0.0, 0.0),
                                           m->r_q_avg[r] + 0.0 + 0.0, (storage - m->r_store_prev[r]) / dt, o);
            m->bal_r[r] = o[0];
        }
        for (int64_t k = 0; k < m->nres; ++k) { /* compute_flow_balance!(::ReservoirModel)  :261-283 */
            const wfo_reservoir* R = &m->res[k];
            double o[2];
            m->res_inflow_avg[k] = R->inflow_cumulative / dt;
            m->res_actevap_avg[k] = R->actevap_cumulative / dt;
            double total_in = m->res_inflow_avg[k] + R->precipitation * R->area;
            double total_out = m->res_outflow_avg[k] + m->res_actevap_avg[k] * R->area;
            wfo_compute_mass_balance_error(total_in, total_out, (R->storage - m->res_store_prev[k]) / dt, o);
            m->bal_res[k] = o[0];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* accessors                                                                             */
/* ------------------------------------------------------------------------------------ */
int64_t wfo_n(const wfo_model* m) { return m->n; }
int64_t wfo_nriv(const wfo_model* m) { return m->nriv; }
int32_t wfo_nlayers(const wfo_model* m) { return m->N; }
int64_t wfo_kinwave_iterations(const wfo_model* m) { return m->kw_iters; }
int32_t wfo_set_mutation(wfo_model* m, int32_t mutation) {
    if (mutation < WFO_MUT_NONE || mutation > WFO_MUT_STALE_RIVER) return -1;
    m->mutation = mutation;
    return 0;
}
int32_t wfo_set_julian_day(wfo_model* m, int32_t julian_day) {
    if (julian_day < 1 || julian_day > 365) return -1;
    for (int64_t k = 0; k < m->nres; ++k) m->res[k].hq_col = julian_day - 1; /* update_index_hq! */
    return 0;
}

void wfo_last_substeps(const wfo_model* m, int32_t* out3) {
    for (int k = 0; k < 3; ++k) out3[k] = m->nsub[k];
}

int32_t wfo_topology(const wfo_model* m, int64_t* land_order, int64_t* land_down, int64_t* river_order,
                     int64_t* river_down, int64_t* river_land_index) {
    if (land_order) memcpy(land_order, m->order, sizeof(int64_t) * (size_t)m->n);
    if (land_down) memcpy(land_down, m->down, sizeof(int64_t) * (size_t)m->n);
    if (river_order) memcpy(river_order, m->rorder, sizeof(int64_t) * (size_t)m->nriv);
    if (river_down) memcpy(river_down, m->rdown, sizeof(int64_t) * (size_t)m->nriv);
    if (river_land_index) memcpy(river_land_index, m->river_land, sizeof(int64_t) * (size_t)m->nriv);
    return 0;
}

#define SOILF(field)                                                  \
    do {                                                              \
        for (int64_t i = 0; i < n; ++i) dst[i] = m->soil[i].field;    \
        return 0;                                                     \
    } while (0)
#define ARR(arr, len)                                                 \
    do {                                                              \
        memcpy(dst, (arr), sizeof(double) * (size_t)(len));           \
        return 0;                                                     \
    } while (0)

int32_t wfo_get(const wfo_model* m, int32_t var, double* dst, int64_t count) {
    const int64_t n = m->n, nr = m->nriv;
    const int N = m->N;
    int64_t need = n;
    if (var == WFH_USTORELAYERDEPTH || var == WFH_VWC) need = (int64_t)N * n;
    if ((var >= WFH_RIVER_Q && var <= WFH_RIVER_INWATER) || var == WFH_BALANCE_RIVER_ERROR || var == WFH_RIVER_ALPHA ||
        (var >= WFH_FLOODPLAIN_STORAGE && var <= WFH_RIVER_Q_CHANNEL_AVERAGE))
        need = nr;
    if (var >= WFH_RESERVOIR_STORAGE && var <= WFH_BALANCE_RESERVOIR_ERROR) {
        if (count != m->nres) return WFH_ERR_ARG;
        for (int64_t k = 0; k < m->nres; ++k) {
            const wfo_reservoir* R = &m->res[k];
            double v = 0.0;
            switch (var) {
            case WFH_RESERVOIR_STORAGE: v = R->storage; break;
            case WFH_RESERVOIR_WATERLEVEL: v = R->waterlevel; break;
            case WFH_RESERVOIR_OUTFLOW: v = R->outflow; break;
            case WFH_RESERVOIR_OUTFLOW_AVERAGE: v = m->res_outflow_avg[k]; break;
            case WFH_RESERVOIR_INFLOW_AVERAGE: v = m->res_inflow_avg[k]; break;
            case WFH_RESERVOIR_ACTEVAP_AVERAGE: v = m->res_actevap_avg[k]; break;
            case WFH_RESERVOIR_ABSTRACTION_AVERAGE: v = m->res_abs_avg[k]; break;
            case WFH_RESERVOIR_PRECIPITATION: v = R->precipitation; break;
            case WFH_RESERVOIR_EVAPORATION: v = R->evaporation; break;
            default: v = m->bal_res[k]; break;
            }
            dst[k] = v;
        }
        return 0;
    }
    if (count != need) return WFH_ERR_ARG;
    switch (var) {
    case WFH_CANOPY_STORAGE: ARR(m->canopy, n);
    case WFH_SNOW_STORAGE: ARR(m->snow_s, n);
    case WFH_SNOW_WATER: ARR(m->snow_w, n);
    case WFH_SOIL_TEMPERATURE: SOILF(tsoil);
    case WFH_SATWATERDEPTH: SOILF(satwaterdepth);
    case WFH_USTORELAYERDEPTH:
        for (int k = 0; k < N; ++k)
            for (int64_t i = 0; i < n; ++i) dst[(int64_t)k * n + i] = m->soil[i].ustore[k];
        return 0;
    case WFH_WATER_TABLE_DEPTH: SOILF(zi);
    case WFH_USTOREDEPTH: SOILF(ustoredepth);
    case WFH_TRANSFER: SOILF(transfer);
    case WFH_RECHARGE: SOILF(recharge);
    case WFH_ACTUAL_INFILTRATION: SOILF(actual_infiltration);
    case WFH_INFILTRATION_EXCESS: SOILF(infiltration_excess);
    case WFH_SATURATION_EXCESS: SOILF(saturation_excess);
    case WFH_TRANSPIRATION: SOILF(transpiration);
    case WFH_SOIL_EVAPORATION: SOILF(soil_evaporation);
    case WFH_ACTUAL_CAPILLARY_FLUX: SOILF(capflux);
    case WFH_ACTUAL_LEAKAGE: SOILF(leakage);
    case WFH_ACTUAL_EVAPOTRANSPIRATION: SOILF(aet);
    case WFH_INTERCEPTION_RATE: ARR(m->interception_rate, n);
    case WFH_NET_RUNOFF: SOILF(net_runoff);
    case WFH_RUNOFF: SOILF(runoff);
    case WFH_NET_RUNOFF_RIVER: ARR(m->net_runoff_river, n);
    case WFH_EXFILTRATION: SOILF(exfiltration);
    case WFH_TOTAL_STORAGE: ARR(m->total_storage, n);
    case WFH_SNOW_WATER_EQUIVALENT: ARR(m->swe, n);
    case WFH_VWC:
        for (int k = 0; k < N; ++k)
            for (int64_t i = 0; i < n; ++i) dst[(int64_t)k * n + i] = k < m->soil[i].nl ? m->soil[i].vwc[k] : NaN;
        return 0;
    case WFH_VWC_ROOT_ZONE: SOILF(vwc_root);
    case WFH_ROOT_ZONE_STORAGE: SOILF(root_zone_storage);
    case WFH_SSF_Q: ARR(m->ssf_q, n);
    case WFH_SSF_QIN_AVERAGE: ARR(m->ssf_qin_avg, n);
    case WFH_SSF_Q_AVERAGE: ARR(m->ssf_q_avg, n);
    case WFH_SSF_TO_RIVER_AVERAGE: ARR(m->ssf_tor_avg, n);
    case WFH_SSF_STORAGE: ARR(m->ssf_storage, n);
    case WFH_SSF_QMAX: ARR(m->qmax, n);
    case WFH_LAND_Q: ARR(m->ol_q, n);
    case WFH_LAND_QIN: ARR(m->ol_qin, n);
    case WFH_LAND_H: ARR(m->ol_h, n);
    case WFH_LAND_STORAGE: ARR(m->ol_storage, n);
    case WFH_LAND_Q_AVERAGE: ARR(m->ol_q_avg, n);
    case WFH_LAND_QIN_AVERAGE: ARR(m->ol_qin_avg, n);
    case WFH_LAND_TO_RIVER_AVERAGE: ARR(m->ol_tor_avg, n);
    case WFH_LAND_INWATER: ARR(m->ol_inwater, n);
    case WFH_RIVER_Q: ARR(m->r_q, nr);
    case WFH_RIVER_QIN: ARR(m->r_qin, nr);
    case WFH_RIVER_H: ARR(m->r_h, nr);
    case WFH_RIVER_STORAGE: ARR(m->r_storage, nr);
    case WFH_RIVER_Q_AVERAGE: ARR(m->r_q_avg, nr);
    case WFH_RIVER_QIN_AVERAGE: ARR(m->r_qin_avg, nr);
    case WFH_RIVER_INWATER: ARR(m->r_inwater, nr);
    case WFH_BALANCE_LAND_ERROR: ARR(m->bal_land, n);
    case WFH_BALANCE_LAND_RELERROR: ARR(m->bal_land_rel, n);
    case WFH_BALANCE_SSF_ERROR: ARR(m->bal_ssf, n);
    case WFH_BALANCE_OVERLAND_ERROR: ARR(m->bal_ol, n);
    case WFH_BALANCE_RIVER_ERROR: ARR(m->bal_r, nr);
    case WFH_LAND_ALPHA: ARR(m->alpha_land, n);
    case WFH_RIVER_ALPHA: ARR(m->r_alpha, nr);
    case WFH_THETA_FC: SOILF(theta_fc);
    case WFH_SURFACE_FLOW_WIDTH: ARR(m->sfw, n);
    case WFH_FLOW_FRACTION_TO_RIVER: ARR(m->fftr, n);
    case WFH_SNOW_IN: ARR(m->snow_in, n);
    case WFH_SNOW_OUT: ARR(m->snow_out, n);
    case WFH_FLOODPLAIN_STORAGE: if (!m->fp) return WFH_ERR_ARG; ARR(m->fp_storage, nr);
    case WFH_FLOODPLAIN_H: if (!m->fp) return WFH_ERR_ARG; ARR(m->fp_h, nr);
    case WFH_FLOODPLAIN_Q: if (!m->fp) return WFH_ERR_ARG; ARR(m->fp_q, nr);
    case WFH_FLOODPLAIN_Q_AVERAGE: if (!m->fp) return WFH_ERR_ARG; ARR(m->fp_q_avg, nr);
    case WFH_FLOODPLAIN_QIN_AVERAGE: if (!m->fp) return WFH_ERR_ARG; ARR(m->fp_qin_avg, nr);
    case WFH_RIVER_Q_CHANNEL_AVERAGE: if (!m->fp) return WFH_ERR_ARG; ARR(m->r_qch_avg, nr);
    default: return WFH_ERR_ARG;
    }
}

/* set_value on prognostic states (BMI.set_value  bmi.jl:243-245) */
int32_t wfo_set(wfo_model* m, int32_t var, const double* src, int64_t count) {
    const int64_t n = m->n, nr = m->nriv;
    switch (var) {
    case WFH_CANOPY_STORAGE: if (count != n) return WFH_ERR_ARG; memcpy(m->canopy, src, 8 * (size_t)n); return 0;
    case WFH_SNOW_STORAGE: if (count != n) return WFH_ERR_ARG; memcpy(m->snow_s, src, 8 * (size_t)n); return 0;
    case WFH_SNOW_WATER: if (count != n) return WFH_ERR_ARG; memcpy(m->snow_w, src, 8 * (size_t)n); return 0;
    case WFH_SOIL_TEMPERATURE:
        if (count != n) return WFH_ERR_ARG;
        for (int64_t i = 0; i < n; ++i) m->soil[i].tsoil = src[i];
        return 0;
    case WFH_SATWATERDEPTH:
        if (count != n) return WFH_ERR_ARG;
        for (int64_t i = 0; i < n; ++i) m->soil[i].satwaterdepth = src[i];
        return 0;
    case WFH_USTORELAYERDEPTH:
        if (count != (int64_t)m->N * n) return WFH_ERR_ARG;
        for (int k = 0; k < m->N; ++k)
            for (int64_t i = 0; i < n; ++i) m->soil[i].ustore[k] = src[(int64_t)k * n + i];
        return 0;
    case WFH_SSF_Q: if (count != n) return WFH_ERR_ARG; memcpy(m->ssf_q, src, 8 * (size_t)n); return 0;
    case WFH_LAND_Q: if (count != n) return WFH_ERR_ARG; memcpy(m->ol_q, src, 8 * (size_t)n); return 0;
    case WFH_LAND_H: if (count != n) return WFH_ERR_ARG; memcpy(m->ol_h, src, 8 * (size_t)n); return 0;
    case WFH_RIVER_Q: if (count != nr) return WFH_ERR_ARG; memcpy(m->r_q, src, 8 * (size_t)nr); return 0;
    case WFH_RIVER_H: if (count != nr) return WFH_ERR_ARG; memcpy(m->r_h, src, 8 * (size_t)nr); return 0;
    case WFH_RIVER_STORAGE: if (count != nr) return WFH_ERR_ARG; memcpy(m->r_storage, src, 8 * (size_t)nr); return 0;
    case WFH_LAND_STORAGE: if (count != n) return WFH_ERR_ARG; memcpy(m->ol_storage, src, 8 * (size_t)n); return 0;
    case WFH_USTOREDEPTH: /* with the total soil water storage it completes (after SATWATERDEPTH) */
        if (count != n) return WFH_ERR_ARG;
        for (int64_t i = 0; i < n; ++i) {
            m->soil[i].ustoredepth = src[i];
            m->soil[i].total_soil_water_storage = m->soil[i].satwaterdepth + src[i];
        }
        return 0;
    case WFH_FLOODPLAIN_STORAGE:
        if (count != nr || !m->fp) return WFH_ERR_ARG;
        memcpy(m->fp_storage, src, 8 * (size_t)nr);
        return 0;
    case WFH_RESERVOIR_STORAGE:
    case WFH_RESERVOIR_WATERLEVEL:
        if (count != m->nres) return WFH_ERR_ARG;
        for (int64_t k = 0; k < m->nres; ++k) {
            if (var == WFH_RESERVOIR_STORAGE) m->res[k].storage = src[k];
            else m->res[k].waterlevel = src[k];
        }
        return 0;
    default: return WFH_ERR_ARG;
    }
}

int32_t wfo_math(int32_t op, const double* x, const double* y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (op) {
            case 0: out[i] = wfh_exp(x[i]); break;
            case 1: out[i] = wfh_log(x[i]); break;
            case 2: out[i] = wfh_pow(x[i], y[i]); break;
            case 3: out[i] = wfh_cbrt(x[i]); break;
            case 4: out[i] = wfh_cld(x[i], y[i]); break;
            case 5: out[i] = wfh_round_sig12(x[i]); break;
            case 6: out[i] = wfh_min(x[i], y[i]); break;
            case 7: out[i] = wfh_max(x[i], y[i]); break;
            case 8: out[i] = wfh_pow10i((int)x[i]); break;
            default: return -1;
        }
    }
    return 0;
}
