// Device (and host) restatement of the per-cell physics of the Wflow.jl `sbm` step, written for
// register-resident columns: one thread owns one cell, the N soil layers live in fixed-size
// arrays, parameters come from structure-of-arrays HBM buffers.
//
// The arithmetic follows the reference expression by expression (same association order, no
// FMA contraction: the library is built with -ffp-contract=off) so that results are bit-identical
// to a host build of the same algorithm. Citations are to the reference Julia sources.
#pragma once

#include "wfh_math.h"

#define WFH_HD __host__ __device__ inline

namespace wfh {

constexpr double NaN_ = __builtin_nan("");

// ---- canopy interception ---------------------------------------------------------------
// rainfall_interception_gash  Wflow/src/vegetation/rainfall_interception.jl:9-70
WFH_HD void interception_gash(double smax, double er, double gap, double precip, double max_evap, double dt,
                              double& throughfall, double& interception, double& stemflow) {
    if (smax > 0.0) {
        double fs, fi, psat;
        if (gap < 1.0 / 1.1) {
            fs = 0.1 * gap;
            fi = 1.0 - 1.1 * gap;
            psat = (er > fi) ? 0.0 : -smax / (er * dt) * wfh_log(1.0 - er / fi);
        } else {
            fs = 1.0 - gap;
            fi = 0.0;
            psat = 0.0;
        }
        if (precip > psat) {
            const double iwet = fi * psat - smax / dt;
            const double isat = er * (precip - psat);
            const double idry = smax / dt;
            interception = iwet + isat + idry;
        } else {
            interception = fi * precip;
        }
        stemflow = fs * precip;
        throughfall = precip - interception - stemflow;
        if (interception > max_evap) {
            const double drain = interception - max_evap;
            interception = max_evap;
            throughfall += drain;
        }
    } else {
        throughfall = precip;
        interception = 0.0;
        stemflow = 0.0;
    }
}

// rainfall_interception_modrut  Wflow/src/vegetation/rainfall_interception.jl:78-130
WFH_HD void interception_modrut(double precip, double pot_evap, double gap, double smax, double dt,
                                double& storage, double& throughfall, double& evap, double& stemflow) {
    double fs, pcan;
    if (gap < 1.0 / 1.1) {
        fs = 0.1 * gap;
        pcan = (1.0 - gap - fs) * precip;
    } else {
        fs = 1.0 - gap;
        pcan = 0.0;
    }
    stemflow = fs * precip;
    throughfall = gap * precip;
    if (storage > smax) {
        const double drain = storage - smax;
        storage = smax;
        throughfall += drain / dt;
    }
    storage += pcan * dt;
    const double max_evap = storage / dt;
    if (pot_evap > max_evap) {
        evap = max_evap;
        storage = 0.0;
    } else {
        evap = pot_evap;
        storage -= evap * dt;
    }
    if (storage > smax) {
        const double drain = storage - smax;
        storage = smax;
        throughfall += drain / dt;
    }
}

// ---- HBV snow --------------------------------------------------------------------------
// precipitation_hbv (rfcf = sfcf = 1) + snowpack_hbv (cfr = 0.05)
// Wflow/src/snow/snow_process.jl:90-116 and :26-70
WFH_HD void snow_hbv(double precip, double temp, double tti, double tt, double ttm, double cfmax, double whc,
                     double dt, double& ss, double& sw, double& swe, double& runoff) {
    double rainfrac;
    if (tti == 0.0) {
        rainfrac = temp > tt ? 1.0 : 0.0;
    } else {
        const double frac = (temp - (tt - tti / 2.0)) / tti;
        rainfrac = wfh_clamp(frac, 0.0, 1.0);
    }
    const double snowfrac = 1.0 - rainfrac;
    const double sp = snowfrac * 1.0 * precip;
    const double lp = rainfrac * 1.0 * precip;
    if (temp > ttm) {
        const double pot = cfmax * (temp - ttm);
        const double melt = wfh_min(pot, ss / dt);
        ss -= melt * dt;
        sw += melt * dt;
    } else {
        const double pot_refr = cfmax * 0.05 * (ttm - temp);
        const double refr = wfh_min(pot_refr * dt, sw);
        ss += refr;
        sw -= refr;
    }
    ss = wfh_max(ss, 0.0);
    sw = wfh_max(sw, 0.0);
    ss += sp * dt;
    sw += lp * dt;
    const double maxw = ss * whc;
    if (sw > maxw) {
        runoff = (sw - maxw) / dt;
        sw = maxw;
    } else {
        runoff = 0.0;
    }
    swe = sw + ss;
}

// ---- soil helpers ------------------------------------------------------------------------
// scurve  Wflow/src/utils.jl:27-30
WFH_HD double scurve(double x, double a, double b, double c) { return 1.0 / (b + wfh_exp(-c * (x - a))); }

// bounded_power  Wflow/src/utils.jl:1070-1076
WFH_HD double bounded_power(double base, double power) { return base > 1.0 ? 1.0 : wfh_pow(base, power); }

// lower_bound_drainable_porosity  Wflow/src/utils.jl:1134-1136
WFH_HD double drainable_porosity(double ts, double tfc) { return wfh_max(ts - tfc, 0.02); }

// set_layerthickness  Wflow/src/utils.jl:390-404 (NaN marks an inactive layer); returns the
// number of active layers (number_of_active_layers, utils.jl:406-409)
template <int N>
WFH_HD int set_layerthickness(double ref, const double* cum, const double* thick, double* out) {
    int na = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double t = thick[i] * NaN_;
        if (ref > cum[i + 1])
            t = thick[i];
        else if (ref - cum[i] > 0.0)
            t = ref - cum[i];
        out[i] = t;
        na += !wfh_isnan(t);
    }
    return na;
}

// Soil layer geometry from the soil thickness (SbmSoilParameters, Wflow/src/soil/soil.jl:333-349,
// :500-503): the configured thicknesses plus a NaN-thickness last layer, clipped at the soil
// thickness; cumulative depths start at 0. cfg_thick[N-1] must be NaN and cfg_cum its cumsum.
template <int N>
WFH_HD int soil_layers(double soil_thickness, const double* cfg_cum, const double* cfg_thick, double* thick,
                       double* cum) {
    const int nl = set_layerthickness<N>(soil_thickness, cfg_cum, cfg_thick, thick);
    cum[0] = 0.0;
    cum[1] = thick[0];
#pragma unroll
    for (int k = 1; k < N; ++k) cum[k + 1] = cum[k] + thick[k];
    return nl;
}

// unsatzone_flow_layer  Wflow/src/soil/soil_process.jl:51-92
WFH_HD void unsatzone_flow_layer(double& usd, double kv_z, double l_sat, double c, double dt, double& sum_ast) {
    if (usd <= 0.0) {
        usd = 0.0;
        sum_ast = 0.0;
        return;
    }
    const double st_sat = wfh_max(0.0, usd - l_sat);
    double st = kv_z * bounded_power(usd / l_sat, c);
    sum_ast = wfh_min(st, st_sat / dt);
    usd -= sum_ast * dt;
    const double remainder = wfh_min((st - sum_ast) * dt, usd);
    const double its_d = wfh_cld(remainder, 2e-4);
    const long its = (long)its_d;
    for (long i = 0; i < its; ++i) {
        st = (kv_z / its_d) * bounded_power(usd / l_sat, c);
        const double st_max = usd / dt;
        if (st < st_max) {
            usd -= st * dt;
            sum_ast += st;
        } else {
            usd = 0.0;
            sum_ast += st_max;
            break;
        }
    }
}

// head_brooks_corey  Wflow/src/soil/soil_process.jl:113-130
WFH_HD double head_brooks_corey(double vwc, double ts, double tr, double c, double hb) {
    const double lam = 2.0 / (c - 3.0);
    if (lam > 0.0) return hb / wfh_pow(vwc / (ts - tr), 1.0 / lam);
    return hb;
}

// feddes_h3  Wflow/src/soil/soil_process.jl:166-176 (from_SI(., MM_PER_DAY), units.jl:55-67)
WFH_HD double feddes_h3(double h3_high, double h3_low, double tpot) {
    const double td = tpot / ((1.0 * (1.0 / 86400.0)) * 1e-3);
    if (td <= 1.0) return h3_low;
    if (td < 5.0) return h3_low + (h3_high - h3_low) * (td - 1.0) / (5.0 - 1.0);
    return h3_high;
}

// rwu_reduction_feddes  Wflow/src/soil/soil_process.jl:183-200
WFH_HD double rwu_reduction_feddes(double h, double h1, double h2, double h3, double h4, double alpha_h1) {
    if (h < h4) return 0.0;
    if (h < h3) return (h - h4) / (h3 - h4);
    if (alpha_h1 == 0.0) {
        if (h < h2) return 1.0;
        if (h < h1) return (h1 - h) / (h1 - h2);
        return 0.0;
    }
    return 1.0;
}

// ---- kinematic wave ------------------------------------------------------------------------
// kinematic_wave  Wflow/src/routing/surface/surface_process.jl:24-70
//
// Split in two so that a routing loop can hoist everything that does not depend on the
// upstream inflow q_in (the fifth root of q_prev above all) out of the dependent chain:
// kw_prepare(q_prev, ...) followed by kw_solve(q_in, ...) performs exactly the operations of the
// reference, in its order.
struct KwPrep {
    double dt_dx;  // dt / dx
    double u_prev; // q_prev^(1/5) (0 for negative q_prev)
    double a_u3;   // alpha * u_prev * u_prev * u_prev
    double lat;    // dt * q_lat
    double q_prev, q_lat;
};

WFH_HD KwPrep kw_prepare(double q_prev, double q_lat, double alpha, double dt, double dx) {
    KwPrep p;
    p.dt_dx = dt / dx;
    p.u_prev = q_prev >= 0.0 ? wfh_pow(q_prev, 0.2) : 0.0;
    p.a_u3 = alpha * p.u_prev * p.u_prev * p.u_prev;
    p.lat = dt * q_lat;
    p.q_prev = q_prev;
    p.q_lat = q_lat;
    return p;
}

WFH_HD void kw_solve(double q_in, const KwPrep& p, double alpha, double& q, double& area) {
    if (q_in + p.q_prev + p.q_lat == 0.0) {
        q = 0.0;
        area = 0.0;
        return;
    }
    const double dt_dx = p.dt_dx;
    const double C = dt_dx * q_in + p.a_u3 + p.lat;
    double u = p.u_prev > 0.0 ? p.u_prev : wfh_cbrt(C / alpha);
    const double c1 = 5.0 * dt_dx, c2 = 3.0 * alpha;
    // The reference stops on |f| <= 1e-12 or after 3000 updates. For large C (|C| above about
    // 1e4) the rounding noise of f exceeds 1e-12, the test can never pass and the iteration
    // settles on a fixed point or a two-cycle of adjacent doubles until the limit. The update
    // is a fixed function of u (every other input is constant), so once a value repeats the
    // rest of the sequence is known: the value after update 3000 is taken from the cycle
    // instead of running the remaining updates. Every residual test the shortcut skips would
    // have evaluated one of the cycle's values, all of which already failed it.
    double u_pp = NaN_;  // the value two updates back
    for (int k = 0; k < 3000; ++k) {
        const double u2 = u * u;
        const double u3 = u2 * u;
        const double f_u = u3 * (dt_dx * u2 + alpha) - C;
        if (wfh_fabs(f_u) <= 1.0e-12) break;
        const double df = u2 * (c1 * u2 + c2);
        const double u_c = u;
        u -= f_u / df;
        if (wfh_isnan(u) || u <= 0.0) u = WFH_KIN_WAVE_MIN_FLOW_QROOT;
        const int n = k + 1;  // updates done; u is the n-th iterate
        if (u == u_c) break;  // fixed point: every later iterate is u
        if (u == u_pp) {      // two-cycle u_c, u, u_c, u, ...: iterate 3000 is u if 3000 - n is even
            if ((3000 - n) & 1) u = u_c;
            break;
        }
        u_pp = u_c;
    }
    u = wfh_max(u, WFH_KIN_WAVE_MIN_FLOW_QROOT);
    const double u3 = u * u * u;
    q = u3 * u * u;
    area = alpha * u3;
}

// The Newton iteration of kw_solve from the start u (the same updates, stopping rule and cycle
// shortcut); returns the final u (before the KIN_WAVE_MIN_FLOW floor) and the number of updates
// the reference would take (3000 when the shortcut jumps to the limit).
WFH_HD double kw_newton(double u, const double C, const double dt_dx, const double alpha, int& updates) {
    const double c1 = 5.0 * dt_dx, c2 = 3.0 * alpha;
    double u_pp = NaN_;
    updates = 0;
    for (int k = 0; k < 3000; ++k) {
        const double u2 = u * u;
        const double u3 = u2 * u;
        const double f_u = u3 * (dt_dx * u2 + alpha) - C;
        if (wfh_fabs(f_u) <= 1.0e-12) break;
        const double df = u2 * (c1 * u2 + c2);
        const double u_c = u;
        u -= f_u / df;
        if (wfh_isnan(u) || u <= 0.0) u = WFH_KIN_WAVE_MIN_FLOW_QROOT;
        const int n = k + 1;
        updates = n;
        if (u == u_c) {
            updates = 3000;
            break;
        }
        if (u == u_pp) {
            if ((3000 - n) & 1) u = u_c;
            updates = 3000;
            break;
        }
        u_pp = u_c;
    }
    return u;
}

// kw_solve with a proposed start u0 (the learned warm start): u0 is used when it is finite,
// positive and its residual |f(u0)| is smaller than that of the reference's start; otherwise the
// reference's start (u_prev, or cbrt(C / alpha)) is kept. The iteration and its stopping rule
// are the reference's either way. With u0 = NaN the result is kw_solve's, bit for bit. Returns
// the Newton updates taken; `used` tells whether u0 was the start.
WFH_HD int kw_solve_from(double q_in, const KwPrep& p, double alpha, const double u0, double& q, double& area,
                         bool& used) {
    used = false;
    if (q_in + p.q_prev + p.q_lat == 0.0) {
        q = 0.0;
        area = 0.0;
        return 0;
    }
    const double dt_dx = p.dt_dx;
    const double C = dt_dx * q_in + p.a_u3 + p.lat;
    double u = p.u_prev > 0.0 ? p.u_prev : wfh_cbrt(C / alpha);
    if (u0 > 0.0 && u0 < INFINITY) {
        const double r2 = u * u, f_ref = r2 * u * (dt_dx * r2 + alpha) - C;
        const double l2 = u0 * u0, f_l = l2 * u0 * (dt_dx * l2 + alpha) - C;
        if (wfh_fabs(f_l) < wfh_fabs(f_ref)) {
            u = u0;
            used = true;
        }
    }
    int updates;
    u = kw_newton(u, C, dt_dx, alpha, updates);
    u = wfh_max(u, WFH_KIN_WAVE_MIN_FLOW_QROOT);
    const double u3 = u * u * u;
    q = u3 * u * u;
    area = alpha * u3;
    return updates;
}

WFH_HD void kinematic_wave(double q_in, double q_prev, double q_lat, double alpha, double dt, double dx,
                           double& q, double& area) {
    const KwPrep p = kw_prepare(q_prev, q_lat, alpha, dt, dx);
    kw_solve(q_in, p, alpha, q, area);
}

// kw_ssf_newton_raphson  Wflow/src/routing/subsurface/subsurface_process.jl:57-78
WFH_HD double kw_ssf_newton_raphson_c(double q, double C, double dt_dx, double cinv, double df) {
    for (int count = 0;; ++count) {
        const double r = dt_dx * q + cinv * q - C;
        q -= (r / df);
        if (wfh_isnan(q)) q = 0.0;
        q = wfh_max(q, WFH_KIN_WAVE_MIN_FLOW);
        if (wfh_fabs(r) <= 1.0e-12 || count >= 3000) break;
    }
    return q;
}

WFH_HD double kw_ssf_newton_raphson(double q, double C, double cel, double dt, double dx) {
    const double dt_dx = dt / dx;
    const double cinv = 1.0 / cel;
    return kw_ssf_newton_raphson_c(q, C, dt_dx, cinv, dt_dx + cinv);
}

// The part of one subsurface kinematic-wave solve that does not depend on the upstream inflow:
// the celerity at the previous water table (ssf_celerity, KhExponential,
// subsurface_process.jl:5-20) and the Newton-Raphson constants derived from it. A routing loop
// evaluates this before the inflow arrives, off the cell-to-cell dependency chain.
struct SsfPrep {
    double dt_dx;  // dt / dx
    double cinv;   // 1 / celerity
    double df;     // dt / dx + 1 / celerity
    double qp_cel; // q_prev / celerity
};

WFH_HD SsfPrep ssf_prepare(double q_prev, double zi_prev, double slope, double sy, double dt, double dx, double kh0,
                           double f, double z_exp) {
    // ssf_celerity (KhExponentialConstant: the depth is capped at z_exp; INFINITY = KhExponential)
    const double cel = (kh0 * wfh_exp(-f * (zi_prev < z_exp ? zi_prev : z_exp)) * slope) / sy;
    SsfPrep p;
    p.dt_dx = dt / dx;
    p.cinv = 1.0 / cel;
    p.df = p.dt_dx + p.cinv;
    p.qp_cel = q_prev / cel;
    return p;
}

// Unsaturated layers of one column as the subsurface solver sees them.
template <int N>
struct Column {
    double ustore[N];
    double thick[N];   // unsaturated layer thickness (NaN = inactive)
    double act[N];     // actual layer thickness
    double cum[N + 1]; // cumulative layer depth
    double zi;
    int n_unsat;
    double theta_e, theta_fc_r;  // theta_s - theta_r, theta_fc - theta_r
};

// water_table_change  Wflow/src/utils.jl:1090-1131
template <int N>
WFH_HD void water_table_change(const Column<N>& s, double net_flux, double sy, double dt, double& dh, double& exf) {
    if (net_flux <= 0.0) {
        dh = net_flux * dt / sy;
    } else {
        dh = 0.0;
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
            if (k >= s.n_unsat) continue;
            const double cap = wfh_max(s.thick[k] * s.theta_e - s.ustore[k], 0.0) / dt;
            const double fl = wfh_min(net_flux, cap);
            if (cap <= net_flux) {
                dh += s.thick[k];
            } else {
                const double syd = s.theta_e - (s.ustore[k] / s.thick[k]);
                dh += fl * dt / syd;
            }
            net_flux -= fl;
            if (net_flux == 0.0) break;
        }
    }
    exf = wfh_max(net_flux, 0.0);
}

// update_ustorelayerdepth!(soil, zi_prev, water_table_depth, i)  Wflow/src/soil/soil.jl:1213-1259
template <int N>
WFH_HD void update_ustorelayerdepth(Column<N>& s, double zi_prev, double zi) {
    const int nprev = s.n_unsat;
    double t[N];
    const int nnew = set_layerthickness<N>(zi, s.cum, s.act, t);
    if (zi < zi_prev) {
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            if (k < nnew || k > nprev) continue;
            if (wfh_isnan(t[k - 1]))
                s.ustore[k - 1] = 0.0;
            else
                s.ustore[k - 1] = (t[k - 1] / s.thick[k - 1]) * s.ustore[k - 1];
        }
    } else {
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            if (k < nprev || k > nnew) continue;
            const double th_prev = wfh_isnan(s.thick[k - 1]) ? 0.0 : s.thick[k - 1];
            const double d = t[k - 1] - th_prev;
            s.ustore[k - 1] = s.ustore[k - 1] + d * s.theta_fc_r;
        }
    }
    s.n_unsat = nnew;
#pragma unroll
    for (int k = 0; k < N; ++k) s.thick[k] = t[k];
    s.zi = zi;
}

// kinematic_wave_ssf for the exponential Kh profile
// Wflow/src/routing/subsurface/subsurface_process.jl:89-172 (ssf_celerity :6-21)
// zi_prev is the subsurface-flow water table (it can differ from the soil column's after an
// early return, which sets the former to d and leaves the column alone); `pp` is
// ssf_prepare(q_prev, zi_prev, slope, sy, dt, dx, kh0, f).
//
// With defer_update set, the closing update_ustorelayerdepth!(soil, zi_prev, zi_out) of the
// single-step branch is left to the caller (it touches only the column, which nothing in the
// remainder of the solve reads) and the function returns true; the caller must apply it before
// the column is read again. The sub-stepped branch always updates the column itself.
template <int N>
WFH_HD bool kinematic_wave_ssf(double q_in, double q_prev, double zi_prev, double q_net, double slope, double sy,
                               double d, double dt, double dx, double dw, double q_max, double kh0, double f,
                               double z_exp, const SsfPrep& pp, Column<N>& soil, const bool defer_update, double& q_out,
                               double& zi_out, double& exf_out, double& net_out) {
    if (q_in + q_prev == 0.0 && q_net <= 0.0) {
        q_out = 0.0;
        zi_out = d;
        exf_out = 0.0;
        net_out = 0.0;
        return false;
    }
    double q = (q_prev + q_in) / 2.0;
    double C = pp.dt_dx * (q_in + q_net) + pp.qp_cel;
    q = kw_ssf_newton_raphson_c(q, C, pp.dt_dx, pp.cinv, pp.df);
    q = wfh_min(q, (q_max * dw));
    double net = (q_in + q_net - q) / (dw * dx);
    double dh, exf;
    water_table_change<N>(soil, net, sy, dt, dh, exf);
    double zi = zi_prev - dh;
    if (zi > d) {
        const double q_excess = (dw * dx) * sy * (zi - d) / dt;
        q = wfh_max(q - q_excess, WFH_KIN_WAVE_MIN_FLOW);
    }
    zi = wfh_clamp(zi, 0.0, d);
    // its = Int(ceil(round(x; sigdigits = 12))), x = |zi - zi_prev| / 0.1. Rounding to 12
    // significant digits moves x by less than 1e-11 of its value, so every |zi - zi_prev| <=
    // 0.0998 (x <= 0.998 up to one rounding) gives its <= 1, the single-step branch; only
    // larger changes need x and the exact count.
    const double adz = wfh_fabs(zi - zi_prev);
    double its_f = 1.0;
    long its = 1;
    if (adz > 0.0998) {
        its_f = WFH_CEIL(wfh_round_sig12(adz / 0.1));
        its = (long)its_f;
    }
    bool deferred = false;
    if (its > 1) {
        const double dt_s = dt / its_f;
        double q_sum = 0.0, exf_sum = 0.0, net_sum = 0.0;
        for (long i = 0; i < its; ++i) {
            const double cel = (kh0 * wfh_exp(-f * (zi_prev < z_exp ? zi_prev : z_exp)) * slope) / sy;
            C = (dt_s / dx) * q_in + q_prev / cel + q_net * (dt_s / dx);
            q = kw_ssf_newton_raphson(q_prev, C, cel, dt_s, dx);
            q = wfh_min(q, (q_max * dw));
            net = (q_in + q_net - q) / (dw * dx);
            water_table_change<N>(soil, net, sy, dt_s, dh, exf);
            zi = zi_prev - dh;
            if (zi > d) {
                const double q_excess = (dw * dx) * sy * (zi - d) / dt_s;
                q = wfh_max(q - q_excess, WFH_KIN_WAVE_MIN_FLOW);
            }
            zi = wfh_clamp(zi, 0.0, d);
            update_ustorelayerdepth<N>(soil, zi_prev, zi);
            exf_sum += exf;
            net_sum += net;
            q_sum += q;
            q_prev = q;
            zi_prev = zi;
        }
        q = q_sum / its_f;
        exf = exf_sum / its_f;
        net = net_sum / its_f;
    } else if (defer_update) {
        deferred = true;
    } else {
        update_ustorelayerdepth<N>(soil, zi_prev, zi);
    }
    q_out = q;
    zi_out = zi;
    exf_out = exf;
    net_out = net;
    return deferred;
}

}  // namespace wfh
