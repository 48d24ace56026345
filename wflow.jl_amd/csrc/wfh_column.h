// The vertical SBM column of one cell for one step (update_land_hydrology_model!,
// Wflow/src/sbm.jl:82-132): interception, snow, open-water runoff and the soil column, as a pure
// function of the cell's forcing, states and parameters. The single-model kernel and the
// ensemble kernel load their own layouts into these structs and store the results back; the
// arithmetic is written once, in the reference's operation order (no FMA contraction).
#pragma once

#include "wfh_physics.h"

namespace wfh {

// Per-cell parameters of the column (parameters of switched-off processes are zeros).
struct ColParams {
    double gap, smax, kc, er;                       // interception
    double tti, tt, ttm, cfmax, whc, wsoil, cfsoil;  // snow, soil temperature
    double rf, wf, sfrac;                           // river / water / soil fractions
    double ts, tr, tfc, dsoil, kv0, fk, zexp;       // soil hydraulics
    double compf, icap_s, icap_c;                   // infiltration
    double rd, h3h, h3l, h1, h2, h4, ah1, hb, rdp;  // roots, Feddes, Brooks-Corey
    double caph, capn, maxleak;                     // capillary rise, leakage
};

template <int N>
struct ColLayers {
    double kvf[N], cbc[N], rfr[N];  // kvfrac, Brooks-Corey c, root fraction
};

// Model-wide switches and the configured layer geometry.
struct ColConfig {
    double dt;
    bool gash, snow, infred, balance;
    const double* cfg_cum;    // [N + 1]
    const double* cfg_thick;  // [N]
};

// Prognostic states of the column (in: start of the step, out: end of the vertical update).
template <int N>
struct ColState {
    double canopy, ss, sw, tsoil;
    double swd, tsws;  // satwaterdepth (read only here), total soil water storage at the last step
    double us[N];
};

// Fluxes and diagnostics of the vertical update.
struct ColFlux {
    double zi, interception, transfer, recharge, act_inf, inf_exc, sat_exc, transp, sevap, capflux, leak, aet;
    double nrr, rl, aol, swe, store_prev, P;
};

// P, E, T: forcing of the step (P and E already zero over a reservoir's area); wd_land: the
// overland water depth of the last step; wd_river: the river water depth (0 without a river).
template <int N>
WFH_HD void sbm_column(const double P, const double E, const double T, const double wd_land, const double wd_river,
                       const ColParams& p, const ColLayers<N>& L, const ColConfig& cfg, ColState<N>& s, ColFlux& o) {
    const double dt = cfg.dt;
    double us[N];
#pragma unroll
    for (int k = 0; k < N; ++k) us[k] = s.us[k];

    // storage_prev! (Wflow/src/mass_balance.jl:96-115,170-183) from last step's states
    double canopy = s.canopy;
    double ss = cfg.snow ? s.ss : 0.0, sw = cfg.snow ? s.sw : 0.0;
    double store_prev = 0.0;
    if (cfg.balance) {
        const double snow = cfg.snow ? ss + sw : 0.0 + 0.0;
        store_prev = s.tsws + canopy + snow + 0.0 * 0.0 + 0.0;
    }

    // ---- interception: canopy.jl:54-98 (Gash, dt >= 23 h) / :115-143 (modified Rutter) ----
    const double cpe = p.kc * E * (1.0 - p.gap);
    double tf, ir, sf;
    if (cfg.gash) {
        interception_gash(p.smax, p.er, p.gap, P, cpe, dt, tf, ir, sf);
    } else {
        interception_modrut(P, cpe, p.gap, p.smax, dt, canopy, tf, ir, sf);
    }

    // ---- snow: snow.jl:123-177 ----
    double wfs, swe = 0.0;
    if (cfg.snow) {
        double sro;
        snow_hbv(tf + sf, T, p.tti, p.tt, p.ttm, p.cfmax, p.whc, dt, ss, sw, swe, sro);
        wfs = sro + 0.0 * 0.0;  // get_water_flux_surface! with snow (runoff.jl:49-58), no glacier
    } else {
        wfs = tf + sf;  // runoff.jl:37-46
    }

    // ---- open water runoff: runoff.jl:61-111 ----
    const double rr = wfh_min(1.0, p.rf) * wfs;
    const double rl = wfh_min(1.0, p.wf) * wfs;
    const double aor = p.rf * wfh_min(wd_river / dt, E);
    const double aol = p.wf * wfh_min(wd_land / dt, E);
    const double nrr = rr - aor;

    // ---- soil boundary conditions: soil.jl:643-682 ----
    const double pot_trans = wfh_max(0.0, cpe - ir);
    double pot_sevap = p.sfrac * E;
    pot_sevap = pot_sevap - 0.0;
    const double wfs_soil = wfh_max(((wfs + 0.0) - rr) - rl + 0.0 / dt, 0.0);

    // ---- update_soil_water_flow!  soil.jl:1150-1211 ----
    const double ts = p.ts, tr = p.tr, dsoil = p.dsoil;
    double act[N], cum[N + 1];
    const int nl = soil_layers<N>(dsoil, cfg.cfg_cum, cfg.cfg_thick, act, cum);
    const double swc = dsoil * (ts - tr);
    // update_diagnostic_vars!  soil.jl:1400-1436
    double usd = 0.0;
    {
        bool first = true;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k < nl) {
                usd = first ? us[k] : usd + us[k];
                first = false;
            }
    }
    const double zi = wfh_max(0.0, dsoil - s.swd / (ts - tr));
    const double thd = drainable_porosity(ts, p.tfc);
    double drainable = (dsoil - zi) * thd;
    double ucap = swc - s.swd - usd;
    double th[N];
    const int nu = set_layerthickness<N>(zi, cum, act, th);
    // soil_temperature!, infiltration_reduction_factor!  soil.jl:685-730
    double tsoil = s.tsoil;
    if (cfg.snow) tsoil = tsoil + p.wsoil * (T - tsoil);
    double fred = 1.0;
    if (cfg.snow && cfg.infred) {
        const double cf = p.cfsoil;
        const double bb = 1.0 / (1.0 - cf);
        fred = scurve(tsoil, 0.0 + 273.15, bb, 8.0) + cf;
    }
    // infiltration  soil_process.jl:16-41
    double infilt, inf_exc;
    {
        const double soilinf = wfs_soil * (1.0 - p.compf);
        const double pathinf = wfs_soil * p.compf;
        const double ms = wfh_min(p.icap_s * fred, soilinf);
        const double mp = wfh_min(p.icap_c * fred, pathinf);
        infilt = wfh_min(mp + ms, wfh_max(0.0, ucap / dt));
        inf_exc = (soilinf - ms) + (pathinf - mp);
    }
    // unsaturated_zone_flow!  soil.jl:764-804
    double transfer = 0.0;
    if (nu > 0) {
        double z = 0.0, flow = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) {
            if (m >= nu) continue;
            z = (m == 0) ? th[0] : z + th[m];
            const double l_sat = th[m] * (ts - tr);
            const double kvz = L.kvf[m] * p.kv0 * wfh_exp(-p.fk * (z < p.zexp ? z : p.zexp));
            double u = (m == 0) ? us[m] + infilt * dt : us[m] + flow * dt;
            unsatzone_flow_layer(u, kvz, l_sat, L.cbc[m], dt, flow);
            us[m] = u;
        }
        transfer = flow;
    }
    // soil_evaporation!  soil.jl:814-856
    double sevap, ses;
    {
        double pot = pot_sevap;
        double seu;
        if (nu == 0)
            seu = 0.0;
        else if (nu == 1)
            seu = pot * wfh_min(1.0, us[0] / (zi * (ts - tr)));
        else
            seu = pot * wfh_min(1.0, us[0] / (th[0] * (ts - tr)));
        seu = wfh_min(seu, us[0] / dt);
        pot -= seu;
        us[0] = us[0] - seu * dt;
        if (nu == 0 || nu == 1) {
            const double e = pot * wfh_min(1.0, (act[0] - zi) / act[0]);
            ses = wfh_min(e, (act[0] - zi) * thd / dt);
        } else {
            ses = 0.0;
        }
        sevap = seu + ses;
        drainable -= ses * dt;
    }
    // transpiration!  soil.jl:865-975
    double ae_u = 0.0, ae_s, transp;
    {
        const double rd = p.rd;
        const double h3 = feddes_h3(p.h3h, p.h3l, pot_trans);
        double sum_rf = 0.0, rf_low = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k >= nu) continue;
            double rfu;
            if (k == nu - 1 && zi < rd) {
                const double rootlength = wfh_min(act[k], rd - cum[k]);
                rfu = L.rfr[k] * (th[k] / rootlength);
            } else {
                rfu = L.rfr[k];
            }
            sum_rf += rfu;
            rf_low = rfu;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k >= nu) continue;
            const double rfu = (k < nu - 1) ? L.rfr[k] : rf_low;
            const double rfs = rd > 0.0 ? wfh_max((1.0 / sum_rf), 1.0) * rfu : 0.0;
            const double vwc = wfh_max(us[k] / th[k], 1e-7);
            const double head = head_brooks_corey(vwc, ts, tr, L.cbc[k], p.hb);
            const double alpha = rwu_reduction_feddes(head, p.h1, p.h2, h3, p.h4, p.ah1);
            const double availcap = wfh_min(1.0, wfh_max(0.0, (rd - cum[k]) / th[k]));
            const double maxextr = us[k] * availcap / dt;
            const double layer = wfh_min(alpha * rfs * pot_trans, maxextr);
            us[k] = us[k] - layer * dt;
            ae_u += layer;
        }
        const double wetroots = scurve(zi, rd, 1.0, p.rdp);
        const double alpha0 = rwu_reduction_feddes(0.0, p.h1, p.h2, h3, p.h4, p.ah1);
        const double rest = pot_trans - ae_u;
        ae_s = wfh_min(rest * wetroots * alpha0, drainable / dt);
        drainable -= ae_s * dt;
        transp = ae_u + ae_s;
    }
    // actual_infiltration!  soil.jl:987-1017
    double act_inf;
    {
        double excess = 0.0;
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
            if (k >= nu) continue;
            excess = wfh_max(0.0, us[k] - th[k] * (ts - tr));
            us[k] = us[k] - excess;
            if (k > 0) us[k - 1] = us[k - 1] + excess;
        }
        act_inf = infilt - excess / dt;
    }
    const double sat_exc = (wfs_soil - act_inf) - inf_exc;
    // unsaturated_store_depth! + capacity  soil.jl:1194-1196
    {
        bool first = true;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k < nl) {
                usd = first ? us[k] : usd + us[k];
                first = false;
            }
        if (nl <= 0) usd = 0.0;
    }
    ucap = swc - s.swd - usd;
    // capillary_flux!  soil.jl:1050-1111
    double capflux = 0.0;
    if (nu > 0) {
        double kvf_nu = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k == nu - 1) kvf_nu = L.kvf[k];
        const double ksat = kvf_nu * p.kv0 * wfh_exp(-p.fk * (zi < p.zexp ? zi : p.zexp));
        const double mx = wfh_max(0.0, wfh_min(wfh_min(wfh_min(ksat, ae_u), ucap / dt), drainable / dt));
        double cap;
        if (zi > p.rd)
            cap = mx * wfh_pow(1.0 - wfh_min(zi, p.caph) / (p.caph), p.capn);
        else
            cap = 0.0;
        double net = cap;
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
            if (k >= nu) continue;
            const double toadd = wfh_min(net, wfh_max((th[k] * (ts - tr) - us[k]) / dt, 0.0));
            us[k] = us[k] + toadd * dt;
            net -= toadd;
            capflux += toadd;
        }
    }
    // leakage!  soil.jl:1118-1136
    double leak;
    {
        double kvf_nl = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k == nl - 1) kvf_nl = L.kvf[k];
        const double deepksat = kvf_nl * p.kv0 * wfh_exp(-p.fk * (dsoil < p.zexp ? dsoil : p.zexp));
        const double deeptransfer = wfh_min(drainable / dt, deepksat);
        leak = wfh_max(0.0, wfh_min(p.maxleak, deeptransfer));
    }
    const double recharge = (transfer - capflux - leak - ae_s - ses);
    double aet = sevap + transp + aor + aol + 0.0;
    aet = aet + ir;  // sbm.jl:130

    s.canopy = canopy;
    s.ss = ss;
    s.sw = sw;
    s.tsoil = tsoil;
#pragma unroll
    for (int k = 0; k < N; ++k) s.us[k] = us[k];
    o.zi = zi;
    o.interception = ir;
    o.transfer = transfer;
    o.recharge = recharge;
    o.act_inf = act_inf;
    o.inf_exc = inf_exc;
    o.sat_exc = sat_exc;
    o.transp = transp;
    o.sevap = sevap;
    o.capflux = capflux;
    o.leak = leak;
    o.aet = aet;
    o.nrr = nrr;
    o.rl = rl;
    o.aol = aol;
    o.swe = swe;
    o.store_prev = store_prev;
    o.P = P;
}

}  // namespace wfh
