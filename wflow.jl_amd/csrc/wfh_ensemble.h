// wfh_ensemble.h - a parameter ensemble of one catchment stepped as one model (included at the end
// of wflow_hip.hip, whose base model supplies topology, node order and the shared parameters).
//
// Layout: member-innermost. Every per-member array holds cell v, member m at v * M + m (M a
// multiple of 64), so a wave is 64 members of one cell: its loads and stores are contiguous, the
// cell's shared parameters and topology are wave-uniform (scalar loads), and the members never
// wait on one another.
//
// Schedule: the step is a sequence of stages, one launch each, recorded once into a HIP graph and
// replayed every step. A stage is a set of strands, a strand a list of cells that one wave group
// updates in order, each cell completely (vertical column, subsurface, soil storage, overland,
// every river sub-step). A cell's upstream cells lie in earlier stages or earlier in its own
// strand, so a stage needs no synchronisation beyond the launch boundary:
//   - a wide topological level (at least `wide` cells) is one stage whose strands are runs of
//     `strand_cells` cells of that level (independent cells);
//   - consecutive narrow levels form one stage as long as every cell has at most one upstream cell
//     inside the stage; its strands are then the chains (paths) of those levels.
// Every schedule computes the same operations on the same values, so results are bitwise
// independent of it; the autotuner (wfh_ens_set_autotune) times candidate schedules on real steps.

namespace {

constexpr int ENS_MAX_SR = 8;       // river sub-steps per step
constexpr int ENS_MAX_MEMBERS = 4096;
constexpr int ENS_PRE = 10;          // fluxes of the vertical update that the routing part reads
// What a stage launch does for each of its cells: everything, the vertical column only (into
// the pre-buffer), or the routing only (the column taken from the pre-buffer).
enum { ENS_FUSED = 0, ENS_COLUMN = 1, ENS_ROUTE = 2 };

struct EnsDev {
    int n, nriv, M, G;  // cells, river cells, members, member groups of 64
    double dt;
    int gash, snow, infred, balance;
    double cfg_thick[WFH_MAX_LAYERS], cfg_cum[WFH_MAX_LAYERS + 1];
    int S_r;
    double dts_r[2];
    // shared maps (node order)
    Prm gap, smax, kc, er, cfmax, tt, tti, ttm, whc, rf, wf, sfrac;
    Prm ts, tr, tfc, kv0, f, hb, h1, h2, h3h, h3l, h4, ah1, d, icap_c, maxleak, compf, rdp, caph, capn, wsoil, cfsoil,
        rd;
    PrmL c, kvfrac, rootfrac;
    Prm area, fl, fw, slope, sy, khfrac, sfw;
    const int *up_ptr, *up_idx, *river_of;
    const double* up_f;
    const int *rup_ptr, *rup_idx;
    const double *r_fw, *r_fl, *r_area;
    const double *P, *E, *T;  // forcing of the step, node order, shared by the members
    const double *fa, *fb;    // [M] multipliers of kv_0 and f
    // per member [n * M]
    double *canopy, *snow_s, *snow_w, *tsoil, *swd, *tsws, *ustore /* [N][n * M] */, *ssf_q, *ol_q, *ol_h;
    const double *qmax, *alpha;
    // per member [nriv * M]
    double *r_q, *r_h, *r_storage, *r_q_avg;
    const double* r_alpha;
    double* rq_sub;                // [S_r - 1][nriv * M]: discharge of the earlier river sub-steps
    unsigned long long* bal_max;   // [M]: largest |land balance error| of the step (bits of a double)
    const int *strand_ptr, *strand_cell;
    // the vertical update of chain-stage cells, done beforehand in a parallel pass: node -> slot
    // (or -1), and the fluxes the routing part reads, [ENS_PRE][npre * M]
    const int* pre_slot;
    double* pre;
    size_t pre_stride;
    // learned warm start of the kinematic-wave solves (LEARN instantiation): the network, and the
    // gate: a wave consults it only where some member's previous discharge is below mlp_gate
    const MlpOps* mlp;
    double mlp_gate;
    // optional instrumentation: Newton updates per solve ([2 families][16 bins] + learned starts
    // used [2]), and harvesting of (features, target, updates) rows of sampled solves
    unsigned long long* kw_hist;
    float* harvest;
    unsigned long long* harvest_n;
    long long harvest_cap;
    int harvest_every;
};
static_assert(sizeof(EnsDev) + 64 <= 4096, "EnsDev too large for a kernel argument");

__device__ __forceinline__ double ens_ld(const Prm& a, int v) { return a.p[v & a.mask]; }

constexpr int ENS_HIST_BINS = 16;
__device__ __forceinline__ int kw_bin(int updates) {
    if (updates <= 8) return updates;
    if (updates >= 3000) return 15;
    if (updates <= 16) return 9;
    if (updates <= 32) return 10;
    if (updates <= 64) return 11;
    if (updates <= 128) return 12;
    if (updates <= 256) return 13;
    return 14;
}

// Features of a kinematic-wave solve for the warm-start network: q_in, previous q, lateral
// inflow (as discharge), alpha, beta, dx, dt, slope (logarithms where they span decades).
__device__ __forceinline__ void kw_features(const double q_in, const KwPrep& p, const double alpha, const double dt,
                                            const double dx, const double slope, float* f) {
    const float ql = (float)(p.q_lat * dx);
    f[0] = log10f((float)q_in + 1e-9f);
    f[1] = log10f((float)p.q_prev + 1e-9f);
    f[2] = copysignf(log10f(1.0f + fabsf(ql) * 1e6f), ql);
    f[3] = log10f((float)alpha);
    f[4] = 0.6f;
    f[5] = log10f((float)dx);
    f[6] = log10f((float)dt);
    f[7] = log10f((float)slope);
}

// The reference's start of the solve (u_prev, or cbrt(C / alpha) without previous flow).
__device__ __forceinline__ double kw_ref_start(const double q_in, const KwPrep& p, const double alpha) {
    const double C = p.dt_dx * q_in + p.a_u3 + p.lat;
    return p.u_prev > 0.0 ? p.u_prev : wfh_cbrt(C / alpha);
}

// Kernel variants: plain solves; plain solves with the instrumentation (Newton-update histogram,
// harvest of training rows); the learned warm start (with the histogram).
enum { ENS_PLAIN = 0, ENS_INSTR = 1, ENS_LEARN = 2 };

// One kinematic-wave solve of the ensemble (family fam: 0 overland, 1 river). ENS_PLAIN:
// kw_solve. ENS_LEARN: every lane of the wave calls this (the network runs across the wave, 64
// members at once) when some lane's previous discharge is below the gate; the network's start is
// taken where its residual beats the reference start's (kw_solve_from).
// The learned solve as a call of its own (taken only by waves the gate lets through): features,
// the network, its proposed start, then the reference's Newton iteration (kw_solve_from).
struct KwLearned {
    double q, area;
    int updates, used;
};
__device__ __noinline__ KwLearned kw_learned_solve(const MlpOps* W, const double q_in, const double dt_dx,
                                                   const double u_prev, const double a_u3, const double lat,
                                                   const double q_prev, const double q_lat, const double alpha,
                                                   const double dt, const double dx, const double slope) {
    KwPrep p;
    p.dt_dx = dt_dx, p.u_prev = u_prev, p.a_u3 = a_u3, p.lat = lat, p.q_prev = q_prev, p.q_lat = q_lat;
    float f[8];
    kw_features(q_in, p, alpha, dt, dx, slope, f);
    const float y = mlp_forward(*W, f, nullptr);
    const float target = W->y_mean + W->y_std * y;
    const double u0 = kw_ref_start(q_in, p, alpha) * (double)exp10f(0.2f * target);
    KwLearned r;
    bool used;
    r.updates = kw_solve_from(q_in, p, alpha, u0, r.q, r.area, used);
    r.used = used;
    return r;
}

template <int VAR>
__device__ __forceinline__ void ens_kw(const EnsDev& e, const MlpOps* W, const int fam, const unsigned key,
                                       const double q_in, const KwPrep& p, const double alpha, const double dt,
                                       const double dx, const double slope, double& q, double& area,
                                       unsigned* hist) {
    int updates = 0;
    bool used = false;
    if (VAR == ENS_LEARN && __any(p.q_prev < e.mlp_gate)) {
        const KwLearned r =
            kw_learned_solve(W, q_in, p.dt_dx, p.u_prev, p.a_u3, p.lat, p.q_prev, p.q_lat, alpha, dt, dx, slope);
        q = r.q;
        area = r.area;
        updates = r.updates;
        used = r.used != 0;
    } else if (VAR != ENS_PLAIN && (hist || e.harvest)) {
        updates = kw_solve_from(q_in, p, alpha, NaN_, q, area, used);
    } else {
        kw_solve(q_in, p, alpha, q, area);
        return;
    }
    if (hist) {
        atomicAdd(&hist[fam * ENS_HIST_BINS + kw_bin(updates)], 1u);
        if (used) atomicAdd(&hist[2 * ENS_HIST_BINS + fam], 1u);
    }
    if (VAR == ENS_INSTR && e.harvest && q > 0.0 && (key % (unsigned)e.harvest_every) == 0u) {
        const unsigned long long row = atomicAdd(e.harvest_n, 1ull);
        if ((long long)row < e.harvest_cap) {
            float f[8];
            kw_features(q_in, p, alpha, dt, dx, slope, f);
            float* o = e.harvest + row * 10;
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = f[j];
            o[8] = log10f((float)q) - 5.0f * log10f((float)kw_ref_start(q_in, p, alpha));
            o[9] = (float)updates;
        }
    }
}

// The update of cell v (node order) for member m: the same operations, in the same order, as
// k_vertical followed by k_route's families for one cell (one subsurface and one overland
// sub-step, S_r river sub-steps). MODE ENS_COLUMN stops after the vertical column (states
// written, the routing inputs into the pre-buffer), ENS_ROUTE starts from them. Returns |land
// balance error| (0 without the balance or in ENS_COLUMN).
template <int N, int MODE, int VAR>
__device__ __forceinline__ double ens_cell(const EnsDev& e, const int v, const int m, const MlpOps* W,
                                           unsigned* hist) {
    const size_t M = (size_t)e.M;
    const size_t i = (size_t)v * M + m;
    const size_t nM = (size_t)e.n * M;
    const double dt = e.dt;
    const double fa = e.fa[m], fb = e.fb[m];
    // ---- loads: member states (vector), shared parameters (wave-uniform) ----
    const double P = e.P[v];
    ColState<N> s;
    s.canopy = e.canopy[i];
    s.ss = e.snow_s[i];
    s.sw = e.snow_w[i];
#pragma unroll
    for (int k = 0; k < N; ++k) s.us[k] = e.ustore[k * nM + i];
    const double oh0 = e.ol_h[i];
    const int r = e.river_of[v];
    const bool riv = r >= 0;
    const size_t ri = (size_t)(riv ? r : 0) * M + m;
    const double rh0 = e.r_h[ri];
    ColParams p;
    p.ts = ens_ld(e.ts, v), p.tr = ens_ld(e.tr, v), p.tfc = ens_ld(e.tfc, v), p.dsoil = ens_ld(e.d, v);
    p.kv0 = ens_ld(e.kv0, v) * fa, p.fk = ens_ld(e.f, v) * fb, p.zexp = INFINITY;
    ColFlux cf;
    if (MODE != ENS_ROUTE) {
        const double E = e.E[v], T = e.T[v];
        s.tsws = e.tsws[i];
        s.swd = e.swd[i];
        s.tsoil = e.tsoil[i];
        p.gap = ens_ld(e.gap, v), p.smax = ens_ld(e.smax, v), p.kc = ens_ld(e.kc, v), p.er = ens_ld(e.er, v);
        p.tti = ens_ld(e.tti, v), p.tt = ens_ld(e.tt, v), p.ttm = ens_ld(e.ttm, v), p.cfmax = ens_ld(e.cfmax, v);
        p.whc = ens_ld(e.whc, v), p.wsoil = ens_ld(e.wsoil, v), p.cfsoil = ens_ld(e.cfsoil, v);
        p.rf = ens_ld(e.rf, v), p.wf = ens_ld(e.wf, v), p.sfrac = ens_ld(e.sfrac, v);
        p.compf = ens_ld(e.compf, v), p.icap_c = ens_ld(e.icap_c, v);
        p.rd = ens_ld(e.rd, v), p.h3h = ens_ld(e.h3h, v), p.h3l = ens_ld(e.h3l, v);
        p.h1 = ens_ld(e.h1, v), p.h2 = ens_ld(e.h2, v), p.h4 = ens_ld(e.h4, v), p.ah1 = ens_ld(e.ah1, v);
        p.hb = ens_ld(e.hb, v), p.rdp = ens_ld(e.rdp, v), p.caph = ens_ld(e.caph, v), p.capn = ens_ld(e.capn, v);
        p.maxleak = ens_ld(e.maxleak, v);
        ColLayers<N> L;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            L.kvf[k] = ld(e.kvfrac, k, v);
            L.cbc[k] = ld(e.c, k, v);
            L.rfr[k] = ld(e.rootfrac, k, v);
        }
        p.icap_s = p.kv0 * L.kvf[0];  // infiltcap_soil = kv_0 kvfrac[1] (soil.jl)
        // ---- vertical column ----
        const ColConfig cfg{dt, e.gash != 0, e.snow != 0, e.infred != 0, e.balance != 0, e.cfg_cum, e.cfg_thick};
        sbm_column<N>(P, E, T, oh0, riv ? rh0 : 0.0, p, L, cfg, s, cf);
    }
    // the column's states are written now, so that they do not stay live through the routing
    // (the balance below needs only their sum)
    if (MODE != ENS_ROUTE) {
        e.canopy[i] = s.canopy;
        if (e.snow) {
            e.snow_s[i] = s.ss;
            e.snow_w[i] = s.sw;
        }
        e.tsoil[i] = s.tsoil;
    }
    const double snow_bal = e.snow ? s.ss + s.sw : 0.0 + 0.0;
    const double canopy_bal = s.canopy;
    if (MODE == ENS_COLUMN) {
#pragma unroll
        for (int k = 0; k < N; ++k) e.ustore[k * nM + i] = s.us[k];
        double* q = e.pre + (size_t)e.pre_slot[v] * M + m;
        const size_t ps = e.pre_stride;
        q[0] = cf.zi, q[ps] = cf.recharge, q[2 * ps] = cf.sat_exc, q[3 * ps] = cf.rl, q[4 * ps] = cf.inf_exc;
        q[5 * ps] = cf.aol, q[6 * ps] = cf.nrr, q[7 * ps] = cf.aet, q[8 * ps] = cf.leak, q[9 * ps] = cf.store_prev;
        return 0.0;
    }
    if (MODE == ENS_ROUTE) {
        const double* q = e.pre + (size_t)e.pre_slot[v] * M + m;
        const size_t ps = e.pre_stride;
        cf.zi = q[0], cf.recharge = q[ps], cf.sat_exc = q[2 * ps], cf.rl = q[3 * ps], cf.inf_exc = q[4 * ps];
        cf.aol = q[5 * ps], cf.nrr = q[6 * ps], cf.aet = q[7 * ps], cf.leak = q[8 * ps], cf.store_prev = q[9 * ps];
        cf.P = P;
    }
    // routing inputs, loaded after the column (fewer registers live across it)
    const double sq0 = e.ssf_q[i], oq0 = e.ol_q[i];
    const double qmax = e.qmax[i], alpha = e.alpha[i];
    const double rq0 = e.r_q[ri], r_store_prev = e.r_storage[ri], ralpha = e.r_alpha[ri];
    const double area = ens_ld(e.area, v), flen = ens_ld(e.fl, v), fw = ens_ld(e.fw, v), slope = ens_ld(e.slope, v);
    const double sy = ens_ld(e.sy, v), sfw = ens_ld(e.sfw, v);
    // the column's parameters are read again (the same values) rather than kept live through it
    const double kv0 = ens_ld(e.kv0, v) * e.fa[m], fk = ens_ld(e.f, v) * e.fb[m];
    const double kh0 = ens_ld(e.khfrac, v) * kv0;

    // ---- lateral subsurface flow (one sub-step of length dt) ----
    const double ts = ens_ld(e.ts, v), tr = ens_ld(e.tr, v), dsoil = ens_ld(e.d, v), tfc = ens_ld(e.tfc, v);
    Column<N> col;
    soil_layers<N>(dsoil, e.cfg_cum, e.cfg_thick, col.act, col.cum);
#pragma unroll
    for (int k = 0; k < N; ++k) col.ustore[k] = s.us[k];
    col.zi = cf.zi;
    col.n_unsat = set_layerthickness<N>(col.zi, col.cum, col.act, col.thick);
    col.theta_e = ts - tr;
    col.theta_fc_r = tfc - tr;
    const int e0 = e.up_ptr[v], e1 = e.up_ptr[v + 1];
    // Upstream inflow of every family. The upstream cells are complete (earlier stages or earlier
    // in this strand), so on a chain (MODE ROUTE) all of it is read here, in one round trip,
    // instead of one round trip before each solve; each sum is the same left fold in ascending
    // upstream id either way.
    double up_s_qin = 0.0, up_s_tor = 0.0, up_o_tor = 0.0, up_o_qin = 0.0;
    constexpr int PRE_SR = 4;  // river sub-steps whose inflow is read ahead (registers)
    double up_r[PRE_SR];
    const int r_e0 = riv ? e.rup_ptr[r] : 0, r_e1 = riv ? e.rup_ptr[r + 1] : 0;
    const size_t rM = (size_t)e.nriv * M;
    if (MODE == ENS_ROUTE) {
        for (int j = e0; j < e1; ++j) {
            const int u = e.up_idx[j];
            const double fu = e.up_f[j];
            const double qs = e.ssf_q[(size_t)u * M + m], qo = e.ol_q[(size_t)u * M + m];
            up_s_qin = up_s_qin + qs * (1.0 - fu);  // lateral_subsurface_flow.jl:228-240
            up_s_tor = up_s_tor + qs * fu;
            up_o_tor = up_o_tor + qo * fu;          // surface_kinwave.jl:317-326
            up_o_qin = up_o_qin + qo * (1.0 - fu);
        }
#pragma unroll
        for (int k = 0; k < PRE_SR; ++k) up_r[k] = 0.0;
        for (int j = r_e0; j < r_e1; ++j) {
            const size_t ui = (size_t)e.rup_idx[j] * M + m;
#pragma unroll
            for (int k = 0; k < PRE_SR; ++k)
                if (k < e.S_r) up_r[k] = up_r[k] + (k == e.S_r - 1 ? e.r_q : e.rq_sub + (size_t)k * rM)[ui];
        }
    }
    double sq, szi;
    SsfCum cum{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    {
        double qin = 0.0, tor = 0.0;  // fold in ascending upstream id  lateral_subsurface_flow.jl:228-240
        if (MODE == ENS_ROUTE) {
            qin = up_s_qin;
            tor = up_s_tor;
        } else {
            for (int j = e0; j < e1; ++j) {
                const int u = e.up_idx[j];
                const double fu = e.up_f[j];
                const double q = e.ssf_q[(size_t)u * M + m];
                qin = qin + q * (1.0 - fu);
                tor = tor + q * fu;
            }
        }
        const double dts = dt;
        const SsfPrep pp = ssf_prepare(sq0, col.zi, slope, sy, dts, flen, kh0, fk, INFINITY);
        double flx = cf.recharge * area;  // flux!(::RechargeModel) with check_flux
        if (col.zi >= dsoil) flx = wfh_max(0.0, flx);
        cum.rech += flx * dts;
        const double qnb = 0.0 + flx;
        cum.tor += tor * dts;
        double qn, zin, exf, net;
        const double z0 = col.zi;
        if (kinematic_wave_ssf<N>(qin, sq0, z0, qnb, slope, sy, dsoil, dts, flen, fw, qmax, kh0, fk, INFINITY, pp,
                                  col, true, qn, zin, exf, net))
            update_ustorelayerdepth<N>(col, z0, zin);
        sq = qn;
        szi = zin;
        cum.qin += qin * dts;
        cum.q += qn * dts;
        cum.exf += exf * dts;
        cum.qnet += net * area * dts;
    }
    e.ssf_q[i] = sq;
    // ---- average_flux_vars!, update_soil_water_storage!  (as ssf_finish) ----
    const double s_exf_avg = cum.exf / dt, s_qin_avg = cum.qin / dt, s_q_avg = cum.q / dt;
    const double s_tor_avg = cum.tor / dt;
    double usd = 0.0;
    {
        bool first = true;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (k < col.n_unsat) {
                usd = first ? col.ustore[k] : usd + col.ustore[k];
                first = false;
            }
    }
    (void)szi;
    const double sbm_runoff = wfh_max(0.0, s_exf_avg + cf.sat_exc + cf.rl + cf.inf_exc);
    const double swd = (dsoil - col.zi) * (ts - tr);
    const double net_runoff = sbm_runoff - cf.aol;
#pragma unroll
    for (int k = 0; k < N; ++k) e.ustore[k * nM + i] = col.ustore[k];
    e.swd[i] = swd;
    e.tsws[i] = swd + usd;
    double bal = 0.0;
    if (e.balance) {  // land balance  mass_balance.jl:215-253
        const double t_in = s_qin_avg / area + (cf.P + 0.0) + 0.0 + 0.0 / area;
        const double vout = net_runoff + cf.aet + cf.nrr + cf.leak + 0.0;
        const double t_out = s_q_avg / area + vout + 0.0 + 0.0 / area;
        const double storage = (swd + usd) + canopy_bal + snow_bal + 0.0 * 0.0 + 0.0;
        const double rate = (storage - cf.store_prev) / dt;
        bal = wfh_fabs(t_in - t_out - rate);
    }
    const double o_inwater = (net_runoff + 0.0) * area + 0.0;

    // ---- overland kinematic wave (one sub-step of length dt) ----
    double o_tor_avg;
    {
        double tor = 0.0, qin_all = 0.0;  // surface_kinwave.jl:317-326
        if (MODE == ENS_ROUTE) {
            tor = up_o_tor;
            qin_all = up_o_qin;
        } else {
            for (int j = e0; j < e1; ++j) {
                const int u = e.up_idx[j];
                const double fu = e.up_f[j];
                const double q = e.ol_q[(size_t)u * M + m];
                tor = tor + q * fu;
                qin_all = qin_all + q * (1.0 - fu);
            }
        }
        const double dts = dt;
        const double oqlat = o_inwater / flen;
        const KwPrep op = kw_prepare(oq0, oqlat, alpha, dts, flen);
        const double o_tor_cum = 0.0 + tor * dts;
        double qin = 0.0;
        if (sfw > 0.0) qin = qin_all;
        double qn, ar;
        ens_kw<VAR>(e, W, 0, (unsigned)v * 2654435761u ^ (unsigned)m * 40503u, qin, op, alpha, dts, flen, slope, qn,
                      ar, hist);
        // (without a surface flow width the depth is unchanged: read again rather than kept live)
        const double oh = sfw > 0.0 ? ar / sfw : e.ol_h[i];
        e.ol_q[i] = qn;
        e.ol_h[i] = oh;
        o_tor_avg = o_tor_cum / dt;
    }

    // ---- river kinematic wave (S_r sub-steps) ----
    if (riv) {
        const double rfl = e.r_fl[r], rfw = e.r_fw[r], rarea = e.r_area[r];
        const double r_inwater = (s_tor_avg + o_tor_avg + cf.nrr * area) + 0.0 * rarea;
        const double rqlat = r_inwater / rfl;
        double inflow = 0.0 / rfl;
        inflow -= 0.0 / rfl;
        const double rqlat_all = rqlat + inflow;
        // every step has at least one river sub-step, which sets h and the storage
        double rq = rq0, rh = 0.0, r_storage = 0.0, r_q_cum = 0.0;
        for (int k = 0; k < e.S_r; ++k) {
            const double dts = (k == e.S_r - 1) ? e.dts_r[1] : e.dts_r[0];
            double sumq = 0.0;  // surface_kinwave.jl:520
            if (MODE == ENS_ROUTE && k < PRE_SR) {
#pragma unroll
                for (int kk = 0; kk < PRE_SR; ++kk)
                    if (kk == k) sumq = up_r[kk];
            } else {
                const double* qk = (k == e.S_r - 1) ? e.r_q : e.rq_sub + (size_t)k * rM;
                for (int j = r_e0; j < r_e1; ++j) sumq = sumq + qk[(size_t)e.rup_idx[j] * M + m];
            }
            const KwPrep rp = kw_prepare(rq, rqlat_all, ralpha, dts, rfl);
            const double qin = 0.0 + sumq;
            double qn, ar;
            ens_kw<VAR>(e, W, 1, ((unsigned)v * 2654435761u ^ (unsigned)m * 40503u) + 97u * (unsigned)(k + 1), qin, rp,
                          ralpha, dts, rfl, slope, qn, ar, hist);
            rq = qn;
            rh = ar / rfw;
            r_storage = rfl * ar;
            r_q_cum += qn * dts;
            if (k < e.S_r - 1) e.rq_sub[(size_t)k * rM + ri] = qn;
        }
        e.r_q[ri] = rq;
        e.r_h[ri] = rh;
        e.r_storage[ri] = r_storage;
        e.r_q_avg[ri] = r_q_cum / dt;
    }
    return bal;
}

template <int N, int MODE, int VAR>
__global__ void __launch_bounds__(256, 2) k_ens_stage(EnsDev e, int s0, int ns) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    // the network's operands in LDS (before any wave leaves: every wave takes the barrier)
    __shared__ unsigned hist_s[2 * ENS_HIST_BINS + 2];
    const MlpOps* W = VAR == ENS_LEARN ? e.mlp : nullptr;  // read from L2 by the (rare) network calls
    unsigned* hist = nullptr;
    if (VAR != ENS_PLAIN && e.kw_hist) {  // (before any wave leaves: every wave takes the barrier)
        for (int k = threadIdx.x; k < 2 * ENS_HIST_BINS + 2; k += blockDim.x) hist_s[k] = 0u;
        hist = hist_s;
        __syncthreads();
    }
    if (w < ns * e.G) {
        const int st = __builtin_amdgcn_readfirstlane(s0 + w / e.G);
        const int g = __builtin_amdgcn_readfirstlane(w % e.G);
        const int m = g * WAVE + lane;
        const int c0 = e.strand_ptr[st], c1 = e.strand_ptr[st + 1];
        double worst = 0.0;
        for (int c = c0; c < c1; ++c) {
            const int v = __builtin_amdgcn_readfirstlane(e.strand_cell[c]);
            worst = wfh_max(worst, ens_cell<N, MODE, VAR>(e, v, m, W, hist));
        }
        if (e.balance && worst > 0.0) atomicMax(&e.bal_max[m], (unsigned long long)wfh_bits(worst));
    }
    if (hist) {
        __syncthreads();
        for (int k = threadIdx.x; k < 2 * ENS_HIST_BINS + 2; k += blockDim.x)
            if (hist_s[k]) atomicAdd(&e.kw_hist[k], (unsigned long long)hist_s[k]);
    }
}

__global__ void k_ens_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Cold-start states and member parameters. Base arrays are the base model's (node order); the
// derived parameters follow wfh_create's expressions with the member's kv_0 = kv_0 fa[m],
// f = f fb[m] and Manning n = n fc[m].
struct EnsInit {
    const double *canopy, *snow_s, *snow_w, *tsoil, *swd, *tsws, *ustore, *ol_q, *ol_h;
    const double *r_q, *r_h, *r_storage;
    const double *n_land, *sqrt_slope, *sfw_pow, *n_river, *sqrt_rslope, *wp_pow;
    const double *fc;
};

template <int N>
__global__ void k_ens_init(EnsDev e, EnsInit b) {
    const size_t M = (size_t)e.M;
    const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t nM = (size_t)e.n * M;
    if (idx >= nM) return;
    const int v = (int)(idx / M), m = (int)(idx % M);
    const size_t i = idx;
    e.canopy[i] = b.canopy[v];
    e.snow_s[i] = b.snow_s[v];
    e.snow_w[i] = b.snow_w[v];
    e.tsoil[i] = b.tsoil[v];
    e.swd[i] = b.swd[v];
    e.tsws[i] = b.tsws[v];
    for (int k = 0; k < N; ++k) e.ustore[k * nM + i] = b.ustore[(size_t)k * e.n + v];
    e.ol_q[i] = b.ol_q[v];
    e.ol_h[i] = b.ol_h[v];
    // initialize_lateral_ssf_model!  utils.jl:916-937 (exponential profile)
    const double kv0 = ens_ld(e.kv0, v) * e.fa[m], fk = ens_ld(e.f, v) * e.fb[m];
    const double kh0 = ens_ld(e.khfrac, v) * kv0;
    const double slope = ens_ld(e.slope, v), d = ens_ld(e.d, v), ts = ens_ld(e.ts, v), tr = ens_ld(e.tr, v);
    const double zi0 = wfh_max(0.0, d - b.swd[v] / (ts - tr));
    const_cast<double*>(e.qmax)[i] = ((kh0 * slope) / fk) * (1.0 - wfh_exp(-fk * d));
    e.ssf_q[i] = ((kh0 * slope) / fk) * (wfh_exp(-fk * zi0) - wfh_exp(-fk * d)) * ens_ld(e.fw, v);
    // ManningFlowParameters  surface_kinwave.jl:32-44
    const double at = wfh_pow((b.n_land[v] * b.fc[m]) / b.sqrt_slope[v], 0.6);
    const_cast<double*>(e.alpha)[i] = at * b.sfw_pow[v];
}

__global__ void k_ens_init_riv(EnsDev e, EnsInit b) {
    const size_t M = (size_t)e.M;
    const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (idx >= (size_t)e.nriv * M) return;
    const int r = (int)(idx / M), m = (int)(idx % M);
    e.r_q[idx] = b.r_q[r];
    e.r_h[idx] = b.r_h[r];
    e.r_storage[idx] = b.r_storage[r];
    e.r_q_avg[idx] = 0.0;
    const double at = wfh_pow((b.n_river[r] * b.fc[m]) / b.sqrt_rslope[r], 0.6);
    const_cast<double*>(e.r_alpha)[idx] = at * b.wp_pow[r];
}

// flags a non-finite forcing value in the ensemble's own copy and replaces the cell's three
// values by finite stand-ins, so that no NaN or infinity reaches the column's iterations (the
// step's results are reported unusable)
__global__ void k_ens_check_forcing(double* P, double* E, double* T, int n, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && wfh_nonfinite3(P[i], E[i], T[i])) {
        __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        P[i] = 0.0;
        E[i] = 0.0;
        T[i] = 283.15;
    }
}

// one member's field of a member-innermost array, into land-index order
__global__ void k_ens_take(double* out, const double* src, const int* node_of_index, int count, int M, int m) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < count) out[j] = src[(size_t)node_of_index[j] * M + m];
}

}  // namespace

struct EnsSchedule {
    int block = 256;       // threads per block
    int wide = 1024;       // levels with at least this many cells are stages of their own
    int strand_cells = 4;  // cells per strand in those stages
    std::vector<int> strand_ptr, strand_cell;  // host copy
    std::vector<std::pair<int, int>> stages;   // (first strand, strand count)
    std::vector<int> modes;                    // ENS_FUSED / ENS_COLUMN / ENS_ROUTE per stage
    std::vector<int> pre_slot;                 // node -> slot of a chain-stage cell, or -1
    int npre = 0;
    bool built = false;
    int* d_ptr = nullptr;
    int* d_cell = nullptr;
    int* d_pre_slot = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double ms = -1.0;  // measured step time (autotuner)
};

struct wfh_ensemble {
    wfh_handle* base = nullptr;
    int M = 0, N = 0;
    EnsDev e{};
    std::vector<void*> allocs;
    std::string err;
    std::vector<int> level;                   // node -> topological level
    std::vector<std::vector<int>> ups;        // node -> upstream nodes
    std::vector<int> down;                    // node -> downstream node or -1
    int max_level = 0;
    std::vector<EnsSchedule> cand;            // schedules (the autotuner's candidates)
    double* pre = nullptr;                    // pre-buffer shared by the schedules, [ENS_PRE][pre_cap * M]
    int pre_cap = 0;
    int cur = 0;                              // schedule in use
    bool tuning = false;
    int tune_next = 0, tune_trials = 1, tune_done = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    double timed_ms = 0.0;
    int64_t timed_steps = 0;
    bool forcing_set = false;
    bool broken = false;                      // non-finite forcing was stepped
    unsigned long long* hist_buf = nullptr;   // Newton update histogram (allocated when first enabled)
    unsigned long long* harvest_ctr = nullptr;
    int* d_index_node = nullptr;              // land index -> node (for per-member reads)
    double *fP = nullptr, *fE = nullptr, *fT = nullptr;
};

namespace {

int32_t efail(wfh_ensemble* x, int32_t code, const std::string& msg) {
    if (x) x->err = msg;
    return code;
}

#define EHIP(x, call)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return efail((x), WFH_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
T* ealloc(wfh_ensemble* x, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
    x->allocs.push_back(p);
    return (T*)p;
}

// The stages of one schedule (see the header comment of this file). The vertical column of the
// cells of chain stages runs first, in a parallel stage of its own (ENS_COLUMN), so that the
// chains carry the routing only.
void ens_build_schedule(const wfh_ensemble* x, EnsSchedule& s) {
    if (s.built) return;
    s.built = true;
    const int n = x->e.n;
    std::vector<std::vector<int>> bylevel(x->max_level + 1);
    for (int v = 0; v < n; ++v) bylevel[x->level[v]].push_back(v);
    s.strand_ptr.assign(1, 0);
    s.strand_cell.clear();
    s.stages.clear();
    s.modes.clear();
    std::vector<int> stage_of(n, -1);
    int L = 0;
    const int K = std::max(1, s.strand_cells);
    while (L <= x->max_level) {
        const int sidx = (int)s.stages.size();
        const int first = (int)s.strand_ptr.size() - 1;
        if ((int)bylevel[L].size() >= s.wide) {
            const std::vector<int>& cells = bylevel[L];
            for (size_t a = 0; a < cells.size(); a += K) {
                for (size_t b = a; b < std::min(cells.size(), a + (size_t)K); ++b) {
                    s.strand_cell.push_back(cells[b]);
                    stage_of[cells[b]] = sidx;
                }
                s.strand_ptr.push_back((int)s.strand_cell.size());
            }
            ++L;
            s.modes.push_back(ENS_FUSED);
        } else {
            // narrow levels L .. L2-1 while every cell has at most one upstream cell in the stage
            int L2 = L;
            while (L2 <= x->max_level && (int)bylevel[L2].size() < s.wide) {
                bool ok = true;
                for (int v : bylevel[L2]) {
                    int inside = 0;
                    for (int u : x->ups[v]) inside += stage_of[u] == sidx;
                    if (inside > 1) {
                        ok = false;
                        break;
                    }
                }
                if (!ok && L2 > L) break;
                for (int v : bylevel[L2]) stage_of[v] = sidx;
                ++L2;
            }
            // strands: from every cell without an upstream cell in the stage, down the chain
            for (int l = L; l < L2; ++l)
                for (int v : bylevel[l]) {
                    bool head = true;
                    for (int u : x->ups[v]) head = head && stage_of[u] != sidx;
                    if (!head) continue;
                    for (int w = v; w >= 0 && stage_of[w] == sidx; w = x->down[w]) s.strand_cell.push_back(w);
                    s.strand_ptr.push_back((int)s.strand_cell.size());
                }
            L = L2;
            s.modes.push_back(ENS_ROUTE);
        }
        s.stages.push_back({first, (int)s.strand_ptr.size() - 1 - first});
    }
    // the column pass over the chain stages' cells, first
    s.pre_slot.assign(n, -1);
    std::vector<int> pre;
    for (size_t k = 0; k < s.stages.size(); ++k) {
        if (s.modes[k] != ENS_ROUTE) continue;
        for (int t = s.stages[k].first; t < s.stages[k].first + s.stages[k].second; ++t)
            for (int c = s.strand_ptr[t]; c < s.strand_ptr[t + 1]; ++c) {
                s.pre_slot[s.strand_cell[c]] = (int)pre.size();
                pre.push_back(s.strand_cell[c]);
            }
    }
    s.npre = (int)pre.size();
    if (s.npre > 0) {
        const int first = (int)s.strand_ptr.size() - 1;
        for (size_t a = 0; a < pre.size(); a += K) {
            for (size_t b = a; b < std::min(pre.size(), a + (size_t)K); ++b) s.strand_cell.push_back(pre[b]);
            s.strand_ptr.push_back((int)s.strand_cell.size());
        }
        s.stages.insert(s.stages.begin(), {first, (int)s.strand_ptr.size() - 1 - first});
        s.modes.insert(s.modes.begin(), ENS_COLUMN);
    }
}

template <int N>
void ens_launch_stages(wfh_ensemble* x, const EnsSchedule& s, hipStream_t st) {
    EnsDev e = x->e;
    e.strand_ptr = s.d_ptr;
    e.strand_cell = s.d_cell;
    e.pre_slot = s.d_pre_slot;
    e.pre = x->pre;
    e.pre_stride = (size_t)x->pre_cap * x->M;
    if (e.balance) hipLaunchKernelGGL(k_ens_zero_u64, dim3((x->M + 255) / 256), dim3(256), 0, st, e.bal_max, x->M);
    const int wpb = s.block / WAVE;
    for (size_t k = 0; k < s.stages.size(); ++k) {
        const auto& sg = s.stages[k];
        const int waves = sg.second * e.G;
        const dim3 grid((waves + wpb - 1) / wpb), block(s.block);
        const int var = e.mlp ? ENS_LEARN : (e.kw_hist || e.harvest) ? ENS_INSTR : ENS_PLAIN;
        if (s.modes[k] == ENS_COLUMN)
            hipLaunchKernelGGL((k_ens_stage<N, ENS_COLUMN, ENS_PLAIN>), grid, block, 0, st, e, sg.first, sg.second);
        else if (s.modes[k] == ENS_ROUTE) {
            auto kr = var == ENS_LEARN   ? k_ens_stage<N, ENS_ROUTE, ENS_LEARN>
                      : var == ENS_INSTR ? k_ens_stage<N, ENS_ROUTE, ENS_INSTR>
                                         : k_ens_stage<N, ENS_ROUTE, ENS_PLAIN>;
            hipLaunchKernelGGL(kr, grid, block, 0, st, e, sg.first, sg.second);
        } else {
            auto kf = var == ENS_LEARN   ? k_ens_stage<N, ENS_FUSED, ENS_LEARN>
                      : var == ENS_INSTR ? k_ens_stage<N, ENS_FUSED, ENS_INSTR>
                                         : k_ens_stage<N, ENS_FUSED, ENS_PLAIN>;
            hipLaunchKernelGGL(kf, grid, block, 0, st, e, sg.first, sg.second);
        }
    }
}

int32_t ens_prepare(wfh_ensemble* x, EnsSchedule& s) {
    if (s.exec) return WFH_OK;
    ens_build_schedule(x, s);
    if (s.npre > x->pre_cap) return efail(x, WFH_ERR_STATE, "internal error: pre-buffer smaller than a schedule");
    if (!s.d_ptr) {
        s.d_ptr = ealloc<int>(x, s.strand_ptr.size());
        s.d_cell = ealloc<int>(x, s.strand_cell.size());
        s.d_pre_slot = ealloc<int>(x, s.pre_slot.size());
        if (!s.d_ptr || !s.d_cell || !s.d_pre_slot) return efail(x, WFH_ERR_HIP, "hipMalloc schedule");
        EHIP(x, hipMemcpy(s.d_pre_slot, s.pre_slot.data(), s.pre_slot.size() * sizeof(int), hipMemcpyHostToDevice));
        EHIP(x, hipMemcpy(s.d_ptr, s.strand_ptr.data(), s.strand_ptr.size() * sizeof(int), hipMemcpyHostToDevice));
        EHIP(x, hipMemcpy(s.d_cell, s.strand_cell.data(), s.strand_cell.size() * sizeof(int),
                          hipMemcpyHostToDevice));
    }
    // the step's launches, recorded once into a graph
    hipStream_t st = x->base->stream;
    EHIP(x, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    switch (x->N) {
        case 2: ens_launch_stages<2>(x, s, st); break;
        case 3: ens_launch_stages<3>(x, s, st); break;
        case 4: ens_launch_stages<4>(x, s, st); break;
        case 5: ens_launch_stages<5>(x, s, st); break;
        default: ens_launch_stages<6>(x, s, st); break;
    }
    EHIP(x, hipStreamEndCapture(st, &s.graph));
    EHIP(x, hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    return WFH_OK;
}

// The recorded graphs bake in the kernel arguments: drop them when an option changes.
void ens_invalidate(wfh_ensemble* x) {
    hipStreamSynchronize(x->base->stream);
    for (EnsSchedule& c : x->cand) {
        if (c.exec) hipGraphExecDestroy(c.exec);
        if (c.graph) hipGraphDestroy(c.graph);
        c.exec = nullptr;
        c.graph = nullptr;
    }
}

}  // namespace

extern "C" int32_t wfh_ens_create(const wfh_config* cfg, const wfh_static* st, int32_t members,
                                  const double* kv0_factor, const double* f_factor, const double* n_factor,
                                  wfh_ensemble** out) {
    if (!out) return WFH_ERR_ARG;
    *out = nullptr;
    if (!cfg || !st || !kv0_factor || !f_factor || !n_factor) {
        g_create_error = "null argument";
        return WFH_ERR_ARG;
    }
    if (members < WAVE || members % WAVE != 0 || members > ENS_MAX_MEMBERS) {
        g_create_error = "members must be a positive multiple of 64, at most 4096";
        return WFH_ERR_ARG;
    }
    for (int m = 0; m < members; ++m)
        if (!(kv0_factor[m] > 0.0) || !(f_factor[m] > 0.0) || !(n_factor[m] > 0.0) || !std::isfinite(kv0_factor[m]) ||
            !std::isfinite(f_factor[m]) || !std::isfinite(n_factor[m])) {
            g_create_error = "member factors must be finite and positive";
            return WFH_ERR_NONFINITE;
        }
    if (cfg->kinematic_wave_adaptive || cfg->floodplain_1d || cfg->snow_gravitational_transport || st->nres > 0 ||
        st->z_exp) {
        g_create_error = "ensemble: adaptive sub-steps, floodplain, snow transport, reservoirs and the "
                         "exponential-constant profile are not supported";
        return WFH_ERR_ARG;
    }
    if (cfg->land_dt < cfg->dt || cfg->subsurface_dt < cfg->dt) {
        g_create_error = "ensemble: one overland and one subsurface sub-step per step (land_dt, subsurface_dt >= dt)";
        return WFH_ERR_ARG;
    }
    wfh_handle* base = nullptr;
    int32_t rc = wfh_create(cfg, st, &base);
    if (rc) return rc;
    if (base->S_r > ENS_MAX_SR || base->S_s != 1 || base->S_l != 1) {
        wfh_destroy(base);
        g_create_error = "ensemble: at most 8 river sub-steps per step";
        return WFH_ERR_ARG;
    }
    wfh_ensemble* x = new wfh_ensemble();
    x->base = base;
    x->M = members;
    x->N = base->N;
    const int n = (int)base->n, nriv = (int)base->nriv, M = members;
    const Dev& D = base->d;
    EnsDev& e = x->e;
    auto bail = [&](int32_t code, const std::string& msg) {
        g_create_error = msg;
        wfh_ens_destroy(x);
        return code;
    };
    e.n = n;
    e.nriv = nriv;
    e.M = M;
    e.G = M / WAVE;
    e.dt = D.dt;
    e.gash = D.gash;
    e.snow = D.snow;
    e.infred = D.infred;
    e.balance = D.balance;
    std::memcpy(e.cfg_thick, D.cfg_thick, sizeof e.cfg_thick);
    std::memcpy(e.cfg_cum, D.cfg_cum, sizeof e.cfg_cum);
    e.S_r = base->S_r;
    e.dts_r[0] = base->dts_r[0];
    e.dts_r[1] = base->dts_r[1];
    e.gap = D.gap, e.smax = D.smax, e.kc = D.kc, e.er = D.er, e.cfmax = D.cfmax, e.tt = D.tt, e.tti = D.tti;
    e.ttm = D.ttm, e.whc = D.whc, e.rf = D.rf, e.wf = D.wf, e.sfrac = D.sfrac, e.ts = D.ts, e.tr = D.tr;
    e.tfc = D.tfc, e.kv0 = D.kv0, e.f = D.f, e.hb = D.hb, e.h1 = D.h1, e.h2 = D.h2, e.h3h = D.h3h, e.h3l = D.h3l;
    e.h4 = D.h4, e.ah1 = D.ah1, e.d = D.d, e.icap_c = D.icap_c, e.maxleak = D.maxleak, e.compf = D.compf;
    e.rdp = D.rdp, e.caph = D.caph, e.capn = D.capn, e.wsoil = D.wsoil, e.cfsoil = D.cfsoil, e.rd = D.rd;
    e.c = D.c, e.kvfrac = D.kvfrac, e.rootfrac = D.rootfrac;
    e.area = D.area, e.fl = D.fl, e.fw = D.fw, e.slope = D.slope, e.sy = D.sy, e.sfw = D.sfw;
    e.up_ptr = D.up_ptr, e.up_idx = D.up_idx, e.up_f = D.up_f, e.river_of = D.river_of;
    e.rup_ptr = D.rup_ptr, e.rup_idx = D.rup_idx;
    e.r_fw = D.r_fw, e.r_fl = D.r_fl, e.r_area = D.r_area;
    // khfrac in node order (the base model holds kh0 = khfrac kv_0 only)
    std::vector<double> khf(n), nl(n), sqs(n), sfwp(n), nr(std::max(nriv, 1)), sqr(std::max(nriv, 1)),
        wpp(std::max(nriv, 1));
    const double apow = (2.0 / 3.0) * 0.6;
    {
        // surface flow width as wfh_create derives it (domain.jl:218-261)
        std::vector<double> sfw_host(n);
        if (hipMemcpy(sfw_host.data(), D.sfw.p, (D.sfw.mask ? n : 1) * sizeof(double), hipMemcpyDeviceToHost) !=
            hipSuccess)
            return bail(WFH_ERR_HIP, "ensemble: reading the base model's surface flow width");
        for (int q = 0; q < n; ++q) {
            const int64_t i = base->land_of[q];
            khf[q] = st->khfrac[i];
            nl[q] = st->land_manning_n[i];
            sqs[q] = std::sqrt(wfh_clamp(st->land_slope[i], 0.00001, INFINITY));
            sfwp[q] = wfh_pow(D.sfw.mask ? sfw_host[q] : sfw_host[0], apow);
        }
        for (int r = 0; r < nriv; ++r) {
            nr[r] = st->river_manning_n[r];
            sqr[r] = std::sqrt(wfh_clamp(st->river_slope[r], 0.00001, INFINITY));
            wpp[r] = wfh_pow(st->river_width[r] + st->river_bankfull_depth[r], apow);
        }
    }
    // topology on the host (node order): levels, upstream lists, downstream
    x->ups.assign(n, {});
    x->down.assign(n, -1);
    x->level.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t w = base->land_down[i];
        if (w >= 0) {
            x->down[base->node_of[i]] = base->node_of[w];
            x->ups[base->node_of[w]].push_back(base->node_of[i]);
        }
    }
    for (int q = 0; q < n; ++q) {  // node order is topological
        int lv = 0;
        for (int u : x->ups[q]) lv = std::max(lv, x->level[u] + 1);
        x->level[q] = lv;
        x->max_level = std::max(x->max_level, lv);
    }
    const size_t nM = (size_t)n * M, rM = (size_t)std::max(nriv, 1) * M;
    double** big[] = {&e.canopy, &e.snow_s, &e.snow_w, &e.tsoil, &e.swd, &e.tsws, &e.ssf_q, &e.ol_q, &e.ol_h};
    for (double** pp : big)
        if (!(*pp = ealloc<double>(x, nM))) return bail(WFH_ERR_HIP, "ensemble: hipMalloc member states");
    if (!(e.ustore = ealloc<double>(x, nM * x->N))) return bail(WFH_ERR_HIP, "ensemble: hipMalloc member states");
    double* qmax = ealloc<double>(x, nM);
    double* alpha = ealloc<double>(x, nM);
    if (!qmax || !alpha) return bail(WFH_ERR_HIP, "ensemble: hipMalloc member parameters");
    e.qmax = qmax;
    e.alpha = alpha;
    double** rv[] = {&e.r_q, &e.r_h, &e.r_storage, &e.r_q_avg};
    for (double** pp : rv)
        if (!(*pp = ealloc<double>(x, rM))) return bail(WFH_ERR_HIP, "ensemble: hipMalloc river states");
    double* ra = ealloc<double>(x, rM);
    e.rq_sub = ealloc<double>(x, rM * std::max(1, e.S_r - 1));
    e.bal_max = ealloc<unsigned long long>(x, M);
    if (!ra || !e.rq_sub || !e.bal_max) return bail(WFH_ERR_HIP, "ensemble: hipMalloc river states");
    e.r_alpha = ra;
    double* fa = ealloc<double>(x, M);
    double* fb = ealloc<double>(x, M);
    double* fc = ealloc<double>(x, M);
    double* khd = ealloc<double>(x, n);
    x->fP = ealloc<double>(x, n);
    x->fE = ealloc<double>(x, n);
    x->fT = ealloc<double>(x, n);
    x->d_index_node = ealloc<int>(x, n);
    double *nld = ealloc<double>(x, n), *sqsd = ealloc<double>(x, n), *sfwpd = ealloc<double>(x, n);
    double *nrd = ealloc<double>(x, nr.size()), *sqrd = ealloc<double>(x, sqr.size()), *wppd = ealloc<double>(x, wpp.size());
    if (!fa || !fb || !fc || !khd || !x->fP || !x->fE || !x->fT || !x->d_index_node || !nld || !sqsd || !sfwpd ||
        !nrd || !sqrd || !wppd)
        return bail(WFH_ERR_HIP, "ensemble: hipMalloc");
    std::vector<int> idx_node(n);
    for (int64_t i = 0; i < n; ++i) idx_node[i] = base->node_of[i];
    struct Up {
        void* d;
        const void* h;
        size_t bytes;
    } ups[] = {{fa, kv0_factor, M * sizeof(double)},        {fb, f_factor, M * sizeof(double)},
               {fc, n_factor, M * sizeof(double)},          {khd, khf.data(), n * sizeof(double)},
               {x->d_index_node, idx_node.data(), n * sizeof(int)}, {nld, nl.data(), n * sizeof(double)},
               {sqsd, sqs.data(), n * sizeof(double)},      {sfwpd, sfwp.data(), n * sizeof(double)},
               {nrd, nr.data(), nr.size() * sizeof(double)}, {sqrd, sqr.data(), sqr.size() * sizeof(double)},
               {wppd, wpp.data(), wpp.size() * sizeof(double)}};
    for (const Up& u : ups)
        if (hipMemcpy(u.d, u.h, u.bytes, hipMemcpyHostToDevice) != hipSuccess)
            return bail(WFH_ERR_HIP, "ensemble: upload");
    e.fa = fa;
    e.fb = fb;
    e.khfrac = Prm{khd, 0xFFFFFFFFu};
    e.mlp = nullptr;
    e.mlp_gate = INFINITY;
    e.kw_hist = nullptr;
    e.harvest = nullptr;
    e.harvest_n = nullptr;
    e.harvest_cap = 0;
    e.harvest_every = 1;
    e.P = x->fP;
    e.E = x->fE;
    e.T = x->fT;
    EnsInit b{D.canopy, D.snow_s, D.snow_w, D.tsoil, D.swd, D.tsws, D.ustore, D.ol_q, D.ol_h, D.r_q, D.r_h,
              D.r_storage, nld, sqsd, sfwpd, nrd, sqrd, wppd, fc};
    hipStream_t s = base->stream;
    const unsigned gi = (unsigned)((nM + 255) / 256), gr = (unsigned)(((size_t)nriv * M + 255) / 256);
    switch (x->N) {
        case 2: hipLaunchKernelGGL(k_ens_init<2>, dim3(gi), dim3(256), 0, s, e, b); break;
        case 3: hipLaunchKernelGGL(k_ens_init<3>, dim3(gi), dim3(256), 0, s, e, b); break;
        case 4: hipLaunchKernelGGL(k_ens_init<4>, dim3(gi), dim3(256), 0, s, e, b); break;
        case 5: hipLaunchKernelGGL(k_ens_init<5>, dim3(gi), dim3(256), 0, s, e, b); break;
        default: hipLaunchKernelGGL(k_ens_init<6>, dim3(gi), dim3(256), 0, s, e, b); break;
    }
    if (nriv > 0) hipLaunchKernelGGL(k_ens_init_riv, dim3(gr), dim3(256), 0, s, e, b);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return bail(WFH_ERR_HIP, "ensemble: initialisation kernels");
    if (hipEventCreate(&x->ev0) != hipSuccess || hipEventCreate(&x->ev1) != hipSuccess)
        return bail(WFH_ERR_HIP, "hipEventCreate");
    // default schedule, and the autotuner's candidates: block size, wide-level threshold, cells
    // per strand of a wide level
    const int blocks[] = {256, 64}, wides[] = {1024, 256, 8192}, strands[] = {4, 1, 16};
    for (int w : wides)
        for (int k : strands)
            for (int bl : blocks) {
                EnsSchedule c;
                c.block = bl;
                c.wide = w;
                c.strand_cells = k;
                x->cand.push_back(c);
            }
    x->cur = 0;
    for (EnsSchedule& c : x->cand) {
        ens_build_schedule(x, c);
        x->pre_cap = std::max(x->pre_cap, c.npre);
    }
    if (x->pre_cap > 0 && !(x->pre = ealloc<double>(x, (size_t)ENS_PRE * x->pre_cap * M)))
        return bail(WFH_ERR_HIP, "ensemble: hipMalloc pre-buffer");
    rc = ens_prepare(x, x->cand[0]);
    if (rc) return bail(rc, x->err);
    *out = x;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_set_forcing(wfh_ensemble* x, const double* P, const double* E, const double* T,
                                       int32_t memkind) {
    if (!x) return WFH_ERR_ARG;
    if (!P || !E || !T) return efail(x, WFH_ERR_ARG, "null forcing pointer");
    const int n = x->e.n;
    hipStream_t s = x->base->stream;
    EHIP(x, hipSetDevice(x->base->cfg.device));
    const double* src[3] = {P, E, T};
    double* dst[3] = {x->fP, x->fE, x->fT};
    if (memkind == WFH_MEM_HOST) {
        std::vector<double> tmp(n);
        for (int k = 0; k < 3; ++k) {
            for (int q = 0; q < n; ++q) tmp[q] = src[k][x->base->land_of[q]];
            EHIP(x, hipMemcpyAsync(dst[k], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
            EHIP(x, hipStreamSynchronize(s));
        }
    } else if (memkind == WFH_MEM_DEVICE) {
        for (int k = 0; k < 3; ++k)
            hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, dst[k], src[k], x->base->d_land_of, n);
        EHIP(x, hipGetLastError());
    } else if (memkind == WFH_MEM_DEVICE_NODE_ORDER) {
        for (int k = 0; k < 3; ++k)
            EHIP(x, hipMemcpyAsync(dst[k], src[k], n * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        return efail(x, WFH_ERR_ARG, "bad memkind");
    }
    hipLaunchKernelGGL(k_ens_check_forcing, dim3((n + 255) / 256), dim3(256), 0, s, x->fP, x->fE, x->fT, n,
                       x->base->d.nonfinite_flag);
    EHIP(x, hipGetLastError());
    x->forcing_set = true;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_step(wfh_ensemble* x) {
    if (!x) return WFH_ERR_ARG;
    if (x->broken) return efail(x, WFH_ERR_STATE, "the ensemble had non-finite forcing; its states are not usable");
    if (!x->forcing_set) return efail(x, WFH_ERR_STATE, "wfh_ens_set_forcing must be called before wfh_ens_step");
    hipStream_t s = x->base->stream;
    EHIP(x, hipSetDevice(x->base->cfg.device));
    int which = x->cur;
    const bool tune = x->tuning;
    if (tune) which = x->tune_next;
    int32_t rc = ens_prepare(x, x->cand[which]);
    if (rc) return rc;
    const bool timed = tune || x->timing;
    if (timed) EHIP(x, hipEventRecord(x->ev0, s));
    EHIP(x, hipGraphLaunch(x->cand[which].exec, s));
    if (timed) EHIP(x, hipEventRecord(x->ev1, s));
    if (timed) {
        EHIP(x, hipEventSynchronize(x->ev1));
        float ms = 0.0f;
        EHIP(x, hipEventElapsedTime(&ms, x->ev0, x->ev1));
        if (x->timing) {
            x->timed_ms += ms;
            ++x->timed_steps;
        }
        if (tune) {
            EnsSchedule& c = x->cand[which];
            c.ms = c.ms < 0.0 ? ms : std::min(c.ms, (double)ms);
            if (++x->tune_next >= (int)x->cand.size()) {
                x->tune_next = 0;
                if (++x->tune_done >= x->tune_trials) {
                    x->tuning = false;
                    int best = 0;
                    for (int k = 1; k < (int)x->cand.size(); ++k)
                        if (x->cand[k].ms < x->cand[best].ms) best = k;
                    x->cur = best;
                }
            }
        }
    }
    x->forcing_set = false;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_set_autotune(wfh_ensemble* x, int32_t trials) {
    if (!x) return WFH_ERR_ARG;
    if (trials < 0) return efail(x, WFH_ERR_ARG, "trials must be >= 0");
    x->tuning = trials > 0;
    x->tune_trials = std::max(1, trials);
    x->tune_next = 0;
    x->tune_done = 0;
    for (EnsSchedule& c : x->cand) c.ms = -1.0;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_set_schedule(wfh_ensemble* x, int32_t block, int32_t wide, int32_t strand_cells) {
    if (!x) return WFH_ERR_ARG;
    if ((block != 64 && block != 128 && block != 256) || wide < 1 || strand_cells < 1)
        return efail(x, WFH_ERR_ARG, "block must be 64, 128 or 256; wide and strand_cells >= 1");
    for (int k = 0; k < (int)x->cand.size(); ++k)
        if (x->cand[k].block == block && x->cand[k].wide == wide && x->cand[k].strand_cells == strand_cells) {
            x->cur = k;
            x->tuning = false;
            return WFH_OK;
        }
    EnsSchedule c;
    c.block = block;
    c.wide = wide;
    c.strand_cells = strand_cells;
    ens_build_schedule(x, c);
    if (c.npre > x->pre_cap)
        return efail(x, WFH_ERR_ARG, "this schedule has more chain-stage cells than the pre-buffer holds; use a "
                                     "larger `wide` threshold");
    x->cand.push_back(c);
    x->cur = (int)x->cand.size() - 1;
    x->tuning = false;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_get_schedule(wfh_ensemble* x, int32_t* block, int32_t* wide, int32_t* strand_cells,
                                        int32_t* stages, int32_t* tuning) {
    if (!x) return WFH_ERR_ARG;
    EnsSchedule& c = x->cand[x->cur];
    int32_t rc = ens_prepare(x, c);
    if (rc) return rc;
    if (block) *block = c.block;
    if (wide) *wide = c.wide;
    if (strand_cells) *strand_cells = c.strand_cells;
    if (stages) *stages = (int32_t)c.stages.size();
    if (tuning) *tuning = x->tuning ? 1 : 0;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_sync(wfh_ensemble* x) {
    if (!x) return WFH_ERR_ARG;
    EHIP(x, hipStreamSynchronize(x->base->stream));
    int flag = 0;
    EHIP(x, hipMemcpy(&flag, x->base->d.nonfinite_flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        x->broken = true;
        return efail(x, WFH_ERR_NONFINITE, "non-finite forcing (precipitation, potential evaporation or temperature)");
    }
    return WFH_OK;
}

// Member m's field of a state or river output, in land-index (or river-index) order.
extern "C" int32_t wfh_ens_get(wfh_ensemble* x, int32_t var, int32_t member, double* dst, int64_t count) {
    if (!x || !dst) return WFH_ERR_ARG;
    if (member < 0 || member >= x->M) return efail(x, WFH_ERR_ARG, "member out of range");
    const EnsDev& e = x->e;
    const int n = e.n, nriv = e.nriv;
    const size_t nM = (size_t)n * x->M;
    const double* src = nullptr;
    int layers = 1;
    bool river = false;
    switch (var) {
        case WFH_CANOPY_STORAGE: src = e.canopy; break;
        case WFH_SNOW_STORAGE: src = e.snow_s; break;
        case WFH_SNOW_WATER: src = e.snow_w; break;
        case WFH_SOIL_TEMPERATURE: src = e.tsoil; break;
        case WFH_SATWATERDEPTH: src = e.swd; break;
        case WFH_USTORELAYERDEPTH: src = e.ustore; layers = x->N; break;
        case WFH_SSF_Q: src = e.ssf_q; break;
        case WFH_SSF_QMAX: src = e.qmax; break;
        case WFH_LAND_Q: src = e.ol_q; break;
        case WFH_LAND_H: src = e.ol_h; break;
        case WFH_LAND_ALPHA: src = e.alpha; break;
        case WFH_RIVER_Q: src = e.r_q; river = true; break;
        case WFH_RIVER_H: src = e.r_h; river = true; break;
        case WFH_RIVER_STORAGE: src = e.r_storage; river = true; break;
        case WFH_RIVER_Q_AVERAGE: src = e.r_q_avg; river = true; break;
        case WFH_RIVER_ALPHA: src = e.r_alpha; river = true; break;
        default: return efail(x, WFH_ERR_ARG, "variable not held per ensemble member");
    }
    const int64_t need = river ? nriv : (int64_t)layers * n;
    if (count != need) return efail(x, WFH_ERR_ARG, "count does not match the variable");
    hipStream_t s = x->base->stream;
    EHIP(x, hipSetDevice(x->base->cfg.device));
    int32_t rc = wfh_ens_sync(x);
    if (rc) return rc;
    if (need == 0) return WFH_OK;
    double* tmp = nullptr;
    EHIP(x, hipMalloc(&tmp, need * sizeof(double)));
    if (river) {
        // river index r is member-innermost at r * M + m already
        std::vector<int> idx(nriv);
        for (int r = 0; r < nriv; ++r) idx[r] = r;
        int* di = nullptr;
        if (hipMalloc(&di, nriv * sizeof(int)) != hipSuccess) {
            hipFree(tmp);
            return efail(x, WFH_ERR_HIP, "hipMalloc");
        }
        hipMemcpy(di, idx.data(), nriv * sizeof(int), hipMemcpyHostToDevice);
        hipLaunchKernelGGL(k_ens_take, dim3((nriv + 255) / 256), dim3(256), 0, s, tmp, src, di, nriv, x->M, member);
        hipStreamSynchronize(s);
        hipFree(di);
    } else {
        for (int k = 0; k < layers; ++k)
            hipLaunchKernelGGL(k_ens_take, dim3((n + 255) / 256), dim3(256), 0, s, tmp + (size_t)k * n,
                               src + (size_t)k * nM, x->d_index_node, n, x->M, member);
    }
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpy(dst, tmp, need * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e1 != hipSuccess || e2 != hipSuccess) return efail(x, WFH_ERR_HIP, "ensemble read-back");
    return WFH_OK;
}

// Largest |land balance error| of the last step, per member (water_mass_balance only).
extern "C" int32_t wfh_ens_balance(wfh_ensemble* x, double* max_abs_error) {
    if (!x || !max_abs_error) return WFH_ERR_ARG;
    if (!x->e.balance) return efail(x, WFH_ERR_STATE, "the ensemble was created without water_mass_balance");
    std::vector<unsigned long long> b(x->M);
    EHIP(x, hipSetDevice(x->base->cfg.device));
    EHIP(x, hipStreamSynchronize(x->base->stream));
    EHIP(x, hipMemcpy(b.data(), x->e.bal_max, x->M * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int m = 0; m < x->M; ++m) max_abs_error[m] = wfh_from_bits(b[m]);
    return WFH_OK;
}

extern "C" int32_t wfh_ens_set_timing(wfh_ensemble* x, int32_t enable) {
    if (!x) return WFH_ERR_ARG;
    x->timing = enable != 0;
    x->timed_ms = 0.0;
    x->timed_steps = 0;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_get_timing(wfh_ensemble* x, double* step_ms) {
    if (!x || !step_ms) return WFH_ERR_ARG;
    *step_ms = x->timed_steps ? x->timed_ms / (double)x->timed_steps : 0.0;
    return WFH_OK;
}

// Algorithmic HBM bytes of one ensemble step: member states read and written once, the member
// parameters read once, upstream discharge read once per edge, shared per-cell data once.
extern "C" int32_t wfh_ens_step_bytes(wfh_ensemble* x, int64_t* bytes) {
    if (!x || !bytes) return WFH_ERR_ARG;
    const int64_t n = x->e.n, nr = x->e.nriv, M = x->M, N = x->N;
    int64_t E = 0, Er = 0;
    for (int v = 0; v < n; ++v) E += (int64_t)x->ups[v].size();
    for (int64_t r = 0; r < nr; ++r) Er += x->base->river_down[r] >= 0;
    const int64_t states = 4 + (x->e.snow ? 2 : 0) + 1 + N + 3;  // canopy, tsoil, swd, tsws; snow; ustore; ssf, ol q/h
    int64_t b = 8 * n * M * (2 * states + 2);                     // + qmax, alpha
    b += 8 * E * M * 2;                                            // upstream subsurface and overland q
    b += 8 * nr * M * (4 + 4 + x->e.S_r * 1) + 8 * Er * M * x->e.S_r;  // river states, sub-step q, folds
    b += n * (8 * 40 + 4 * 4);                                     // shared parameters and topology
    *bytes = b;
    return WFH_OK;
}

// The learned warm start of the kinematic-wave solves (NULL: off): consulted by a wave where some
// member's previous discharge is below `gate` (INFINITY: every solve).
extern "C" int32_t wfh_ens_set_learned(wfh_ensemble* x, wfh_mlp* mlp, double gate) {
    if (!x) return WFH_ERR_ARG;
    if (mlp && mlp->device != x->base->cfg.device) return efail(x, WFH_ERR_ARG, "network on another device");
    if (!(gate >= 0.0)) return efail(x, WFH_ERR_ARG, "gate must be >= 0");
    ens_invalidate(x);
    x->e.mlp = mlp ? mlp->ops : nullptr;
    x->e.mlp_gate = gate;
    return WFH_OK;
}

// Harvest (features [8], target, Newton updates) rows of the plain solves of the following steps
// into the device buffer rows [cap][10], one solve in `every` (by a hash of cell, member and
// sub-step); NULL stops. The target is log10(q) - 5 log10(reference start u).
extern "C" int32_t wfh_ens_set_harvest(wfh_ensemble* x, float* rows, int64_t cap, int32_t every) {
    if (!x) return WFH_ERR_ARG;
    if (rows && (cap <= 0 || every < 1)) return efail(x, WFH_ERR_ARG, "cap must be > 0 and every >= 1");
    EHIP(x, hipSetDevice(x->base->cfg.device));
    if (!x->harvest_ctr && !(x->harvest_ctr = ealloc<unsigned long long>(x, 1)))
        return efail(x, WFH_ERR_HIP, "hipMalloc");
    ens_invalidate(x);
    hipLaunchKernelGGL(k_ens_zero_u64, dim3(1), dim3(64), 0, x->base->stream, x->harvest_ctr, 1);
    EHIP(x, hipGetLastError());
    x->e.harvest = rows;
    x->e.harvest_n = x->harvest_ctr;
    x->e.harvest_cap = cap;
    x->e.harvest_every = every;
    return WFH_OK;
}

extern "C" int32_t wfh_ens_harvest_count(wfh_ensemble* x, int64_t* rows) {
    if (!x || !rows) return WFH_ERR_ARG;
    unsigned long long v = 0;
    EHIP(x, hipSetDevice(x->base->cfg.device));
    EHIP(x, hipStreamSynchronize(x->base->stream));
    if (x->harvest_ctr) EHIP(x, hipMemcpy(&v, x->harvest_ctr, sizeof v, hipMemcpyDeviceToHost));
    *rows = (int64_t)std::min<unsigned long long>(v, (unsigned long long)std::max<int64_t>(x->e.harvest_cap, 0));
    return WFH_OK;
}

// Histogram of Newton updates per kinematic-wave solve over the steps since it was enabled:
// hist [2][16] (overland, river; bins 0..8 updates, 9-16, 17-32, 33-64, 65-128, 129-256,
// 257-2999, 3000) then learned starts taken [2]. enable = 1 (re)starts counting, 0 stops; out
// (34 values) may be NULL.
extern "C" int32_t wfh_ens_kw_histogram(wfh_ensemble* x, int32_t enable, uint64_t* out) {
    if (!x) return WFH_ERR_ARG;
    EHIP(x, hipSetDevice(x->base->cfg.device));
    const int nb = 2 * ENS_HIST_BINS + 2;
    if (out) {
        EHIP(x, hipStreamSynchronize(x->base->stream));
        if (x->hist_buf)
            EHIP(x, hipMemcpy(out, x->hist_buf, nb * sizeof(uint64_t), hipMemcpyDeviceToHost));
        else
            std::fill(out, out + nb, (uint64_t)0);
    }
    if (enable == 1) {
        if (!x->hist_buf && !(x->hist_buf = ealloc<unsigned long long>(x, nb))) return efail(x, WFH_ERR_HIP, "hipMalloc");
        hipLaunchKernelGGL(k_ens_zero_u64, dim3(1), dim3(64), 0, x->base->stream, x->hist_buf, nb);
        EHIP(x, hipGetLastError());
        if (!x->e.kw_hist) {
            ens_invalidate(x);
            x->e.kw_hist = x->hist_buf;
        }
    } else if (enable == 0 && x->e.kw_hist) {
        ens_invalidate(x);
        x->e.kw_hist = nullptr;
    }
    return WFH_OK;
}

extern "C" const char* wfh_ens_last_error(const wfh_ensemble* x) {
    if (!x) return g_create_error.c_str();
    return x->err.c_str();
}

extern "C" void wfh_ens_destroy(wfh_ensemble* x) {
    if (!x) return;
    if (x->base && x->base->stream) hipStreamSynchronize(x->base->stream);
    for (EnsSchedule& c : x->cand) {
        if (c.exec) hipGraphExecDestroy(c.exec);
        if (c.graph) hipGraphDestroy(c.graph);
    }
    for (void* p : x->allocs) hipFree(p);
    if (x->ev0) hipEventDestroy(x->ev0);
    if (x->ev1) hipEventDestroy(x->ev1);
    if (x->base) wfh_destroy(x->base);
    delete x;
}
