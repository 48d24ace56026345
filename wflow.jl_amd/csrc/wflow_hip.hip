// wflow_hip.hip - MI355X (gfx950) implementation of the Wflow.jl `sbm` time step behind the C ABI
// declared in include/wflow_hip.h.
//
// Device layout
//   * Land cells are renumbered in topological-level order (level = longest upstream path in
//     cells, headwaters first; ascending land index inside a level). Every per-cell array is a
//     structure-of-arrays FP64 buffer in that order, so both kernels read and write coalesced.
//   * Layered soil variables are layer-major: a[k * n + p].
//   * Parameter maps whose values are all equal are stored once (mask 0) and broadcast.
//   * Upstream lists are int32 CSR in node order; inside a list the upstream cells keep
//     ascending land index, the order in which the reference folds `sum_at`
//     (Wflow/src/utils.jl:472-477).
//
// Kernels per step
//   K1 k_vertical  one thread per cell: interception, snow, open-water runoff, SBM soil column
//                  (update_land_hydrology_model!, Wflow/src/sbm.jl:82-132).
//   K2 k_route     persistent dataflow kernel over all routing of the step: lateral subsurface
//                  flow (kinwave_subsurface_update!), update_soil_water_storage!, overland and
//                  river kinematic waves for every fixed sub-step, total storage and the water
//                  balance. One lane owns one cell for the whole step. A lane waits only on the
//                  upstream cells' outflow of the same sub-step, published into single-consumer
//                  slots: a slot holds a sentinel NaN until the upstream lane stores its q
//                  (relaxed agent-scope atomic store); the downstream lane polls it, folds it into
//                  its inflow sum and resets it to the sentinel. Wave tickets are handed out in
//                  level order, so a lane only ever waits on lanes that already hold a ticket; the
//                  kernel therefore drains without any grid-wide barrier. Waits are bounded: a
//                  lane that spins past the limit raises an abort flag that every wave polls, and
//                  the host reports WFH_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wflow_hip.h"
#include "wfh_math.h"
#include "wfh_physics.h"
#include "wfh_column.h"

using namespace wfh;

namespace {

constexpr uint64_t SLOT_EMPTY = 0xFFF40000DEADBEEFULL;  // signalling NaN, never produced by arithmetic
constexpr int WAVE = 64;
constexpr int ROUTE_BLOCK = 256;
// Waves per SIMD the routing kernel is compiled for: two (at most 256 registers per lane) up to
// four soil layers, so that one wave's memory round trips and slot polls overlap the other's
// arithmetic; one for five and six layers, whose column would spill at two.
constexpr int route_waves(const int nlayers) { return nlayers <= 4 ? 2 : 1; }
// Polls without progress (each with a sleep) before a wave gives up (RouteArgs::spin_limit). A
// wave idle this long raises the abort flag; every other wave reads the flag after at most 256
// idle polls of its own and leaves the kernel.
constexpr long SPIN_LIMIT = 2000000;

struct Prm {  // parameter map, broadcast when uniform
    const double* p;
    uint32_t mask;
};
struct PrmL {  // layered parameter map [N][n]
    const double* p;
    uint32_t mask;
    uint32_t kstride;
};
__device__ __forceinline__ double ld(const Prm& a, int i) { return a.p[i & a.mask]; }
__device__ __forceinline__ double ld(const PrmL& a, int k, int i) { return a.p[k * a.kstride + (i & a.mask)]; }

// Accesses of the arrays a step writes. COH (the multi-step run, where a cell's next step may run
// on another wave, CU or XCD while the kernel is still running) makes them agent-scope relaxed
// atomics, which bypass the non-coherent caches, as the outflow slots do; otherwise plain.
template <bool COH>
__device__ __forceinline__ double ldv(const double* p) {
    if constexpr (COH)
        return wfh_from_bits(__hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    else
        return *p;
}
template <bool COH>
__device__ __forceinline__ void stv(double* p, const double v) {
    if constexpr (COH)
        __hip_atomic_store((uint64_t*)p, wfh_bits(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

struct Dev {
    int n, nriv;
    // forcing (node order)
    const double *P, *E, *T;
    // ---- static maps (node order) ----
    Prm gap, smax, kc, er, cfmax, tt, tti, ttm, whc, rf, wf, sfrac;
    Prm ts, tr, tfc, kv0, f, hb, h1, h2, h3h, h3l, h4, ah1, d, icap_s, icap_c, maxleak, compf, rdp, caph, capn, wsoil,
        cfsoil, rd;
    PrmL c, kvfrac, rootfrac;
    Prm area, fl, fw, slope, sy, kh0, qmax, alpha, sfw, elev;
    Prm zexp;  // depth below which Ksat stays constant (INFINITY: exponential profile)
    // topology
    const int* up_ptr;   // [n+1]
    const int* up_idx;   // [E] node ids of upstream cells
    const double* up_f;  // [E] flow_fraction_to_river of the upstream cell
    const int* river_of; // [n] river id or -1
    const uint8_t* flags;  // bit0: has a land downstream, bit1: river cell, bit2: has a river downstream,
                           // bit3: in a reservoir's area (no precipitation / evaporation on land)
    const int* rup_ptr;  // [nriv+1]
    const int* rup_idx;  // [Er] river ids
    const int* rup_node; // [Er] node ids of the upstream river cells
    const int* down;     // [n] downstream node id or -1
    // river static (river order)
    const double *r_alpha, *r_fw, *r_fl, *r_area;
    // ---- land states ----
    double *canopy, *snow_s, *snow_w, *tsoil, *swd, *ustore, *tsws;
    double *ssf_q, *ol_q, *ol_h, *ol_storage;
    double *r_q, *r_h, *r_storage;
    // ---- K1 outputs ----
    double *zi, *interception, *swe, *transfer, *recharge, *act_inf, *inf_excess, *sat_excess, *transpiration,
        *soil_evap, *capflux, *leakage, *aet, *net_runoff_river, *runoff_land, *aeow_l, *store_prev;
    // ---- K2 outputs ----
    double *ustoredepth, *runoff, *net_runoff, *exfilt, *vwc, *vwc_root, *rzs;
    double *ssf_qin_avg, *ssf_q_avg, *ssf_tor_avg, *ssf_storage;
    double *ol_qin, *ol_q_avg, *ol_qin_avg, *ol_tor_avg, *ol_inwater;
    double *r_qin, *r_q_avg, *r_qin_avg, *r_inwater;
    double *total_storage, *bal_land, *bal_land_rel, *bal_ssf, *bal_ol, *bal_r;
    // ---- scalars ----
    double cfg_thick[WFH_MAX_LAYERS], cfg_cum[WFH_MAX_LAYERS + 1];
    double dt;
    int gash, snow, infred, balance;
    // ---- reservoirs (nres > 0) ----
    int nres;
    const int* res_of_r;       // [nriv] reservoir whose outlet is the river cell, or -1
    const double* rup_res;     // [Er] 1 when the edge's upstream river cell is a reservoir outlet
    const int *res_cov_ptr, *res_cov_idx;  // nodes of each reservoir's area, ascending land index
    const int* res_type;       // [nres] 1 rating curve, 2 free weir, 3 modified Puls, 4 simple
    const int* res_sct;        // [nres] storage curve: 1 linear, 2 interpolation of the S-H table
    const int *res_sh_ptr, *res_hq_ptr;  // [nres+1] rows of each reservoir's S-H / H-Q table
    const double *res_sh_h, *res_sh_s;   // S-H rows
    const double *res_hq_h, *res_hq_q;   // H-Q rows; q is [rows][365], one column per julian day
    const double* res_shx;     // [nres][4] minimum / maximum of the S-H table's S and H
    const double* res_hq_ext;  // [nres][732]: minimum / maximum of H, then of each day's Q column
                               // (rating-curve reservoirs only)
    int res_hq_col;            // julian day - 1
    const double *res_area, *res_demand, *res_maxrel, *res_maxsto, *res_tff, *res_tmf, *res_thr, *res_rcc;
    const double* res_rce;     // [nres] free weir rating curve exponent
    int* abort_flag;           // the routing kernels' abort flag (RouteArgs::abort_flag)
    int* nonfinite_flag;       // set by the vertical kernel when a cell's forcing is not finite
    const int* chain_land;     // [nchunks] 1: a chain chunk of the land lists (see run_chain)
    const int* chain_rl;       // [nchunks] 1: all 64 cells are river cells and a chain of the river lists
    const int* chain_sup;      // [nchunks] 1: a chain chunk of the snow transport's lists (sup_ptr)
    int res_linked;            // some free weir is linked to a lower reservoir
    const int *res_lower, *res_upper;  // [nres] linked lower / upper reservoir, or -1
    uint64_t* res_xfer;        // [nres] single-consumer slot: the linked free weir's reverse flow of
                               // the sub-step into this reservoir (+inf: none)
    const double *res_ext, *res_obs;
    double *res_storage, *res_wl, *res_outflow, *res_P, *res_E;
    double* res_cum;           // [4][nres] sums of the step: inflow, outflow, actual evaporation, abstraction
    double *res_outflow_avg, *res_inflow_avg, *res_actevap_avg, *res_abs_avg, *res_store_prev, *bal_res;
    // ---- lateral snow transport ----
    int snow_lat;
    const int *sup_ptr, *sup_idx;  // [n+1], [E]: all upstream cells, in the reference's topological order
    double *snow_in, *snow_out;    // [n] (zeros without transport)
    // ---- 1-D floodplain (floodplain_1d), indexed by river cell ----
    int fp, fp_nd;
    const double* fp_depth;        // [fp_nd] flood depths, 0 first
    const double* fp_prof;         // [nriv][4][fp_nd] storage, width, flow area, wetted perimeter
    const double *fp_n, *fp_slope, *r_bfs, *r_bfd;  // Manning n, slope, bankfull storage and depth
    double *fp_storage, *fp_h, *fp_q, *fp_qin, *fp_exch, *fp_q_cum, *fp_qin_cum, *fp_q_avg, *fp_qin_avg;
    double* r_qch_avg;             // channel discharge average
    double *fp_rs, *fp_sprev;      // channel storage after the last sub-step; channel + floodplain at step start
};
// Dev is passed by value to every kernel: it must fit the kernel argument space
static_assert(sizeof(Dev) + 512 <= 4096, "Dev too large for a kernel argument");

struct RouteArgs {
    uint64_t* slot_ssf;  // [S_s][n]
    uint64_t* slot_ol;   // [S_l][n]
    uint64_t* slot_r;    // [S_r][nriv]
    unsigned long long* counters;  // [2] wave tickets, alternating per launch
    int* abort_flag;
    long spin_limit;  // SPIN_LIMIT (wfh_test_route_stall lowers it)
    int parity;
    int nchunks;
    // multi-step run (k_route<N, false, true>): steps of the launch, per chunk the number of its
    // steps completed, the forcing frames of the steps ([nsteps][n], node order) and optionally
    // each step's river discharge average ([nsteps][nriv], river order)
    int nsteps;
    int* chunk_done;
    const double *run_P, *run_E, *run_T;
    double* run_q_avg;
    int S_s, S_l, S_r;
    double dts_s[2], dts_l[2], dts_r[2];  // sub-step lengths: first (S-1) steps and the last one
};

// State of the adaptive sub-steps (one launch per sub-step and family), kept between launches.
struct AdArgs {
    double* szi;           // [n] subsurface water table depth
    double* scum;          // [6][n] subsurface sums (SsfCum order: qin, q, exf, qnet, tor, rech)
    double* ocum;          // [3][n] overland sums: q, qin, to_river
    double* rcum;          // [2][nriv] river sums (river-position order): q, qin
    const int* rp_r;       // [nriv] river index at each river position
    const int* rp_up_ptr;  // [nriv+1]
    const int* rp_up_idx;  // [Er] river positions of the upstream river cells, ascending river index
    const int* rp_down;    // [nriv] river position downstream, or -1
    const double* rp_up_res;  // [Er] 1 when the edge's upstream river cell is a reservoir outlet
    const int* rp_node;    // [nriv] node of the river cell at each river position
    const int *fp_sup_ptr, *fp_sup_idx;  // [nriv+1], [Er]: all upstream river positions, in the
                                         // order of the reference's river topological sort
    const int* chain_r;    // [nchunks_r] 1: a chain chunk (see run_chain)
    int nchunks_r;
};

// Order-statistic selection over 64-bit keys (the bit patterns of positive doubles order as
// the values): one 8-bit digit per pass from the top, then the next larger key.
struct SelCtl {
    unsigned long long count;      // eligible values
    unsigned long long rank;       // rank still to find among the keys matching prefix/mask
    unsigned long long prefix, mask;
    unsigned long long bin_count;  // keys equal to the selected one (after the last pass)
    unsigned long long next;       // smallest key above the selected one
    unsigned long long minkey;     // smallest key (minimum reduction)
    unsigned int hist[256];
};

// ============================================================================================
// K1: vertical column update (update_land_hydrology_model!, Wflow/src/sbm.jl:82-132)
// ============================================================================================
// The values of one cell's vertical update that the routing part of the step reads next.
template <int N>
struct VertOut {
    double us[N];
    double zi, recharge, sat_exc, runoff_l, inf_exc, aeow_l, nrr, aet, leak, canopy, snow_s, snow_w, store_prev, P;
};

// The vertical update of cell i (node order): reads the step's forcing and the cell's states,
// writes every state and flux of the vertical part back, and returns what routing reads.
template <int N, bool COH = false>
__device__ __forceinline__ void vertical_cell(const Dev& d, const int i, VertOut<N>& o) {
    const int n = d.n;
    // Every input of the cell is loaded here, in one block, unconditionally (parameters of
    // switched-off processes are allocated as zeros): loads spread over the branches below
    // would each be waited for at the branch's end, one memory round trip after another.
    const double P_raw = d.P[i], E_raw = d.E[i], T_raw = d.T[i];
    // Non-finite forcing: flagged for wfh_sync, and the column runs on finite stand-ins instead
    // (the step's results are reported unusable), so that no NaN or infinity reaches the
    // column's iterations (the Brooks-Corey sub-steps are counted from the water amounts).
    const bool bad = wfh_nonfinite3(P_raw, E_raw, T_raw);
    if (bad) __hip_atomic_store(d.nonfinite_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double P_in = bad ? 0.0 : P_raw, E_in = bad ? 0.0 : E_raw, T = bad ? 283.15 : T_raw;
    const uint8_t fl = d.nres > 0 ? d.flags[i] : 0;
    ColState<N> s;
    s.canopy = ldv<COH>(d.canopy + i);
    s.ss = ldv<COH>(d.snow_s + i);
    s.sw = ldv<COH>(d.snow_w + i);
    s.tsws = ldv<COH>(d.tsws + i);
    ColParams p;
    p.gap = ld(d.gap, i), p.smax = ld(d.smax, i), p.kc = ld(d.kc, i), p.er = ld(d.er, i);
    p.tti = ld(d.tti, i), p.tt = ld(d.tt, i), p.ttm = ld(d.ttm, i), p.cfmax = ld(d.cfmax, i);
    p.whc = ld(d.whc, i), p.wsoil = ld(d.wsoil, i), p.cfsoil = ld(d.cfsoil, i);
    const int r = d.river_of[i];
    const double wd_land = ldv<COH>(d.ol_h + i);
    const double r_h0 = ldv<COH>(d.r_h + (r >= 0 ? r : 0));
    p.rf = ld(d.rf, i), p.wf = ld(d.wf, i), p.sfrac = ld(d.sfrac, i);
    p.ts = ld(d.ts, i), p.tr = ld(d.tr, i), p.tfc = ld(d.tfc, i), p.dsoil = ld(d.d, i);
    p.kv0 = ld(d.kv0, i), p.fk = ld(d.f, i), p.zexp = ld(d.zexp, i);
    s.swd = ldv<COH>(d.swd + i), s.tsoil = ldv<COH>(d.tsoil + i);
    p.compf = ld(d.compf, i), p.icap_s = ld(d.icap_s, i), p.icap_c = ld(d.icap_c, i);
    p.rd = ld(d.rd, i), p.h3h = ld(d.h3h, i), p.h3l = ld(d.h3l, i);
    p.h1 = ld(d.h1, i), p.h2 = ld(d.h2, i), p.h4 = ld(d.h4, i), p.ah1 = ld(d.ah1, i), p.hb = ld(d.hb, i);
    p.rdp = ld(d.rdp, i), p.caph = ld(d.caph, i), p.capn = ld(d.capn, i), p.maxleak = ld(d.maxleak, i);
    // precipitation and potential evaporation over a reservoir's area go to the reservoir
    // (update_forcing!  io.jl:125-143)
    const double P = (fl & 8) ? 0.0 : P_in, E = (fl & 8) ? 0.0 : E_in;
    ColLayers<N> L;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        s.us[k] = ldv<COH>(d.ustore + (k * n + i));
        L.kvf[k] = ld(d.kvfrac, k, i);
        L.cbc[k] = ld(d.c, k, i);
        L.rfr[k] = ld(d.rootfrac, k, i);
    }
    const ColConfig cfg{d.dt, d.gash != 0, d.snow != 0, d.infred != 0, d.balance != 0, d.cfg_cum, d.cfg_thick};
    ColFlux f;
    sbm_column<N>(P, E, T, wd_land, r >= 0 ? r_h0 : 0.0, p, L, cfg, s, f);

    // ---- write back ----
    if (d.balance) stv<COH>(d.store_prev + (i), f.store_prev);
    stv<COH>(d.canopy + (i), s.canopy);
    if (d.snow) {
        stv<COH>(d.swe + (i), f.swe);
        stv<COH>(d.snow_s + (i), s.ss);
        stv<COH>(d.snow_w + (i), s.sw);
    }
    stv<COH>(d.tsoil + (i), s.tsoil);
#pragma unroll
    for (int k = 0; k < N; ++k) stv<COH>(d.ustore + (k * n + i), s.us[k]);
    stv<COH>(d.zi + (i), f.zi);
    stv<COH>(d.interception + (i), f.interception);
    stv<COH>(d.transfer + (i), f.transfer);
    stv<COH>(d.recharge + (i), f.recharge);
    stv<COH>(d.act_inf + (i), f.act_inf);
    stv<COH>(d.inf_excess + (i), f.inf_exc);
    stv<COH>(d.sat_excess + (i), f.sat_exc);
    stv<COH>(d.transpiration + (i), f.transp);
    stv<COH>(d.soil_evap + (i), f.sevap);
    stv<COH>(d.capflux + (i), f.capflux);
    stv<COH>(d.leakage + (i), f.leak);
    stv<COH>(d.aet + (i), f.aet);
    stv<COH>(d.net_runoff_river + (i), f.nrr);
    stv<COH>(d.runoff_land + (i), f.rl);
    stv<COH>(d.aeow_l + (i), f.aol);
#pragma unroll
    for (int k = 0; k < N; ++k) o.us[k] = s.us[k];
    o.zi = f.zi;
    o.recharge = f.recharge;
    o.sat_exc = f.sat_exc;
    o.runoff_l = f.rl;
    o.inf_exc = f.inf_exc;
    o.aeow_l = f.aol;
    o.nrr = f.nrr;
    o.aet = f.aet;
    o.leak = f.leak;
    o.canopy = s.canopy;
    o.snow_s = s.ss;
    o.snow_w = s.sw;
    o.store_prev = f.store_prev;
    o.P = f.P;
}

template <int N>
__global__ void __launch_bounds__(256) k_vertical(Dev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    VertOut<N> o;
    vertical_cell<N>(d, i, o);
}

// ============================================================================================
// K2: routing dataflow kernel
// ============================================================================================
__device__ __forceinline__ void slot_put(uint64_t* s, double v) {
    __hip_atomic_store(s, wfh_bits(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool slot_take(uint64_t* s, double& v) {
    const uint64_t b = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (b == SLOT_EMPTY) return false;
    __hip_atomic_store(s, SLOT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = wfh_from_bits(b);
    return true;
}

// Folds of the upstream discharge in the reference's left-fold order (sum_at, utils.jl:472-477).
//   FOLD_SSF: acc1 = sum q (1 - f)  (q_in),   acc2 = sum q f  (to_river)   lateral_subsurface_flow.jl:228-240
//   FOLD_OL:  acc1 = sum q f (to_river),      acc2 = sum q (1 - f) (qin)    surface_kinwave.jl:317-326
//   FOLD_RIV: acc1 = sum q                                                  surface_kinwave.jl:520
//   FOLD_SNOW: acc1 = own material + material passed on by each upstream cell, in the reference's
//              topological order of the upstream cells (accucapacityflux!  routing/utils.jl:82-108)
//   FOLD_RIV_RES (a model with reservoirs): acc1 = sum q over the non-reservoir upstream cells,
//             acc2 = outflow of an upstream reservoir (f = 1 on that edge)  surface_kinwave.jl:478,520
enum { FOLD_SSF = 0, FOLD_OL = 1, FOLD_RIV = 2, FOLD_RIV_RES = 3, FOLD_SNOW = 4 };

template <int FOLD>
__device__ __forceinline__ void fold_in(double& acc1, double& acc2, double q, double f) {
    if (FOLD == FOLD_SSF) {
        acc1 = acc1 + q * (1.0 - f);
        acc2 = acc2 + q * f;
    } else if (FOLD == FOLD_OL) {
        acc1 = acc1 + q * f;
        acc2 = acc2 + q * (1.0 - f);
    } else if (FOLD == FOLD_RIV_RES && f != 0.0) {
        acc2 = q;  // a reservoir's outflow is assigned to the downstream inflow (surface_kinwave.jl:478)
    } else {
        acc1 = acc1 + q;
    }
}

// One family of routing sub-steps (all subsurface, all overland or all river sub-steps of the
// model step) for the 64 cells of a wave. Every lane runs the same code; the wave advances a
// diagonal wavefront over (cell, sub-step): a lane computes sub-step k once all upstream cells
// have produced their sub-step-k outflow and, for multi-step families, its in-wave downstream
// cell has computed sub-step k-1 (so the outflow it is about to replace has been used).
//
// In-wave exchange goes through a 64-entry table in LDS: each round every lane writes its
// (next sub-step, last outflow) pair, then reads the entries of its in-wave upstream cells with
// independent loads and keeps the values that belong to its current sub-step. Outflow to a
// consumer in another wave goes through that consumer's global slot (one per edge and
// sub-step); such slots are polled once per round and picked up in the next round, so the poll
// latency overlaps the round's work. The upstream outflow is folded in ascending edge order,
// the reference's sum_at order, as far as it has arrived. Returns false when the wave has to
// stop (a wait exceeded its limit, or another wave raised the abort flag).
template <int FOLD, class Compute>
__device__ __forceinline__ bool run_family(const RouteArgs& a, const int lane, const int c, const bool part,
                                           const int S, const int* __restrict__ up_node,
                                           const int* __restrict__ up_slot, const double* __restrict__ up_f,
                                           const int e0, const int e1, uint64_t* slots, const size_t stride,
                                           const int my_slot, const bool pub_cross, const int cons_lane,
                                           int* lds_k, double* lds_q, Compute& compute, const double acc0 = 0.0) {
    // The first PF upstream edges are buffered in registers; further edges (a cell with more
    // than PF upstream cells) are taken one per round in order, after the buffered ones.
    constexpr int PF = 4;
    int k = part ? 0 : S;
    int e = e0;
    double acc1 = acc0, acc2 = 0.0, qv = 0.0;  // acc0: start of the first sub-step's fold
    const int npf = (e1 - e0) < PF ? (e1 - e0) : PF;
    int en[PF], es[PF];
    double ef[PF], buf[PF];
    uint64_t v[PF];
    bool pol[PF];
    // Every lane loads PF entries (unused ones at a valid index); the empty asm takes them all as
    // operands, so the compiler issues the loads together and waits once, instead of moving each
    // into the per-lane branch that uses it and waiting for them one by one.
    int un[PF], us[PF];
    double uf[PF];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const int ej = j < npf ? e0 + j : (npf > 0 ? e0 : 0);
        un[j] = up_node[ej];
        us[j] = up_slot[ej];
        uf[j] = (FOLD != FOLD_RIV && FOLD != FOLD_SNOW) ? up_f[ej] : 0.0;
    }
    static_assert(PF == 4, "the operand list below names four edges");
    asm volatile("" : "+v"(un[0]), "+v"(un[1]), "+v"(un[2]), "+v"(un[3]), "+v"(us[0]), "+v"(us[1]), "+v"(us[2]),
                 "+v"(us[3]), "+v"(uf[0]), "+v"(uf[1]), "+v"(uf[2]), "+v"(uf[3]));
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const bool use = j < npf;
        en[j] = use ? un[j] : c * WAVE;
        es[j] = use ? us[j] : 0;
        ef[j] = use ? uf[j] : 0.0;
        buf[j] = 0.0;
        v[j] = SLOT_EMPTY;
        pol[j] = false;
    }
    unsigned have = 0;  // bit j: buf[j] holds edge e0 + j's outflow for sub-step k
    long spin = 0;
    while (__any(k < S)) {
        bool act = k < S;
        // (1) publish this lane's state to the wave's table
        lds_k[lane] = k;
        lds_q[lane] = qv;
        __builtin_amdgcn_wave_barrier();
        // (2) pick up last round's polls; gather the in-wave upstream outflow that is ready. The
        // table entries of all PF producers are read unconditionally (unused edges point at
        // lane 0), so the reads overlap and one wait covers them.
        int tk[PF];
        double tq[PF];
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            tk[j] = lds_k[en[j] & (WAVE - 1)];
            tq[j] = lds_q[en[j] & (WAVE - 1)];
        }
        bool got = false;
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const bool from_slot = pol[j] && v[j] != SLOT_EMPTY;
            if (from_slot)
                __hip_atomic_store(slots + (size_t)k * stride + es[j], SLOT_EMPTY, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            const bool from_wave = act && j < npf && (en[j] >> 6) == c && !(have & (1u << j)) && tk[j] == k + 1;
            buf[j] = from_slot ? wfh_from_bits(v[j]) : (from_wave ? tq[j] : buf[j]);
            have |= (from_slot || from_wave) ? (1u << j) : 0u;
            got = got || from_slot || from_wave;
        }
        // (3) issue this round's polls for the cross-wave edges still missing
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            pol[j] = act && j < npf && (en[j] >> 6) != c && !(have & (1u << j));
            v[j] = SLOT_EMPTY;
            if (pol[j])
                v[j] = __hip_atomic_load(slots + (size_t)k * stride + es[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // (4) fold the buffered prefix in edge order, then at most one further edge
        if (act) {
#pragma unroll
            for (int j = 0; j < PF; ++j)
                if (e == e0 + j && (have & (1u << j))) {
                    fold_in<FOLD>(acc1, acc2, buf[j], ef[j]);
                    ++e;
                }
            if (e >= e0 + PF && e < e1) {
                const int u = up_node[e];
                const double f = (FOLD != FOLD_RIV && FOLD != FOLD_SNOW) ? up_f[e] : 0.0;
                double q = 0.0;
                bool took = false;
                if ((u >> 6) == c) {
                    if (lds_k[u & (WAVE - 1)] == k + 1) {
                        q = lds_q[u & (WAVE - 1)];
                        took = true;
                    }
                } else {
                    took = slot_take(slots + (size_t)k * stride + up_slot[e], q);
                }
                if (took) {
                    fold_in<FOLD>(acc1, acc2, q, f);
                    ++e;
                    got = true;
                }
            }
        }
        // (5) compute the ready lanes. Lanes whose in-wave upstream cells complete here are
        // served within the same round: the producers republish, the consumers gather, fold and
        // compute in turn, so a chain inside the wave advances several cells per round. Every
        // pass computes at least one sub-step, so the passes end.
        bool any_done = false;
        for (int pass = 0; pass < WAVE * S; ++pass) {
            const int kcons = lds_k[cons_lane & (WAVE - 1)];
            const bool bp = S == 1 || cons_lane < 0 || kcons >= k;
            const bool can = act && e == e1 && bp;
            __builtin_amdgcn_wave_barrier();
            if (!__any(can)) break;
            any_done = true;
            if (can) {
                qv = compute(k, acc1, acc2);
                if (pub_cross) slot_put(slots + (size_t)k * stride + my_slot, qv);
                ++k;
                e = e0;
                acc1 = 0.0;
                acc2 = 0.0;
                have = 0;
            }
            act = k < S;
            lds_k[lane] = k;
            lds_q[lane] = qv;
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                tk[j] = lds_k[en[j] & (WAVE - 1)];
                tq[j] = lds_q[en[j] & (WAVE - 1)];
            }
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                const bool from_wave =
                    act && j < npf && (en[j] >> 6) == c && !(have & (1u << j)) && tk[j] == k + 1;
                buf[j] = from_wave ? tq[j] : buf[j];
                have |= from_wave ? (1u << j) : 0u;
            }
            if (act) {
#pragma unroll
                for (int j = 0; j < PF; ++j)
                    if (e == e0 + j && (have & (1u << j))) {
                        fold_in<FOLD>(acc1, acc2, buf[j], ef[j]);
                        ++e;
                    }
            }
        }
        if (any_done || __any(got)) {
            spin = 0;
        } else {
            __builtin_amdgcn_s_sleep(1);
            ++spin;
            if ((spin & 255) == 0 && __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                return false;
            if (spin > a.spin_limit) {
                __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    return true;
}

// Multi-step run: waits until chunk `ca` and chunk `cb` (per lane) have completed `s` steps. Both
// hold lower tickets than the waiting wave (earlier steps), so the wait ends; it is bounded like
// every other wait, and the loads after it are issued only once it has ended.
__device__ __forceinline__ bool wait_chunks(const RouteArgs& a, const int ca, const int cb, const int s) {
    long spin = 0;
    for (;;) {
        const int va = __hip_atomic_load(a.chunk_done + ca, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int vb = __hip_atomic_load(a.chunk_done + cb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!__any(va < s || vb < s)) break;
        __builtin_amdgcn_s_sleep(1);
        ++spin;
        if ((spin & 255) == 0 && __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        if (spin > a.spin_limit) {
            __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    asm volatile("" ::: "memory");
    return true;
}

__device__ __forceinline__ double readlane_f64(const double v, const int l) {
    const uint64_t b = wfh_bits(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return wfh_from_bits(((uint64_t)hi << 32) | lo);
}

// A family of one sub-step over a chain chunk: all 64 lanes valid, at most 4 upstream cells
// each, and lane l's only upstream cell inside the chunk is lane l-1 (checked on the host). The
// lanes first take their inflow from other chunks; then the lanes fold and solve one after the
// other, each handing its outflow to the next through a register broadcast. The folds, in edge
// order, and the solves are those of run_family, so the results are the same; what goes is the
// per-hop exchange through LDS and the wave votes of run_family's in-round passes.
template <int FOLD, class Compute>
__device__ __forceinline__ bool run_chain(const RouteArgs& a, const int lane, const int c,
                                          const int* __restrict__ up_node, const int* __restrict__ up_slot,
                                          const double* __restrict__ up_f, const int e0, const int e1,
                                          uint64_t* slots, const int my_slot, const bool pub_cross, Compute& compute,
                                          const double acc0 = 0.0) {
    constexpr int PF = 4;
    const int ne = e1 - e0;
    int en[PF], es[PF];
    double ef[PF], buf[PF];
    bool have[PF];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const bool use = j < ne;
        const int ej = use ? e0 + j : (ne > 0 ? e0 : 0);  // (edge arrays hold at least one entry)
        en[j] = use ? up_node[ej] : c * WAVE;
        es[j] = use ? up_slot[ej] : 0;
        ef[j] = (use && FOLD != FOLD_RIV && FOLD != FOLD_SNOW) ? up_f[ej] : 0.0;
        buf[j] = 0.0;
        have[j] = !use || (en[j] >> 6) == c;
    }
    long spin = 0;
    for (;;) {
        bool missing = false;
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            if (!have[j]) have[j] = slot_take(slots + es[j], buf[j]);
            missing = missing || !have[j];
        }
        if (!__any(missing)) break;
        __builtin_amdgcn_s_sleep(1);
        ++spin;
        if ((spin & 255) == 0 && __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        if (spin > a.spin_limit) {
            __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    double prev = 0.0, qv = 0.0;
    for (int l = 0; l < WAVE; ++l) {
        if (lane == l) {
            double acc1 = acc0, acc2 = 0.0;
#pragma unroll
            for (int j = 0; j < PF; ++j)
                if (j < ne) fold_in<FOLD>(acc1, acc2, (en[j] >> 6) == c ? prev : buf[j], ef[j]);
            qv = compute(0, acc1, acc2);
            if (pub_cross) slot_put(slots + my_slot, qv);
        }
        prev = readlane_f64(qv, l);
    }
    return true;
}

__device__ __forceinline__ double shfl_up1_f64(const double v) {
    const uint64_t b = wfh_bits(v);
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)b, 1u), hi = (uint32_t)__shfl_up((int)(uint32_t)(b >> 32), 1u);
    return wfh_from_bits(((uint64_t)hi << 32) | lo);
}

// A family of S <= 4 sub-steps over a chain chunk (as run_chain) as a diagonal wavefront: in
// round r lane l computes sub-step r - l, with lane l-1's outflow of the same sub-step (computed
// in round r-1) shifted over from the neighbouring lane, and the inflow from other chunks of all
// S sub-steps taken beforehand. Each lane computes its sub-steps in order, and its neighbour has
// read an outflow before it is replaced, so the results are run_family's.
template <int FOLD, class Compute>
__device__ __forceinline__ bool run_chain_diag(const RouteArgs& a, const int lane, const int c, const int S,
                                               const int* __restrict__ up_node, const int* __restrict__ up_slot,
                                               const double* __restrict__ up_f, const int e0, const int e1,
                                               uint64_t* slots, const size_t stride, const int my_slot,
                                               const bool pub_cross, Compute& compute) {
    constexpr int PF = 4, SM = 4;
    const int ne = e1 - e0;
    int en[PF], es[PF];
    double ef[PF], buf[PF][SM];
    bool have[PF][SM];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const bool use = j < ne;
        const int ej = use ? e0 + j : (ne > 0 ? e0 : 0);
        en[j] = use ? up_node[ej] : c * WAVE;
        es[j] = use ? up_slot[ej] : 0;
        ef[j] = (use && FOLD != FOLD_RIV && FOLD != FOLD_SNOW) ? up_f[ej] : 0.0;
#pragma unroll
        for (int k = 0; k < SM; ++k) {
            buf[j][k] = 0.0;
            have[j][k] = !use || (en[j] >> 6) == c || k >= S;
        }
    }
    long spin = 0;
    for (;;) {
        bool missing = false;
#pragma unroll
        for (int j = 0; j < PF; ++j)
#pragma unroll
            for (int k = 0; k < SM; ++k) {
                if (!have[j][k]) have[j][k] = slot_take(slots + (size_t)k * stride + es[j], buf[j][k]);
                missing = missing || !have[j][k];
            }
        if (!__any(missing)) break;
        __builtin_amdgcn_s_sleep(1);
        ++spin;
        if ((spin & 255) == 0 && __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        if (spin > a.spin_limit) {
            __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    double qv = 0.0;
    for (int r = 0; r < WAVE + S - 1; ++r) {
        const double from_prev = shfl_up1_f64(qv);
        const int k = r - lane;
        if (k >= 0 && k < S) {
            double acc1 = 0.0, acc2 = 0.0;
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                double bk = buf[j][0];
#pragma unroll
                for (int kk = 1; kk < SM; ++kk) bk = k == kk ? buf[j][kk] : bk;
                if (j < ne) fold_in<FOLD>(acc1, acc2, (en[j] >> 6) == c ? from_prev : bk, ef[j]);
            }
            qv = compute(k, acc1, acc2);
            if (pub_cross) slot_put(slots + (size_t)k * stride + my_slot, qv);
        }
    }
    return true;
}

// interpolate_linear(x, xp, fp)  reservoir.jl:364-377 for a table of n rows whose xp is
// non-decreasing (checked at create). The minimum and maximum of xp and fp come precomputed; the
// last row with xp <= x is found by bisection (xp[lo] <= x < xp[hi] throughout), which for a
// non-decreasing xp is the row the reference's findall takes. The interpolation keeps the
// reference's operation order.
__device__ __forceinline__ double interp_table(const double x, const double* xp, const double* fp, const int fs,
                                               const int n, const double xmin, const double xmax,
                                               const double fmin, const double fmax) {
    if (wfh_isnan(x)) return x;  // (the reference raises: findall finds no row)
    if (x <= xmin) return fmin;
    if (x >= xmax) return fmax;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= x) lo = mid;
        else hi = mid;
    }
    const double x1 = xp[lo], x2 = xp[lo + 1], f1 = fp[lo * fs], f2 = fp[(lo + 1) * fs];
    return f1 * (1.0 - (x - x1) / (x2 - x1)) + f2 * (x - x1) / (x2 - x1);
}

// update_reservoir_model! at the river cell that is reservoir ri's outlet (node p), for its
// river discharge q_v of a sub-step of length dts: external inflow / abstraction
// (surface_kinwave.jl:441-489), the outflow of the rating-curve, free-weir (no linked lower
// reservoir), modified Puls, simple or observed-outflow update and the linear or interpolated
// storage curve (reservoir.jl:389-635). The reservoir's state
// lives in memory and only this lane touches it. first / last: the step's first / last river
// sub-step (set_reservoir_vars!, average_reservoir_vars!, the reservoir balance). Returns the
// outflow, which becomes the downstream river cell's inflow.
// LINK (linked free weirs, one launch per sub-step): a reservoir with a linked free weir upstream
// first takes that weir's reverse flow of the sub-step from its slot and applies it to its
// storage, water level and cumulative outflow, as the weir's update does in the reference
// (reservoir.jl:539-550); the weir reads the lower reservoir's level of the previous sub-step.
template <bool LINK = false>
__device__ __forceinline__ double reservoir_step(const Dev& d, const int ri, const int p, const double q_v,
                                              const double dts, const bool first, const bool last) {
    const int nr = d.nres;
    const double area = d.res_area[ri];
    double storage0 = d.res_storage[ri], wl0 = d.res_wl[ri];
    double cin = first ? 0.0 : d.res_cum[ri], cout = first ? 0.0 : d.res_cum[nr + ri];
    double cev = first ? 0.0 : d.res_cum[2 * nr + ri], cab = first ? 0.0 : d.res_cum[3 * nr + ri];
    if (first && d.balance) d.res_store_prev[ri] = storage0;
    if (LINK && d.res_upper[ri] >= 0) {
        // the weir is upstream, so it has computed; its slot store becomes visible shortly
        double xfer;
        long spin = 0;
        while (!slot_take(d.res_xfer + ri, xfer)) {
            if (++spin > SPIN_LIMIT ||
                ((spin & 255) == 0 && __hip_atomic_load(d.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(d.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return 0.0 / 0.0;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (xfer != INFINITY) {
            const double lower_res_storage = storage0 + xfer * dts;
            if (d.res_sct[ri] == 2) {
                const int r0 = d.res_sh_ptr[ri];
                const double* ext = d.res_shx + (size_t)ri * 4;
                wl0 = interp_table(lower_res_storage, d.res_sh_s + r0, d.res_sh_h + r0, 1, d.res_sh_ptr[ri + 1] - r0,
                                   ext[0], ext[1], ext[2], ext[3]);
            } else {
                wl0 = wl0 + (lower_res_storage - storage0) / area;
            }
            cout -= xfer * dts;
            storage0 = lower_res_storage;
        }
    }
    // update_inflow!  surface_kinwave.jl:772-799: the outlet cell's overland and subsurface flow
    const double io = d.ol_q_avg[p], is = d.ssf_q_avg[p];
    const double inflow_ext = d.res_ext[ri];
    double inflow;
    if (inflow_ext < 0.0) {
        const double abstraction = wfh_min(-inflow_ext, (storage0 / dts) * 0.98);
        cab += abstraction * dts;
        inflow = -abstraction;
    } else {
        inflow = inflow_ext;
    }
    const double net_inflow = q_v + io + is + inflow;
    const double precipitation = d.res_P[ri] * area;
    const double available_storage = storage0 + (net_inflow + precipitation) * dts;
    const double potential_evaporation = d.res_E[ri] * area;
    const double evaporation = wfh_min(available_storage / dts, potential_evaporation);
    const double obs = d.res_obs[ri], maxsto = d.res_maxsto[ri];
    double outflow, storage;
    double xfer_out = INFINITY;  // reverse flow into a linked lower reservoir (+inf: none)
    if (!wfh_isnan(obs)) {  // update_reservoir_outflow_obs  reservoir.jl:555-575
        const double storage_input = wfh_max(storage0 / dts + precipitation - evaporation + net_inflow, 0.0);
        outflow = wfh_min(obs, storage_input);
        storage = (storage_input - outflow) * dts;
        if (!wfh_isnan(maxsto)) {
            const double overflow = wfh_max(0.0, (storage - maxsto) / dts);
            storage -= overflow * dts;
            outflow += overflow;
        }
    } else if (d.res_type[ri] == 1) {  // update_reservoir_hq  reservoir.jl:457-483
        const double storage_input = storage0 / dts + precipitation - evaporation + net_inflow;
        const int r0 = d.res_hq_ptr[ri], col = d.res_hq_col;
        const double* ext = d.res_hq_ext + (size_t)ri * 732;
        outflow = interp_table(wl0, d.res_hq_h + r0, d.res_hq_q + (size_t)r0 * 365 + col, 365,
                               d.res_hq_ptr[ri + 1] - r0, ext[0], ext[1], ext[2 + 2 * col], ext[3 + 2 * col]);
        outflow = wfh_min(outflow, storage_input);
        storage = (storage_input - outflow) * dts;
        const double overflow = wfh_max(0.0, (storage - maxsto) / dts);
        storage = wfh_min(storage, maxsto);
        outflow += overflow;
    } else if (d.res_type[ri] == 2) {  // update_reservoir_free_weir  reservoir.jl:485-553
        const double storage_input = wfh_max(storage0 / dts + precipitation - evaporation + net_inflow, 0.0);
        const double thr = d.res_thr[ri];
        const int lo = LINK ? d.res_lower[ri] : -1;  // (ri's own lower reservoir)
        const double wl_lo = lo >= 0 ? d.res_wl[lo] : 0.0;
        const double diff_wl = lo >= 0 ? wl0 - wl_lo : 0.0;
        outflow = 0.0;
        if (diff_wl >= 0.0) {
            if (wl0 > thr) {
                const double dh = wl0 - thr;
                outflow = d.res_rcc[ri] * wfh_pow(dh, d.res_rce[ri]);
                const double maxflow = dh * area / dts;
                outflow = wfh_min(outflow, maxflow);
            }
        } else if (wl_lo > thr) {  // reverse flow from the linked lower reservoir
            const double dh = wl_lo - thr;
            outflow = -d.res_rcc[ri] * wfh_pow(dh, d.res_rce[ri]);
            const double maxflow = dh * d.res_area[lo] / dts;
            outflow = wfh_max(outflow, -maxflow);
        }
        storage = (storage_input - outflow) * dts;
        if (diff_wl < 0.0) xfer_out = outflow;
    } else if (d.res_type[ri] == 3) {  // update_reservoir_modified_puls  reservoir.jl:427-456
        const double res_factor = area / (dts * sqrt(d.res_rcc[ri]));
        const double si_factor = storage0 / dts + precipitation - evaporation + net_inflow;
        const double si_factor_adj = si_factor - area * d.res_thr[ri] / dts;
        outflow = 0.0;
        if (si_factor_adj > 0.0) {
            const double qst = -res_factor + sqrt((res_factor * res_factor + 4.0 * si_factor_adj));
            if (qst > 0.0) outflow = 0.25 * (qst * qst);
        }
        outflow = wfh_min(outflow, si_factor);
        storage = (si_factor - outflow) * dts;
    } else {  // update_reservoir_simple  reservoir.jl:389-421
        storage = storage0 + (net_inflow + precipitation - evaporation) * dts;
        storage = wfh_max(storage, 0.0);
        const double fill_fraction = storage / maxsto;
        const double fac = 1.0 / (1.0 + wfh_exp(-30.0 * (fill_fraction - d.res_tmf[ri])));  // scurve
        const double demand_release = wfh_min(fac * d.res_demand[ri], storage / dts);
        storage -= demand_release * dts;
        const double release_wanted = wfh_max(0.0, (storage - maxsto * d.res_tff[ri]) / dts);
        const double overflow_q = wfh_max(0.0, (storage - maxsto) / dts);
        const double release_realized = wfh_min(release_wanted, overflow_q + d.res_maxrel[ri] - demand_release);
        storage -= release_realized * dts;
        outflow = release_realized + demand_release;
    }
    // every sub-step a linked weir hands its lower reservoir a value, whichever update ran (an
    // observed outflow takes the place of the weir's)
    if (LINK && d.res_lower[ri] >= 0) slot_put(d.res_xfer + d.res_lower[ri], xfer_out);
    if (d.res_sct[ri] == 2) {  // interpolate_linear(storage, S, H)  reservoir.jl:613-621
        const int r0 = d.res_sh_ptr[ri];
        const double* ext = d.res_shx + (size_t)ri * 4;
        d.res_wl[ri] = interp_table(storage, d.res_sh_s + r0, d.res_sh_h + r0, 1, d.res_sh_ptr[ri + 1] - r0, ext[0],
                                    ext[1], ext[2], ext[3]);
    } else {
        d.res_wl[ri] = wl0 + (storage - storage0) / area;
    }
    d.res_storage[ri] = storage;
    d.res_outflow[ri] = outflow;
    cin += net_inflow * dts;
    cout += outflow * dts;
    cev += evaporation / area * dts;
    d.res_cum[ri] = cin;
    d.res_cum[nr + ri] = cout;
    d.res_cum[2 * nr + ri] = cev;
    d.res_cum[3 * nr + ri] = cab;
    if (last) {
        const double dt = d.dt;
        // average_reservoir_vars!  surface_kinwave.jl:244-258
        d.res_outflow_avg[ri] = cout / dt;
        d.res_inflow_avg[ri] = cin / dt;
        d.res_abs_avg[ri] = cab / dt;
        if (d.balance) {  // compute_flow_balance!(::ReservoirModel)  mass_balance.jl:261-283
            const double ev_avg = cev / dt;
            d.res_actevap_avg[ri] = ev_avg;
            const double total_in = cin / dt + d.res_P[ri] * area;
            const double total_out = cout / dt + ev_avg * area;
            d.bal_res[ri] = total_in - total_out - (storage - d.res_store_prev[ri]) / dt;
        }
    }
    return outflow;
}

// Mean precipitation and potential evaporation over each reservoir's area (update_forcing!
// io.jl:125-143): a left fold over the area's cells in ascending land index, divided by their
// count. One thread per reservoir.
__global__ void k_res_forcing(Dev d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.nres) return;
    double sp = 0.0, se = 0.0;
    const int e0 = d.res_cov_ptr[k], e1 = d.res_cov_ptr[k + 1];
    for (int e = e0; e < e1; ++e) {
        const int q = d.res_cov_idx[e];
        sp += d.P[q];
        se += d.E[q];
    }
    d.res_P[k] = sp / (double)(e1 - e0);
    d.res_E[k] = se / (double)(e1 - e0);
}

// Per-cell inputs of the closing subsurface bookkeeping of a step: soil parameters and the
// vertical update's outputs that enter the soil storage update and the balances.
struct SsfCellIn {
    double ts, tr, rd, dsoil, area, sy;
    double sat_exc, runoff_l, inf_exc, aeow, nrr;
    double P_in, aet, leak, canopy, snow_s, snow_w, store_prev;
    double snow_in, snow_out;  // lateral snow transport (0 without it)
};
// Sums over the subsurface sub-steps of flux x sub-step length.
struct SsfCum {
    double qin, q, exf, qnet, tor, rech;
};
struct SsfEnd {
    double usd, swd, s_tor_avg, o_inwater;
};

// average_flux_vars! (groundwater.jl:630-636), update_soil_water_storage! (soil.jl:1294-1392),
// the subsurface and land balances (mass_balance.jl:215-253, 470-490) and the overland lateral
// inflow (surface_kinwave.jl:740-766) of cell p after its last subsurface sub-step.
template <int N, bool COH = false>
__device__ __forceinline__ SsfEnd ssf_finish(const Dev& d, const int p, const Column<N>& col, const SsfCellIn& c,
                                             const SsfCum& cum, const double sq, const double szi) {
    const int n = d.n;
    const double dt = d.dt;
    const double ts = c.ts, tr = c.tr, rd = c.rd, dsoil = c.dsoil, area = c.area;
    SsfEnd o;
    const double s_exf_avg = cum.exf / dt;
    const double s_qin_avg = cum.qin / dt;
    const double s_q_avg = cum.q / dt;
    const double s_qnet_avg = cum.qnet / dt;
    const double rech_avg = cum.rech / dt;
    o.s_tor_avg = cum.tor / dt;
    stv<COH>(d.ssf_q + (p), sq);
    stv<COH>(d.ssf_qin_avg + (p), s_qin_avg);
    stv<COH>(d.ssf_q_avg + (p), s_q_avg);
    stv<COH>(d.ssf_tor_avg + (p), o.s_tor_avg);
    stv<COH>(d.ssf_storage + (p), c.sy * (dsoil - szi) * area);
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
    const double sbm_runoff = wfh_max(0.0, s_exf_avg + c.sat_exc + c.runoff_l + c.inf_exc);
    int nl = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) nl += !wfh_isnan(col.act[k]);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double v = NaN_;
        if (k < nl) {
            if (k < col.n_unsat)
                v = (col.ustore[k] + (col.act[k] - col.thick[k]) * (ts - tr)) / col.act[k] + tr;
            else
                v = ts;
        }
        stv<COH>(d.vwc + (k * n + p), v);
    }
    double rs_unsat = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k < col.n_unsat) rs_unsat += wfh_min(1.0, (wfh_max(0.0, rd - col.cum[k]) / col.thick[k])) * col.ustore[k];
    const double rs_sat = wfh_max(0.0, rd - col.zi) * (ts - tr);
    const double rzs = rs_sat + rs_unsat;
    const double swd = (dsoil - col.zi) * (ts - tr);
    stv<COH>(d.ustoredepth + (p), usd);
    stv<COH>(d.swd + (p), swd);
    stv<COH>(d.tsws + (p), swd + usd);
    stv<COH>(d.exfilt + (p), s_exf_avg);
    stv<COH>(d.runoff + (p), sbm_runoff);
    stv<COH>(d.rzs + (p), rzs);
    stv<COH>(d.vwc_root + (p), rzs / rd + tr);
    const double net_runoff = sbm_runoff - c.aeow;
    stv<COH>(d.net_runoff + (p), net_runoff);
    stv<COH>(d.zi + (p), col.zi);
#pragma unroll
    for (int k = 0; k < N; ++k) stv<COH>(d.ustore + (k * n + p), col.ustore[k]);
    if (d.balance) {
        double fin = 0.0, fout = 0.0;  // mass_balance.jl:470-490
        if (rech_avg > 0.0)
            fin += rech_avg;
        else
            fout -= rech_avg;
        stv<COH>(d.bal_ssf + (p), (s_qin_avg + fin) - (s_q_avg + fout) - s_qnet_avg);
        // land balance  mass_balance.jl:215-253
        const double t_in = s_qin_avg / area + (c.P_in + 0.0) + c.snow_in + 0.0 / area;
        const double vout = net_runoff + c.aet + c.nrr + c.leak + 0.0;
        const double t_out = s_q_avg / area + vout + c.snow_out + 0.0 / area;
        const double snow = d.snow ? c.snow_s + c.snow_w : 0.0 + 0.0;
        const double storage = (swd + usd) + c.canopy + snow + 0.0 * 0.0 + 0.0;
        const double rate = (storage - c.store_prev) / dt;
        const double err = t_in - t_out - rate;
        const double avg = (t_in + t_out) / 2.0;
        stv<COH>(d.bal_land + (p), err);
        stv<COH>(d.bal_land_rel + (p), avg == 0.0 ? 0.0 : err / avg);
    }
    // overland lateral inflow  surface_kinwave.jl:740-766
    o.o_inwater = (net_runoff + 0.0) * area + 0.0;
    stv<COH>(d.ol_inwater + (p), o.o_inwater);
    o.usd = usd;
    o.swd = swd;
    return o;
}

// average_flow_vars! of the overland model, its balance (mass_balance.jl:382-399) and the river
// lateral inflow of a river cell (surface_kinwave.jl:710-734) after cell p's last overland
// sub-step; returns the river lateral inflow (0 for a cell without river).
template <bool COH = false>
__device__ __forceinline__ double ol_finish(const Dev& d, const int p, const int r, const double oq, const double oh,
                                            const double o_storage, const double oqin, const double o_q_cum,
                                            const double o_tor_cum, const double o_qin_cum, const double o_inwater,
                                            const double o_store_prev, const double s_tor_avg, const double nrr,
                                            const double area, const double rarea) {
    const double dt = d.dt;
    const double q_avg = o_q_cum / dt;
    const double o_tor_avg = o_tor_cum / dt;
    const double qin_avg = o_qin_cum / dt;
    stv<COH>(d.ol_q + (p), oq);
    stv<COH>(d.ol_h + (p), oh);
    stv<COH>(d.ol_storage + (p), o_storage);
    stv<COH>(d.ol_qin + (p), oqin);
    stv<COH>(d.ol_q_avg + (p), q_avg);
    stv<COH>(d.ol_tor_avg + (p), o_tor_avg);
    stv<COH>(d.ol_qin_avg + (p), qin_avg);
    if (d.balance) stv<COH>(d.bal_ol + (p), (o_inwater + qin_avg) - q_avg - (o_storage - o_store_prev) / dt);
    double r_inwater = 0.0;
    if (r >= 0) {
        r_inwater = (s_tor_avg + o_tor_avg + nrr * area) + 0.0 * rarea;
        stv<COH>(d.r_inwater + (r), r_inwater);
    }
    return r_inwater;
}

// average_flow_vars! of the river model and its balance (mass_balance.jl:290-313) after river
// cell r's last sub-step.
template <bool COH = false>
__device__ __forceinline__ void riv_finish(const Dev& d, const int r, const double rq, const double rh,
                                           const double r_storage, const double rqin, const double r_q_cum,
                                           const double r_qin_cum, const double r_inwater, const double r_store_prev) {
    const double dt = d.dt;
    const double q_avg = r_q_cum / dt, qin_avg = r_qin_cum / dt;
    stv<COH>(d.r_q + (r), rq);
    stv<COH>(d.r_h + (r), rh);
    stv<COH>(d.r_storage + (r), r_storage);
    stv<COH>(d.r_qin + (r), rqin);
    stv<COH>(d.r_q_avg + (r), q_avg);
    stv<COH>(d.r_qin_avg + (r), qin_avg);
    if (d.balance)
        stv<COH>(d.bal_r + (r), (r_inwater + qin_avg + wfh_max(0.0, 0.0)) - (q_avg + 0.0 + 0.0) - (r_storage - r_store_prev) / dt);
}

// update_total_water_storage!  sbm.jl:143-182
template <bool COH = false>
__device__ __forceinline__ void total_storage_finish(const Dev& d, const int p, const bool riv, const double rh,
                                                     const double rfw, const double rfl, const double area,
                                                     const double snow_s, const double snow_w, const double canopy,
                                                     const double usd, const double swd, const double oh,
                                                     const double rfrac) {
    double tot = 0.0;
    if (riv) tot = (rh * rfw * rfl) / (area);
    const double ss = d.snow ? snow_s : 0.0, sw = d.snow ? snow_w : 0.0;
    tot = tot + ((((ss + sw) + 0.0 * 0.0) + canopy) + 0.0);
    const double sub = usd + swd;
    const double lat = oh * (1.0 - rfrac);
    tot += sub + lat;
    stv<COH>(d.total_storage + (p), tot);
}

// EXT: the model has reservoirs or lateral snow transport (a separate instantiation keeps their
// code, and its registers, out of the kernel that runs without them).
// FUSE: the wave performs each cell's vertical update itself at the start of the chunk (one
// launch per step instead of the vertical kernel before this one), so that the vertical work runs
// beside the long chains of the network instead of before them.
// MULTI: a run of a.nsteps steps in one launch (no EXT). Ticket t is step t / nchunks of chunk
// t % nchunks, so tickets go out step after step; a chunk starts a step once it has completed the
// previous one and every chunk it hands outflow to has completed the previous one too (which
// emptied the slots it fills). The wave then does the cell's vertical update of that step as
// well, from the step's forcing frame, and the arrays the steps hand on are accessed coherently.
// A chunk's step can thus run while the long chains of the previous step are still draining.
template <int N, bool EXT, bool MULTI = false, bool FUSE = false>
__global__ void __launch_bounds__(ROUTE_BLOCK) __attribute__((amdgpu_waves_per_eu(route_waves(N)))) k_route(Dev d,
                                                                                                    RouteArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ int x_k[ROUTE_BLOCK];     // per-wave exchange tables of the routing families
    __shared__ double x_q[ROUTE_BLOCK];
    // the wave's table base, wave-uniform (a scalar register rather than a vector one per lane)
    const int wbase = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~(WAVE - 1)));
    int* const lds_k = x_k + wbase;
    double* const lds_q = x_q + wbase;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[a.parity ^ 1] = 0ull;  // next launch's counter
    const int n = d.n;
    const double dt = d.dt;
    const unsigned long long tickets = (unsigned long long)a.nchunks * (MULTI ? (unsigned long long)a.nsteps : 1ull);
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(&a.counters[a.parity], 1ull);
        t = __shfl(t, 0);
        if (t >= tickets) break;
        const int c = MULTI ? (int)(t % (unsigned long long)a.nchunks) : (int)t;
        const int step = MULTI ? (int)(t / (unsigned long long)a.nchunks) : 0;
        const int p = c * WAVE + lane;
        const bool valid = p < n;
        const int q_ = valid ? p : 0;  // clamped index for loads of inactive lanes
        // Loads are issued for every lane (inactive lanes read cell 0) and selected afterwards:
        // a load under a per-lane condition becomes a branch, and the compiler then waits for
        // it before the join instead of overlapping it with the others.
        const int dn_ = d.down[q_], e0_ = d.up_ptr[q_], e1_ = d.up_ptr[q_ + 1];
        const uint8_t fl_ = d.flags[q_];
        const int dn = valid ? dn_ : -1;
        const bool cons_in = dn >= 0 && (dn >> 6) == c;
        // in-wave land consumer: a reservoir outlet (flag bit 0 clear) feeds no land cell
        const bool lcons_in = EXT ? (cons_in && (fl_ & 1)) : cons_in;
        const int e0 = e0_, e1 = valid ? e1_ : e0_;
        const uint8_t fl = valid ? fl_ : 0;
        if constexpr (MULTI) {
            if (!wait_chunks(a, c, dn >= 0 ? (dn >> 6) : c, step)) return;
        }

        // The cell's vertical update (the first half of the step): done here at the start of the
        // chunk (FUSE, MULTI), or its state outputs read back from the vertical kernel's.
        VertOut<N> vo;
        if constexpr (FUSE || MULTI) {
            if (valid) {
                if constexpr (MULTI) {
                    Dev dv = d;
                    dv.P = a.run_P + (size_t)step * n;
                    dv.E = a.run_E + (size_t)step * n;
                    dv.T = a.run_T + (size_t)step * n;
                    vertical_cell<N, true>(dv, p, vo);
                } else {
                    vertical_cell<N>(d, p, vo);
                }
            } else {
#pragma unroll
                for (int k = 0; k < N; ++k) vo.us[k] = 0.0;
                vo.zi = vo.recharge = vo.sat_exc = vo.runoff_l = vo.inf_exc = vo.aeow_l = vo.nrr = 0.0;
                vo.aet = vo.leak = vo.canopy = vo.snow_s = vo.snow_w = vo.store_prev = vo.P = 0.0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < N; ++k) vo.us[k] = d.ustore[k * n + q_];
            vo.zi = d.zi[q_];
            vo.recharge = d.recharge[q_];
        }
        // ---------------- lateral subsurface flow ----------------
        const double area = ld(d.area, q_);
        const double dsoil = ld(d.d, q_), ts = ld(d.ts, q_), tr = ld(d.tr, q_), tfc = ld(d.tfc, q_);
        const double sy = ld(d.sy, q_), slope = ld(d.slope, q_), flen = ld(d.fl, q_), fw = ld(d.fw, q_);
        const double qmax = ld(d.qmax, q_), kh0 = ld(d.kh0, q_), fk = ld(d.f, q_);
        const double zexp = EXT ? ld(d.zexp, q_) : INFINITY;
        const double rech = vo.recharge;
        Column<N> col;
        soil_layers<N>(dsoil, d.cfg_cum, d.cfg_thick, col.act, col.cum);
#pragma unroll
        for (int k = 0; k < N; ++k) col.ustore[k] = vo.us[k];
        col.zi = vo.zi;
        col.n_unsat = set_layerthickness<N>(col.zi, col.cum, col.act, col.thick);
        col.theta_e = ts - tr;
        col.theta_fc_r = tfc - tr;
        double sq = ldv<MULTI>(d.ssf_q + q_), szi = col.zi;  // sbm_model.jl:80 hands the soil water table over
        SsfCum cum{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // Inflow-independent part of the first sub-step's solve, evaluated before the chain
        // reaches this cell; the closing soil-column update of the last sub-step is applied
        // after the family (nothing in the family reads the column after that solve).
        SsfPrep pp = ssf_prepare(sq, szi, slope, sy, a.S_s == 1 ? a.dts_s[1] : a.dts_s[0], flen, kh0, fk, zexp);
        bool col_upd = false;
        double col_z0 = 0.0, col_z1 = 0.0;
        auto ssf_compute = [&](int k, double qin, double tor) -> double {
            const double dts = (k == a.S_s - 1) ? a.dts_s[1] : a.dts_s[0];
            if (k > 0) pp = ssf_prepare(sq, szi, slope, sy, dts, flen, kh0, fk, zexp);
            // flux!(::RechargeModel) with check_flux  boundary_conditions.jl:12-21,219-236
            double flx = rech * area;
            if (szi >= dsoil) flx = wfh_max(0.0, flx);
            cum.rech += flx * dts;
            const double qnb = 0.0 + flx;
            cum.tor += tor * dts;
            double qn, zin, exf, net;
            if (kinematic_wave_ssf<N>(qin, sq, szi, qnb, slope, sy, dsoil, dts, flen, fw, qmax, kh0, fk, zexp, pp,
                                      col, k == a.S_s - 1, qn, zin, exf, net)) {
                col_upd = true;
                col_z0 = szi;
                col_z1 = zin;
            }
            sq = qn;
            szi = zin;
            cum.qin += qin * dts;
            cum.q += qn * dts;
            cum.exf += exf * dts;
            cum.qnet += net * area * dts;
            return qn;
        };
        const bool lchain = d.chain_land[c] != 0;  // wave-uniform
        if (lchain && a.S_s == 1
                ? !run_chain<FOLD_SSF>(a, lane, c, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ssf, p,
                                       (fl & 1) && !lcons_in, ssf_compute)
                : !run_family<FOLD_SSF>(a, lane, c, valid, a.S_s, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ssf,
                                        (size_t)n, p, (fl & 1) && !lcons_in, lcons_in ? (dn & 63) : -1, lds_k, lds_q,
                                        ssf_compute))
            return;
        if (col_upd) update_ustorelayerdepth<N>(col, col_z0, col_z1);

        // Inputs of the later phases (soil storage, balances, overland flow), loaded once the
        // subsurface family is done, so that they take no registers while it runs (with FUSE,
        // the vertical fluxes this lane stored at the start of the chunk). Every one of these
        // arrays is written, in this kernel, only by the lane that reads it, after the read.
        if constexpr (!MULTI) {
            vo.sat_exc = d.sat_excess[q_];
            vo.runoff_l = d.runoff_land[q_];
            vo.inf_exc = d.inf_excess[q_];
            vo.aeow_l = d.aeow_l[q_];
            vo.nrr = d.net_runoff_river[q_];
            vo.canopy = d.canopy[q_];
            const double P_ = d.P[q_], aet_ = d.aet[q_], leak_ = d.leakage[q_], sp_ = d.store_prev[q_];
            const double ss_ = d.snow_s[q_], sw_ = d.snow_w[q_];
            vo.P = d.balance ? ((EXT && (fl & 8)) ? 0.0 : P_) : 0.0;
            vo.aet = d.balance ? aet_ : 0.0;
            vo.leak = d.balance ? leak_ : 0.0;
            vo.store_prev = d.balance ? sp_ : 0.0;
            vo.snow_s = d.snow ? ss_ : 0.0;
            vo.snow_w = d.snow ? sw_ : 0.0;
        }
        const double rd = ld(d.rd, q_);
        const double snow_in = EXT ? d.snow_in[q_] : 0.0, snow_out = EXT ? d.snow_out[q_] : 0.0;
        const double sat_exc = vo.sat_exc, runoff_l = vo.runoff_l, inf_exc = vo.inf_exc;
        const double aeow = vo.aeow_l, nrr = vo.nrr, canopy = vo.canopy;
        const double P_in = vo.P, aet = vo.aet;
        const double leak = vo.leak, store_prev = vo.store_prev;
        const double snow_s = vo.snow_s, snow_w = vo.snow_w;
        const double alpha = ld(d.alpha, q_), sfw = ld(d.sfw, q_), rfrac = ld(d.rf, q_);
        const double oq0 = ldv<MULTI>(d.ol_q + q_), oh0 = ldv<MULTI>(d.ol_h + q_),
                     o_store_prev = ldv<MULTI>(d.ol_storage + q_);

        // ---------------- average_flux_vars!, update_soil_water_storage! ----------------
        SsfEnd se{0.0, 0.0, 0.0, 0.0};
        if (valid)
            se = ssf_finish<N, MULTI>(d, p, col,
                               SsfCellIn{ts, tr, rd, dsoil, area, sy, sat_exc, runoff_l, inf_exc, aeow, nrr, P_in, aet,
                                         leak, canopy, snow_s, snow_w, store_prev, snow_in, snow_out},
                               cum, sq, szi);
        const double o_inwater = se.o_inwater, s_tor_avg = se.s_tor_avg, usd = se.usd, swd = se.swd;

        // ---------------- overland kinematic wave ----------------
        const double oqlat = o_inwater / flen;
        double oq = oq0, oh = oh0, o_storage = 0.0, oqin = 0.0;
        double o_q_cum = 0.0, o_qin_cum = 0.0, o_tor_cum = 0.0;
        KwPrep op = kw_prepare(oq, oqlat, alpha, a.S_l == 1 ? a.dts_l[1] : a.dts_l[0], flen);
        auto ol_compute = [&](int k, double tor, double qin_all) -> double {
            const double dts = (k == a.S_l - 1) ? a.dts_l[1] : a.dts_l[0];
            if (k > 0) op = kw_prepare(oq, oqlat, alpha, dts, flen);
            o_tor_cum += tor * dts;
            double qin = 0.0;
            if (sfw > 0.0) qin = qin_all;
            double qn, ar;
            kw_solve(qin, op, alpha, qn, ar);
            oq = qn;
            if (sfw > 0.0) oh = ar / sfw;
            o_storage = flen * sfw * oh;
            o_q_cum += qn * dts;
            o_qin_cum += qin * dts;
            oqin = qin;
            return qn;
        };
        if (lchain && a.S_l == 1
                ? !run_chain<FOLD_OL>(a, lane, c, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ol, p,
                                      (fl & 1) && !lcons_in, ol_compute)
                : !run_family<FOLD_OL>(a, lane, c, valid, a.S_l, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ol,
                                       (size_t)n, p, (fl & 1) && !lcons_in, lcons_in ? (dn & 63) : -1, lds_k, lds_q,
                                       ol_compute))
            return;
        // inputs of the river phase, loaded once the overland family is done
        const int r_ = d.river_of[q_];
        const int r = valid ? r_ : -1;
        const bool riv = r >= 0;
        const int rr = riv ? r : 0;
        const int re0 = d.rup_ptr[rr], re1_ = d.rup_ptr[riv ? rr + 1 : rr], re1 = riv ? re1_ : re0;
        const double rfl = d.r_fl[rr], rfw = d.r_fw[rr], ralpha = d.r_alpha[rr], rarea = d.r_area[rr];
        const int rres_ = EXT ? d.res_of_r[rr] : -1;
        const int rres = riv ? rres_ : -1;
        const double rq0 = ldv<MULTI>(d.r_q + rr), rh0 = ldv<MULTI>(d.r_h + rr),
                     r_store_prev = ldv<MULTI>(d.r_storage + rr);
        double r_inwater = 0.0;
        if (valid)
            r_inwater = ol_finish<MULTI>(d, p, r, oq, oh, o_storage, oqin, o_q_cum, o_tor_cum, o_qin_cum, o_inwater,
                                         o_store_prev, s_tor_avg, nrr, area, rarea);

        // ---------------- river kinematic wave ----------------
        const bool rcons_in = riv && (fl & 4) && cons_in;  // a river cell drains to the land downstream
        const double rqlat = r_inwater / rfl;
        double rq = rq0, rh = rh0, r_storage = 0.0, rqin = 0.0;
        double r_q_cum = 0.0, r_qin_cum = 0.0;
        double inflow = 0.0 / rfl;  // reservoir / lake / external inflow terms (none in this model)
        inflow -= 0.0 / rfl;
        const double rqlat_all = rqlat + inflow;
        KwPrep rp = kw_prepare(rq, rqlat_all, ralpha, a.S_r == 1 ? a.dts_r[1] : a.dts_r[0], rfl);
        auto riv_compute = [&](int k, double sumq, double res_out) -> double {
            const double dts = (k == a.S_r - 1) ? a.dts_r[1] : a.dts_r[0];
            if (k > 0) rp = kw_prepare(rq, rqlat_all, ralpha, dts, rfl);
            const double qin = (EXT ? res_out : 0.0) + sumq;
            double qn, ar;
            kw_solve(qin, rp, ralpha, qn, ar);
            rq = qn;
            rh = ar / rfw;
            r_storage = rfl * ar;
            r_q_cum += qn * dts;
            r_qin_cum += qin * dts;
            rqin = qin;
            if (EXT && rres >= 0) return reservoir_step(d, rres, p, qn, dts, k == 0, k == a.S_r - 1);
            return qn;
        };
        if (d.chain_rl[c] != 0 && a.S_r <= 4
                ? !run_chain_diag<EXT ? FOLD_RIV_RES : FOLD_RIV>(a, lane, c, a.S_r, d.rup_node, d.rup_idx, d.rup_res,
                                                                 re0, re1, a.slot_r, (size_t)d.nriv, rr,
                                                                 riv && (fl & 4) && !rcons_in, riv_compute)
                : !run_family<EXT ? FOLD_RIV_RES : FOLD_RIV>(a, lane, c, riv, a.S_r, d.rup_node, d.rup_idx,
                                                             d.rup_res, re0, re1, a.slot_r, (size_t)d.nriv, rr,
                                                             riv && (fl & 4) && !rcons_in,
                                                             rcons_in ? (dn & 63) : -1, lds_k, lds_q, riv_compute))
            return;
        if (riv) riv_finish<MULTI>(d, r, rq, rh, r_storage, rqin, r_q_cum, r_qin_cum, r_inwater, r_store_prev);
        if (MULTI && riv && a.run_q_avg) a.run_q_avg[(size_t)step * d.nriv + r] = r_q_cum / dt;

        // ---------------- update_total_water_storage!  sbm.jl:143-182 ----------------
        if (valid)
            total_storage_finish<MULTI>(d, p, riv, rh, rfw, rfl, area, snow_s, snow_w, canopy, usd, swd, oh, rfrac);
        if constexpr (MULTI) {
            // every store of the chunk's step has completed before the step is marked done
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            asm volatile("" ::: "memory");
            if (lane == 0) __hip_atomic_store(a.chunk_done + c, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
// ============================================================================================
// Adaptive sub-steps (kinematic_wave__adaptive_time_step_flag): each sub-step's length is a
// quantile (overland, river) or minimum (subsurface) over the state the previous sub-step left
// in every cell, so sub-steps of one family cannot overlap; every sub-step is one dataflow
// launch over the cells (the river one over river cells only), with the carried state in
// memory. The loop over sub-steps runs on the host (update_subsurface_flow_model!
// lateral_subsurface_flow.jl:279-304, update_overland_flow_model! surface_kinwave.jl:347-385,
// update_river_flow_model! :613-662).
// ============================================================================================
#define AD_TICKET(NCHUNKS)                                                       \
    unsigned long long t_ = 0;                                                   \
    if (lane == 0) t_ = atomicAdd(&a.counters[a.parity], 1ull);                  \
    t_ = __shfl(t_, 0);                                                          \
    if (t_ >= (unsigned long long)(NCHUNKS)) break;                              \
    const int c = (int)t_;

// One subsurface sub-step of length dts (first: the step's first sub-step, last: its last).
template <int N>
__global__ void __launch_bounds__(ROUTE_BLOCK) k_ad_ssf(Dev d, RouteArgs a, AdArgs x, double dts, int first,
                                                        int last) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ int x_k[ROUTE_BLOCK];
    __shared__ double x_q[ROUTE_BLOCK];
    int* const lds_k = x_k + (threadIdx.x & ~(WAVE - 1));
    double* const lds_q = x_q + (threadIdx.x & ~(WAVE - 1));
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[a.parity ^ 1] = 0ull;
    const int n = d.n;
    for (;;) {
        AD_TICKET(a.nchunks)
        const int p = c * WAVE + lane;
        const bool valid = p < n;
        const int q_ = valid ? p : 0;
        const int dn_ = d.down[q_], e0 = d.up_ptr[q_], e1_ = d.up_ptr[q_ + 1];
        const uint8_t fl_ = d.flags[q_];
        const int dn = valid ? dn_ : -1;
        const uint8_t fl = valid ? fl_ : 0;
        const bool cons_in = dn >= 0 && (dn >> 6) == c && (fl & 1);  // a reservoir outlet feeds no land cell
        const int e1 = valid ? e1_ : e0;
        const double area = ld(d.area, q_);
        const double dsoil = ld(d.d, q_), ts = ld(d.ts, q_), tr = ld(d.tr, q_), tfc = ld(d.tfc, q_);
        const double sy = ld(d.sy, q_), slope = ld(d.slope, q_), flen = ld(d.fl, q_), fw = ld(d.fw, q_);
        const double qmax = ld(d.qmax, q_), kh0 = ld(d.kh0, q_), fk = ld(d.f, q_), zexp = ld(d.zexp, q_);
        const double rech = d.recharge[q_];
        Column<N> col;
        soil_layers<N>(dsoil, d.cfg_cum, d.cfg_thick, col.act, col.cum);
#pragma unroll
        for (int k = 0; k < N; ++k) col.ustore[k] = d.ustore[k * n + q_];
        col.zi = d.zi[q_];
        col.n_unsat = set_layerthickness<N>(col.zi, col.cum, col.act, col.thick);
        col.theta_e = ts - tr;
        col.theta_fc_r = tfc - tr;
        double sq = d.ssf_q[q_], szi = x.szi[q_];
        SsfCum cum{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (!first) {
            cum.qin = x.scum[q_];
            cum.q = x.scum[(size_t)n + q_];
            cum.exf = x.scum[2 * (size_t)n + q_];
            cum.qnet = x.scum[3 * (size_t)n + q_];
            cum.tor = x.scum[4 * (size_t)n + q_];
            cum.rech = x.scum[5 * (size_t)n + q_];
        }
        const SsfPrep pp = ssf_prepare(sq, szi, slope, sy, dts, flen, kh0, fk, zexp);
        bool col_upd = false;
        double col_z0 = 0.0, col_z1 = 0.0;
        auto ssf_compute = [&](int, double qin, double tor) -> double {
            // flux!(::RechargeModel) with check_flux  boundary_conditions.jl:12-21,219-236
            double flx = rech * area;
            if (szi >= dsoil) flx = wfh_max(0.0, flx);
            cum.rech += flx * dts;
            const double qnb = 0.0 + flx;
            cum.tor += tor * dts;
            double qn, zin, exf, net;
            if (kinematic_wave_ssf<N>(qin, sq, szi, qnb, slope, sy, dsoil, dts, flen, fw, qmax, kh0, fk, zexp, pp, col,
                                      true,
                                      qn, zin, exf, net)) {
                col_upd = true;
                col_z0 = szi;
                col_z1 = zin;
            }
            sq = qn;
            szi = zin;
            cum.qin += qin * dts;
            cum.q += qn * dts;
            cum.exf += exf * dts;
            cum.qnet += net * area * dts;
            return qn;
        };
        if (d.chain_land[c] != 0
                ? !run_chain<FOLD_SSF>(a, lane, c, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ssf, p,
                                       (fl & 1) && !cons_in, ssf_compute)
                : !run_family<FOLD_SSF>(a, lane, c, valid, 1, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ssf,
                                        (size_t)n, p, (fl & 1) && !cons_in, cons_in ? (dn & 63) : -1, lds_k, lds_q,
                                        ssf_compute))
            return;
        if (col_upd) update_ustorelayerdepth<N>(col, col_z0, col_z1);
        if (valid && !last) {
            d.ssf_q[p] = sq;
            x.szi[p] = szi;
            d.zi[p] = col.zi;
#pragma unroll
            for (int k = 0; k < N; ++k) d.ustore[k * n + p] = col.ustore[k];
            x.scum[p] = cum.qin;
            x.scum[(size_t)n + p] = cum.q;
            x.scum[2 * (size_t)n + p] = cum.exf;
            x.scum[3 * (size_t)n + p] = cum.qnet;
            x.scum[4 * (size_t)n + p] = cum.tor;
            x.scum[5 * (size_t)n + p] = cum.rech;
        } else if (valid) {
            const double P_in = d.balance && !(d.flags[p] & 8) ? d.P[p] : 0.0, aet = d.balance ? d.aet[p] : 0.0;
            const double leak = d.balance ? d.leakage[p] : 0.0, sp = d.balance ? d.store_prev[p] : 0.0;
            const double ss = d.snow ? d.snow_s[p] : 0.0, sw = d.snow ? d.snow_w[p] : 0.0;
            ssf_finish<N>(d, p, col,
                          SsfCellIn{ts, tr, ld(d.rd, p), dsoil, area, sy, d.sat_excess[p], d.runoff_land[p],
                                    d.inf_excess[p], d.aeow_l[p], d.net_runoff_river[p], P_in, aet, leak, d.canopy[p],
                                    ss, sw, sp, d.snow_in[p], d.snow_out[p]},
                          cum, sq, szi);
        }
    }
}

// One overland sub-step of length dts.
__global__ void __launch_bounds__(ROUTE_BLOCK) k_ad_ol(Dev d, RouteArgs a, AdArgs x, double dts, int first,
                                                       int last) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ int x_k[ROUTE_BLOCK];
    __shared__ double x_q[ROUTE_BLOCK];
    int* const lds_k = x_k + (threadIdx.x & ~(WAVE - 1));
    double* const lds_q = x_q + (threadIdx.x & ~(WAVE - 1));
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[a.parity ^ 1] = 0ull;
    const int n = d.n;
    for (;;) {
        AD_TICKET(a.nchunks)
        const int p = c * WAVE + lane;
        const bool valid = p < n;
        const int q_ = valid ? p : 0;
        const int dn_ = d.down[q_], e0 = d.up_ptr[q_], e1_ = d.up_ptr[q_ + 1];
        const uint8_t fl_ = d.flags[q_];
        const int dn = valid ? dn_ : -1;
        const uint8_t fl = valid ? fl_ : 0;
        const bool cons_in = dn >= 0 && (dn >> 6) == c && (fl & 1);  // a reservoir outlet feeds no land cell
        const int e1 = valid ? e1_ : e0;
        const double alpha = ld(d.alpha, q_), sfw = ld(d.sfw, q_), flen = ld(d.fl, q_);
        double oq = d.ol_q[q_], oh = d.ol_h[q_];
        const double o_inwater = d.ol_inwater[q_];
        double o_q_cum = 0.0, o_qin_cum = 0.0, o_tor_cum = 0.0;
        if (!first) {
            o_q_cum = x.ocum[q_];
            o_qin_cum = x.ocum[(size_t)n + q_];
            o_tor_cum = x.ocum[2 * (size_t)n + q_];
        }
        const double oqlat = o_inwater / flen;
        double o_storage = 0.0, oqin = 0.0;
        const KwPrep op = kw_prepare(oq, oqlat, alpha, dts, flen);
        auto ol_compute = [&](int, double tor, double qin_all) -> double {
            o_tor_cum += tor * dts;
            double qin = 0.0;
            if (sfw > 0.0) qin = qin_all;
            double qn, ar;
            kw_solve(qin, op, alpha, qn, ar);
            oq = qn;
            if (sfw > 0.0) oh = ar / sfw;
            o_storage = flen * sfw * oh;
            o_q_cum += qn * dts;
            o_qin_cum += qin * dts;
            oqin = qin;
            return qn;
        };
        if (d.chain_land[c] != 0
                ? !run_chain<FOLD_OL>(a, lane, c, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ol, p,
                                      (fl & 1) && !cons_in, ol_compute)
                : !run_family<FOLD_OL>(a, lane, c, valid, 1, d.up_idx, d.up_idx, d.up_f, e0, e1, a.slot_ol,
                                       (size_t)n, p, (fl & 1) && !cons_in, cons_in ? (dn & 63) : -1, lds_k, lds_q,
                                       ol_compute))
            return;
        if (valid && !last) {
            d.ol_q[p] = oq;
            d.ol_h[p] = oh;
            x.ocum[p] = o_q_cum;
            x.ocum[(size_t)n + p] = o_qin_cum;
            x.ocum[2 * (size_t)n + p] = o_tor_cum;
        } else if (valid) {
            const int r = d.river_of[p];
            const double o_store_prev = d.ol_storage[p];
            ol_finish(d, p, r, oq, oh, o_storage, oqin, o_q_cum, o_tor_cum, o_qin_cum, o_inwater, o_store_prev,
                      d.ssf_tor_avg[p], d.net_runoff_river[p], ld(d.area, p), r >= 0 ? d.r_area[r] : 0.0);
        }
    }
}

// ============================================================================================
// 1-D floodplain of the kinematic-wave river (floodplain_1d). One launch per river sub-step does
// the channel-floodplain exchange of each cell, the river sub-step and the floodplain
// accumulation (k_ad_riv<RES, true>); k_fp_qin follows, k_fp_finish after the last sub-step.
// ============================================================================================
// interpolation_indices  floodplain.jl:272-284, 0-based: the last index with v[i] <= x (the
// first when none), and the next one except at the end
__device__ __forceinline__ void fp_indices(const double x, const double* v, const int n, int& i1, int& i2) {
    int a = 0;
    for (int i = 0; i < n; ++i)
        if (v[i] <= x) a = i;
    i1 = a;
    i2 = a == n - 1 ? a : a + 1;
}

// compute_floodplain_flow_area  floodplain.jl:330-341 (compute_flood_flow_area :286-300) on the
// width (W) and flow area (A) rows of a cell's profile
__device__ __forceinline__ double fp_flow_area(const double* W, const double* A, const double* depth, const double h,
                                               const int i1, const int i2) {
    const double channel_area = W[0] * h;
    const double delta_h = h - depth[i1];
    const double flood_flow_area = A[i1] + (W[i2] * delta_h);
    return wfh_max(flood_flow_area - channel_area, 0.0);
}

// river_channel_floodplain_exchange!  surface_kinwave.jl:572-601 for river cell r: the channel
// storage is the previous sub-step's (the previous step's at the first sub-step). Returns the
// floodplain storage and depth after the exchange and the exchange flux (the river's extra
// lateral inflow times the flow length).
__device__ __forceinline__ void fp_exchange(const Dev& d, const int r, const bool first, const double dts, double& fs,
                                            double& fh, double& exch) {
    const int nd = d.fp_nd;
    const double rs = first ? d.r_storage[r] : d.fp_rs[r];
    fs = d.fp_storage[r];
    const double storage_total = rs + fs;
    double delta_river_storage;
    if (storage_total > d.r_bfs[r]) {
        const double flood_storage = storage_total - d.r_bfs[r];
        const double* S = d.fp_prof + (size_t)r * 4 * nd;
        const double* W = S + nd;
        const double rfl = d.r_fl[r];
        int i1, i2;
        fp_indices(flood_storage, S, nd, i1, i2);  // compute_flood_depth  floodplain.jl:316-327
        const double delta_A = (flood_storage - S[i1]) / rfl;
        const double delta_h = delta_A / W[i2];
        const double h = d.fp_depth[i1] + delta_h;
        const double river_storage = (d.r_bfd[r] + h) * d.r_fw[r] * rfl;
        delta_river_storage = river_storage - rs;
        fs = wfh_max(storage_total - river_storage, 0.0);
        fh = fs > 0.0 ? h : 0.0;
    } else {
        delta_river_storage = wfh_max(storage_total - rs, 0.0);
        fh = 0.0;
        fs = 0.0;
    }
    exch = delta_river_storage / dts;
}

// The Manning flow capacity of update_floodplain_model! (surface_kinwave.jl:395-425) for river
// cell r at flood depth h: zero unless this and the downstream cell's floodplain flow areas at
// this depth exceed 1e-5 m2; rd = the downstream river cell or -1.
__device__ __forceinline__ double fp_capacity(const Dev& d, const int r, const int rd, const double h) {
    if (!(h > 0.0)) return 0.0;
    const int nd = d.fp_nd;
    int i1, i2;
    fp_indices(h, d.fp_depth, nd, i1, i2);
    const double* Wv = d.fp_prof + ((size_t)r * 4 + 1) * nd;
    const double flow_area = fp_flow_area(Wv, Wv + nd, d.fp_depth, h, i1, i2);
    double flow_area_ds = flow_area;
    if (rd >= 0) {
        const double* Wd = d.fp_prof + ((size_t)rd * 4 + 1) * nd;
        flow_area_ds = fp_flow_area(Wd, Wd + nd, d.fp_depth, h, i1, i2);
    }
    if (!(flow_area > 1.0e-05 && flow_area_ds > 1.0e-05)) return 0.0;
    const double wetted_perimeter = Wv[2 * nd + i1] + 2.0 * (h - d.fp_depth[i1]);
    const double hydraulic_radius = flow_area / wetted_perimeter;
    // manning_flow  surface_process.jl:166-169
    return wfh_cbrt(hydraulic_radius * hydraulic_radius) * sqrt(d.fp_slope[r]) * flow_area / d.fp_n[r];
}

// One river sub-step of length dts, over the river cells in river-position order (river cells
// sorted by node, a topological order of the river graph).
// FP (a floodplain model): each lane first does its cell's channel-floodplain exchange, which
// enters the river sub-step as lateral inflow; after the river family the wave runs the
// floodplain accumulation (accucapacityflux! of update_floodplain_model!, routing/utils.jl:82-108)
// down the unfiltered river graph as a second family: a cell adds what its upstream cells pass
// on, in the reference's river topological order, to its floodplain storage, passes on
// min(storage / dt, capacity) and keeps the rest.
template <bool RES, bool FP, bool LINK>
__global__ void __launch_bounds__(ROUTE_BLOCK) k_ad_riv(Dev d, RouteArgs a, AdArgs x, double dts, int first,
                                                        int last) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ int x_k[ROUTE_BLOCK];
    __shared__ double x_q[ROUTE_BLOCK];
    int* const lds_k = x_k + (threadIdx.x & ~(WAVE - 1));
    double* const lds_q = x_q + (threadIdx.x & ~(WAVE - 1));
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[a.parity ^ 1] = 0ull;
    const int nr = d.nriv;
    for (;;) {
        AD_TICKET(x.nchunks_r)
        const int rp = c * WAVE + lane;
        const bool valid = rp < nr;
        const int q_ = valid ? rp : 0;
        const int r_ = x.rp_r[q_], rdn_ = x.rp_down[q_], e0 = x.rp_up_ptr[q_], e1_ = x.rp_up_ptr[q_ + 1];
        const int r = r_;
        const int rdn = valid ? rdn_ : -1;
        const bool cons_in = rdn >= 0 && (rdn >> 6) == c;
        const int e1 = valid ? e1_ : e0;
        const double ralpha = d.r_alpha[r], rfw = d.r_fw[r], rfl = d.r_fl[r];
        double rq = d.r_q[r], rh = d.r_h[r];
        const double r_inwater = d.r_inwater[r];
        double r_q_cum = 0.0, r_qin_cum = 0.0;
        if (!first) {
            r_q_cum = x.rcum[q_];
            r_qin_cum = x.rcum[(size_t)nr + q_];
        }
        const double rqlat = r_inwater / rfl;
        double fs = 0.0, fh = 0.0, fexch = 0.0;
        if (FP) fp_exchange(d, r, first != 0, dts, fs, fh, fexch);
        double inflow = 0.0 / rfl;  // reservoir / lake / external inflow terms (none in this model)
        inflow -= 0.0 / rfl;
        if (FP) inflow += fexch / rfl;  // channel-floodplain exchange (surface_kinwave.jl:534-536)
        const double rqlat_all = rqlat + inflow;
        double r_storage = 0.0, rqin = 0.0;
        const KwPrep rpp = kw_prepare(rq, rqlat_all, ralpha, dts, rfl);
        const int rres_ = RES ? d.res_of_r[r] : -1, pnode = RES ? x.rp_node[q_] : 0;
        const int rres = valid ? rres_ : -1;
        auto riv_compute = [&](int, double sumq, double res_out) -> double {
            const double qin = (RES ? res_out : 0.0) + sumq;
            double qn, ar;
            kw_solve(qin, rpp, ralpha, qn, ar);
            rq = qn;
            rh = ar / rfw;
            r_storage = rfl * ar;
            r_q_cum += qn * dts;
            r_qin_cum += qin * dts;
            rqin = qin;
            if (RES && rres >= 0) return reservoir_step<LINK>(d, rres, pnode, qn, dts, first != 0, last != 0);
            return qn;
        };
        const bool chain = x.chain_r[c] != 0;  // wave-uniform
        if (chain ? !run_chain<RES ? FOLD_RIV_RES : FOLD_RIV>(a, lane, c, x.rp_up_idx, x.rp_up_idx, x.rp_up_res, e0, e1,
                                                                a.slot_r, rp, rdn >= 0 && !cons_in, riv_compute)
                  : !run_family<RES ? FOLD_RIV_RES : FOLD_RIV>(a, lane, c, valid, 1, x.rp_up_idx, x.rp_up_idx,
                                                                 x.rp_up_res, e0, e1, a.slot_r, (size_t)nr, rp,
                                                                 rdn >= 0 && !cons_in, cons_in ? (rdn & 63) : -1,
                                                                 lds_k, lds_q, riv_compute))
            return;
        if (FP) {
            const double cap = fp_capacity(d, r, rdn_ >= 0 ? x.rp_r[rdn_] : -1, fh);
            double s_new = fs, flux = 0.0;
            auto pass = [&](int, double material, double) -> double {
                const double flux_val = wfh_min(material / dts, cap);
                const double update = flux_val * dts;
                s_new = material - update;
                flux = flux_val;
                return update;
            };
            // the floodplain lists hold the same upstream cells as the river's (in another order),
            // so a chain chunk of the river is one of the floodplain too
            const int f0 = x.fp_sup_ptr[q_], f1 = valid ? x.fp_sup_ptr[q_ + 1] : f0;
            if (chain ? !run_chain<FOLD_SNOW>(a, lane, c, x.fp_sup_idx, x.fp_sup_idx, nullptr, f0, f1, a.slot_ssf, rp,
                                              rdn >= 0 && !cons_in, pass, fs)
                      : !run_family<FOLD_SNOW>(a, lane, c, valid, 1, x.fp_sup_idx, x.fp_sup_idx, nullptr, f0, f1,
                                               a.slot_ssf, (size_t)nr, rp, rdn >= 0 && !cons_in,
                                               cons_in ? (rdn & 63) : -1, lds_k, lds_q, pass, fs))
                return;
            if (valid) {
                if (first) {
                    d.fp_q_cum[r] = 0.0;
                    d.fp_qin_cum[r] = 0.0;
                    if (d.balance) d.fp_sprev[r] = d.r_storage[r] + d.fp_storage[r];  // get_storage  mass_balance.jl:132-140
                }
                d.fp_rs[r] = r_storage;
                d.fp_h[r] = fh;
                d.fp_exch[r] = fexch;
                d.fp_storage[r] = s_new;
                d.fp_q[r] = flux;
                d.fp_q_cum[r] = (first ? 0.0 : d.fp_q_cum[r]) + flux * dts;
            }
        }
        if (valid && !last) {
            d.r_q[r] = rq;
            d.r_h[r] = rh;
            x.rcum[rp] = r_q_cum;
            x.rcum[(size_t)nr + rp] = r_qin_cum;
        } else if (valid) {
            riv_finish(d, r, rq, rh, r_storage, rqin, r_q_cum, r_qin_cum, r_inwater, d.r_storage[r]);
        }
    }
}

// flux_in!  routing/utils.jl:162-167: floodplain inflow = sum of the upstream cells' floodplain
// flow over the upstream lists without reservoir outlets, ascending river index
__global__ void k_fp_qin(Dev d, AdArgs x, double dts) {
    const int rp = blockIdx.x * blockDim.x + threadIdx.x;
    if (rp >= d.nriv) return;
    double s = 0.0;
    for (int e = x.rp_up_ptr[rp]; e < x.rp_up_ptr[rp + 1]; ++e)
        if (x.rp_up_res[e] == 0.0) s = s + d.fp_q[x.rp_r[x.rp_up_idx[e]]];
    const int r = x.rp_r[rp];
    d.fp_qin[r] = s;
    d.fp_qin_cum[r] += s * dts;
}

// After the last river sub-step (surface_kinwave.jl:652-659): the river averages become channel
// plus floodplain, and the river balance counts the floodplain storage (mass_balance.jl:296-313)
__global__ void k_fp_finish(Dev d) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nriv) return;
    const double dt = d.dt;
    const double fq_avg = d.fp_q_cum[r] / dt, fqin_avg = d.fp_qin_cum[r] / dt;
    const double qch = d.r_q_avg[r];
    const double q_avg = qch + fq_avg, qin_avg = d.r_qin_avg[r] + fqin_avg;
    d.fp_q_avg[r] = fq_avg;
    d.fp_qin_avg[r] = fqin_avg;
    d.r_qch_avg[r] = qch;
    d.r_q_avg[r] = q_avg;
    d.r_qin_avg[r] = qin_avg;
    if (d.balance)
        d.bal_r[r] = (d.r_inwater[r] + qin_avg + wfh_max(0.0, 0.0)) - (q_avg + 0.0 + 0.0) -
                     ((d.r_storage[r] + d.fp_storage[r]) - d.fp_sprev[r]) / dt;
}

// update_total_water_storage! after the adaptive river sub-steps.
__global__ void k_ad_total(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    const int r = d.river_of[p];
    const bool riv = r >= 0;
    const int rr = riv ? r : 0;
    total_storage_finish(d, p, riv, riv ? d.r_h[rr] : 0.0, riv ? d.r_fw[rr] : 0.0, riv ? d.r_fl[rr] : 0.0,
                         ld(d.area, p), d.snow_s[p], d.snow_w[p], d.canopy[p], d.ustoredepth[p], d.swd[p], d.ol_h[p],
                         ld(d.rf, p));
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, o), hi = __shfl_xor((unsigned)(v >> 32), o);
        const unsigned long long w = ((unsigned long long)hi << 32) | lo;
        v = w < v ? w : v;
    }
    return v;
}

// ============================================================================================
// Lateral snow transport (snow_gravitational_transport__flag)  surface_process.jl:9-19
// ============================================================================================
// One dataflow launch after the vertical kernel: snow storage, then snow water, each passed down
// the unfiltered flow graph with accucapacityflux! (routing/utils.jl:82-108); a cell adds what its
// upstream cells pass on, in the reference's topological order, to its own material, passes on
// min(material / dt, capacity) and keeps the rest. The capacities come from the state before
// the transport (TAN80 = 5.67, snow_storage_max = 10 m).
__global__ void __launch_bounds__(ROUTE_BLOCK) k_snow_transport(Dev d, RouteArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ int x_k[ROUTE_BLOCK];
    __shared__ double x_q[ROUTE_BLOCK];
    int* const lds_k = x_k + (threadIdx.x & ~(WAVE - 1));
    double* const lds_q = x_q + (threadIdx.x & ~(WAVE - 1));
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[a.parity ^ 1] = 0ull;
    const int n = d.n;
    const double dt = d.dt;
    for (;;) {
        AD_TICKET(a.nchunks)
        const int p = c * WAVE + lane;
        const bool valid = p < n;
        const int q_ = valid ? p : 0;
        const int dn_ = d.down[q_], e0 = d.sup_ptr[q_], e1_ = d.sup_ptr[q_ + 1];
        const int dn = valid ? dn_ : -1;
        const bool cons_in = dn >= 0 && (dn >> 6) == c;
        const int e1 = valid ? e1_ : e0;
        const double ss = d.snow_s[q_], sw = d.snow_w[q_], slope = ld(d.slope, q_);
        const double frac = wfh_min(0.5, slope / 5.67) * wfh_min(1.0, ss / 10.0);
        const double cap_s = frac * ss / dt, cap_w = sw * frac / dt;
        double ss_new = ss, sw_new = sw, flux_s = 0.0, flux_w = 0.0;
        auto pass_s = [&](int, double material, double) -> double {
            const double flux_val = wfh_min(material / dt, cap_s);
            const double update = flux_val * dt;
            ss_new = material - update;
            flux_s = flux_val;
            return update;
        };
        const bool chain = d.chain_sup[c] != 0;  // wave-uniform
        if (chain ? !run_chain<FOLD_SNOW>(a, lane, c, d.sup_idx, d.sup_idx, nullptr, e0, e1, a.slot_ssf, p,
                                          dn >= 0 && !cons_in, pass_s, ss)
                  : !run_family<FOLD_SNOW>(a, lane, c, valid, 1, d.sup_idx, d.sup_idx, nullptr, e0, e1, a.slot_ssf,
                                           (size_t)n, p, dn >= 0 && !cons_in, cons_in ? (dn & 63) : -1, lds_k, lds_q,
                                           pass_s, ss))
            return;
        auto pass_w = [&](int, double material, double) -> double {
            const double flux_val = wfh_min(material / dt, cap_w);
            const double update = flux_val * dt;
            sw_new = material - update;
            flux_w = flux_val;
            return update;
        };
        if (chain ? !run_chain<FOLD_SNOW>(a, lane, c, d.sup_idx, d.sup_idx, nullptr, e0, e1, a.slot_ol, p,
                                          dn >= 0 && !cons_in, pass_w, sw)
                  : !run_family<FOLD_SNOW>(a, lane, c, valid, 1, d.sup_idx, d.sup_idx, nullptr, e0, e1, a.slot_ol,
                                           (size_t)n, p, dn >= 0 && !cons_in, cons_in ? (dn & 63) : -1, lds_k, lds_q,
                                           pass_w, sw))
            return;
        if (valid) {
            d.snow_s[p] = ss_new;
            d.snow_w[p] = sw_new;
            d.snow_out[p] = flux_s + flux_w;
        }
    }
}

// flux_in!  routing/utils.jl:162-167: snow_in = sum of snow_out over the upstream lists (without
// reservoir outlets), ascending id.
__global__ void k_snow_in(Dev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    double s = 0.0;
    for (int e = d.up_ptr[p]; e < d.up_ptr[p + 1]; ++e) s = s + d.snow_out[d.up_idx[e]];
    d.snow_in[p] = s;
}

// ---- stable time steps ----
__global__ void k_sel_reset(SelCtl* s) {
    s->hist[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        s->count = 0ull;
        s->rank = 0ull;
        s->prefix = 0ull;
        s->mask = 0ull;
        s->bin_count = 0ull;
        s->next = ~0ull;
        s->minkey = ~0ull;
    }
}

// Courant time step flow_length / celerity of every cell with q > KIN_WAVE_MIN_FLOW as a key
// (all-ones for the others), and their count.  surface_kinwave.jl:684-693
__global__ void k_kw_keys(const double* __restrict__ q, Prm alpha, Prm flen, int m, uint64_t* keys, SelCtl* s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool elig = false;
    if (i < m) {
        const double qi = q[i];
        uint64_t key = ~0ull;
        if (qi > WFH_KIN_WAVE_MIN_FLOW) {
            const double cel = 1.0 / (ld(alpha, i) * 0.6 * wfh_pow(qi, (0.6 - 1.0)));
            key = wfh_bits(ld(flen, i) / cel);
            elig = true;
        }
        keys[i] = key;
    }
    const unsigned long long b = __ballot(elig);
    if ((threadIdx.x & (WAVE - 1)) == 0 && b) atomicAdd(&s->count, (unsigned long long)__popcll(b));
}

// Rank of the lower order statistic of quantile!(v, p) (0-based), from the count.
__global__ void k_sel_init(SelCtl* s, double p) {
    const unsigned long long k = s->count;
    if (k >= 2) {
        double g;
        s->rank = (unsigned long long)(wfh_quantile_index((int64_t)k, p, &g) - 1);
    } else {
        s->rank = 0ull;
    }
}

__global__ void __launch_bounds__(256) k_sel_hist(const uint64_t* __restrict__ keys, int m, SelCtl* s, int shift) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long mask = s->mask, prefix = s->prefix;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        const uint64_t key = keys[i];
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&s->hist[threadIdx.x], h[threadIdx.x]);
}

// The digit whose bin holds the wanted rank; clears the histogram for the next pass.
__global__ void k_sel_pick(SelCtl* s, int shift) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = s->hist[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long r = s->rank;
        unsigned long long cum = 0ull;  // keys in the bins below B
        int B = -1;
        for (int b = 0; b < 256; ++b) {
            if (r < cum + h[b]) {
                B = b;
                break;
            }
            cum += h[b];
        }
        if (B < 0) {  // rank beyond the keys (no eligible value): any bin, the result is unused
            B = 255;
            cum = r;
        }
        s->rank = r - cum;
        s->prefix |= (unsigned long long)B << shift;
        s->mask |= 0xFFull << shift;
        s->bin_count = h[B];
    }
    s->hist[threadIdx.x] = 0u;
}

// Smallest key above the selected one, needed when the selected key's multiplicity ends at it.
__global__ void __launch_bounds__(256) k_sel_next(const uint64_t* __restrict__ keys, int m, SelCtl* s) {
    if (s->rank + 1 < s->bin_count) return;
    const unsigned long long a = s->prefix;
    unsigned long long best = ~0ull;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        const uint64_t key = keys[i];
        if (key > a && key < best) best = key;
    }
    best = wave_min_u64(best);
    if ((threadIdx.x & (WAVE - 1)) == 0 && best != ~0ull) atomicMin(&s->next, best);
}

// Subsurface Courant time steps: minimum of flow_length / ssf_celerity over the cells with a
// positive water table depth (reduced per wave, then one atomic per wave).
// lateral_subsurface_flow.jl:314-344
// Grid-stride over the cells with one minimum and one count per lane, reduced per wave: the
// minimum and the count do not depend on the order, and a fixed grid keeps the atomics on the
// two shared words to one pair per wave.
__global__ void __launch_bounds__(256) k_ssf_min(Dev d, AdArgs x, SelCtl* s) {
    unsigned long long key = ~0ull;
    unsigned cnt = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < d.n; i += gridDim.x * 256) {
        const double z = x.szi[i];
        if (z > 0.0) {
            const double ze = ld(d.zexp, i);
            const double cel = (ld(d.kh0, i) * wfh_exp(-ld(d.f, i) * (z < ze ? z : ze)) * ld(d.slope, i)) / ld(d.sy, i);
            const unsigned long long k = wfh_bits(ld(d.fl, i) / cel);
            key = k < key ? k : key;
            ++cnt;
        }
    }
    key = wave_min_u64(key);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & (WAVE - 1)) == 0 && cnt) {
        atomicMin(&s->minkey, key);
        atomicAdd(&s->count, (unsigned long long)cnt);
    }
}

// --------------------------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t* p, uint64_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_gather(double* out, const double* in, const int* idx, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}
__global__ void k_scatter(double* out, const double* in, const int* idx, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[idx[i]] = in[i];
}

}  // namespace

// ============================================================================================
// host side
// ============================================================================================
struct wfh_handle {
    wfh_config cfg{};
    int64_t n = 0, nriv = 0;
    int N = 0;
    int S_s = 0, S_l = 0, S_r = 0;
    double dts_s[2]{}, dts_l[2]{}, dts_r[2]{};
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;  // 3 events per timed step: start, after K1, after K2
    size_t timed_steps = 0;
    bool timing = false;
    bool forcing_set = false, broken = false;
    bool broken_nonfinite = false;  // broken by non-finite forcing only: a loaded state repairs it
    int* run_chunk_done = nullptr;  // [nchunks] steps completed per chunk (wfh_run)
    int parity = 0;
    int route_grid = 0;
    int route_grid_max = 0;  // resident blocks of the routing kernel on this device
    // graph mode: the fixed-sub-step step's launches recorded into a graph per starting ticket
    // parity and replayed; the autotuner times candidate routing grids on real steps
    bool graph_mode = false;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    std::vector<int> tune_grids;
    std::vector<double> tune_ms;
    int tune_next = 0, tune_left = 0;
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    int nchunks = 0;
    int max_level = 0;
    int thin_level = 0;
    std::string err;
    // host topology (land index order unless stated)
    std::vector<int64_t> land_order, land_down, river_order, river_down, river_land;
    std::vector<int> node_of;   // land index -> node
    std::vector<int> land_of;   // node -> land index
    // device allocations
    std::vector<void*> allocs;
    Dev d{};
    RouteArgs ra{};
    double *fP = nullptr, *fE = nullptr, *fT = nullptr;  // node-order forcing buffers
    int* d_land_of = nullptr;
    // host forcing: uploaded in the caller's cell order on a copy stream, then gathered into the
    // node order on the model's stream (allocated at the first host forcing)
    double* hf_tmp = nullptr;              // [3][n] device
    hipStream_t copy_stream = nullptr;
    hipEvent_t hf_copied = nullptr, hf_gathered = nullptr;
    double* get_tmp = nullptr;             // [N][n] device: host wfh_get scatters into cell order here
    // adaptive sub-steps
    AdArgs ad{};
    SelCtl* sel = nullptr;       // device
    SelCtl* sel_host = nullptr;  // pinned read-back
    uint64_t* keys = nullptr;    // [max(n, nriv)]
    int32_t nsub[3]{};           // sub-steps of the last step
    int64_t nres = 0;
    std::vector<double> fftr_land;  // flow fraction to river, land index order
    bool ext = false;               // reservoirs, lateral snow transport or an exponential-constant profile
    bool fuse = false;              // the routing kernel does the vertical update (launch_step)
};

namespace {

const int PCR_DX[10] = {0, -1, 0, 1, -1, 0, 1, -1, 0, 1};
const int PCR_DY[10] = {0, -1, -1, -1, 0, 0, 0, 1, 1, 1};

int32_t fail(wfh_handle* h, int32_t code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return fail((h), WFH_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
T* dalloc(wfh_handle* h, size_t count, hipError_t& e) {
    void* p = nullptr;
    e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) h->allocs.push_back(p);
    return (T*)p;
}

// flowgraph  Wflow/src/routing/utils.jl:6-33 over sorted raster indices; -1 = pit
void flowgraph(const int64_t* idx, int64_t n, const uint8_t* ldd, int64_t nx, int64_t ny, std::vector<int64_t>& down,
               std::vector<uint8_t>& ldd_out) {
    down.assign(n, -1);
    ldd_out.assign(n, 5);
    for (int64_t i = 0; i < n; ++i) {
        const int l = ldd[i];
        ldd_out[i] = (uint8_t)l;
        if (l == 5) continue;
        if (l < 1 || l > 9) {
            ldd_out[i] = 5;
            continue;
        }
        const int64_t x = idx[i] % nx + PCR_DX[l], y = idx[i] / nx + PCR_DY[l];
        if (x < 0 || y < 0 || x >= nx || y >= ny) {
            ldd_out[i] = 5;
            continue;
        }
        const int64_t t = x + y * nx;
        const int64_t* it = std::lower_bound(idx, idx + n, t);
        if (it == idx + n || *it != t) {
            ldd_out[i] = 5;
            continue;
        }
        down[i] = it - idx;
    }
}

// Graphs.topological_sort_by_dfs for a graph with at most one out-neighbour per vertex.
// Returns false on a cycle (is_cyclic check of flowgraph).
bool toposort_dfs(const std::vector<int64_t>& down, std::vector<int64_t>& order) {
    const int64_t n = (int64_t)down.size();
    std::vector<uint8_t> color(n, 0);
    std::vector<int64_t> post;
    post.reserve(n);
    std::vector<int64_t> stack;
    for (int64_t v = 0; v < n; ++v) {
        if (color[v]) continue;
        stack.clear();
        stack.push_back(v);
        color[v] = 1;
        while (!stack.empty()) {
            const int64_t u = stack.back();
            const int64_t w = down[u];
            if (w >= 0 && color[w] == 1) return false;
            if (w >= 0 && color[w] == 0) {
                color[w] = 1;
                stack.push_back(w);
            } else {
                color[u] = 2;
                post.push_back(u);
                stack.pop_back();
            }
        }
    }
    order.assign(post.rbegin(), post.rend());
    return true;
}

double flow_length_of(int ldd, double xl, double yl) {  // get_flow_length  utils.jl:418-430
    if (ldd == 2 || ldd == 8) return yl;
    if (ldd == 4 || ldd == 6) return xl;
    return std::hypot(xl, yl);
}
double flow_width_of(int ldd, double xl, double yl) {  // get_flow_width  utils.jl:439-450
    if (ldd == 2 || ldd == 8) return xl;
    if (ldd == 4 || ldd == 6) return yl;
    return (xl * yl) / std::hypot(xl, yl);
}

// upload a parameter map (land order in) in node order, broadcast when uniform
int32_t put_prm(wfh_handle* h, const std::vector<double>& v, Prm& out) {
    const size_t n = v.size();
    bool uniform = n > 0;
    for (size_t i = 1; i < n && uniform; ++i) uniform = wfh_bits(v[i]) == wfh_bits(v[0]);
    hipError_t e;
    if (uniform) {
        double* p = dalloc<double>(h, 1, e);
        if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
        HIPCHK(h, hipMemcpy(p, v.data(), sizeof(double), hipMemcpyHostToDevice));
        out = Prm{p, 0u};
    } else {
        std::vector<double> tmp(n);
        for (size_t q = 0; q < n; ++q) tmp[q] = v[h->land_of[q]];
        double* p = dalloc<double>(h, n, e);
        if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
        HIPCHK(h, hipMemcpy(p, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
        out = Prm{p, 0xFFFFFFFFu};
    }
    return WFH_OK;
}

// layered map [N][n] in land order
int32_t put_prml(wfh_handle* h, const std::vector<double>& v, int N, PrmL& out) {
    const size_t n = (size_t)h->n;
    bool uniform = true;
    for (int k = 0; k < N && uniform; ++k)
        for (size_t i = 1; i < n && uniform; ++i) uniform = wfh_bits(v[k * n + i]) == wfh_bits(v[k * n]);
    hipError_t e;
    if (uniform) {
        std::vector<double> tmp(N);
        for (int k = 0; k < N; ++k) tmp[k] = v[k * n];
        double* p = dalloc<double>(h, N, e);
        if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
        HIPCHK(h, hipMemcpy(p, tmp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        out = PrmL{p, 0u, 1u};
    } else {
        std::vector<double> tmp((size_t)N * n);
        for (int k = 0; k < N; ++k)
            for (size_t q = 0; q < n; ++q) tmp[k * n + q] = v[k * n + h->land_of[q]];
        double* p = dalloc<double>(h, (size_t)N * n, e);
        if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
        HIPCHK(h, hipMemcpy(p, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
        out = PrmL{p, 0xFFFFFFFFu, (uint32_t)n};
    }
    return WFH_OK;
}

// device array in node order from land-order host values
int32_t put_state(wfh_handle* h, const std::vector<double>& v, double*& out, int layers = 1) {
    const size_t n = (size_t)h->n;
    hipError_t e;
    out = dalloc<double>(h, n * layers, e);
    if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
    if (!v.empty()) {
        std::vector<double> tmp(n * layers);
        for (int k = 0; k < layers; ++k)
            for (size_t q = 0; q < n; ++q) tmp[k * n + q] = v[k * n + h->land_of[q]];
        HIPCHK(h, hipMemcpy(out, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        HIPCHK(h, hipMemset(out, 0, n * layers * sizeof(double)));
    }
    return WFH_OK;
}

int32_t put_river(wfh_handle* h, const std::vector<double>& v, double*& out) {
    hipError_t e;
    out = dalloc<double>(h, (size_t)h->nriv, e);
    if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
    if (h->nriv > 0) {
        if (!v.empty())
            HIPCHK(h, hipMemcpy(out, v.data(), h->nriv * sizeof(double), hipMemcpyHostToDevice));
        else
            HIPCHK(h, hipMemset(out, 0, h->nriv * sizeof(double)));
    }
    return WFH_OK;
}

template <class T>
int32_t put_raw(wfh_handle* h, const std::vector<T>& v, const T*& out) {
    hipError_t e;
    T* p = dalloc<T>(h, v.size(), e);
    if (e != hipSuccess) return fail(h, WFH_ERR_HIP, "hipMalloc");
    if (!v.empty()) HIPCHK(h, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    out = p;
    return WFH_OK;
}

int32_t launch_fill(wfh_handle* h, uint64_t* p, uint64_t v, size_t n) {
    if (n == 0) return WFH_OK;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, h->stream, p, v, n);
    HIPCHK(h, hipGetLastError());
    return WFH_OK;
}

int sub_steps(double dt, double dt_fixed, double dts[2]) {
    // while t < dt: dt_s = check_timestepsize(dt_fixed, t, dt) (routing/timestepping.jl:11-17)
    double t = 0.0;
    int k = 0;
    dts[0] = dts[1] = 0.0;
    while (t < dt) {
        double s = dt_fixed;
        if (t + s > dt) s = dt - t;
        if (k == 0) dts[0] = s;
        dts[1] = s;
        t += s;
        ++k;
    }
    return k;
}

struct ProgState {
    double** p;
    int layers;
    bool river;
};

}  // namespace

#define REQF(field)                                                                \
    do {                                                                           \
        if (!st->field) {                                                          \
            h->err = "wfh_static." #field " is NULL";                              \
            int32_t rc_ = WFH_ERR_ARG;                                             \
            wfh_destroy_keep_error(h, out);                                        \
            return rc_;                                                            \
        }                                                                          \
    } while (0)

static thread_local std::string g_create_error;

static void wfh_destroy_keep_error(wfh_handle* h, wfh_handle** out) {
    g_create_error = h->err;
    wfh_destroy(h);
    *out = nullptr;
}

template <int N>
static void launch_kernels(wfh_handle* h);

// interpolate_linear(x, xp, fp)  reservoir.jl:364-377 as the reference writes it, for the
// storage curve at create (initialize_storage, get_maximum_storage); fp read with stride fs.
static double interp_host(double x, const double* xp, const double* fp, int64_t n, int64_t fs) {
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
    if (i1 < 0 || i1 + 1 >= n) return std::nan("");
    const int64_t i2 = i1 + 1;
    return fp[i1 * fs] * (1.0 - (x - xp[i1]) / (xp[i2] - xp[i1])) + fp[i2 * fs] * (x - xp[i1]) / (xp[i2] - xp[i1]);
}

extern "C" int32_t wfh_create(const wfh_config* cfg, const wfh_static* st, wfh_handle** out) {
    if (!out) return WFH_ERR_ARG;
    *out = nullptr;
    if (!cfg || !st) {
        g_create_error = "null argument";
        return WFH_ERR_ARG;
    }
    if (cfg->abi_version != WFH_ABI_VERSION) {
        g_create_error = "ABI version mismatch";
        return WFH_ERR_ARG;
    }
    wfh_handle* h = new wfh_handle();
    *out = h;
    h->cfg = *cfg;
    const int64_t n = st->n;
    const int N = cfg->n_config_layers + 1;
    const bool adaptive = cfg->kinematic_wave_adaptive != 0;
    const bool fp = cfg->floodplain_1d != 0;
    if (n < 1 || n > (int64_t)INT32_MAX / 2 || N < 2 || N > 6 || !(cfg->dt > 0.0) ||
        (!adaptive && !(cfg->land_dt > 0.0 && cfg->river_dt > 0.0 && cfg->subsurface_dt > 0.0)) ||
        (adaptive && !(cfg->subsurface_alpha_coefficient > 0.0 && cfg->subsurface_alpha_coefficient < 1e300))) {
        h->err = "unsupported sizes (1 <= n < 2^30, 1 <= soil_layer__thickness count <= 5), non-positive time step "
                 "or non-positive subsurface alpha coefficient";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_ARG;
    }
    REQF(cell_index); REQF(ldd); REQF(x_length); REQF(y_length); REQF(land_slope); REQF(land_manning_n);
    REQF(river_location); REQF(river_width_land); REQF(river_length_land); REQF(water_fraction); REQF(elevation);
    REQF(canopy_gap_fraction); REQF(max_canopy_storage); REQF(rooting_depth); REQF(crop_coefficient);
    REQF(theta_s); REQF(theta_r); REQF(kv_0); REQF(f); REQF(air_entry_pressure); REQF(h1); REQF(h2);
    REQF(h3_high); REQF(h3_low); REQF(h4); REQF(alpha_h1); REQF(soil_thickness); REQF(infiltcap_compacted);
    REQF(max_leakage); REQF(compacted_fraction); REQF(rootdistpar); REQF(cap_hmax); REQF(cap_n); REQF(w_soil);
    REQF(cf_soil); REQF(brooks_corey_c); REQF(kvfrac); REQF(khfrac);
    const bool gash = cfg->dt >= 23.0 * 3600.0;
    if (gash) REQF(evap_precip_ratio);
    if (cfg->snow) { REQF(cfmax); REQF(tt); REQF(tti); REQF(ttm); REQF(whc); }
    for (int64_t i = 1; i < n; ++i)
        if (st->cell_index[i] <= st->cell_index[i - 1]) {
            h->err = "cell_index must be strictly ascending";
            wfh_destroy_keep_error(h, out);
            return WFH_ERR_ARG;
        }
    int64_t nriv = 0;
    for (int64_t i = 0; i < n; ++i) nriv += st->river_location[i] != 0;
    if (nriv != st->nriv) {
        h->err = "nriv does not match river_location";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_ARG;
    }
    if (nriv > 0 && (!st->river_slope || !st->river_manning_n || !st->river_bankfull_depth || !st->river_length ||
                     !st->river_width)) {
        h->err = "river arrays are NULL";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_ARG;
    }
    h->n = n;
    h->nriv = nriv;
    h->N = N;
    int32_t rc;

    // ---------------- topology (network.jl:48-258) ----------------
    std::vector<uint8_t> ldd2;
    flowgraph(st->cell_index, n, st->ldd, st->nx, st->ny, h->land_down, ldd2);
    if (!toposort_dfs(h->land_down, h->land_order)) {
        h->err = "One or more cycles detected in flow graph";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_CYCLE;
    }
    std::vector<int64_t> land_river(n, -1);
    for (int64_t i = 0; i < n; ++i)
        if (st->river_location[i]) {
            land_river[i] = (int64_t)h->river_land.size();
            h->river_land.push_back(i);
        }
    {
        std::vector<int64_t> ridx(nriv);
        std::vector<uint8_t> rldd(nriv), rldd2;
        for (int64_t r = 0; r < nriv; ++r) {
            ridx[r] = st->cell_index[h->river_land[r]];
            rldd[r] = ldd2[h->river_land[r]];
        }
        flowgraph(ridx.data(), nriv, rldd.data(), st->nx, st->ny, h->river_down, rldd2);
        if (!toposort_dfs(h->river_down, h->river_order)) {
            h->err = "One or more cycles detected in river flow graph";
            wfh_destroy_keep_error(h, out);
            return WFH_ERR_CYCLE;
        }
    }
    // reservoirs (NetworkReservoir, network.jl:296-360): outlets on river cells, in ascending
    // land (= river) index; the reservoir's area from reservoir_coverage (the outlet alone without)
    const int64_t nres = st->nres;
    std::vector<int> res_of_land(n, -1), res_of_river(nriv, -1), res_cov_id(n, 0);
    if (nres < 0 || (nres > 0 && (!st->reservoir_outlet || !st->reservoir_outflow_type || !st->reservoir_area ||
                                  !st->reservoir_waterlevel))) {
        h->err = "nres < 0, or reservoir_outlet / outflow_type / area / waterlevel missing";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_ARG;
    }
    {
        std::vector<char> fed(nriv, 0);
        for (int64_t k = 0; k < nres; ++k) {
            const int64_t i = st->reservoir_outlet[k];
            const bool ok_i = i >= 0 && i < n && (k == 0 || i > st->reservoir_outlet[k - 1]) && land_river[i] >= 0;
            const int64_t rd = ok_i ? h->river_down[land_river[i]] : -1;
            const int32_t t = st->reservoir_outflow_type[k];
            const int32_t sct = st->reservoir_storage_curve_type ? st->reservoir_storage_curve_type[k] : 1;
            if (!ok_i || rd < 0 || fed[rd] || t < 1 || t > 4 || (sct != 1 && sct != 2)) {
                h->err = "reservoir " + std::to_string(k) +
                         ": outlet must be a river cell with a river cell downstream (outlets ascending, one "
                         "reservoir per downstream cell), outflow type 1..4 and storage curve type 1 or 2";
                wfh_destroy_keep_error(h, out);
                return WFH_ERR_ARG;
            }
            // S-H and H-Q tables: present, non-empty, finite, H and S non-decreasing
            bool tab_ok = true;
            if (sct == 2) {
                const int64_t* sp = st->reservoir_sh_ptr;
                tab_ok = sp && st->reservoir_sh_h && st->reservoir_sh_s && sp[k + 1] > sp[k];
                for (int64_t j = tab_ok ? sp[k] : 0; tab_ok && j < sp[k + 1]; ++j)
                    tab_ok = std::isfinite(st->reservoir_sh_h[j]) && std::isfinite(st->reservoir_sh_s[j]) &&
                             (j == sp[k] || (st->reservoir_sh_h[j] >= st->reservoir_sh_h[j - 1] &&
                                             st->reservoir_sh_s[j] >= st->reservoir_sh_s[j - 1]));
            }
            if (t == 1) {
                const int64_t* hp = st->reservoir_hq_ptr;
                tab_ok = tab_ok && hp && st->reservoir_hq_h && st->reservoir_hq_q && hp[k + 1] > hp[k];
                for (int64_t j = tab_ok ? hp[k] : 0; tab_ok && j < hp[k + 1]; ++j) {
                    tab_ok = std::isfinite(st->reservoir_hq_h[j]) &&
                             (j == hp[k] || st->reservoir_hq_h[j] >= st->reservoir_hq_h[j - 1]);
                    for (int c = 0; tab_ok && c < 365; ++c) tab_ok = std::isfinite(st->reservoir_hq_q[j * 365 + c]);
                }
            }
            if (!tab_ok) {
                h->err = "reservoir " + std::to_string(k) +
                         ": missing, empty, non-finite or decreasing S-H / H-Q table";
                wfh_destroy_keep_error(h, out);
                return WFH_ERR_ARG;
            }
            fed[rd] = 1;
            res_of_land[i] = (int)k;
            res_of_river[land_river[i]] = (int)k;
        }
        std::vector<int64_t> cnt(nres, 0);
        for (int64_t i = 0; i < n && nres > 0; ++i) {
            const int64_t id = st->reservoir_coverage ? st->reservoir_coverage[i] : res_of_land[i] + 1;
            if (id < 0 || id > nres) {
                h->err = "reservoir_coverage out of range";
                wfh_destroy_keep_error(h, out);
                return WFH_ERR_ARG;
            }
            res_cov_id[i] = (int)id;
            if (id > 0) ++cnt[id - 1];
        }
        for (int64_t k = 0; k < nres; ++k)
            if (cnt[k] == 0) {
                h->err = "reservoir " + std::to_string(k) + " has an empty area";
                wfh_destroy_keep_error(h, out);
                return WFH_ERR_ARG;
            }
    }
    // linked free weirs (lower_reservoir_ind, reservoir.jl:152-159): the lower reservoir's outlet
    // lies downstream of the weir's on the river, and one weir links to a reservoir at most
    std::vector<int> res_lower(std::max<int64_t>(nres, 1), -1), res_upper(std::max<int64_t>(nres, 1), -1);
    bool linked = false;
    for (int64_t k = 0; k < nres && st->reservoir_lower; ++k) {
        const int32_t lw = st->reservoir_lower[k];
        if (lw == 0) continue;
        bool ok = lw >= 1 && lw <= nres && lw - 1 != k && st->reservoir_outflow_type[k] == 2 && res_upper[lw - 1] < 0;
        if (ok) {
            ok = false;
            const int64_t target = land_river[st->reservoir_outlet[lw - 1]];
            for (int64_t rr = h->river_down[land_river[st->reservoir_outlet[k]]]; rr >= 0; rr = h->river_down[rr])
                if (rr == target) {
                    ok = true;
                    break;
                }
        }
        if (!ok) {
            h->err = "reservoir " + std::to_string(k) + ": linked lower reservoir must be another reservoir whose "
                     "outlet lies downstream on the river, linked from one free weir only";
            wfh_destroy_keep_error(h, out);
            return WFH_ERR_ARG;
        }
        res_lower[k] = lw - 1;
        res_upper[lw - 1] = (int)k;
        linked = true;
    }
    h->nres = nres;
    if (fp) {  // FloodPlainProfile inputs (floodplain.jl:48-147, 380-402)
        bool ok = st->floodplain_ndepth >= 1 && st->floodplain_depth && st->floodplain_volume &&
                  st->floodplain_manning_n && st->floodplain_slope && st->river_bankfull_depth;
        for (int64_t j = 0; ok && j < st->floodplain_ndepth; ++j)
            ok = std::isfinite(st->floodplain_depth[j]) && st->floodplain_depth[j] > (j ? st->floodplain_depth[j - 1] : 0.0);
        for (int64_t j = 0; ok && j < (int64_t)st->floodplain_ndepth * nriv; ++j) ok = std::isfinite(st->floodplain_volume[j]);
        if (!ok) {
            h->err = "floodplain: flood depths (ascending, > 0), finite volumes, manning_n, slope and "
                     "river_bankfull_depth needed";
            wfh_destroy_keep_error(h, out);
            return WFH_ERR_ARG;
        }
    }
    // levels (longest upstream path) in DFS topological order, then the level-major node order
    std::vector<int32_t> level(n, 0);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t v = h->land_order[k], w = h->land_down[v];
        if (w >= 0) level[w] = std::max(level[w], level[v] + 1);
    }
    h->max_level = 0;
    for (int64_t i = 0; i < n; ++i) h->max_level = std::max(h->max_level, level[i]);
    {
        // Levels holding at least one wave of cells are laid out level by level (ascending land
        // index inside a level). From the first level with fewer cells than a wave onwards the
        // cells are laid out in an upstream-first depth-first postorder that visits the deepest
        // upstream branch last, so a main stem and the cells feeding it stay contiguous and a
        // wave holds a long piece of one chain. Both parts are topological orders and every
        // cell of the first part precedes every cell of the second.
        std::vector<int64_t> per_level(h->max_level + 1, 0);
        for (int64_t i = 0; i < n; ++i) per_level[level[i]]++;
        int thin = h->max_level + 1;
        for (int l = 0; l <= h->max_level; ++l)
            if (per_level[l] < WAVE) {
                thin = l;
                break;
            }
        h->thin_level = thin;
        std::vector<int64_t> cnt(h->max_level + 2, 0);
        for (int64_t i = 0; i < n; ++i)
            if (level[i] < thin) cnt[level[i] + 1]++;
        for (int l = 0; l <= h->max_level; ++l) cnt[l + 1] += cnt[l];
        h->land_of.assign(n, 0);
        h->node_of.assign(n, 0);
        int64_t q_thin = 0;
        for (int64_t i = 0; i < n; ++i)
            if (level[i] < thin) {
                const int64_t q = cnt[level[i]]++;
                h->land_of[q] = (int)i;
                h->node_of[i] = (int)q;
                ++q_thin;
            }
        // upstream lists (ascending id) for the DFS
        std::vector<int64_t> uptr(n + 1, 0), uidx;
        for (int64_t i = 0; i < n; ++i)
            if (h->land_down[i] >= 0) uptr[h->land_down[i] + 1]++;
        for (int64_t i = 0; i < n; ++i) uptr[i + 1] += uptr[i];
        uidx.resize(uptr[n]);
        {
            std::vector<int64_t> fillp(uptr.begin(), uptr.end() - 1);
            for (int64_t i = 0; i < n; ++i)
                if (h->land_down[i] >= 0) uidx[fillp[h->land_down[i]]++] = i;
        }
        // subtree depth (longest upstream path) is `level`; roots of the thin part are thin cells
        // whose downstream cell is absent (pits): a thin cell only drains to thin cells
        std::vector<std::pair<int64_t, int>> stack;  // (cell, next child position)
        int64_t q = q_thin;
        for (int64_t root = 0; root < n; ++root) {
            if (level[root] < thin || h->land_down[root] >= 0) continue;
            stack.push_back({root, 0});
            while (!stack.empty()) {
                auto& top = stack.back();
                const int64_t v = top.first;
                // children in thin part, deepest (largest level) last; ties by ascending id
                std::vector<int64_t> kids;
                for (int64_t e = uptr[v]; e < uptr[v + 1]; ++e)
                    if (level[uidx[e]] >= thin) kids.push_back(uidx[e]);
                std::stable_sort(kids.begin(), kids.end(),
                                 [&](int64_t a, int64_t b) { return level[a] < level[b]; });
                if (top.second < (int)kids.size()) {
                    const int64_t u = kids[top.second++];
                    stack.push_back({u, 0});
                } else {
                    h->land_of[q] = (int)v;
                    h->node_of[v] = (int)q;
                    ++q;
                    stack.pop_back();
                }
            }
        }
        if (q != n) {
            h->err = "internal error: node order does not cover every cell";
            wfh_destroy_keep_error(h, out);
            return WFH_ERR_ARG;
        }
    }
    // upstream CSR in node order, lists in ascending land index
    std::vector<int> up_ptr(n + 1, 0), up_idx;
    std::vector<int> river_of(n, -1);
    std::vector<uint8_t> flags(n, 0);
    {
        std::vector<std::vector<int>> ups(n);
        // a reservoir outlet's overland and subsurface flow go to the reservoir: it is left out
        // of its downstream cell's upstream list (filter_upstream_nodes  utils.jl:61-71)
        for (int64_t i = 0; i < n; ++i)
            if (h->land_down[i] >= 0 && res_of_land[i] < 0) ups[h->land_down[i]].push_back((int)i);  // ascending i
        up_idx.reserve(n);
        for (int64_t q = 0; q < n; ++q) {
            const int v = h->land_of[q];
            for (int u : ups[v]) up_idx.push_back(h->node_of[u]);
            up_ptr[q + 1] = (int)up_idx.size();
            if (h->land_down[v] >= 0 && res_of_land[v] < 0) flags[q] |= 1;
            if (res_cov_id[v] > 0) flags[q] |= 8;
            if (land_river[v] >= 0) {
                flags[q] |= 2;
                river_of[q] = (int)land_river[v];
                if (h->river_down[land_river[v]] >= 0) flags[q] |= 4;
            }
        }
    }
    std::vector<int> rup_ptr(nriv + 1, 0), rup_idx, rup_node, down_node(n, -1);
    std::vector<double> rup_res;
    {
        std::vector<std::vector<int>> ups(nriv);
        for (int64_t r = 0; r < nriv; ++r)
            if (h->river_down[r] >= 0) ups[h->river_down[r]].push_back((int)r);
        for (int64_t r = 0; r < nriv; ++r) {
            for (int u : ups[r]) {
                rup_idx.push_back(u);
                rup_node.push_back(h->node_of[h->river_land[u]]);
                rup_res.push_back(res_of_river[u] >= 0 ? 1.0 : 0.0);
            }
            rup_ptr[r + 1] = (int)rup_idx.size();
        }
        for (int64_t q = 0; q < n; ++q) {
            const int64_t w = h->land_down[h->land_of[q]];
            down_node[q] = w >= 0 ? h->node_of[w] : -1;
        }
        // a river cell's river downstream is the river cell at its land downstream
        for (int64_t r = 0; r < nriv; ++r)
            if (h->river_down[r] >= 0 && h->river_land[h->river_down[r]] != h->land_down[h->river_land[r]]) {
                h->err = "river drainage differs from land drainage";
                wfh_destroy_keep_error(h, out);
                return WFH_ERR_ARG;
            }
    }

    // ---------------- land parameters (domain.jl:218-366) ----------------
    std::vector<double> area(n), fw(n), fl(n), slope(n), rf(n), wf(n), sfw(n), fftr(n, 0.0), alpha(n);
    for (int64_t i = 0; i < n; ++i) {
        const double xl = st->x_length[i], yl = st->y_length[i];
        area[i] = xl * yl;
        fw[i] = flow_width_of(ldd2[i], xl, yl);
        fl[i] = flow_length_of(ldd2[i], xl, yl);
        slope[i] = wfh_clamp(st->land_slope[i], 0.00001, INFINITY);
        rf[i] = st->river_location[i] ? wfh_min((st->river_length_land[i] * st->river_width_land[i]) / (area[i]), 1.0)
                                      : 0.0;
        wf[i] = wfh_max(st->water_fraction[i] - rf[i], 0.0);
        const double land_area = (1.0 - rf[i]) * area[i];
        sfw[i] = st->river_location[i] ? land_area / fl[i] : fw[i];
    }
    {  // get_flow_fraction_to_river  utils.jl:493-510 (every upstream j of a river cell i)
        for (int64_t j = 0; j < n; ++j) {
            const int64_t i = h->land_down[j];
            if (i >= 0 && st->river_location[i] && ldd2[j] != ldd2[i]) fftr[j] = slope[j] / (slope[i] + slope[j]);
        }
    }
    const double apow = (2.0 / 3.0) * 0.6;
    for (int64_t i = 0; i < n; ++i) {  // ManningFlowParameters  surface_kinwave.jl:32-44
        const double at = wfh_pow(st->land_manning_n[i] / std::sqrt(slope[i]), 0.6);
        alpha[i] = at * wfh_pow(sfw[i], apow);
    }
    std::vector<double> r_alpha(nriv), r_fw(nriv), r_fl(nriv), r_area(nriv);
    for (int64_t r = 0; r < nriv; ++r) {
        const double s = wfh_clamp(st->river_slope[r], 0.00001, INFINITY);
        r_fw[r] = st->river_width[r];
        r_fl[r] = st->river_length[r];
        const double wp = r_fw[r] + st->river_bankfull_depth[r];
        const double at = wfh_pow(st->river_manning_n[r] / std::sqrt(s), 0.6);
        r_alpha[r] = at * wfh_pow(wp, apow);
        r_area[r] = area[h->river_land[r]];
    }
    // ---------------- soil (soil.jl:333-615) ----------------
    double cfg_thick[WFH_MAX_LAYERS], cfg_cum[WFH_MAX_LAYERS + 1];
    for (int k = 0; k < N - 1; ++k) cfg_thick[k] = cfg->soil_layer_thickness_mm[k] * 1e-3;
    cfg_thick[N - 1] = NaN_;
    cfg_cum[0] = 0.0;
    cfg_cum[1] = cfg_thick[0];
    for (int k = 1; k < N; ++k) cfg_cum[k + 1] = cfg_cum[k] + cfg_thick[k];
    std::vector<double> tfc(n), rootfrac((size_t)N * n), rd(n), sy(n), kh0(n), qmax(n), q0(n), swd0(n), zi0(n),
        sfrac(n), icap_s(n), storage0(n);
    for (int64_t i = 0; i < n; ++i) {
        double act[WFH_MAX_LAYERS], cum[WFH_MAX_LAYERS + 1];
        int nl = 0;
        {
            for (int k = 0; k < N; ++k) act[k] = cfg_thick[k] * NaN_;
            for (int k = 0; k < N; ++k) {
                if (st->soil_thickness[i] > cfg_cum[k + 1])
                    act[k] = cfg_thick[k];
                else if (st->soil_thickness[i] - cfg_cum[k] > 0.0)
                    act[k] = st->soil_thickness[i] - cfg_cum[k];
                nl += !wfh_isnan(act[k]);
            }
            cum[0] = 0.0;
            cum[1] = act[0];
            for (int k = 1; k < N; ++k) cum[k + 1] = cum[k] + act[k];
        }
        const double ts = st->theta_s[i], tr = st->theta_r[i];
        if (st->theta_fc) {
            tfc[i] = st->theta_fc[i];
        } else {  // field_capacity  soil_process.jl:138-160 with vwc_brooks_corey :99-106 at h = -1 m
            double fc = 0.0, total = 0.0;
            for (int k = 0; k < nl; ++k) {
                const double hb = st->air_entry_pressure[i];
                const double c = st->brooks_corey_c[(size_t)k * n + i];
                double v;
                if (-1.0 < hb)
                    v = (ts - tr) * wfh_pow(hb / -1.0, 2.0 / (c - 3.0)) + tr;
                else
                    v = ts;
                fc += v * act[k];
                total += act[k];
            }
            tfc[i] = fc / total;
        }
        const double rdu = st->rooting_depth[i];
        for (int k = 0; k < N; ++k) {  // default root fraction (unclipped rooting depth)  soil.jl:526-548
            double v;
            if (st->rootfraction)
                v = st->rootfraction[(size_t)k * n + i];
            else if (rdu > 0.0)
                v = ((rdu - cum[k]) >= act[k]) ? act[k] / rdu : wfh_max((rdu - cum[k]) / rdu, 0.0);
            else
                v = 0.0;
            rootfrac[(size_t)k * n + i] = v;
        }
        const double d = st->soil_thickness[i];
        const double swc = d * (ts - tr);
        rd[i] = wfh_min(d * 0.99, rdu);  // sbm.jl:57-58
        swd0[i] = 0.85 * swc;             // soil.jl:164
        zi0[i] = wfh_max(0.0, d - swd0[i] / (ts - tr));
        sfrac[i] = wfh_max(st->canopy_gap_fraction[i] - wf[i] - rf[i] - 0.0, 0.0);  // soil_fraction!
        icap_s[i] = st->kv_0[i] * st->kvfrac[i];
        // lateral subsurface flow (lateral_subsurface_flow.jl:81-171, utils.jl:916-937)
        sy[i] = wfh_max(ts - tfc[i], 0.02);
        kh0[i] = st->khfrac[i] * st->kv_0[i];
        const double fk = st->f[i];
        const double ze = st->z_exp ? st->z_exp[i] : INFINITY;
        if (ze < INFINITY) {  // initialize_lateral_ssf_model!(::KhExponentialConstant)  utils.jl:940-975
            const double q_constant = kh0[i] * wfh_exp(-fk * ze) * slope[i] * (d - ze);
            qmax[i] = ((kh0[i] * slope[i]) / fk) * (1.0 - wfh_exp(-fk * ze)) + q_constant;
            if (zi0[i] < ze)
                q0[i] = (((kh0[i] * slope[i]) / fk) * (wfh_exp(-fk * zi0[i]) - wfh_exp(-fk * ze)) + q_constant) * fw[i];
            else
                q0[i] = kh0[i] * wfh_exp(-fk * zi0[i]) * slope[i] * (d - zi0[i]) * fw[i];
        } else {
            qmax[i] = ((kh0[i] * slope[i]) / fk) * (1.0 - wfh_exp(-fk * d));
            q0[i] = ((kh0[i] * slope[i]) / fk) * (wfh_exp(-fk * zi0[i]) - wfh_exp(-fk * d)) * fw[i];
        }
    }

    // ---------------- node order, uploads ----------------
    if (hipSetDevice(cfg->device) != hipSuccess) {
        h->err = "hipSetDevice failed";
        wfh_destroy_keep_error(h, out);
        return WFH_ERR_HIP;
    }
    auto V = [&](const double* p) { return std::vector<double>(p, p + n); };
    Dev& D = h->d;
    D.n = (int)n;
    D.nriv = (int)nriv;
    D.dt = cfg->dt;
    D.gash = gash;
    D.snow = cfg->snow != 0;
    D.infred = cfg->soil_infiltration_reduction != 0;
    D.balance = cfg->water_mass_balance != 0;
    for (int k = 0; k < WFH_MAX_LAYERS; ++k) D.cfg_thick[k] = k < N ? cfg_thick[k] : NaN_;
    for (int k = 0; k <= WFH_MAX_LAYERS; ++k) D.cfg_cum[k] = k <= N ? cfg_cum[k] : NaN_;
    std::vector<double> zero(n, 0.0);
    if ((rc = put_prm(h, V(st->canopy_gap_fraction), D.gap))) goto done;
    if ((rc = put_prm(h, V(st->max_canopy_storage), D.smax))) goto done;
    if ((rc = put_prm(h, V(st->crop_coefficient), D.kc))) goto done;
    if ((rc = put_prm(h, gash ? V(st->evap_precip_ratio) : zero, D.er))) goto done;
    if ((rc = put_prm(h, D.snow ? V(st->cfmax) : zero, D.cfmax))) goto done;
    if ((rc = put_prm(h, D.snow ? V(st->tt) : zero, D.tt))) goto done;
    if ((rc = put_prm(h, D.snow ? V(st->tti) : zero, D.tti))) goto done;
    if ((rc = put_prm(h, D.snow ? V(st->ttm) : zero, D.ttm))) goto done;
    if ((rc = put_prm(h, D.snow ? V(st->whc) : zero, D.whc))) goto done;
    if ((rc = put_prm(h, rf, D.rf))) goto done;
    if ((rc = put_prm(h, wf, D.wf))) goto done;
    if ((rc = put_prm(h, sfrac, D.sfrac))) goto done;
    if ((rc = put_prm(h, V(st->theta_s), D.ts))) goto done;
    if ((rc = put_prm(h, st->z_exp ? V(st->z_exp) : std::vector<double>(n, INFINITY), D.zexp))) goto done;
    if ((rc = put_prm(h, V(st->theta_r), D.tr))) goto done;
    if ((rc = put_prm(h, tfc, D.tfc))) goto done;
    if ((rc = put_prm(h, V(st->kv_0), D.kv0))) goto done;
    if ((rc = put_prm(h, V(st->f), D.f))) goto done;
    if ((rc = put_prm(h, V(st->air_entry_pressure), D.hb))) goto done;
    if ((rc = put_prm(h, V(st->h1), D.h1))) goto done;
    if ((rc = put_prm(h, V(st->h2), D.h2))) goto done;
    if ((rc = put_prm(h, V(st->h3_high), D.h3h))) goto done;
    if ((rc = put_prm(h, V(st->h3_low), D.h3l))) goto done;
    if ((rc = put_prm(h, V(st->h4), D.h4))) goto done;
    if ((rc = put_prm(h, V(st->alpha_h1), D.ah1))) goto done;
    if ((rc = put_prm(h, V(st->soil_thickness), D.d))) goto done;
    if ((rc = put_prm(h, icap_s, D.icap_s))) goto done;
    if ((rc = put_prm(h, V(st->infiltcap_compacted), D.icap_c))) goto done;
    if ((rc = put_prm(h, V(st->max_leakage), D.maxleak))) goto done;
    if ((rc = put_prm(h, V(st->compacted_fraction), D.compf))) goto done;
    if ((rc = put_prm(h, V(st->rootdistpar), D.rdp))) goto done;
    if ((rc = put_prm(h, V(st->cap_hmax), D.caph))) goto done;
    if ((rc = put_prm(h, V(st->cap_n), D.capn))) goto done;
    if ((rc = put_prm(h, V(st->w_soil), D.wsoil))) goto done;
    if ((rc = put_prm(h, V(st->cf_soil), D.cfsoil))) goto done;
    if ((rc = put_prm(h, rd, D.rd))) goto done;
    if ((rc = put_prml(h, std::vector<double>(st->brooks_corey_c, st->brooks_corey_c + (size_t)N * n), N, D.c)))
        goto done;
    if ((rc = put_prml(h, std::vector<double>(st->kvfrac, st->kvfrac + (size_t)N * n), N, D.kvfrac))) goto done;
    if ((rc = put_prml(h, rootfrac, N, D.rootfrac))) goto done;
    if ((rc = put_prm(h, area, D.area))) goto done;
    if ((rc = put_prm(h, fl, D.fl))) goto done;
    if ((rc = put_prm(h, fw, D.fw))) goto done;
    if ((rc = put_prm(h, slope, D.slope))) goto done;
    if ((rc = put_prm(h, sy, D.sy))) goto done;
    if ((rc = put_prm(h, kh0, D.kh0))) goto done;
    if ((rc = put_prm(h, qmax, D.qmax))) goto done;
    if ((rc = put_prm(h, alpha, D.alpha))) goto done;
    if ((rc = put_prm(h, sfw, D.sfw))) goto done;
    if ((rc = put_prm(h, V(st->elevation), D.elev))) goto done;
    {
        // topology
        std::vector<double> up_f(up_idx.size());
        for (size_t e = 0; e < up_idx.size(); ++e) up_f[e] = fftr[h->land_of[up_idx[e]]];
        if ((rc = put_raw<int>(h, up_ptr, D.up_ptr))) goto done;
        if ((rc = put_raw<int>(h, up_idx, D.up_idx))) goto done;
        {  // chain chunks of the land lists (run_chain): full, at most 4 upstream cells per lane,
           // and lane l's only upstream cell inside the chunk is lane l-1
            const int64_t nch = (n + WAVE - 1) / WAVE;
            std::vector<int> chain((size_t)nch, 0);
            for (int64_t c = 0; c < nch; ++c) {
                bool ok = (c + 1) * WAVE <= n;
                for (int l = 0; ok && l < WAVE; ++l) {
                    const int64_t q = c * WAVE + l;
                    int inside = 0;
                    ok = up_ptr[q + 1] - up_ptr[q] <= 4;
                    for (int e = up_ptr[q]; ok && e < up_ptr[q + 1]; ++e)
                        if (up_idx[e] / WAVE == c) {
                            ++inside;
                            ok = up_idx[e] == q - 1;
                        }
                    ok = ok && inside == (l > 0 ? 1 : 0);
                }
                chain[c] = ok ? 1 : 0;
            }
            if ((rc = put_raw<int>(h, chain, D.chain_land))) goto done;
            // the same for the river lists over the node order: every cell a river cell
            std::vector<int> chain_r((size_t)nch, 0);
            for (int64_t c = 0; c < nch; ++c) {
                bool ok = (c + 1) * WAVE <= n;
                for (int l = 0; ok && l < WAVE; ++l) {
                    const int64_t q = c * WAVE + l;
                    const int r = river_of[q];
                    ok = r >= 0;
                    if (!ok) break;
                    int inside = 0;
                    ok = rup_ptr[r + 1] - rup_ptr[r] <= 4;
                    for (int e = rup_ptr[r]; ok && e < rup_ptr[r + 1]; ++e)
                        if (rup_node[e] / WAVE == c) {
                            ++inside;
                            ok = rup_node[e] == q - 1;
                        }
                    ok = ok && inside == (l > 0 ? 1 : 0);
                }
                chain_r[c] = ok ? 1 : 0;
            }
            if ((rc = put_raw<int>(h, chain_r, D.chain_rl))) goto done;
        }
        if ((rc = put_raw<double>(h, up_f, D.up_f))) goto done;
        if ((rc = put_raw<int>(h, river_of, D.river_of))) goto done;
        if ((rc = put_raw<uint8_t>(h, flags, D.flags))) goto done;
        if ((rc = put_raw<int>(h, rup_ptr, D.rup_ptr))) goto done;
        if ((rc = put_raw<int>(h, rup_idx, D.rup_idx))) goto done;
        if ((rc = put_raw<int>(h, rup_node, D.rup_node))) goto done;
        if ((rc = put_raw<double>(h, rup_res, D.rup_res))) goto done;
        if ((rc = put_raw<int>(h, res_of_river, D.res_of_r))) goto done;
        h->fftr_land = fftr;
        // lateral snow transport: every upstream cell (the flow graph itself), ordered by the
        // reference's topological order (the order accucapacityflux! passes material on in)
        D.snow_lat = (cfg->snow_gravitational_transport && cfg->snow) ? 1 : 0;
        h->ext = nres > 0 || D.snow_lat || st->z_exp != nullptr;
        {
            std::vector<int> sup_ptr(n + 1, 0), sup_idx;
            if (D.snow_lat) {
                std::vector<int64_t> pos(n);
                for (int64_t k = 0; k < n; ++k) pos[h->land_order[k]] = k;
                std::vector<std::vector<int>> ups(n);
                for (int64_t i = 0; i < n; ++i)
                    if (h->land_down[i] >= 0) ups[h->land_down[i]].push_back((int)i);
                sup_idx.reserve(n);
                for (int64_t q = 0; q < n; ++q) {
                    std::vector<int>& u = ups[h->land_of[q]];
                    std::sort(u.begin(), u.end(), [&](int x, int y) { return pos[x] < pos[y]; });
                    for (int w : u) sup_idx.push_back(h->node_of[w]);
                    sup_ptr[q + 1] = (int)sup_idx.size();
                }
            }
            if ((rc = put_raw<int>(h, sup_ptr, D.sup_ptr))) goto done;
            if ((rc = put_raw<int>(h, sup_idx, D.sup_idx))) goto done;
            {  // chain chunks of these lists (run_chain), as for the land lists
                const int64_t nch = (n + WAVE - 1) / WAVE;
                std::vector<int> chain((size_t)nch, 0);
                for (int64_t c = 0; c < nch && D.snow_lat; ++c) {
                    bool ok = (c + 1) * WAVE <= n;
                    for (int l = 0; ok && l < WAVE; ++l) {
                        const int64_t q = c * WAVE + l;
                        int inside = 0;
                        ok = sup_ptr[q + 1] - sup_ptr[q] <= 4;
                        for (int e = sup_ptr[q]; ok && e < sup_ptr[q + 1]; ++e)
                            if (sup_idx[e] / WAVE == c) {
                                ++inside;
                                ok = sup_idx[e] == q - 1;
                            }
                        ok = ok && inside == (l > 0 ? 1 : 0);
                    }
                    chain[c] = ok ? 1 : 0;
                }
                if ((rc = put_raw<int>(h, chain, D.chain_sup))) goto done;
            }
            if ((rc = put_state(h, {}, D.snow_in))) goto done;
            if ((rc = put_state(h, {}, D.snow_out))) goto done;
        }
        // 1-D floodplain: the profile of every river cell (FloodPlainProfile  floodplain.jl:48-147),
        // Manning n, slope clamped at 1e-5 (:395-396), bankfull storage (surface_kinwave.jl:83);
        // the floodplain starts empty (FloodPlainVariables  floodplain.jl:319-341)
        D.fp = fp ? 1 : 0;
        if (fp) {
            const int nd = st->floodplain_ndepth + 1;
            D.fp_nd = nd;
            std::vector<double> depth(nd, 0.0), prof((size_t)nriv * 4 * nd, 0.0), seg(nd);
            for (int j = 1; j < nd; ++j) depth[j] = st->floodplain_depth[j - 1];
            for (int64_t r = 0; r < nriv; ++r) {
                double *S = &prof[(size_t)r * 4 * nd], *W = S + nd, *A = W + nd, *P = A + nd;
                const double fl = st->river_length[r];
                std::fill(seg.begin(), seg.end(), 0.0);
                W[0] = st->river_width[r];
                for (int j = 0; j < nd - 1; ++j) {
                    const double hseg = depth[j + 1] - depth[j];
                    const double v0 = j ? st->floodplain_volume[(size_t)(j - 1) * nriv + r] : 0.0;
                    W[j + 1] = (st->floodplain_volume[(size_t)j * nriv + r] - v0) / (hseg * fl);
                    if (W[j + 1] < W[j]) W[j + 1] = W[j];
                    A[j + 1] = W[j + 1] * hseg;
                    P[j + 1] = (W[j + 1] - W[j]) + 2.0 * hseg;
                    seg[j + 1] = A[j + 1] * fl;
                    if (j == 0) P[0] = P[1] - 2.0 * hseg;
                }
                for (int j = 2; j < nd; ++j) P[j] = P[j - 1] + P[j];
                for (int j = 1; j < nd; ++j) A[j] = A[j - 1] + A[j];
                S[0] = seg[0];
                for (int j = 1; j < nd; ++j) S[j] = S[j - 1] + seg[j];
            }
            std::vector<double> fpn(nriv), fps(nriv), bfs(nriv), bfd(nriv);
            for (int64_t r = 0; r < nriv; ++r) {
                fpn[r] = st->floodplain_manning_n[r];
                fps[r] = wfh_clamp(st->floodplain_slope[r], 0.00001, INFINITY);
                bfd[r] = st->river_bankfull_depth[r];
                bfs[r] = st->river_bankfull_depth[r] * st->river_width[r] * st->river_length[r];
            }
            if ((rc = put_raw<double>(h, depth, D.fp_depth))) goto done;
            if ((rc = put_raw<double>(h, prof, D.fp_prof))) goto done;
            if ((rc = put_raw<double>(h, fpn, D.fp_n))) goto done;
            if ((rc = put_raw<double>(h, fps, D.fp_slope))) goto done;
            if ((rc = put_raw<double>(h, bfs, D.r_bfs))) goto done;
            if ((rc = put_raw<double>(h, bfd, D.r_bfd))) goto done;
            const std::vector<double> zr((size_t)std::max<int64_t>(nriv, 1), 0.0);
            double** fo[] = {&D.fp_storage, &D.fp_h, &D.fp_q, &D.fp_qin, &D.fp_exch, &D.fp_q_cum, &D.fp_qin_cum,
                             &D.fp_q_avg, &D.fp_qin_avg, &D.r_qch_avg, &D.fp_rs, &D.fp_sprev};
            for (double** o : fo) {
                const double* tmp = nullptr;
                if ((rc = put_raw<double>(h, zr, tmp))) goto done;
                *o = const_cast<double*>(tmp);
            }
        }
        if ((rc = put_raw<int>(h, down_node, D.down))) goto done;
        if ((rc = put_raw<double>(h, r_alpha, D.r_alpha))) goto done;
        if ((rc = put_raw<double>(h, r_fw, D.r_fw))) goto done;
        if ((rc = put_raw<double>(h, r_fl, D.r_fl))) goto done;
        if ((rc = put_raw<double>(h, r_area, D.r_area))) goto done;
        const int* lo = nullptr;
        if ((rc = put_raw<int>(h, h->land_of, lo))) goto done;
        h->d_land_of = const_cast<int*>(lo);
    }
    {
        // cold-start states (soil.jl:150-201, canopy/snow zeros, routing zeros)
        std::vector<double> tsoil0(n, 10.0 + 273.15), ssf_storage0(n);
        for (int64_t i = 0; i < n; ++i) ssf_storage0[i] = sy[i] * (st->soil_thickness[i] - zi0[i]) * area[i];
        if ((rc = put_state(h, {}, D.canopy))) goto done;
        if ((rc = put_state(h, {}, D.snow_s))) goto done;
        if ((rc = put_state(h, {}, D.snow_w))) goto done;
        if ((rc = put_state(h, tsoil0, D.tsoil))) goto done;
        if ((rc = put_state(h, swd0, D.swd))) goto done;
        if ((rc = put_state(h, {}, D.ustore, N))) goto done;
        if ((rc = put_state(h, swd0, D.tsws))) goto done;
        if ((rc = put_state(h, q0, D.ssf_q))) goto done;
        if ((rc = put_state(h, {}, D.ol_q))) goto done;
        if ((rc = put_state(h, {}, D.ol_h))) goto done;
        if ((rc = put_state(h, {}, D.ol_storage))) goto done;
        if ((rc = put_river(h, {}, D.r_q))) goto done;
        if ((rc = put_river(h, {}, D.r_h))) goto done;
        if ((rc = put_river(h, {}, D.r_storage))) goto done;
        double** outs[] = {&D.zi, &D.interception, &D.swe, &D.transfer, &D.recharge, &D.act_inf, &D.inf_excess,
                           &D.sat_excess, &D.transpiration, &D.soil_evap, &D.capflux, &D.leakage, &D.aet,
                           &D.net_runoff_river, &D.runoff_land, &D.aeow_l, &D.store_prev, &D.ustoredepth, &D.runoff,
                           &D.net_runoff, &D.exfilt, &D.vwc_root, &D.rzs, &D.ssf_qin_avg, &D.ssf_q_avg,
                           &D.ssf_tor_avg, &D.ssf_storage, &D.ol_qin, &D.ol_q_avg, &D.ol_qin_avg, &D.ol_tor_avg,
                           &D.ol_inwater, &D.total_storage, &D.bal_land, &D.bal_land_rel, &D.bal_ssf, &D.bal_ol};
        for (double** o : outs)
            if ((rc = put_state(h, {}, *o))) goto done;
        if ((rc = put_state(h, zi0, D.zi))) goto done;
        if ((rc = put_state(h, ssf_storage0, D.ssf_storage))) goto done;
        if ((rc = put_state(h, {}, D.vwc, N))) goto done;
        double** routs[] = {&D.r_qin, &D.r_q_avg, &D.r_qin_avg, &D.r_inwater, &D.bal_r};
        for (double** o : routs)
            if ((rc = put_river(h, {}, *o))) goto done;
        // reservoirs (reservoir.jl:47-300): parameters, the linear storage curve's initial
        // storage (initialize_storage  :343-362) and each area's nodes in ascending land index
        {
            const size_t nr = (size_t)std::max<int64_t>(nres, 1);
            D.nres = (int)nres;
            auto res_vec = [&](const double* src, double dflt) {
                std::vector<double> v(nr, dflt);
                for (int64_t k = 0; k < nres; ++k) v[k] = src ? src[k] : dflt;
                return v;
            };
            const double nan = std::nan("");
            std::vector<int> rtype(nr, 4), cov_ptr(nres + 1, 0), cov_idx;
            for (int64_t k = 0; k < nres; ++k) rtype[k] = st->reservoir_outflow_type[k];
            for (int64_t i = 0; i < n; ++i)
                if (res_cov_id[i] > 0) ++cov_ptr[res_cov_id[i]];
            for (int64_t k = 0; k < nres; ++k) cov_ptr[k + 1] += cov_ptr[k];
            cov_idx.resize(nres > 0 ? cov_ptr[nres] : 0);
            {
                std::vector<int> fill(nr, 0);
                for (int64_t i = 0; i < n; ++i)
                    if (res_cov_id[i] > 0) {
                        const int k = res_cov_id[i] - 1;
                        cov_idx[cov_ptr[k] + fill[k]++] = h->node_of[i];
                    }
            }
            std::vector<double> area_r = res_vec(st->reservoir_area, 1.0), wl = res_vec(st->reservoir_waterlevel, 0.0);
            std::vector<double> sto(nr, 0.0), maxsto = res_vec(st->reservoir_maximum_storage, nan);
            // storage curves and rating curves (validated above), their extremes, the initial
            // storage (initialize_storage  reservoir.jl:342-362) and the maximum storage of a
            // rating-curve reservoir (get_maximum_storage  :324-340, set at :175)
            std::vector<int> sct(nr, 1), sh_ptr(nres + 1, 0), hq_ptr(nres + 1, 0);
            std::vector<double> sh_h, sh_s, hq_h, hq_q, shx(4 * nr, 0.0), hq_ext(732 * nr, 0.0);
            for (int64_t k = 0; k < nres; ++k) {
                sct[k] = st->reservoir_storage_curve_type ? st->reservoir_storage_curve_type[k] : 1;
                if (sct[k] == 2) {
                    const int64_t a = st->reservoir_sh_ptr[k], b = st->reservoir_sh_ptr[k + 1];
                    sh_h.insert(sh_h.end(), st->reservoir_sh_h + a, st->reservoir_sh_h + b);
                    sh_s.insert(sh_s.end(), st->reservoir_sh_s + a, st->reservoir_sh_s + b);
                    double* e = &shx[4 * k];
                    e[0] = e[1] = st->reservoir_sh_s[a];
                    e[2] = e[3] = st->reservoir_sh_h[a];
                    for (int64_t j = a + 1; j < b; ++j) {
                        e[0] = wfh_min(e[0], st->reservoir_sh_s[j]);
                        e[1] = wfh_max(e[1], st->reservoir_sh_s[j]);
                        e[2] = wfh_min(e[2], st->reservoir_sh_h[j]);
                        e[3] = wfh_max(e[3], st->reservoir_sh_h[j]);
                    }
                    sto[k] = interp_host(wl[k], st->reservoir_sh_h + a, st->reservoir_sh_s + a, b - a, 1);
                } else {
                    sto[k] = area_r[k] * wl[k];
                }
                sh_ptr[k + 1] = (int)sh_h.size();
                if (rtype[k] == 1) {
                    const int64_t a = st->reservoir_hq_ptr[k], b = st->reservoir_hq_ptr[k + 1];
                    hq_h.insert(hq_h.end(), st->reservoir_hq_h + a, st->reservoir_hq_h + b);
                    hq_q.insert(hq_q.end(), st->reservoir_hq_q + a * 365, st->reservoir_hq_q + b * 365);
                    double* e = &hq_ext[732 * k];
                    e[0] = e[1] = st->reservoir_hq_h[a];
                    for (int64_t j = a + 1; j < b; ++j) {
                        e[0] = wfh_min(e[0], st->reservoir_hq_h[j]);
                        e[1] = wfh_max(e[1], st->reservoir_hq_h[j]);
                    }
                    for (int c = 0; c < 365; ++c) {
                        double lo = st->reservoir_hq_q[a * 365 + c], hi = lo;
                        for (int64_t j = a + 1; j < b; ++j) {
                            lo = wfh_min(lo, st->reservoir_hq_q[j * 365 + c]);
                            hi = wfh_max(hi, st->reservoir_hq_q[j * 365 + c]);
                        }
                        e[2 + 2 * c] = lo;
                        e[3 + 2 * c] = hi;
                    }
                    maxsto[k] = sct[k] == 2 ? interp_host(e[1], st->reservoir_sh_h + st->reservoir_sh_ptr[k],
                                                          st->reservoir_sh_s + st->reservoir_sh_ptr[k],
                                                          st->reservoir_sh_ptr[k + 1] - st->reservoir_sh_ptr[k], 1)
                                            : area_r[k] * e[1];
                }
                hq_ptr[k + 1] = (int)hq_h.size();
            }
            if (sh_h.size() > (size_t)INT32_MAX / 2 || hq_h.size() > (size_t)INT32_MAX / 365) {
                rc = fail(h, WFH_ERR_ARG, "reservoir tables too large");
                goto done;
            }
            D.res_hq_col = 0;
            if ((rc = put_raw<int>(h, sct, D.res_sct))) goto done;
            if ((rc = put_raw<int>(h, sh_ptr, D.res_sh_ptr))) goto done;
            if ((rc = put_raw<int>(h, hq_ptr, D.res_hq_ptr))) goto done;
            if ((rc = put_raw<double>(h, sh_h, D.res_sh_h))) goto done;
            if ((rc = put_raw<double>(h, sh_s, D.res_sh_s))) goto done;
            if ((rc = put_raw<double>(h, hq_h, D.res_hq_h))) goto done;
            if ((rc = put_raw<double>(h, hq_q, D.res_hq_q))) goto done;
            if ((rc = put_raw<double>(h, shx, D.res_shx))) goto done;
            if ((rc = put_raw<double>(h, hq_ext, D.res_hq_ext))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_rating_curve_exponent, nan), D.res_rce))) goto done;
            D.res_linked = linked ? 1 : 0;
            if ((rc = put_raw<int>(h, res_lower, D.res_lower))) goto done;
            if ((rc = put_raw<int>(h, res_upper, D.res_upper))) goto done;
            {
                hipError_t ex;
                D.res_xfer = dalloc<uint64_t>(h, nr, ex);  // emptied once the stream exists
                if (ex != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc reservoir slots"); goto done; }
            }
            std::vector<double> zeros(nr, 0.0), zeros4(4 * nr, 0.0), nans(nr, nan);
            if ((rc = put_raw<int>(h, rtype, D.res_type))) goto done;
            if ((rc = put_raw<int>(h, cov_ptr, D.res_cov_ptr))) goto done;
            if ((rc = put_raw<int>(h, cov_idx, D.res_cov_idx))) goto done;
            if ((rc = put_raw<double>(h, area_r, D.res_area))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_demand, nan), D.res_demand))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_maximum_release, nan), D.res_maxrel))) goto done;
            if ((rc = put_raw<double>(h, maxsto, D.res_maxsto))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_target_full_fraction, nan), D.res_tff))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_target_minimum_fraction, nan), D.res_tmf))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_threshold, nan), D.res_thr))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_rating_curve_coefficient, nan), D.res_rcc))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_external_inflow, 0.0), D.res_ext))) goto done;
            if ((rc = put_raw<double>(h, res_vec(st->reservoir_outflow_obs, nan), D.res_obs))) goto done;
            double** rs[] = {&D.res_storage, &D.res_wl, &D.res_outflow, &D.res_P, &D.res_E, &D.res_outflow_avg,
                             &D.res_inflow_avg, &D.res_actevap_avg, &D.res_abs_avg, &D.res_store_prev, &D.bal_res};
            const std::vector<double>* init[] = {&sto, &wl, &nans, &zeros, &zeros, &zeros,
                                                 &zeros, &zeros, &zeros, &zeros, &zeros};
            for (size_t j = 0; j < sizeof(rs) / sizeof(rs[0]); ++j) {
                const double* tmp = nullptr;
                if ((rc = put_raw<double>(h, *init[j], tmp))) goto done;
                *rs[j] = const_cast<double*>(tmp);
            }
            const double* cum = nullptr;
            if ((rc = put_raw<double>(h, zeros4, cum))) goto done;
            D.res_cum = const_cast<double*>(cum);
        }
        hipError_t e;
        h->fP = dalloc<double>(h, n, e);
        h->fE = dalloc<double>(h, n, e);
        h->fT = dalloc<double>(h, n, e);
        if (e != hipSuccess) {
            rc = fail(h, WFH_ERR_HIP, "hipMalloc forcing");
            goto done;
        }
    }
    {
        // routing sub-steps (fixed kinematic-wave time steps, config_structure.jl:84-88); the
        // adaptive launches take one sub-step each
        if (adaptive) {
            h->S_s = h->S_l = h->S_r = 1;
            for (int k = 0; k < 2; ++k) h->dts_s[k] = h->dts_l[k] = h->dts_r[k] = cfg->dt;
        } else {
            h->S_s = sub_steps(cfg->dt, cfg->subsurface_dt, h->dts_s);
            h->S_l = sub_steps(cfg->dt, cfg->land_dt, h->dts_l);
            h->S_r = sub_steps(cfg->dt, cfg->river_dt, h->dts_r);
        }
        hipError_t e;
        RouteArgs& A = h->ra;
        A.slot_ssf = dalloc<uint64_t>(h, (size_t)h->S_s * n, e);
        if (e != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc slots"); goto done; }
        A.slot_ol = dalloc<uint64_t>(h, (size_t)h->S_l * n, e);
        if (e != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc slots"); goto done; }
        A.slot_r = dalloc<uint64_t>(h, (size_t)h->S_r * std::max<int64_t>(nriv, 1), e);
        if (e != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc slots"); goto done; }
        A.counters = dalloc<unsigned long long>(h, 2, e);
        A.abort_flag = dalloc<int>(h, 2, e);
        A.spin_limit = SPIN_LIMIT;
        h->d.abort_flag = A.abort_flag;
        h->d.nonfinite_flag = A.abort_flag + 1;
        if (e != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc counters"); goto done; }
        A.S_s = h->S_s;
        A.S_l = h->S_l;
        A.S_r = h->S_r;
        for (int k = 0; k < 2; ++k) {
            A.dts_s[k] = h->dts_s[k];
            A.dts_l[k] = h->dts_l[k];
            A.dts_r[k] = h->dts_r[k];
        }
        h->nchunks = (int)((n + WAVE - 1) / WAVE);
        A.nchunks = h->nchunks;
        if (hipStreamCreate(&h->stream) != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipStreamCreate"); goto done; }
        if ((rc = launch_fill(h, A.slot_ssf, SLOT_EMPTY, (size_t)h->S_s * n))) goto done;
        if ((rc = launch_fill(h, h->d.res_xfer, SLOT_EMPTY, (size_t)std::max<int64_t>(nres, 1)))) goto done;
        if ((rc = launch_fill(h, A.slot_ol, SLOT_EMPTY, (size_t)h->S_l * n))) goto done;
        if ((rc = launch_fill(h, A.slot_r, SLOT_EMPTY, (size_t)h->S_r * std::max<int64_t>(nriv, 1)))) goto done;
        if ((rc = launch_fill(h, (uint64_t*)A.counters, 0, 2))) goto done;
        if (hipMemsetAsync(A.abort_flag, 0, 2 * sizeof(int), h->stream) != hipSuccess) {
            rc = fail(h, WFH_ERR_HIP, "hipMemsetAsync");
            goto done;
        }
        // persistent grid: as many resident blocks as the device holds, at most one per 4 chunks
        int dev = cfg->device, cus = 0, per_cu = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        hipError_t oe = hipSuccess;
        switch (N) {
            case 2: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_route<2, false>, ROUTE_BLOCK, 0); break;
            case 3: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_route<3, false>, ROUTE_BLOCK, 0); break;
            case 4: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_route<4, false>, ROUTE_BLOCK, 0); break;
            case 5: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_route<5, false>, ROUTE_BLOCK, 0); break;
            default: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_route<6, false>, ROUTE_BLOCK, 0); break;
        }
        if (oe != hipSuccess || per_cu < 1) per_cu = 1;
        if (cus < 1) cus = 1;
        const int want = (h->nchunks + (ROUTE_BLOCK / WAVE) - 1) / (ROUTE_BLOCK / WAVE);
        h->route_grid = std::max(1, std::min(want, per_cu * cus));
        h->route_grid_max = h->route_grid;
        // One launch per step (the vertical update inside the routing kernel) for fixed
        // sub-steps at two routing waves per SIMD; lateral snow transport keeps the vertical
        // kernel, because the transport runs between the two halves of the step.
        h->fuse = !D.snow_lat && !adaptive && !fp && !linked && route_waves(N) == 2;
        if (adaptive || fp || linked) {
            // river positions: river cells sorted by node, a topological order of the river graph
            // because every river edge is a land edge and node order is topological
            AdArgs& X = h->ad;
            std::vector<std::pair<int, int>> nr(nriv);
            for (int64_t r = 0; r < nriv; ++r) nr[r] = {h->node_of[h->river_land[r]], (int)r};
            std::sort(nr.begin(), nr.end());
            std::vector<int> rp_r(nriv), rpos(nriv), rp_ptr(nriv + 1, 0), rp_idx, rp_down(nriv, -1), rp_node(nriv);
            std::vector<double> rp_res;
            for (int64_t k = 0; k < nriv; ++k) {
                rp_r[k] = nr[k].second;
                rpos[nr[k].second] = (int)k;
            }
            std::vector<std::vector<int>> ups(nriv);
            for (int64_t r = 0; r < nriv; ++r)
                if (h->river_down[r] >= 0) ups[h->river_down[r]].push_back((int)r);  // ascending river index
            for (int64_t k = 0; k < nriv; ++k) {
                const int r = rp_r[k];
                for (int u : ups[r]) {
                    if (rpos[u] >= k) {
                        rc = fail(h, WFH_ERR_CYCLE, "internal error: river positions are not topological");
                        goto done;
                    }
                    rp_idx.push_back(rpos[u]);
                    rp_res.push_back(res_of_river[u] >= 0 ? 1.0 : 0.0);
                }
                rp_node[k] = h->node_of[h->river_land[r]];
                rp_ptr[k + 1] = (int)rp_idx.size();
                rp_down[k] = h->river_down[r] >= 0 ? rpos[h->river_down[r]] : -1;
            }
            if ((rc = put_raw<int>(h, rp_r, X.rp_r))) goto done;
            if ((rc = put_raw<int>(h, rp_ptr, X.rp_up_ptr))) goto done;
            if ((rc = put_raw<int>(h, rp_idx, X.rp_up_idx))) goto done;
            if ((rc = put_raw<int>(h, rp_down, X.rp_down))) goto done;
            if ((rc = put_raw<double>(h, rp_res, X.rp_up_res))) goto done;
            if ((rc = put_raw<int>(h, rp_node, X.rp_node))) goto done;
            {  // floodplain accumulation: upstream positions ordered by the river topological order
                std::vector<int64_t> opos(nriv);
                for (int64_t k = 0; k < nriv; ++k) opos[h->river_order[k]] = k;
                std::vector<int> sp(nriv + 1, 0), si;
                for (int64_t k = 0; k < nriv && fp; ++k) {
                    std::vector<int>& u = ups[rp_r[k]];
                    std::vector<int> v(u.begin(), u.end());
                    std::sort(v.begin(), v.end(), [&](int a_, int b_) { return opos[a_] < opos[b_]; });
                    for (int w : v) si.push_back(rpos[w]);
                    sp[k + 1] = (int)si.size();
                }
                if ((rc = put_raw<int>(h, sp, X.fp_sup_ptr))) goto done;
                if ((rc = put_raw<int>(h, si, X.fp_sup_idx))) goto done;
            }
            X.nchunks_r = (int)((nriv + WAVE - 1) / WAVE);
            {  // chain chunks (run_chain): full, at most 4 upstream cells per lane, and lane l's only
               // upstream cell inside the chunk is lane l-1
                std::vector<int> chain((size_t)std::max(X.nchunks_r, 1), 0);
                for (int c = 0; c < X.nchunks_r; ++c) {
                    bool ok = (int64_t)(c + 1) * WAVE <= nriv;
                    for (int l = 0; ok && l < WAVE; ++l) {
                        const int p = c * WAVE + l;
                        int inside = 0;
                        ok = rp_ptr[p + 1] - rp_ptr[p] <= 4;
                        for (int e = rp_ptr[p]; ok && e < rp_ptr[p + 1]; ++e)
                            if (rp_idx[e] / WAVE == c) {
                                ++inside;
                                ok = rp_idx[e] == p - 1;
                            }
                        ok = ok && inside == (l > 0 ? 1 : 0);
                    }
                    chain[c] = ok ? 1 : 0;
                }
                if ((rc = put_raw<int>(h, chain, X.chain_r))) goto done;
            }
            hipError_t ea = hipSuccess, eb;
            X.szi = dalloc<double>(h, n, eb); if (eb != hipSuccess) ea = eb;
            X.scum = dalloc<double>(h, 6 * (size_t)n, eb); if (eb != hipSuccess) ea = eb;
            X.ocum = dalloc<double>(h, 3 * (size_t)n, eb); if (eb != hipSuccess) ea = eb;
            X.rcum = dalloc<double>(h, 2 * (size_t)std::max<int64_t>(nriv, 1), eb); if (eb != hipSuccess) ea = eb;
            h->keys = dalloc<uint64_t>(h, (size_t)std::max<int64_t>(n, nriv), eb); if (eb != hipSuccess) ea = eb;
            h->sel = dalloc<SelCtl>(h, 1, eb); if (eb != hipSuccess) ea = eb;
            if (ea != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "hipMalloc adaptive sub-step state"); goto done; }
            if (hipHostMalloc((void**)&h->sel_host, sizeof(SelCtl), hipHostMallocDefault) != hipSuccess) {
                h->sel_host = nullptr;
                rc = fail(h, WFH_ERR_HIP, "hipHostMalloc");
                goto done;
            }
        }
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = fail(h, WFH_ERR_HIP, "init sync"); goto done; }
    }
    D.P = h->fP;
    D.E = h->fE;
    D.T = h->fT;
    return WFH_OK;
done:
    wfh_destroy_keep_error(h, out);
    return rc;
}

constexpr size_t MAX_TIMED_STEPS = 100000;

static hipEvent_t* timing_events(wfh_handle* h) {
    if (!h->timing || h->timed_steps >= MAX_TIMED_STEPS) return nullptr;
    const size_t need = 3 * (h->timed_steps + 1);
    while (h->ev.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->ev.push_back(e);
    }
    return &h->ev[3 * h->timed_steps++];
}

// lateral_snow_transport! and flux_in! (between the vertical update and the routing)
static void launch_snow_transport(wfh_handle* h, const Dev& d, RouteArgs a, int& parity) {
    a.parity = parity;
    parity ^= 1;
    hipLaunchKernelGGL(k_snow_transport, dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, d, a);
    hipLaunchKernelGGL(k_snow_in, dim3((unsigned)((h->n + 255) / 256)), dim3(256), 0, h->stream, d);
}
static void launch_snow_transport(wfh_handle* h) { launch_snow_transport(h, h->d, h->ra, h->parity); }

// The kernels of one fixed-sub-step step on the handle's stream, with the routing scratch
// (slots, tickets, abort flag) that `a` and `d.abort_flag` point to; `parity` selects the ticket
// counter of the first dataflow launch and is advanced per launch.
template <int N>
static void launch_step(wfh_handle* h, const Dev& d, const RouteArgs& base, int& parity, hipEvent_t* ev) {
    const int n = (int)h->n;
    if (ev) hipEventRecord(ev[0], h->stream);
    if (h->nres > 0)
        hipLaunchKernelGGL(k_res_forcing, dim3((unsigned)((h->nres + 63) / 64)), dim3(64), 0, h->stream, d);
    if (!h->fuse) hipLaunchKernelGGL(k_vertical<N>, dim3((n + 255) / 256), dim3(256), 0, h->stream, d);
    if (ev) hipEventRecord(ev[1], h->stream);
    if (d.snow_lat) launch_snow_transport(h, d, base, parity);
    RouteArgs a = base;
    a.parity = parity;
    parity ^= 1;
    bool fused = false;
    if constexpr (route_waves(N) == 2) {
        if (h->fuse && h->ext)
            hipLaunchKernelGGL((k_route<N, true, false, true>), dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream,
                               d, a);
        else if (h->fuse)
            hipLaunchKernelGGL((k_route<N, false, false, true>), dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream,
                               d, a);
        fused = h->fuse;
    }
    if (fused) {
    } else if (h->ext) {
        hipLaunchKernelGGL((k_route<N, true>), dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, d, a);
    } else {
        hipLaunchKernelGGL((k_route<N, false>), dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, d, a);
    }
    if (ev) hipEventRecord(ev[2], h->stream);
}

template <int N>
static void launch_kernels(wfh_handle* h) {
    launch_step<N>(h, h->d, h->ra, h->parity, timing_events(h));
}

// ---- adaptive sub-steps ----
constexpr int MAX_ADAPTIVE_SUBSTEPS = 1000000;  // per family and step: a guard, the reference has none

static int32_t read_sel(wfh_handle* h) {
    HIPCHK(h, hipMemcpyAsync(h->sel_host, h->sel, sizeof(SelCtl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int flags[2] = {0, 0};
    HIPCHK(h, hipMemcpy(flags, h->ra.abort_flag, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (flags[1]) {
        h->broken = h->broken_nonfinite = true;
        return fail(h, WFH_ERR_NONFINITE,
                    "non-finite forcing (precipitation, potential evaporation or temperature) in a step; the "
                    "states are not usable: load a saved state to continue");
    }
    const int flag = flags[0];
    if (flag) {
        h->broken = true;
        return fail(h, WFH_ERR_DEVICE, "routing kernel aborted: a dependency wait exceeded its spin limit");
    }
    return WFH_OK;
}

// stable_timestep of the subsurface model  lateral_subsurface_flow.jl:314-344
static int32_t ssf_stable_timestep(wfh_handle* h, double& dts) {
    hipLaunchKernelGGL(k_sel_reset, dim3(1), dim3(256), 0, h->stream, h->sel);
    hipLaunchKernelGGL(k_ssf_min, dim3((unsigned)std::min<int64_t>((h->n + 255) / 256, 2048)), dim3(256), 0, h->stream,
                       h->d, h->ad, h->sel);
    HIPCHK(h, hipGetLastError());
    int32_t rc = read_sel(h);
    if (rc) return rc;
    const double dt_min = h->sel_host->count > 0 ? wfh_from_bits(h->sel_host->minkey) : 0.5;
    dts = dt_min * h->cfg.subsurface_alpha_coefficient;
    return WFH_OK;
}

// stable_timestep of a kinematic-wave model over m cells  surface_kinwave.jl:674-704: the
// quantile's two order statistics come from an 8-pass radix selection on the device.
static int32_t kw_stable_timestep(wfh_handle* h, const double* q, Prm alpha, Prm flen, int m, double p, double& dts) {
    if (m == 0) {
        dts = 600.0;
        return WFH_OK;
    }
    const unsigned g1 = (unsigned)((m + 255) / 256);
    const unsigned gs = std::min<unsigned>(g1, 1024u);
    hipLaunchKernelGGL(k_sel_reset, dim3(1), dim3(256), 0, h->stream, h->sel);
    hipLaunchKernelGGL(k_kw_keys, dim3(g1), dim3(256), 0, h->stream, q, alpha, flen, m, h->keys, h->sel);
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(1), 0, h->stream, h->sel, p);
    for (int shift = 56; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(k_sel_hist, dim3(gs), dim3(256), 0, h->stream, (const uint64_t*)h->keys, m, h->sel, shift);
        hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(256), 0, h->stream, h->sel, shift);
    }
    hipLaunchKernelGGL(k_sel_next, dim3(gs), dim3(256), 0, h->stream, (const uint64_t*)h->keys, m, h->sel);
    HIPCHK(h, hipGetLastError());
    int32_t rc = read_sel(h);
    if (rc) return rc;
    const SelCtl& c = *h->sel_host;
    if (c.count == 0) {
        dts = 600.0;
    } else if (c.count == 1) {
        dts = wfh_from_bits(c.prefix);
    } else {
        double g;
        wfh_quantile_index((int64_t)c.count, p, &g);
        const double a = wfh_from_bits(c.prefix);
        const double b = c.rank + 1 < c.bin_count ? a : wfh_from_bits(c.next);
        dts = wfh_quantile_interp(a, b, g);
    }
    return WFH_OK;
}

// check_timestepsize  routing/timestepping.jl:11-17
static double check_timestepsize(double ts, double t, double end) { return t + ts > end ? end - t : ts; }

template <int N>
static int32_t launch_adaptive(wfh_handle* h) {
    const int n = (int)h->n, nriv = (int)h->nriv;
    const double dt = h->cfg.dt;
    hipEvent_t* ev = timing_events(h);
    if (ev) hipEventRecord(ev[0], h->stream);
    if (h->nres > 0)
        hipLaunchKernelGGL(k_res_forcing, dim3((unsigned)((h->nres + 63) / 64)), dim3(64), 0, h->stream, h->d);
    hipLaunchKernelGGL(k_vertical<N>, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d);
    if (ev) hipEventRecord(ev[1], h->stream);
    if (h->d.snow_lat) launch_snow_transport(h);
    const int nch_r = h->ad.nchunks_r;
    const int grid_r = std::max(1, std::min(h->route_grid, (nch_r + (ROUTE_BLOCK / WAVE) - 1) / (ROUTE_BLOCK / WAVE)));
    auto args = [&]() {
        RouteArgs a = h->ra;
        a.parity = h->parity;
        h->parity ^= 1;
        return a;
    };
    int32_t rc;
    // subsurface: the water table handed over by the soil (sbm_model.jl:74-81)
    HIPCHK(h, hipMemcpyAsync(h->ad.szi, h->d.zi, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    for (int fam = 0; fam < 3; ++fam) {
        double t = 0.0;
        int k = 0;
        while (t < dt) {
            double dts;
            rc = WFH_OK;
            if (!h->cfg.kinematic_wave_adaptive)  // fixed sub-steps (a floodplain model)
                dts = fam == 0 ? h->cfg.subsurface_dt : fam == 1 ? h->cfg.land_dt : h->cfg.river_dt;
            else if (fam == 0)
                rc = ssf_stable_timestep(h, dts);
            else if (fam == 1)
                rc = kw_stable_timestep(h, h->d.ol_q, h->d.alpha, h->d.fl, n, 0.02, dts);
            else
                rc = kw_stable_timestep(h, h->d.r_q, Prm{h->d.r_alpha, 0xFFFFFFFFu}, Prm{h->d.r_fl, 0xFFFFFFFFu}, nriv,
                                        0.05, dts);
            if (rc) return rc;
            dts = check_timestepsize(dts, t, dt);
            if (!(dts > 0.0) || ++k > MAX_ADAPTIVE_SUBSTEPS)
                return fail(h, WFH_ERR_STATE, "adaptive sub-step length not positive or too many sub-steps");
            const double tn = t + dts;
            const int first = k == 1, last = !(tn < dt);
            if (fam == 0)
                hipLaunchKernelGGL(k_ad_ssf<N>, dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, h->d, args(), h->ad,
                                   dts, first, last);
            else if (fam == 1)
                hipLaunchKernelGGL(k_ad_ol, dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, h->d, args(), h->ad,
                                   dts, first, last);
            else if (nriv > 0) {
                const bool res = h->nres > 0, fp = h->d.fp != 0, link = h->d.res_linked != 0;
                auto kr = res ? (link ? (fp ? k_ad_riv<true, true, true> : k_ad_riv<true, false, true>)
                                      : (fp ? k_ad_riv<true, true, false> : k_ad_riv<true, false, false>))
                              : (fp ? k_ad_riv<false, true, false> : k_ad_riv<false, false, false>);
                hipLaunchKernelGGL(kr, dim3(grid_r), dim3(ROUTE_BLOCK), 0, h->stream, h->d, args(), h->ad, dts, first,
                                   last);
                if (fp)  // flux_in! of the floodplain flow this sub-step produced
                    hipLaunchKernelGGL(k_fp_qin, dim3((unsigned)((nriv + 255) / 256)), dim3(256), 0, h->stream, h->d,
                                       h->ad, dts);
            }
            HIPCHK(h, hipGetLastError());
            t = tn;
        }
        h->nsub[fam] = k;
    }
    if (h->d.fp && nriv > 0)
        hipLaunchKernelGGL(k_fp_finish, dim3((unsigned)((nriv + 255) / 256)), dim3(256), 0, h->stream, h->d);
    hipLaunchKernelGGL(k_ad_total, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d);
    if (ev) hipEventRecord(ev[2], h->stream);
    HIPCHK(h, hipGetLastError());
    return read_sel(h);
}

extern "C" int32_t wfh_set_forcing(wfh_handle* h, const double* P, const double* E, const double* T,
                                   int32_t memkind) {
    if (!h) return WFH_ERR_ARG;
    if (!P || !E || !T) return fail(h, WFH_ERR_ARG, "null forcing pointer");
    const int n = (int)h->n;
    if (memkind == WFH_MEM_HOST) {
        // The upload overlaps the previous step's kernels: it goes to a buffer of its own on the
        // copy stream, which first waits until the previous step has gathered from that buffer.
        // The call returns once the copies are done, so the caller may reuse its arrays.
        if (!h->hf_tmp) {
            hipError_t e;
            h->hf_tmp = dalloc<double>(h, (size_t)3 * n, e);
            if (e != hipSuccess) {
                h->hf_tmp = nullptr;
                return fail(h, WFH_ERR_HIP, "hipMalloc host-forcing buffer");
            }
            HIPCHK(h, hipStreamCreate(&h->copy_stream));
            HIPCHK(h, hipEventCreateWithFlags(&h->hf_copied, hipEventDisableTiming));
            HIPCHK(h, hipEventCreateWithFlags(&h->hf_gathered, hipEventDisableTiming));
            HIPCHK(h, hipEventRecord(h->hf_gathered, h->stream));
        }
        const double* src[3] = {P, E, T};
        HIPCHK(h, hipStreamWaitEvent(h->copy_stream, h->hf_gathered, 0));
        for (int k = 0; k < 3; ++k)
            HIPCHK(h, hipMemcpyAsync(h->hf_tmp + (size_t)k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice,
                                     h->copy_stream));
        HIPCHK(h, hipEventRecord(h->hf_copied, h->copy_stream));
        HIPCHK(h, hipStreamSynchronize(h->copy_stream));  // the caller's arrays are free again
        HIPCHK(h, hipStreamWaitEvent(h->stream, h->hf_copied, 0));
        const int g = (n + 255) / 256;
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fP, h->hf_tmp, h->d_land_of, n);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fE, h->hf_tmp + n, h->d_land_of, n);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fT, h->hf_tmp + 2 * (size_t)n, h->d_land_of,
                           n);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(h->hf_gathered, h->stream));
        h->d.P = h->fP;
        h->d.E = h->fE;
        h->d.T = h->fT;
    } else if (memkind == WFH_MEM_DEVICE) {
        const int g = (n + 255) / 256;
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fP, P, h->d_land_of, n);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fE, E, h->d_land_of, n);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(256), 0, h->stream, h->fT, T, h->d_land_of, n);
        HIPCHK(h, hipGetLastError());
        h->d.P = h->fP;
        h->d.E = h->fE;
        h->d.T = h->fT;
    } else if (memkind == WFH_MEM_DEVICE_NODE_ORDER && h->graph_mode) {
        HIPCHK(h, hipMemcpyAsync(h->fP, P, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->fE, E, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->fT, T, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        h->d.P = h->fP;
        h->d.E = h->fE;
        h->d.T = h->fT;
    } else if (memkind == WFH_MEM_DEVICE_NODE_ORDER) {
        h->d.P = P;
        h->d.E = E;
        h->d.T = T;
    } else {
        return fail(h, WFH_ERR_ARG, "bad memkind");
    }
    h->forcing_set = true;
    return WFH_OK;
}

extern "C" int32_t wfh_step(wfh_handle* h) {
    if (!h) return WFH_ERR_ARG;
    if (h->broken) return fail(h, WFH_ERR_STATE, "handle is in an error state (a previous step failed)");
    if (!h->forcing_set) return fail(h, WFH_ERR_STATE, "wfh_set_forcing must be called before wfh_step");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    if (h->cfg.kinematic_wave_adaptive || h->d.fp || h->d.res_linked) {  // one launch per sub-step
        switch (h->N) {
            case 2: return launch_adaptive<2>(h);
            case 3: return launch_adaptive<3>(h);
            case 4: return launch_adaptive<4>(h);
            case 5: return launch_adaptive<5>(h);
            default: return launch_adaptive<6>(h);
        }
    }
    const bool tuning = h->tune_left > 0;
    if (tuning) {
        h->route_grid = h->tune_grids[h->tune_next];
        HIPCHK(h, hipEventRecord(h->tune_ev[0], h->stream));
    }
    if (h->graph_mode && !tuning) {
        const int p0 = h->parity;
        if (!h->graph_exec[p0]) {
            int p = p0;
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            switch (h->N) {
                case 2: launch_step<2>(h, h->d, h->ra, p, nullptr); break;
                case 3: launch_step<3>(h, h->d, h->ra, p, nullptr); break;
                case 4: launch_step<4>(h, h->d, h->ra, p, nullptr); break;
                case 5: launch_step<5>(h, h->d, h->ra, p, nullptr); break;
                default: launch_step<6>(h, h->d, h->ra, p, nullptr); break;
            }
            HIPCHK(h, hipStreamEndCapture(h->stream, &h->graph[p0]));
            HIPCHK(h, hipGraphInstantiate(&h->graph_exec[p0], h->graph[p0], nullptr, nullptr, 0));
        }
        HIPCHK(h, hipGraphLaunch(h->graph_exec[p0], h->stream));
        h->parity ^= h->d.snow_lat ? 0 : 1;  // as launch_step advances it (one or two dataflow launches)
    } else {
        switch (h->N) {
            case 2: launch_kernels<2>(h); break;
            case 3: launch_kernels<3>(h); break;
            case 4: launch_kernels<4>(h); break;
            case 5: launch_kernels<5>(h); break;
            default: launch_kernels<6>(h); break;
        }
    }
    HIPCHK(h, hipGetLastError());
    if (tuning) {
        HIPCHK(h, hipEventRecord(h->tune_ev[1], h->stream));
        HIPCHK(h, hipEventSynchronize(h->tune_ev[1]));
        float ms = 0.0f;
        HIPCHK(h, hipEventElapsedTime(&ms, h->tune_ev[0], h->tune_ev[1]));
        double& best = h->tune_ms[h->tune_next];
        best = best < 0.0 ? ms : std::min(best, (double)ms);
        h->tune_next = (h->tune_next + 1) % (int)h->tune_grids.size();
        if (--h->tune_left == 0) {
            int b = 0;
            for (int k = 1; k < (int)h->tune_grids.size(); ++k)
                if (h->tune_ms[k] < h->tune_ms[b]) b = k;
            h->route_grid = h->tune_grids[b];
        }
    }
    h->nsub[0] = h->S_s;
    h->nsub[1] = h->S_l;
    h->nsub[2] = h->S_r;
    return WFH_OK;
}

static void drop_graphs(wfh_handle* h) {
    for (int k = 0; k < 2; ++k) {
        if (h->graph_exec[k]) hipGraphExecDestroy(h->graph_exec[k]);
        if (h->graph[k]) hipGraphDestroy(h->graph[k]);
        h->graph_exec[k] = nullptr;
        h->graph[k] = nullptr;
    }
}

// Graph mode (fixed sub-steps): the step's launches are recorded once per starting ticket parity
// and replayed by every following wfh_step. Forcing bound in node order is then copied into the
// library's buffers (a graph keeps the pointers it was recorded with).
extern "C" int32_t wfh_set_graph(wfh_handle* h, int32_t enable) {
    if (!h) return WFH_ERR_ARG;
    if (enable && (h->cfg.kinematic_wave_adaptive || h->d.fp || h->d.res_linked))
        return fail(h, WFH_ERR_ARG, "graph mode covers fixed sub-steps only (no adaptive, floodplain or linked weirs)");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graphs(h);
    h->graph_mode = enable != 0;
    // forcing bound in node order is not what a graph reads (its own buffers): it must be set again
    if (h->graph_mode && h->forcing_set && (h->d.P != h->fP || h->d.E != h->fE || h->d.T != h->fT))
        h->forcing_set = false;
    return WFH_OK;
}

// Autotuner of the routing kernel's persistent grid: the next `trials` x candidates steps run the
// candidate grids (all resident blocks, 3/4, 1/2, 1/4) in turn, each step timed; the fastest is
// kept. Every grid computes the same results (the schedule changes which wave takes which chunk,
// not the operations). trials = 0 stops tuning and keeps the current grid.
extern "C" int32_t wfh_autotune(wfh_handle* h, int32_t trials) {
    if (!h) return WFH_ERR_ARG;
    if (trials < 0) return fail(h, WFH_ERR_ARG, "trials must be >= 0");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (!h->tune_ev[0]) {
        HIPCHK(h, hipEventCreate(&h->tune_ev[0]));
        HIPCHK(h, hipEventCreate(&h->tune_ev[1]));
    }
    drop_graphs(h);
    const int g = h->route_grid_max;
    h->tune_grids.clear();
    for (int f : {4, 3, 2, 1}) {
        const int gg = std::max(1, g * f / 4);
        if (std::find(h->tune_grids.begin(), h->tune_grids.end(), gg) == h->tune_grids.end()) h->tune_grids.push_back(gg);
    }
    h->tune_ms.assign(h->tune_grids.size(), -1.0);
    h->tune_next = 0;
    h->tune_left = trials * (int)h->tune_grids.size();
    return WFH_OK;
}

// The routing grid in use (blocks of 256 threads) and whether the autotuner is still running.
extern "C" int32_t wfh_get_schedule(wfh_handle* h, int32_t* route_blocks, int32_t* max_blocks, int32_t* tuning,
                                    int32_t* graph_mode) {
    if (!h) return WFH_ERR_ARG;
    if (route_blocks) *route_blocks = h->route_grid;
    if (max_blocks) *max_blocks = h->route_grid_max;
    if (tuning) *tuning = h->tune_left > 0;
    if (graph_mode) *graph_mode = h->graph_mode;
    return WFH_OK;
}

// ---- multi-step run (wfh_run) ----
// __global__ fill of the per-chunk step counters (a kernel, not a memset, so that the run can be
// recorded into a graph by a caller)
__global__ void k_fill_i32(int* p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int N>
static void launch_run(wfh_handle* h, RouteArgs a) {
    a.parity = h->parity;
    h->parity ^= 1;
    hipLaunchKernelGGL((k_route<N, false, true>), dim3(h->route_grid), dim3(ROUTE_BLOCK), 0, h->stream, h->d, a);
}

// nsteps steps of update_model! in one launch (the time loop of run!, Wflow.jl:324-372, for the
// forcing frames given): bitwise what nsteps calls of wfh_set_forcing + wfh_step give, with the
// steps overlapping on the device. Fixed sub-steps, no reservoirs, lateral snow transport,
// exponential-constant profile, floodplain or adaptive sub-steps.
// Switched off: the first device run of this kernel (a 48 x 48 catchment, 12 steps) did not finish
// within minutes, and its cause is not established (the ticket and wait protocol is deadlock-free
// in its CPU emulation, and every wait is bounded). wfh_run reports WFH_ERR_STATE until the cause
// is found; the kernel is kept for that work.
constexpr bool RUN_ENABLED = false;

extern "C" int32_t wfh_run(wfh_handle* h, int32_t nsteps, const double* P, const double* E, const double* T,
                           double* river_q_avg) {
    if (!h) return WFH_ERR_ARG;
    if (!RUN_ENABLED)
        return fail(h, WFH_ERR_STATE, "wfh_run is switched off: its device run did not finish and the cause is "
                                      "not established; use wfh_step per step");
    if (h->broken) return fail(h, WFH_ERR_STATE, "handle is in an error state (a previous step failed)");
    if (nsteps < 1 || !P || !E || !T) return fail(h, WFH_ERR_ARG, "nsteps >= 1 and forcing frames required");
    if (h->cfg.kinematic_wave_adaptive || h->d.fp || h->d.res_linked || h->ext)
        return fail(h, WFH_ERR_ARG,
                    "wfh_run covers fixed sub-steps without reservoirs, lateral snow transport, the "
                    "exponential-constant profile or a floodplain");
    if ((unsigned long long)nsteps * (unsigned long long)h->nchunks > (1ull << 40))
        return fail(h, WFH_ERR_ARG, "too many steps for one run");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    if (!h->run_chunk_done) {
        void* p = nullptr;
        HIPCHK(h, hipMalloc(&p, (size_t)h->nchunks * sizeof(int)));
        h->allocs.push_back(p);
        h->run_chunk_done = (int*)p;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3((h->nchunks + 255) / 256), dim3(256), 0, h->stream, h->run_chunk_done, 0,
                       h->nchunks);
    RouteArgs a = h->ra;
    a.nsteps = nsteps;
    a.chunk_done = h->run_chunk_done;
    a.run_P = P;
    a.run_E = E;
    a.run_T = T;
    a.run_q_avg = river_q_avg;
    hipEvent_t* ev = timing_events(h);
    if (ev) HIPCHK(h, hipEventRecord(ev[0], h->stream));
    if (ev) HIPCHK(h, hipEventRecord(ev[1], h->stream));
    switch (h->N) {
        case 2: launch_run<2>(h, a); break;
        case 3: launch_run<3>(h, a); break;
        case 4: launch_run<4>(h, a); break;
        case 5: launch_run<5>(h, a); break;
        default: launch_run<6>(h, a); break;
    }
    HIPCHK(h, hipGetLastError());
    if (ev) HIPCHK(h, hipEventRecord(ev[2], h->stream));
    return WFH_OK;
}

// Testing only: the routing kernel's spin limit, and optionally a wait that can never be met. The
// first upstream edge of cell `node` (node order) is redirected to the outflow slot of a cell in
// another chunk that publishes nothing, so the cell's wave idles until the limit, raises the abort
// flag, and every other wave leaves the kernel; wfh_sync then reports WFH_ERR_DEVICE. The edge
// stays redirected: the handle is for destroying afterwards.
extern "C" int32_t wfh_test_route_stall(wfh_handle* h, int64_t spin_limit, int64_t node) {
    if (!h) return WFH_ERR_ARG;
    if (spin_limit < 256) return fail(h, WFH_ERR_ARG, "spin_limit must be >= 256");
    if (node >= h->n) return fail(h, WFH_ERR_ARG, "node out of range");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graphs(h);
    h->ra.spin_limit = spin_limit;
    if (node < 0) return WFH_OK;
    const int64_t n = h->n;
    std::vector<int> up_ptr(n + 1), down(n);
    std::vector<uint8_t> flags(n);
    HIPCHK(h, hipMemcpy(up_ptr.data(), h->d.up_ptr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(down.data(), h->d.down, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(flags.data(), h->d.flags, n * sizeof(uint8_t), hipMemcpyDeviceToHost));
    const int e0 = up_ptr[node];
    if (e0 == up_ptr[node + 1]) return fail(h, WFH_ERR_ARG, "the cell has no upstream cell");
    int from = -1;  // a cell of another chunk whose land outflow goes to no slot
    for (int64_t q = 0; q < n && from < 0; ++q) {
        const bool cons_in = down[q] >= 0 && (down[q] / WAVE) == (int)(q / WAVE);
        if (q / WAVE != node / WAVE && (!(flags[q] & 1) || cons_in)) from = (int)q;
    }
    if (from < 0) return fail(h, WFH_ERR_ARG, "no cell of another chunk leaves its slot empty");
    HIPCHK(h, hipMemcpy(const_cast<int*>(h->d.up_idx) + e0, &from, sizeof(int), hipMemcpyHostToDevice));
    return WFH_OK;
}

// ---- allocation-free step (wfh_step_noalloc) ----
// The routing scratch of the fixed-sub-step step (ticket counters, abort flag, one outflow slot per
// cell and sub-step) taken from a fixed device array instead of the handle's allocations. Every
// GPU the library runs on holds its own copy for the life of the process; its device address is
// looked up once per device (hipGetSymbolAddress) and only that address is used from the host.
// Layout in 64-bit words: [0], [1] ticket counters, [2] abort flag (low 32 bits), [3..31] unused,
// then slot_ssf [S_s][n], slot_ol [S_l][n], slot_r [S_r][max(nriv, 1)].
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr size_t NOALLOC_HEAD_WORDS = 32;
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_BYTES / sizeof(uint64_t)];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Every call starts from a cleared scratch: the previous call (of any handle on this GPU) leaves
// its counters behind, and a call that aborted may leave slots filled.
__global__ void k_scratch_init(uint64_t* s, size_t words) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        s[i] = i < NOALLOC_HEAD_WORDS ? 0ull : SLOT_EMPTY;
}
// hands an abort raised in the fixed scratch on to the handle's own flag, which wfh_sync reads
__global__ void k_flag_or(int* dst, const int* src) {
    if (*src) *dst = 1;
}
}  // namespace

extern "C" int32_t wfh_step_noalloc(wfh_handle* h) {
    if (!h) return WFH_ERR_ARG;
    if (h->broken) return fail(h, WFH_ERR_STATE, "handle is in an error state (a previous step failed)");
    if (!h->forcing_set) return fail(h, WFH_ERR_STATE, "wfh_set_forcing must be called before wfh_step_noalloc");
    if (h->cfg.kinematic_wave_adaptive || h->d.fp || h->d.res_linked)
        return fail(h, WFH_ERR_ARG,
                    "wfh_step_noalloc covers fixed sub-steps only (no adaptive sub-steps, floodplain or linked "
                    "free weirs)");
    const int dev = h->cfg.device;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(h, WFH_ERR_ARG, "device ordinal out of range");
    const size_t nslot = (size_t)(h->S_s + h->S_l) * (size_t)h->n + (size_t)h->S_r * (size_t)std::max<int64_t>(h->nriv, 1);
    const size_t words = NOALLOC_HEAD_WORDS + nslot;
    if (words * sizeof(uint64_t) > NOALLOC_SCRATCH_BYTES)
        return fail(h, WFH_ERR_ARG,
                    "wfh_step_noalloc: the model needs " + std::to_string(words * sizeof(uint64_t)) +
                        " bytes of routing scratch, above the fixed limit of " +
                        std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes; use wfh_step");
    HIPCHK(h, hipSetDevice(dev));
    uint64_t* s = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) return fail(h, WFH_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
            g_scratch_ptr[dev] = (uint64_t*)p;
        }
        s = g_scratch_ptr[dev];
    }
    const int g = (int)std::min<size_t>((words + 255) / 256, 8192);
    hipLaunchKernelGGL(k_scratch_init, dim3(g), dim3(256), 0, h->stream, s, words);
    RouteArgs a = h->ra;
    a.counters = (unsigned long long*)s;
    a.abort_flag = (int*)(s + 2);
    a.slot_ssf = s + NOALLOC_HEAD_WORDS;
    a.slot_ol = a.slot_ssf + (size_t)h->S_s * h->n;
    a.slot_r = a.slot_ol + (size_t)h->S_l * h->n;
    Dev d = h->d;
    d.abort_flag = a.abort_flag;
    int parity = 0;
    switch (h->N) {
        case 2: launch_step<2>(h, d, a, parity, nullptr); break;
        case 3: launch_step<3>(h, d, a, parity, nullptr); break;
        case 4: launch_step<4>(h, d, a, parity, nullptr); break;
        case 5: launch_step<5>(h, d, a, parity, nullptr); break;
        default: launch_step<6>(h, d, a, parity, nullptr); break;
    }
    hipLaunchKernelGGL(k_flag_or, dim3(1), dim3(1), 0, h->stream, h->ra.abort_flag, (const int*)a.abort_flag);
    HIPCHK(h, hipGetLastError());
    h->nsub[0] = h->S_s;
    h->nsub[1] = h->S_l;
    h->nsub[2] = h->S_r;
    return WFH_OK;
}

extern "C" int32_t wfh_last_substeps(wfh_handle* h, int32_t* subsurface, int32_t* land, int32_t* river) {
    if (!h) return WFH_ERR_ARG;
    if (subsurface) *subsurface = h->nsub[0];
    if (land) *land = h->nsub[1];
    if (river) *river = h->nsub[2];
    return WFH_OK;
}

// The column is a kernel argument (part of Dev), so it applies to steps launched after the call.
extern "C" int32_t wfh_set_julian_day(wfh_handle* h, int32_t julian_day) {
    if (!h) return WFH_ERR_ARG;
    if (julian_day < 1 || julian_day > 365) return fail(h, WFH_ERR_ARG, "julian day outside 1..365");
    if (h->d.res_hq_col != julian_day - 1) {
        if (h->graph_mode) hipStreamSynchronize(h->stream);
        drop_graphs(h);
    }
    h->d.res_hq_col = julian_day - 1;
    return WFH_OK;
}

extern "C" int32_t wfh_sync(wfh_handle* h) {
    if (!h) return WFH_ERR_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int flag[2] = {0, 0};
    HIPCHK(h, hipMemcpy(flag, h->ra.abort_flag, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (flag[1]) {
        h->broken = h->broken_nonfinite = true;
        return fail(h, WFH_ERR_NONFINITE,
                    "non-finite forcing (precipitation, potential evaporation or temperature) in a step; the "
                    "states are not usable: load a saved state to continue");
    }
    if (flag[0]) {
        h->broken = true;
        return fail(h, WFH_ERR_DEVICE, "routing kernel aborted: a dependency wait exceeded its spin limit");
    }
    return WFH_OK;
}

namespace {
struct VarInfo {
    double* p;
    int layers;
    bool river;
    bool settable;
    bool res = false;  // one value per reservoir
};

VarInfo var_info(wfh_handle* h, int32_t v) {
    Dev& D = h->d;
    const int N = h->N;
    switch (v) {
        case WFH_CANOPY_STORAGE: return {D.canopy, 1, false, true};
        case WFH_SNOW_STORAGE: return {D.snow_s, 1, false, true};
        case WFH_SNOW_WATER: return {D.snow_w, 1, false, true};
        case WFH_SOIL_TEMPERATURE: return {D.tsoil, 1, false, true};
        case WFH_SATWATERDEPTH: return {D.swd, 1, false, true};
        case WFH_USTORELAYERDEPTH: return {D.ustore, N, false, true};
        case WFH_WATER_TABLE_DEPTH: return {D.zi, 1, false, false};
        case WFH_USTOREDEPTH: return {D.ustoredepth, 1, false, false};
        case WFH_TRANSFER: return {D.transfer, 1, false, false};
        case WFH_RECHARGE: return {D.recharge, 1, false, false};
        case WFH_ACTUAL_INFILTRATION: return {D.act_inf, 1, false, false};
        case WFH_INFILTRATION_EXCESS: return {D.inf_excess, 1, false, false};
        case WFH_SATURATION_EXCESS: return {D.sat_excess, 1, false, false};
        case WFH_TRANSPIRATION: return {D.transpiration, 1, false, false};
        case WFH_SOIL_EVAPORATION: return {D.soil_evap, 1, false, false};
        case WFH_ACTUAL_CAPILLARY_FLUX: return {D.capflux, 1, false, false};
        case WFH_ACTUAL_LEAKAGE: return {D.leakage, 1, false, false};
        case WFH_ACTUAL_EVAPOTRANSPIRATION: return {D.aet, 1, false, false};
        case WFH_INTERCEPTION_RATE: return {D.interception, 1, false, false};
        case WFH_NET_RUNOFF: return {D.net_runoff, 1, false, false};
        case WFH_RUNOFF: return {D.runoff, 1, false, false};
        case WFH_NET_RUNOFF_RIVER: return {D.net_runoff_river, 1, false, false};
        case WFH_EXFILTRATION: return {D.exfilt, 1, false, false};
        case WFH_TOTAL_STORAGE: return {D.total_storage, 1, false, false};
        case WFH_SNOW_WATER_EQUIVALENT: return {D.swe, 1, false, false};
        case WFH_VWC: return {D.vwc, N, false, false};
        case WFH_VWC_ROOT_ZONE: return {D.vwc_root, 1, false, false};
        case WFH_ROOT_ZONE_STORAGE: return {D.rzs, 1, false, false};
        case WFH_SSF_Q: return {D.ssf_q, 1, false, true};
        case WFH_SSF_QIN_AVERAGE: return {D.ssf_qin_avg, 1, false, false};
        case WFH_SSF_Q_AVERAGE: return {D.ssf_q_avg, 1, false, false};
        case WFH_SSF_TO_RIVER_AVERAGE: return {D.ssf_tor_avg, 1, false, false};
        case WFH_SSF_STORAGE: return {D.ssf_storage, 1, false, false};
        case WFH_LAND_Q: return {D.ol_q, 1, false, true};
        case WFH_LAND_QIN: return {D.ol_qin, 1, false, false};
        case WFH_LAND_H: return {D.ol_h, 1, false, true};
        case WFH_LAND_STORAGE: return {D.ol_storage, 1, false, false};
        case WFH_LAND_Q_AVERAGE: return {D.ol_q_avg, 1, false, false};
        case WFH_LAND_QIN_AVERAGE: return {D.ol_qin_avg, 1, false, false};
        case WFH_LAND_TO_RIVER_AVERAGE: return {D.ol_tor_avg, 1, false, false};
        case WFH_LAND_INWATER: return {D.ol_inwater, 1, false, false};
        case WFH_RIVER_Q: return {D.r_q, 1, true, true};
        case WFH_RIVER_QIN: return {D.r_qin, 1, true, false};
        case WFH_RIVER_H: return {D.r_h, 1, true, true};
        case WFH_RIVER_STORAGE: return {D.r_storage, 1, true, false};
        case WFH_RIVER_Q_AVERAGE: return {D.r_q_avg, 1, true, false};
        case WFH_RIVER_QIN_AVERAGE: return {D.r_qin_avg, 1, true, false};
        case WFH_RIVER_INWATER: return {D.r_inwater, 1, true, false};
        case WFH_BALANCE_LAND_ERROR: return {D.bal_land, 1, false, false};
        case WFH_BALANCE_LAND_RELERROR: return {D.bal_land_rel, 1, false, false};
        case WFH_BALANCE_SSF_ERROR: return {D.bal_ssf, 1, false, false};
        case WFH_BALANCE_OVERLAND_ERROR: return {D.bal_ol, 1, false, false};
        case WFH_BALANCE_RIVER_ERROR: return {D.bal_r, 1, true, false};
        case WFH_RIVER_ALPHA: return {const_cast<double*>(D.r_alpha), 1, true, false};
        case WFH_RESERVOIR_STORAGE: return {D.res_storage, 1, false, true, true};
        case WFH_RESERVOIR_WATERLEVEL: return {D.res_wl, 1, false, true, true};
        case WFH_RESERVOIR_OUTFLOW: return {D.res_outflow, 1, false, false, true};
        case WFH_RESERVOIR_OUTFLOW_AVERAGE: return {D.res_outflow_avg, 1, false, false, true};
        case WFH_RESERVOIR_INFLOW_AVERAGE: return {D.res_inflow_avg, 1, false, false, true};
        case WFH_RESERVOIR_ACTEVAP_AVERAGE: return {D.res_actevap_avg, 1, false, false, true};
        case WFH_RESERVOIR_ABSTRACTION_AVERAGE: return {D.res_abs_avg, 1, false, false, true};
        case WFH_RESERVOIR_PRECIPITATION: return {D.res_P, 1, false, false, true};
        case WFH_RESERVOIR_EVAPORATION: return {D.res_E, 1, false, false, true};
        case WFH_BALANCE_RESERVOIR_ERROR: return {D.bal_res, 1, false, false, true};
        case WFH_SNOW_IN: return {D.snow_in, 1, false, false};
        case WFH_SNOW_OUT: return {D.snow_out, 1, false, false};
        case WFH_FLOODPLAIN_STORAGE: return {D.fp_storage, 1, true, true};
        case WFH_FLOODPLAIN_H: return {D.fp_h, 1, true, false};
        case WFH_FLOODPLAIN_Q: return {D.fp_q, 1, true, false};
        case WFH_FLOODPLAIN_Q_AVERAGE: return {D.fp_q_avg, 1, true, false};
        case WFH_FLOODPLAIN_QIN_AVERAGE: return {D.fp_qin_avg, 1, true, false};
        case WFH_RIVER_Q_CHANNEL_AVERAGE: return {D.r_qch_avg, 1, true, false};
        default: return {nullptr, 0, false, false};
    }
}
}  // namespace

// parameters stored as broadcast-or-node-order maps
static bool prm_var(wfh_handle* h, int32_t v, Prm& p) {
    switch (v) {
        case WFH_LAND_ALPHA: p = h->d.alpha; return true;
        case WFH_THETA_FC: p = h->d.tfc; return true;
        case WFH_SURFACE_FLOW_WIDTH: p = h->d.sfw; return true;
        case WFH_SSF_QMAX: p = h->d.qmax; return true;
        default: return false;
    }
}

// Device-memory variants of wfh_get / wfh_set: land variables are permuted between the library's
// node order and the caller's land-index order by a gather / scatter kernel (or copied as they
// are for WFH_MEM_DEVICE_NODE_ORDER); river and reservoir variables are copied. Both complete
// before returning.
static int32_t get_device(wfh_handle* h, int32_t var, double* dst, int64_t count, int32_t memkind);
static int32_t set_device(wfh_handle* h, int32_t var, const double* src, int64_t count, int32_t memkind);

extern "C" int32_t wfh_get(wfh_handle* h, int32_t var, double* dst, int64_t count, int32_t memkind) {
    if (!h || !dst) return WFH_ERR_ARG;
    if (memkind == WFH_MEM_DEVICE || memkind == WFH_MEM_DEVICE_NODE_ORDER)
        return get_device(h, var, dst, count, memkind);
    if (memkind != WFH_MEM_HOST) return fail(h, WFH_ERR_ARG, "bad memkind");
    int32_t rc = wfh_sync(h);
    if (rc) return rc;
    const int64_t n = h->n;
    Prm pm;
    if (prm_var(h, var, pm)) {
        if (count != n) return fail(h, WFH_ERR_ARG, "count mismatch");
        std::vector<double> tmp(pm.mask ? n : 1);
        HIPCHK(h, hipMemcpy(tmp.data(), pm.p, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < n; ++q) dst[h->land_of[q]] = pm.mask ? tmp[q] : tmp[0];
        return WFH_OK;
    }
    if (var == WFH_FLOW_FRACTION_TO_RIVER) {
        if (count != n) return fail(h, WFH_ERR_ARG, "count mismatch");
        std::copy(h->fftr_land.begin(), h->fftr_land.end(), dst);
        return WFH_OK;
    }
    VarInfo vi = var_info(h, var);
    if (!vi.p) return fail(h, WFH_ERR_ARG, "unknown variable");
    if (vi.res) {
        if (count != h->nres) return fail(h, WFH_ERR_ARG, "count mismatch");
        if (h->nres) HIPCHK(h, hipMemcpy(dst, vi.p, h->nres * sizeof(double), hipMemcpyDeviceToHost));
        return WFH_OK;
    }
    if (vi.river) {
        if (count != h->nriv) return fail(h, WFH_ERR_ARG, "count mismatch");
        if (h->nriv) HIPCHK(h, hipMemcpy(dst, vi.p, h->nriv * sizeof(double), hipMemcpyDeviceToHost));
        return WFH_OK;
    }
    if (count != n * vi.layers) return fail(h, WFH_ERR_ARG, "count mismatch");
    // into the reference's cell order on the device, then one copy
    if (!h->get_tmp) {
        hipError_t e;
        h->get_tmp = dalloc<double>(h, (size_t)n * h->N, e);
        if (e != hipSuccess) {
            h->get_tmp = nullptr;
            return fail(h, WFH_ERR_HIP, "hipMalloc read-back buffer");
        }
    }
    for (int k = 0; k < vi.layers; ++k)
        hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->get_tmp + k * n,
                           vi.p + k * n, h->d_land_of, (int)n);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(dst, h->get_tmp, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return WFH_OK;
}

extern "C" int32_t wfh_set(wfh_handle* h, int32_t var, const double* src, int64_t count, int32_t memkind) {
    if (!h || !src) return WFH_ERR_ARG;
    if (memkind == WFH_MEM_DEVICE || memkind == WFH_MEM_DEVICE_NODE_ORDER)
        return set_device(h, var, src, count, memkind);
    if (memkind != WFH_MEM_HOST) return fail(h, WFH_ERR_ARG, "bad memkind");
    VarInfo vi = var_info(h, var);
    if (!vi.p || !vi.settable) return fail(h, WFH_ERR_ARG, "variable is not settable");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int64_t n = h->n;
    if (vi.res) {
        if (count != h->nres) return fail(h, WFH_ERR_ARG, "count mismatch");
        if (h->nres) HIPCHK(h, hipMemcpy(vi.p, src, h->nres * sizeof(double), hipMemcpyHostToDevice));
        return WFH_OK;
    }
    if (vi.river) {
        if (count != h->nriv) return fail(h, WFH_ERR_ARG, "count mismatch");
        if (h->nriv) HIPCHK(h, hipMemcpy(vi.p, src, h->nriv * sizeof(double), hipMemcpyHostToDevice));
        if (var == WFH_RIVER_H) {  // keep storage consistent (sbm_model.jl:174-176)
            std::vector<double> s(h->nriv), fw(h->nriv), fl(h->nriv);
            HIPCHK(h, hipMemcpy(fw.data(), h->d.r_fw, h->nriv * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(h, hipMemcpy(fl.data(), h->d.r_fl, h->nriv * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t r = 0; r < h->nriv; ++r) s[r] = src[r] * fw[r] * fl[r];
            HIPCHK(h, hipMemcpy(h->d.r_storage, s.data(), h->nriv * sizeof(double), hipMemcpyHostToDevice));
        }
        return WFH_OK;
    }
    if (count != n * vi.layers) return fail(h, WFH_ERR_ARG, "count mismatch");
    std::vector<double> tmp((size_t)n * vi.layers);
    for (int k = 0; k < vi.layers; ++k)
        for (int64_t q = 0; q < n; ++q) tmp[k * n + q] = src[k * n + h->land_of[q]];
    HIPCHK(h, hipMemcpy(vi.p, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    return WFH_OK;
}

static int32_t get_device(wfh_handle* h, int32_t var, double* dst, int64_t count, int32_t memkind) {
    int32_t rc = wfh_sync(h);
    if (rc) return rc;
    const int64_t n = h->n;
    const bool node = memkind == WFH_MEM_DEVICE_NODE_ORDER;
    VarInfo vi = var_info(h, var);
    if (!vi.p) {  // derived parameters: through host memory
        std::vector<double> tmp((size_t)count);
        if ((rc = wfh_get(h, var, tmp.data(), count, WFH_MEM_HOST))) return rc;
        if (node && count == n) {
            std::vector<double> t2((size_t)n);
            for (int64_t q = 0; q < n; ++q) t2[q] = tmp[h->land_of[q]];
            tmp.swap(t2);
        }
        HIPCHK(h, hipMemcpy(dst, tmp.data(), (size_t)count * sizeof(double), hipMemcpyHostToDevice));
        return WFH_OK;
    }
    const int64_t want = vi.res ? h->nres : vi.river ? h->nriv : n * vi.layers;
    if (count != want) return fail(h, WFH_ERR_ARG, "count mismatch");
    if (count == 0) return WFH_OK;
    if (vi.res || vi.river || node) {
        HIPCHK(h, hipMemcpyAsync(dst, vi.p, (size_t)count * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else {
        for (int k = 0; k < vi.layers; ++k)
            hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, dst + k * n,
                               vi.p + k * n, h->d_land_of, (int)n);
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return WFH_OK;
}

static int32_t set_device(wfh_handle* h, int32_t var, const double* src, int64_t count, int32_t memkind) {
    VarInfo vi = var_info(h, var);
    if (!vi.p || !vi.settable) return fail(h, WFH_ERR_ARG, "variable is not settable");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int64_t n = h->n;
    const bool node = memkind == WFH_MEM_DEVICE_NODE_ORDER;
    const int64_t want = vi.res ? h->nres : vi.river ? h->nriv : n * vi.layers;
    if (count != want) return fail(h, WFH_ERR_ARG, "count mismatch");
    if (var == WFH_RIVER_H) {  // keeps the river storage consistent: through host memory
        std::vector<double> tmp((size_t)count);
        HIPCHK(h, hipMemcpy(tmp.data(), src, (size_t)count * sizeof(double), hipMemcpyDeviceToHost));
        return wfh_set(h, var, tmp.data(), count, WFH_MEM_HOST);
    }
    if (count == 0) return WFH_OK;
    if (vi.res || vi.river || node) {
        HIPCHK(h, hipMemcpyAsync(vi.p, src, (size_t)count * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else {
        for (int k = 0; k < vi.layers; ++k)
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, vi.p + k * n,
                               src + k * n, h->d_land_of, (int)n);
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return WFH_OK;
}

extern "C" int32_t wfh_get_topology(wfh_handle* h, int64_t* land_order, int64_t* land_down, int64_t* river_order,
                                    int64_t* river_down, int64_t* river_land_index, int32_t* max_level) {
    if (!h) return WFH_ERR_ARG;
    if (land_order) std::copy(h->land_order.begin(), h->land_order.end(), land_order);
    if (land_down) std::copy(h->land_down.begin(), h->land_down.end(), land_down);
    if (river_order) std::copy(h->river_order.begin(), h->river_order.end(), river_order);
    if (river_down) std::copy(h->river_down.begin(), h->river_down.end(), river_down);
    if (river_land_index) std::copy(h->river_land.begin(), h->river_land.end(), river_land_index);
    if (max_level) *max_level = h->max_level;
    return WFH_OK;
}

extern "C" int32_t wfh_get_node_order(wfh_handle* h, int64_t* land_index_of_node) {
    if (!h || !land_index_of_node) return WFH_ERR_ARG;
    for (int64_t q = 0; q < h->n; ++q) land_index_of_node[q] = h->land_of[q];
    return WFH_OK;
}

// checkpoint: prognostic states in node order (states.jl:73-170 set plus the storages that the
// next step's balance reads)
static std::vector<std::pair<double*, size_t>> state_blocks(wfh_handle* h) {
    Dev& D = h->d;
    const size_t n = (size_t)h->n, r = (size_t)h->nriv;
    const size_t nr = (size_t)h->nres;
    return {{D.canopy, n},     {D.snow_s, n}, {D.snow_w, n}, {D.tsoil, n},          {D.swd, n},
            {D.ustore, n * h->N}, {D.tsws, n}, {D.ssf_q, n}, {D.ol_q, n},           {D.ol_h, n},
            {D.ol_storage, n}, {D.r_q, r},     {D.r_h, r},      {D.r_storage, r},
            {D.res_storage, nr}, {D.res_wl, nr}, {D.res_outflow, nr}, {D.fp_storage, D.fp ? r : 0}};
}

extern "C" int64_t wfh_state_size(const wfh_handle* h) {
    if (!h) return -1;
    size_t b = 4 * sizeof(int64_t);
    for (auto& s : state_blocks(const_cast<wfh_handle*>(h))) b += s.second * sizeof(double);
    return (int64_t)b;
}

extern "C" int32_t wfh_save_state(wfh_handle* h, void* buf, int64_t bytes) {
    if (!h || !buf || bytes != wfh_state_size(h)) return WFH_ERR_ARG;
    int32_t rc = wfh_sync(h);
    if (rc) return rc;
    int64_t* hdr = (int64_t*)buf;
    hdr[0] = 0x5746484853544154LL;
    hdr[1] = h->n;
    hdr[2] = h->nriv;
    hdr[3] = h->N;
    char* p = (char*)buf + 4 * sizeof(int64_t);
    for (auto& s : state_blocks(h)) {
        if (s.second) HIPCHK(h, hipMemcpy(p, s.first, s.second * sizeof(double), hipMemcpyDeviceToHost));
        p += s.second * sizeof(double);
    }
    return WFH_OK;
}

extern "C" int32_t wfh_load_state(wfh_handle* h, const void* buf, int64_t bytes) {
    if (!h || !buf || bytes != wfh_state_size(h)) return WFH_ERR_ARG;
    const int64_t* hdr = (const int64_t*)buf;
    if (hdr[0] != 0x5746484853544154LL || hdr[1] != h->n || hdr[2] != h->nriv || hdr[3] != h->N)
        return fail(h, WFH_ERR_ARG, "state buffer does not match this model");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const char* p = (const char*)buf + 4 * sizeof(int64_t);
    for (auto& s : state_blocks(h)) {
        if (s.second) HIPCHK(h, hipMemcpy(s.first, p, s.second * sizeof(double), hipMemcpyHostToDevice));
        p += s.second * sizeof(double);
    }
    if (h->broken && h->broken_nonfinite) {  // the routing itself completed: its scratch is clean
        const int zero = 0;
        HIPCHK(h, hipMemcpy(h->d.nonfinite_flag, &zero, sizeof(int), hipMemcpyHostToDevice));
        h->broken = h->broken_nonfinite = false;
    }
    return WFH_OK;
}

extern "C" int32_t wfh_set_timing(wfh_handle* h, int32_t enable) {
    if (!h) return WFH_ERR_ARG;
    h->timing = enable != 0;
    h->timed_steps = 0;
    return WFH_OK;
}

extern "C" int32_t wfh_get_timing(wfh_handle* h, double* vertical_ms, double* routing_ms) {
    if (!h) return WFH_ERR_ARG;
    if (h->timed_steps == 0) return fail(h, WFH_ERR_STATE, "no timed step recorded");
    int32_t rc = wfh_sync(h);
    if (rc) return rc;
    double sa = 0.0, sb = 0.0;
    for (size_t k = 0; k < h->timed_steps; ++k) {
        float a = 0.f, b = 0.f;
        HIPCHK(h, hipEventElapsedTime(&a, h->ev[3 * k], h->ev[3 * k + 1]));
        HIPCHK(h, hipEventElapsedTime(&b, h->ev[3 * k + 1], h->ev[3 * k + 2]));
        sa += a;
        sb += b;
    }
    if (vertical_ms) *vertical_ms = sa / h->timed_steps;
    if (routing_ms) *routing_ms = sb / h->timed_steps;
    return WFH_OK;
}

extern "C" const char* wfh_last_error(const wfh_handle* h) {
    if (!h) return g_create_error.c_str();
    return h->err.c_str();
}

extern "C" void wfh_destroy(wfh_handle* h) {
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (hipEvent_t ev : h->tune_ev)
        if (ev) hipEventDestroy(ev);
    for (void* p : h->allocs) hipFree(p);
    if (h->sel_host) hipHostFree(h->sel_host);
    for (hipEvent_t ev : h->ev)
        if (ev) hipEventDestroy(ev);
    if (h->copy_stream) hipStreamSynchronize(h->copy_stream);
    if (h->hf_copied) hipEventDestroy(h->hf_copied);
    if (h->hf_gathered) hipEventDestroy(h->hf_gathered);
    if (h->copy_stream) hipStreamDestroy(h->copy_stream);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int32_t wfh_kernel_bytes(wfh_handle* h, int64_t* vertical_bytes, int64_t* routing_bytes) {
    if (!h) return WFH_ERR_ARG;
    const Dev& D = h->d;
    const int64_t n = h->n, nr = h->nriv, N = h->N;
    const int64_t E = n - [&] {  // edges = cells with a downstream
        int64_t pits = 0;
        for (int64_t v : h->land_down) pits += v < 0;
        return pits;
    }();
    int64_t Er = 0;
    for (int64_t v : h->river_down) Er += v >= 0;
    auto P = [&](const Prm& p) { return p.mask ? 8 * n : 8; };
    auto PL = [&](const PrmL& p) { return p.mask ? 8 * n * N : 8 * N; };
    const int64_t A = 8 * n;
    // ---- K1 ----
    int64_t k1 = 3 * A;                                   // forcing
    k1 += A * (2 + N) + A + A;                            // canopy, tsoil, swd, ustore[N], ol_h (read)
    k1 += 4 * n + 8 * nr;                                 // river_of, r_h
    if (D.snow) k1 += 2 * A;                              // snow_s, snow_w
    if (D.balance) k1 += A;                               // tsws
    k1 += P(D.gap) + P(D.smax) + P(D.kc) + P(D.rf) + P(D.wf) + P(D.sfrac) + P(D.ts) + P(D.tr) + P(D.tfc) +
          P(D.d) + P(D.kv0) + P(D.f) + P(D.hb) + P(D.h1) + P(D.h2) + P(D.h3h) + P(D.h3l) + P(D.h4) + P(D.ah1) +
          P(D.icap_s) + P(D.icap_c) + P(D.compf) + P(D.rd) + P(D.rdp) + P(D.caph) + P(D.capn) + P(D.maxleak) +
          PL(D.c) + PL(D.kvfrac) + PL(D.rootfrac);
    if (D.gash) k1 += P(D.er);
    if (D.snow) k1 += P(D.cfmax) + P(D.tt) + P(D.tti) + P(D.ttm) + P(D.whc) + P(D.wsoil) + P(D.cfsoil);
    k1 += A * (2 + N) + A * 16;                           // canopy, tsoil, ustore[N], 16 fluxes
    if (D.snow) k1 += 3 * A;                              // snow_s, snow_w, swe
    if (D.balance) k1 += A;                               // store_prev
    // ---- K2 ----
    int64_t k2 = A * (N + 1) + A * 4 + A * 5;             // ustore[N], zi, ssf_q, recharge, ol q/h/storage; 5 K1 fluxes
    k2 += A * 2;                                          // canopy, snow (total storage; 1 word without snow)
    if (D.balance) k2 += 4 * A;                           // P, aet, leakage, store_prev
    k2 += 4 * (n + 1) + 4 * E + 8 * E + n + 4 * n;        // up_ptr, up_idx, up_f, flags, river_of
    k2 += P(D.d) + P(D.ts) + P(D.tr) + P(D.rd) + P(D.slope) + P(D.sy) + P(D.fl) + P(D.fw) + P(D.qmax) +
          P(D.kh0) + P(D.f) + P(D.area) + P(D.alpha) + P(D.sfw) + P(D.rf) + P(D.tfc);
    k2 += 8 * E * (h->S_s + h->S_l) + 8 * Er * h->S_r;    // upstream discharge per edge and sub-step
    k2 += 8 * nr * 7 + 4 * (nr + 1) + 4 * Er;             // river states + static, river CSR
    k2 += A * 14 + A * 2 * N;                             // subsurface + soil outputs, ustore[N], vwc[N]
    k2 += A * 9;                                          // overland outputs + total storage
    k2 += 8 * nr * 7;                                     // river outputs
    if (D.balance) k2 += 4 * A + 8 * nr;                  // balance errors
    if (h->fuse) {
        // one kernel: nothing the vertical update writes needs reading back, so none of those
        // reads is algorithmic (the launch keeps ustore[N], zi and recharge in registers and
        // re-reads the 5 fluxes, canopy, snow and balance inputs: traffic above this count)
        k2 += k1 - (A * (N + 1) + A + A * 5 + A * 2 + (D.balance ? 4 * A : 0));
        k1 = 0;
    }
    if (vertical_bytes) *vertical_bytes = k1;
    if (routing_bytes) *routing_bytes = k2;
    return WFH_OK;
}

// ============================================================================================
// Parameter ensembles (member-innermost layout, staged launches replayed from a HIP graph)
// ============================================================================================
#include "wfh_mlp.h"
#include "wfh_ensemble.h"
