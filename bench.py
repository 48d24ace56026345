#!/usr/bin/env python3
"""Throughput of the Wflow.jl `sbm` time step (update_model!) on one MI355X.

Workload: 4096 x 4096 cells (16.8 M) from the synthetic generator, dx 250 m, hourly steps,
river cells where the upstream area exceeds 2000 cells, FP64 states. A "step" is one model
time step of the whole grid: the vertical column kernel plus the routing kernel (lateral
subsurface flow, overland and river kinematic waves with their fixed sub-steps, storage and
water-balance updates). Inputs are resident in HBM before timing starts: the forcing for the
timed steps is generated on the GPU beforehand (a ring of distinct hourly frames, in the
library's node order) and bound without copies.

Output: one JSON line with the metric, the roofline of the dominant kernel (algorithmic bytes
per launch from the library / mean HIP-event duration of that kernel over the timed steps) and
a CPU baseline (the FP64 oracle timed on this host over a bounded sample of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wflow.jl_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("grid", "batch", "ensemble"), default="grid",
                    help="grid: one nx^2 catchment at dx 250 m; batch: 64 catchments of log-uniform size "
                         "(1e4 .. 2e6 cells, dx 1 km, river threshold 100 cells) in one model; ensemble: "
                         "--ens-members parameter members of one --ens-nx^2 catchment (kv_0, f, n lognormal "
                         "sigma 0.3), member-innermost, autotuned schedule")
    ap.add_argument("--ens-members", type=int, default=256)
    ap.add_argument("--ens-nx", type=int, default=2048)
    ap.add_argument("--ens-tune-trials", type=int, default=1)
    ap.add_argument("--learned", action="store_true",
                    help="ensemble workload: also time the learned Newton warm start (trained first)")
    ap.add_argument("--learn-harvest-steps", type=int, default=6)
    ap.add_argument("--learn-epochs", type=int, default=4)
    ap.add_argument("--learn-gate", type=float, default=1e-12,
                    help="a wave consults the network where some member's previous discharge is below this "
                         "(inf: every solve)")
    ap.add_argument("--no-variants", action="store_true",
                    help="grid workload: skip the batched (ensemble) and batched+learned lines attached to it")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--tune-trials", type=int, default=1,
                    help="batch workload: routing-grid autotuner trials per candidate during the warm-up")
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--frames", type=int, default=24, help="distinct forcing frames kept in HBM")
    ap.add_argument("--spinup", type=int, default=-1,
                    help="model steps run before the warm-up (each with its own forcing frame), so that the "
                         "timed steps end at the configuration's last hour (720 for the grid, 480 for the batch); "
                         "-1: that default, 0: a cold start")
    ap.add_argument("--cpu-sample-cells", type=int, default=200_000, help="cells per CPU thread")
    ap.add_argument("--cpu-threads", type=int, default=16, help="CPU threads of the baseline (at most the cores)")
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reservoirs", type=int, default=0,
                    help="place this many reservoirs on the grid's river cells (simple / modified Puls)")
    ap.add_argument("--reservoir-curves", action="store_true",
                    help="with --reservoirs: cycle the outflow types 1-4 (rating curves, free weirs) and give every "
                         "third reservoir an S-H storage table")
    ap.add_argument("--floodplain", action="store_true",
                    help="1-D floodplain beside every river cell (floodplain_1d); one launch per river sub-step")
    ap.add_argument("--snow-transport", action="store_true", help="lateral snow transport (gravitational)")
    ap.add_argument("--host-forcing", action="store_true",
                    help="hand the forcing over from host memory every step (WFH_MEM_HOST: permutation into the "
                         "node order and a PCIe upload inside the timed region); not the headline line")
    ap.add_argument("--adaptive", action="store_true",
                    help="adaptive kinematic-wave sub-steps (kinematic_wave__adaptive_time_step_flag) instead of "
                         "the fixed ones; not the headline line: no roofline, every sub-step waits for the last")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_routing_4096.json"))
    ap.add_argument("--ens-traffic-file", default=os.path.join(ROOT, "profiles", "traffic_ensemble256_2048.json"),
                    help="HBM bytes per ensemble step from counter runs (roofline.traffic of the ensemble)")
    ap.add_argument("--fp64-file", default=os.path.join(ROOT, "profiles", "fp64_4096.json"),
                    help="FP64 VALU FLOPs per launch from a counter run (SQ_INSTS_VALU_*_F64) of the same workload")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="after the timed steps, run one elementwise kernel over this many MiB (reads and "
                         "writes each byte once: known traffic, for scaling hardware byte counters)")
    return ap.parse_args()


STATE_VARS_RESTART = None  # filled in main (Var members)


def cpu_baseline(catchment, frames_host, args, states=None, hour0=0):
    """The oracle on complete basins of the same catchment; basins exchange no water, so a set of
    whole basins is an exact sub-problem. The largest basin (the main stem) is always in the
    sample; further basins are drawn until every thread (up to 16, the host share of one GPU) has
    about --cpu-sample-cells cells. Each group of basins is stepped by its own oracle model in its
    own thread (the C oracle releases the interpreter lock), the way the reference spreads
    independent subbasins over its threads. With `states` (land / river order arrays of the GPU
    model at the timed region's start) the oracle starts from the same state and steps the same
    forcing frames as the timed GPU steps. The single-thread rate of one group is reported too."""
    import threading

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from wflowhip import Var, _abi, synth

    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    lab = synth.basin_ids(catchment.arrays["ldd"], catchment.nx, catchment.ny, catchment.arrays["cell_index"])
    ids, counts = np.unique(lab, return_counts=True)
    big = int(np.argmax(counts))
    rng = np.random.default_rng(5)
    order = [k for k in rng.permutation(len(ids)) if k != big]
    want = args.cpu_sample_cells * max(1, threads - 1)
    take, total = [big], 0
    for k in order:
        if total + counts[k] > want and total > 0:
            continue
        take.append(k)
        total += counts[k]
        if total >= 0.9 * want:
            break
    # the largest basin alone in its group, the rest split greedily into groups of similar size
    groups = [[ids[big]]] + [[] for _ in range(threads - 1)]
    sizes = [0] * threads
    for k in sorted(take[1:], key=lambda k: -counts[k]):
        g = 1 + int(np.argmin(sizes[1:])) if threads > 1 else 0
        groups[g].append(ids[k])
        sizes[g] += counts[k]
    groups = [g for g in groups if g]
    masks = [np.isin(lab, np.array(g)) for g in groups]
    kw = dict(adaptive=args.adaptive, snow_transport=args.snow_transport)
    subs = [catchment.subset(mk) for mk in masks]
    riv = catchment.arrays["river_location"].astype(bool)
    rmasks = [mk[riv] for mk in masks]

    def make(j):
        mdl = O.OracleModel(subs[j], **kw)
        if states is not None:
            for var, a in states.items():
                if _abi.Var(var) in _abi.RIVER_VARS:
                    mdl.set(var, a[rmasks[j]])
                elif _abi.Var(var) in _abi.LAYERED_VARS:
                    mdl.set(var, a.reshape(-1, catchment.n)[:, masks[j]].ravel())
                else:
                    mdl.set(var, a[masks[j]])
        return mdl

    models = [make(j) for j in range(len(subs))]
    frames = [[tuple(f[mk] for f in frames_host[k]) for k in range(len(frames_host))] for mk in masks]

    def run(j):
        for k in range(len(frames_host)):
            models[j].step(*frames[j][k])

    # single-thread rate on the last (random-basin) group, then all groups at once
    js = len(models) - 1
    t0 = time.perf_counter()
    run(js)
    t_single = time.perf_counter() - t0
    models[js].close()
    models[js] = make(js)
    ths = [threading.Thread(target=run, args=(j,)) for j in range(len(models))]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    for mdl in models:
        mdl.close()
    cells = int(sum(int(mk.sum()) for mk in masks))
    steps = len(frames_host)
    return {
        "value": cells * steps / wall,
        "unit": "cell-timesteps/s",
        "cores": len(models),
        "kind": "port",
        "single_thread_value": int(masks[js].sum()) * steps / t_single,
        "sample": f"{len(take)} complete basins ({cells} cells, including the largest basin of "
                  f"{int(counts[big])} cells, which holds the main stem) of the same workload in {len(models)} "
                  f"groups, one thread each, {steps} steps, C oracle (-O2), update_model! only"
                  + (f", started from the GPU model's state at model hour {hour0} and stepped over the same "
                     f"forcing as the timed GPU steps" if states is not None else ", cold start")
                  + (" (adaptive sub-steps of each group itself)" if args.adaptive else ""),
    }


def _ens_steps(ens, frames, k0, steps, memkind):
    for j in range(steps):
        ens.set_forcing(*frames[(k0 + j) % len(frames)], memkind=memkind)
        ens.update_model()
    ens.sync()


def train_warm_start(args, log):
    """The learned Newton start: rows harvested from plain solves of another catchment's ensemble
    (a 1024^2 catchment, seed 41, 64 members, hourly), then trained on the GPU."""
    import numpy as np

    from wflowhip import learned as LW, synth

    tc = synth.make_catchment(1024, cell=250.0, dt=3600.0, river_threshold=2000, seed=41)
    f = np.exp(0.3 * np.random.default_rng(41).standard_normal((3, 64)))
    t0 = time.time()
    rows = LW.harvest(tc, f, args.learn_harvest_steps, lambda k: synth.forcing_numpy(tc, 48 + k), every=16,
                      cap=4_000_000)
    log(f"# harvested {rows.shape[0]} rows in {time.time() - t0:.1f}s")
    t0 = time.time()
    net, losses = LW.fit(rows, epochs=args.learn_epochs)
    log(f"# trained {args.learn_epochs} epochs in {time.time() - t0:.1f}s, mse {losses[0]:.3f} -> {losses[-1]:.3f}")
    return net, {"harvested_rows": int(rows.shape[0]), "epochs": args.learn_epochs,
                 "mse_first_last": [losses[0], losses[-1]],
                 "source": "plain solves of a 64-member ensemble of another 1024^2 catchment (seed 41), "
                           f"{args.learn_harvest_steps} hourly steps, every 16th solve"}


def ensemble_run(args, t0, learned=False):
    """Parameter ensemble of one catchment (wfh_ens_*): every member's update_model! in one graph
    replay per step; the first steps of the warm-up run the autotuner's candidate schedules. With
    `learned`, the same ensemble then runs as many steps with the learned warm start."""
    import numpy as np
    import torch

    from wflowhip import Ensemble, _abi, synth
    from wflowhip import learned as LW

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    net, train_info = (None, None)
    if learned:
        net, train_info = train_warm_start(args, log)
    c = synth.make_catchment(args.ens_nx, cell=250.0, dt=3600.0, river_threshold=2000, seed=4)
    M = args.ens_members
    rng = np.random.default_rng(44)
    fa, fb, fc = np.exp(0.3 * rng.standard_normal((3, M)))
    workload = (f"{M}-member parameter ensemble of one {args.ens_nx}x{args.ens_nx} catchment (kv_0, f, Manning n "
                "scaled per member, lognormal sigma 0.3), dx 250 m, hourly, river threshold 2000 cells, sbm + "
                "kinematic wave (land 1, river 4, subsurface 1 sub-steps), member-innermost, staged launches "
                "replayed from a HIP graph, autotuned schedule")
    log(f"# {workload}: n={c.n} nriv={c.nriv} longest path={c.max_depth} ({time.time() - t0:.1f}s)")
    ens = Ensemble(c, fa, fb, fc, water_mass_balance=True)
    log(f"# ensemble created ({time.time() - t0:.1f}s)")
    from wflowhip import Model
    base = Model(c)  # node order only
    node_t = torch.as_tensor(base.node_order(), device="cuda")
    base.close()
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[node_t].contiguous()
    nframes = max(1, min(args.frames, args.steps + args.warmup))
    frames = [synth.forcing_torch(c, k, elev) for k in range(nframes)]
    torch.cuda.synchronize()
    NODE = _abi.WFH_MEM_DEVICE_NODE_ORDER
    ens.set_autotune(args.ens_tune_trials)
    k = 0
    while ens.schedule()["tuning"] or k < args.warmup:
        _ens_steps(ens, frames, k, 1, NODE)
        k += 1
    sched = ens.schedule()
    log(f"# {k} warm-up steps (autotuner): schedule {sched} ({time.time() - t0:.1f}s)")

    def timed(label):
        nonlocal k
        _ens_steps(ens, frames, k, 1, NODE)  # records the graph of this variant
        k += 1
        ens.set_timing(True)
        t1 = time.perf_counter()
        _ens_steps(ens, frames, k, args.steps, NODE)
        elapsed = time.perf_counter() - t1
        k += args.steps
        step_ms = ens.timing()
        ens.set_timing(False)
        bal = float(ens.balance().max())
        log(f"# {label}: {args.steps} steps in {elapsed:.3f}s ({step_ms:.2f} ms per graph replay); max |land "
            f"balance error| {bal:.2e}")
        # Newton-update histogram over two further (untimed) steps
        ens.kw_histogram(True)
        _ens_steps(ens, frames, k, 2, NODE)
        k += 2
        hist, used = ens.kw_histogram(False)
        upd = [LW.mean_updates(hist[0]), LW.mean_updates(hist[1])]
        return elapsed, step_ms, bal, hist, used, upd

    elapsed, step_ms, bal, hist, used, upd = timed("plain")
    nbytes = ens.step_bytes()
    achieved = nbytes / (step_ms * 1e-3) / 1e9
    result = {
        "metric": "member-cell-timesteps/s FP64 SBM+kinwave (batched ensemble GPU step)",
        "value": M * c.n * args.steps / elapsed,
        "unit": "cell-timesteps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": k, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (fractal DEM, seeded parameters and member factors, GPU-generated forcing)",
        "config": {"workload": workload, "members": M, "cells": c.n, "river_cells": c.nriv,
                   "longest_flow_path": c.max_depth, "schedule": sched, "step_bytes": nbytes,
                   "max_abs_land_balance_error": bal,
                   "newton_updates_per_solve": {"overland": upd[0], "river": upd[1]},
                   "newton_update_histogram": {"overland": hist[0].tolist(), "river": hist[1].tolist()}},
        "roofline": {"bound": "hbm", "kernel": "k_ens_stage (all stages of a step)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    if os.path.exists(args.ens_traffic_file):
        tf = json.load(open(args.ens_traffic_file))
        if tf.get("members") == M and tf.get("n") == c.n:
            result["roofline"]["traffic"] = tf["bytes_per_step"]
    ens.close()
    if learned:
        # plain and learned side by side over the same model hours: two ensembles of half the
        # members each (both fit in HBM together), stepped alternately on the same forcing
        Mh = M // 2
        pair = [Ensemble(c, fa[:Mh], fb[:Mh], fc[:Mh], water_mass_balance=True) for _ in range(2)]
        pair[1].set_learned(net, args.learn_gate)
        for e_ in pair:
            e_.set_schedule(sched["block"], sched["wide"], sched["strand_cells"])
        for j in range(args.warmup):
            for e_ in pair:
                _ens_steps(e_, frames, j, 1, NODE)
        ms = [0.0, 0.0]
        for e_ in pair:
            e_.set_timing(True)
        for j in range(args.warmup, args.warmup + args.steps):
            for i_, e_ in enumerate(pair):
                _ens_steps(e_, frames, j, 1, NODE)
        ms = [e_.timing() for e_ in pair]
        hists = []
        for e_ in pair:
            e_.set_timing(False)
            e_.kw_histogram(True)
        for j in range(args.warmup + args.steps, args.warmup + args.steps + 2):
            for e_ in pair:
                _ens_steps(e_, frames, j, 1, NODE)
        hists = [e_.kw_histogram(False) for e_ in pair]
        bal_l = float(pair[1].balance().max())
        log(f"# side by side, {Mh} members each, same hours: plain {ms[0]:.2f} ms, learned {ms[1]:.2f} ms per step")
        upd = [[LW.mean_updates(h[0][f_]) for f_ in (0, 1)] for h in hists]
        result["learned"] = {
            "members": Mh, "ms_per_step_plain": ms[0], "ms_per_step_learned": ms[1],
            "value": Mh * c.n / (ms[1] * 1e-3), "value_plain": Mh * c.n / (ms[0] * 1e-3),
            "ratio_vs_plain_ensemble": ms[0] / ms[1],
            "max_abs_land_balance_error": bal_l,
            "newton_updates_per_solve": {"plain": {"overland": upd[0][0], "river": upd[0][1]},
                                         "learned": {"overland": upd[1][0], "river": upd[1][1]}},
            "newton_update_histogram": {"plain": {"overland": hists[0][0][0].tolist(),
                                                  "river": hists[0][0][1].tolist()},
                                        "learned": {"overland": hists[1][0][0].tolist(),
                                                    "river": hists[1][0][1].tolist()}},
            "learned_starts_taken": {"overland": int(hists[1][1][0]), "river": int(hists[1][1][1]),
                                     "solves": [int(hists[1][0][0].sum()), int(hists[1][0][1].sum())]},
            "training": train_info, "gate_q_prev": args.learn_gate,
            "note": "two ensembles of half the members stepped alternately over the same model hours "
                    "(graph-replay time per step from HIP events)",
        }
        for e_ in pair:
            e_.close()
        net.close()
    return result, c, frames, node_t


def calibration_kernel(args):
    """One elementwise kernel of known traffic for scaling the HBM counters (--calibration-copy-mib)."""
    if args.calibration_copy_mib <= 0:
        return
    import torch

    nb = args.calibration_copy_mib << 20
    src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.neg(src, out=dst)  # one elementwise kernel: reads nb bytes, writes nb bytes
    torch.cuda.synchronize()
    print(f"# calibration kernel (neg): {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
    del src, dst


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(a model is one GPU's basins; see the sharding notes of the design)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from wflowhip import Model, Var, _abi, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t0 = time.time()
    if args.workload == "ensemble":
        result, c, frames, node_t = ensemble_run(args, t0, learned=args.learned)
        calibration_kernel(args)
        if not args.no_cpu_baseline:
            node = node_t.cpu().numpy()
            inv = np.empty_like(node)
            inv[node] = np.arange(node.size)
            frames_host = [tuple(f.cpu().numpy()[inv] for f in frames[k]) for k in range(args.cpu_sample_steps)]
            result["cpu_baseline"] = cpu_baseline(c, frames_host, args)
            result["cpu_baseline"]["sample"] += " (member 0, the unperturbed catchment; a member is one model)"
        else:
            result["cpu_baseline"] = None
        print(json.dumps(result), flush=True)
        return
    if args.workload == "grid":
        c = synth.make_catchment(args.nx, cell=250.0, dt=3600.0, river_threshold=2000, seed=1)
        workload = (f"{args.nx}x{args.nx} cells, dx 250 m, hourly, river threshold 2000 cells, "
                    "sbm + kinematic wave (land 1, river 4, subsurface 1 sub-steps)")
    else:
        rng = np.random.default_rng(2)
        sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e6), args.members))
        members = [synth.make_catchment(int(round(math.sqrt(s))), cell=1000.0, dt=3600.0, river_threshold=100,
                                        seed=1000 + m) for m, s in enumerate(sizes)]
        c = synth.make_batch(members)
        workload = (f"{args.members} catchments in one model, {min(m.n for m in members)} .. "
                    f"{max(m.n for m in members)} cells each (log-uniform), dx 1 km, hourly, river threshold "
                    "100 cells, sbm + kinematic wave (land 1, river 4, subsurface 1 sub-steps)")
    print(f"# {workload}: n={c.n} nriv={c.nriv} longest path={c.max_depth} "
          f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    if args.adaptive:
        workload = workload.replace("(land 1, river 4, subsurface 1 sub-steps)", "(adaptive sub-steps)")
    if args.reservoirs:
        synth.add_reservoirs(c, args.reservoirs, seed=1, radius=2, curves=args.reservoir_curves)
        workload += f", {args.reservoirs} reservoirs" + (" (outflow types 1-4, S-H tables)" if args.reservoir_curves
                                                          else "")
    if args.floodplain:
        synth.add_floodplain(c, seed=1)
        workload += ", 1-D floodplain (5 flood depths, bankfull depth x 0.02)"
    if args.snow_transport:
        workload += ", lateral snow transport"
    m = Model(c, water_mass_balance=True, adaptive=args.adaptive, snow_transport=args.snow_transport)
    print(f"# model created ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    node = m.node_order()
    node_t = torch.as_tensor(node, device="cuda")
    elev = torch.as_tensor(c.arrays["elevation"], device="cuda")[node_t].contiguous()
    # spin-up: the configuration's run up to the timed hours (each step its own forcing frame)
    plain_cfg = not (args.adaptive or args.floodplain or args.reservoirs or args.snow_transport or args.host_forcing)
    run_len = 720 if args.workload == "grid" else 480
    h0 = args.spinup if args.spinup >= 0 else (max(0, run_len - args.warmup - args.steps) if plain_cfg else 0)
    for k in range(h0):
        fr = synth.forcing_torch(c, k, elev)
        torch.cuda.synchronize()
        m.set_forcing(*fr, memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        m.update_model()
        m.sync()
        if (k + 1) % 100 == 0:
            print(f"# spin-up: {k + 1} of {h0} steps ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    if h0:
        workload += f", timed over model hours {h0 + args.warmup}-{h0 + args.warmup + args.steps} after a {h0}-step spin-up"
    nframes = max(1, min(max(args.frames, args.steps + args.warmup), 64))
    frames = [synth.forcing_torch(c, h0 + k, elev) for k in range(nframes)]
    torch.cuda.synchronize()
    print(f"# {nframes} forcing frames on the GPU ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    if args.host_forcing:  # the same frames in host memory, in the reference's cell order
        inv_h = np.empty_like(node)
        inv_h[node] = np.arange(node.size)
        frames_h = [tuple(np.ascontiguousarray(f.cpu().numpy()[inv_h]) for f in fr) for fr in frames]
        workload += ", forcing handed over from host memory every step (PCIe-inclusive)"

    def step(k):
        if args.host_forcing:
            m.set_forcing(*frames_h[k % nframes])
        else:
            m.set_forcing(*frames[k % nframes], memkind=_abi.WFH_MEM_DEVICE_NODE_ORDER)
        m.update_model()

    tuned = None
    if args.workload == "batch" and plain_cfg and args.tune_trials > 0:
        # the routing-grid autotuner runs on the first warm-up steps (real steps of the model)
        m.autotune(args.tune_trials)
    k = 0
    while k < args.warmup or (tuned is None and m.schedule()["tuning"]):
        step(k)
        k += 1
    if args.workload == "batch" and plain_cfg and args.tune_trials > 0:
        tuned = m.schedule()
        print(f"# autotuned routing grid: {tuned} after {k} warm-up steps", file=sys.stderr, flush=True)
    args.warmup = k
    m.sync()
    torch.cuda.synchronize()
    states = None
    if not args.no_cpu_baseline and plain_cfg:  # the state the CPU baseline starts from
        states = {int(v): m.get(v) for v in (Var.CANOPY_STORAGE, Var.SNOW_STORAGE, Var.SNOW_WATER,
                                              Var.SOIL_TEMPERATURE, Var.SATWATERDEPTH, Var.USTOREDEPTH,
                                              Var.USTORELAYERDEPTH, Var.SSF_Q, Var.LAND_Q, Var.LAND_H,
                                              Var.LAND_STORAGE, Var.RIVER_Q, Var.RIVER_H, Var.RIVER_STORAGE)}
    m.set_timing(True)
    t1 = time.perf_counter()
    subs = []
    for k in range(args.warmup, args.warmup + args.steps):
        step(k)
        if args.adaptive or args.floodplain:
            subs.append(m.last_substeps())
    m.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    k1_ms, k2_ms = m.timing()
    k1_bytes, k2_bytes = m.kernel_bytes()
    m.set_timing(False)
    graph_replay = None
    if args.workload == "batch" and plain_cfg:
        # the same steps replayed from the HIP graph of the step (graph mode), after the timed region
        m.set_graph(True)
        step(args.warmup + args.steps)  # records the graphs
        step(args.warmup + args.steps + 1)
        m.sync()
        tg = time.perf_counter()
        for k in range(args.warmup + args.steps + 2, args.warmup + 2 * args.steps + 2):
            step(k)
        m.sync()
        graph_replay = {"ms_per_step": (time.perf_counter() - tg) / args.steps * 1e3, "schedule": m.schedule()}
        m.set_graph(False)
        print(f"# graph replay: {graph_replay['ms_per_step']:.3f} ms per step", file=sys.stderr, flush=True)
    bal = float(np.abs(m.get(Var.BALANCE_LAND_ERROR)).max())
    # the plain model runs the vertical update inside the routing launch (no vertical bytes)
    fused = k1_bytes == 0
    kernels = ("routing launch with the vertical update" if fused else
               f"vertical {k1_ms:.3f} ms, routing")
    print(f"# {args.steps} steps in {elapsed:.3f}s: {kernels} {k2_ms:.3f} ms per step; "
          f"max |land balance error| {bal:.2e}", file=sys.stderr, flush=True)

    if k2_ms >= k1_ms:
        dom, dom_ms, dom_bytes = "k_route", k2_ms, k2_bytes
    else:
        dom, dom_ms, dom_bytes = "k_vertical", k1_ms, k1_bytes
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_file):
        try:
            tf = json.load(open(args.traffic_file))
            # the counters were taken on the plain configuration of this grid
            if (plain_cfg and tf.get("kernel") == dom and tf.get("n") == c.n
                    and bool(tf.get("vertical_fused")) == fused):
                traffic = tf.get("bytes_per_launch")
        except Exception:
            traffic = None
    metric = ("cell-timesteps/s FP64 SBM+kinwave (plain GPU step)" if args.workload == "grid" else
              "cell-timesteps/s FP64 SBM+kinwave (batched GPU step)")
    if args.adaptive:
        metric = metric.replace(" step)", " step, adaptive sub-steps)")
    if args.reservoirs or args.snow_transport or args.floodplain:
        metric = metric.replace(" step", " step with reservoirs / floodplain / snow transport", 1)
    if args.host_forcing:
        metric = metric.replace(" step", " step, host forcing", 1)
    result = {
        "metric": metric,
        "value": c.n * args.steps / elapsed,
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (fractal DEM, seeded parameters, GPU-generated forcing)",
        "config": {"workload": workload,
                   "cells": c.n, "river_cells": c.nriv, "longest_flow_path": c.max_depth,
                   "kernel_ms": {"k_vertical": k1_ms, "k_route": k2_ms},
                   "kernel_bytes": {"k_vertical": k1_bytes, "k_route": k2_bytes},
                   "vertical_fused": fused,
                   **({"autotuned_schedule": tuned, "graph_replay": graph_replay} if graph_replay else {})},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    # FP64 VALU rate: FLOPs per launch from the counter run of the same workload over this run's
    # live kernel times
    if os.path.exists(args.fp64_file) and plain_cfg:
        try:
            ff = json.load(open(args.fp64_file))
            if ff.get("n") == c.n and bool(ff.get("vertical_fused")) == fused:
                result["fp64_valu"] = {
                    k: {"achieved": ff[k]["flops_per_launch"] / (ms * 1e-3) / 1e12, "peak": ff["peak_tflops"],
                        "unit": "TFLOP/s", "frac": ff[k]["flops_per_launch"] / (ms * 1e-3) / 1e12 / ff["peak_tflops"]}
                    for k, ms in (("k_route", k2_ms), ("k_vertical", k1_ms)) if k in ff and ms > 0.05}
        except Exception:
            pass
    if args.adaptive or args.floodplain:
        # the routing time covers all sub-step launches and the host read-backs between them
        result["roofline"] = None
        result["config"]["kernel_ms"] = {"k_vertical": k1_ms, "routing": k2_ms}
        del result["config"]["kernel_bytes"]
        result["config"]["substeps_mean"] = [float(np.mean([s[i] for s in subs])) for i in range(3)]
        result["config"]["substeps_max"] = [int(max(s[i] for s in subs)) for i in range(3)]
    calibration_kernel(args)
    if not args.no_cpu_baseline:
        frames_host = []
        inv = np.empty_like(node)
        inv[node] = np.arange(node.size)
        k0 = args.warmup if states is not None else 0
        for k in range(k0, min(k0 + args.cpu_sample_steps, nframes)):
            frames_host.append(tuple(f.cpu().numpy()[inv] for f in frames[k]))
        result["cpu_baseline"] = cpu_baseline(c, frames_host, args, states, h0 + k0)
    else:
        result["cpu_baseline"] = None
    m.close()
    if args.workload == "grid" and not args.no_variants and not (args.adaptive or args.floodplain or args.reservoirs
                                                                  or args.snow_transport or args.host_forcing):
        # the metric's comparison: the batched form (parameter ensemble, configs' combined workload)
        # and the batched form with the learned warm start, against this plain line
        del frames
        torch.cuda.empty_cache()
        ens, _, _, _ = ensemble_run(args, t0, learned=True)
        plain = result["value"]
        result["variants"] = {
            "batched": {"value": ens["value"], "unit": ens["unit"], "ms_per_step": ens["ms_per_step"],
                        "ratio_vs_plain": ens["value"] / plain, "workload": ens["config"]["workload"],
                        "schedule": ens["config"]["schedule"], "roofline": ens["roofline"],
                        "newton_updates_per_solve": ens["config"]["newton_updates_per_solve"]},
            "batched_learned": {"value": ens["learned"]["value"], "unit": ens["unit"],
                                "members": ens["learned"]["members"],
                                "ms_per_step": ens["learned"]["ms_per_step_learned"],
                                "ratio_vs_plain": ens["learned"]["value"] / plain,
                                "ratio_vs_batched": ens["learned"]["ratio_vs_plain_ensemble"],
                                "newton_updates_per_solve": ens["learned"]["newton_updates_per_solve"],
                                "learned_starts_taken": ens["learned"]["learned_starts_taken"],
                                "training": ens["learned"]["training"]},
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
