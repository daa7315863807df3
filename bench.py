#!/usr/bin/env python
"""Coupled cell-steps/s of the Frehg2 coupled step on one MI355X.

Workload: the coupled 512x512 surface over 512x512x30 soil (dx = 1 m, dz = 0.05 m, dt = 1 s,
50 mm/h rain on a dry tilted bed with a Gaussian random field, lognormal per-cell Ks), inputs
generated from seeds by frehg_amd.workloads. One step = Orchestrator::step, coupled branch:
SWE step, SW<->GW exchange, Richards step with its Krylov solves.

Each Richards step is the Picard loop to max |h^{m+1} - h^m| <= 1e-8 m (--picard-max iterates at
most); the reference's own predictor step (one iterate) is measured as well and reported under
predictor_step. Four GPU paths run from the same initial state for warmup + timed steps:
  plain        FP64 Jacobi-PCG from x0 = 0 in every iterate (the reference's solver semantics)
  learned      FP32-inner / FP64-refined PCG; the first iterate and its Krylov guess are
               h^n + d^{n-1} + MLP correction (d^{n-1}: the previous step's solve increment;
               zero-guess fallback), later iterates start from the previous one
  extrapolated the same from h^n + d^{n-1} (no network)
  previous     the same from h^n
The MLP is trained beforehand (not timed) on pairs (state, converged increment minus the previous
one) logged from plain runs of two other seeded bed/Ks realisations of the same workload. `value`
is the learned path's throughput; every path's final head is checked against the plain path's.
Kernel durations come from HIP events on the backend's stream during the timed region.
The CPU baseline is the repository's FP64 oracle (a restatement of the reference algorithm)
timed on a bounded sample of the same workload on this host's cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# L2<->fabric bytes of one FP64 CG SpMV launch (k_spmv_pair<double>) on the 512x512x30 system
# per cell, from two rocprofv3 --pmc runs of this script with --copy-probe (FETCH_SIZE x2.000 and
# WRITE_SIZE x1.000, the scale factors that a 1 GiB read + 1 GiB write elementwise kernel needs).
SPMV64_TRAFFIC_BYTES_PER_CELL = 554.7e6 / 7864320
F16_PEAK_TFLOPS = 2500.0  # dense MFMA, FP16 (the learned-guess inference) and BF16 alike


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def interior3(a, nx, ny, nz):
    return a.reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]


KERNEL_TYPES = ("spmv64", "spmv32", "update", "assemble", "correct", "refine", "guess")
# kernel types timed inside the timed region (the roofline kernels); the others are timed over
# further, untimed steps, since timing events make a launch slower to issue
LIVE_TYPES = ("spmv64", "spmv32", "guess")
BREAKDOWN_STEPS = 10


def run_gpu(case, abi, gpu, path, warmup, steps, grab_after=None, model=None, picard_max=1, gate=0.0):
    """Warmup + timed coupled steps; returns timing, solver counters and kernel timings."""
    ctx = gpu.Context(0)
    if path == "plain":
        case.setup_gpu(ctx, abi.PREC_FP64, abi.GUESS_ZERO, picard_max=picard_max)
    elif path == "plain_previous":
        case.setup_gpu(ctx, abi.PREC_FP64, abi.GUESS_PREVIOUS, picard_max=picard_max)
    elif path == "learned":
        ctx.mlp_set_params(*model)
        ctx.mlp_gate(gate)
        case.setup_gpu(ctx, abi.PREC_MIXED, abi.GUESS_LEARNED, picard_max=picard_max)
    elif path == "extrapolated":
        case.setup_gpu(ctx, abi.PREC_MIXED, abi.GUESS_EXTRAPOLATED, picard_max=picard_max)
    else:
        case.setup_gpu(ctx, abi.PREC_MIXED, abi.GUESS_PREVIOUS, picard_max=picard_max)
    grabbed = None
    for s in range(warmup):
        ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
        if grab_after is not None and s + 1 == grab_after:
            grabbed = ctx.re_get("h")
    ctx.synchronize()
    kt_ids = dict(zip(KERNEL_TYPES, (abi.KT_GW_SPMV, abi.KT_GW_SPMV32, abi.KT_GW_UPDATE, abi.KT_GW_ASSEMBLE,
                                     abi.KT_GW_CORRECT, abi.KT_GW_REFINE, abi.KT_GW_GUESS)))
    ctx.profile_select([kt_ids[k] for k in LIVE_TYPES])
    ctx.profile_enable(True)
    ctx.profile_reset()
    iters = refines = fallbacks = picards = 0
    mass_err = 0.0
    t0 = time.perf_counter()
    for s in range(steps):
        st = ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
        iters += st.gw.solve.iterations
        refines += st.gw.solve.refinements
        fallbacks += st.gw.guess_fallback
        picards += st.gw.picard_iterations
        mass_err = max(mass_err, abs(st.gw.mass_residual) / st.gw.mass_before)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    live = {k: ctx.profile_read(kt_ids[k]) for k in LIVE_TYPES}
    # per-kernel-type breakdown: further steps with every type timed (not part of the timing)
    ctx.profile_select(kt_ids.values())
    ctx.profile_reset()
    for s in range(BREAKDOWN_STEPS):
        ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
    ctx.synchronize()
    kt = {k: ctx.profile_read(v) for k, v in kt_ids.items()}
    ctx.profile_enable(False)
    inferred = ctx.mlp_inferred_cells() if path == "learned" else None
    shape = ctx.spmv_shape(abi.PREC_FP64 if path.startswith("plain") else abi.PREC_MIXED)
    h = ctx.re_get("h")
    depth = ctx.swe_get("dept")
    sw_vol, gw_vol = ctx.water_volume()
    ctx.close()
    return {"elapsed": elapsed, "iters": iters, "refines": refines, "fallbacks": fallbacks, "picards": picards,
            "kt": kt, "live": live, "h": h, "depth": depth, "inferred": inferred,
            "grabbed": grabbed, "volume": sw_vol + gw_vol, "spmv_shape": shape, "mass_err": mass_err}


def train_model(abi, gpu, workloads, args, seeds, log_steps, stride, train_steps, batch, picard_max=1):
    """Log (features, converged increment) pairs from plain FP64 runs of other seeded realisations
    of the workload, then train the MLP on the GPU. Returns (params, norm) and a summary."""
    ctx = gpu.Context(0)
    ctx.mlp_init(17)
    ncell = args.nx * args.ny * args.nz
    ctx.mlp_log(stride, capacity=len(seeds) * log_steps * ((ncell + stride - 1) // stride))
    t0 = time.perf_counter()
    for sb, sk in seeds:
        case = workloads.CoupledCase(nx=args.nx, ny=args.ny, nz=args.nz, seed_bed=sb, seed_ks=sk)
        case.setup_gpu(ctx, abi.PREC_FP64, abi.GUESS_ZERO, picard_max=picard_max)
        for _ in range(log_steps):
            ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
    ctx.synchronize()
    t_log = time.perf_counter() - t0
    npairs = ctx.mlp_logged()
    t0 = time.perf_counter()
    first, last = ctx.mlp_train(train_steps, batch=batch, lr=1e-3, seed=29)
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    model = ctx.mlp_get_params()
    ctx.close()
    return model, {"pairs": npairs, "realisations": [list(x) for x in seeds], "steps_logged_each": log_steps,
                   "cell_stride": stride, "adam_steps": train_steps, "batch": batch, "lr": 1e-3,
                   "loss_first": round(first, 5), "loss_last": round(last, 5),
                   "log_seconds": round(t_log, 2), "train_seconds": round(t_train, 2)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(case, abi, steps, picard_max=1):
    """The FP64 oracle on the same workload for `steps` coupled steps (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    nthreads = ol.lib().orc_num_threads() if hasattr(ol.lib(), "orc_num_threads") else os.cpu_count()
    s = ol.OracleSwe(case.nx, case.ny, case.dx, case.dx, case.swe)
    s.set_bathymetry(case.bed_with_offset())
    s.initialize_state(-1.0e3)
    s.set_boundary_polygons(case.boundary_regions())
    r = ol.OracleRe(case.nx, case.ny, case.nz, case.re_params(picard_max=picard_max))
    r.set_soil([case.soil], None, True, case.ks.ravel())
    r.set_initial(h=case.h0)
    oc = ol.OracleCoupled(s, r, case.cp, case.t_end)
    t0 = time.perf_counter()
    for _ in range(steps):
        oc.step(case.dt, case.rain)
    elapsed = time.perf_counter() - t0
    return {"elapsed": elapsed, "threads": nthreads, "h": r.get("h")}


def bench_ensemble(args):
    """The batched-ensemble configuration: 96 uneven Richards boxes (nx, ny in [32, 128], nz in [20, 120]),
    mixed precision, previous-step guess, 30 s steps, every step replayed from captured graphs."""
    from frehg_amd import abi, gpu, workloads
    log("[bench] generating the 96-box ensemble")
    case = workloads.EnsembleCase(nbox=96, seed=3)
    case_ex = workloads.EnsembleCase(nbox=96, seed=3, guess=abi.GUESS_EXTRAPOLATED)
    steps, warmup = args.steps, args.warmup
    res = {}
    for mode in ("graph", "eager", "graph_extrapolated"):
        run_case = case_ex if mode == "graph_extrapolated" else case
        use_graph = mode != "eager"
        ctx = gpu.Context(0)
        run_case.setup_gpu(ctx)
        for _ in range(warmup):
            ctx.ens_step(use_graph=use_graph)
        ctx.synchronize()
        worst = 0.0
        its = refs = extra = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            st = ctx.ens_step(use_graph=use_graph)
            worst = max(worst, st.max_rel_mass_residual)
            its = max(its, st.max_iterations)
            refs = max(refs, st.max_refinements)
            extra += st.extra_launches
        ctx.synchronize()
        el = time.perf_counter() - t0
        res[mode] = {"value": round(case.cells * steps / el, 1), "ms_per_step": round(el / steps * 1e3, 3),
                     "max_inner_iterations": its, "max_refinements": refs, "eager_launches_past_budget": extra,
                     "max_rel_mass_balance_error": worst}
        if mode == "graph":
            b = int(np.argmin([nx * ny * nz for nx, ny, nz in case.dims]))
            hb = ctx.ens_get(b, "h")
        ctx.close()
        log(f"[bench]   {mode}: {res[mode]}")
    # parity spot check: the smallest box against the FP64 oracle over the same steps
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    nx, ny, nz = case.dims[b]
    p = abi.ReParams.from_buffer_copy(case.params[b])
    p.solver = abi.solver_config(1e-10)
    r = ol.OracleRe(nx, ny, nz, p)
    r.set_soil([case.soils[b]], None, True, case.ks[b])
    r.set_initial(h=np.full(nx * ny * nz, case.h0[b]))
    for _ in range(warmup + steps):
        r.step()
    err = float(np.max(np.abs(interior3(hb, nx, ny, nz) - interior3(r.get("h"), nx, ny, nz))))
    out = {"metric": "ensemble cell-steps/s (96 uneven Richards boxes, graph-replayed steps)",
           "value": res["graph"]["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": res["graph"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic (seeded per-box soils and lognormal Ks fields)",
           "config": {"workload": "96 Richards boxes, nx,ny in [32,128], nz in [20,120], dx=dy=0.5 m, dz=0.05 m, "
                                  "top flux 1e-6 m/s, dt 30 s, FP32-in-FP64 refinement, previous-step guess",
                      "cells": case.cells, "seed": 3},
           "graph": res["graph"], "eager": res["eager"],
           "graph_extrapolated_start": res["graph_extrapolated"],
           "parity_smallest_box": {"box": b, "dims": case.dims[b], "max_head_err_vs_fp64_oracle_m": err}}
    print(json.dumps(out), flush=True)


def bench_infiltration(args):
    """The 40x40x50 infiltration benchmark in full: 360 Picard steps of 10 s (free-drainage bottom),
    the plain FP64 GPU path against the FP64 oracle over the same 360 steps on this host's cores."""
    from frehg_amd import abi, gpu, workloads
    case = workloads.InfiltrationCase()
    ctx = gpu.Context(0)
    ctx.re_init(case.nx, case.ny, case.nz, case.re_params())
    ctx.re_set_initial(h=case.h0)
    ctx.synchronize()
    pic = its = 0
    worst = 0.0
    t0 = time.perf_counter()
    for _ in range(case.steps):
        st = ctx.re_step()
        pic += st.picard_iterations
        its += st.solve.iterations
        worst = max(worst, abs(st.mass_residual) / st.mass_before)
    ctx.synchronize()
    el = time.perf_counter() - t0
    hg = ctx.re_get("h")
    ctx.close()
    log(f"[bench] GPU: {el:.3f} s for {case.steps} steps")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    r = ol.OracleRe(case.nx, case.ny, case.nz, case.re_params())
    r.set_initial(h=case.h0)
    t0 = time.perf_counter()
    for _ in range(case.steps):
        r.step()
    cel = time.perf_counter() - t0
    log(f"[bench] CPU oracle: {cel:.1f} s for {case.steps} steps")
    err = float(np.max(np.abs(interior3(hg, case.nx, case.ny, case.nz) - interior3(r.get("h"), case.nx, case.ny,
                                                                                       case.nz))))
    nthreads = ol.lib().orc_num_threads()
    out = {"metric": "Richards cell-steps/s, 40x40x50 infiltration benchmark (360 Picard steps of 10 s)",
           "value": round(case.cells * case.steps / el, 1), "unit": "cell-steps/s", "n_gpus": 1,
           "steps": case.steps, "warmup": 0, "ms_per_step": round(el / case.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (the configured homogeneous column)",
           "config": {"workload": "GW 40x40x50, dx=dy=0.1 m, dz=0.02 m, top Dirichlet -0.75 m, free-drainage bottom, "
                                  "dt 10 s x 360, Picard to 1e-8 m (at most 50 iterates), FP64 CG rtol 1e-10"},
           "picard_iters_per_step": pic / case.steps, "krylov_iters_per_step": its / case.steps,
           "max_rel_mass_balance_error": worst, "max_head_err_vs_fp64_oracle_m": err,
           "cpu_baseline": {"value": round(case.cells * case.steps / cel, 1), "unit": "cell-steps/s",
                            "cores": nthreads, "kind": "port", "cpu_model": cpu_model(),
                            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                            "sample": f"all {case.steps} steps ({cel:.1f} s), FP64 oracle, Jacobi-PCG, same Picard "
                                      "settings"}}
    print(json.dumps(out), flush=True)


def bench_large(args):
    """The large coupled configuration, flow only: 1024x1024 surface at dx = 2 m over 1024x1024x64 soil at
    dz = 0.04 m (bed seed 11, Ks seed 12), Picard to 1e-8 m, dt 1 s: plain FP64, extrapolated + mixed
    and learned + mixed (network trained on two other seeded realisations of this size); FP64 water
    ledger every step; the plain path after one step against the FP64 oracle."""
    from frehg_amd import abi, gpu, workloads
    log("[bench] generating the 1024x1024x64 inputs")
    case = workloads.CoupledCase.large()
    steps, warmup = args.steps, args.warmup
    a = argparse.Namespace(nx=case.nx, ny=case.ny, nz=case.nz)

    class Large:  # train_model builds its realisations through workloads.CoupledCase(nx, ny, nz, ...)
        CoupledCase = staticmethod(lambda nx, ny, nz, seed_bed, seed_ks: workloads.CoupledCase.large(
            seed_bed=seed_bed, seed_ks=seed_ks))
    log("[bench] plain FP64 path")
    plain = run_gpu(case, abi, gpu, "plain", warmup, steps, grab_after=1, picard_max=args.picard_max)
    log(f"[bench]   {plain['elapsed']:.3f} s")
    ext = run_gpu(case, abi, gpu, "extrapolated", warmup, steps, picard_max=args.picard_max)
    log(f"[bench]   extrapolated {ext['elapsed']:.3f} s")
    model, training = train_model(abi, gpu, Large, a, seeds=((21, 22), (23, 24)), log_steps=6, stride=32,
                                  train_steps=args.train_steps, batch=65536, picard_max=args.picard_max)
    fast = run_gpu(case, abi, gpu, "learned", warmup, steps, model=model, picard_max=args.picard_max, gate=args.gate)
    log(f"[bench]   learned {fast['elapsed']:.3f} s")
    parity = None
    if not args.no_cpu:
        log("[bench] FP64 oracle, one coupled step of the same workload")
        cb = cpu_baseline(case, abi, 1, picard_max=args.picard_max)
        parity = float(np.max(np.abs(interior3(cb["h"], case.nx, case.ny, case.nz) -
                                     interior3(plain["grabbed"], case.nx, case.ny, case.nz))))
        log(f"[bench]   {cb['elapsed']:.1f} s; GPU FP64 vs oracle max |dh| = {parity:.3e} m")
    hp = interior3(plain["h"], case.nx, case.ny, case.nz)

    def row(r):
        return {"value": round(case.cells * steps / r["elapsed"], 1),
                "ms_per_step": round(r["elapsed"] / steps * 1e3, 3),
                "picard_iters_per_step": r["picards"] / steps, "iters_per_step": r["iters"] / steps,
                "max_rel_mass_balance_error": r["mass_err"],
                "max_head_err_vs_plain_fp64_m": float(np.max(np.abs(hp - interior3(r["h"], case.nx, case.ny,
                                                                                   case.nz))))}
    out = {"metric": "coupled cell-steps/s, 1024x1024 + 1024x1024x64 (flow only)",
           "value": round(case.cells * steps / fast["elapsed"], 1), "unit": "cell-steps/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(fast["elapsed"] / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32+f16",
           "data": "synthetic (seeded Gaussian random fields: bed sigma 0.05 m / 20 m, ln Ks sigma 1 / 10 m)",
           "config": {"workload": "coupled SW 1024x1024 (dx 2 m) + GW 1024x1024x64 (dz 0.04 m), dt 1 s, 50 mm/h rain, "
                                  "critical-depth outflow on the low-x edge, Picard to 1e-8 m; no solute tracer",
                      "cells_per_step": case.cells},
           "learned_mixed": row(fast), "extrapolated_mixed": row(ext), "plain_fp64": row(plain),
           "speedup_vs_plain_fp64": round(plain["elapsed"] / fast["elapsed"], 3),
           "gpu_fp64_max_head_err_vs_oracle_after_1_step_m": parity, "training": training,
           "spmv_shape_cells_per_lane_kchunk": plain["spmv_shape"]}
    print(json.dumps(out), flush=True)


def bench_learned_eval(args):
    """Held-out evaluation of the learned guess on the coupled 512x512 + 512x512x30 workload: train on
    pairs logged from plain FP64 runs of 20 seeded bed/Ks realisations (seed r for the bed, r + 100
    for Ks, r = 101..120; about 2e7 pairs), then run held-out realisations r = 301..305 (none of their
    bed or Ks seeds is used in training) with FP64 Jacobi-PCG at the same tolerance from x0 = 0, from
    x0 = h^n + the previous step's solve increment + the MLP's correction, from x0 = h^n and from
    x0 = h^n + the previous step's solve increment, reporting Krylov (and Picard) iterations per step
    and the head difference to the zero guess: once with the reference's predictor step (one Picard
    iterate) and once with the Picard loop to 1e-8 m, where the starting head is also the loop's
    first iterate."""
    from frehg_amd import abi, gpu, workloads
    train_seeds = [(r, r + 100) for r in range(101, 121)]
    eval_seeds = [(r, r + 100) for r in range(301, 306)]
    ncell = args.nx * args.ny * args.nz
    log_steps = 11  # the first step of each run logs no pairs (no previous solve increment)
    stride = max(1, (len(train_seeds) * (log_steps - 1) * ncell) // 20_000_000)
    log(f"[bench] training on {len(train_seeds)} realisations, {log_steps} logged steps each, cell stride {stride}")
    model, training = train_model(abi, gpu, workloads, args, seeds=train_seeds, log_steps=log_steps, stride=stride,
                                  train_steps=args.train_steps, batch=65536)
    log(f"[bench]   {training['pairs']} pairs, loss {training['loss_first']} -> {training['loss_last']}")
    per_seed = []
    for sb, sk in eval_seeds:
        case = workloads.CoupledCase(nx=args.nx, ny=args.ny, nz=args.nz, seed_bed=sb, seed_ks=sk)
        row = {"seeds": [sb, sk]}
        heads = {}
        runs = (("zero", abi.GUESS_ZERO, 1), ("learned", abi.GUESS_LEARNED, 1), ("previous", abi.GUESS_PREVIOUS, 1),
                ("extrapolated", abi.GUESS_EXTRAPOLATED, 1), ("picard_zero", abi.GUESS_ZERO, 50),
                ("picard_learned", abi.GUESS_LEARNED, 50), ("picard_previous", abi.GUESS_PREVIOUS, 50),
                ("picard_extrapolated", abi.GUESS_EXTRAPOLATED, 50))
        for name, guess, picard_max in runs:
            ctx = gpu.Context(0)
            if guess == abi.GUESS_LEARNED:
                ctx.mlp_set_params(*model)
                ctx.mlp_gate(args.gate)
            case.setup_gpu(ctx, abi.PREC_FP64, guess, picard_max=picard_max)
            its = pic = fb = 0
            for _ in range(args.steps):
                st = ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
                its += st.gw.solve.iterations
                pic += st.gw.picard_iterations
                fb += st.gw.guess_fallback
            heads[name] = ctx.re_get("h")
            ctx.close()
            row[name] = {"krylov_iters_per_step": round(its / args.steps, 3),
                         "picard_iters_per_step": round(pic / args.steps, 3), "guess_fallbacks": fb}
        for name in ("learned", "previous", "extrapolated"):
            row[name]["max_head_diff_vs_zero_guess_m"] = float(np.max(np.abs(heads[name] - heads["zero"])))
        for name in ("picard_learned", "picard_previous", "picard_extrapolated"):
            row[name]["max_head_diff_vs_zero_guess_m"] = float(np.max(np.abs(heads[name] - heads["picard_zero"])))
        log(f"[bench]   held-out {sb}/{sk}: {row}")
        per_seed.append(row)
    names = [r[0] for r in runs]
    mean = {name: round(float(np.mean([r[name]["krylov_iters_per_step"] for r in per_seed])), 3) for name in names}
    mean_picard = {name: round(float(np.mean([r[name]["picard_iters_per_step"] for r in per_seed])), 3)
                   for name in names if name.startswith("picard")}
    out = {"metric": "Krylov iterations per coupled step on held-out realisations, learned vs zero initial guess",
           "value": mean["learned"], "unit": "iterations/step", "n_gpus": 1, "steps": args.steps, "warmup": 0,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f16",
           "data": "synthetic (seeded Gaussian random fields: bed sigma 0.05 m / 20 m, ln Ks sigma 1 / 10 m)",
           "config": {"workload": f"coupled SW {args.nx}x{args.ny} + GW {args.nx}x{args.ny}x{args.nz}, dt 1 s, "
                                  "FP64 Jacobi-PCG at rtol 1e-10 for every guess",
                      "train_realisations": [list(x) for x in train_seeds],
                      "held_out_realisations": [list(x) for x in eval_seeds]},
           "mean_krylov_iters_per_step": mean, "mean_picard_iters_per_step": mean_picard,
           "picard": {"max": 50, "tol_m": 1e-8}, "per_realisation": per_seed, "training": training}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=30)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--log-steps", type=int, default=11,
                    help="coupled steps logged per training realisation (the first logs no pairs)")
    ap.add_argument("--train-realisations", type=int, default=20,
                    help="seeded bed/Ks realisations the learned guess is trained on (bed seed r, Ks seed r + 100, "
                         "r = 101, 102, ...; about 2e7 logged pairs in all)")
    ap.add_argument("--picard-max", type=int, default=50,
                    help="Picard iterations per Richards step at most, to max |dh| <= 1e-8 m (1: the "
                         "reference's predictor step, which is also measured and reported as predictor_step)")
    ap.add_argument("--gate", type=float, default=1e-6,
                    help="learned-guess inference gate [m]: cells whose previous solve increment is at most "
                         "this take the extrapolated start without the network (0: the network everywhere)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run an elementwise 1 GiB -> 1 GiB kernel (torch), the reference point that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    ap.add_argument("--workload", choices=("coupled", "ensemble", "learned-eval", "large", "infiltration"),
                    default="coupled",
                    help="coupled: the 512x512 + 512x512x30 coupled step (default); ensemble: 96 uneven boxes; "
                         "learned-eval: the learned guess trained on 20 realisations and evaluated on 5 "
                         "held-out ones (iterations per step against the zero guess); large: 1024x1024 + "
                         "1024x1024x64 flow only; infiltration: the 40x40x50 column, 360 steps, GPU and CPU")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); the multi-GPU "
              "sharding of the coupled step is a design, not a build.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "learned-eval":
        bench_learned_eval(args)
        return
    if args.workload == "large":
        bench_large(args)
        return
    if args.workload == "infiltration":
        bench_infiltration(args)
        return

    if args.copy_probe:
        import torch
        n = (1 << 30) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.add(a, 1.0, out=b)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b
        log("[bench] copy probe: 3 elementwise kernels, 1 GiB read + 1 GiB written each")

    from frehg_amd import abi, gpu, workloads
    log(f"[bench] generating inputs {args.nx}x{args.ny}x{args.nz}")
    case = workloads.CoupledCase(nx=args.nx, ny=args.ny, nz=args.nz)
    n3 = case.nx * case.ny * case.nz
    cpu_steps = 0 if args.no_cpu else max(1, min(args.cpu_steps, args.warmup))

    log("[bench] plain FP64 path")
    plain = run_gpu(case, abi, gpu, "plain", args.warmup, args.steps, grab_after=cpu_steps or None,
                    picard_max=args.picard_max)
    log(f"[bench]   {plain['elapsed']:.3f} s, {plain['iters']} Krylov iterations")
    seeds = [(r, r + 100) for r in range(101, 101 + args.train_realisations)]
    stride = max(1, (len(seeds) * (args.log_steps - 1) * n3) // 20_000_000)
    log(f"[bench] training the learned guess on {len(seeds)} other seeded realisations (not timed)")
    model, training = train_model(abi, gpu, workloads, args, seeds=seeds, log_steps=args.log_steps, stride=stride,
                                  train_steps=args.train_steps, batch=65536, picard_max=args.picard_max)
    log(f"[bench]   {training['pairs']} pairs, loss {training['loss_first']} -> {training['loss_last']}")
    log(f"[bench] learned path (FP32-in-FP64 refinement, MLP guess, gate {args.gate:g} m)")
    fast = run_gpu(case, abi, gpu, "learned", args.warmup, args.steps, model=model, picard_max=args.picard_max,
                   gate=args.gate)
    log(f"[bench]   {fast['elapsed']:.3f} s, {fast['iters']} inner iterations, {fast['refines']} refinements, "
        f"{fast['fallbacks']} fallbacks")
    log("[bench] previous-step path (FP32-in-FP64 refinement, x0 = h^n)")
    prev = run_gpu(case, abi, gpu, "previous", args.warmup, args.steps, picard_max=args.picard_max)
    log(f"[bench]   {prev['elapsed']:.3f} s, {prev['iters']} inner iterations, {prev['refines']} refinements")
    log("[bench] extrapolated path (FP32-in-FP64 refinement, x0 = h^n + the previous step's solve increment)")
    ext = run_gpu(case, abi, gpu, "extrapolated", args.warmup, args.steps, picard_max=args.picard_max)
    log(f"[bench]   {ext['elapsed']:.3f} s, {ext['iters']} inner iterations, {ext['refines']} refinements")
    log("[bench] plain FP64 path with warm starts (every Krylov solve from x0 = h^n / h^m)")
    plain_prev = run_gpu(case, abi, gpu, "plain_previous", args.warmup, args.steps, picard_max=args.picard_max)
    log(f"[bench]   {plain_prev['elapsed']:.3f} s, {plain_prev['iters']} Krylov iterations")
    predictor = None
    if args.picard_max > 1:
        log("[bench] the reference's predictor step (one Picard iterate): plain, learned, previous, extrapolated")
        p1 = run_gpu(case, abi, gpu, "plain", args.warmup, args.steps, picard_max=1)
        f1 = run_gpu(case, abi, gpu, "learned", args.warmup, args.steps, model=model, picard_max=1, gate=args.gate)
        v1 = run_gpu(case, abi, gpu, "previous", args.warmup, args.steps, picard_max=1)
        e1 = run_gpu(case, abi, gpu, "extrapolated", args.warmup, args.steps, picard_max=1)
        h1 = interior3(p1["h"], case.nx, case.ny, case.nz)

        def summary(r, inner):
            out = {"value": round(case.cells * args.steps / r["elapsed"], 1),
                   "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 4),
                   ("inner_iters_per_step" if inner else "krylov_iters_per_step"): r["iters"] / args.steps}
            if inner:
                out["refinements_per_step"] = r["refines"] / args.steps
                out["max_head_err_vs_plain_fp64_m"] = float(np.max(np.abs(h1 - interior3(r["h"], case.nx, case.ny,
                                                                                          case.nz))))
            return out
        predictor = {"plain_fp64": summary(p1, False), "learned_mixed": summary(f1, True),
                     "previous_mixed": summary(v1, True), "extrapolated_mixed": summary(e1, True),
                     "speedup_vs_plain_fp64": round(p1["elapsed"] / f1["elapsed"], 3)}
        log(f"[bench]   {predictor}")

    hp = interior3(plain["h"], case.nx, case.ny, case.nz)
    hf = interior3(fast["h"], case.nx, case.ny, case.nz)
    head_err = float(np.max(np.abs(hp - hf)))
    head_err_prev = float(np.max(np.abs(hp - interior3(prev["h"], case.nx, case.ny, case.nz))))
    head_err_ext = float(np.max(np.abs(hp - interior3(ext["h"], case.nx, case.ny, case.nz))))
    head_err_pp = float(np.max(np.abs(hp - interior3(plain_prev["h"], case.nx, case.ny, case.nz))))

    def depth_err(r):
        a = r["depth"].reshape(case.ny + 2, case.nx + 2)[1:-1, 1:-1]
        b = plain["depth"].reshape(case.ny + 2, case.nx + 2)[1:-1, 1:-1]
        return float(np.max(np.abs(a - b)))

    def roof(kt_ms_n, bytes_per_cell):
        ms, n = kt_ms_n
        if n == 0:
            return None
        avg_s = ms / n / 1e3
        achieved = bytes_per_cell * n3 / avg_s / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_us": round(avg_s * 1e6, 2),
                "launches": n, "bytes_per_cell": bytes_per_cell}

    # fused PCG SpMV kernel: reads z, p_prev, D, ax, ay, az; writes p, q (8 words per cell)
    roof64 = roof(plain["live"]["spmv64"], 64)
    if roof64 and (args.nx, args.ny, args.nz) == (512, 512, 30):
        roof64["traffic"] = round(SPMV64_TRAFFIC_BYTES_PER_CELL * n3)
        roof64["traffic_unit"] = "bytes per launch (rocprofv3 PMC, scaled)"
    roof32 = roof(fast["live"]["spmv32"], 32)
    roof_guess = None
    ms, n = fast["live"]["guess"]
    if n:
        # algorithmic MLP forward over the cells the network ran on (the gate sends the others
        # to the extrapolated start; the count is the last step's, steps differ little)
        flops = 2 * (11 * 64 + 64 * 64 + 64) * fast["inferred"]
        tf = flops / (ms / n / 1e3) / 1e12
        roof_guess = {"bound": "mfma", "achieved": round(tf, 2), "peak": F16_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(tf / F16_PEAK_TFLOPS, 5), "avg_launch_us": round(ms / n * 1e3, 2),
                      "launches": n, "flop_per_cell": 9728, "cells_inferred": fast["inferred"],
                      "cells_total": n3}

    cpu = None
    if cpu_steps:
        log(f"[bench] CPU baseline: FP64 oracle, {cpu_steps} coupled steps of the same workload")
        cb = cpu_baseline(case, abi, cpu_steps, picard_max=args.picard_max)
        parity = float(np.max(np.abs(interior3(cb["h"], case.nx, case.ny, case.nz) -
                                     interior3(plain["grabbed"], case.nx, case.ny, case.nz))))
        cpu = {"value": round(case.cells * cpu_steps / cb["elapsed"], 1), "unit": "cell-steps/s",
               "cores": cb["threads"], "kind": "port", "cpu_model": cpu_model(),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
               "sample": f"first {cpu_steps} coupled steps of the same {args.nx}x{args.ny}x{args.nz} workload "
                         f"({cb['elapsed']:.1f} s), FP64 oracle with Jacobi-PCG at the reference rtol, the same "
                         f"Picard settings as the plain path",
               "gpu_fp64_max_head_err_vs_oracle_m": parity}
        log(f"[bench]   {cb['elapsed']:.1f} s on {cb['threads']} threads; GPU FP64 vs oracle max |dh| = {parity:.3e} m")

    # the ceiling of the SpMV's access mix on this device: six reads and two writes per cell over
    # arrays of the SpMV's size, no neighbours (frehg_stream_probe)
    probe_gbs, probe_ms = gpu.stream_probe(case.nx * case.ny * case.nz, reps=50)
    stream = {"achieved": round(probe_gbs, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(probe_gbs / 8000.0, 4),
              "avg_launch_us": round(probe_ms * 1e3, 2), "bytes_per_cell": 64,
              "spmv_frac_of_probe": round(roof64["achieved"] / probe_gbs, 4) if roof64 else None}
    log(f"[bench] streaming probe (6 reads + 2 writes per cell): {probe_gbs:.0f} GB/s = {probe_gbs / 80:.1f}% of peak")

    value = case.cells * args.steps / fast["elapsed"]
    plain_value = case.cells * args.steps / plain["elapsed"]
    out = {
        "metric": "coupled cell-steps/s at <=1e-6 m head err: learned+mixed vs plain FP64 GPU",
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fast["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32+f16",
        "data": "synthetic (seeded Gaussian random fields: bed sigma 0.05 m / 20 m, ln Ks sigma 1 / 10 m)",
        "config": {"workload": f"coupled SW {args.nx}x{args.ny} + GW {args.nx}x{args.ny}x{args.nz}, dx 1 m, "
                               f"dz 0.05 m, dt 1 s, 50 mm/h rain, critical-depth outflow on the low-x edge, "
                               f"closed sides",
                   "cells_per_step": case.cells, "fast_path": "fp32-in-fp64 refinement, learned (MLP) guess",
                   "richards_step": (f"Picard loop to max |dh| <= 1e-8 m (at most {args.picard_max} iterates); "
                                     "plain path: FP64 PCG from x0 = 0 in every iterate; learned path: h^n + the "
                                     "previous step's solve increment + the MLP's correction as the first iterate "
                                     "and Krylov guess, later iterates warm-started"
                                     if args.picard_max > 1 else "the reference's predictor step (one iterate)"),
                   "gw_rtol": 1e-10, "sw_rtol": 1e-12},
        "plain_fp64": {"value": round(plain_value, 1), "ms_per_step": round(plain["elapsed"] / args.steps * 1e3, 4),
                       "krylov_iters_per_step": plain["iters"] / args.steps,
                       "picard_iters_per_step": plain["picards"] / args.steps},
        "plain_fp64_warm_start": {"value": round(case.cells * args.steps / plain_prev["elapsed"], 1),
                                  "ms_per_step": round(plain_prev["elapsed"] / args.steps * 1e3, 4),
                                  "krylov_iters_per_step": plain_prev["iters"] / args.steps,
                                  "picard_iters_per_step": plain_prev["picards"] / args.steps,
                                  "max_head_err_vs_plain_fp64_m": head_err_pp,
                                  "note": "FP64 PCG, every solve from x0 = h^n (then h^m): the plain method with "
                                          "warm starts"},
        "fast_picard_iters_per_step": fast["picards"] / args.steps,
        "speedup_vs_plain_fp64": round(value / plain_value, 3),
        "speedup_vs_plain_fp64_warm_start": round(plain_prev["elapsed"] / fast["elapsed"], 3),
        "speedup_learned_vs_extrapolated": round(ext["elapsed"] / fast["elapsed"], 3),
        "max_head_err_vs_plain_fp64_m": head_err,
        "max_depth_err_vs_plain_fp64_m": depth_err(fast),
        "learned_gate_m": args.gate,
        "fast_inner_iters_per_step": fast["iters"] / args.steps,
        "fast_refinements_per_step": fast["refines"] / args.steps,
        "fast_guess_fallbacks": fast["fallbacks"],
        "max_rel_mass_balance_error": {"learned": fast["mass_err"], "plain_fp64": plain["mass_err"],
                                       "previous": prev["mass_err"], "extrapolated": ext["mass_err"]},
        "previous_guess_mixed": {"value": round(case.cells * args.steps / prev["elapsed"], 1),
                                 "ms_per_step": round(prev["elapsed"] / args.steps * 1e3, 4),
                                 "picard_iters_per_step": prev["picards"] / args.steps,
                                 "inner_iters_per_step": prev["iters"] / args.steps,
                                 "refinements_per_step": prev["refines"] / args.steps,
                                 "max_head_err_vs_plain_fp64_m": head_err_prev,
                                 "max_depth_err_vs_plain_fp64_m": depth_err(prev)},
        "extrapolated_guess_mixed": {"value": round(case.cells * args.steps / ext["elapsed"], 1),
                                     "ms_per_step": round(ext["elapsed"] / args.steps * 1e3, 4),
                                     "picard_iters_per_step": ext["picards"] / args.steps,
                                     "inner_iters_per_step": ext["iters"] / args.steps,
                                     "refinements_per_step": ext["refines"] / args.steps,
                                     "max_head_err_vs_plain_fp64_m": head_err_ext,
                                     "max_depth_err_vs_plain_fp64_m": depth_err(ext)},
        "training": training,
        "predictor_step": predictor,
        "roofline": roof64,
        "roofline_kernel": "FP64 CG SpMV of the plain path (k_spmv_pair<double> at the autotuned shape): "
                           "fused p = z + beta p_prev, 7-point q = A p, partial p.q (64 B per cell)",
        "roofline_fp32_spmv": roof32,
        "stream_probe_6r2w": stream,
        "roofline_mlp_guess": roof_guess,
        "spmv_shape_cells_per_lane_kchunk": {"plain_fp64": plain["spmv_shape"], "fast_fp32": fast["spmv_shape"]},
        "kernel_ms_per_step_fast": {k: round(v[0] / BREAKDOWN_STEPS, 4) for k, v in fast["kt"].items()},
        "kernel_ms_per_step_plain": {k: round(v[0] / BREAKDOWN_STEPS, 4) for k, v in plain["kt"].items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
