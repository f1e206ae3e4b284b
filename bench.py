#!/usr/bin/env python3
"""Benchmark of the PSTM hot path on one MI355X.

Workload (the 3-D PSTM configuration): 1,000,000 traces x 1,500 samples at
4 ms, midpoints uniform over a 200 x 200 grid of 25 m bins, offsets uniform in
0-4000 m, azimuths uniform; v_rms(x,y,t0) = 1500 + 600 t0 + 200 sin(x/2000)
cos(y/3000); 200 point diffractors (30 Hz Ricker) plus Gaussian noise at
20 dB; image 200 x 200 x 1,500 samples (t0 every 4 ms, one-way tau every
2 ms); aperture 1500 m (midpoint-to-bin distance).

One step = one pass of the hot path over the resident survey: build the
plan (midpoint binning and per-tile aperture lists), rho-filter the traces
(2.5-D half derivative) with the trace-domain part of the weight (sample n
times sqrt(n dt)) fused in, quantise them to FP16 with per-trace FP32
scales, and migrate (LINEAR interpolation, FAST math, FP32 accumulation,
weight tau per update). The FP16 form with the per-update weight tau*sqrt(t)
(no gain), the plain FP32 form (FP32 traces, per-update weight) and the FP32
form of the headline method (FP32 gained traces) are timed the same way and
reported beside the headline, with the FP16/FP32 ratio of each method.

value = updates / s of the headline form, where one update is one (trace,
in-aperture bin, output sample) triple counted by the plan. Every form's
image is compared with the FP64 oracle of the per-update weight (the
reference's operator with LINEAR interpolation) on 48 bins scattered over the
central 40 x 40-bin sub-volume, the edges and the corners (or on a central
N x N sub-volume, --oracle-bins N) (rel-L2), and the headline form also with
its own FP64 operator on its stored values. The reference's CPU loop (the oracle's binary32
restatement of src/psktmCPU.c) is timed on a bounded sample of the same
workload on the host cores. Only rank 0 of one GPU runs.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "seismic-migration_amd")]

METRIC = "G trace-sample updates/s: FP16-trace vs FP32 plain, rel-L2≤1e-3 vs FP64"
# Algorithmic VALU work per update (the roofline's unit): 2 FMA
# (tau'^2 + A s', tau'^2 + B s'), 3 v_sqrt_f32, 1 add, 1 index bound (the
# validity test of the sample index: a v_min clamp), 1 cvt, 1 fract,
# 1 lshl_add (LDS address), 1 FMA (interpolation), 1 FMA (accumulate). Tiles
# whose traveltimes are bounded inside the padded LDS rows run the loop
# compiled without the clamp (11 issued instructions per update); the
# roofline keeps the algorithm's 12.
VALU_OPS_PER_UPDATE = 12
# The gained-trace form (sqrt(t) applied per trace sample by the rho filter)
# has no per-update weight sqrt: 11 algorithmic VALU instructions, 2 of them
# v_sqrt_f32.
VALU_OPS_PER_UPDATE_GAINED = 11
FORM_DTYPES = {"f16g": "f16", "f16": "f16", "bf16": "bf16", "f32": "f32", "f32g": "f32"}
# Conventional FLOP count per update (FMA = 2, sqrt = 1, add = 1).
FLOPS_PER_UPDATE = 12
FLOPS_PER_UPDATE_GAINED = 10      # 2 FMA, 2 sqrt, add, interpolation FMA, add
# MI355X: 256 CUs, 2.4 GHz, 64 FP32 lanes per CU per clock without packed
# math; published FP32 vector peak (packed FMA) 157.3 TFLOP/s.
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
FP32_PUBLISHED_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--traces", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64 comparison")
    ap.add_argument("--cpu-traces", type=int, default=10000,
                    help="traces in the CPU-baseline sample")
    ap.add_argument("--oracle-bins", type=int, default=0,
                    help="0: compare 48 bins scattered over the central 40x40 sub-volume, the "
                         "edges and the corners with FP64; N > 0: the central N x N bins")
    ap.add_argument("--forms", default="f16g,f16,f32,f32g",
                    help="timed forms: f16g (FP16, gained traces; the headline), f16, "
                         "bf16, f32 (FP32 plain), f32g (FP32, gained traces); the first is "
                         "the headline")
    ap.add_argument("--workload", default="pstm3d",
                    choices=["pstm3d", "batched", "lsm", "line2d", "rho", "rates"],
                    help="pstm3d: the headline 3-D workload (default); batched: 64 offset-class "
                         "gathers in one persistent launch; lsm: least-squares 2-D line, 20 CG "
                         "iterations; line2d: the 2-D line, reference-exact GPU form vs the "
                         "reference CPU loop; rho: the rho filter alone on the headline's "
                         "traces; rates: gfx950 VALU, v_sqrt_f32, LDS gather and "
                         "LDS atomic rates behind the roofline figures")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache (batched workload)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def calibrate_copy(nbytes):
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4}), flush=True)


def scattered_bins(nx, ny, seed=20261205):
    """48 image bins for the FP64 comparison: 24 in the central 40 x 40-bin
    sub-volume (bins 80-119), the 4 corners, the 4 edge midpoints, 8 more on
    the edges and 8 anywhere."""
    import numpy as np
    rng = np.random.default_rng(seed)
    c = nx // 2 - 20
    bins = [(int(i), int(j)) for i, j in rng.integers(c, c + 40, (24, 2))]
    bins += [(0, 0), (0, ny - 1), (nx - 1, 0), (nx - 1, ny - 1)]
    bins += [(0, ny // 2), (nx - 1, ny // 2), (nx // 2, 0), (nx // 2, ny - 1)]
    for k in range(8):
        t = int(rng.integers(1, nx - 1))
        bins.append([(0, t), (nx - 1, t), (t, 0), (t, ny - 1)][k % 4])
    bins += [(int(i), int(j)) for i, j in zip(rng.integers(0, nx, 8), rng.integers(0, ny, 8))]
    return bins


def host_cpu():
    """CPU model and logical CPUs of this host (recorded beside the CPU baseline)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload != "pstm3d":
        if args.gpus != 1:
            print(json.dumps({"error": "one GPU only"}))
            sys.exit(2)
        {"batched": bench_batched, "lsm": bench_lsm, "line2d": bench_line2d,
         "rho": bench_rho, "rates": bench_rates}[args.workload](args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this PSTM path runs on one GPU; multi-GPU sharding is a "
                                   "design note (trace-sharded partial images + one all-reduce), "
                                   "not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import psktm
    from psktm import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    props = torch.cuda.get_device_properties(dev)
    log(f"device {props.name} {getattr(props, 'gcnArchName', '')}")

    t0 = time.time()
    s = synth.diffractor_3d(args.traces, device="cuda")
    torch.cuda.synchronize()
    log(f"synthetic survey: {args.traces} traces x {s.traces.shape[1]} in {time.time() - t0:.1f} s")
    grid = psktm.Grid(*s.grid)
    ns, dt = s.traces.shape[1], s.dt
    dw_us = dt * 1e6
    survey = psktm.Survey(s.mid_x, s.mid_y, s.off_x, s.off_y, ns, dt)
    image = torch.empty(grid.shape, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def one_step(form, ev=None):
        plan = psktm.Plan(grid, survey, aperture=s.aperture)
        gained = form.endswith("g")
        if FORM_DTYPES[form] == "f32":
            tr = psktm.rho_filter(s.traces, dw_us, 2, gain=gained)
            sc = None
        else:
            dtype = torch.float16 if FORM_DTYPES[form] == "f16" else torch.bfloat16
            tr, sc = psktm.rho_filter(s.traces, dw_us, 2, out_dtype=dtype, gain=gained)
        if ev is not None:
            ev[0].record(stream)
        psktm.migrate(plan, tr, s.vel, image, scale=sc, gained=gained)
        if ev is not None:
            ev[1].record(stream)
        return plan, tr, sc

    plan0 = psktm.Plan(grid, survey, aperture=s.aperture)
    updates = plan0.updates()
    plan_bytes = plan0.nbytes()
    plan0.close()
    log(f"updates per migration {updates:.4e}; plan holds {plan_bytes / 2**30:.2f} GiB")

    results = {}
    for form in [f.strip() for f in args.forms.split(",") if f.strip()]:
        for _ in range(args.warmup):
            p, _, _ = one_step(form)
            torch.cuda.synchronize()
            p.close()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        plans = []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for k in range(args.steps):
            p, _, _ = one_step(form, evs[k])
            plans.append(p)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        for p in plans:
            p.close()
        kern_ms = [a.elapsed_time(b) for a, b in evs]
        results[form] = dict(wall=wall, ms_per_step=1e3 * wall / args.steps,
                             kernel_ms=float(np.mean(kern_ms)),
                             rate=updates * args.steps / wall)
        log(f"{form}: {results[form]['ms_per_step']:.1f} ms/step, migrate kernel "
            f"{results[form]['kernel_ms']:.1f} ms, {results[form]['rate'] / 1e9:.1f} G updates/s")

    # ---- accuracy vs the FP64 oracle ------------------------------------------
    acc = {}
    acc_bins = None
    if not args.no_oracle:
        import oracle
        nb = args.oracle_bins
        if nb > 0:
            c0x, c0y = grid.nx // 2 - nb // 2, grid.ny // 2 - nb // 2
            bins = [(c0x + i, c0y + j) for i in range(nb) for j in range(nb)]
            acc_bins = f"central {nb}x{nb} bins (x {c0x}-{c0x + nb - 1}, y {c0y}-{c0y + nb - 1})"
        else:
            bins = scattered_bins(grid.nx, grid.ny)
            acc_bins = ("48 bins: 24 in the central 40x40 sub-volume (80-119), 4 corners, "
                        "12 on the edges, 8 anywhere")
        f32_traces = psktm.rho_filter(s.traces, dw_us, 2)
        host = lambda t: t.cpu().numpy()
        og = oracle.Grid(*s.grid)
        okw = dict(vel3d=True, aperture=s.aperture, interp=oracle.LINEAR)
        hmx, hmy, hox, hoy, hvel = (host(a) for a in (s.mid_x, s.mid_y, s.off_x, s.off_y, s.vel))
        t1 = time.time()
        htr = host(f32_traces)
        ref = None
        for c in range(0, len(bins), 64):        # progress lines on the large sub-volumes
            part = oracle.pstm_f64_bins(og, hmx, hmy, hox, hoy, htr, hvel, dt,
                                        bins[c:c + 64], **okw)
            ref = part if ref is None else np.concatenate([ref, part])
            log(f"FP64 oracle: {min(c + 64, len(bins))}/{len(bins)} bins ({time.time() - t1:.0f} s)")
        del htr
        log(f"FP64 oracle on {len(bins)} bins x {grid.nt} in {time.time() - t1:.1f} s")
        bi = torch.as_tensor([b[0] for b in bins], device=dev)
        bj = torch.as_tensor([b[1] for b in bins], device=dev)
        plan = psktm.Plan(grid, survey, aperture=s.aperture)
        for form in ("f16g", "f32g", "f32", "f16", "bf16"):
            gained = form.endswith("g")
            if FORM_DTYPES[form] == "f32":
                src = psktm.trace_gain(f32_traces, dt) if gained else f32_traces
                img = psktm.migrate(plan, src, s.vel, gained=gained)
            else:
                dtype = torch.float16 if FORM_DTYPES[form] == "f16" else torch.bfloat16
                if form == "f16g":     # the headline path: rho filter with the fused gain
                    q, sc = psktm.rho_filter(s.traces, dw_us, 2, out_dtype=dtype, gain=True)
                else:
                    q, sc = psktm.quantize(f32_traces, dtype)
                img = psktm.migrate(plan, q, s.vel, scale=sc, gained=gained)
            sub = img[bi, bj].double().cpu().numpy()
            acc[form] = dict(rel_l2=float(np.linalg.norm(sub - ref) / np.linalg.norm(ref)),
                             max_abs=float(np.max(np.abs(sub - ref))),
                             max_abs_ref=float(np.max(np.abs(ref))))
            if form == "f16g":
                # its own FP64 operator (weight tau) on the dequantised stored values
                deq = (q.float() * sc[:, None]).cpu().numpy()
                own = oracle.pstm_f64_bins(og, hmx, hmy, hox, hoy, deq, hvel, dt, bins,
                                           gained=True, **okw) if nb == 0 else None
                del deq
                if own is not None:
                    acc[form]["rel_l2_vs_own_fp64_on_stored"] = float(
                        np.linalg.norm(sub - own) / np.linalg.norm(own))
            log(f"{form}: rel-L2 vs FP64 {acc[form]['rel_l2']:.3e}")
        plan.close()
        del f32_traces

    # ---- CPU baseline: the reference loop on a bounded sample ---------------
    cpu = None
    if not args.no_cpu:
        import oracle
        k1 = min(args.cpu_traces, args.traces)
        mx, my = s.mid_x[:k1].cpu().numpy(), s.mid_y[:k1].cpu().numpy()
        hx, hy = s.off_x[:k1].cpu().numpy(), s.off_y[:k1].cpu().numpy()
        trh = s.traces[:k1].cpu().numpy()
        vel = s.vel.cpu().numpy()
        xs = np.arange(grid.nx, dtype=np.float32) * np.float32(grid.dx)
        ys = np.arange(grid.ny, dtype=np.float32) * np.float32(grid.dy)
        r2 = np.float32(s.aperture) * np.float32(s.aperture)
        pairs = 0
        for k in range(k1):
            ex = mx[k] - xs
            ey = my[k] - ys
            pairs += int(np.count_nonzero((ex * ex)[:, None] + (ey * ey)[None, :] <= r2))
        cores = len(os.sched_getaffinity(0))
        t1 = time.perf_counter()
        oracle.pstm_f32(oracle.Grid(*s.grid), mx, my, hx, hy, trh, vel, dt, vel3d=True,
                        aperture=s.aperture, threads=cores)
        el = time.perf_counter() - t1
        cpu = {"value": pairs * grid.nt / el / 1e9, "unit": "G trace-sample updates/s",
               "cores": cores, "kind": "port", "host": host_cpu(),
               "sample": f"first {k1} traces of the workload, reference TRUNC binary32 loop "
                         f"with the same 1500 m aperture ({pairs * grid.nt:.3e} updates, "
                         f"{el:.1f} s)"}
        log(f"CPU baseline {cpu['value']:.3f} G updates/s on {cores} threads ({el:.1f} s)")

    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # runs of this same workload (scaled by the copy-kernel calibration)
    head_form = next(iter(results))
    head = results[head_form]
    traffic = None
    tp = os.path.join(ROOT, "profiles", "migrate_fwd_3d_1M_traces_traffic.json")
    if os.path.exists(tp):
        prof = json.load(open(tp))
        if prof["workload"]["traces"] == args.traces and \
                prof["workload"].get("form") == head_form and \
                abs(prof["updates_per_launch"] - updates) <= 1e-9 * updates:
            traffic = prof["hbm_bytes_per_launch"]

    f32 = results.get("f32")
    f16 = results.get("f16")
    f32g = results.get("f32g")
    f16g = results.get("f16g")
    kern_s = head["kernel_ms"] / 1e3
    per_update = VALU_OPS_PER_UPDATE_GAINED if head_form.endswith("g") else VALU_OPS_PER_UPDATE
    valu_ops = updates * per_update / kern_s
    flops = updates * (FLOPS_PER_UPDATE_GAINED if head_form.endswith("g")
                       else FLOPS_PER_UPDATE) / kern_s
    line = {
        "metric": METRIC,
        "value": head["rate"] / 1e9,
        "unit": "G trace-sample updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3-D PSTM: 1e6 traces x 1500 samples (4 ms) -> 200x200 bins x 1500 "
                        "tau, aperture 1500 m, v(x,y,t0); step = plan + rho filter (fused "
                        "sqrt(n dt) gain for gained forms) + quantise + migration (LINEAR, "
                        "FAST)",
            "traces": args.traces, "samples": ns, "image": list(grid.shape),
            "aperture_m": s.aperture, "headline_form": head_form,
            "trace_storage": {"f16g": "f16 + per-trace f32 scale, samples gained by sqrt(n dt)",
                              "f16": "f16 + per-trace f32 scale", "bf16": "bf16 + per-trace "
                              "f32 scale", "f32": "f32", "f32g": "f32, gained"}[head_form],
            "updates_per_step": updates,
        },
        "forms": {k: {"G_updates_per_s": v["rate"] / 1e9, "ms_per_step": v["ms_per_step"],
                      "migrate_kernel_ms": v["kernel_ms"]} for k, v in results.items()},
        "fp16_over_fp32": (head["rate"] / f32["rate"]) if f32 else None,
        "fp16_over_fp32_same_method": {
            "gained_traces": (f16g["rate"] / f32g["rate"]) if (f16g and f32g) else None,
            "per_update_weight": (f16["rate"] / f32["rate"]) if (f16 and f32) else None},
        "accuracy_vs_fp64": acc or None,
        "accuracy_bins": acc_bins,
        "roofline": {
            "bound": "valu",
            "achieved": valu_ops / 1e12,
            "peak": VALU_PEAK_LANE_OPS / 1e12,
            "unit": "T VALU lane-op/s",
            "frac": valu_ops / VALU_PEAK_LANE_OPS,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
            "hbm_GBps": (traffic / kern_s / 1e9) if traffic else None,
            "kernel": "migrate_fwd_kernel",
            "form": head_form,
            "per_update": (f"{per_update} VALU instructions ({3 - (per_update < 12)} v_sqrt_f32; "
                           f"{per_update - 1} issued where the index clamp is not needed)"),
            # the same rate counted at the loop's issued instructions without
            # the clamp (every tile of this workload proves its bound)
            "frac_at_issued_loop": valu_ops * (per_update - 1) / per_update / VALU_PEAK_LANE_OPS,
            "fp32_published": {"achieved_tflops": flops / 1e12,
                               "peak_tflops": FP32_PUBLISHED_TFLOPS,
                               "frac": flops / 1e12 / FP32_PUBLISHED_TFLOPS},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Secondary workloads (their own JSON line; the default run is pstm3d).
# ---------------------------------------------------------------------------

def bench_batched(args):
    """64 offset-class gathers (lognormal 2,000-60,000 traces, patches of
    32-128 bins at 25 m x 1,000 tau, 1,000 samples at 4 ms, aperture 1500 m),
    FP16 traces + FP32 scales, gained by sqrt(n dt) in the rho filter (as the
    headline form), migrated by ONE persistent launch replayed from a HIP
    graph, tile shape from the autotuner; compared with launching the same
    gathers one plan at a time."""
    import numpy as np
    import torch

    import psktm
    from psktm import synth
    from psktm.tune import CANDIDATES_3D, Autotuner, signature, time_gpu

    dev = torch.device("cuda:0")
    t0 = time.time()
    gd = synth.offset_class_gathers(64, device="cuda")
    torch.cuda.synchronize()
    log(f"64 gathers, {sum(gd.counts)} traces x {gd.traces.shape[1]} in {time.time() - t0:.1f} s")
    ns = gd.traces.shape[1]
    q, sc = psktm.rho_filter(gd.traces, gd.dt * 1e6, 2, out_dtype=torch.float16, gain=True)
    survey = psktm.Survey(gd.mid_x, gd.mid_y, gd.off_x, gd.off_y, ns, gd.dt)
    gathers = [(psktm.Grid(*g), f, c) for g, f, c in zip(gd.grids, gd.first, gd.counts)]

    # autotune on every fourth gather (a quarter of the work)
    sub = gathers[::4]
    key = signature("batch", ns=ns, nt=gd.grids[0][6], dtype="f16", gained=True,
                    aperture=gd.aperture, gathers=len(gathers), seed=gd.seed)
    tuner = Autotuner(args.tune_cache)

    def trial(c):
        b = psktm.Batch(sub, survey, aperture=gd.aperture, tile_x=c[0], tile_y=c[1],
                        traces_per_stage=c[2])
        t = time_gpu(lambda: b.migrate(q, gd.vel, scale=sc, gained=True), repeats=1)
        log(f"  tune {c}: {t * 1e3:.1f} ms")
        b.close()
        return t

    # the three quickest trials timed again on all 64 gathers: the quarter's
    # ranking sits within the trials' spread
    def whole(c):
        b = psktm.Batch(gathers, survey, aperture=gd.aperture, tile_x=c[0], tile_y=c[1],
                        traces_per_stage=c[2])
        out = torch.empty(int(sum(b.sizes)), dtype=torch.float32, device=dev)
        t = time_gpu(lambda: b.migrate(q, gd.vel, out, scale=sc, use_graph=True, gained=True),
                     repeats=1)
        log(f"  whole batch {c}: {t * 1e3:.1f} ms")
        b.close()
        return t

    t1 = time.time()
    choice = tuner.choose(key, CANDIDATES_3D, trial, refine_fn=whole, refine_top=3)
    log(f"autotuner choice {choice} ({time.time() - t1:.1f} s, cache {tuner.cache_path})")
    batch = psktm.Batch(gathers, survey, aperture=gd.aperture, tile_x=choice[0],
                        tile_y=choice[1], traces_per_stage=choice[2])
    updates = batch.updates()
    images = torch.empty(int(sum(batch.sizes)), dtype=torch.float32, device=dev)
    for _ in range(max(1, args.warmup)):
        batch.migrate(q, gd.vel, images, scale=sc, use_graph=True, gained=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        batch.migrate(q, gd.vel, images, scale=sc, use_graph=True, gained=True)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t2) / args.steps
    # the same gathers launched one plan at a time: with one workgroup per
    # tile (the batch kernel's summation order, so images must be identical)
    # and with the library's choice (split tile lists for small gathers)
    def sequential_run(splits):
        plans = []
        for g, f, c in gathers:
            sl = slice(f, f + c)
            s = psktm.Survey(gd.mid_x[sl].contiguous(), gd.mid_y[sl].contiguous(),
                             gd.off_x[sl].contiguous(), gd.off_y[sl].contiguous(), ns, gd.dt)
            plans.append((psktm.Plan(g, s, aperture=gd.aperture, tile_x=choice[0],
                                     tile_y=choice[1], traces_per_stage=choice[2],
                                     splits=splits), sl))
        imgs = [torch.empty(g.shape, dtype=torch.float32, device=dev) for g, _, _ in gathers]

        def sequential():
            for (p, sl), img in zip(plans, imgs):
                psktm.migrate(p, q[sl], gd.vel, img, scale=sc[sl], gained=True)

        sequential()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(args.steps):
            sequential()
        torch.cuda.synchronize()
        return (time.perf_counter() - t3) / args.steps, imgs

    seq, seq_imgs = sequential_run(1)
    same = all(torch.equal(batch.image(images, k), seq_imgs[k]) for k in range(len(gathers)))
    del seq_imgs
    seq_split, _ = sequential_run(0)
    line = {
        "metric": "G trace-sample updates/s (batched offset-class gathers, one persistent launch)",
        "value": updates / wall / 1e9, "unit": "G trace-sample updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 offset-class gathers, lognormal(ln 11000, 0.8) traces clipped "
                               "to 2,000-60,000, patches 32-128 bins at 25 m x 1,000 tau, "
                               "1,000 samples at 4 ms, aperture 1500 m, FP16 traces gained by "
                               "sqrt(n dt)",
                   "traces": int(sum(gd.counts)), "updates_per_step": updates,
                   "tile": list(choice), "graph": True},
        "sequential_ms": seq * 1e3, "persistent_over_sequential": seq / wall,
        "images_equal_sequential": bool(same),
        "sequential_split_ms": seq_split * 1e3,
        "persistent_over_sequential_split": seq_split / wall,
        "autotuner": tuner.cache.get(key),
    }
    print(json.dumps(line), flush=True)


def _lsm_line(dev, n_cmp=512, n_off=64, ns=2000, dt=0.002, nt=2000, seed=None):
    import numpy as np
    import torch

    import psktm
    from psktm import synth
    seed = synth.SEED0 + 4 if seed is None else seed
    d = synth.line_2d(seed=seed, n_cmp=n_cmp, n_off=n_off, ns=ns, nt=nt, dt=dt)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)
    tau = np.arange(nt, dtype=np.float32) * np.float32(d.grid[8])
    vel = t(1500.0 + 500.0 * 2.0 * tau.astype(np.float64))
    survey = psktm.Survey(t(d.mid_x), t(d.mid_y), t(d.off_x), t(d.off_y), ns, dt)
    grid = psktm.Grid(*d.grid)
    # reflectivity: 8 random-dip interfaces + 50 point scatterers
    rng = np.random.default_rng(seed)
    m = np.zeros(grid.shape, np.float32)
    for _ in range(8):
        l0, slope, a = rng.integers(100, nt - 100), rng.uniform(-1.0, 1.0), rng.choice([-1, 1])
        for i in range(grid.nx):
            l = int(round(l0 + slope * (i - grid.nx / 2)))
            if 0 < l < nt:
                m[i, 0, l] += a
    for _ in range(50):
        m[rng.integers(0, grid.nx), 0, rng.integers(50, nt - 50)] += rng.choice([-2.0, 2.0])
    return grid, survey, vel, t(m)


def bench_lsm(args):
    """Least-squares 2-D PSTM: 512 CMPs x 64 offsets (32,768 traces x 2,000
    samples at 2 ms), image 512 x 2,000 (tau step 1 ms), data = demigration of
    8 random-dip interfaces + 50 point scatterers, 20 CGLS iterations with FP64
    inner products (precise accumulation). Dot-product test and finite
    differences on a random 16-CMP slice. The same least squares on the gained
    operator (the trace gain sqrt(n dt) applied to the samples, weight tau per
    update in both kernels) is reported beside it."""
    import numpy as np
    import torch

    import psktm
    from psktm import lsm

    dev = torch.device("cuda:0")
    grid, survey, vel, m_true = _lsm_line(dev)
    plan = psktm.Plan(grid, survey, precise=True)
    upd = plan.updates()
    data = psktm.demigrate(plan, m_true, vel)
    torch.cuda.synchronize()
    # timing of the operators
    tf = time.perf_counter(); psktm.migrate(plan, data, vel); torch.cuda.synchronize()
    tf = time.perf_counter() - tf
    ta = time.perf_counter(); psktm.demigrate(plan, m_true, vel); torch.cuda.synchronize()
    ta = time.perf_counter() - ta
    t0 = time.perf_counter()
    m, hist = lsm.cgls(plan, vel, data, iters=20)
    cg_s = time.perf_counter() - t0
    mig = psktm.migrate(plan, data, vel)
    cos = lambda a, b: float((a * b).sum() / (a.norm() * b.norm()))
    # dot-product test and finite differences on a random 16-CMP slice
    rng = np.random.default_rng(5)
    c0 = int(rng.integers(0, 512 - 16))
    sg, ss, sv, _ = _lsm_line(dev, n_cmp=16)
    # shift the slice to CMPs c0..c0+15 (same offsets, same time axes)
    shift = np.float32(12.5 * c0)
    ss = psktm.Survey(ss.mid_x + shift, ss.mid_y, ss.off_x, ss.off_y, ss.n_samples, ss.dt)
    sg = psktm.Grid(sg.nx, float(shift), sg.dx, sg.ny, sg.y0, sg.dy, sg.nt, sg.tau0, sg.dtau)
    sp = psktm.Plan(sg, ss, precise=True)
    g = torch.Generator(device="cpu").manual_seed(7)
    x = torch.randn((ss.n_traces, ss.n_samples), generator=g).to(dev)
    y = torch.randn(sg.shape, generator=g).to(dev)
    lhs = psktm.dot_f64(psktm.migrate(sp, x, sv).reshape(-1), y.reshape(-1)).item()
    rhs = psktm.dot_f64(x.reshape(-1), psktm.demigrate(sp, y, sv).reshape(-1)).item()
    mm = y.clone().requires_grad_(True)
    dd = torch.randn((ss.n_traces, ss.n_samples), generator=g).to(dev)
    e = torch.randn(sg.shape, generator=g).to(dev)
    f = lambda v: 0.5 * ((lsm.Demigrate.apply(v, sp, sv) - dd).double() ** 2).sum()
    L = f(mm); L.backward()
    gd = float((mm.grad.double() * e.double()).sum())
    with torch.no_grad():
        fd = float((f(mm + 0.5 * e) - f(mm - 0.5 * e)) / 1.0)
    # the same least squares on the gained operator (trace gain + gained
    # migration, psktm_demigrate_gained): one v_sqrt_f32 fewer per update
    gdata = psktm.demigrate(plan, m_true, vel, gained=True)
    gtr = psktm.trace_gain(gdata, survey.dt)
    torch.cuda.synchronize()
    tgf = time.perf_counter(); psktm.migrate(plan, gtr, vel, gained=True); torch.cuda.synchronize()
    tgf = time.perf_counter() - tgf
    tga = time.perf_counter(); psktm.demigrate(plan, m_true, vel, gained=True); torch.cuda.synchronize()
    tga = time.perf_counter() - tga
    t0 = time.perf_counter()
    gm, ghist = lsm.cgls(plan, vel, gdata, iters=20, gained=True)
    gcg_s = time.perf_counter() - t0
    glhs = psktm.dot_f64(psktm.migrate(sp, psktm.trace_gain(x, ss.dt), sv, gained=True).reshape(-1),
                         y.reshape(-1)).item()
    grhs = psktm.dot_f64(x.reshape(-1),
                         psktm.demigrate(sp, y, sv, gained=True).reshape(-1)).item()
    gained = {
        "cg_iterations_per_s": 20.0 / gcg_s, "ms_per_iteration": gcg_s / 20 * 1e3,
        "migrate_ms": tgf * 1e3, "demigrate_ms": tga * 1e3,
        "migrate_G_updates_per_s": upd / tgf / 1e9, "demigrate_G_updates_per_s": upd / tga / 1e9,
        "residual_monotone": all(ghist[k + 1] <= ghist[k] * (1 + 1e-9)
                                 for k in range(len(ghist) - 1)),
        "residual_norm2_first_last": [ghist[0], ghist[-1]],
        "cos_image_vs_model": cos(gm, m_true),
        "dot_product_test_rel": abs(glhs - grhs) / abs(glhs),
    }
    line = {
        "metric": "least-squares PSTM: 20 CGLS iterations (FP64 inner products)",
        "value": 20.0 / cg_s, "unit": "CG iterations/s", "n_gpus": 1, "steps": 20, "warmup": 0,
        "ms_per_step": cg_s / 20 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2-D line 512 CMPs x 64 offsets 0-3150 m, 32,768 traces x 2,000 "
                               "samples (2 ms), image 512 x 2,000 (1 ms), v = 1500 + 500 t0",
                   "updates_per_operator": upd},
        "migrate_ms": tf * 1e3, "demigrate_ms": ta * 1e3,
        "migrate_G_updates_per_s": upd / tf / 1e9, "demigrate_G_updates_per_s": upd / ta / 1e9,
        "residual_norm2": hist, "residual_monotone": all(hist[k + 1] <= hist[k] * (1 + 1e-9)
                                                         for k in range(len(hist) - 1)),
        "cos_image_vs_model": {"cgls": cos(m, m_true), "migration": cos(mig, m_true)},
        "dot_product_test": {"slice_first_cmp": c0, "lhs": lhs, "rhs": rhs,
                             "rel": abs(lhs - rhs) / abs(lhs)},
        "finite_difference": {"fd": fd, "grad_dot": gd, "rel": abs(fd - gd) / abs(gd)},
        "gained_operator": gained,
    }
    print(json.dumps(line), flush=True)


def bench_line2d(args):
    """The 2-D line (64 CMPs x 32 offsets, 1,000 samples at 4 ms): the
    reference-exact GPU form (bit-identical to the reference loop) timed
    against the reference's CPU loop on the host (full workload, best of 3),
    plus the fast LINEAR form."""
    import numpy as np
    import torch

    import oracle
    import psktm
    from psktm import synth

    dev = torch.device("cuda:0")
    d = synth.line_2d()
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)
    grid = psktm.Grid(*d.grid)
    ns = d.traces.shape[1]
    survey = psktm.Survey(t(d.mid_x), t(d.mid_y), t(d.off_x), t(d.off_y), ns, d.dt)
    tr, vel = t(d.traces), t(d.vel)
    updates = float(d.traces.shape[0]) * grid.nx * grid.ny * grid.nt
    img = psktm.migrate_exact(grid, survey, tr, vel)
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(3, args.steps)):
        t0 = time.perf_counter()
        psktm.migrate_exact(grid, survey, tr, vel, img)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    exact_s = min(ts)
    plan = psktm.Plan(grid, survey)
    psktm.migrate(plan, tr, vel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(3, args.steps)):
        fast = psktm.migrate(plan, tr, vel)
    torch.cuda.synchronize()
    fast_s = (time.perf_counter() - t0) / max(3, args.steps)
    cores = len(os.sched_getaffinity(0))
    cpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        ref = oracle.pstm_f32(oracle.Grid(*d.grid), d.mid_x, d.mid_y, d.off_x, d.off_y, d.traces,
                              d.vel, d.dt, threads=cores)
        cpu.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    oracle.ktm_migration_racy(oracle.Grid(*d.grid), d.mid_x, d.mid_y, d.off_x, d.off_y,
                              d.traces, d.vel, d.dt, threads=cores)
    racy = time.perf_counter() - t0
    bit = bool(np.array_equal(img.cpu().numpy().view(np.uint32), ref.view(np.uint32)))
    line = {
        "metric": "G trace-sample updates/s (2-D line, reference-exact GPU form)",
        "value": updates / exact_s / 1e9, "unit": "G trace-sample updates/s", "n_gpus": 1,
        "steps": len(ts), "warmup": 1, "ms_per_step": exact_s * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2-D line 64 CMPs x 32 offsets (2,048 traces x 1,000 samples), "
                               "64 x 1,000 image, no aperture", "updates": updates},
        "bit_identical_to_reference_loop": bit,
        "fast_linear_G_updates_per_s": updates / fast_s / 1e9,
        "cpu_baseline": {"value": updates / min(cpu) / 1e9, "unit": "G trace-sample updates/s",
                         "cores": cores, "kind": "port", "host": host_cpu(),
                         "sample": "full 2-D line, best of 3, race-free restatement"},
        "cpu_racy_as_written_G_updates_per_s": updates / racy / 1e9,
    }
    print(json.dumps(line), flush=True)


def bench_rho(args):
    """The rho filter alone (scripts/ktmApplyRhoFilter.m, opt 2) on the
    headline's traces (1e6 x 1,500 samples at 4 ms): FP16 out with the fused
    sqrt(n dt) gain (what the headline migrates), then FP32 out. Per call:
    HIP events between back-to-back calls on the launch stream, against the
    HBM roofline (4 B read and 2 or 4 B written per sample, plus one 4 B
    scale per trace for FP16). Parity: the first 2,000 traces' FP32 output
    against the oracle's restatement; CPU baseline: the same restatement
    (numpy FFTs, one thread) on the first 20,000 traces."""
    import numpy as np
    import torch

    import oracle
    import psktm
    from psktm import synth

    s = synth.diffractor_3d(args.traces, device="cuda")
    tr = s.traces
    n, m = tr.shape
    dw_us = s.dt * 1e6
    stream = torch.cuda.current_stream()
    steps = max(args.steps, 10)
    forms = {}
    for form, kw, out_b in (("f16g", dict(out_dtype=torch.float16, gain=True), 2),
                            ("f32", {}, 4)):
        for _ in range(max(args.warmup, 2)):
            psktm.rho_filter(tr, dw_us, 2, **kw)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(stream)
        for k in range(steps):
            psktm.rho_filter(tr, dw_us, 2, **kw)
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]
        nbytes = n * m * (4 + out_b) + (4 * n if out_b == 2 else 0)
        mean = float(np.mean(ms))
        forms[form] = {"ms_per_call": mean, "best_ms": float(min(ms)), "hbm_bytes": nbytes,
                       "GB_per_s": nbytes / mean / 1e6, "frac_of_8TB_per_s": nbytes / mean / 8e9}
        log(f"rho {form}: {mean:.3f} ms per call (best {min(ms):.3f}), "
            f"{nbytes / mean / 1e6:.0f} GB/s")
    k = min(n, 2000) // 2 * 2
    got = psktm.rho_filter(tr[:k].contiguous(), dw_us, 2).cpu().numpy().astype(np.float64)
    host = tr[:20000].cpu().numpy()
    ref = oracle.apply_rho_filter(host[:k], dw_us, 2)      # even count: no trace rotation
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    t0 = time.perf_counter()
    for c in range(0, host.shape[0], 1000):
        oracle.apply_rho_filter(host[c:c + 1000], dw_us, 2)
    cpu_s = time.perf_counter() - t0
    head = forms["f16g"]
    traffic = None
    tp = os.path.join(ROOT, "profiles", "rho_filter_1M_traces_traffic.json")
    if os.path.exists(tp):
        prof = json.load(open(tp))
        if prof["workload"]["traces"] == n and prof["workload"]["samples"] == m:
            traffic = prof["hbm_bytes_per_launch"]
    line = {
        "metric": "rho filter: trace samples per second (opt 2, FP16 gained out)",
        "value": n * m / head["ms_per_call"] / 1e6, "unit": "G samples/s", "n_gpus": 1,
        "steps": steps, "warmup": max(args.warmup, 2), "ms_per_step": head["ms_per_call"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rho filter, {n} traces x {m} samples (4 ms), 2048-point FFTs"},
        "roofline": {"bound": "hbm", "achieved": head["GB_per_s"], "peak": 8000.0,
                     "unit": "GB/s", "frac": head["frac_of_8TB_per_s"],
                     "traffic": traffic},
        "forms": forms,
        "parity": {"traces": k, "rel_l2_fp32_vs_oracle": rel},
        "cpu_baseline": {"value": host.shape[0] * m / cpu_s / 1e9, "unit": "G samples/s",
                         "cores": 1, "kind": "port", "host": host_cpu(),
                         "sample": f"first {host.shape[0]} traces, numpy restatement in "
                                   "1,000-trace blocks (FP64 FFTs)"},
    }
    print(json.dumps(line), flush=True)


def bench_rates(args):
    """gfx950 microbenchmarks (seismic-migration_amd/csrc/rates/rates.hip):
    8 independent FMA or v_sqrt_f32 chains per lane, v_sqrt_f32 mixed with
    FMAs, random ds_read_b64 gathers, and LDS atomic adds (FP32, FP64, u32, u64)
    with lane-consecutive, 4-apart (the adjoint's lanes own 4 consecutive tau
    samples) and random slots. 2,048 workgroups of 256 lanes (8 waves per
    SIMD); the time is the second of two launches."""
    import ctypes

    import torch
    torch.cuda.init()
    lib = ctypes.CDLL(os.path.join(ROOT, "seismic-migration_amd", "psktm", "librates.so"))
    lib.psktm_rates_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_float)]
    blocks, lanes = 2048, 2048 * 256

    def run(kind, iters):
        ms = ctypes.c_float(0)
        rc = lib.psktm_rates_run(kind, blocks, iters, ctypes.byref(ms))
        if rc:
            raise RuntimeError(f"rates kind {kind}: rc {rc}")
        return ms.value / 1e3

    res = {}
    t = run(0, 65536)
    res["v_fma_f32_lane_ops_per_s"] = lanes * 8 * 65536 / t
    for kind, name in ((5, "v_fma_f32_vgpr"), (6, "v_add_f32_vgpr"), (7, "v_mul_f32_vgpr")):
        for iters in (4096, 65536):
            res[f"{name}_iters{iters}_lane_ops_per_s"] = lanes * 8 * iters / run(kind, iters)
    for kind, name, fmas in ((1, "sqrt_alone", 0), (2, "sqrt_with_8_fma", 8),
                             (3, "sqrt_with_64_fma", 64)):
        t = run(kind, 2048)
        res[f"v_sqrt_f32_{name}_per_s"] = lanes * 8 * 2048 / t
        # issue slots: FMA-equivalent time of 8 sqrt (+ their v_add) and the FMAs
        res[f"{name}_fma_slots_per_8_sqrt"] = (t / (lanes * 2048)) * \
            res["v_fma_f32_lane_ops_per_s"] - fmas
    t = run(4, 2048)
    res["ds_read_b64_random_per_s"] = lanes * 8 * 2048 / t
    res["ds_read_b64_random_TBps"] = res["ds_read_b64_random_per_s"] * 8 / 1e12
    for ty, tname in ((1, "f32"), (2, "f64"), (3, "u32"), (4, "u64")):
        for p, pname in ((0, "consecutive"), (1, "stride4"), (2, "random")):
            t = run(10 * ty + p, 512)
            res[f"ds_add_{tname}_{pname}_per_s"] = lanes * 8 * 512 / t
            log(f"ds_add_{tname} {pname}: {res[f'ds_add_{tname}_{pname}_per_s']:.3e}/s")
    # issue rates of the migration loop's instruction types at 2-8 resident
    # waves per SIMD (dynamic LDS per 256-lane workgroup limits residency)
    lib.psktm_rates_issue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_float)]
    ops = ["v_add_f32", "v_fma_f32", "v_cvt_i32_f32+v_cvt_f32_i32", "v_fract_f32",
           "v_lshl_add_u32", "v_fma_mix_f32", "v_fma_f32_sgpr", "v_add_f32+v_sqrt_f32"]
    issue = {}
    for op, name in enumerate(ops):
        for w in (2, 4, 6, 8):
            ms = ctypes.c_float(0)
            iters = 4096
            rc = lib.psktm_rates_issue(op, 256 * w * 4, iters, 160 * 1024 // w - 1024,
                                       ctypes.byref(ms))
            if rc:
                raise RuntimeError(f"issue probe {name}: rc {rc}")
            per = 16 if op == 2 else 8            # two conversions per chain step
            issue[f"{name}@{w}w"] = 256 * w * 4 * 256 * per * iters / (ms.value / 1e3)
            log(f"issue {name} at {w} waves/SIMD: {issue[f'{name}@{w}w']:.3e} lane-instr/s")
    res["issue_lane_instr_per_s"] = issue
    nominal = 256 * 64 * 2.4e9
    line = {"metric": "gfx950 rates (VALU, v_sqrt_f32, LDS gathers and atomics)",
            "value": res["v_fma_f32_lane_ops_per_s"] / nominal, "unit":
            "v_fma_f32 lane-ops/s over 256 CU x 64 lanes x 2.4 GHz", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic",
            "config": {"workload": "microbenchmarks, 2,048 workgroups x 256 lanes"},
            "rates": res}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
