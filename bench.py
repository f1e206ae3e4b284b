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
(2.5-D half derivative), quantise them to FP16 with per-trace FP32 scales,
and migrate (LINEAR interpolation, FAST math, FP32 accumulation). The plain
FP32 form is the same pass with FP32 traces and no quantisation; its rate is
reported beside the headline for the FP16/FP32 ratio.

value = updates / s of the FP16 form, where one update is one (trace,
in-aperture bin, output sample) triple counted by the plan. The image of the
FP16, BF16 and FP32 forms is compared with the FP64 oracle on a central
sub-volume of bins (rel-L2). The reference's CPU loop (the oracle's binary32
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
# Per-update VALU instruction count of migrate_fwd_kernel's inner loop as
# compiled for gfx950: 2 FMA (tau'^2 + A s', tau'^2 + B s'), 3 v_sqrt_f32,
# 1 add, 1 min (clamp), 1 cvt, 1 fract, 1 lshl_add (LDS address), 1 FMA
# (interpolation), 1 FMA (accumulate).
VALU_OPS_PER_UPDATE = 12
# Conventional FLOP count per update (FMA = 2, sqrt = 1, add = 1).
FLOPS_PER_UPDATE = 12
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
    ap.add_argument("--cpu-traces", type=int, default=4000,
                    help="traces in the CPU-baseline sample")
    ap.add_argument("--oracle-bins", type=int, default=4,
                    help="side of the central sub-volume compared with FP64")
    ap.add_argument("--forms", default="f16,f32", help="timed forms (f16, f32, bf16)")
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


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
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
        if form == "f32":
            tr = psktm.rho_filter(s.traces, dw_us, 2)
            sc = None
        else:
            dtype = torch.float16 if form == "f16" else torch.bfloat16
            tr, sc = psktm.rho_filter(s.traces, dw_us, 2, out_dtype=dtype)
        if ev is not None:
            ev[0].record(stream)
        psktm.migrate(plan, tr, s.vel, image, scale=sc)
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

    # ---- accuracy vs the FP64 oracle on a central sub-volume ---------------
    acc = {}
    if not args.no_oracle:
        import oracle
        nb = args.oracle_bins
        c0x, c0y = grid.nx // 2 - nb // 2, grid.ny // 2 - nb // 2
        bins = (c0x, c0x + nb, c0y, c0y + nb)
        f32_traces = psktm.rho_filter(s.traces, dw_us, 2)
        # traces that can reach the sub-volume (midpoint within the aperture
        # of its bin rectangle)
        x_lo, x_hi = c0x * grid.dx, (c0x + nb - 1) * grid.dx
        y_lo, y_hi = c0y * grid.dy, (c0y + nb - 1) * grid.dy
        ex = torch.clamp(torch.maximum(x_lo - s.mid_x, s.mid_x - x_hi), min=0)
        ey = torch.clamp(torch.maximum(y_lo - s.mid_y, s.mid_y - y_hi), min=0)
        sel = torch.nonzero(ex * ex + ey * ey <= (s.aperture * 1.001) ** 2).squeeze(1)
        host = lambda t: t.index_select(0, sel).cpu().numpy()
        t1 = time.time()
        ref = oracle.pstm_f64(oracle.Grid(*s.grid), host(s.mid_x), host(s.mid_y), host(s.off_x),
                              host(s.off_y), host(f32_traces), s.vel.cpu().numpy(), dt,
                              vel3d=True, aperture=s.aperture, interp=oracle.LINEAR, bins=bins)
        log(f"FP64 oracle on {nb}x{nb} bins x {grid.nt} ({len(sel)} traces) in {time.time() - t1:.1f} s")
        plan = psktm.Plan(grid, survey, aperture=s.aperture)
        for form in ("f32", "f16", "bf16"):
            if form == "f32":
                img = psktm.migrate(plan, f32_traces, s.vel)
            else:
                dtype = torch.float16 if form == "f16" else torch.bfloat16
                q, sc = psktm.quantize(f32_traces, dtype)
                img = psktm.migrate(plan, q, s.vel, scale=sc)
            sub = img[bins[0]:bins[1], bins[2]:bins[3]].double().cpu().numpy()
            acc[form] = dict(rel_l2=float(np.linalg.norm(sub - ref) / np.linalg.norm(ref)),
                             max_abs=float(np.max(np.abs(sub - ref))),
                             max_abs_ref=float(np.max(np.abs(ref))))
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
               "cores": cores, "kind": "port",
               "sample": f"first {k1} traces of the workload, reference TRUNC binary32 loop "
                         f"with the same 1500 m aperture ({pairs * grid.nt:.3e} updates, "
                         f"{el:.1f} s)"}
        log(f"CPU baseline {cpu['value']:.3f} G updates/s on {cores} threads ({el:.1f} s)")

    head = results.get("f16") or next(iter(results.values()))
    f32 = results.get("f32")
    kern_s = head["kernel_ms"] / 1e3
    valu_ops = updates * VALU_OPS_PER_UPDATE / kern_s
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
                        "tau, aperture 1500 m, v(x,y,t0); step = plan + rho filter + FP16 "
                        "quantise + migration (LINEAR, FAST)",
            "traces": args.traces, "samples": ns, "image": list(grid.shape),
            "aperture_m": s.aperture, "trace_storage": "f16 + per-trace f32 scale",
            "updates_per_step": updates,
        },
        "forms": {k: {"G_updates_per_s": v["rate"] / 1e9, "ms_per_step": v["ms_per_step"],
                      "migrate_kernel_ms": v["kernel_ms"]} for k, v in results.items()},
        "fp16_over_fp32": (head["rate"] / f32["rate"]) if f32 else None,
        "accuracy_vs_fp64": acc or None,
        "roofline": {
            "bound": "valu",
            "achieved": valu_ops / 1e12,
            "peak": VALU_PEAK_LANE_OPS / 1e12,
            "unit": "T VALU lane-op/s",
            "frac": valu_ops / VALU_PEAK_LANE_OPS,
            "traffic": None,
            "kernel": "migrate_fwd_kernel",
            "per_update": f"{VALU_OPS_PER_UPDATE} VALU instructions (3 v_sqrt_f32)",
            "fp32_published": {"achieved_tflops": updates * FLOPS_PER_UPDATE / kern_s / 1e12,
                               "peak_tflops": FP32_PUBLISHED_TFLOPS,
                               "frac": updates * FLOPS_PER_UPDATE / kern_s / 1e12 / FP32_PUBLISHED_TFLOPS},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
