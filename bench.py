#!/usr/bin/env python3
"""Benchmark: mixed-precision vs FP64 sph-to-grid of 1M clustered particles onto 256^3 on one MI355X.

Workload (fixed seed, synthetic): 1,000,000 particles sampled from a log-normal density field (Gaussian
random field on 256^3, P(k) ~ k^-3, sigma_ln_rho = 1.5), equal masses, h = 1.2 (1/(nbar rho))^(1/3)
clipped to [0.5, 50] cells, velocities from the field's gradient (rms 5 km/s per axis) plus N(0, 2 km/s),
M4 kernel, domain [0,256]^3 with unit cells, grid 256^3.  Inputs and outputs are resident in HBM for the
timed region.  A step is one full kiwi_hip_sph_to_grid call (binning, sort, gather, mean velocity).
Before each precision's warmup the library's autotuner (kiwi_hip_autotune) picks the cell kernel's
launch shape on this workload (untimed; the shape never changes results; --no-autotune keeps the
default, --tuning G,K sets one), and the chosen shape is reported in config.launch_shape.

The metric is cell-integrals/s: (particle, cell) pairs whose cell intersects the particle's support
sphere, divided by the step time.  `value` is the mixed-precision rate; the FP64 rate, the speed ratio
and the mixed-vs-FP64 relative total-mass error are reported in the same line.

Mixed precision computes particles whose support radius spans at least the FP64-fallback threshold in
cells (kiwi_hip_set_fp64_fallback; default 16) in FP64, in a second cell kernel beside the FP32 one; the
line also reports the same mixed step without the fallback.

roofline: the FP32 cell-integral kernel (k_particles<float>: lattice values and per-pair cell sums),
VALU-bound.  `achieved` is the operations the kernel executes per launch -- the per-unit counts of
kiwi_amd/flops_per_unit.json (kiwi_amd/csrc/flop_count.cpp instantiates the kernel's lattice math over
an operation-counting scalar; +, -, x, /, sqrt and each transcendental count 1, an FMA 2) times the
launch's work counters (lattice values by region and cells, less the FP64 fallback's) -- divided by the
kernel's duration, measured with HIP events on the stream the kernel runs on; `frac` = achieved / the
FP32 vector peak (157.3 TFLOP/s).  `ref_equiv` is a different quantity, kept apart: the reference
algorithm's operation count for the same work (the pinned mean per bounding-box pair of
kiwi_amd/flops_per_pair.json, from oracle/flops_per_pair.py) over the same duration; the lattice
restatement does ~1/80 of those operations, so it exceeds the peak and is no roofline fraction.
`traffic` is the HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE passes at the launch shape
named beside it (profiles/k_particles_m4_1M_256_traffic.json; a constant read from that file, not
measured in this run).
cpu_baseline: the CPU oracle (a bit-exact restatement of the reference's loop) on a bounded sample of
the same workload (every 128th particle, same grid), on this host's cores, with the oracle/reference
time ratio measured on identical inputs on the build machine (oracle/cpu_calibration.json).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}  # MI355X dense vector peaks (FP32, FP64)
CPU_SAMPLE_EVERY = 128


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cell_kernel_flops(st, fpu, part="main"):
    """Algorithmic FLOPs of a cell-kernel launch: lattice values by region and cell stencils (the wall-plane
    constants come from the line kernel's axis tables).  part "main": the kernel on the main stream (in
    mixed precision the FP32 one, i.e. the totals less the FP64 fallback's); "fallback": the fallback's."""
    fb_pts, fb_pairs = list(st.fp64_fallback_points), st.fp64_fallback_pairs
    if part == "fallback":
        pts, pairs = fb_pts, fb_pairs
    else:
        pts = [st.lattice_points[r] - fb_pts[r] for r in range(3)]
        pairs = st.pairs_sphere - fb_pairs
    return sum(pts[r] * fpu["point"][r] for r in range(3)) + pairs * fpu["cell"]


def calibration():
    """The oracle/reference time ratio on the same sample at all cores, from oracle/cpu_calibration.json."""
    path = os.path.join(ROOT, "oracle", "cpu_calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        cal = json.load(f)
    rows = [r for r in cal["rows"] if r["case"] == "lognormal/16"] or cal["rows"]
    r = rows[0]
    return {"oracle_over_reference": r["oracle_over_ref"], "case": r["case"], "threads": r["threads"],
            "reference_pairs_sphere_per_s": r["ref_pairs_sphere_per_s"], "host": cal["host"]}


def line_kernel_flops(st, fpu):
    return sum(st.line_evals[r] * fpu["line"][r] for r in range(3))


def run_precision(ctx, hip, dev, dom, res, out, prec, steps, warmup, autotune):
    """Times `steps` calls; returns (host ms per step, mean stats of the device times, last stats, shape)."""
    import torch
    shape = ctx.get_tuning()
    if autotune:  # untimed: the launch shape (never the results) is chosen on this workload
        shape, ms = ctx.autotune([{"parts": dev, "domain": dom, "res": res, "precision": prec, "out": out}])
        log(f"  autotuned launch shape: item_grab {shape[0]}, cursors {shape[1]} ({ms:.2f} ms per call)")
    for _ in range(warmup):
        ctx.sph_to_grid(dev, dom, res, precision=prec, out=out)
    torch.cuda.synchronize()
    gms, kms, fms, lms, pms, ams, stats = [], [], [], [], [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        *_, stats = ctx.sph_to_grid(dev, dom, res, precision=prec, out=out)
        gms.append(stats.device_ms_gather)
        kms.append(stats.device_ms_gather_kernel)
        fms.append(stats.device_ms_gather_fallback)
        lms.append(stats.device_ms_lines)
        pms.append(stats.device_ms_prep)
        ams.append(stats.device_ms_accum)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"  device ms per step: prep {np.mean(pms):.2f}, lines {np.mean(lms):.2f}, cells {np.mean(gms):.2f} "
        f"(main kernel {np.mean(kms):.2f}, FP64 fallback kernel {np.mean(fms):.2f}), accum {np.mean(ams):.2f}")
    times = {"cells": float(np.mean(gms)), "kernel": float(np.mean(kms)), "fallback": float(np.mean(fms))}
    return (t1 - t0) * 1e3 / steps, times, stats, shape


def cpu_baseline(parts, dom, res, hip, ctx):
    from oracle import oracle as O
    sample = {k: np.ascontiguousarray(v[::CPU_SAMPLE_EVERY]) for k, v in parts.items()}
    *_, gst = ctx.sph_to_grid(sample, dom, res, precision=hip.FP64)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    O.sph_to_grid(sample, dom, *res, threads=threads)
    dt = time.perf_counter() - t0
    out = {"value": gst.pairs_sphere / dt, "unit": "cell-integrals/s", "cores": threads, "kind": "port",
           "sample": f"every {CPU_SAMPLE_EVERY}th particle ({len(sample['x'])} particles) of the same "
                     f"workload onto the same {res[0]}^3 grid, FP64 oracle (bit-exact restatement of the "
                     f"reference loop), {dt:.1f} s, {gst.pairs_sphere} pairs"}
    cal = calibration()
    if cal:
        out["calibration"] = cal
        out["reference_estimate"] = out["value"] * cal["oracle_over_reference"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--tuning", default=None, help="item_grab,cursors: use this launch shape (no autotuning)")
    ap.add_argument("--fp64-fallback", type=float, default=None,
                    help="mixed precision's FP64 fallback threshold, support radius in cells (<= 0: off; "
                         "default: the library's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the mapping shards by grid slab across GPUs; that path is "
              f"not built here); got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import torch
    from kiwi_amd import hip, workloads

    t = time.perf_counter()
    parts, dom, res = workloads.lognormal_field(n_particles=args.particles, n=args.grid, seed=2)
    log(f"workload generated in {time.perf_counter() - t:.1f} s: {args.particles} particles -> {res}")
    with open(os.path.join(ROOT, "kiwi_amd", "flops_per_unit.json")) as f:
        fpu = json.load(f)
    with open(os.path.join(ROOT, "kiwi_amd", "flops_per_pair.json")) as f:
        f_pair = json.load(f)["flops_per_bbox_pair"]

    dev = {k: torch.from_numpy(v).to("cuda:0") for k, v in parts.items()}
    nc = res[0] * res[1] * res[2]
    out64 = [torch.zeros(nc, dtype=torch.float64, device="cuda:0") for _ in range(4)]
    out32 = [torch.zeros(nc, dtype=torch.float64, device="cuda:0") for _ in range(4)]
    ctx = hip.Context(0)
    if args.fp64_fallback is not None:
        ctx.set_fp64_fallback(args.fp64_fallback)
    fb_cells = ctx.get_fp64_fallback()
    if args.tuning:
        ctx.set_tuning(*(int(v) for v in args.tuning.split(",")))
        args.no_autotune = True

    ms64, g64, st64, shape64 = run_precision(ctx, hip, dev, dom, res, out64, hip.FP64, args.steps, args.warmup,
                                             not args.no_autotune)
    log(f"fp64: {ms64:.2f} ms/step, cell kernel {g64['kernel']:.2f} ms, pairs {st64.pairs_sphere}")
    ms32, g32, st32, shape32 = run_precision(ctx, hip, dev, dom, res, out32, hip.MIXED, args.steps, args.warmup,
                                             not args.no_autotune)
    log(f"mixed: {ms32:.2f} ms/step, cell kernels {g32['cells']:.2f} ms")

    m64 = out64[0].cpu().numpy()
    m32 = out32[0].cpu().numpy()
    tot64 = float(np.sum(m64))
    mass_err = abs(float(np.sum(m32)) - tot64) / tot64
    l1_err = float(np.sum(np.abs(m32 - m64))) / float(np.sum(np.abs(m64)))  # per-cell error, mass weighted
    field_err = {name: float(torch.max(torch.abs(a - b)).item() / max(torch.max(torch.abs(b)).item(), 1e-300))
                 for name, a, b in zip(("mass", "px", "py", "pz"), out32, out64)}
    log(f"mixed vs fp64: total-mass rel err {mass_err:.3e}, mass L1 rel err {l1_err:.3e}, "
        f"per-field max err/max {field_err}")

    # the same mixed step without the FP64 fallback, for comparison (untimed for `value`)
    no_fb = None
    if fb_cells > 0:
        ctx.set_fp64_fallback(0.0)
        ms_nf, g_nf, st_nf, _ = run_precision(ctx, hip, dev, dom, res, out32, hip.MIXED, max(3, args.steps // 2), 1,
                                              False)
        m_nf = out32[0].cpu().numpy()
        no_fb = {"ms_per_step": ms_nf, "cells_ms": g_nf["cells"],
                 "total_mass_rel_err": abs(float(np.sum(m_nf)) - tot64) / tot64,
                 "mass_l1_rel_err": float(np.sum(np.abs(m_nf - m64))) / float(np.sum(np.abs(m64)))}
        ctx.set_fp64_fallback(fb_cells)
        log(f"mixed without the FP64 fallback: {ms_nf:.2f} ms/step, mass L1 rel err {no_fb['mass_l1_rel_err']:.3e}")

    pairs = st32.pairs_sphere
    value = pairs / (ms32 * 1e-3)
    fl32 = cell_kernel_flops(st32, fpu)  # the FP32 cell kernel's share of the work
    fl_fb = cell_kernel_flops(st32, fpu, "fallback")
    fl64 = cell_kernel_flops(st64, fpu)
    alg = f_pair * st32.pairs_bbox  # reference-algorithm operations for the step's cell work
    ref_equiv = alg / (g32["cells"] * 1e-3) / 1e12
    exe = fl32 / (g32["kernel"] * 1e-3) / 1e12
    traffic, traffic_shape = None, None
    tpath = os.path.join(ROOT, "profiles", "k_particles_m4_1M_256_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic, traffic_shape = tj.get("mixed_bytes_per_launch"), tj.get("launch_shape")
    lines_kernel_ms = st32.device_ms_lines_kernel

    result = {
        "metric": "cell-integrals/s, 1M part->256^3: mixed vs FP64 on GPU at <=1e-6 mass err",
        "value": value, "unit": "cell-integrals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (FP64 accumulation)", "data": "synthetic (seeded log-normal field)",
        "config": {"workload": f"sph-to-grid, {args.particles} log-normal clustered particles -> "
                               f"{res[0]}^3, M4, h 0.5-50 cells", "particles": args.particles,
                   "grid": list(res), "kernel": "M4", "parallelism": "single GPU",
                   "launch_shape": {"fp64": {"item_grab": shape64[0], "cursors": shape64[1]},
                                    "mixed": {"item_grab": shape32[0], "cursors": shape32[1]},
                                    "autotuned": not args.no_autotune}},
        "fp64": {"value": pairs / (ms64 * 1e-3), "ms_per_step": ms64, "cells_ms": g64["kernel"],
                 "roofline_frac": fl64 / (g64["kernel"] * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"],
                 "roofline_peak_tflops": PEAK_TFLOPS["fp64"],
                 "ref_equiv_tflops": f_pair * st64.pairs_bbox / (g64["kernel"] * 1e-3) / 1e12},
        "fp64_fallback": {
            "support_cells": fb_cells,
            "note": "mixed precision computes particles whose support radius spans >= support_cells cells in "
                    "FP64 (their FP32 lattice differences lose up to ~4e-2 of the particle at h = 30 cells); "
                    "that cell kernel runs on the context's third stream beside the FP32 one, from the end of the "
                    "item layout",
            "items_share": st32.fp64_fallback_blocks / max(st32.particle_blocks, 1),
            "pairs_share": st32.fp64_fallback_pairs / max(st32.pairs_sphere, 1),
            "kernel_ms": g32["fallback"], "flops_per_launch": fl_fb,
            "tflops": fl_fb / (g32["fallback"] * 1e-3) / 1e12 if g32["fallback"] > 0 else None,
            "without_fallback": no_fb},
        "mixed_over_fp64": ms64 / ms32,
        "mixed_total_mass_rel_err": mass_err,
        "mixed_field_err_over_max": field_err, "mixed_mass_l1_rel_err": l1_err,
        "pairs_sphere": pairs, "pairs_bbox": st32.pairs_bbox, "particle_blocks": st32.particle_blocks,
        "cells_ms": g32["cells"], "prep_ms": st32.device_ms_prep,
        "lines_exposed_ms": st32.device_ms_lines, "lines_kernel_ms": lines_kernel_ms,
        "accum_ms": st32.device_ms_accum,
        "lines_tflops": (line_kernel_flops(st32, fpu) / (lines_kernel_ms * 1e-3) / 1e12) if lines_kernel_ms > 0 else None,
        "work": {"lattice_points": list(st32.lattice_points), "line_evals": list(st32.line_evals),
                 "plane_evals": st32.plane_evals},
        "roofline": {"bound": "valu", "achieved": exe, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
                     "frac": exe / PEAK_TFLOPS["mixed"], "traffic": traffic, "traffic_launch_shape": traffic_shape,
                     "kernel": "k_particles<float>", "kernel_ms": g32["kernel"], "flops_per_launch": fl32},
        "ref_equiv": {"tflops": ref_equiv, "flops_per_launch": alg, "flops_per_bbox_pair": f_pair,
                      "note": "reference-algorithm operations for the same work over the kernel time; not a "
                              "roofline fraction"},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(parts, dom, res, hip, ctx)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.3e} cell-integrals/s")
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
