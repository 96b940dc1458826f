"""Measurement of the batched sph-to-grid (kiwi_hip_sph_to_grid_batch) on its benchmark workload; not
collected by pytest.

Workload (kiwi_amd.workloads.cutout_batch): 64 independent cutouts, particle counts log-uniform in
[1e3, 2e5], the largest grid axis log-uniform in [32, 192] cells and the others 1 to 4 times shorter,
log-normal density (sigma_ln_rho 1.5), kernels alternating between the cubic spline and Wendland C2
(generic path), FP32 bulk (mixed precision) for both.  Particles and outputs resident on the device.

The library's autotuner first picks the cell kernel's launch shape on the batch (untimed;
--no-autotune keeps the default).  Then times the batch (one call) against the same problems as 64
single calls, checks that both give bit-identical sums, times the batch again in graph mode
(kiwi_hip_set_graphs: one captured HIP graph replayed per kernel path after the sizing pass) and checks
it against the eager batch bit for bit, and prints one JSON line.

    python tests/perf_batch.py [--problems 64] [--steps 3] [--fp64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import workloads  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--fp64", action="store_true", help="every problem in FP64 (default: mixed precision)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    args = ap.parse_args()
    import torch
    torch.cuda.init()
    from kiwi_amd import hip

    t0 = time.perf_counter()
    batch = workloads.cutout_batch(n_problems=args.problems)
    dev = torch.device("cuda:0")
    probs = []
    for parts, dom, res, kname in batch:
        nc = int(np.prod(res))
        probs.append({"parts": {k: torch.from_numpy(np.ascontiguousarray(parts[k])).to(dev) for k in workloads.COLS},
                      "domain": dom, "res": res, "kernel": hip.kernel_spec(kname),
                      "precision": hip.FP64 if args.fp64 else hip.MIXED,
                      "out": tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))})
    n_parts = sum(len(b[0]["x"]) for b in batch)
    n_cells = sum(int(np.prod(b[2])) for b in batch)
    log(f"{len(probs)} problems, {n_parts} particles, {n_cells} cells ({time.perf_counter() - t0:.1f} s)")
    ctx = hip.Context(0)
    shape = ctx.get_tuning()
    if not args.no_autotune:  # untimed; the launch shape never changes results
        shape, tms = ctx.autotune(probs)
        log(f"autotuned launch shape: item_grab {shape[0]}, cursors {shape[1]} ({tms:.1f} ms per batch)")

    def run_batch():
        return ctx.sph_to_grid_batch(probs)

    def run_single():
        return [ctx.sph_to_grid(q["parts"], q["domain"], q["res"], kernel=q["kernel"], precision=q["precision"],
                                out=q["out"]) for q in probs]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            a = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        return 1e3 * float(np.median(ts)), r

    ms_batch, rb = timed(run_batch)
    batch_out = [tuple(o.clone() for o in q["out"]) for q in probs]
    ms_single, rs = timed(run_single)
    identical = all(bool(torch.equal(a, b)) for q, bo in zip(probs, batch_out) for a, b in zip(q["out"], bo))
    ctx.set_graphs(True)
    ms_graph, _ = timed(run_batch)
    ctx.set_graphs(False)
    graph_identical = all(bool(torch.equal(a, b)) for q, bo in zip(probs, batch_out) for a, b in zip(q["out"], bo))
    pairs = sum(r[4].pairs_sphere for r in rb)
    pairs_bbox = sum(r[4].pairs_bbox for r in rb)
    prec = "FP64" if args.fp64 else "mixed"
    res = {"workload": f"{len(probs)} cutouts, {n_parts} particles, {n_cells} cells, M4 / WC2 {prec}",
           "ms_batch": ms_batch, "ms_single_calls": ms_single, "speedup": ms_single / ms_batch,
           "ms_batch_graph": ms_graph, "graph_bit_identical": graph_identical,
           "pairs_sphere": pairs, "pairs_bbox": pairs_bbox,
           "cell_integrals_per_s_batch": pairs / (ms_batch * 1e-3),
           "cell_integrals_per_s_single": pairs / (ms_single * 1e-3), "bit_identical": identical,
           "launch_shape": {"item_grab": shape[0], "cursors": shape[1], "autotuned": not args.no_autotune}}
    log(f"batch {ms_batch:.2f} ms, single calls {ms_single:.2f} ms (x{ms_single / ms_batch:.2f}), graph mode "
        f"{ms_graph:.2f} ms; {pairs} sphere pairs; identical: {identical}, graph identical: {graph_identical}")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
