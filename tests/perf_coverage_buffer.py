"""Measurement of the coverage-grid buffer (kiwi_hip_coverage_buffer) on the sph-to-grid benchmark's
particles; not collected by pytest.  Workload: 1e6 particles of the log-normal density field of
kiwi_amd/workloads.py (256^3 box, h 0.5-25), inputs resident on the device, at the reference's default
parameters (spacing 0.05 of the box, radius factor 2: a 20^3 lattice) and at a fine lattice (spacing
0.01, factor 1: 100^3 candidates, some uncovered).  CPU baseline: the reference's own
add_coverage_grid_buffer (oracle/_ref/libkiwi_pts.so, one thread) on the default-parameter case, whose
output must equal the GPU's.

    python tests/perf_coverage_buffer.py [--particles 1000000] [--steps 10] [--warmup 2] [--no-cpu-baseline]

Prints progress to stderr and one JSON line to stdout.
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
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    import torch
    torch.cuda.init()
    from kiwi_amd import hip
    from oracle import oracle as O

    parts, dom, _ = workloads.lognormal_field(n_particles=args.particles, n=256, seed=2)
    cols = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in ("x", "y", "z", "h")}
    dev = torch.device("cuda:0")
    dcols = {k: torch.from_numpy(v).to(dev) for k, v in cols.items()}
    ctx = hip.Context(0)
    rec = {"metric": "coverage-buffer particles/s", "unit": "particles/s",
           "config": {"workload": f"coverage-grid buffer of {args.particles} log-normal particles (256^3 box, "
                                  f"h 0.5-25)", "particles": args.particles}, "cases": {}}
    for name, fr, fa in (("default", 0.05, 2.0), ("fine", 0.01, 1.0)):
        nx, ny, nz, _ = hip.cartesian_lattice(dom, fr)
        cap = nx * ny * nz
        out = tuple(torch.zeros(cap, dtype=torch.float64, device=dev) for _ in range(3)) + (
            torch.zeros(cap, dtype=torch.int32, device=dev),)
        for _ in range(args.warmup):
            ctx.coverage_buffer(dcols, dom, fr, fa, out=out)
        torch.cuda.synchronize()
        times, dms = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            st = ctx.coverage_buffer(dcols, dom, fr, fa, out=out)[2]
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
            dms.append(st.device_ms)
        ms = 1e3 * float(np.median(times))
        log(f"{name}: spacing {fr}, factor {fa}: {ms:.3f} ms/call (device {np.median(dms):.3f} ms), "
            f"{st.kept} of {st.candidates} kept, {st.particle_columns} particle columns")
        case = {"spacing_fraction": fr, "factor": fa, "ms_per_call": ms, "device_ms": float(np.median(dms)),
                "candidates": st.candidates, "kept": st.kept, "particle_columns": st.particle_columns,
                "value": args.particles / (ms * 1e-3)}
        if name == "default" and not args.no_cpu_baseline and os.path.exists(O.PTS_SO):
            t = time.perf_counter()
            rid, rxyz = O.ref_coverage_buffer(cols["x"], cols["y"], cols["z"], cols["h"], dom, fr, fa)
            tc = time.perf_counter() - t
            gid = out[3].cpu().numpy()[: st.kept]
            gxyz = torch.stack(out[:3], 1).cpu().numpy()[: st.kept]
            same = bool(np.array_equal(gid, rid) and np.array_equal(gxyz, rxyz))
            log(f"  reference add_coverage_grid_buffer: {tc:.2f} s, same output: {same}")
            assert same
            rec["cpu_baseline"] = {"value": args.particles / tc, "unit": "particles/s", "cores": 1, "kind": "reference",
                                   "sample": f"the whole default-parameter case, the reference's own "
                                             f"add_coverage_grid_buffer, one thread, {tc:.2f} s"}
        rec["cases"][name] = case
    rec["value"] = rec["cases"]["default"]["value"]
    rec["ms_per_call"] = rec["cases"]["default"]["ms_per_call"]
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
