"""Measurement of the generic-kernel path of kiwi_hip_sph_to_grid on the benchmark field; not collected
by pytest.

Workload: the 1M-particle log-normal field of kiwi_amd.workloads.lognormal_field() onto 256^3 (the
sph-to-grid benchmark's particles and grid) with a generic kernel (default Wendland C2), FP64 and mixed
precision, particles and outputs resident on the device.  Prints one JSON line: per precision the
call time, the library's stage times (HIP events), sphere pairs per second, and the mixed result's
total-mass and per-field errors against FP64.

    python tests/perf_generic.py [--kernel WC2] [--steps 3]
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
    ap.add_argument("--kernel", default="WC2")
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args()
    import torch
    torch.cuda.init()
    from kiwi_amd import hip

    parts, dom, res = workloads.lognormal_field()
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(parts[k])).to(dev) for k in workloads.COLS}
    nc = int(np.prod(res))
    ctx = hip.Context(0)
    k = hip.kernel_spec(args.kernel)
    out = {}
    res_json = {"workload": f"sph-to-grid, {len(parts['x'])} log-normal particles -> 256^3, {args.kernel} (generic path)"}
    for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        o = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
        ctx.sph_to_grid(t, dom, res, kernel=k, precision=prec, out=o)
        torch.cuda.synchronize()
        ts, st = [], None
        for _ in range(args.steps):
            a = time.perf_counter()
            st = ctx.sph_to_grid(t, dom, res, kernel=k, precision=prec, out=o)[4]
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        ms = 1e3 * float(np.median(ts))
        out[name] = o
        res_json[name] = {"ms_per_call": ms, "cells_ms": st.device_ms_gather, "lines_ms": st.device_ms_lines,
                          "accum_ms": st.device_ms_accum, "prep_ms": st.device_ms_prep,
                          "pairs_sphere": st.pairs_sphere, "cell_integrals_per_s": st.pairs_sphere / (ms * 1e-3)}
        log(f"{name}: {ms:.1f} ms per call, cells {st.device_ms_gather:.1f} ms, lines {st.device_ms_lines:.1f} ms")
    f64 = [x.cpu().numpy() for x in out["fp64"]]
    f32 = [x.cpu().numpy() for x in out["mixed"]]
    res_json["mixed_over_fp64"] = res_json["fp64"]["ms_per_call"] / res_json["mixed"]["ms_per_call"]
    res_json["mixed_total_mass_rel_err"] = float(abs(f32[0].sum() - f64[0].sum()) / f64[0].sum())
    res_json["mixed_field_err_over_max"] = {
        n: float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for n, a, b in zip(("mass", "px", "py", "pz"), f32, f64)}
    print(json.dumps(res_json), flush=True)


if __name__ == "__main__":
    main()
