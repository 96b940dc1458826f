"""Measurement of kiwi_hip_sph_to_grid on the small uniform case; not collected by pytest.

Workload: the 2,000-particle uniform cube of tests/golden/c1_snapshot.npz onto 32^3 (M4, its stored
domain), particles and outputs resident on the device, FP64 and mixed precision, eager calls and graph
mode (kiwi_hip_set_graphs).  The problem is latency-bound (846,189 sphere pairs), so the figure is the
wall time of a whole call including its launches and the host synchronisation that ends it.  Prints
one JSON line with the median call time and the sphere and bounding-box pairs per second.

    python tests/perf_uniform_2k_32.py [--calls 200]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

COLS = ("x", "y", "z", "vx", "vy", "vz", "m", "h")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calls", type=int, default=200)
    args = ap.parse_args()
    import torch
    torch.cuda.init()
    from kiwi_amd import hip

    c1 = np.load(os.path.join(ROOT, "tests", "golden", "c1_snapshot.npz"))
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(c1[k], dtype=np.float64)).to(dev) for k in COLS}
    dom = np.asarray(c1["domain"], dtype=np.float64)
    res = (32, 32, 32)
    nc = int(np.prod(res))
    ctx = hip.Context(0)
    out = {"workload": "sph-to-grid, 2000 uniform particles -> 32^3, M4 (tests/golden/c1_snapshot.npz)",
           "calls": args.calls}
    for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        for graphs in (False, True):
            ctx.set_graphs(graphs)
            o = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
            st = None
            for _ in range(5):
                st = ctx.sph_to_grid(t, dom, res, precision=prec, out=o)[4]
            torch.cuda.synchronize()
            ts = []
            for _ in range(args.calls):
                a = time.perf_counter()
                ctx.sph_to_grid(t, dom, res, precision=prec, out=o)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - a)
            ms = 1e3 * float(np.median(ts))
            key = f"{name}_{'graph' if graphs else 'eager'}"
            out[key] = {"ms_per_call": ms, "pairs_sphere": st.pairs_sphere, "pairs_bbox": st.pairs_bbox,
                        "pairs_sphere_per_s": st.pairs_sphere / (ms * 1e-3),
                        "pairs_bbox_per_s": st.pairs_bbox / (ms * 1e-3)}
            log(f"{key}: {ms:.3f} ms per call, {st.pairs_sphere / (ms * 1e-3):.3e} sphere pairs/s")
    ctx.set_graphs(False)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
