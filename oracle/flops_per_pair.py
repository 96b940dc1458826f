"""Pin the reference algorithm's operations per (particle, cell) pair on the metric workload.

Runs the oracle's operation-counting instantiation of the cubic cell loop (oracle_count_m4) over a
deterministic subsample of the 1M-particle log-normal workload onto its 256^3 grid and writes the
mean per bounding-box pair to kiwi_amd/flops_per_pair.json.  The throughput figure a kernel is
credited with is this mean times the bounding-box pairs its launch covers.

    python oracle/flops_per_pair.py [--every 97]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[0] = ROOT  # run as a script: import the package, not oracle/oracle.py


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--every", type=int, default=97, help="take every k-th particle")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    from kiwi_amd import workloads
    from oracle import oracle as O
    parts, dom, res = workloads.lognormal_field()
    cols = [np.ascontiguousarray(parts[c][:: args.every]) for c in ("x", "y", "z", "h")]
    L = O.lib()
    L.oracle_count_m4.restype = ctypes.c_int
    pd = ctypes.POINTER(ctypes.c_double)
    L.oracle_count_m4.argtypes = [ctypes.c_longlong, pd, pd, pd, pd, pd, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
    out = (ctypes.c_longlong * 6)()
    d = np.ascontiguousarray(dom, dtype=np.float64)
    t0 = time.perf_counter()
    rc = L.oracle_count_m4(len(cols[0]), *[c.ctypes.data_as(pd) for c in cols], d.ctypes.data_as(pd), *res,
                           args.threads, out)
    assert rc == 0
    pairs, add, mul, div, sq, tr = list(out)
    per = {k: v / pairs for k, v in (("add", add), ("mul", mul), ("div", div), ("sqrt", sq), ("transcendental", tr))}
    total = sum(per.values())
    result = {
        "workload": f"every {args.every}th particle of the 1M log-normal workload onto 256^3 "
                    f"({len(cols[0])} particles, {pairs} bounding-box pairs)",
        "convention": "each +, -, *, /, sqrt and transcendental call = 1 op (sincos = 2); "
                      "comparisons, fabs, selects = 0; literal constants = 0",
        "per_bbox_pair": per, "flops_per_bbox_pair": total, "seconds": time.perf_counter() - t0,
    }
    path = os.path.join(ROOT, "kiwi_amd", "flops_per_pair.json")
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    print(json.dumps(result, indent=1))


if __name__ == "__main__":
    main()
