"""Measurement of the cartesian-map resample (kiwi_hip_cartesian_map) on the Voronoi-map benchmark
workload; not collected by pytest.  The sites are the final tessellation's cells of tests/
perf_voronoi_map.py (2e5 log-normal particles plus the coverage-grid buffer, one Lloyd step), built
through the reference's vendored voro++ (oracle/_ref/libkiwi_voro.so), each with a random mapping.dat
row; the grid is res^3 over the domain (default 256^3), inputs and outputs resident on the device.

Checks: 2,000 sampled centres against the oracle's brute-force nearest site.  CPU baseline: voro++'s own
find_voronoi_cell (the reference's cell loop, src/cartesian_map.cpp:166-172) on a sample of the centres,
one thread, timed as the difference of two shim calls that build the same container, with and without
the sample queries.

    python tests/perf_cartesian_map.py [--particles 200000] [--res 256] [--steps 5] [--warmup 2]
                                       [--sample 1000000]

Prints progress to stderr and one JSON line to stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import perf_voronoi_map as PV  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def centres(dom, res, flat):
    """Cell centres of the given flat (ix res + iy) res + iz indices."""
    iz = flat % res
    iy = (flat // res) % res
    ix = flat // (res * res)
    d = [(dom[2 * a + 1] - dom[2 * a]) / res for a in range(3)]
    return np.stack([dom[0] + (ix + 0.5) * d[0], dom[2] + (iy + 0.5) * d[1], dom[4] + (iz + 0.5) * d[2]], 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=200_000)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sample", type=int, default=1_000_000, help="centres in the CPU-baseline sample (0: none)")
    args = ap.parse_args()

    import torch
    torch.cuda.init()
    from kiwi_amd import hip
    from oracle import oracle as O

    _, _, cells, geo = PV.build(args.particles)
    dom = cells["domain"]
    sites = np.ascontiguousarray(geo["site"])
    ids = np.ascontiguousarray(geo["cell_id"], dtype=np.int32)
    n = len(ids)
    rng = np.random.default_rng(5)
    S = {"x": sites[:, 0].copy(), "y": sites[:, 1].copy(), "z": sites[:, 2].copy(), "id": ids,
         "density": rng.uniform(0.1, 2.0, n), "vx": rng.normal(size=n), "vy": rng.normal(size=n), "vz": rng.normal(size=n)}
    dev = torch.device("cuda:0")
    dS = {k: torch.from_numpy(v).to(dev) for k, v in S.items()}
    res = args.res
    nc = res ** 3
    out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4)) + (
        torch.zeros(nc, dtype=torch.int32, device=dev),)
    ctx = hip.Context(0)
    for _ in range(args.warmup):
        ctx.cartesian_map(dS, dom, res, out=out)
    torch.cuda.synchronize()
    times, prep, query = [], [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        st = ctx.cartesian_map(dS, dom, res, out=out)[5]
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        prep.append(st.device_ms_prep)
        query.append(st.device_ms_query)
    ms = 1e3 * float(np.median(times))
    qms = float(np.median(query))
    log(f"GPU: {ms:.3f} ms/call ({n} sites -> {res}^3): prep {np.median(prep):.3f} ms, query {qms:.3f} ms")

    cid = out[4].cpu().numpy()
    pick = rng.choice(nc, 2000, replace=False)
    want = ids[O.nearest_site(sites, centres(dom, res, pick), chunk=64)]
    same = bool(np.array_equal(cid[pick], want))
    dens = out[0].cpu().numpy()
    by_id = np.zeros(ids.max() + 1, np.int64)
    by_id[ids] = np.arange(n)
    rows_ok = bool(np.array_equal(dens, S["density"][by_id[cid]]))
    log(f"sampled centres match the oracle's nearest site: {same}; rows: {rows_ok}")
    assert same and rows_ok

    cpu = None
    if args.sample and os.path.exists(O.VORO_SO):
        q = centres(dom, res, rng.choice(nc, args.sample, replace=False))
        t = time.perf_counter()
        O.tessellate(ids, sites, dom, np.zeros((0, 3)))
        t0 = time.perf_counter() - t
        t = time.perf_counter()
        g = O.tessellate(ids, sites, dom, q)
        t1 = time.perf_counter() - t
        assert np.all(g["home"] >= 0)
        cpu = {"value": args.sample / max(t1 - t0, 1e-9), "unit": "cells/s", "cores": 1, "kind": "reference",
               "sample": f"{args.sample} random centres of the same grid, voro++ find_voronoi_cell (the "
                         f"reference's vendored voro++), one thread: {t1:.2f} s with the queries minus "
                         f"{t0:.2f} s for the same container without"}
        log(f"CPU voro++ find_voronoi_cell: {cpu['value']:.3e} cells/s")

    out_bytes = nc * (4 * 8 + 4)
    rec = {"metric": "cartesian-map cells/s", "value": nc / (ms * 1e-3), "unit": "cells/s",
           "ms_per_call": ms, "device_ms": {"prep": float(np.median(prep)), "query": qms},
           "config": {"workload": f"cartesian map of the Voronoi-map benchmark's {n} sites (2e5 log-normal "
                                  f"particles + coverage buffer, one Lloyd step) onto {res}^3",
                      "sites": n, "res": res},
           "query_output_gbs": out_bytes / (qms * 1e-3) / 1e9,
           "sampled_match": same, "cpu_baseline": cpu}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
