"""Measurement of the Voronoi tessellation (kiwi_hip_voronoi_tessellate); not collected by pytest.

Workloads: `map` (default) -- the Voronoi map benchmark's sites, 2e5 particles of the log-normal density
field of kiwi_amd/workloads.py (256^3 box, sigma_ln_rho 1.5) plus the coverage-grid buffer
(src/point_set.cpp:167-272; empty for this field); `clustered` -- 1e5 sites in [0,4]x[0,3]x[0,2], 70% in
four Gaussian clusters (sigma 0.15) over a uniform background, the clustered case of
tests/test_gpu_tessellate.py.  Each is tessellated in its box, then one Lloyd step (every site moved to its cell's centroid,
kiwi_hip_voronoi_centroids; src/point_set.cpp:275-306) and the final tessellation: what the reference's
`map` stage runs with voro++ before its particle loop (src/tessellation.cpp:23-61,
src/mapping.cpp:222-269).  Sites resident on the device.  The CPU baseline is the reference's vendored
voro++ (oracle/_ref/libkiwi_voro.so, one thread, as the reference runs it) on the same sites; its
geometry is compared with the GPU's (cells, order, faces and neighbours but for faces below 1e-12 of the
extent squared).

    python tests/perf_tessellate.py [--workload map|clustered] [--steps 3] [--no-cpu]

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
sys.path.insert(0, os.path.join(ROOT, "tests"))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("map", "clustered"), default="map")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the voro++ baseline and comparison")
    args = ap.parse_args()
    import torch
    torch.cuda.init()
    from kiwi_amd import hip

    if args.workload == "map":
        from kiwi_amd import workloads
        from perf_voronoi_map import coverage_buffer
        parts, dom, _ = workloads.lognormal_field(n_particles=200_000, n=256, seed=2)
        pos = np.stack([parts["x"], parts["y"], parts["z"]], 1).astype(np.float64)
        buf = coverage_buffer(pos, np.asarray(parts["h"], dtype=np.float64), dom).reshape(-1, 3)
        sites = np.ascontiguousarray(np.concatenate([pos, buf]))
        dom = [float(v) for v in dom]
        label = (f"Voronoi tessellation of the Voronoi map's sites: {len(sites)} (2e5 log-normal particles, 256^3 "
                 f"field, sigma_ln_rho 1.5, + {len(buf)} coverage-buffer sites)")
    else:
        rng = np.random.default_rng(9)
        lo, hi = np.zeros(3), np.array([4.0, 3.0, 2.0])
        centres = rng.uniform(lo + 0.5, hi - 0.5, size=(4, 3))
        sites = np.vstack([centres[rng.integers(4, size=70000)] + rng.normal(scale=0.15, size=(70000, 3)),
                           rng.uniform(lo, hi, size=(30000, 3))])
        sites = np.ascontiguousarray(np.clip(sites, lo, hi))
        dom = [0.0, 4.0, 0.0, 3.0, 0.0, 2.0]
        label = ("Voronoi tessellation of 1e5 clustered sites (four Gaussian clusters, sigma 0.15, 70%; uniform "
                 "background) in [0,4]x[0,3]x[0,2]")
    log(f"{len(sites)} sites")
    dev = torch.device("cuda:0")
    dsites = torch.from_numpy(sites).to(dev)
    ctx = hip.Context(0)

    def lloyd():
        """tessellate, move every site to its centroid, tessellate again"""
        c0, i0 = ctx.voronoi_tessellate(dsites, dom)
        s0 = torch.from_numpy(sites[c0["cell_site"].cpu().numpy()]).to(dev)
        cen, _ = ctx.voronoi_centroids(c0, s0, out=(torch.empty_like(s0), torch.empty(len(s0), dtype=torch.float64,
                                                                                          device=dev)))
        c1, i1 = ctx.voronoi_tessellate(cen, dom)
        return c0, i0, cen, c1, i1

    lloyd()
    torch.cuda.synchronize()
    one, step = [], []
    for _ in range(args.steps):
        a = time.perf_counter()
        c0, i0 = ctx.voronoi_tessellate(dsites, dom)
        torch.cuda.synchronize()
        one.append(time.perf_counter() - a)
        a = time.perf_counter()
        c0, i0, cen, c1, i1 = lloyd()
        torch.cuda.synchronize()
        step.append(time.perf_counter() - a)
    ms1, ms_step = 1e3 * float(np.median(one)), 1e3 * float(np.median(step))
    # home cells of the sites' own positions in the final tessellation (the map's particles, for `map`)
    ctx.voronoi_home_cells(dsites)
    torch.cuda.synchronize()
    th = []
    for _ in range(args.steps):
        a = time.perf_counter()
        home, unres = ctx.voronoi_home_cells(dsites)
        torch.cuda.synchronize()
        th.append(time.perf_counter() - a)
    ms_home = 1e3 * float(np.median(th))
    nc, nf = len(c0["cell_site"]), len(c0["face_nbr"])
    log(f"GPU: tessellation {ms1:.1f} ms (device: order/buckets {i0['device_ms_prep']:.2f}, count "
        f"{i0['device_ms_count']:.2f}, fill {i0['device_ms_fill']:.2f}); {nc} cells, {nf} faces, "
        f"{len(c0['vert_xyz'])} face vertices; Lloyd step + final tessellation {ms_step:.1f} ms; home cells of "
        f"{len(sites)} points {ms_home:.2f} ms")
    res = {"workload": label + ", then one Lloyd step and the final tessellation",
           "cells": nc, "faces": nf, "face_vertices": int(len(c0["vert_xyz"])),
           "gpu_ms_tessellation": ms1, "gpu_ms_lloyd_step_and_final": ms_step, "gpu_ms_home_cells": ms_home,
           "device_ms": {"prep": i0["device_ms_prep"], "count": i0["device_ms_count"], "fill": i0["device_ms_fill"]},
           "cells_per_s": nc / (ms1 * 1e-3)}
    if not args.no_cpu:
        from oracle import oracle as O
        from geometry_check import assert_same_cells
        a = time.perf_counter()
        r = O.tessellate(np.arange(len(sites)), sites, np.asarray(dom, dtype=np.float64), np.zeros((0, 3)))
        cpu = time.perf_counter() - a
        g = {k: v.cpu().numpy() for k, v in c0.items() if k != "domain"}
        ext = float(np.max(np.abs(dom)))
        assert_same_cells(g, r, 1e-12 * ext ** 2)
        s0 = torch.from_numpy(np.ascontiguousarray(sites[g["cell_site"]])).to(dev)
        _, vol = ctx.voronoi_centroids(c0, s0)  # device tensors, as the cells
        dv = float(np.max(np.abs(vol.cpu().numpy() - r["volume"])))
        log(f"voro++ (one thread): {cpu:.2f} s; same cells and faces; max volume difference {dv:.2e} "
            f"(extent^3 {ext ** 3:.2e})")
        res.update({"cpu_baseline": {"value": cpu * 1e3, "unit": "ms per tessellation", "cores": 1,
                                     "kind": "reference", "sample": "the same sites, voro++ container as the "
                                                                    "reference builds it"},
                    "same_geometry_as_voro": True, "max_volume_diff": dv, "speedup_vs_voro": cpu * 1e3 / ms1})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
