"""Measurement of the whole Voronoi map stage on the GPU, with no host geometry; not collected by pytest.

Workload: the Voronoi map benchmark (tests/perf_voronoi_map.py): 2e5 particles of the log-normal density
field of kiwi_amd/workloads.py (256^3 box, sigma_ln_rho 1.5), M4, FP64.  The stage as the reference runs
it (src/point_set.cpp:74-371, src/mapping.cpp:222-506): the coverage-grid buffer, the tessellation of
the particles and buffer sites, one Lloyd step, the final tessellation, every particle's home cell and
the map -- here kiwi_hip_coverage_buffer, kiwi_hip_voronoi_tessellate, kiwi_hip_voronoi_centroids,
kiwi_hip_voronoi_tessellate, kiwi_hip_voronoi_home_cells and kiwi_hip_voronoi_map, everything resident
on the device.  The same stage with the geometry and home cells from the reference's vendored voro++
(oracle/_ref/libkiwi_voro.so, one thread) is timed on the host for its geometry part and gives the
reference sums the GPU stage is compared with (same cell order; the cells differ by the rounding of the
Lloyd step's centroids).

    python tests/perf_map_stage.py [--steps 3]

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
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args()
    import torch
    torch.cuda.init()
    from kiwi_amd import hip, workloads
    from oracle import oracle as O
    from perf_voronoi_map import coverage_buffer

    parts, dom, _ = workloads.lognormal_field(n_particles=200_000, n=256, seed=2)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in workloads.COLS}
    dom = [float(v) for v in dom]
    pos = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    dev = torch.device("cuda:0")
    dparts = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    dpos = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    ctx = hip.Context(0)

    def stage():
        ids, buf, _ = ctx.coverage_buffer(dparts, dom)
        buf = torch.as_tensor(np.asarray(buf, dtype=np.float64).reshape(-1, 3), device=dev)
        sites = torch.cat([dpos, buf]) if len(buf) else dpos
        c0, _ = ctx.voronoi_tessellate(sites, dom)
        s0 = sites[c0["cell_site"].long()]
        cen, _ = ctx.voronoi_centroids(c0, s0)
        c1, _ = ctx.voronoi_tessellate(cen, dom)
        home, unres = ctx.voronoi_home_cells(dpos)
        nc = len(c1["cell_site"])
        out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
        st = ctx.voronoi_map(dparts, home, c1, out=out)[4]
        return out, st, nc, int(len(buf)), unres

    stage()
    torch.cuda.synchronize()
    ts = []
    for _ in range(args.steps):
        a = time.perf_counter()
        out, st, nc, nbuf, unres = stage()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    ms = 1e3 * float(np.median(ts))
    log(f"GPU stage: {ms:.1f} ms ({nc} cells, {nbuf} buffer sites, {unres} unresolved particles, map visits "
        f"{st.pairs_visited})")

    # the same stage with voro++ geometry and home cells (the coverage buffer from the GPU's, which equals the
    # reference's own function)
    a = time.perf_counter()
    buf = coverage_buffer(pos, parts["h"], np.asarray(dom)).reshape(-1, 3)
    sites = np.concatenate([pos, buf])
    g0 = O.tessellate(np.arange(1, len(sites) + 1), sites, np.asarray(dom), np.zeros((0, 3)))
    g = O.tessellate(g0["cell_id"], g0["centroid"], np.asarray(dom), pos)
    t_voro = time.perf_counter() - a
    cells = {k: np.ascontiguousarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    cells["vert_xyz"] = np.ascontiguousarray(g["vert_xyz"].reshape(-1))
    cells["domain"] = dom
    ref = ctx.voronoi_map(parts, g["home"].astype(np.int32), cells)
    got = [o.cpu().numpy() for o in out]
    err = {n: float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for n, a, b in zip(("mass", "px", "py", "pz"), got, ref[:4])}
    mass_rel = float(abs(got[0].sum() - ref[0].sum()) / ref[0].sum())
    log(f"voro++ buffer + two tessellations + home cells (host, one thread): {t_voro:.2f} s; GPU stage vs voro++ "
        f"geometry: field error over max {err}, total mass {mass_rel:.2e}, visits {st.pairs_visited} vs "
        f"{ref[4].pairs_visited}")
    print(json.dumps({"workload": "Voronoi map stage, 2e5 log-normal particles (256^3 field, sigma_ln_rho 1.5), "
                                  "M4, FP64: coverage buffer, tessellation, one Lloyd step, final tessellation, "
                                  "home cells and map, all on the GPU",
                      "cells": nc, "gpu_ms_stage": ms, "host_s_voro_geometry": t_voro,
                      "field_err_over_max_vs_voro_geometry": err, "total_mass_rel_diff": mass_rel,
                      "visits": [int(st.pairs_visited), int(ref[4].pairs_visited)]}), flush=True)


if __name__ == "__main__":
    main()
