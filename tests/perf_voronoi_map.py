"""Measurement of the Voronoi map (kiwi_hip_voronoi_map) on its benchmark workload; not collected by
pytest.  The tessellation is host input, as in the reference (voro++ stays on the host), and comes from
the reference's vendored voro++ through oracle/_ref/libkiwi_voro.so, so this harness lives with the
tests; the CPU baseline is the oracle's BFS (oracle_voronoi_map, one thread) on a bounded sample.

Workload: 2e5 particles from the log-normal density field of kiwi_amd/workloads.py (256^3 box,
sigma_ln_rho 1.5), sites = the particle positions plus the coverage-grid buffer (candidates on a
lattice of spacing 0.05 x the smallest extent, kept where no particle lies within 2 h; src/point_set.cpp:
167-272), then one Lloyd step (each site moved to its cell's centroid; src/point_set.cpp:275-306), M4
kernel, FP64; then the same map in mixed precision (face integrals in FP32), timed and compared with
the FP64 result; then the inverse-mapping test (kiwi_hip_inverse_mapping) on the same inputs with the
FP64 map's cell masses, in both precisions.

    python tests/perf_voronoi_map.py [--particles 200000] [--steps 5] [--warmup 2] [--sample 3000]

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


def coverage_buffer(pos, h, dom, spacing_fraction=0.05, factor=2.0):
    """Lattice candidates not within factor * h_i of any particle (add_coverage_grid_buffer)."""
    from scipy.spatial import cKDTree
    lo, hi = dom[0::2], dom[1::2]
    ext = hi - lo
    sp = spacing_fraction * ext.min()
    n = [max(1, int(round(e / sp))) for e in ext]
    axes = [lo[a] + sp * (np.arange(n[a]) + 0.5) for a in range(3)]
    cand = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    tree = cKDTree(pos)
    keep = []
    for c, near in zip(cand, tree.query_ball_point(cand, factor * h.max())):
        near = np.asarray(near, dtype=np.int64)
        covered = near.size > 0 and np.any(np.sum((pos[near] - c) ** 2, axis=1) <= (factor * h[near]) ** 2)
        if not covered:
            keep.append(c)
    return np.array(keep).reshape(-1, 3)


def build(n_particles, seed=2, field=256):
    from oracle import oracle as O
    t = time.perf_counter()
    parts, dom, _ = workloads.lognormal_field(n_particles=n_particles, n=field, seed=seed)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in workloads.COLS}
    pos = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    buf = coverage_buffer(pos, parts["h"], dom)
    sites = np.concatenate([pos, buf])
    ids = np.arange(1, len(sites) + 1)
    log(f"{n_particles} particles, {len(buf)} buffer sites ({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    g0 = O.tessellate(ids, sites, dom, np.zeros((0, 3)))
    g = O.tessellate(g0["cell_id"], g0["centroid"], dom, pos)  # one Lloyd step, then the final tessellation
    log(f"tessellated twice: {len(g['volume'])} cells, {len(g['face_nbr'])} faces, {len(g['vert_xyz'])} face "
        f"vertices ({time.perf_counter() - t:.1f} s)")
    cells = {k: np.ascontiguousarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    cells["vert_xyz"] = np.ascontiguousarray(g["vert_xyz"].reshape(-1))
    cells["domain"] = dom
    return parts, g["home"].astype(np.int32), cells, g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=200_000)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sample", type=int, default=3000, help="particles in the CPU-baseline sample (0: none)")
    args = ap.parse_args()

    import torch
    torch.cuda.init()
    from kiwi_amd import hip
    from oracle import oracle as O

    parts, home, cells, geo = build(args.particles)
    dev = torch.device("cuda:0")
    dparts = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    dcells = {k: torch.from_numpy(v).to(dev) for k, v in cells.items() if k != "domain"}
    dcells["domain"] = cells["domain"]
    dhome = torch.from_numpy(home).to(dev)
    nc = len(cells["cell_face_ptr"]) - 1
    out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
    ctx = hip.Context(0)
    for _ in range(args.warmup):
        ctx.voronoi_map(dparts, dhome, dcells, out=out)
    torch.cuda.synchronize()
    times, st = [], None
    for _ in range(args.steps):
        t = time.perf_counter()
        st = ctx.voronoi_map(dparts, dhome, dcells, out=out)[4]
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    ms = 1e3 * float(np.median(times))
    log(f"GPU: {ms:.2f} ms/call (prep {st.device_ms_prep:.2f}, faces {st.device_ms_faces:.2f}, accum "
        f"{st.device_ms_accum:.2f}); candidates {st.pairs_candidate}, visited {st.pairs_visited}, face "
        f"integrals {st.face_evals}, unmapped {st.unmapped}, unresolved {st.unresolved}")
    mass = out[0].cpu().numpy()
    f64 = [o.cpu().numpy() for o in out]
    # mixed precision: the face integrals in FP32
    out32 = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
    for _ in range(args.warmup):
        ctx.voronoi_map(dparts, dhome, dcells, precision=hip.MIXED, out=out32)
    torch.cuda.synchronize()
    times32, st32 = [], None
    for _ in range(args.steps):
        t = time.perf_counter()
        st32 = ctx.voronoi_map(dparts, dhome, dcells, precision=hip.MIXED, out=out32)[4]
        torch.cuda.synchronize()
        times32.append(time.perf_counter() - t)
    ms32 = 1e3 * float(np.median(times32))
    f32 = [o.cpu().numpy() for o in out32]
    mixed = {"ms_per_call": ms32, "device_ms": {"prep": st32.device_ms_prep, "faces": st32.device_ms_faces,
                                                "accum": st32.device_ms_accum},
             "same_visits": bool(st32.pairs_visited == st.pairs_visited and st32.face_evals == st.face_evals),
             "total_mass_rel_err": float(abs(f32[0].sum() - f64[0].sum()) / f64[0].sum()),
             "field_err_over_max": {n: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
                                    for n, a, b in zip(("mass", "px", "py", "pz"), f32, f64)}}
    log(f"mixed: {ms32:.2f} ms/call (faces {st32.device_ms_faces:.2f}); total mass {mixed['total_mass_rel_err']:.2e}, "
        f"fields {mixed['field_err_over_max']}")
    # the inverse-mapping test on the same inputs, forward cell masses from the FP64 map
    cm = out[0].clone()
    drec = torch.zeros(len(parts["x"]), dtype=torch.float64, device=dev)
    dcov = torch.zeros_like(drec)
    inverse = {}
    for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        ctx.inverse_mapping(dparts, dhome, dcells, cm, precision=prec, out=(drec, dcov))
        torch.cuda.synchronize()
        ti = []
        for _ in range(args.steps):
            t = time.perf_counter()
            ctx.inverse_mapping(dparts, dhome, dcells, cm, precision=prec, out=(drec, dcov))
            torch.cuda.synchronize()
            ti.append(time.perf_counter() - t)
        cv = dcov.cpu().numpy()
        inverse[name] = {"ms_per_call": 1e3 * float(np.median(ti)), "coverage_min": float(cv.min()),
                         "coverage_max": float(cv.max()),
                         "recovered_over_original": float(drec.sum().item() / parts["m"].sum())}
    log(f"inverse-mapping test: {inverse}")
    res = {"workload": "voronoi map, 2e5 log-normal particles (256^3 field, sigma_ln_rho 1.5) + coverage-grid "
                       "buffer, 1 Lloyd step, M4, FP64 (and mixed)",
           "particles": args.particles, "cells": nc, "faces": len(cells["face_nbr"]),
           "ms_per_call": ms, "device_ms": {"prep": st.device_ms_prep, "faces": st.device_ms_faces,
                                            "accum": st.device_ms_accum},
           "pairs_candidate": st.pairs_candidate, "pairs_visited": st.pairs_visited, "face_evals": st.face_evals,
           "cell_integrals_per_s": st.pairs_visited / (ms * 1e-3),
           "face_integrals_per_s": st.face_evals / (ms * 1e-3),
           "face_kernel_face_integrals_per_s": st.face_evals / (st.device_ms_faces * 1e-3),
           "mass_total_over_particles": float(mass.sum() / parts["m"].sum()), "mixed": mixed,
           "inverse_mapping": inverse}
    if args.sample > 0:
        k = min(args.sample, args.particles)
        sub = {c: np.ascontiguousarray(v[:k]) for c, v in parts.items()}
        pid = np.arange(1, k + 1, dtype=np.int32)
        t = time.perf_counter()
        ref = O.voronoi_map(sub, pid, home[:k], cells["cell_face_ptr"], cells["face_vert_ptr"], cells["vert_xyz"],
                            cells["face_nbr"])
        cpu_s = time.perf_counter() - t
        got = ctx.voronoi_map(sub, home[:k].copy(), cells)
        err = max(float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) for a, b in zip(got[:4], ref[:4]))
        res["cpu_baseline"] = {"value": float(ref[4][4] / cpu_s), "unit": "cell integrals/s", "cores": 1,
                               "kind": "port", "sample": f"first {k} particles of the same workload "
                                                         f"({int(ref[4][4])} visited pairs, {cpu_s:.1f} s)"}
        res["sample_max_rel_err_vs_oracle"] = err
        log(f"CPU oracle: {k} particles in {cpu_s:.1f} s; GPU vs oracle on the sample: {err:.2e} of field max")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
