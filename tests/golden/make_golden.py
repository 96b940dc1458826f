#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference KIWI built from source in this
container (oracle/_ref, by `make -f oracle/ref.mk`).  Runs only where /root/reference exists; the
fixtures it writes are plain numbers (inputs and the reference's outputs), committed to the repo.

  c1_snapshot.npz       the 2,000-particle uniform-cube input (x y z vx vy vz m h rho) + domain
  c1_ref_grid.npz       `kiwi sph-to-grid` (reference CLI, OMP_NUM_THREADS=1) on it: 32^3 VGTCGRD2 fields
  c1_ref_stderr.txt     the reference CLI's log lines for that run (counters)
  m4_vertex.npz         (h, phi, r0, R0) -> full_integral_new, random + boundary cases
  generic_vertex.npz    (kernel, H, phi, r0, R0) -> generic_full_integral for all six kernels
  face.npz              random planar faces -> compute_face_integral(_generic) and out_vert
  m4_gridded.npz        (h, phi, r0, R0) -> gridded_integral on the generated lookup table
  kernel_grids.npz      `kiwi sph-to-grid` (reference CLI, OMP_NUM_THREADS=1) on every 4th C1 particle
                        onto 20^3 for each kernel on the generic path (M4 forced generic, M4 with
                        support 2.5h, M5, M6, WC2, WC4, WC6) and for the tabulated cubic path

  voronoi_map.npz       `kiwi map` (reference CLI, OMP_NUM_THREADS=1, no buffer, no Lloyd step: the sites
                        are the particles) on every 8th C1 particle, M4: the per-cell volume, density and
                        velocity of mapping.dat, with the tessellation of the same sites by the reference's
                        vendored voro++ as CSR arrays and each particle's home cell

  inverse_mapping.npz   `kiwi inverse-mapping-test` (reference CLI, OMP_NUM_THREADS=1) chained from the same
                        `kiwi map` run: per particle (by id) m_original, m_recovered, coverage_fraction,
                        and the totals line of inverse-mapping-test.dat

  cartesian_map.npz     `kiwi cartesian-map` (reference CLI, resolution 20) chained from the same `kiwi map` run:
                        the density, velocity and Voronoi cell id fields of cartesian_map.dat

  realdata.npz          the reference's only real-data fixture, tests/data/brick_03927_cutout.ascii (a
                        SPLASH export of a Phantom run: 7,300 gas particles as the reference's own reader
                        keeps them), with the domain rule of the reference's tests/real_data_test.py
                        (bounding box padded by 3 x 1.02 x max h, whole 1 pc cells): `kiwi sph-to-grid`
                        (OMP_NUM_THREADS=1) for the cubic spline, the cubic spline through the generic
                        integral, the tabulated cubic, the quintic spline and Wendland C6, plus the cubic
                        spline on 0.2 pc cells over the same padded box; and `kiwi map` with the
                        coverage_grid buffer and one Lloyd step (the test's Voronoi checks): the point set
                        after the buffer (points-buff.dat), after the Lloyd step (points-final.dat) and
                        mapping.dat; and `kiwi map` with neither (the sites are the particles, so the
                        tessellation here is the reference's own): mapping.dat

  python tests/golden/make_golden.py [--only kernel_grids | --only voronoi | --only inverse | --only cmap
                                    | --only coverage | --only realdata]
"""
import ctypes
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from kiwi_amd import workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def read_vgtcgrd2(path):
    with open(path, "rb") as f:
        raw = f.read()
    assert raw[:8] == b"VGTCGRD2"
    nx, ny, nz = np.frombuffer(raw, dtype="<i4", count=3, offset=8)
    dom = np.frombuffer(raw, dtype="<f8", count=6, offset=20)
    n = int(nx) * int(ny) * int(nz)
    body = np.frombuffer(raw, dtype="<f8", count=4 * n, offset=68).reshape(4, n)
    return (int(nx), int(ny), int(nz)), dom, body


def run_ref_sph_to_grid(parts, domain, res, kernel=None, threads=1, extra=None, tabulated=False):
    with tempfile.TemporaryDirectory() as td:
        snap = os.path.join(td, "snap.txt")
        workloads.write_plain_ascii(parts, snap)
        cfg = {"run_name": "golden", "input_snapshot": snap, "input_snapshot_format": "plain_ascii",
               "work_folder": os.path.join(td, "work"),
               "domain": dict(zip(["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"], map(float, domain))),
               "cartesian_grid": {"resolution": int(res[0]), "resolution_x": int(res[0]),
                                  "resolution_y": int(res[1]), "resolution_z": int(res[2])}}
        if kernel:
            cfg["kernel"] = kernel
        if tabulated:
            cfg["cartesian_grid"]["use_tabulated_integral"] = True
        if extra:
            cfg.update(extra)
        cp = os.path.join(td, "cfg.json")
        with open(cp, "w") as f:
            json.dump(cfg, f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([O.REF_CLI, "sph-to-grid", cp], env=env, capture_output=True, text=True, check=True)
        shape, dom, body = read_vgtcgrd2(os.path.join(td, "work", "snap", "cartesian_grid.dat"))
        return shape, dom, body, r.stderr


# (label, kernel config) of the generic-path grids
KERNEL_GRID_CASES = [("M4_generic", {"name": "m4", "generic_path": True}),
                     ("M4_support2.5", {"name": "m4", "support_radius_in_h": 2.5}),
                     ("M5", {"name": "m5"}), ("M6", {"name": "m6"}), ("WC2", {"name": "wc2"}),
                     ("WC4", {"name": "wc4"}), ("WC6", {"name": "wc6"})]


def kernel_grids():
    """Reference grids for the generic path, one per kernel case (plain numbers, committed)."""
    c1 = np.load(os.path.join(OUT, "c1_snapshot.npz"))
    parts = {k: np.ascontiguousarray(c1[k][::4]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    domain = np.array(c1["domain"])
    res = (20, 20, 20)
    out = {"domain": domain, "res": np.array(res), **{f"p_{k}": v for k, v in parts.items()}}
    for label, kcfg in KERNEL_GRID_CASES + [("M4_tabulated", None)]:
        shape, dom, body, _ = run_ref_sph_to_grid(parts, domain, res, kernel=kcfg, tabulated=(label == "M4_tabulated"))
        assert tuple(shape) == res
        for i, f in enumerate(("density", "vx", "vy", "vz")):
            out[f"{label}_{f}"] = body[i]
    np.savez_compressed(os.path.join(OUT, "kernel_grids.npz"), **out)


def run_ref_map(parts, domain, threads=1, inverse=False, cmap_res=0, buffer=None):
    with tempfile.TemporaryDirectory() as td:
        snap = os.path.join(td, "snap.txt")
        workloads.write_plain_ascii(parts, snap)
        cfg = {"run_name": "golden", "input_snapshot": snap, "input_snapshot_format": "plain_ascii",
               "work_folder": os.path.join(td, "work"),
               "domain": dict(zip(["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"], map(float, domain))),
               "buffer": {"enabled": False}, "lloyd": {"iterations": 0}}
        if buffer:
            cfg["buffer"] = {"enabled": True, "strategy": "coverage_grid", "coverage_grid": buffer}
        if cmap_res:
            cfg["cartesian_map"] = {"resolution": int(cmap_res)}
        cp = os.path.join(td, "cfg.json")
        with open(cp, "w") as f:
            json.dump(cfg, f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([O.REF_CLI, "map", cp], env=env, capture_output=True, text=True, check=True)
        rows = np.loadtxt(os.path.join(td, "work", "snap", "mapping.dat"))
        if cmap_res:
            subprocess.run([O.REF_CLI, "cartesian-map", cp], env=env, capture_output=True, text=True, check=True)
            with open(os.path.join(td, "work", "snap", "cartesian_map.dat"), "rb") as f:
                raw = f.read()
            assert raw[:8] == b"VGTCMAP1"
            res = int(np.frombuffer(raw, np.int32, 1, 8)[0])
            n3 = res ** 3
            off = 12 + 48
            fields = np.frombuffer(raw, np.float64, 4 * n3, off).reshape(4, n3)
            ids = np.frombuffer(raw, np.int32, n3, off + 32 * n3)
            return rows, res, fields, ids
        if not inverse:
            return rows, r.stderr
        subprocess.run([O.REF_CLI, "inverse-mapping-test", cp], env=env, capture_output=True, text=True,
                       check=True)
        inv = np.loadtxt(os.path.join(td, "work", "snap", "inverse-mapping-test.dat"), skiprows=1, max_rows=1)
        prow = np.loadtxt(os.path.join(td, "work", "snap", "inverse-mapping-test.dat"), skiprows=2, ndmin=2)
        return rows, inv, prow


def voronoi():
    """The reference's Voronoi map on a small site set, with the geometry the GPU path consumes."""
    c1 = np.load(os.path.join(OUT, "c1_snapshot.npz"))
    parts = {k: np.ascontiguousarray(c1[k][::8]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    domain = np.array(c1["domain"])
    rows, log = run_ref_map(parts, domain)
    n = len(parts["x"])
    pts = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    ids = np.arange(1, n + 1)  # plain_ascii ids are the 1-based row numbers (src/snapshot.cpp:67-68)
    g = O.tessellate(ids, pts, domain, pts)
    by_id = {int(r[0]): r for r in rows}
    ref = np.array([by_id[int(i)] for i in g["cell_id"]])
    np.savez_compressed(os.path.join(OUT, "voronoi_map.npz"), domain=domain, pid=ids.astype(np.int32),
                        **{f"p_{k}": v for k, v in parts.items()}, **g,
                        ref_volume=ref[:, 4], ref_density=ref[:, 5], ref_vx=ref[:, 6], ref_vy=ref[:, 7],
                        ref_vz=ref[:, 8], ref_site=ref[:, 1:4])


def inverse():
    """The reference's inverse-mapping test on the voronoi() setup (same particles, sites, domain)."""
    c1 = np.load(os.path.join(OUT, "c1_snapshot.npz"))
    parts = {k: np.ascontiguousarray(c1[k][::8]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    domain = np.array(c1["domain"])
    _, totals, prow = run_ref_map(parts, domain, inverse=True)
    order = np.argsort(prow[:, 0])
    prow = prow[order]
    np.savez_compressed(os.path.join(OUT, "inverse_mapping.npz"), pid=prow[:, 0].astype(np.int32),
                        m_original=prow[:, 1], m_recovered=prow[:, 2], coverage=prow[:, 3],
                        total_original=totals[0], total_grid=totals[1], total_recovered=totals[2])


def cmap():
    """The reference's cartesian-map resample on the voronoi() setup, resolution 20."""
    c1 = np.load(os.path.join(OUT, "c1_snapshot.npz"))
    parts = {k: np.ascontiguousarray(c1[k][::8]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    domain = np.array(c1["domain"])
    _, res, fields, ids = run_ref_map(parts, domain, cmap_res=20)
    np.savez_compressed(os.path.join(OUT, "cartesian_map.npz"), res=res, density=fields[0], vx=fields[1],
                        vy=fields[2], vz=fields[3], cell_id=ids.astype(np.int32))


def coverage():
    """The reference's coverage-grid buffer (add_coverage_grid_buffer, src/point_set.cpp:208-273): the
    buffer sites of mapping.dat (ids past the particle count, with their positions) for two setups with
    a mix of covered and uncovered candidates: the voronoi() particles (spacing 0.06, radius factor 1.2)
    and 600 particles of a small log-normal field, h spread over a decade (spacing 0.04, factor 0.6)."""
    c1 = np.load(os.path.join(OUT, "c1_snapshot.npz"))
    cases = {}
    parts = {k: np.ascontiguousarray(c1[k][::8]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    cases["c1"] = (parts, np.array(c1["domain"]), {"spacing_fraction": 0.06, "coverage_radius_factor": 1.2})
    lp, ldom, _ = workloads.lognormal_field(n_particles=600, n=32, seed=4, h_min=0.3, h_max=6.0)
    lp = {k: np.ascontiguousarray(lp[k], dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    lp["rho"] = np.ones_like(lp["m"])
    cases["lognormal"] = (lp, np.asarray(ldom, dtype=np.float64), {"spacing_fraction": 0.04,
                                                                   "coverage_radius_factor": 0.6})
    out = {}
    for name, (pp, dom, prm) in cases.items():
        with tempfile.TemporaryDirectory() as td:  # the particles as the reference parsed them
            snap = os.path.join(td, "snap.txt")
            workloads.write_plain_ascii(pp, snap)
            back = np.loadtxt(snap, ndmin=2)
        rows, _ = run_ref_map(pp, dom, buffer=prm)
        n = len(pp["x"])
        buf = rows[rows[:, 0] > n]
        buf = buf[np.argsort(buf[:, 0])]
        print(f"coverage {name}: {n} particles, {len(buf)} buffer sites", file=sys.stderr)
        out.update({f"{name}_domain": dom, f"{name}_spacing_fraction": prm["spacing_fraction"],
                    f"{name}_factor": prm["coverage_radius_factor"], f"{name}_snapshot": back,
                    f"{name}_buffer_id": buf[:, 0].astype(np.int32), f"{name}_buffer_xyz": buf[:, 1:4]})
    np.savez_compressed(os.path.join(OUT, "coverage_buffer.npz"), **out)


REAL_SNAPSHOT = "/root/reference/tests/data/brick_03927_cutout.ascii"
REAL_CASES = [("cubic", None, False), ("cubic_generic", {"generic_path": True}, False), ("cubic_tab", None, True),
              ("quintic_spline", {"name": "quintic_spline"}, False), ("wendland_c6", {"name": "wendland_c6"}, False)]


def real_domain(parts, cell):
    """The reference test's padded domain (tests/real_data_test.py:112-119): bounding box padded by
    3 x 1.02 x max h, whole cells of `cell` per axis."""
    pad = 3.0 * float(parts["h"].max()) * 1.02
    dom, res = [], []
    for c in ("x", "y", "z"):
        a, b = float(parts[c].min()) - pad, float(parts[c].max()) + pad
        n = max(1, int(np.ceil((b - a) / cell)))
        dom += [a, a + n * cell]
        res.append(n)
    return np.array(dom), tuple(res)


def realdata():
    """The reference CLI on its own real-data cutout (module docstring)."""
    parts = O.ref_load_snapshot(REAL_SNAPSHOT)
    cols = ("x", "y", "z", "vx", "vy", "vz", "m", "h")
    out = {f"p_{k}": parts[k] for k in cols}
    out["p_id"] = parts["id"]
    dom, res = real_domain(parts, 1.0)
    out["domain"], out["res"] = dom, np.array(res)
    for label, kcfg, tab in REAL_CASES:
        shape, gdom, body = run_ref_file(REAL_SNAPSHOT, dom, res, kernel=kcfg, tabulated=tab)
        assert tuple(shape) == res and np.array_equal(gdom, dom)
        for i, f in enumerate(("density", "vx", "vy", "vz")):
            out[f"{label}_{f}"] = body[i]
        print(f"realdata {label}: {res}", file=sys.stderr)
    fdom, fres = real_domain(parts, 0.2)
    shape, _, body = run_ref_file(REAL_SNAPSHOT, fdom, fres)
    out["fine_domain"], out["fine_res"] = fdom, np.array(fres)
    for i, f in enumerate(("density", "vx", "vy", "vz")):
        out[f"fine_{f}"] = body[i]
    print(f"realdata fine cubic: {fres}", file=sys.stderr)
    buff, final, rows = run_ref_file_map(REAL_SNAPSHOT, dom)
    out["buff_id"], out["buff_xyz"] = buff[:, 0].astype(np.int32), buff[:, 1:4]
    out["final_id"], out["final_xyz"] = final[:, 0].astype(np.int32), final[:, 1:4]
    out["map_rows"] = rows
    _, _, plain = run_ref_file_map(REAL_SNAPSHOT, dom, buffer=False)
    out["map_plain_rows"] = plain
    print(f"realdata map: {len(buff)} sites, {len(rows)} rows", file=sys.stderr)
    np.savez_compressed(os.path.join(OUT, "realdata.npz"), **out)


def run_ref_file(snapshot, domain, res, kernel=None, tabulated=False):
    with tempfile.TemporaryDirectory() as td:
        cfg = {"run_name": "real", "input_snapshot": snapshot, "work_folder": os.path.join(td, "work"),
               "domain": dict(zip(["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"], map(float, domain))),
               "cartesian_grid": {"resolution_x": int(res[0]), "resolution_y": int(res[1]),
                                  "resolution_z": int(res[2]), "use_tabulated_integral": bool(tabulated)}}
        if kernel:
            cfg["kernel"] = kernel
        cp = os.path.join(td, "cfg.json")
        with open(cp, "w") as f:
            json.dump(cfg, f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([O.REF_CLI, "sph-to-grid", cp], env=env, capture_output=True, text=True, check=True)
        base = os.path.splitext(os.path.basename(snapshot))[0]
        return read_vgtcgrd2(os.path.join(td, "work", base, "cartesian_grid.dat"))


def run_ref_file_map(snapshot, domain, buffer=True):
    with tempfile.TemporaryDirectory() as td:
        cfg = {"run_name": "real", "input_snapshot": snapshot, "work_folder": os.path.join(td, "work"),
               "domain": dict(zip(["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"], map(float, domain))),
               "buffer": {"enabled": bool(buffer), "strategy": "coverage_grid"},
               "lloyd": {"iterations": 1 if buffer else 0, "save_each_iteration": False}}
        cp = os.path.join(td, "cfg.json")
        with open(cp, "w") as f:
            json.dump(cfg, f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([O.REF_CLI, "map", cp], env=env, capture_output=True, text=True, check=True)
        d = os.path.join(td, "work", os.path.splitext(os.path.basename(snapshot))[0])
        return (np.loadtxt(os.path.join(d, "points-buff.dat"), ndmin=2) if buffer else None,
                np.loadtxt(os.path.join(d, "points-final.dat"), ndmin=2),
                np.loadtxt(os.path.join(d, "mapping.dat"), ndmin=2))


def main():
    if not O.have_ref() or not os.path.exists(O.REF_CLI):
        sys.exit("oracle/_ref is not built (make -f oracle/ref.mk)")
    if "--only" in sys.argv:
        what = sys.argv[sys.argv.index("--only") + 1]
        {"kernel_grids": kernel_grids, "voronoi": voronoi, "inverse": inverse, "cmap": cmap,
         "coverage": coverage, "realdata": realdata}[what]()
        return
    R = O.ref()
    rng = np.random.default_rng(20251203)

    # --- C1: the 2,000-particle CPU-runnable sph-to-grid case --------------------------------------
    parts, domain, res = workloads.uniform_cube()
    np.savez_compressed(os.path.join(OUT, "c1_snapshot.npz"), domain=domain,
                        **{k: parts[k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")})
    # re-read what we wrote so the fixture is exactly the text the reference parsed
    shape, dom, body, log = run_ref_sph_to_grid(parts, domain, res)
    np.savez_compressed(os.path.join(OUT, "c1_ref_grid.npz"), shape=np.array(shape), domain=dom,
                        density=body[0], vx=body[1], vy=body[2], vz=body[3])
    with open(os.path.join(OUT, "c1_ref_stderr.txt"), "w") as f:
        f.write(log)

    # --- cubic vertex integral: random, plus the r0/R0/phi guards and region boundaries -----------
    n = 6000
    h = rng.uniform(0.05, 3.0, n)
    r0 = h * rng.uniform(0.0, 2.6, n)
    R0 = h * rng.uniform(0.0, 2.6, n)
    phi = rng.uniform(0.0, 0.5 * np.pi, n)
    # exact guard and boundary cases
    r0[:50] = 0.0
    R0[50:100] = 0.0
    phi[100:150] = 0.0
    r0[150:200] = h[150:200]            # r0 == h
    r0[200:250] = 2.0 * h[200:250]      # r0 == 2h
    R0[250:300] = np.sqrt(np.maximum(h[250:300] ** 2 - r0[250:300] ** 2, 0.0))  # line distance == h
    r0[300:400] = h[300:400] * 1e-6     # near the plane
    phi[400:500] = 0.5 * np.pi * (1 - 1e-9)
    val = np.array([R.ref_m4_vertex(*t) for t in zip(h, phi, r0, R0)])
    np.savez_compressed(os.path.join(OUT, "m4_vertex.npz"), h=h, phi=phi, r0=r0, R0=R0, value=val)

    # --- generic vertex integral, all six kernels ---------------------------------------------------
    n = 3000
    kid = rng.integers(0, 6, n).astype(np.int32)
    H = rng.uniform(0.2, 3.0, n)
    r0 = H * rng.uniform(0.0, 1.3, n)
    R0 = H * rng.uniform(0.0, 1.3, n)
    phi = rng.uniform(0.0, 0.5 * np.pi, n)
    val = np.array([R.ref_generic_vertex(int(k), *t) for k, t in zip(kid, zip(H, phi, r0, R0))])
    np.savez_compressed(os.path.join(OUT, "generic_vertex.npz"), kernel=kid, H=H, phi=phi, r0=r0, R0=R0, value=val)

    # --- whole faces: random convex polygons (3..7 vertices) in random planes, cubic and generic ----
    nf = 1500
    rows = []
    for i in range(nf):
        nv = int(rng.integers(3, 8))
        ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
        rad = rng.uniform(0.3, 1.5)
        e1 = rng.normal(size=3)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(e1, rng.normal(size=3))
        e2 /= np.linalg.norm(e2)
        c = rng.uniform(-1.5, 1.5, 3)
        verts = c + rad * (np.cos(ang)[:, None] * e1 + np.sin(ang)[:, None] * e2)
        if rng.uniform() < 0.5:
            verts = verts[::-1]
        p = rng.uniform(-1.0, 1.0, 3)
        hh = rng.uniform(0.1, 1.2)
        kk = int(rng.integers(0, 6))
        v = np.ascontiguousarray(verts.reshape(-1))
        ov = ctypes.c_int(0)
        fm = R.ref_m4_face(p[0], p[1], p[2], hh, O._p(v), nv, ctypes.byref(ov))
        ovg = ctypes.c_int(0)
        K = ["M4", "M5", "M6", "WC2", "WC4", "WC6"][kk]
        radk = {"M4": 2.0, "M5": 2.5, "M6": 3.0}.get(K, 2.0)
        fg = R.ref_generic_face(kk, p[0], p[1], p[2], radk * hh, O._p(v), nv, ctypes.byref(ovg))
        rows.append((nv, verts, p, hh, kk, fm, ov.value, fg, ovg.value))
    nvs = np.array([r[0] for r in rows], dtype=np.int32)
    verts = np.concatenate([r[1].reshape(-1) for r in rows])
    np.savez_compressed(os.path.join(OUT, "face.npz"), nverts=nvs, verts=verts,
                        p=np.array([r[2] for r in rows]), h=np.array([r[3] for r in rows]),
                        kernel=np.array([r[4] for r in rows], dtype=np.int32),
                        m4=np.array([r[5] for r in rows]), m4_out=np.array([r[6] for r in rows], dtype=np.int32),
                        generic=np.array([r[7] for r in rows]),
                        generic_out=np.array([r[8] for r in rows], dtype=np.int32))

    # --- tabulated cubic path -----------------------------------------------------------------------
    n = 3000
    h = rng.uniform(0.1, 3.0, n)
    r0 = h * rng.uniform(0.0, 2.5, n)
    R0 = h * rng.uniform(0.0, 3.0, n)
    phi = rng.uniform(0.0, 0.5 * np.pi, n)
    val = np.array([R.ref_m4_gridded(*t) for t in zip(h, phi, r0, R0)])
    np.savez_compressed(os.path.join(OUT, "m4_gridded.npz"), h=h, phi=phi, r0=r0, R0=R0, value=val)
    kernel_grids()
    voronoi()
    inverse()
    cmap()
    coverage()
    realdata()
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
