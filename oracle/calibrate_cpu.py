#!/usr/bin/env python3
"""Times the reference CLI (oracle/_ref/kiwi, built from the reference's sources by oracle/ref.mk) and
the CPU oracle on identical sph-to-grid inputs on this machine, and writes oracle/cpu_calibration.json.

The GPU host runs only the oracle (the reference build is not part of the measurement there); the
ratio oracle time / reference time measured here converts the oracle's rate on that host into an
estimate of the reference's.

Cases (M4, FP64, the reference's default kernel path):
  c1            the 2,000-particle uniform cube -> 32^3 (tests/golden/c1_snapshot.npz)
  lognormal/16  every 16th particle of the 1M log-normal field -> 256^3 (62,500 particles)
  lognormal/128 every 128th particle of the same field -> 256^3 (7,813 particles)

The reference's time is the wall time of `kiwi sph-to-grid` (plain_ascii read, the particle loop,
normalisation, the VGTCGRD2 write); the parse and write are measured apart from it on a run with no
particles in the grid and subtracted (`ref_overhead_s`).  The oracle's time is its sph-to-grid call
alone.  Throughput is the reference's own pair count (bounding-box pairs) and the sphere pairs of the
same run, per second.

  python oracle/calibrate_cpu.py [--quick]
"""
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:] = [ROOT] + [q for q in sys.path if os.path.abspath(q or '.') != os.path.dirname(os.path.abspath(__file__))]
from kiwi_amd import workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "oracle", "cpu_calibration.json")


def ref_run(parts, domain, res, threads):
    """Wall seconds of the reference CLI's sph-to-grid on these particles."""
    with tempfile.TemporaryDirectory() as td:
        snap = os.path.join(td, "snap.txt")
        workloads.write_plain_ascii(parts, snap)
        cfg = {"run_name": "calib", "input_snapshot": snap, "input_snapshot_format": "plain_ascii",
               "work_folder": os.path.join(td, "work"),
               "domain": dict(zip(["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"], map(float, domain))),
               "cartesian_grid": {"resolution": int(res[0]), "resolution_x": int(res[0]),
                                  "resolution_y": int(res[1]), "resolution_z": int(res[2])}}
        cp = os.path.join(td, "cfg.json")
        with open(cp, "w") as f:
            json.dump(cfg, f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        subprocess.run([O.REF_CLI, "sph-to-grid", cp], env=env, capture_output=True, text=True, check=True)
        return time.perf_counter() - t0


def oracle_run(parts, domain, res, threads):
    t0 = time.perf_counter()
    *_, st = O.sph_to_grid(parts, domain, *res, threads=threads)
    return time.perf_counter() - t0, st


def sphere_pairs(parts, domain, res):
    """Pairs whose cell's closest point lies within 2h of the particle (the throughput unit)."""
    nx, ny, nz = res
    d = [(domain[1] - domain[0]) / nx, (domain[3] - domain[2]) / ny, (domain[5] - domain[4]) / nz]
    lo = [domain[0], domain[2], domain[4]]
    n = [nx, ny, nz]
    tot = 0
    for i in range(len(parts["x"])):
        hh = parts["h"][i]
        if not hh > 0:
            continue
        H = 2.0 * hh
        c = [parts["x"][i], parts["y"][i], parts["z"][i]]
        gaps = []
        for a in range(3):
            i0 = max(int(np.floor((c[a] - H - lo[a]) / d[a])), 0)
            i1 = min(int(np.floor((c[a] + H - lo[a]) / d[a])), n[a] - 1)
            if i0 > i1:
                break
            idx = np.arange(i0, i1 + 1)
            cl, ch = lo[a] + idx * d[a], lo[a] + (idx + 1) * d[a]
            g = np.where(cl > c[a], cl - c[a], np.where(ch < c[a], c[a] - ch, 0.0))
            gaps.append(g * g)
        if len(gaps) < 3:
            continue
        s = gaps[0][:, None, None] + gaps[1][None, :, None] + gaps[2][None, None, :]
        tot += int(np.count_nonzero(s < H * H))
    return tot


def main():
    quick = "--quick" in sys.argv
    if not os.path.exists(O.REF_CLI):
        sys.exit("oracle/_ref/kiwi is not built (make -f oracle/ref.mk)")
    ncpu = os.cpu_count() or 1
    cases = []
    c1 = np.load(os.path.join(ROOT, "tests", "golden", "c1_snapshot.npz"))
    p1 = {k: np.ascontiguousarray(c1[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")}
    cases.append(("c1", p1, np.array(c1["domain"]), (32, 32, 32), [ncpu, 1]))
    if not quick:
        parts, dom, res = workloads.lognormal_field(n_particles=1_000_000, n=256, seed=2)
        for every, thr in ((16, [ncpu]), (128, [1])):
            sub = {k: np.ascontiguousarray(v[::every]) for k, v in parts.items()}
            sub["rho"] = np.ones_like(sub["m"])
            cases.append((f"lognormal/{every}", sub, np.asarray(dom, dtype=np.float64), res, thr))
    rows = []
    for name, parts, dom, res, threads in cases:
        # the CLI's parse / write overhead: the same file with every particle moved outside the grid
        far = dict(parts)
        far["x"] = parts["x"] + 1e6
        overhead = ref_run(far, dom, res, threads[0])
        sp = sphere_pairs(parts, dom, res)
        for t in threads:
            tr = ref_run(parts, dom, res, t) - overhead
            to, st = oracle_run(parts, dom, res, t)
            row = {"case": name, "particles": len(parts["x"]), "grid": list(res), "threads": t,
                   "pairs_bbox": int(st.pairs_bbox), "pairs_sphere": sp,
                   "ref_s": tr, "ref_overhead_s": overhead, "oracle_s": to, "oracle_over_ref": to / tr,
                   "ref_pairs_sphere_per_s": sp / tr, "oracle_pairs_sphere_per_s": sp / to,
                   "ref_pairs_bbox_per_s": st.pairs_bbox / tr}
            print(json.dumps(row), flush=True)
            rows.append(row)
    info = {"host": f"{ncpu} cores, " + open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0]
            .strip(" :\t"), "reference_build": "oracle/ref.mk (g++ -O2 -fopenmp, the reference's Release -O2)",
            "rows": rows}
    with open(OUT, "w") as f:
        json.dump(info, f, indent=1)
    print("written", OUT)


if __name__ == "__main__":
    main()
