"""Measurement of the host snapshot readers (libkiwi_io) against the reference's own reader
(load_sph_snapshot through oracle/_ref/libkiwi_pts.so, one thread); not collected by pytest.
Workload: the sph-to-grid benchmark's 1e6 log-normal particles written as plain_ascii (%.17g, as the
reference's write_plain_ascii does), as a splash_ascii file with the same columns plus itype, and as a
tagged Phantom dump (positions and velocities real*8, h real*4, itype int*1).

    python tests/perf_io.py [--particles 1000000] [--threads 0] [--dir /tmp]

Prints one JSON line: seconds and particles/s for both readers and formats; the outputs are checked
equal bit for bit.
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

from kiwi_amd import io, workloads  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--dir", default="/tmp")
    args = ap.parse_args()
    from oracle import oracle as O
    from test_io import LABELS, phantom_dump, splash_text

    parts, _, _ = workloads.lognormal_field(n_particles=args.particles, n=256, seed=2)
    parts["rho"] = np.asarray(parts.get("rho", np.ones_like(parts["x"])), dtype=np.float64)
    plain = os.path.join(args.dir, f"kiwi_io_plain_{os.getpid()}.txt")
    splash = os.path.join(args.dir, f"kiwi_io_splash_{os.getpid()}.txt")
    workloads.write_plain_ascii(parts, plain)
    cols = np.stack([parts[k] for k in ("x", "y", "z", "m", "h", "rho", "vx", "vy", "vz")] +
                    [np.zeros_like(parts["x"])], 1)
    with open(splash, "w") as f:
        f.write(splash_text(LABELS, cols, itype=np.ones(len(cols), int)))
    phantom = os.path.join(args.dir, f"kiwi_io_phantom_{os.getpid()}.bin")
    arrays = [(c, 7, parts[c]) for c in ("x", "y", "z", "vx", "vy", "vz")] + [("h", 6, parts["h"]),
                                                                             ("itype", 1, np.ones(len(cols), int))]
    with open(phantom, "wb") as f:
        f.write(phantom_dump([(len(cols), arrays), (0, [])], npartoftype=(len(cols), 0)))
    threads = io.set_threads(args.threads)
    rec = {"metric": "snapshot read particles/s", "particles": args.particles, "threads": threads, "cases": {}}
    try:
        for name, path in (("plain_ascii", plain), ("splash_ascii", splash), ("phantom_binary", phantom)):
            t = time.perf_counter()
            a = io.load_snapshot(path)
            ta = time.perf_counter() - t
            t = time.perf_counter()
            b = O.ref_load_snapshot(path, cap=args.particles + 16)
            tb = time.perf_counter() - t
            same = all(np.array_equal(a[k], b[k]) for k in a)
            print(f"{name}: {os.path.getsize(path) / 1e6:.0f} MB, libkiwi_io {ta:.3f} s ({threads} threads), reference "
                  f"{tb:.3f} s, same {same}", file=sys.stderr, flush=True)
            assert same
            rec["cases"][name] = {"mb": os.path.getsize(path) / 1e6, "seconds": ta, "reference_seconds": tb,
                                  "value": args.particles / ta, "reference_value": args.particles / tb}
    finally:
        os.remove(plain)
        os.remove(splash)
        os.remove(phantom)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
