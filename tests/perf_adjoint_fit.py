"""Measurement of the adjoint (kiwi_hip_sph_to_grid_adjoint) on its benchmark workload; not collected by
pytest.

Workload: 2e5 particles from the log-normal density field of kiwi_amd/workloads.py on a 128^3 box
(sigma_ln_rho 1.5), M4 kernel, grid 128^3.  Loss L = sum_c (rho_c - rho_target_c)^2, rho_target from a
copy with 5% Gaussian mass perturbations.
  1. times: forward FP64 and mixed, adjoint FP64 and mixed;
  2. finite-difference checks of the FP64 adjoint on 1,000 random particles (m, x, y, z, h): central
     differences of the FP64 forward (fourth order for x, y, z, h), evaluated by replacing the
     particle's own contribution (a one-particle forward at the shifted value) in the full grid;
  3. the mixed adjoint against the FP64 adjoint (per gradient, relative to its largest value);
  4. 200 steps of mass recovery (Adam on the masses through torch.autograd; cubic spline only).
--kernel K times and checks the adjoint of another kernel (M5, M6, WC2, WC4, WC6 or M4g, the cubic
spline on the generic path) on the same workload.

    python tests/perf_adjoint_fit.py [--particles 200000] [--grid 128] [--fd 1000] [--steps 200] [--kernel M4]

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=200_000)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--fd", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--kernel", default="M4", help="M4 (analytic), M4g (generic path), M5, M6, WC2, WC4, WC6")
    args = ap.parse_args()

    import torch
    torch.cuda.init()
    from kiwi_amd import hip
    from kiwi_amd.autograd import sph_to_grid

    kern = hip.kernel_spec("M4", generic_path=True) if args.kernel == "M4g" else hip.kernel_spec(args.kernel)
    if args.kernel != "M4" and args.steps:
        sys.exit("the mass-recovery fit runs the cubic spline only: pass --steps 0 with --kernel")
    parts, dom, res = workloads.lognormal_field(n_particles=args.particles, n=args.grid, seed=4)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in workloads.COLS}
    n = len(parts["x"])
    dev = torch.device("cuda:0")
    ctx = hip.Context(0)
    nc = int(np.prod(res))
    V = float(np.prod((dom[1::2] - dom[0::2]) / np.array(res)))
    rng = np.random.default_rng(12)
    m_true = parts["m"].copy()
    pert = dict(parts)
    pert["m"] = m_true * (1 + 0.05 * rng.standard_normal(n))
    t = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    tp = dict(t)
    tp["m"] = torch.from_numpy(pert["m"]).to(dev)

    def forward(tt, precision=hip.FP64):
        out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
        ctx.sph_to_grid(tt, dom, res, kernel=kern, precision=precision, out=out)
        return out

    target = forward(tp)[0] / V
    res_json = {"workload": f"adjoint: {n} log-normal particles onto {res[0]}^3, {args.kernel}, "
                            f"L = sum (rho - rho_target)^2"}

    # 1. times
    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            a = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        return 1e3 * float(np.median(ts))

    mass = forward(t)[0]
    gm = (2 * (mass / V - target) / V).contiguous()
    zero = torch.zeros(nc, dtype=torch.float64, device=dev)
    gcells = (gm, zero, zero, zero)
    outs = {f: torch.zeros(n, dtype=torch.float64, device=dev) for f in hip.GRAD_FIELDS}
    outs32 = {f: torch.zeros(n, dtype=torch.float64, device=dev) for f in hip.GRAD_FIELDS}
    times = {"forward_fp64_ms": timed(lambda: forward(t)), "forward_mixed_ms": timed(lambda: forward(t, hip.MIXED)),
             "adjoint_fp64_ms": timed(lambda: ctx.sph_to_grid_adjoint(t, dom, res, gcells, kernel=kern, out=outs)),
             "adjoint_mixed_ms": timed(lambda: ctx.sph_to_grid_adjoint(t, dom, res, gcells, kernel=kern,
                                                                       precision=hip.MIXED, out=outs32))}
    st = ctx.sph_to_grid_adjoint(t, dom, res, gcells, kernel=kern, out=outs)[1]
    ctx.sph_to_grid_adjoint(t, dom, res, gcells, kernel=kern, precision=hip.MIXED, out=outs32)
    torch.cuda.synchronize()
    res_json.update(times)
    res_json["adjoint_cells"] = st.cells
    res_json["adjoint_lattice_points"] = st.lattice_points
    log(f"times: {times}; adjoint cells {st.cells}, lattice points {st.lattice_points}")
    g64 = {f: outs[f].cpu().numpy() for f in hip.GRAD_FIELDS}
    g32 = {f: outs32[f].cpu().numpy() for f in hip.GRAD_FIELDS}

    # 2. finite differences on random particles
    idx = rng.choice(n, min(args.fd, n), replace=False)
    one = {k: torch.empty(1, dtype=torch.float64, device=dev) for k in t}
    single_out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))

    def contribution(i, f=None, value=None):
        for k in t:
            one[k].copy_(t[k][i:i + 1])
        if f is not None:
            one[f].fill_(value)
        ctx.sph_to_grid(one, dom, res, kernel=kern, out=single_out)
        return single_out[0].clone()

    fd_err = {}
    a0 = time.perf_counter()
    for f in ("m", "x", "y", "z", "h"):
        scale = float(np.sqrt(np.mean(g64[f] ** 2)))
        worst = 0.0
        for i in idx:
            i = int(i)
            base = float(parts[f][i])
            own = contribution(i)
            if f == "m":  # L is quadratic in m: the central difference is exact up to rounding
                eps = 1e-6 * abs(base)
                rp = (mass - own + contribution(i, f, base + eps)) / V - target
                rm = (mass - own + contribution(i, f, base - eps)) / V - target
                fd = float(torch.sum(rp * rp - rm * rm)) / (2 * eps)
            else:  # fourth-order central difference
                eps = 1e-3 * float(parts["h"][i])
                d = []
                for k in (1, 2):
                    rp = (mass - own + contribution(i, f, base + k * eps)) / V - target
                    rm = (mass - own + contribution(i, f, base - k * eps)) / V - target
                    d.append(float(torch.sum(rp * rp - rm * rm)))
                fd = (8 * d[0] - d[1]) / (12 * eps)
            worst = max(worst, abs(g64[f][i] - fd) / max(abs(fd), 1e-2 * scale))
        fd_err[f] = worst
        log(f"FD d/d{f}: worst relative error {worst:.2e} over {len(idx)} particles")
    res_json["fd_particles"] = int(len(idx))
    res_json["fd_worst_rel_err"] = fd_err
    res_json["fd_seconds"] = time.perf_counter() - a0

    # 3. mixed against FP64
    res_json["mixed_vs_fp64_rel_err"] = {f: float(np.max(np.abs(g32[f] - g64[f])) / max(np.abs(g64[f]).max(), 1e-300))
                                         for f in hip.GRAD_FIELDS}
    log(f"mixed vs fp64: {res_json['mixed_vs_fp64_rel_err']}")

    # 4. mass recovery
    # start from the perturbed masses, fit the density of the true ones
    truth = forward(t)[0] / V
    m = tp["m"].clone().requires_grad_(True)
    opt = torch.optim.Adam([m], lr=0.01 * float(m_true.mean()))
    hist = []
    a0 = time.perf_counter()
    for step in range(args.steps):
        opt.zero_grad()
        rho = sph_to_grid(m, t["x"], t["y"], t["z"], t["h"], t["vx"], t["vy"], t["vz"], ctx, dom, res)[0] / V
        loss = torch.sum((rho - truth) ** 2)
        loss.backward()
        opt.step()
        if step % 50 == 0 or step == args.steps - 1:
            err = float(torch.sqrt(torch.mean(((m.detach() - t["m"]) / t["m"]) ** 2)))
            hist.append({"step": step, "loss": float(loss.detach()), "rms_rel_mass_err": err})
            log(f"fit step {step}: loss {float(loss.detach()):.3e}, rms relative mass error {err:.4f}")
    torch.cuda.synchronize()
    res_json["fit"] = {"steps": args.steps, "seconds": time.perf_counter() - a0, "history": hist}
    print(json.dumps(res_json), flush=True)


if __name__ == "__main__":
    main()
