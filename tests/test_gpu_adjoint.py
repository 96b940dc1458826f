"""GPU checks of kiwi_hip_sph_to_grid_adjoint (no reference counterpart): the gradients of a loss of the
per-cell sums with respect to every particle's m, x, y, z, h, vx, vy, vz.

Pinned by:
  - central finite differences (fourth order for x, y, z, h) of the FP64 forward (kiwi_hip_sph_to_grid
    itself) on random particles of the C1 workload, relative 1e-5;
  - exact invariants of the forward: on a domain padded by the largest support every particle's cells
    hold exactly its mass, so for L = total mass dL/dm = 1 and dL/dx = dL/dh = 0, and for L = total
    x-momentum dL/dvx = m and dL/dm = vx;
  - the single-precision lattice (KIWI_MIXED) against FP64: 2e-4 of the largest value for the mass and
    velocity gradients, 3e-2 for the position and h gradients (FP32 rounding, see the test);
  - tests/test_lattice_math.py checks the lattice gradient itself against quadrature of the face flux,
    and the generic kernels' lattice gradient against central differences of their lattice value;
  - the generic path (M4 forced generic, M5, M6, WC2, WC4, WC6): finite differences of its own FP64 forward
    on random particles (1e-5), the total-mass invariants on a padded domain, and mixed against FP64 for
    the kernels the forward runs in FP32 (WC2).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = ("x", "y", "z", "vx", "vy", "vz", "m", "h")


def say(msg):
    print(msg, flush=True)


def c1():
    parts, dom, res = workloads.uniform_cube()
    return {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in COLS}, dom, res


class Loss:
    """L = sum_c w_c (mass_c / V - t_c)^2 + sum_c u_c . p_c  and its gradient with respect to the sums.  The
    momentum weights are scaled so that both terms give gradients of comparable size."""

    def __init__(self, ctx, parts, dom, res, seed=3, kernel=None):
        rng = np.random.default_rng(seed)
        self.ctx, self.dom, self.res, self.kernel = ctx, dom, res, kernel
        self.V = np.prod((dom[1::2] - dom[0::2]) / np.array(res))
        pert = dict(parts)
        pert["m"] = parts["m"] * (1 + 0.05 * rng.standard_normal(len(parts["m"])))
        self.t = ctx.sph_to_grid(pert, dom, res, kernel=kernel)[0] / self.V
        nc = int(np.prod(res))
        self.w = rng.uniform(0.5, 1.5, nc)
        g_mass = 2 * np.abs(self.t).max() * 0.05 / self.V
        vmax = max(np.abs(parts[c]).max() for c in ("vx", "vy", "vz"))
        self.u = [rng.standard_normal(nc) * g_mass / vmax for _ in range(3)]

    def terms(self, parts):
        """per-cell terms of L (a finite difference of L sums the per-cell differences, so cells the
        perturbed particle does not reach cancel exactly)"""
        mass, px, py, pz, _ = self.ctx.sph_to_grid(parts, self.dom, self.res, kernel=self.kernel)
        r = mass / self.V - self.t
        return self.w * r * r + self.u[0] * px + self.u[1] * py + self.u[2] * pz

    def cell_grads(self, parts):
        mass = self.ctx.sph_to_grid(parts, self.dom, self.res, kernel=self.kernel)[0]
        return (2 * self.w * (mass / self.V - self.t) / self.V, self.u[0], self.u[1], self.u[2])


def test_adjoint_matches_finite_differences(hip_ctx):
    say("case: C1 2000 particles -> 32^3, FP64 adjoint vs central differences on 40 particles")
    parts, dom, res = c1()
    L = Loss(hip_ctx, parts, dom, res)
    g, st = hip_ctx.sph_to_grid_adjoint(parts, dom, res, L.cell_grads(parts))
    assert st.skipped_outside == 0 and st.cells > 0
    rng = np.random.default_rng(7)
    idx = rng.choice(len(parts["x"]), 40, replace=False)
    for f in ("m", "x", "y", "z", "h", "vx"):
        scale = np.sqrt(np.mean(g[f] ** 2))
        worst = 0.0
        for i in idx:
            base = parts[f][i]

            def diff(step):
                lp, lm = dict(parts), dict(parts)
                lp[f] = parts[f].copy()
                lm[f] = parts[f].copy()
                lp[f][i] = base + step
                lm[f][i] = base - step
                return np.sum(L.terms(lp) - L.terms(lm))

            if f in ("m", "vx"):  # L is quadratic in m and linear in v: central differences are exact
                eps = 1e-6 * abs(base) if f == "m" else 1e-4
                fd = diff(eps) / (2 * eps)
            else:  # fourth-order central difference
                eps = 1e-3 * parts["h"][i]
                fd = (8 * diff(eps) - diff(2 * eps)) / (12 * eps)
            err = abs(g[f][i] - fd) / max(abs(fd), 1e-2 * scale)
            worst = max(worst, err)
        say(f"  d/d{f}: worst relative error {worst:.2e} (rms gradient {scale:.3e})")
        assert worst < 1e-5, f


def test_adjoint_invariants(hip_ctx):
    say("case: C1, L = total mass and L = total x-momentum")
    parts, dom, res = c1()
    nc = int(np.prod(res))
    one, zero = np.ones(nc), np.zeros(nc)
    g, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, (one, zero, zero, zero))
    np.testing.assert_allclose(g["m"], 1.0, rtol=0, atol=1e-12)
    m_over_h = parts["m"] / parts["h"]
    for f in ("x", "y", "z", "h"):
        assert np.max(np.abs(g[f]) / m_over_h) < 1e-11, f
    for f in ("vx", "vy", "vz"):
        assert np.max(np.abs(g[f])) == 0.0
    g, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, (zero, one, zero, zero))
    np.testing.assert_allclose(g["vx"], parts["m"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["m"], parts["vx"], rtol=0, atol=1e-12 * np.abs(parts["vx"]).max())
    for f in ("x", "y", "z", "h"):
        assert np.max(np.abs(g[f]) / (m_over_h * np.abs(parts["vx"]).max())) < 1e-11, f


def test_adjoint_mixed_close_to_fp64(hip_ctx):
    say("case: C1, KIWI_MIXED adjoint vs FP64 adjoint")
    parts, dom, res = c1()
    L = Loss(hip_ctx, parts, dom, res)
    gc = L.cell_grads(parts)
    g64, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc)
    g32, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc, precision=hip.MIXED)
    errs = {}
    for f in hip.GRAD_FIELDS:
        scale = np.abs(g64[f]).max()
        errs[f] = np.max(np.abs(g32[f] - g64[f])) / scale
        say(f"  d/d{f}: max |mixed - fp64| / max|fp64| = {errs[f]:.2e}")
    # FP32 lattice values and gradients carry ~1e-6 absolute rounding per lattice point (the closed form
    # cancels internally at radii ~h-2h).  Mass and velocity gradients sum the values themselves and
    # stay within ~4e-5 of their maximum; position and h gradients are small differences of lattice
    # gradients and come to ~1e-2.  The FP64 adjoint is the accurate mode (finite differences above).
    for f in ("m", "vx", "vy", "vz"):
        assert errs[f] < 2e-4, (f, errs[f])
    for f in ("x", "y", "z", "h"):
        assert errs[f] < 3e-2, (f, errs[f])


def test_adjoint_deterministic_and_device_outputs(hip_ctx):
    import torch
    say("case: C1 adjoint twice (bit-identical), host vs device buffers")
    parts, dom, res = c1()
    L = Loss(hip_ctx, parts, dom, res)
    gc = L.cell_grads(parts)
    a, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc)
    b, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc)
    for f in hip.GRAD_FIELDS:
        assert np.array_equal(a[f], b[f]), f
    dev = torch.device("cuda:0")
    dparts = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    dgc = tuple(torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in gc)
    out = {f: torch.zeros(len(parts["x"]), dtype=torch.float64, device=dev) for f in hip.GRAD_FIELDS}
    hip_ctx.sph_to_grid_adjoint(dparts, dom, res, dgc, out=out)
    torch.cuda.synchronize()
    for f in hip.GRAD_FIELDS:
        assert np.array_equal(a[f], out[f].cpu().numpy()), f


def test_autograd_function_and_edge_cases(hip_ctx):
    import torch
    from kiwi_amd.autograd import sph_to_grid
    say("case: torch.autograd.Function backward == direct adjoint; skipped and h <= 0 particles; kernels")
    parts, dom, res = c1()
    parts["h"][:3] = 0.0
    parts["x"][3:6] = dom[1] + 5.0  # support box misses the grid
    L = Loss(hip_ctx, parts, dom, res)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in parts.items()}
    mass, px, py, pz = sph_to_grid(t["m"], t["x"], t["y"], t["z"], t["h"], t["vx"], t["vy"], t["vz"], hip_ctx, dom,
                                   res)
    tt = torch.from_numpy(L.t).to(dev)
    w = torch.from_numpy(L.w).to(dev)
    u = [torch.from_numpy(v).to(dev) for v in L.u]
    loss = torch.sum(w * (mass / L.V - tt) ** 2) + torch.sum(u[0] * px) + torch.sum(u[1] * py) + torch.sum(u[2] * pz)
    loss.backward()
    g, st = hip_ctx.sph_to_grid_adjoint(parts, dom, res, L.cell_grads(parts))
    assert st.skipped_outside == 3
    for f in hip.GRAD_FIELDS:
        np.testing.assert_allclose(t[f].grad.cpu().numpy(), g[f], rtol=1e-12, atol=1e-14 * np.abs(g[f]).max())
        assert np.all(g[f][:6] == 0.0)
    with pytest.raises(hip.KiwiError):  # the tabulated integral's table lookup is not differentiated
        hip_ctx.sph_to_grid_adjoint(parts, dom, res, L.cell_grads(parts), kernel=hip.kernel_spec("M4", tabulated=True))


def test_mass_recovery_by_gradient_descent(hip_ctx):
    """A small version of the fit: recover 5%-perturbed particle masses from the target density by
    gradient descent (Adam) on L = sum (rho - rho_target)^2 through the adjoint."""
    import torch
    from kiwi_amd.autograd import sph_to_grid
    say("case: 20000 log-normal particles -> 48^3, 60 gradient steps on the masses")
    parts, dom, res = workloads.lognormal_field(n_particles=20000, n=48, seed=5)
    parts = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in COLS}
    rng = np.random.default_rng(11)
    m_true = parts["m"].copy()
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    V = float(np.prod((dom[1::2] - dom[0::2]) / np.array(res)))
    target = sph_to_grid(t["m"], t["x"], t["y"], t["z"], t["h"], t["vx"], t["vy"], t["vz"], hip_ctx, dom, res)[0] / V
    m = torch.from_numpy(m_true * (1 + 0.05 * rng.standard_normal(len(m_true)))).to(dev).requires_grad_(True)
    opt = torch.optim.Adam([m], lr=0.01 * float(m_true.mean()))
    first = None
    for step in range(60):
        opt.zero_grad()
        rho = sph_to_grid(m, t["x"], t["y"], t["z"], t["h"], t["vx"], t["vy"], t["vz"], hip_ctx, dom, res)[0] / V
        loss = torch.sum((rho - target) ** 2)
        loss.backward()
        opt.step()
        first = first if first is not None else float(loss.detach())
    final = float(loss.detach())
    err0 = 0.05
    err = float(torch.sqrt(torch.mean(((m.detach().cpu() - torch.from_numpy(m_true)) / torch.from_numpy(m_true)) ** 2)))
    say(f"  loss {first:.3e} -> {final:.3e}; rms relative mass error 0.05 -> {err:.4f}")
    # the loss falls by > 100x; masses that the density field cannot separate (particles sharing
    # cells) stay partly unresolved, so the mass error falls by less
    assert final < 0.01 * first
    assert err < 0.7 * err0


GENERIC = [("WC2", False), ("M5", False), ("M6", False), ("WC4", False), ("WC6", False), ("M4", True)]


def small_cloud(seed=21):
    """600 particles of varying h in a padded box, 24^3 cells"""
    rng = np.random.default_rng(seed)
    n = 600
    parts = {"x": rng.uniform(-1, 1, n), "y": rng.uniform(-1, 1, n), "z": rng.uniform(-1, 1, n),
             "vx": rng.standard_normal(n), "vy": rng.standard_normal(n), "vz": rng.standard_normal(n),
             "m": rng.uniform(0.5, 1.5, n) * 1e-3, "h": rng.uniform(0.08, 0.3, n)}
    parts = {k: np.ascontiguousarray(v) for k, v in parts.items()}
    dom = workloads.padded_domain(parts, support_in_h=3.0)  # the widest default support (M6: 3h)
    return parts, dom, (24, 24, 24)


@pytest.mark.parametrize("name,generic", GENERIC)
def test_generic_adjoint_matches_finite_differences(hip_ctx, name, generic):
    say(f"case: generic {name}{' (forced generic)' if generic else ''}, 600 particles -> 24^3, FP64 adjoint vs "
        f"central differences on 12 particles")
    parts, dom, res = small_cloud()
    kern = hip.kernel_spec(name, generic_path=generic)
    L = Loss(hip_ctx, parts, dom, res, kernel=kern)
    g, st = hip_ctx.sph_to_grid_adjoint(parts, dom, res, L.cell_grads(parts), kernel=kern)
    assert st.skipped_outside == 0 and st.cells > 0
    rng = np.random.default_rng(8)
    idx = rng.choice(len(parts["x"]), 12, replace=False)
    for f in ("m", "x", "y", "z", "h", "vy"):
        scale = np.sqrt(np.mean(g[f] ** 2))
        worst = 0.0
        for i in idx:
            base = parts[f][i]

            def diff(step):
                lp, lm = dict(parts), dict(parts)
                lp[f] = parts[f].copy()
                lm[f] = parts[f].copy()
                lp[f][i] = base + step
                lm[f][i] = base - step
                return np.sum(L.terms(lp) - L.terms(lm))

            if f in ("m", "vy"):
                eps = 1e-6 * abs(base) if f == "m" else 1e-4
                fd = diff(eps) / (2 * eps)
            else:
                eps = 1e-3 * parts["h"][i]
                fd = (8 * diff(eps) - diff(2 * eps)) / (12 * eps)
            worst = max(worst, abs(g[f][i] - fd) / max(abs(fd), 1e-2 * scale))
        say(f"  d/d{f}: worst relative error {worst:.2e}")
        assert worst < 1e-5, (name, f, worst)
    # the total-mass invariants on the padded domain: dM/dm = 1, dM/dx = dM/dh = 0
    nc = int(np.prod(res))
    one, zero = np.ones(nc), np.zeros(nc)
    g, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, (one, zero, zero, zero), kernel=kern)
    np.testing.assert_allclose(g["m"], 1.0, rtol=0, atol=1e-11)
    m_over_h = parts["m"] / parts["h"]
    for f in ("x", "y", "z", "h"):
        assert np.max(np.abs(g[f]) / m_over_h) < 1e-10, (name, f)


def test_generic_adjoint_mixed_close_to_fp64(hip_ctx):
    say("case: WC2 (FP32 lattice in mixed precision), mixed adjoint vs FP64 adjoint")
    parts, dom, res = small_cloud(seed=22)
    kern = hip.kernel_spec("WC2")
    L = Loss(hip_ctx, parts, dom, res, kernel=kern)
    gc = L.cell_grads(parts)
    g64, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc, kernel=kern)
    g32, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc, kernel=kern, precision=hip.MIXED)
    for f in hip.GRAD_FIELDS:
        e = np.max(np.abs(g32[f] - g64[f])) / np.abs(g64[f]).max()
        say(f"  d/d{f}: max |mixed - fp64| / max|fp64| = {e:.2e}")
        assert e < (2e-4 if f in ("m", "vx", "vy", "vz") else 3e-2), (f, e)
    # the adjoint has no FP64 fallback, so M5 (FP32 in the forward only beside it) runs in FP64
    k5 = hip.kernel_spec("M5")
    a, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc, kernel=k5)
    b, _ = hip_ctx.sph_to_grid_adjoint(parts, dom, res, gc, kernel=k5, precision=hip.MIXED)
    for f in hip.GRAD_FIELDS:
        assert np.array_equal(a[f], b[f])
