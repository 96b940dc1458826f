"""Synthetic SPH particle sets for the parity tests and the benchmark (fixed seeds, numpy only).

uniform_cube      2,000 particles uniform in [-1,1]^3 pc, m = 1e-3 Msun, h = 1.2 (m/rho_32)^(1/3) where
                  rho_32 = 32 m / (4/3 pi r_32^3) from the distance r_32 to the 32nd nearest neighbour,
                  velocities N(0, 1 km/s) per axis; domain padded by max 2h, Cartesian grid 32^3.
lognormal_field   N particles drawn from a log-normal density field: a periodic Gaussian random field
                  on an n^3 lattice with P(k) ~ k^-3 (k != 0), scaled to sigma_ln_rho, ln rho = delta -
                  sigma^2/2; particles are sampled multinomially over lattice cells in proportion to rho,
                  uniform inside the cell; equal masses; h = 1.2 (1/(nbar rho))^(1/3) cells clipped to
                  [h_min, h_max]; velocities = the finite-difference gradient of delta at the particle's
                  cell, scaled to v_rms per axis, plus N(0, noise) per axis.  Domain [0, n]^3 with unit
                  cells.
lognormal_box     the same for a box of nx x ny x nz unit cells: the field is generated on the enclosing
                  cube and the particles are sampled in the box.
cutout_batch      a batch of independent cutouts: particle counts log-uniform in [n_lo, n_hi], the
                  largest grid axis log-uniform in [g_lo, g_hi] cells and the others 1 to 4 times shorter,
                  kernels alternating between the cubic spline (mixed precision) and Wendland C2 (generic
                  path, FP64).
"""
import numpy as np

COLS = ("x", "y", "z", "vx", "vy", "vz", "m", "h")


def uniform_cube(n=2000, seed=1, nn=32):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, size=(n, 3))
    m = np.full(n, 1e-3)
    # distance to the nn-th nearest neighbour (brute force; n is small)
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    r_nn = np.sqrt(np.partition(d2, nn, axis=1)[:, nn])  # column 0 is the particle itself
    rho = nn * m / (4.0 / 3.0 * np.pi * r_nn ** 3)
    h = 1.2 * (m / rho) ** (1.0 / 3.0)
    vel = rng.normal(0.0, 1.0, size=(n, 3))
    parts = dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2], m=m, h=h)
    parts["rho"] = rho
    pad = 2.0 * h.max()
    domain = np.array([-1 - pad, 1 + pad, -1 - pad, 1 + pad, -1 - pad, 1 + pad])
    return parts, domain, (32, 32, 32)


def gaussian_field(n, sigma, seed, slope=-3.0):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n)).astype(np.float64)
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)  # sqrt(P), P ~ k^slope
    amp[0, 0, 0] = 0.0
    delta = np.fft.irfftn(fk * amp, s=(n, n, n), axes=(0, 1, 2))
    delta *= sigma / delta.std()
    return delta


def lognormal_field(n_particles=1_000_000, n=256, sigma=1.5, seed=2, h_min=0.5, h_max=50.0, v_rms=5.0,
                    noise=2.0, hfact=1.2):
    rng = np.random.default_rng(seed + 1000)
    delta = gaussian_field(n, sigma, seed)
    rho = np.exp(delta - 0.5 * sigma * sigma)
    w = rho.reshape(-1)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cell = np.searchsorted(cdf, rng.uniform(0.0, 1.0, n_particles), side="right")
    cell = np.minimum(cell, n ** 3 - 1)
    ix, rem = np.divmod(cell, n * n)
    iy, iz = np.divmod(rem, n)
    pos = np.stack([ix, iy, iz], 1) + rng.uniform(0.0, 1.0, size=(n_particles, 3))
    nbar = n_particles / float(n ** 3)
    h = hfact * (1.0 / (nbar * w[cell])) ** (1.0 / 3.0)
    h = np.clip(h, h_min, h_max)
    vel = np.empty((n_particles, 3))
    for a in range(3):
        g = 0.5 * (np.roll(delta, -1, axis=a) - np.roll(delta, 1, axis=a))
        g = g.reshape(-1)
        vel[:, a] = g[cell] * (v_rms / g.std())
    vel += rng.normal(0.0, noise, size=vel.shape)
    m = np.full(n_particles, 1.0 / n_particles)
    parts = dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), vx=vel[:, 0].copy(),
                 vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), m=m, h=h)
    domain = np.array([0.0, float(n), 0.0, float(n), 0.0, float(n)])
    return parts, domain, (n, n, n)


def lognormal_box(n_particles, dims, sigma=1.5, seed=2, h_min=0.5, h_max=50.0, v_rms=5.0, noise=2.0, hfact=1.2):
    nx, ny, nz = (int(d) for d in dims)
    n = max(nx, ny, nz)
    rng = np.random.default_rng(seed + 1000)
    delta = gaussian_field(n, sigma, seed)[:nx, :ny, :nz]
    rho = np.exp(delta - 0.5 * sigma * sigma)
    w = rho.reshape(-1)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cell = np.minimum(np.searchsorted(cdf, rng.uniform(0.0, 1.0, n_particles), side="right"), w.size - 1)
    ix, rem = np.divmod(cell, ny * nz)
    iy, iz = np.divmod(rem, nz)
    pos = np.stack([ix, iy, iz], 1) + rng.uniform(0.0, 1.0, size=(n_particles, 3))
    nbar = n_particles / float(w.size)
    h = np.clip(hfact * (1.0 / (nbar * w[cell])) ** (1.0 / 3.0), h_min, h_max)
    vel = np.empty((n_particles, 3))
    for a in range(3):
        g = (0.5 * (np.roll(delta, -1, axis=a) - np.roll(delta, 1, axis=a))).reshape(-1)
        vel[:, a] = g[cell] * (v_rms / g.std())
    vel += rng.normal(0.0, noise, size=vel.shape)
    m = np.full(n_particles, 1.0 / n_particles)
    parts = dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), vx=vel[:, 0].copy(),
                 vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), m=m, h=h)
    domain = np.array([0.0, float(nx), 0.0, float(ny), 0.0, float(nz)])
    return parts, domain, (nx, ny, nz)


def cutout_batch(n_problems=64, seed=3, n_lo=1_000, n_hi=200_000, g_lo=32, g_hi=192):
    """A list of (parts, domain, res, kernel_name) problems (see the module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n_problems):
        npart = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        big = int(round(np.exp(rng.uniform(np.log(g_lo), np.log(g_hi)))))
        dims = [big] + [max(8, int(round(big / rng.uniform(1.0, 4.0)))) for _ in range(2)]
        dims = [dims[i] for i in rng.permutation(3)]
        parts, dom, res = lognormal_box(npart, dims, seed=seed * 1000 + q)
        out.append((parts, dom, res, "M4" if q % 2 == 0 else "WC2"))
    return out


def padded_domain(parts, support_in_h=2.0):
    """The conservation variant: the particles' bounding box padded by the maximum support per side."""
    pad = support_in_h * float(np.max(parts["h"]))
    return np.array([parts["x"].min() - pad, parts["x"].max() + pad, parts["y"].min() - pad,
                     parts["y"].max() + pad, parts["z"].min() - pad, parts["z"].max() + pad])


def subsample(parts, every):
    return {k: np.ascontiguousarray(v[::every]) for k, v in parts.items()}


def write_plain_ascii(parts, path):
    """The reference's plain_ascii snapshot: 9 columns x y z vx vy vz m h rho, one particle per line."""
    rho = parts.get("rho", np.zeros_like(parts["x"]))
    arr = np.stack([parts[c] for c in COLS] + [rho], 1)
    np.savetxt(path, arr, fmt="%.17g")
