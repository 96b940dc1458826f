"""Mixed precision's FP64 fallback (kiwi_hip_set_fp64_fallback): particles whose support radius spans at
least the threshold in cells compute their lattice values, line and plane constants and cell sums in FP64;
their per-pair sums are stored as float like the FP32 items' and accumulated in FP64.

Checks:
  * every particle on the fallback: the mixed call equals the FP64 call up to the float rounding of each
    per-pair sum (2^-24 of the pair's sum), cell by cell;
  * isolated particles with h = 10-30 cells on a grid holding their whole support: the cell sums add up to
    the particle's mass within 1e-5 with the fallback (the FP32 lattice alone loses ~1e-3 there);
  * a split workload (both classes present): total mass within 1e-6 of FP64, every field within 1e-5 of
    its maximum, repeat calls bit-identical, a batch equal to its single calls bit for bit;
  * the threshold's round trip and its argument check.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip, workloads  # noqa: E402

pytestmark = pytest.mark.gpu

COLS = workloads.COLS


def say(msg):
    print(msg, flush=True)


def rel_max(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


@pytest.fixture
def fb_ctx(hip_ctx):
    """The shared context with its threshold restored after each test."""
    saved = hip_ctx.get_fp64_fallback()
    yield hip_ctx
    hip_ctx.set_fp64_fallback(saved)


def test_threshold_round_trip(fb_ctx):
    say("case: fp64 fallback threshold set/get, NaN rejected")
    assert fb_ctx.get_fp64_fallback() == hip.FP64_FALLBACK_CELLS
    fb_ctx.set_fp64_fallback(3.5)
    assert fb_ctx.get_fp64_fallback() == 3.5
    with pytest.raises(hip.KiwiError):
        fb_ctx.set_fp64_fallback(float("nan"))
    assert fb_ctx.get_fp64_fallback() == 3.5


def test_all_particles_on_fallback_equal_fp64_to_float_rounding(fb_ctx):
    say("case: 20,000 log-normal particles -> 48^3, every particle on the FP64 fallback vs FP64 mode")
    parts, _, _ = workloads.lognormal_field(n_particles=20000, n=32, seed=7, h_max=6.0)
    dom = workloads.padded_domain(parts)
    res = (48, 48, 48)
    f = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.FP64)
    fb_ctx.set_fp64_fallback(1e-9)
    m = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    st = m[4]
    assert st.fp64_fallback_blocks == st.particle_blocks > 0
    assert st.pairs_sphere == f[4].pairs_sphere
    # each cell is a sum of pair sums Msum m (v) >= 0 up to FP64 round-off; each pair sum was rounded
    # to float once: |mixed - fp64| <= 2^-24 sum |pair term|, and sum |pair term| <= the cell's mass
    # (velocity fields: the cell's sum of Msum m |v|, bounded here by mass x max |v|)
    vmax = max(np.max(np.abs(parts[c])) for c in ("vx", "vy", "vz"))
    tol = 2.0 ** -24 * 1.01
    assert np.all(np.abs(m[0] - f[0]) <= tol * f[0] + 1e-14 * f[0].max())
    for k in (1, 2, 3):
        assert np.all(np.abs(m[k] - f[k]) <= tol * vmax * f[0] + 1e-14 * np.abs(f[k]).max())
    say(f"  mass max err / max {rel_max(m[0], f[0]):.2e}")


@pytest.mark.parametrize("h_cells", [10.0, 20.0, 30.0])
def test_isolated_large_h_mass(fb_ctx, h_cells):
    say(f"case: one particle with h = {h_cells} cells, grid holding its support, mixed with and without "
        f"the fallback")
    rng = np.random.default_rng(int(h_cells))
    n = int(4 * h_cells) + 6
    c = n / 2.0 + rng.uniform(-0.5, 0.5, 3)
    parts = {"x": np.array([c[0]]), "y": np.array([c[1]]), "z": np.array([c[2]]), "vx": np.array([1.0]),
             "vy": np.array([-2.0]), "vz": np.array([0.5]), "m": np.array([1.0]), "h": np.array([h_cells])}
    dom = np.array([0.0, n, 0.0, n, 0.0, n])
    res = (n, n, n)
    f = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.FP64)
    on = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    fb_ctx.set_fp64_fallback(0.0)
    off = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    e64 = abs(f[0].sum() - 1.0)
    e_on = abs(on[0].sum() - 1.0)
    e_off = abs(off[0].sum() - 1.0)
    l1_on = float(np.abs(on[0] - f[0]).sum())
    l1_off = float(np.abs(off[0] - f[0]).sum())
    say(f"  mass error: FP64 {e64:.1e}, mixed+fallback {e_on:.1e} (L1 vs FP64 {l1_on:.1e}), "
        f"FP32 lattice alone {e_off:.1e} (L1 {l1_off:.1e})")
    assert on[4].fp64_fallback_blocks == on[4].particle_blocks
    assert off[4].fp64_fallback_blocks == 0
    assert e64 <= 1e-10  # FP64 lattice sums over ~2e6 cells
    assert e_on <= 1e-5 and l1_on <= 1e-5
    assert l1_on < l1_off


def test_split_workload(fb_ctx):
    say("case: 50,000 log-normal particles -> 64^3 with the threshold inside the h range: both classes")
    parts, dom, res = workloads.lognormal_field(n_particles=50000, n=64, seed=5)
    hc = 2.0 * parts["h"]  # support radius in cells (unit cells)
    thr = float(np.quantile(hc, 0.7))
    fb_ctx.set_fp64_fallback(thr)
    f = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.FP64)
    m = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    st = m[4]
    share = st.fp64_fallback_blocks / st.particle_blocks
    err = abs(m[0].sum() - f[0].sum()) / f[0].sum()
    l1 = float(np.abs(m[0] - f[0]).sum() / f[0].sum())
    fields = [rel_max(x, y) for x, y in zip(m[:4], f[:4])]
    say(f"  threshold {thr:.2f} cells: {share:.1%} of the items in FP64; total mass rel err {err:.2e}, "
        f"mass L1 {l1:.2e}, per-field max err / max " + ", ".join(f"{e:.1e}" for e in fields))
    assert 0.0 < share < 1.0
    assert err <= 1e-6 and max(fields) <= 1e-5
    again = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    for a, b in zip(again[:4], m[:4]):
        assert np.array_equal(a, b)
    # the fallback lowers the mass-weighted error against the FP32 lattice alone
    fb_ctx.set_fp64_fallback(0.0)
    o = fb_ctx.sph_to_grid(parts, dom, res, precision=hip.MIXED)
    l1_off = float(np.abs(o[0] - f[0]).sum() / f[0].sum())
    say(f"  mass L1 without the fallback {l1_off:.2e}")
    assert l1 < l1_off
    # a batch of two problems on this path equals its single calls bit for bit
    fb_ctx.set_fp64_fallback(thr)
    half = {k: np.ascontiguousarray(v[::2]) for k, v in parts.items()}
    probs = [{"parts": parts, "domain": dom, "res": res, "precision": hip.MIXED},
             {"parts": half, "domain": dom, "res": (40, 40, 40), "precision": hip.MIXED}]
    batch = fb_ctx.sph_to_grid_batch(probs)
    single = [fb_ctx.sph_to_grid(p["parts"], p["domain"], p["res"], precision=hip.MIXED) for p in probs]
    for b, s1 in zip(batch, single):
        for u, v in zip(b[:4], s1[:4]):
            assert np.array_equal(u, v)
