"""GPU parity of kiwi_hip_voronoi_centroids, the Lloyd step (lloyd_step, src/point_set.cpp:275-302),
against the reference's own voro++ (voronoicell_base::centroid and volume through the test-side shim
oracle/_ref/libkiwi_voro.so) on random, clustered and lattice-aligned site sets.

Tolerance: centroids within 1e-12 of the domain's extent and volumes within 1e-12 of the mean cell
volume (the two decompositions into tetrahedra are equal in exact arithmetic; in double they differ by
rounding).  The volumes add up to voro++'s total within 1e-12 of the domain (and fill the domain
within 1e-10, voro++'s own tiling), and a second tessellation from the GPU's centroids has the same
cells as one from voro++'s.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.VORO_SO):
        pytest.fail(f"voro++ shim not built at {O.VORO_SO}: these tests tessellate their inputs with it")
    return O


def tessellation(sites, dom):
    O = oracle()
    g = O.tessellate(np.arange(1, len(sites) + 1), sites, dom, np.zeros((0, 3)))
    cells = {k: np.ascontiguousarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    cells["vert_xyz"] = np.ascontiguousarray(g["vert_xyz"].reshape(-1))
    cells["domain"] = dom
    return g, cells


def sites_of(kind, n, rng, dom):
    lo, hi = dom[0::2], dom[1::2]
    if kind == "uniform":
        return rng.uniform(lo, hi, size=(n, 3))
    if kind == "clustered":
        c = rng.uniform(lo + 0.2 * (hi - lo), hi - 0.2 * (hi - lo), size=(5, 3))
        s = c[rng.integers(5, size=n)] + rng.normal(scale=0.05, size=(n, 3)) * (hi - lo)
        return np.clip(s, lo + 1e-9, hi - 1e-9)
    k = int(round(n ** (1 / 3)))  # a lattice of cell centres: every vertex is shared by 8 cells
    ax = [lo[a] + (hi[a] - lo[a]) * (np.arange(k) + 0.5) / k for a in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)


@pytest.mark.parametrize("kind,n", [("uniform", 3000), ("clustered", 5000), ("lattice", 1000)])
def test_centroids_match_voro(hip_ctx, kind, n):
    say(f"case: Lloyd centroids, {kind} sites, n={n}")
    rng = np.random.default_rng(7)
    dom = np.array([-1.0, 2.0, 0.0, 1.5, -0.5, 0.5])
    sites = sites_of(kind, n, rng, dom)
    g, cells = tessellation(sites, dom)
    cen, vol = hip_ctx.voronoi_centroids(cells, np.ascontiguousarray(g["site"]))
    ext = float(np.max(dom[1::2] - dom[0::2]))
    dc = float(np.max(np.abs(cen - g["centroid"])))
    mean_vol = float(np.prod(dom[1::2] - dom[0::2])) / len(vol)
    dv = float(np.max(np.abs(vol - g["volume"])))
    say(f"  {len(vol)} cells: centroid max diff / extent {dc / ext:.2e}, volume max diff / mean {dv / mean_vol:.2e}")
    assert dc <= 1e-12 * ext and dv <= 1e-12 * mean_vol
    V = np.prod(dom[1::2] - dom[0::2])
    assert abs(vol.sum() - g["volume"].sum()) <= 1e-12 * V
    assert abs(vol.sum() - V) <= 1e-10 * V  # voro++'s cells tile the box to its own tolerance
    if kind == "lattice":  # the lattice is its own Lloyd fixed point
        assert np.max(np.abs(cen - g["site"])) <= 1e-13 * ext


def test_device_arrays_and_second_tessellation(hip_ctx):
    import torch
    say("case: Lloyd step on device arrays; the next tessellation from the GPU's centroids")
    O = oracle()
    rng = np.random.default_rng(8)
    dom = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    sites = sites_of("clustered", 4000, rng, dom)
    g, cells = tessellation(sites, dom)
    dev = torch.device("cuda:0")
    dcells = {k: torch.from_numpy(v).to(dev) for k, v in cells.items() if k != "domain"}
    dcells["domain"] = dom
    dsites = torch.from_numpy(np.ascontiguousarray(g["site"])).to(dev)
    out = (torch.zeros((len(g["volume"]), 3), dtype=torch.float64, device=dev),
           torch.zeros(len(g["volume"]), dtype=torch.float64, device=dev))
    hip_ctx.voronoi_centroids(dcells, dsites, out=out)
    hc, hv = hip_ctx.voronoi_centroids(cells, np.ascontiguousarray(g["site"]))
    assert np.array_equal(out[0].cpu().numpy(), hc) and np.array_equal(out[1].cpu().numpy(), hv)
    a = O.tessellate(g["cell_id"], hc, dom, np.zeros((0, 3)))
    b = O.tessellate(g["cell_id"], g["centroid"], dom, np.zeros((0, 3)))
    assert np.array_equal(a["cell_id"], b["cell_id"])
    assert np.array_equal(a["face_nbr"], b["face_nbr"]) and np.array_equal(a["face_vert_ptr"], b["face_vert_ptr"])
    np.testing.assert_allclose(a["volume"], b["volume"], rtol=1e-10, atol=0)


def test_malformed_offsets_rejected(hip_ctx):
    say("case: Lloyd step on malformed CSR offsets")
    dom = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    g, cells = tessellation(np.random.default_rng(9).uniform(0, 1, size=(50, 3)), dom)
    bad = dict(cells)
    bad["cell_face_ptr"] = cells["cell_face_ptr"].copy()
    bad["cell_face_ptr"][3] = len(cells["face_nbr"]) + 10
    with pytest.raises(hip.KiwiError) as e:
        hip_ctx.voronoi_centroids(bad, np.ascontiguousarray(g["site"]))
    assert e.value.code == -1
