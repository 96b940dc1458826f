"""GPU parity of kiwi_hip_voronoi_tessellate (the HIP Voronoi tessellation) against voro++, the
reference's vendored copy (oracle/_ref/libkiwi_voro.so), on the container the reference builds
(build_cell_geometry, src/mapping.cpp:222-269; run_tessellation_stage, src/tessellation.cpp:23-61):

  - the same cells in the same order (voro++'s blocks), the same face count and neighbour list per cell
    (exact), every face's vertex cycle equal to voro++'s up to its starting vertex (within 1e-12 of the
    domain extent) on the small cases;
  - volumes and centroids (kiwi_hip_voronoi_centroids over the GPU geometry) within 1e-12 of voro++'s
    (extent^3 and extent), and the volumes summing to the box;
  - the lattice's exact cubes (tools/test_lattice_grid.cpp:95-198);
  - the Voronoi map over the GPU geometry equal to the map over voro++'s (1e-12 / 1e-11 of the field
    maximum, the same visits), also with the home cells from kiwi_hip_voronoi_home_cells, which equal
    voro++'s find_voronoi_cell exactly.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip  # noqa: E402
from geometry_check import assert_same_cells, assert_same_cycles, volumes, voro_tolerance_bounds  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.VORO_SO):
        pytest.fail(f"voro++ shim not built at {O.VORO_SO}")
    return O


def host(cells):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)) for k, v in cells.items()}


def check_case(ctx, pts, dom, cycles=True, min_area=None, tol_bounds=False):
    import torch
    O = oracle()
    dev = torch.device("cuda:0")
    cells, info = ctx.voronoi_tessellate(torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float64)).to(dev), dom)
    g = host(cells)
    r = O.tessellate(np.arange(len(pts)), pts, np.asarray(dom, dtype=np.float64), np.zeros((0, 3)))
    say(f"  cells {len(g['cell_site'])} (voro++ {len(r['cell_id'])}), faces {len(g['face_nbr'])}, dropped "
        f"{info['n_dropped']}, device ms {info['device_ms_prep']:.2f} / {info['device_ms_count']:.2f} / "
        f"{info['device_ms_fill']:.2f}")
    ext = max(abs(v) for v in dom)
    assert_same_cells(g, r, None if min_area is None else min_area * ext ** 2)
    if cycles:
        assert_same_cycles(g, r, 1e-12 * ext)
    nc = len(g["cell_site"])
    assert info["n_dropped"] == len(pts) - nc
    if nc:
        site = torch.from_numpy(np.ascontiguousarray(pts[g["cell_site"]])).to(dev)
        cen, vol = ctx.voronoi_centroids(cells, site, out=(torch.empty((nc, 3), dtype=torch.float64, device=dev),
                                                            torch.empty(nc, dtype=torch.float64, device=dev)))
        cen, vol = cen.cpu().numpy(), vol.cpu().numpy()
        bv, bc = (voro_tolerance_bounds(pts[g["cell_site"]], r) if tol_bounds else (0.0, 0.0))
        dv = np.abs(vol - r["volume"])
        dc = np.max(np.abs(cen - r["centroid"]), axis=1)
        assert np.all(dv <= 1e-12 * ext ** 3 + bv), (np.argmax(dv - bv), dv.max())
        assert np.all(dc <= 1e-12 * ext + bc), (np.argmax(dc - bc), dc.max())
        box = np.prod(np.asarray(dom)[1::2] - np.asarray(dom)[::2])
        assert abs(vol.sum() - box) <= 1e-11 * box
    return cells, g, r


@pytest.fixture(scope="module")
def ctx():
    import torch
    torch.cuda.init()
    c = hip.Context(0)
    yield c
    c.close()


def test_uniform_points(ctx):
    say("case: 2000 uniform sites, unit cube")
    check_case(ctx, np.random.default_rng(1).random((2000, 3)), [0, 1, 0, 1, 0, 1])


def test_clustered_points_offset_box(ctx):
    say("case: 3000 sites with a dense cluster, box [-3,5]x[0,2]x[1,4]")
    rng = np.random.default_rng(2)
    n = 3000
    pts = np.column_stack([rng.uniform(-3, 5, n), rng.uniform(0, 2, n), rng.uniform(1, 4, n)])
    pts[:400] = pts[:400] * 0.05 + [0.0, 1.0, 2.0]
    check_case(ctx, pts, [-3, 5, 0, 2, 1, 4])


def test_lattice_cubes(ctx):
    say("case: 6^3 lattice")
    ax = (np.arange(6) + 0.5) / 6
    pts = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T.copy()
    _, g, _ = check_case(ctx, pts, [0, 1, 0, 1, 0, 1])
    assert np.all(np.diff(g["cell_face_ptr"]) == 6) and np.all(np.diff(g["face_vert_ptr"]) == 4)
    np.testing.assert_allclose(volumes(g, pts[g["cell_site"]]), 1 / 216, rtol=1e-13)


def test_many_faced_cell(ctx):
    say("case: a site ringed by 150 sites (one cell of > 64 faces)")
    k = np.arange(150) + 0.5
    phi = np.arccos(1 - 2 * k / 150)
    th = np.pi * (1 + 5 ** 0.5) * k
    ring = 0.5 + 0.3 * np.column_stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)])
    outer = np.random.default_rng(3).random((149, 3))
    outer = outer[np.linalg.norm(outer - 0.5, axis=1) > 0.45]
    pts = np.vstack([[[0.5, 0.5, 0.5]], ring, outer])
    _, g, _ = check_case(ctx, pts, [0, 1, 0, 1, 0, 1])
    assert np.diff(g["cell_face_ptr"]).max() > 64


def test_dropped_and_boundary_sites(ctx):
    say("case: sites outside the box, on its faces and corners")
    rng = np.random.default_rng(4)
    pts = rng.random((400, 3)) * 2.0
    pts[:20, 0] = 2.0
    pts[20:30, 1] = 0.0
    pts[30:35] = [[2.0, 2.0, 2.0], [0.0, 0.0, 0.0], [2.0, 0.0, 2.0], [0.0, 2.0, 0.0], [1.0, 2.0, 1.0]]
    pts[35:60] = rng.uniform(-1.0, 3.0, (25, 3))
    check_case(ctx, pts, [0, 2, 0, 2, 0, 2])


@pytest.mark.parametrize("n", [0, 1, 2, 5])
def test_few_sites(ctx, n):
    say(f"case: {n} sites")
    check_case(ctx, np.random.default_rng(10 + n).random((n, 3)), [0, 1, 0, 1, 0, 1])


def test_host_input_and_repeat_identical(ctx):
    say("case: host sites vs device sites, two calls")
    import torch
    pts = np.random.default_rng(7).random((1500, 3))
    a, _ = ctx.voronoi_tessellate(pts, [0, 1, 0, 1, 0, 1])
    b, _ = ctx.voronoi_tessellate(torch.from_numpy(pts).cuda(), [0, 1, 0, 1, 0, 1])
    for k in ("cell_face_ptr", "face_vert_ptr", "vert_xyz", "face_nbr", "cell_site"):
        assert torch.equal(a[k], b[k]), k


def test_errors(ctx):
    say("case: non-finite site, bad domain")
    pts = np.random.default_rng(8).random((10, 3))
    pts[3, 1] = np.nan
    with pytest.raises(hip.KiwiError):
        ctx.voronoi_tessellate(pts, [0, 1, 0, 1, 0, 1])
    with pytest.raises(hip.KiwiError):
        ctx.voronoi_tessellate(pts[:2] * 0 + 0.5, [0, 1, 1, 1, 0, 1])


def test_clustered_1e5_sites(ctx):
    """1e5 sites in four Gaussian clusters over a uniform background: the same cells and the same faces
    and neighbours as voro++ (vectorised comparison) but for faces below 1e-12 of the extent squared
    (near-cospherical sites: 2-4 faces of about 1e-14 exist on one side only); volumes and centroids
    within 1e-12 of the extent plus what voro++'s fixed plane tolerance allows a cell (for the smallest
    cells of the clusters, 1e-5 of their volume; measured differences reach 0.11 and 0.28 of the bound)."""
    say("case: 1e5 clustered sites")
    rng = np.random.default_rng(9)
    lo, hi = np.zeros(3), np.array([4.0, 3.0, 2.0])
    centres = rng.uniform(lo + 0.5, hi - 0.5, size=(4, 3))
    pts = np.vstack([centres[rng.integers(4, size=70000)] + rng.normal(scale=0.15, size=(70000, 3)),
                     rng.uniform(lo, hi, size=(30000, 3))])
    pts = np.clip(pts, lo, hi)
    check_case(ctx, pts, [0, 4, 0, 3, 0, 2], cycles=False, min_area=1e-12, tol_bounds=True)


def test_map_over_gpu_geometry_matches_voro(ctx):
    """The Voronoi map (FP64, M4) over the GPU tessellation equals the map over voro++'s: the same visits
    and cell sums within 1e-12 of each field's maximum."""
    say("case: voronoi map over GPU vs voro++ geometry, 1500 sites, 1200 particles")
    import torch
    O = oracle()
    rng = np.random.default_rng(11)
    dom = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    sites = rng.uniform(-1, 1, size=(1500, 3))
    pts = rng.uniform(-1, 1, size=(1200, 3))
    parts = {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "z": pts[:, 2].copy(), "vx": rng.normal(size=1200),
             "vy": rng.normal(size=1200), "vz": rng.normal(size=1200), "m": rng.uniform(0.5, 1.5, 1200),
             "h": 0.17 * rng.uniform(0.5, 2.0, 1200)}
    parts = {k: np.ascontiguousarray(v) for k, v in parts.items()}
    geo = O.tessellate(np.arange(len(sites)), sites, dom, pts)
    home = geo["home"].astype(np.int32)
    vcells = {k: np.ascontiguousarray(geo[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    vcells["vert_xyz"] = np.ascontiguousarray(geo["vert_xyz"].reshape(-1))
    vcells["domain"] = dom
    ref = ctx.voronoi_map(parts, home, vcells)
    gcells, _ = ctx.voronoi_tessellate(torch.from_numpy(sites).cuda(), dom)
    g = host(gcells)
    np.testing.assert_array_equal(g["cell_site"], geo["cell_id"])
    g["vert_xyz"] = np.ascontiguousarray(g["vert_xyz"].reshape(-1))
    got = ctx.voronoi_map(parts, home, g)
    for a, b, name in zip(got[:4], ref[:4], ("mass", "px", "py", "pz")):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), name
    assert (got[4].pairs_visited, got[4].face_evals) == (ref[4].pairs_visited, ref[4].face_evals)


def test_centroids_of_device_cells_without_outputs(ctx):
    """voronoi_centroids on device cells with no output arrays returns device tensors (the library writes
    its outputs where the cells live), equal to the call with explicit outputs."""
    say("case: Lloyd step on device cells, outputs allocated by the wrapper")
    import torch
    pts = np.random.default_rng(12).random((800, 3))
    cells, _ = ctx.voronoi_tessellate(torch.from_numpy(pts).cuda(), [0, 1, 0, 1, 0, 1])
    site = torch.from_numpy(np.ascontiguousarray(pts[cells["cell_site"].cpu().numpy()])).cuda()
    cen, vol = ctx.voronoi_centroids(cells, site)
    assert cen.is_cuda and vol.is_cuda
    out = (torch.empty_like(site), torch.empty(len(site), dtype=torch.float64, device=site.device))
    ctx.voronoi_centroids(cells, site, out=out)
    assert torch.equal(cen, out[0]) and torch.equal(vol, out[1])
    assert abs(float(vol.sum()) - 1.0) <= 1e-12


def test_home_cells_match_voro(ctx):
    """kiwi_hip_voronoi_home_cells against voro++'s find_voronoi_cell resolved to cell indices (the map
    stage's home cells, src/mapping.cpp:434-440): points inside, outside, on the box's faces and on
    sites; host and device inputs give the same."""
    say("case: home cells, 1200 clustered sites, 3000 points")
    import torch
    O = oracle()
    rng = np.random.default_rng(13)
    dom = np.array([-1.0, 2.0, 0.0, 1.0, 0.0, 1.5])
    lo, hi = dom[0::2], dom[1::2]
    sites = np.clip(rng.normal(loc=[0.5, 0.5, 0.7], scale=0.3, size=(1200, 3)), lo, hi)
    pts = rng.uniform(lo, hi, size=(3000, 3))
    pts[:100] = hi + rng.uniform(0.01, 0.5, size=(100, 3))  # outside
    pts[100:150, 0] = hi[0]  # on the x max face: held or not by the rounding of voro++'s block index
    pts[150:200, 1] = lo[1]  # on the y min face
    pts[200:250] = sites[rng.integers(1200, size=50)]  # on sites
    r = O.tessellate(np.arange(len(sites)), sites, dom, pts)
    ctx.voronoi_tessellate(torch.from_numpy(sites).cuda(), dom)
    hd, und = ctx.voronoi_home_cells(torch.from_numpy(pts).cuda())
    hh, unh = ctx.voronoi_home_cells(pts)
    np.testing.assert_array_equal(hd.cpu().numpy(), r["home"])
    np.testing.assert_array_equal(hh, r["home"])
    assert und == unh == int(np.sum(r["home"] < 0)) >= 100


def test_map_on_gpu_tessellation_and_home_cells(ctx):
    """The map stage with no voro++: tessellation, home cells and map on the GPU, against the map over
    voro++'s geometry and home cells: the same visits, sums within 1e-11 of each field's maximum (the
    vertices agree to rounding and a face's cycle may start at another vertex, which rounds the face
    integrals differently; measured 1.1e-12 on the momentum, whose signed sums cancel)."""
    say("case: GPU tessellation + home cells + map vs voro++ inputs, 2000 sites, 1500 particles")
    import torch
    O = oracle()
    rng = np.random.default_rng(14)
    dom = np.array([0.0, 2.0, 0.0, 2.0, 0.0, 2.0])
    sites = rng.uniform(0, 2, size=(2000, 3))
    pts = rng.uniform(0, 2, size=(1500, 3))
    parts = {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "z": pts[:, 2].copy(), "vx": rng.normal(size=1500),
             "vy": rng.normal(size=1500), "vz": rng.normal(size=1500), "m": rng.uniform(0.5, 1.5, 1500),
             "h": 0.2 * rng.uniform(0.5, 2.0, 1500)}
    parts = {k: np.ascontiguousarray(v) for k, v in parts.items()}
    geo = O.tessellate(np.arange(len(sites)), sites, dom, pts)
    vcells = {k: np.ascontiguousarray(geo[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr")}
    vcells["vert_xyz"] = np.ascontiguousarray(geo["vert_xyz"].reshape(-1))
    vcells["domain"] = dom
    ref = ctx.voronoi_map(parts, geo["home"].astype(np.int32), vcells)
    dev = torch.device("cuda:0")
    gcells, _ = ctx.voronoi_tessellate(torch.from_numpy(sites).to(dev), dom)
    dparts = {k: torch.from_numpy(v).to(dev) for k, v in parts.items()}
    home, un = ctx.voronoi_home_cells(torch.from_numpy(pts).to(dev))
    assert un == 0
    nc = len(gcells["cell_site"])
    out = tuple(torch.zeros(nc, dtype=torch.float64, device=dev) for _ in range(4))
    got = ctx.voronoi_map(dparts, home, gcells, out=out)
    for a, b, name in zip(got[:4], ref[:4], ("mass", "px", "py", "pz")):
        a = a.cpu().numpy()
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b)), name
    assert (got[4].pairs_visited, got[4].face_evals) == (ref[4].pairs_visited, ref[4].face_evals)


def test_map_sites_2e5(ctx):
    """The Voronoi map benchmark's sites at full size: 2e5 particles of the log-normal field (256^3 box,
    sigma_ln_rho 1.5; kiwi_amd.workloads.lognormal_field, seed 2), tessellated: the same cells, order,
    faces and neighbours as voro++ (faces below 1e-12 of the extent squared aside), volumes and
    centroids within voro++'s tolerance bounds, the volumes summing to the box, and the particles' home
    cells equal to voro++'s."""
    say("case: 2e5 log-normal map sites, 256^3 box")
    import torch
    from kiwi_amd import workloads
    O = oracle()
    parts, dom, _ = workloads.lognormal_field(n_particles=200_000, n=256, seed=2)
    pts = np.ascontiguousarray(np.stack([parts["x"], parts["y"], parts["z"]], 1).astype(np.float64))
    dom = [float(v) for v in dom]
    _, g, r = check_case(ctx, pts, dom, cycles=False, min_area=1e-12, tol_bounds=True)
    assert len(g["cell_site"]) == 200_000
    rng = np.random.default_rng(15)
    q = rng.uniform(np.asarray(dom[0::2]), np.asarray(dom[1::2]), size=(20000, 3))
    rh = O.tessellate(np.arange(len(pts)), pts, np.asarray(dom), q)["home"]
    home, _ = ctx.voronoi_home_cells(torch.from_numpy(q).cuda())
    np.testing.assert_array_equal(home.cpu().numpy(), rh)


def test_more_cells_than_workspace_slots(ctx):
    """3e5 sites: more cells than the 262,144 workspace slots, so the fill pass computes the cells again
    instead of reading the count pass's; the same cells, faces and neighbours as voro++ and volumes
    summing to the box."""
    say("case: 3e5 uniform sites (fill pass recomputes)")
    pts = np.random.default_rng(16).random((300_000, 3)) * [3.0, 2.0, 1.0]
    _, g, _ = check_case(ctx, pts, [0, 3, 0, 2, 0, 1], cycles=False, min_area=1e-12, tol_bounds=True)
    assert len(g["cell_site"]) == 300_000
