"""GPU checks of kiwi_hip_cartesian_map, the cartesian-map resample (run_cartesian_map_stage,
src/cartesian_map.cpp:84-201): against the reference CLI's cartesian_map.dat (tests/golden/
cartesian_map.npz) and against voro++'s own find_voronoi_cell (the reference's vendored voro++ through
oracle/_ref/libkiwi_voro.so) for every cell centre of random site sets.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import hip  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")


def say(msg):
    print(msg, flush=True)


def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.VORO_SO):
        pytest.fail(f"voro++ shim not built at {O.VORO_SO}: these tests tessellate their inputs with it "
                    f"(make -f oracle/ref.mk where the reference sources exist)")
    return O


def centres(domain, res):
    d = np.asarray(domain, dtype=np.float64)
    ax = [d[2 * a] + (np.arange(res) + 0.5) * ((d[2 * a + 1] - d[2 * a]) / res) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def site_rows(rng, sites, ids):
    n = len(sites)
    return {"x": np.ascontiguousarray(sites[:, 0]), "y": np.ascontiguousarray(sites[:, 1]),
            "z": np.ascontiguousarray(sites[:, 2]), "id": np.ascontiguousarray(ids, dtype=np.int32),
            "density": rng.uniform(0.1, 2.0, n), "vx": rng.normal(size=n), "vy": rng.normal(size=n),
            "vz": rng.normal(size=n)}


def test_matches_reference_cli(hip_ctx):
    say("case: 250 C1 sites, resolution 20, vs the reference CLI's cartesian_map.dat")
    g = np.load(os.path.join(GOLD, "voronoi_map.npz"))
    gc = np.load(os.path.join(GOLD, "cartesian_map.npz"))
    site = g["site"]
    sites = {"x": np.ascontiguousarray(site[:, 0]), "y": np.ascontiguousarray(site[:, 1]),
             "z": np.ascontiguousarray(site[:, 2]), "id": np.ascontiguousarray(g["cell_id"], dtype=np.int32),
             "density": np.ascontiguousarray(g["ref_density"]), "vx": np.ascontiguousarray(g["ref_vx"]),
             "vy": np.ascontiguousarray(g["ref_vy"]), "vz": np.ascontiguousarray(g["ref_vz"])}
    d, vx, vy, vz, cid, st = hip_ctx.cartesian_map(sites, g["domain"], int(gc["res"]))
    assert np.array_equal(cid, gc["cell_id"])
    for a, f in ((d, "density"), (vx, "vx"), (vy, "vy"), (vz, "vz")):
        assert np.array_equal(a, gc[f]), f
    assert st.unresolved == 0 and st.unmapped == 0


def random_sites(rng, n, dom, clustered, outside=0):
    lo, hi = dom[0::2], dom[1::2]
    if clustered:  # three clumps, clipped onto the faces
        c = rng.uniform(lo + 0.2 * (hi - lo), hi - 0.2 * (hi - lo), size=(3, 3))
        p = np.clip(c[rng.integers(3, size=n)] + rng.normal(scale=0.08, size=(n, 3)) * (hi - lo), lo, hi)
    else:
        p = rng.uniform(lo, hi, size=(n, 3))
    if outside:  # sites beyond a face take no part
        k = rng.choice(n, outside, replace=False)
        ax = rng.integers(3, size=outside)
        p[k, ax] = np.where(rng.random(outside) < 0.5, lo[ax] - 0.01, hi[ax] + 0.01)
    return p


@pytest.mark.parametrize("seed,n,res,clustered,outside", [(1, 500, 24, False, 0), (2, 3000, 40, True, 25),
                                                          (3, 40, 17, False, 3), (4, 1, 9, False, 0)])
def test_random_sites_match_oracle(hip_ctx, seed, n, res, clustered, outside):
    from oracle import oracle as O
    say(f"case: {n} {'clustered' if clustered else 'uniform'} sites ({outside} outside), resolution {res}, vs oracle")
    rng = np.random.default_rng(seed)
    dom = np.array([-1.0, 2.0, 0.0, 1.5, -0.5, 0.5])
    sites = random_sites(rng, n, dom, clustered, outside)
    ids = rng.permutation(np.arange(5, 5 + n))
    S = site_rows(rng, sites, ids)
    S["mapped"] = np.ascontiguousarray((rng.random(n) < 0.9).astype(np.int32))
    d, vx, vy, vz, cid, st = hip_ctx.cartesian_map(S, dom, res)
    rows = np.stack([S["density"], S["vx"], S["vy"], S["vz"]], 1)
    want, wid, unres, unm = O.cartesian_map(sites, ids, rows, dom, res, S["mapped"].astype(bool))
    assert np.array_equal(cid, wid)
    for k, (a, f) in enumerate(((d, "density"), (vx, "vx"), (vy, "vy"), (vz, "vz"))):
        assert np.array_equal(a, want[:, k]), f
    assert st.unresolved == unres and st.unmapped == unm
    say(f"  prep {st.device_ms_prep:.3f} ms, query {st.device_ms_query:.3f} ms")


@pytest.mark.parametrize("clustered", [False, True])
def test_tessellation_cells_match_voro(hip_ctx, clustered):
    """With the tessellation's own cells as sites, every centre's id is the cell voro++'s
    find_voronoi_cell reports."""
    O = oracle()
    say(f"case: {'clustered' if clustered else 'uniform'} 2000 points, resolution 30, vs voro++ find_voronoi_cell")
    rng = np.random.default_rng(11 + clustered)
    dom = np.array([0.0, 1.0, -2.0, 0.0, 0.0, 3.0])
    pts = random_sites(rng, 2000, dom, clustered)
    geo = O.tessellate(rng.permutation(np.arange(1, 2001)), pts, dom, centres(dom, 30))
    assert np.all(geo["home"] >= 0)
    S = site_rows(rng, geo["site"], geo["cell_id"])
    cid = hip_ctx.cartesian_map(S, dom, 30)[4]
    assert np.array_equal(cid, geo["cell_id"][geo["home"]])


def test_large_sample_is_nearest(hip_ctx):
    """200,000 clustered sites onto 128^3: every sampled centre's site is a nearest site (brute force over
    all sites for 1,000 sampled centres), and the fields are the chosen site's row."""
    say("case: 200000 clustered sites, resolution 128, sampled brute force")
    rng = np.random.default_rng(21)
    dom = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    n, res = 200_000, 128
    sites = random_sites(rng, n, dom, True)
    S = site_rows(rng, sites, np.arange(1, n + 1))
    d, vx, vy, vz, cid, st = hip_ctx.cartesian_map(S, dom, res)
    q = centres(dom, res)
    pick = rng.choice(res ** 3, 1000, replace=False)
    from oracle import oracle as O
    assert np.array_equal(cid[pick], O.nearest_site(sites, q[pick], chunk=128) + 1)
    assert np.array_equal(d, S["density"][cid - 1]) and np.array_equal(vz, S["vz"][cid - 1])
    assert st.unresolved == 0 and st.unmapped == 0
    say(f"  prep {st.device_ms_prep:.3f} ms, query {st.device_ms_query:.3f} ms")


def test_unmapped_empty_device_and_errors(hip_ctx):
    import torch
    say("case: unmapped cells, no sites, device inputs/outputs, bad arguments")
    rng = np.random.default_rng(9)
    dom = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    sites = rng.uniform(0, 1, size=(200, 3))
    S = site_rows(rng, sites, np.arange(1, 201))
    ref = hip_ctx.cartesian_map(S, dom, 16)
    S["mapped"] = np.ascontiguousarray((np.arange(200) % 3 != 0).astype(np.int32))
    d, vx, vy, vz, cid, st = hip_ctx.cartesian_map(S, dom, 16)
    unm = (cid - 1) % 3 == 0
    assert st.unmapped == int(unm.sum()) > 0
    assert np.all(d[unm] == 0) and np.array_equal(d[~unm], ref[0][~unm]) and np.array_equal(cid, ref[4])
    empty = {k: v[:0] for k, v in S.items()}
    e = hip_ctx.cartesian_map(empty, dom, 8)
    assert e[5].unresolved == 512 and np.all(e[4] == 0) and np.all(e[0] == 0)
    # every site outside the domain: as no sites
    far = {k: v.copy() for k, v in S.items()}
    far["x"] += 2.0
    assert hip_ctx.cartesian_map(far, dom, 8)[5].unresolved == 512
    dev = torch.device("cuda:0")
    Sd = {k: torch.from_numpy(v).to(dev) for k, v in S.items()}
    outd = tuple(torch.zeros(16 ** 3, dtype=torch.float64, device=dev) for _ in range(4)) + (
        torch.zeros(16 ** 3, dtype=torch.int32, device=dev),)
    hip_ctx.cartesian_map(Sd, dom, 16, out=outd)
    torch.cuda.synchronize()
    assert np.array_equal(outd[0].cpu().numpy(), d) and np.array_equal(outd[4].cpu().numpy(), cid)
    with pytest.raises(hip.KiwiError, match="resolution"):
        hip_ctx.cartesian_map(S, dom, 0)
    with pytest.raises(hip.KiwiError, match="2048"):
        hip_ctx.cartesian_map(S, dom, 4096)
    with pytest.raises(hip.KiwiError, match="domain"):
        hip_ctx.cartesian_map(S, [0.0, 1.0, 1.0, 1.0, 0.0, 1.0], 8)
    with pytest.raises(ValueError):
        hip_ctx.cartesian_map(S, dom, 8, out=outd)  # out of the wrong size
    bad = dict(S)
    bad["x"] = S["x"].copy()
    bad["x"][7] = np.nan
    with pytest.raises(hip.KiwiError, match="non-finite"):
        hip_ctx.cartesian_map(bad, dom, 8)
    assert hip_ctx.cartesian_map(S, dom, 4)[5].unresolved == 0


def test_exact_ties_duplicates_and_bucket_edges(hip_ctx):
    """Sites on a lattice whose points sit exactly between cell centres (every centre equidistant from
    several sites), duplicated sites, and sites on the bucket edges of the search grid: the chosen site
    is the oracle's (the lowest index among exact ties) for every centre."""
    from oracle import oracle as O
    say("case: lattice-aligned sites with exact ties, duplicates, bucket-edge sites")
    rng = np.random.default_rng(13)
    dom = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    ax = np.arange(0, 9) / 8.0
    lat = np.stack([v.ravel() for v in np.meshgrid(ax, ax, ax, indexing="ij")], 1)
    for sites in (lat, lat[rng.permutation(len(lat))], np.vstack([lat[::3], lat[::3], lat[1::5]])):
        n = len(sites)
        S = site_rows(rng, np.ascontiguousarray(sites), np.arange(1, n + 1))
        for res in (16, 24):
            cid = hip_ctx.cartesian_map(S, dom, res)[4]
            _, want, _, _ = O.cartesian_map(sites, np.arange(1, n + 1), np.zeros((n, 4)), dom, res)
            assert np.array_equal(cid, want), (n, res)
    # sites exactly on the search grid's bucket edges (one site a bucket: edges at k / cbrt(n))
    n = 343
    e = np.arange(7) / 7.0
    edge = np.stack([v.ravel() for v in np.meshgrid(e, e, e, indexing="ij")], 1)
    S = site_rows(rng, np.ascontiguousarray(edge), np.arange(1, n + 1))
    for res in (20, 33):
        cid = hip_ctx.cartesian_map(S, dom, res)[4]
        _, want, _, _ = O.cartesian_map(edge, np.arange(1, n + 1), np.zeros((n, 4)), dom, res)
        assert np.array_equal(cid, want), res
