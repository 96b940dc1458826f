"""The oracle pinned: against the reference's own numbers (golden fixtures written by the reference built
from source; oracle/_ref when it is built here), against closed forms and against brute-force quadrature.

These run on the CPU (no GPU needed)."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

KEYS = ["M4", "M5", "M6", "WC2", "WC4", "WC6"]
RADKERN = {"M4": 2.0, "M5": 2.5, "M6": 3.0, "WC2": 2.0, "WC4": 2.0, "WC6": 2.0}


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(O.ORACLE_SO):
        O.build()


def box_faces(lo, hi):
    """The 6 faces of an axis-aligned box, in the winding the face walk needs
    (src/cartesian_grid.cpp:54-72)."""
    xm, ym, zm = lo
    xp, yp, zp = hi
    return [[xp, ym, zp, xp, yp, zp, xp, yp, zm, xp, ym, zm], [xm, yp, zm, xm, yp, zp, xm, ym, zp, xm, ym, zm],
            [xp, yp, zm, xp, yp, zp, xm, yp, zp, xm, yp, zm], [xm, ym, zp, xp, ym, zp, xp, ym, zm, xm, ym, zm],
            [xm, yp, zp, xp, yp, zp, xp, ym, zp, xm, ym, zp], [xp, ym, zm, xp, yp, zm, xm, yp, zm, xm, ym, zm]]


def box_mass(k, p, h, lo, hi):
    return sum(O.face(k, p, h, np.array(f))[0] for f in box_faces(lo, hi))


# ------------------------------------------------------------------ golden vectors of the reference
def test_m4_vertex_golden():
    g = np.load(os.path.join(GOLD, "m4_vertex.npz"))
    L = O.lib()
    got = np.array([L.oracle_m4_vertex(*t) for t in zip(g["h"], g["phi"], g["r0"], g["R0"])])
    assert np.array_equal(got, g["value"])  # bit-exact (same libm on this host)


def test_generic_vertex_golden():
    g = np.load(os.path.join(GOLD, "generic_vertex.npz"))
    L = O.lib()
    got = np.array([L.oracle_generic_vertex(int(k), *t)
                    for k, t in zip(g["kernel"], zip(g["H"], g["phi"], g["r0"], g["R0"]))])
    np.testing.assert_allclose(got, g["value"], rtol=1e-13, atol=1e-15)


def test_m4_gridded_golden():
    g = np.load(os.path.join(GOLD, "m4_gridded.npz"))
    L = O.lib()
    got = np.array([L.oracle_m4_gridded(*t) for t in zip(g["h"], g["phi"], g["r0"], g["R0"])])
    np.testing.assert_allclose(got, g["value"], rtol=1e-13, atol=1e-16)


def test_face_golden():
    g = np.load(os.path.join(GOLD, "face.npz"))
    verts = g["verts"]
    off = 0
    for i, nv in enumerate(g["nverts"]):
        v = verts[off:off + 3 * nv]
        off += 3 * nv
        p, h = g["p"][i], g["h"][i]
        m, ov = O.face(O.kernel("M4"), p, h, v)
        assert abs(m - g["m4"][i]) <= 1e-13 + 1e-12 * abs(g["m4"][i])
        assert ov == g["m4_out"][i]
        key = KEYS[g["kernel"][i]]
        mg, ovg = O.face(O.kernel(key), p, h, v)
        assert abs(mg - g["generic"][i]) <= 1e-12 + 1e-11 * abs(g["generic"][i])
        assert ovg == g["generic_out"][i]


def test_c1_grid_matches_reference_cli():
    """The oracle's sph-to-grid loop against the reference CLI's cartesian_grid.dat, 2,000 particles ->
    32^3 (fixture written with OMP_NUM_THREADS=1): identical to the last bit."""
    s = np.load(os.path.join(GOLD, "c1_snapshot.npz"))
    g = np.load(os.path.join(GOLD, "c1_ref_grid.npz"))
    parts = {k: s[k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    dom = s["domain"]
    res = tuple(int(v) for v in g["shape"])
    mass, px, py, pz, st = O.sph_to_grid(parts, dom, *res, threads=1)
    V = ((dom[1] - dom[0]) / res[0]) * ((dom[3] - dom[2]) / res[1]) * ((dom[5] - dom[4]) / res[2])
    assert np.array_equal(mass / V, g["density"])
    pos = mass > 0
    vx = np.where(pos, px / np.where(pos, mass, 1), st.v_mean[0])
    assert np.array_equal(vx, g["vx"])


KERNEL_GRID_CASES = [("M4_generic", "M4", 0.0), ("M4_support2.5", "M4", 2.5), ("M5", "M5", 0.0), ("M6", "M6", 0.0),
                     ("WC2", "WC2", 0.0), ("WC4", "WC4", 0.0), ("WC6", "WC6", 0.0), ("M4_tabulated", "M4", 0.0)]


@pytest.mark.parametrize("label,name,support", KERNEL_GRID_CASES)
def test_generic_grids_match_reference_cli(label, name, support):
    """The oracle's generic and tabulated paths against the reference CLI on every 4th C1 particle onto
    20^3, for each kernel (fixture written with OMP_NUM_THREADS=1): identical to the last bit."""
    g = np.load(os.path.join(GOLD, "kernel_grids.npz"))
    parts = {k: g["p_" + k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    dom = g["domain"]
    res = tuple(int(v) for v in g["res"])
    tab = label == "M4_tabulated"
    k = O.kernel(name, support_in_h=support, generic=not tab, tabulated=tab)
    mass, px, py, pz, st = O.sph_to_grid(parts, dom, *res, k=k, threads=1)
    V = ((dom[1] - dom[0]) / res[0]) * ((dom[3] - dom[2]) / res[1]) * ((dom[5] - dom[4]) / res[2])
    assert np.array_equal(mass / V, g[label + "_density"])
    pos = mass > 0
    for f, q, i in (("vx", px, 0), ("vy", py, 1), ("vz", pz, 2)):
        assert np.array_equal(np.where(pos, q / np.where(pos, mass, 1), st.v_mean[i]), g[f"{label}_{f}"])


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("key", KEYS)
def test_enclosing_box_is_one_and_partitions_add(key):
    """A box enclosing the support has mass 1; two halves add up to the whole, wherever the particle is
    (tools/test_generic_kernel.cpp:261-292 states the same invariants)."""
    rng = np.random.default_rng(KEYS.index(key))
    k = O.kernel(key)
    for _ in range(5):
        p = rng.uniform(-0.5, 0.5, 3)
        h = rng.uniform(0.2, 0.8)
        H = RADKERN[key] * h
        lo, hi = p - H - rng.uniform(0.01, 1, 3), p + H + rng.uniform(0.01, 1, 3)
        assert abs(box_mass(k, p, h, lo, hi) - 1.0) <= 1e-10
        # split along a random axis at a random plane (possibly away from the particle)
        ax = rng.integers(0, 3)
        cut = rng.uniform(lo[ax], hi[ax])
        hi1, lo2 = hi.copy(), lo.copy()
        hi1[ax] = cut
        lo2[ax] = cut
        q = rng.uniform(-1.5, 1.5, 3)  # particle anywhere, also outside the box
        whole = box_mass(k, q, h, lo, hi)
        parts = box_mass(k, q, h, lo, hi1) + box_mass(k, q, h, lo2, hi)
        assert abs(whole - parts) <= 1e-10


def test_guards_are_exact_zeros():
    """r0 = 0, R0 = 0 or phi = 0 give exactly 0 (src/kernel_integral.cpp:357-358, 411-413)."""
    L = O.lib()
    assert L.oracle_m4_vertex(1.0, 0.3, 0.0, 0.5) == 0.0
    assert L.oracle_m4_vertex(1.0, 0.3, 0.5, 0.0) == 0.0
    assert L.oracle_m4_vertex(1.0, 0.0, 0.5, 0.5) == 0.0


def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def _w_shape(key, q):
    p = lambda x, n: np.where(x > 0, np.abs(x) ** n, 0.0)  # noqa: E731
    q = np.asarray(q, dtype=float)
    if key == "M4":
        w = p(1 - q, 3) - 4 * p(0.5 - q, 3)
    elif key == "M5":
        w = p(1 - q, 4) - 5 * p(0.6 - q, 4) + 10 * p(0.2 - q, 4)
    elif key == "M6":
        w = p(1 - q, 5) - 6 * p(2 / 3 - q, 5) + 15 * p(1 / 3 - q, 5)
    elif key == "WC2":
        w = p(1 - q, 4) * (1 + 4 * q)
    elif key == "WC4":
        w = p(1 - q, 6) * (1 + 6 * q + 35 / 3 * q * q)
    else:
        w = p(1 - q, 8) * (1 + 8 * q + 25 * q * q + 32 * q ** 3)
    return np.where(q < 1, w, 0.0)


def _pyramid_quadrature(key, H, r0, R0, phi, n=160):
    """Brute-force kernel mass of the vertex pyramid: integrate W over the triangle (foot, edge foot,
    vertex at azimuth phi) swept from the particle, in spherical shells: mass = int dOmega m(r_b),
    with m(r) = int_0^r W s^2 ds, by Gauss-Legendre in (theta, s)."""
    A3 = {"M4": 16.0, "M5": 30.517578125, "M6": 54.675, "WC2": 10.5, "WC4": 15.46875, "WC6": 21.328125}[key] / np.pi
    xg, wg = _gl(n)
    # azimuth in the wall plane, 0 .. phi
    th = 0.5 * phi * (xg + 1)
    wth = 0.5 * phi * wg
    total = 0.0
    for t, wt in zip(th, wth):
        Rmax = R0 / np.cos(t)  # in-plane distance from the foot to the edge line along azimuth t
        # integrate over in-plane radius rho in [0, Rmax]: dOmega = r0 rho drho / r^3, mass in cone = m(r)
        rho = 0.5 * Rmax * (xg + 1)
        wr = 0.5 * Rmax * wg
        r = np.sqrt(r0 * r0 + rho * rho)
        # m(r) = A3 int_0^{r/H} w(q) q^2 dq
        qq = np.minimum(r / H, 1.0)
        s = 0.5 * qq[:, None] * (xg[None, :] + 1)
        ws = 0.5 * qq[:, None] * wg[None, :]
        m = A3 * np.sum(ws * _w_shape(key, s) * s * s, axis=1)
        total += wt * np.sum(wr * r0 * rho / r ** 3 * m)
    return total


@pytest.mark.parametrize("key", KEYS)
def test_vertex_integral_vs_quadrature(key):
    """Generic (and, for M4, cubic) vertex integrals against brute-force quadrature of the kernel taken
    straight from its definition (the method of tools/test_generic_kernel.cpp:82-112)."""
    rng = np.random.default_rng(100 + KEYS.index(key))
    L = O.lib()
    kid = KEYS.index(key)
    for _ in range(4):
        H = 1.0
        r0 = rng.uniform(0.15, 0.9)
        R0 = rng.uniform(0.15, 0.9)
        phi = rng.uniform(0.2, 1.3)
        ref = _pyramid_quadrature(key, H, r0, R0, phi)
        got = L.oracle_generic_vertex(kid, H, phi, r0, R0)
        assert abs(got - ref) <= 2e-7 * max(abs(ref), 1e-3), (key, r0, R0, phi, got, ref)
        if key == "M4":
            assert abs(L.oracle_m4_vertex(H / 2, phi, r0, R0) - got) <= 1e-12


def test_kernel_tables_regenerate_exactly():
    """kernel_tables.h is exactly what the derivation script writes."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "kiwi_amd", "csrc", "gen_kernel_tables.py"), "--check"])
    assert r.returncode == 0


# ------------------------------------------------------------------ the reference itself, when built here
@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (needs the reference sources)")
def test_oracle_vs_reference_random():
    R = O.ref()
    L = O.lib()
    rng = np.random.default_rng(7)
    for _ in range(3000):
        h = rng.uniform(0.05, 3)
        r0, R0 = h * rng.uniform(0, 3, 2)
        phi = rng.uniform(0, np.pi / 2)
        assert L.oracle_m4_vertex(h, phi, r0, R0) == R.ref_m4_vertex(h, phi, r0, R0)
        kid = int(rng.integers(0, 6))
        assert L.oracle_generic_vertex(kid, 2 * h, phi, r0, R0) == R.ref_generic_vertex(kid, 2 * h, phi, r0, R0)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (needs the reference sources)")
def test_derived_tables_equal_reference_tables():
    """Our exact re-derivation of the kernel tables equals the reference's generated tables."""
    R = O.ref()
    sys.path.insert(0, os.path.join(ROOT, "kiwi_amd", "csrc"))
    import gen_kernel_tables as G
    for kid, (key, name, (breaks, pieces), radkern, hfact) in enumerate(G.KERNELS):
        nreg = ctypes.c_int()
        c4, n4, a, b = np.zeros(4), np.zeros(4), np.zeros(60), np.zeros(256)
        R.ref_kernel_spec(kid, ctypes.byref(nreg), O._p(c4), O._p(n4), O._p(a), O._p(b))
        nr, n2, n3, A, B = G.derive(breaks, pieces)
        assert nr == nreg.value
        np.testing.assert_array_equal(np.array([float(v) for v in breaks] + [0.0] * (4 - len(breaks))), c4)
        np.testing.assert_array_equal([float(n2), float(n3), radkern, hfact], n4)
        np.testing.assert_array_equal(np.array([[float(v) for v in row] for row in A]).reshape(-1), a)
        np.testing.assert_array_equal(np.array([[[float(v) for v in r] for r in blk] for blk in B]).reshape(-1), b)


def test_operation_count_follows_the_computed_path():
    """The counting instantiation behind the reference operation count (oracle/flops_per_pair.py)
    evaluates the same face integrals as the oracle, bit for bit, and counts a plausible number of
    operations per cell (a static estimate from the source: 7.9k-12.1k per pair)."""
    L = O.lib()
    L.oracle_face_counted.restype = ctypes.c_double
    L.oracle_face_counted.argtypes = [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                                              ctypes.POINTER(ctypes.c_longlong)]
    k = O.kernel("M4")
    rng = np.random.default_rng(11)
    per_cell = []
    for _ in range(200):
        p = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.3, 1.5)
        lo = p + rng.uniform(-2.5, 2.0, 3) * h
        hi = lo + rng.uniform(0.1, 0.6, 3) * h
        tot = 0
        for f in box_faces(lo, hi):
            v = np.ascontiguousarray(f, dtype=np.float64)
            cnt = (ctypes.c_longlong * 5)()
            val = L.oracle_face_counted(p[0], p[1], p[2], h, O._p(v), 4, cnt)
            assert val == O.face(k, p, h, v)[0]
            tot += sum(cnt)
        per_cell.append(tot)
    assert 7000 < np.mean(per_cell) < 13000, np.mean(per_cell)
    # a particle in the face plane: every pyramid of the face is flat and the tiers are skipped
    v = np.ascontiguousarray([0.5, -0.5, 0.0, 0.5, 0.5, 0.0, -0.5, 0.5, 0.0, -0.5, -0.5, 0.0])
    cnt = (ctypes.c_longlong * 5)()
    assert L.oracle_face_counted(0.1, 0.2, 0.0, 1.0, O._p(v), 4, cnt) == 0.0
    assert cnt[4] == 16  # only the face walk's own calls remain: 2 acos + 2 sin per edge, 4 edges


# ------------------------------------------------------------------ Voronoi map
def test_voronoi_map_matches_reference_cli():
    """The oracle's Voronoi map (BFS over the tessellation with the face cache, src/mapping.cpp:87-206,
    400-460) against the reference CLI's mapping.dat for the same sites (tests/golden/voronoi_map.npz:
    250 C1 particles as sites, tessellated by the reference's vendored voro++; mapping.dat holds 15
    significant digits)."""
    g = np.load(os.path.join(GOLD, "voronoi_map.npz"))
    parts = {k: g["p_" + k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    mass, px, py, pz, counts = O.voronoi_map(parts, g["pid"], g["home"], g["cell_face_ptr"], g["face_vert_ptr"],
                                             g["vert_xyz"].reshape(-1), g["face_nbr"])
    vol = g["volume"]
    dens = np.where(vol > 0, mass / np.where(vol > 0, vol, 1), 0.0)
    # The tessellation here uses voro++'s own block guess, the CLI's a scaled one: the cell cut order
    # differs, so vertex coordinates agree to rounding only (~1e-16 relative), which moves a few cells'
    # sums by ~1e-12 of the field maximum.
    np.testing.assert_allclose(dens, g["ref_density"], rtol=0, atol=1e-11 * np.abs(g["ref_density"]).max())
    vm = [parts[c].mean() for c in ("vx", "vy", "vz")]
    w = g["ref_density"] / np.abs(g["ref_density"]).max()
    for q, f, i in ((px, "vx", 0), (py, "vy", 1), (pz, "vz", 2)):
        v = np.where(mass > 0, q / np.where(mass > 0, mass, 1), vm[i])
        assert np.max(np.abs(v - g["ref_" + f]) * w) <= 1e-11 * np.abs(g["ref_" + f]).max()
    assert counts[0] == 0  # every particle resolved to a cell
    # the sites tile the domain
    dom = g["domain"]
    assert abs(vol.sum() - np.prod(dom[1::2] - dom[::2])) < 1e-10 * np.prod(dom[1::2] - dom[::2])


def test_inverse_mapping_matches_reference_cli():
    """The oracle's inverse-mapping test (src/mapping.cpp:540-701) against the reference CLI's
    inverse-mapping-test.dat for the setup of voronoi_map.npz (tests/golden/inverse_mapping.npz, 15
    significant digits), with the forward masses density x volume taken from the reference's own
    mapping.dat, as the stage does (:553-559).  Every particle's support lies inside the padded domain,
    so every coverage is 1 to rounding; the per-particle recovered masses pin the sum itself."""
    g = np.load(os.path.join(GOLD, "voronoi_map.npz"))
    gi = np.load(os.path.join(GOLD, "inverse_mapping.npz"))
    parts = {k: g["p_" + k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    assert np.array_equal(gi["pid"], g["pid"])
    cell_mass = g["ref_density"] * g["ref_volume"]
    rec, cov, counts = O.inverse_mapping(parts, g["pid"], g["home"], g["cell_face_ptr"], g["face_vert_ptr"],
                                         g["vert_xyz"].reshape(-1), g["face_nbr"], cell_mass)
    np.testing.assert_allclose(gi["m_original"], parts["m"], rtol=1e-14)
    np.testing.assert_allclose(rec, gi["m_recovered"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(cov, gi["coverage"], rtol=1e-11, atol=0)
    assert abs(rec.sum() - gi["total_recovered"]) <= 1e-12 * gi["total_recovered"]
    assert counts[0] == 0
    # a cell without forward mass drops out of every particle's sum (:656-660)
    cm = cell_mass.copy()
    c0 = int(g["home"][0])
    cm[c0] = 0.0
    rec0, _, _ = O.inverse_mapping(parts, g["pid"], g["home"], g["cell_face_ptr"], g["face_vert_ptr"],
                                   g["vert_xyz"].reshape(-1), g["face_nbr"], cm)
    assert rec0[0] < rec[0] and np.all(rec0 <= rec + 1e-15)


def test_cartesian_map_matches_reference_cli():
    """The oracle's cartesian-map resample (nearest site of every cell centre, src/cartesian_map.cpp:
    150-183) against the reference CLI's cartesian_map.dat at resolution 20 for the setup of
    voronoi_map.npz (tests/golden/cartesian_map.npz): every cell id, and the rows taken from the CLI's
    own mapping.dat, bit for bit."""
    g = np.load(os.path.join(GOLD, "voronoi_map.npz"))
    gc = np.load(os.path.join(GOLD, "cartesian_map.npz"))
    rows = np.stack([g["ref_density"], g["ref_vx"], g["ref_vy"], g["ref_vz"]], 1)
    out, cid, unres, unm = O.cartesian_map(g["site"], g["cell_id"], rows, g["domain"], int(gc["res"]))
    assert np.array_equal(cid, gc["cell_id"])
    for k, f in enumerate(("density", "vx", "vy", "vz")):
        assert np.array_equal(out[:, k], gc[f]), f
    assert unres == 0 and unm == 0
    assert len(np.unique(cid)) > 100  # the grid samples most cells
    # an unmapped site: its cells keep the id and take a zero row (:175-178)
    j = int(np.flatnonzero(g["cell_id"] == cid[0])[0])
    mapped = np.ones(len(rows), bool)
    mapped[j] = False
    out2, cid2, _, unm2 = O.cartesian_map(g["site"], g["cell_id"], rows, g["domain"], int(gc["res"]), mapped)
    hit = cid == cid[0]
    assert np.array_equal(cid2, cid) and unm2 == int(hit.sum()) > 0
    assert np.all(out2[hit] == 0) and np.array_equal(out2[~hit], out[~hit])
    # no sites: every centre unresolved
    assert O.cartesian_map(np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros((0, 4)), g["domain"], 4)[2] == 64


@pytest.mark.skipif(not os.path.exists(O.VORO_SO), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("clustered", [False, True])
def test_cartesian_map_nearest_site_is_voro_cell(clustered):
    """The nearest site in voro++'s distance order is the cell voro++'s own find_voronoi_cell reports
    (the reference's vendored voro++, oracle/_ref/libkiwi_voro.so) for every centre of random uniform and
    clustered site sets in a non-cubic domain.  The sites are the tessellation's cells: the clustered set
    has points clipped onto the max faces, of which voro++ drops those whose block index rounds past the
    last block."""
    rng = np.random.default_rng(7 + clustered)
    dom = np.array([-1.0, 2.0, 0.0, 1.5, -0.5, 0.5])
    lo, hi = dom[0::2], dom[1::2]
    n, res = 700, 18
    if clustered:
        c = rng.uniform(lo + 0.2 * (hi - lo), hi - 0.2 * (hi - lo), size=(3, 3))
        sites = np.clip(c[rng.integers(3, size=n)] + rng.normal(scale=0.08, size=(n, 3)) * (hi - lo), lo, hi)
    else:
        sites = rng.uniform(lo, hi, size=(n, 3))
    ids = rng.permutation(np.arange(5, 5 + n)).astype(np.int32)
    ax = [dom[2 * a] + (np.arange(res) + 0.5) * ((dom[2 * a + 1] - dom[2 * a]) / res) for a in range(3)]
    q = np.stack([v.ravel() for v in np.meshgrid(*ax, indexing="ij")], 1)
    geo = O.tessellate(ids, sites, dom, q)
    assert np.all(geo["home"] >= 0)
    nc = len(geo["cell_id"])
    _, cid, _, _ = O.cartesian_map(geo["site"], geo["cell_id"], np.zeros((nc, 4)), dom, res)
    assert np.array_equal(cid, geo["cell_id"][geo["home"]])
    # sites outside the domain take no part, as in voro++'s pre_container
    far = np.array([[-1.5, 0.7, 0.0], [0.2, 1.6, 0.1], [0.3, 0.4, -0.75]])
    _, cid2, _, _ = O.cartesian_map(np.vstack([far, geo["site"]]), np.concatenate([[1, 2, 3], geo["cell_id"]]),
                                    np.zeros((nc + 3, 4)), dom, res)
    assert np.array_equal(cid2, cid)


def test_coverage_buffer_matches_reference_cli():
    """The oracle's coverage-grid buffer (src/point_set.cpp:166-273) against the buffer sites in the
    reference CLI's mapping.dat (tests/golden/coverage_buffer.npz: ids exact, positions written with 15
    significant digits) for two setups with covered and uncovered candidates."""
    g = np.load(os.path.join(GOLD, "coverage_buffer.npz"))
    for name in ("c1", "lognormal"):
        sn, dom = g[f"{name}_snapshot"], g[f"{name}_domain"]
        ids, xyz = O.coverage_buffer(sn[:, 0], sn[:, 1], sn[:, 2], sn[:, 7], dom, float(g[f"{name}_spacing_fraction"]),
                                     float(g[f"{name}_factor"]))
        assert np.array_equal(ids, g[f"{name}_buffer_id"]), name
        np.testing.assert_allclose(xyz, g[f"{name}_buffer_xyz"], rtol=0, atol=1e-14 * np.abs(dom).max())
        assert 0 < len(ids) < np.prod(O.cartesian_lattice(dom, float(g[f"{name}_spacing_fraction"]))[1])


def coverage_cases(rng):
    """Particle sets for the coverage buffer: clustered with h over two decades in a non-cubic domain,
    with particles outside the domain, h <= 0 and NaN h, and particles placed at exactly r from a
    candidate along an axis."""
    dom = np.array([-1.0, 2.0, 0.0, 1.5, -0.5, 0.5])
    lo, hi = dom[0::2], dom[1::2]
    n = 400
    c = rng.uniform(lo, hi, size=(4, 3))
    p = c[rng.integers(4, size=n)] + rng.normal(scale=0.15, size=(n, 3))
    h = np.exp(rng.uniform(np.log(0.003), np.log(0.12), n))
    h[:5] = [0.0, -0.05, np.nan, 0.3, 0.2]
    sp, _, cand = O.cartesian_lattice(dom, 0.05)
    k = rng.choice(len(cand), 6, replace=False)
    p[-6:] = cand[k]
    p[-6:, 0] += 2.0 * h[-6:]  # a candidate at exactly (or a rounding away from) the radius
    return [(p, h, dom, 0.05, 2.0), (p, h, dom, 0.031, 1.0), (p[:50], h[:50] * 3, dom, 0.08, 2.5)]


@pytest.mark.skipif(not os.path.exists(O.PTS_SO), reason="oracle/_ref not built (needs the reference sources)")
def test_coverage_buffer_matches_reference_function():
    """The oracle's restatement equals the reference's own add_coverage_grid_buffer (oracle/_ref/
    libkiwi_pts.so) bit for bit on edge-case particle sets, and raises as it does with no h > 0."""
    rng = np.random.default_rng(17)
    for p, h, dom, fr, fa in coverage_cases(rng):
        ids, xyz = O.coverage_buffer(p[:, 0], p[:, 1], p[:, 2], h, dom, fr, fa)
        rid, rxyz = O.ref_coverage_buffer(p[:, 0], p[:, 1], p[:, 2], h, dom, fr, fa)
        assert len(ids) > 0 and np.array_equal(ids, rid) and np.array_equal(xyz, rxyz)
    # the lattice sizes of the restatement are the reference's
    for dom, fr in (([0, 1, 0, 1, 0, 1], 0.05), ([0, 1.05, 0, 1, 0, 0.95], 0.1), ([0, 2.5, 0, 1, 0, 3.5], 0.2)):
        _, n, _ = O.cartesian_lattice(dom, fr)
        _, rx = O.ref_coverage_buffer([0.5], [0.5], [0.5], [1e-9], dom, fr, 1.0)
        assert len(rx) == np.prod(n) - 1 or len(rx) == np.prod(n)
    z = np.zeros(3)
    for f in (O.coverage_buffer, O.ref_coverage_buffer):
        with pytest.raises(ValueError, match="all smoothing lengths are <= 0"):
            f(z, z, z, np.array([0.0, -1.0, np.nan]), [0, 1, 0, 1, 0, 1])


def test_voronoi_binary128_arbiter():
    """The binary128 cell sums used to arbitrate Voronoi cells (oracle.voronoi_map_cells_quad): the
    reference's analytic cubic algorithm in binary128 equals the series-free generic path for the same
    M4 kernel (in binary128 to ~1e-15, in double to ~2e-13), while the analytic algorithm in double
    carries the I-series error; at 2x the fixture's smoothing lengths that error is ~1e-11 of the
    field maximum."""
    g = np.load(os.path.join(GOLD, "voronoi_map.npz"))
    parts = {k: g["p_" + k] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    parts["h"] = 2.0 * parts["h"]
    geo = (g["cell_face_ptr"], g["face_vert_ptr"], g["vert_xyz"].reshape(-1), g["face_nbr"])
    gen = O.kernel("M4", generic=True)
    a = O.voronoi_map(parts, g["pid"], g["home"], *geo)
    b = O.voronoi_map(parts, g["pid"], g["home"], *geo, k=gen)
    cells = np.argsort(np.abs(a[0] - b[0]))[-25:]
    qa = O.voronoi_map_cells_quad(parts, g["pid"], g["home"], *geo, cells)
    qb = O.voronoi_map_cells_quad(parts, g["pid"], g["home"], *geo, cells, k=gen)
    for f in range(4):
        s = np.abs(b[f]).max()
        assert np.max(np.abs(qa[:, f] - qb[:, f])) <= 1e-14 * s
        assert np.max(np.abs(b[f][cells] - qb[:, f])) <= 5e-13 * s
    assert np.max(np.abs(a[0][cells] - qa[:, 0])) > 1e-12 * np.abs(b[0]).max()  # the series error is real
    # closing a cell with its lower-index neighbour's copy of a shared face (the GPU map's rule) instead
    # of its own copy changes the sums only by the rounding of voro++'s vertices on this small domain
    qo = O.voronoi_map_cells_quad(parts, g["pid"], g["home"], *geo, cells, share_lower=False)
    for f in range(4):
        d = np.max(np.abs(qa[:, f] - qo[:, f])) / np.abs(b[f]).max()
        assert d <= 1e-13
    assert np.any(qa != qo)  # the rule is applied: some cells are closed with the other copy


def test_realdata_fixture_pins_oracle():
    """The oracle against the reference CLI on the reference's own real-data cutout
    (tests/golden/realdata.npz): every kernel case on the reference test's padded 1 pc grid within 1e-13
    of the field maximum (the CLI writes density = mass / cell volume, divided once more here), and the
    fixture itself satisfies the reference test's conservation thresholds."""
    g = np.load(os.path.join(GOLD, "realdata.npz"))
    parts = {k: np.ascontiguousarray(g["p_" + k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    dom, res = g["domain"], tuple(int(v) for v in g["res"])
    V = np.prod((dom[1::2] - dom[0::2]) / np.array(res))
    cases = (("cubic", O.kernel()), ("cubic_generic", O.kernel("M4", generic=True)),
             ("cubic_tab", O.kernel("M4", tabulated=True)), ("quintic_spline", O.kernel("M6", generic=True)),
             ("wendland_c6", O.kernel("WC6", generic=True)))
    M = parts["m"].sum()
    for label, k in cases:
        m, px, py, pz, st = O.sph_to_grid(parts, dom, *res, k=k, threads=8)
        ref = g[label + "_density"] * V
        assert np.max(np.abs(m - ref)) <= 1e-13 * ref.max(), label
        for p, f in ((px, "vx"), (py, "vy"), (pz, "vz")):
            rp = ref * g[f"{label}_{f}"]
            assert np.max(np.abs(p - rp)) <= 1e-13 * np.abs(rp).max(), (label, f)
        em = abs(ref.sum() - M) / M
        assert em < (1e-3 if label == "cubic_tab" else 1e-9), label


def test_voronoi_arbiter_sharing_rules_on_realdata():
    """The binary128 Voronoi arbiter's face-sharing rules on the reference's real-data cutout (sites = the
    particles, tessellated by the reference's voro++ with its block counts): under the reference's face
    cache (share_lower=2) it equals the oracle's double BFS -- which equals the reference CLI -- to 1e-11 of
    the density maximum, while closing cells with their own copies or with the lower index's copies (the
    GPU's rule) moves the densest clump's cells by ~1e-6: voro++'s two copies of a shared face differ there
    by far more than rounding."""
    if not os.path.exists(O.VORO_SO):
        pytest.skip("voro++ shim not built (needs the reference sources)")
    g = np.load(os.path.join(GOLD, "realdata.npz"))
    parts = {k: np.ascontiguousarray(g["p_" + k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    pos = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    geo = O.tessellate(g["p_id"], pos, g["domain"], pos)
    args = (geo["cell_face_ptr"], geo["face_vert_ptr"], geo["vert_xyz"].reshape(-1), geo["face_nbr"])
    pid = np.arange(1, len(parts["x"]) + 1, dtype=np.int32)
    home = geo["home"].astype(np.int32)
    m = O.voronoi_map(parts, pid, home, *args)[0]
    vol = geo["volume"]
    dmax = np.max(m / vol)
    rows = {int(r[0]): r for r in g["map_plain_rows"]}
    ref = np.array([rows[int(i)][5] for i in geo["cell_id"]])
    assert np.max(np.abs(m / vol - ref)) <= 1e-13 * dmax  # the oracle is the reference CLI here
    cells = np.argsort(m / vol)[-12:]
    err = {}
    for rule in (0, 1, 2):
        q = O.voronoi_map_cells_quad(parts, pid, home, *args, cells, share_lower=rule)[:, 0]
        err[rule] = float(np.max(np.abs(q - m[cells]) / vol[cells]) / dmax)
    assert err[2] <= 1e-11
    assert err[0] > 1e-8 and err[1] > 1e-8
