"""The lattice restatement of the M4 cell integral (kiwi_amd/csrc/m4_lattice.h), checked on the host.

The HIP kernels evaluate Msum(cell) as an 8-corner stencil of lattice values W.  These tests pin that
restatement, compiled for the host from the same header (tests/lattice_host.cpp), against:
  - the reference's cubic algorithm in binary128 (the oracle's quad path, itself the oracle's double
    path instantiated in higher precision, which the golden-vector tests in test_oracle.py pin),
  - closed forms: a box that contains the whole support holds exactly the particle's mass, and
    boxes that tile it add up to it,
  - its own unregrouped forms (the literal sum of the 6 pyramids; the line constants as differences
    of vertex integrals),
  - the single-precision atan against double-precision atan.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
_pd = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lat(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("lattice") / "liblattice_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(HERE, "lattice_host.cpp"),
                    "-o", so], check=True)
    L = ctypes.CDLL(so)
    L.lat_free_both.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _pd]
    L.lat_line_both.argtypes = [ctypes.c_double, ctypes.c_double, _pd]
    L.lat_line_pair.argtypes = [ctypes.c_double, ctypes.c_double, _pd]
    L.lat_box.restype = ctypes.c_double
    L.lat_box.argtypes = [_pd, ctypes.c_double, _pd, _pd, ctypes.c_int]
    L.lat_U.restype = ctypes.c_double
    L.lat_U.argtypes = [ctypes.c_double] * 3
    L.lat_atan_poly_d.restype = ctypes.c_double
    L.lat_atan_poly_d.argtypes = [ctypes.c_double]
    L.lat_atan_poly_f.restype = ctypes.c_float
    L.lat_atan_poly_f.argtypes = [ctypes.c_float]
    L.lat_atan2_pos_f.restype = ctypes.c_float
    L.lat_atan2_pos_f.argtypes = [ctypes.c_float, ctypes.c_float]
    return L


def _box(L, p, h, lo, hi, single=False):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (p, lo, hi)]
    return L.lat_box(a[0].ctypes.data_as(_pd), h, a[1].ctypes.data_as(_pd), a[2].ctypes.data_as(_pd), int(single))


def _region(r2):
    return 1 if r2 < 1 else (2 if r2 < 4 else 3)


def test_closed_form_equals_sum_of_pyramids(lat):
    rng = np.random.default_rng(5)
    pts = list(rng.uniform(0.01, 2.6, size=(400, 3)))
    pts += [(1e-4, 0.7, 0.9), (0.3, 1e-5, 1.2), (1.9, 0.05, 0.04), (0.5, 0.5, 0.5), (1.0, 1.0, 1.0)]
    out = (ctypes.c_double * 2)()
    for ax, ay, az in pts:
        k = _region(ax * ax + ay * ay + az * az)
        lat.lat_free_both(ax, ay, az, k, out)
        assert abs(out[0] - out[1]) <= 1e-12 * max(1.0, abs(out[1])), (ax, ay, az, k, out[0], out[1])


def test_line_constants_shared_polynomials(lat):
    rng = np.random.default_rng(6)
    pairs = list(rng.uniform(0.001, 2.1, size=(600, 2))) + [(0.999, 0.01), (0.01, 0.999), (1.41, 1.41)]
    out = (ctypes.c_double * 4)()
    for da, db in pairs:
        lat.lat_line_both(da, db, out)
        assert abs(out[0] - out[2]) <= 1e-13 and abs(out[1] - out[3]) <= 1e-13, (da, db, list(out))
        if da * da + db * db > 4:
            assert out[0] == 0.0 and out[1] == 0.0


def test_fused_line_pair_equals_the_two_offsets(lat):
    """The line-table entry m4_line_pair(da, db) is m4_line(da, db) + m4_line(db, da): to 1e-14 in double.
    In single precision both forms lose the same cancellation (about 4.5e-7 absolute at worst on these
    pairs); the fused form may not be worse than the two separate offsets.  Both are exactly 0 outside the
    support (l2 >= 4) and for a zero offset."""
    rng = np.random.default_rng(11)
    pairs = list(rng.uniform(0.001, 2.1, size=(3000, 2)))
    pairs += [(0.999, 0.01), (0.01, 0.999), (1.41, 1.41), (0.7071, 0.7071), (1e-4, 0.5), (0.5, 1e-4),
              (1.999, 0.01), (0.0, 0.5), (0.5, 0.0), (1.5, 1.5)]
    out = (ctypes.c_double * 8)()
    worst_fused = worst_sep = 0.0
    for da, db in pairs:
        lat.lat_line_pair(da, db, out)
        v = list(out)
        assert abs(v[0] - v[2]) <= 1e-14 and abs(v[1] - v[3]) <= 1e-14, (da, db, v)
        worst_fused = max(worst_fused, abs(v[4] - v[0]), abs(v[5] - v[1]))
        worst_sep = max(worst_sep, abs(v[6] - v[2]), abs(v[7] - v[3]))
        if da * da + db * db >= 4 or da == 0 or db == 0:
            assert v[0] == v[1] == v[4] == v[5] == 0.0
    assert worst_fused < 1e-6 and worst_fused <= 1.25 * worst_sep, (worst_fused, worst_sep)


@pytest.mark.parametrize("p,h", [((0.13, -0.41, 0.27), 0.8), ((1.77, 0.52, -1.05), 0.35), ((0.0, 0.5, 0.25), 0.6),
                                 ((0.5, 0.5, 0.5), 0.45), ((-0.2, 0.3, 0.71), 1.1)])
def test_box_sums_equal_binary128_reference_algorithm(lat, p, h):
    """Every cell of the particle's bounding box, against the reference's cubic algorithm evaluated in
    binary128.  The cases include a particle on a grid plane (x = 0) and one on a lattice point."""
    dom = np.array([-3.0, 3.0, -2.5, 3.5, -3.0, 3.0])
    res = (24, 24, 24)
    d = (dom[1::2] - dom[::2]) / np.array(res)
    parts = {k: np.array([v]) for k, v in zip(("x", "y", "z"), p)}
    parts.update(vx=np.zeros(1), vy=np.zeros(1), vz=np.zeros(1), m=np.ones(1), h=np.array([h]))
    lo_i = np.maximum(np.floor((np.array(p) - 2 * h - dom[::2]) / d).astype(int), 0)
    hi_i = np.minimum(np.floor((np.array(p) + 2 * h - dom[::2]) / d).astype(int), np.array(res) - 1)
    cells, lat_vals = [], []
    for i in range(lo_i[0], hi_i[0] + 1):
        for j in range(lo_i[1], hi_i[1] + 1):
            for k in range(lo_i[2], hi_i[2] + 1):
                lo = dom[::2] + np.array([i, j, k]) * d
                cells.append((i * res[1] + j) * res[2] + k)
                lat_vals.append(_box(lat, p, h, lo, lo + d))
    q = O.sph_to_grid_cells_quad(parts, dom, *res, np.array(cells))[:, 0]
    lat_vals = np.maximum(np.array(lat_vals), 0.0)  # the reference clamps Msum < 0 to 0
    err = np.abs(lat_vals - q)
    assert err.max() <= 2e-14, (err.max(), cells[int(np.argmax(err))])
    assert abs(q.sum() - 1.0) < 1e-12  # the bounding box holds the whole support


def test_enclosing_box_holds_the_mass_and_tiles_add_up(lat):
    rng = np.random.default_rng(8)
    for _ in range(20):
        p = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.2, 2.0)
        lo = p - 2 * h - rng.uniform(0.01, 1.0, 3)
        hi = p + 2 * h + rng.uniform(0.01, 1.0, 3)
        assert abs(_box(lat, p, h, lo, hi) - 1.0) < 1e-13
        # split the box at random planes through the support: the 8 parts add up to the whole
        cut = p + rng.uniform(-1.5, 1.5, 3) * h
        tot = 0.0
        for sx in range(2):
            for sy in range(2):
                for sz in range(2):
                    a = np.where([sx, sy, sz], cut, lo)
                    b = np.where([sx, sy, sz], hi, cut)
                    tot += _box(lat, p, h, a, b)
        assert abs(tot - 1.0) < 1e-13


def test_single_precision_lattice_close_to_double(lat):
    rng = np.random.default_rng(9)
    worst = 0.0
    for _ in range(300):
        p = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.3, 1.5)
        lo = p + rng.uniform(-2.2, 1.8, 3) * h
        hi = lo + rng.uniform(0.05, 0.8, 3) * h
        worst = max(worst, abs(_box(lat, p, h, lo, hi, single=True) - _box(lat, p, h, lo, hi)))
    assert worst < 2e-6, worst


def test_atan_polynomial_error_bounds(lat):
    t = np.linspace(0.0, 1.0, 200001)[1:]
    exact = np.arctan(t)
    pd = np.array([lat.lat_atan_poly_d(float(v)) for v in t[::10]])
    assert np.max(np.abs(pd - exact[::10]) / exact[::10]) < 4e-8
    t32 = t[::10].astype(np.float32)
    pf = np.array([lat.lat_atan_poly_f(float(v)) for v in t32], dtype=np.float64)
    exact32 = np.arctan(t32.astype(np.float64))  # the float argument's own atan
    assert np.max(np.abs(pf - exact32) / exact32) < 1.5e-7
    rng = np.random.default_rng(10)
    for y, x in rng.uniform(1e-4, 10.0, size=(2000, 2)):
        y32, x32 = float(np.float32(y)), float(np.float32(x))
        assert abs(lat.lat_atan2_pos_f(y32, x32) - math.atan2(y32, x32)) <= 3e-7 * math.atan2(y32, x32)


# ---------------------------------------------------------------------- generic kernels (gen_lattice.h)
KERNELS = ["M4", "M5", "M6", "WC2", "WC4", "WC6"]
RADKERN = {"M4": 2.0, "M5": 2.5, "M6": 3.0, "WC2": 2.0, "WC4": 2.0, "WC6": 2.0}


def _box_faces(lo, hi):
    xm, ym, zm = lo
    xp, yp, zp = hi
    return [[xp, ym, zp, xp, yp, zp, xp, yp, zm, xp, ym, zm], [xm, yp, zm, xm, yp, zp, xm, ym, zp, xm, ym, zm],
            [xp, yp, zm, xp, yp, zp, xm, yp, zp, xm, yp, zm], [xm, ym, zp, xp, ym, zp, xp, ym, zm, xm, ym, zm],
            [xm, yp, zp, xp, yp, zp, xp, ym, zp, xm, ym, zp], [xp, ym, zm, xp, yp, zm, xm, yp, zm, xm, ym, zm]]


def _gen_box(L, kid, p, H, lo, hi):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (p, lo, hi)]
    return L.lat_gen_box(kid, a[0].ctypes.data_as(_pd), H, a[1].ctypes.data_as(_pd), a[2].ctypes.data_as(_pd))


@pytest.mark.parametrize("name", KERNELS)
def test_generic_lattice_boxes_equal_the_generic_face_walk(lat, name):
    """Generic kernels: 8-corner lattice sums against the oracle's generic face walk (the reference's
    src/kernel_integral_generic.cpp algorithm, pinned in test_oracle.py), for random boxes in and
    around the support, and the enclosing box holds the whole mass."""
    lat.lat_gen_box.restype = ctypes.c_double
    lat.lat_gen_box.argtypes = [ctypes.c_int, _pd, ctypes.c_double, _pd, _pd]
    kid = KERNELS.index(name)
    k = O.kernel(name, generic=True)
    rng = np.random.default_rng(20 + kid)
    worst = 0.0
    for _ in range(150):
        p = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.3, 1.2)
        H = RADKERN[name] * h
        lo = p + rng.uniform(-1.1, 0.9, 3) * H
        hi = lo + rng.uniform(0.05, 0.7, 3) * H
        ref = sum(O.face(k, p, h, np.array(f))[0] for f in _box_faces(lo, hi))
        got = _gen_box(lat, kid, p, H, lo, hi)
        worst = max(worst, abs(got - ref))
    assert worst < 1e-12, worst
    p = np.array([0.1, -0.2, 0.3])
    H = RADKERN[name] * 0.7
    assert abs(_gen_box(lat, kid, p, H, p - 1.3 * H, p + 1.2 * H) - 1.0) < 1e-13


@pytest.mark.parametrize("name", ["M4", "WC2"])
def test_generic_single_precision_lattice(lat, name):
    """The mixed-precision generic path (generic cubic spline and Wendland C2: gen_lattice.h in float)
    against the double-precision lattice: values finite and within 2e-6 of the double ones at random
    offsets, including offsets of 1e-6..1e-12 of the support on one or two axes (the basis recurrences
    overflowed single precision there before they were scaled); and, for an isolated particle at
    h = 0.8 and 2.5 cells, cell sums of the single-precision lattice (summed in double, as on the
    device) within 1e-6 of the particle's mass in total, 1e-3 in L1, and 2e-4 of the largest cell."""
    for f in (lat.lat_gen_W, lat.lat_gen_W32):
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    kid = KERNELS.index(name)
    rng = np.random.default_rng(31 + kid)
    worst = 0.0
    for i in range(3000):
        d = rng.uniform(0.0, 1.2, 3)
        if i % 3 == 1:
            d[i % 3] = 10.0 ** rng.uniform(-12, -6)
        if i % 3 == 2:
            d[:2] = 10.0 ** rng.uniform(-12, -6, 2)
        if np.linalg.norm(d) < 1e-3:
            continue
        w64, w32 = lat.lat_gen_W(kid, *d), lat.lat_gen_W32(kid, *d)
        assert np.isfinite(w32), d
        worst = max(worst, abs(w32 - w64))
    # lattice values are sums of O(1) plane, line and pyramid terms (which cancel to ~0 outside the
    # support): the difference is a few single-precision ulps of those terms
    assert worst < 2e-6, worst

    def cells(W, p, hc):
        H = RADKERN[name] * hc
        n = int(np.ceil(2 * H)) + 3
        lo = np.floor(p - H) - 1
        g = [lo[a] + np.arange(n + 1) for a in range(3)]
        U = np.zeros((n + 1,) * 3)
        for i in range(n + 1):
            for j in range(n + 1):
                for k in range(n + 1):
                    dx, dy, dz = (g[0][i] - p[0]) / H, (g[1][j] - p[1]) / H, (g[2][k] - p[2]) / H
                    if dx == 0 or dy == 0 or dz == 0:
                        continue
                    w = W(kid, abs(dx), abs(dy), abs(dz))
                    U[i, j, k] = -w if (dx < 0) != ((dy < 0) != (dz < 0)) else w
        return U[1:, 1:, 1:] + U[1:, :-1, :-1] + U[:-1, 1:, :-1] + U[:-1, :-1, 1:] \
            - U[:-1, :-1, :-1] - U[1:, 1:, :-1] - U[1:, :-1, 1:] - U[:-1, 1:, 1:]

    for hc in (0.8, 2.5):
        p = rng.uniform(0, 1, 3)
        c64, c32 = cells(lat.lat_gen_W, p, hc), cells(lat.lat_gen_W32, p, hc)
        M = c64.sum()
        assert abs(M - 1.0) < 1e-12
        assert abs(c32.sum() - M) < 1e-6
        assert np.abs(c32 - c64).sum() < 1e-3
        assert np.abs(c32 - c64).max() < 2e-4 * np.abs(c64).max()


def test_polygon_face_walk_single_precision(lat):
    """The Voronoi map's mixed mode: the cubic spline's face walk in single precision, on vertex
    coordinates relative to the particle and Newell's plane normal, against the double-precision walk on
    the real Voronoi cells of tests/golden/voronoi_map.npz (h 0.5x to 3x the fixture's).  Faces with
    two nearly coincident vertices (edges 1e-4 h) made the first-three-vertex plane lose 1e-4 of r0 in
    single precision (4.7e-5 per face); with Newell's normal the worst face is ~1e-6, at walls near the
    support edge where the plane and pyramid terms of the vertex integral (O(5)) cancel to O(0.25).
    Bounds: every face within 3e-6, every cell (sum of its faces) within 5e-6, the median face 1e-7."""
    lat.lat_face_walk.restype = ctypes.c_double
    lat.lat_face_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _pd, ctypes.c_double, _pd,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lat.lat_face_walk32.restype = ctypes.c_double
    lat.lat_face_walk32.argtypes = [_pd, ctypes.c_double, _pd, ctypes.c_int]
    g = np.load(os.path.join(HERE, "golden", "voronoi_map.npz"))
    cfp, fvp, vx = g["cell_face_ptr"], g["face_vert_ptr"], g["vert_xyz"].reshape(-1, 3)
    P = np.stack([g["p_x"], g["p_y"], g["p_z"]], 1)
    rng = np.random.default_rng(3)
    face_err, cell_err = [], 0.0
    for _ in range(200):
        i = rng.integers(len(P))
        p = np.ascontiguousarray(P[i])
        h = g["p_h"][i] * rng.uniform(0.5, 3.0)
        for c in np.argsort(np.linalg.norm(g["site"] - p, axis=1))[:12]:
            m64 = m32 = 0.0
            for f in range(cfp[c], cfp[c + 1]):
                v = np.ascontiguousarray(vx[fvp[f]:fvp[f + 1]]).reshape(-1)
                r = ctypes.c_int()
                a = lat.lat_face_walk(0, 0, 2.0, p.ctypes.data_as(_pd), h, v.ctypes.data_as(_pd), len(v) // 3,
                                      ctypes.byref(r))
                b = lat.lat_face_walk32(p.ctypes.data_as(_pd), h, v.ctypes.data_as(_pd), len(v) // 3)
                assert np.isfinite(b)
                face_err.append(abs(a - b))
                m64 += a
                m32 += b
            cell_err = max(cell_err, abs(m64 - m32))
    assert max(face_err) < 3e-6, max(face_err)
    assert np.median(face_err) < 1e-7
    assert cell_err < 5e-6, cell_err


# ---------------------------------------------------------------------- tabulated cubic path (m4_table.h)
def test_vertex_value_and_gridded_lookup_match_reference_golden(lat):
    """m4_table.h: the FP64 vertex integral that fills the lookup table, against the reference's
    full_integral_new (tests/golden/m4_vertex.npz, guards and region boundaries included), and the
    trilinear lookup over a table this header generates, against the reference's gridded_integral over
    its own table (tests/golden/m4_gridded.npz)."""
    for f in ("lat_m4_vertex_value", "lat_m4_gridded"):
        getattr(lat, f).restype = ctypes.c_double
        getattr(lat, f).argtypes = [ctypes.c_double] * 4
    g = np.load(os.path.join(HERE, "golden", "m4_vertex.npz"))
    v = np.array([lat.lat_m4_vertex_value(*t) for t in zip(g["h"], g["phi"], g["r0"], g["R0"])])
    err = np.abs(v - g["value"])
    # The reference's I-series loses up to ~2e-12 for vertices with R0 << r0 (a = R0/r0 ~ 1e-3); brute-force
    # quadrature of the worst such case in this fixture (h 1.119, phi 0.8204, r0 1.9547, R0 0.005752) gives
    # 3.6831441429e-07, this header 3.6831441440e-07 and the reference 3.6831247540e-07.
    assert err.max() < 2.5e-12, err.max()
    assert np.median(err) < 1e-16
    assert np.all(v[g["r0"] == 0] == 0) and np.all(v[g["R0"] == 0] == 0) and np.all(v[g["phi"] == 0] == 0)
    g = np.load(os.path.join(HERE, "golden", "m4_gridded.npz"))
    v = np.array([lat.lat_m4_gridded(*t) for t in zip(g["h"], g["phi"], g["r0"], g["R0"])])
    assert np.max(np.abs(v - g["value"])) < 1e-13


# ---------------------------------------------------------------------- Voronoi face walk (face_walk.h)
@pytest.mark.parametrize("kind,name,generic,tabulated,qtol", [(0, "M4", False, False, 5e-14),
                                                              (1, "M6", True, False, 5e-14),
                                                              (1, "WC2", True, False, 5e-14),
                                                              (1, "WC6", True, False, 3e-13),
                                                              (2, "M4", False, True, None)])
def test_polygon_face_walk_matches_oracle_face_walk(lat, kind, name, generic, tabulated, qtol):
    """face_walk.h (the Voronoi map's per-face integral and reachability test, run by the HIP kernels)
    against the oracle's face walk (src/kernel_integral.cpp:471-647, pinned by the golden grids in
    test_oracle.py) on real Voronoi faces: the cells of tests/golden/voronoi_map.npz near randomly
    chosen particles, with smoothing lengths from 0.5x to 3x the fixture's.  The reachability flag
    must agree exactly and the values to 1e-12 (the reference's acos of R0 / r12 and its cubic I-series
    round differently from the restated geometry here); against the same face walk in binary128 the
    values must lie within `qtol` (WC6: its monomials cancel near the support edge).  The tabulated
    path has no binary128 form: 1e-13 against the oracle's double."""
    lat.lat_face_walk.restype = ctypes.c_double
    lat.lat_face_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _pd, ctypes.c_double, _pd,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    g = np.load(os.path.join(HERE, "golden", "voronoi_map.npz"))
    cfp, fvp, vx = g["cell_face_ptr"], g["face_vert_ptr"], g["vert_xyz"].reshape(-1, 3)
    P = np.stack([g["p_x"], g["p_y"], g["p_z"]], 1)
    k = O.kernel(name, generic=generic, tabulated=tabulated)
    kid = KERNELS.index(name)
    rng = np.random.default_rng(40 + kind + kid)
    worst, worst_q, reach, nfaces = 0.0, 0.0, 0, 0
    for _ in range(120):
        i = rng.integers(len(P))
        p = np.ascontiguousarray(P[i])
        h = g["p_h"][i] * rng.uniform(0.5, 3.0)
        near = np.argsort(np.linalg.norm(g["site"] - p, axis=1))[:12]
        c = int(rng.choice(near))
        for f in range(cfp[c], cfp[c + 1]):
            v = np.ascontiguousarray(vx[fvp[f]:fvp[f + 1]]).reshape(-1)
            r = ctypes.c_int()
            m = lat.lat_face_walk(kind, kid, RADKERN[name], p.ctypes.data_as(_pd), h, v.ctypes.data_as(_pd),
                                  len(v) // 3, ctypes.byref(r))
            mo, ov = O.face(k, p, h, v)
            assert r.value == int(ov != len(v) // 3), (i, c, f)
            worst = max(worst, abs(m - mo))
            if qtol is not None:
                worst_q = max(worst_q, abs(m - O.face_quad(k, p, h, v)))
            reach += r.value
            nfaces += 1
    assert worst < (1e-13 if tabulated else 1e-12), worst
    if qtol is not None:
        assert worst_q < qtol, worst_q
    assert 0 < reach < nfaces  # both outcomes exercised


# ---------------------------------------------------------------------- adjoint (m4_adjoint.h, dual.h)
def _face_flux(a, b, c, n=600):
    """The M4 kernel's integral over the rectangle [0, b] x [0, c] at distance a (units of h), by
    Gauss-Legendre quadrature: the derivative of the box mass W(a, b, c) by a."""
    xg, wg = np.polynomial.legendre.leggauss(n)
    y, z = 0.5 * b * (xg + 1), 0.5 * c * (xg + 1)
    Y, Z = np.meshgrid(y, z, indexing="ij")
    q = np.sqrt(a * a + Y * Y + Z * Z)
    w = np.where(q < 1, 1 - 1.5 * q * q + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0)) / np.pi
    return 0.25 * b * c * np.einsum("i,j,ij->", wg, wg, w)


def test_adjoint_lattice_gradient(lat):
    """m4_adj_point (the adjoint's lattice value and gradient, dual numbers through the forward's closed
    form): the value equals the forward's W, the gradient equals central differences of W and, as an
    independent check, the kernel's flux through the box face by 2-D quadrature; the single-precision
    gradient is within 5e-6 of the double one (cancellation in the free part)."""
    lat.lat_adj_point.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _pd]
    lat.lat_U.restype = ctypes.c_double
    lat.lat_U.argtypes = [ctypes.c_double] * 3

    def adj(a, b, c, single=0):
        o = (ctypes.c_double * 4)()
        lat.lat_adj_point(a, b, c, single, o)
        return np.array(o[:])

    rng = np.random.default_rng(31)
    pts = list(rng.uniform(0.02, 2.4, size=(200, 3))) + [(0.5, 0.5, 0.5), (1e-3, 0.9, 1.1), (0.9, 1.2, 1.9)]
    eps = 1e-5
    for t, (a, b, c) in enumerate(pts):
        g = adj(a, b, c)
        assert abs(g[0] - lat.lat_U(a, b, c)) < 1e-14
        fd = [(lat.lat_U(a + eps, b, c) - lat.lat_U(a - eps, b, c)) / (2 * eps),
              (lat.lat_U(a, b + eps, c) - lat.lat_U(a, b - eps, c)) / (2 * eps),
              (lat.lat_U(a, b, c + eps) - lat.lat_U(a, b, c - eps)) / (2 * eps)]
        assert np.max(np.abs(np.array(fd) - g[1:])) < 1e-9, (a, b, c)
        if t < 40:
            assert abs(_face_flux(a, b, c) - g[1]) < 1e-11
            assert abs(_face_flux(c, a, b) - g[3]) < 1e-11
        assert np.max(np.abs(adj(a, b, c, 1) - g)) < 5e-6


def test_adjoint_lattice_gradient_at_region_crossings(lat):
    """Lattice points whose edge line sits exactly on the region boundaries l^2 = 1 and l^2 = 4 (where
    the line constants take square roots of 0): the gradient stays finite and equals central differences
    in double; in single precision no point near the boundaries gives a non-finite gradient."""
    lat.lat_adj_point.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _pd]
    lat.lat_U.restype = ctypes.c_double
    lat.lat_U.argtypes = [ctypes.c_double] * 3

    def adj(a, b, c, single=0):
        o = (ctypes.c_double * 4)()
        lat.lat_adj_point(a, b, c, single, o)
        return np.array(o[:])

    eps = 1e-6
    for a, b in ((0.6, 0.8), (0.8, 0.6), (1.2, 1.6), (0.0 + 1.0, 1e-300 + 0.0 + 1.7320508075688772)):
        for c in (0.3, 0.9, 1.7):
            g = adj(a, b, c)
            assert np.all(np.isfinite(g)), (a, b, c)
            fd = [(lat.lat_U(a + eps, b, c) - lat.lat_U(a - eps, b, c)) / (2 * eps),
                  (lat.lat_U(a, b + eps, c) - lat.lat_U(a, b - eps, c)) / (2 * eps),
                  (lat.lat_U(a, b, c + eps) - lat.lat_U(a, b, c - eps)) / (2 * eps)]
            assert np.max(np.abs(np.array(fd) - g[1:])) < 1e-7, (a, b, c, g, fd)
    rng = np.random.default_rng(9)
    for _ in range(4000):
        th = rng.uniform(0, np.pi / 2)
        rr = rng.choice([1.0, 2.0])
        a, b = np.float32(rr * np.cos(th)), np.float32(rr * np.sin(th))
        assert np.all(np.isfinite(adj(float(a), float(b), rng.uniform(0.01, 2.5), 1)))


@pytest.mark.parametrize("kid", range(6))
def test_generic_adjoint_lattice_gradient(lat, kid):
    """The generic kernels' lattice gradient (gen_adjoint.h, dual numbers through gen_lattice.h's closed
    form) for M4, M5, M6, WC2, WC4, WC6: the value equals the forward's W (gen_lattice_W) and the gradient
    equals central differences of W, at random offsets inside and beyond the support (units of H),
    including points next to the region boundaries; the single-precision gradient of the kernels the
    mixed mode runs in FP32 (M4, WC2) stays within 1e-4 of the double one."""
    lat.lat_gen_adj_point.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + [ctypes.c_int, _pd]
    lat.lat_gen_W.restype = ctypes.c_double
    lat.lat_gen_W.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3

    def adj(a, b, c, single=0):
        o = (ctypes.c_double * 4)()
        lat.lat_gen_adj_point(kid, a, b, c, single, o)
        return np.array(o[:])

    def W(a, b, c):
        return lat.lat_gen_W(kid, a, b, c)

    rng = np.random.default_rng(40 + kid)
    pts = list(rng.uniform(0.01, 1.2, size=(150, 3))) + [(0.3, 0.3, 0.3), (1e-3, 0.45, 0.5), (0.5, 0.6, 0.55)]
    eps = 1e-5
    worst = 0.0
    for a, b, c in pts:
        g = adj(a, b, c)
        assert abs(g[0] - W(a, b, c)) < 1e-13, (a, b, c)  # summation order (WC6 cancels near the edge)
        fd = np.array([(W(a + eps, b, c) - W(a - eps, b, c)) / (2 * eps),
                       (W(a, b + eps, c) - W(a, b - eps, c)) / (2 * eps),
                       (W(a, b, c + eps) - W(a, b, c - eps)) / (2 * eps)])
        err = np.max(np.abs(fd - g[1:]))
        worst = max(worst, err)
        assert err < 1e-7, (kid, a, b, c, g, fd)
        if kid in (0, 3):
            assert np.max(np.abs(adj(a, b, c, 1) - g)) < 1e-4, (kid, a, b, c)
