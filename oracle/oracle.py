"""ctypes bindings for the CPU oracle (oracle/libkiwi_oracle.so) and, when it has been built here,
for the reference itself (oracle/_ref/libkiwi_ref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libkiwi_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkiwi_ref.so")
REF_CLI = os.path.join(HERE, "_ref", "kiwi")
VORO_SO = os.path.join(HERE, "_ref", "libkiwi_voro.so")
PTS_SO = os.path.join(HERE, "_ref", "libkiwi_pts.so")

KERNELS = ["M4", "M5", "M6", "WC2", "WC4", "WC6"]

_d = ctypes.c_double
_i = ctypes.c_int
_ll = ctypes.c_longlong
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pll = ctypes.POINTER(ctypes.c_longlong)


class OracleKernel(ctypes.Structure):
    _fields_ = [("kernel_id", _i), ("support_in_h", _d), ("generic_path", _i), ("tabulated", _i)]


class OracleStats(ctypes.Structure):
    _fields_ = [("skipped_outside", _ll), ("n_negative", _ll), ("n_negative_warn", _ll),
                ("pairs_bbox", _ll), ("pairs_nonzero", _ll), ("v_mean", _d * 3)]


def build():
    """Compile the oracle (and nothing else) in-tree."""
    import subprocess
    root = os.path.dirname(HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile", "oracle/libkiwi_oracle.so"], cwd=root, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library not built: {ORACLE_SO} (run make -f oracle/Makefile)")
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_m4_vertex.restype = _d
        L.oracle_m4_vertex.argtypes = [_d, _d, _d, _d]
        L.oracle_m4_gridded.restype = _d
        L.oracle_m4_gridded.argtypes = [_d, _d, _d, _d]
        L.oracle_generic_vertex.restype = _d
        L.oracle_generic_vertex.argtypes = [_i, _d, _d, _d, _d]
        L.oracle_face.restype = _d
        L.oracle_face.argtypes = [ctypes.POINTER(OracleKernel), _d, _d, _d, _d, _pd, _i, _pi]
        L.oracle_face_q.restype = _d
        L.oracle_face_q.argtypes = [ctypes.POINTER(OracleKernel), _d, _d, _d, _d, _pd, _i]
        L.oracle_sph_to_grid.restype = _i
        L.oracle_sph_to_grid.argtypes = [_ll] + [_pd] * 8 + [_pd, _i, _i, _i, ctypes.POINTER(OracleKernel), _i,
                                                              _pd, _pd, _pd, _pd, ctypes.POINTER(OracleStats)]
        L.oracle_sph_to_grid_cells_q.restype = _i
        L.oracle_sph_to_grid_cells_q.argtypes = [_ll] + [_pd] * 8 + [_pd, _i, _i, _i, ctypes.POINTER(OracleKernel),
                                                                      _ll, _pll, _pd]
        L.oracle_voronoi_map.restype = _i
        L.oracle_voronoi_map.argtypes = [_ll] + [_pd] * 8 + [_pi, _pi, _ll, _pll, _pll, _pd, _pi,
                                                              ctypes.POINTER(OracleKernel), _pd, _pd, _pd, _pd, _pll]
        L.oracle_inverse_mapping.restype = _i
        L.oracle_inverse_mapping.argtypes = [_ll] + [_pd] * 5 + [_pi, _pi, _ll, _pll, _pll, _pd, _pi,
                                                                  ctypes.POINTER(OracleKernel), _pd, _pd, _pd, _pll]
        L.oracle_voronoi_map_cells_q.restype = _i
        L.oracle_voronoi_map_cells_q.argtypes = [_ll] + [_pd] * 8 + [_pi, _pi, _ll, _pll, _pll, _pd, _pi,
                                                                      ctypes.POINTER(OracleKernel), _ll, _pll, _i, _pd]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    """The reference's own kernel math (built from /root/reference by oracle/ref.mk), or None."""
    global _ref
    if _ref is None:
        if not have_ref():
            return None
        R = ctypes.CDLL(REF_SO)
        R.ref_m4_vertex.restype = _d
        R.ref_m4_vertex.argtypes = [_d, _d, _d, _d]
        R.ref_m4_face.restype = _d
        R.ref_m4_face.argtypes = [_d, _d, _d, _d, _pd, _i, _pi]
        R.ref_generic_vertex.restype = _d
        R.ref_generic_vertex.argtypes = [_i, _d, _d, _d, _d]
        R.ref_generic_face.restype = _d
        R.ref_generic_face.argtypes = [_i, _d, _d, _d, _d, _pd, _i, _pi]
        R.ref_kernel_spec.restype = _i
        R.ref_kernel_spec.argtypes = [_i, _pi, _pd, _pd, _pd, _pd]
        R.ref_m4_gridded.restype = _d
        R.ref_m4_gridded.argtypes = [_d, _d, _d, _d]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(_pd)


def tessellate(ids, pts, domain, particles):
    """Voronoi geometry of the points (reference's vendored voro++, oracle/_ref/libkiwi_voro.so) as CSR
    arrays, each cell's volume and centroid, plus each particle's home cell.  pts: (n, 3); particles:
    (m, 3).  Returns a dict."""
    V = ctypes.CDLL(VORO_SO)
    V.voro_build.restype = ctypes.c_void_p
    V.voro_build.argtypes = [_i, _pi, _pd, _pd, _pd, _pd, _ll, _pd, _pd, _pd, _pll]
    V.voro_fetch.argtypes = [ctypes.c_void_p, _pi, _pd, _pd, _pd, _pll, _pll, _pd, _pi, _pi]
    V.voro_free.argtypes = [ctypes.c_void_p]
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    P = [np.ascontiguousarray(pts[:, k], dtype=np.float64) for k in range(3)]
    Q = [np.ascontiguousarray(particles[:, k], dtype=np.float64) for k in range(3)]
    dom = np.ascontiguousarray(domain, dtype=np.float64)
    sizes = np.zeros(3, dtype=np.int64)
    h = V.voro_build(len(ids), ids.ctypes.data_as(_pi), *[_p(a) for a in P], _p(dom), len(Q[0]), *[_p(a) for a in Q],
                     sizes.ctypes.data_as(_pll))
    nc, nf, nv = (int(v) for v in sizes)
    g = {"cell_id": np.zeros(nc, np.int32), "site": np.zeros((nc, 3)), "volume": np.zeros(nc), "centroid": np.zeros((nc, 3)),
         "cell_face_ptr": np.zeros(nc + 1, np.int64), "face_vert_ptr": np.zeros(nf + 1, np.int64),
         "vert_xyz": np.zeros((nv, 3)), "face_nbr": np.zeros(nf, np.int32), "home": np.zeros(len(Q[0]), np.int32)}
    V.voro_fetch(h, g["cell_id"].ctypes.data_as(_pi), _p(g["site"]), _p(g["volume"]), _p(g["centroid"]),
                 g["cell_face_ptr"].ctypes.data_as(_pll), g["face_vert_ptr"].ctypes.data_as(_pll), _p(g["vert_xyz"]),
                 g["face_nbr"].ctypes.data_as(_pi), g["home"].ctypes.data_as(_pi))
    V.voro_free(h)
    return g


def kernel(name="M4", support_in_h=0.0, generic=False, tabulated=False):
    return OracleKernel(KERNELS.index(name), support_in_h, int(generic), int(tabulated))


def face_quad(k, p, h, verts):
    """oracle face integral with the kernel mathematics in binary128 (analytic cubic or generic path)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    return lib().oracle_face_q(ctypes.byref(k), p[0], p[1], p[2], h, _p(v), len(v) // 3)


def face(k, p, h, verts):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    ov = ctypes.c_int(0)
    m = lib().oracle_face(ctypes.byref(k), p[0], p[1], p[2], h, _p(v), len(v) // 3, ctypes.byref(ov))
    return m, ov.value


def sph_to_grid(parts, domain, nx, ny, nz, k=None, threads=1):
    """parts: dict of float64 arrays x y z vx vy vz m h.  Returns (mass, px, py, pz, stats) un-normalised."""
    k = k or kernel()
    n = len(parts["x"])
    cols = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in ("x", "y", "z", "vx", "vy", "vz", "m", "h")]
    dom = np.ascontiguousarray(domain, dtype=np.float64)
    nc = nx * ny * nz
    out = [np.zeros(nc) for _ in range(4)]
    st = OracleStats()
    rc = lib().oracle_sph_to_grid(n, *[_p(c) for c in cols], _p(dom), nx, ny, nz, ctypes.byref(k), threads,
                                  *[_p(o) for o in out], ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_sph_to_grid failed ({rc})")
    return out[0], out[1], out[2], out[3], st


def sph_to_grid_cells_quad(parts, domain, nx, ny, nz, cells, k=None):
    """Reference sums (mass, px, py, pz) for the listed flat cell indices, with the kernel mathematics
    (cubic or generic path, per k) in binary128.  Returns an array of shape (len(cells), 4)."""
    k = k or kernel()
    cols = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in ("x", "y", "z", "vx", "vy", "vz", "m", "h")]
    dom = np.ascontiguousarray(domain, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    out = np.zeros((len(cells), 4))
    rc = lib().oracle_sph_to_grid_cells_q(len(cols[0]), *[_p(c) for c in cols], _p(dom), nx, ny, nz, ctypes.byref(k),
                                          len(cells), cells.ctypes.data_as(_pll), _p(out))
    if rc != 0:
        raise ValueError("oracle_sph_to_grid_cells_q failed")
    return out


def voronoi_map(parts, pid, home, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr, k=None):
    k = k or kernel()
    n = len(parts["x"])
    cols = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in ("x", "y", "z", "vx", "vy", "vz", "m", "h")]
    pid = np.ascontiguousarray(pid, dtype=np.int32)
    home = np.ascontiguousarray(home, dtype=np.int32)
    cfp = np.ascontiguousarray(cell_face_ptr, dtype=np.int64)
    fvp = np.ascontiguousarray(face_vert_ptr, dtype=np.int64)
    fxyz = np.ascontiguousarray(face_xyz, dtype=np.float64)
    fnb = np.ascontiguousarray(face_nbr, dtype=np.int32)
    ncells = len(cfp) - 1
    out = [np.zeros(ncells) for _ in range(4)]
    counts = np.zeros(6, dtype=np.int64)
    rc = lib().oracle_voronoi_map(n, *[_p(c) for c in cols], pid.ctypes.data_as(_pi), home.ctypes.data_as(_pi),
                                  ncells, cfp.ctypes.data_as(_pll), fvp.ctypes.data_as(_pll), _p(fxyz),
                                  fnb.ctypes.data_as(_pi), ctypes.byref(k), *[_p(o) for o in out],
                                  counts.ctypes.data_as(_pll))
    if rc != 0:
        raise ValueError(f"oracle_voronoi_map failed ({rc})")
    return out[0], out[1], out[2], out[3], counts


def _round_half_away(v):
    """std::round: to the nearest integer, halves away from zero (v - trunc(v) is exact)."""
    t = float(np.trunc(v))
    return int(t + np.sign(v)) if abs(v - t) >= 0.5 else int(t)


def cartesian_lattice(domain, spacing_fraction):
    """The buffer's candidate lattice (generate_cartesian_lattice, src/point_set.cpp:166-195): spacing =
    fraction x the smallest extent; n_a = max(1, round(extent_a / spacing)) (round half away from zero);
    candidate (i, j, k) at min + spacing (index + 0.5), k fastest.  Returns (spacing, (nx, ny, nz), xyz)."""
    d = np.asarray(domain, dtype=np.float64)
    ext = d[1::2] - d[0::2]
    sp = spacing_fraction * min(ext[0], ext[1], ext[2])
    n = [max(1, _round_half_away(e / sp)) for e in ext]
    ax = [d[2 * a] + sp * (np.arange(n[a]) + 0.5) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return sp, tuple(n), np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def _bin_keys(v, bin_size):
    """Per-axis hash key of the reference's coverage hash: floor(v / bin) as int, biased by 2^19 and
    clamped to [0, 2^20) (bin_of and bin_key, src/point_set.cpp:29-35, 217-222)."""
    b = np.floor(np.asarray(v) / bin_size)
    b = np.clip(b, -2.0 ** 30, 2.0 ** 30).astype(np.int64)
    return np.clip(b + (1 << 19), 0, (1 << 20) - 1)


def coverage_buffer(x, y, z, h, domain, spacing_fraction=0.05, factor=2.0, chunk=2048):
    """The coverage-grid buffer (add_coverage_grid_buffer, src/point_set.cpp:208-273) restated with numpy:
    a lattice candidate is covered when some particle whose hash bin is one of the 27 around the
    candidate's (the reference's search set; bin = factor x h_max, h_max = the largest h, NaN ignored)
    lies within r = factor h of it: (dx^2 + dy^2) + dz^2 <= r * r, dx = particle - candidate, all in
    double without fused multiply-add.  The uncovered candidates are kept in lattice order with ids
    n + 1, n + 2, ...  Returns (ids, xyz); raises ValueError as the reference throws."""
    x, y, z, h = (np.asarray(a, dtype=np.float64) for a in (x, y, z, h))
    hm = float(np.fmax.reduce(np.concatenate([[0.0], h])))  # std::max from 0: a NaN never takes over
    if not hm > 0.0:
        raise ValueError("add_coverage_grid_buffer: all smoothing lengths are <= 0")
    bin_size = factor * hm
    sp, n, cand = cartesian_lattice(domain, spacing_fraction)
    pk = [_bin_keys(a, bin_size) for a in (x, y, z)]
    r = factor * h
    r2 = r * r
    covered = np.zeros(len(cand), bool)
    for c0 in range(0, len(cand), chunk):
        c = cand[c0:c0 + chunk]
        near = np.ones((len(c), len(x)), bool)
        for a in range(3):
            ck = _bin_keys(c[:, a], bin_size)
            lo_k = _bin_keys(np.floor(c[:, a] / bin_size) - 1.0, 1.0)  # key of (bin - 1), (bin + 1)
            hi_k = _bin_keys(np.floor(c[:, a] / bin_size) + 1.0, 1.0)
            near &= (pk[a][None, :] == ck[:, None]) | (pk[a][None, :] == lo_k[:, None]) | (pk[a][None, :] == hi_k[:, None])
        dx = x[None, :] - c[:, 0:1]
        dy = y[None, :] - c[:, 1:2]
        dz = z[None, :] - c[:, 2:3]
        covered[c0:c0 + chunk] = np.any(near & ((dx * dx + dy * dy) + dz * dz <= r2[None, :]), axis=1)
    keep = np.flatnonzero(~covered)
    return (len(x) + 1 + np.arange(len(keep))).astype(np.int32), cand[keep]


def ref_coverage_buffer(x, y, z, h, domain, spacing_fraction=0.05, factor=2.0):
    """The reference's own add_coverage_grid_buffer through oracle/_ref/libkiwi_pts.so (particle ids
    1..n).  Returns (ids, xyz); raises ValueError with the reference's message."""
    L = ctypes.CDLL(PTS_SO)
    L.ref_cartesian_lattice.argtypes = [_pd, _d, _pi, _pd]
    L.ref_coverage_buffer.restype = _ll
    L.ref_coverage_buffer.argtypes = [_ll, _pd, _pd, _pd, _pd, _pd, _d, _d, _pd, _pd, _pd, _pi, _ll, ctypes.c_char_p, _i]
    cols = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, h)]
    dom = np.ascontiguousarray(domain, dtype=np.float64)
    n3 = np.zeros(3, np.int32)
    sp = ctypes.c_double()
    cap = L.ref_cartesian_lattice(_p(dom), float(spacing_fraction), n3.ctypes.data_as(_pi), ctypes.byref(sp))
    o = [np.zeros(max(cap, 1)) for _ in range(3)]
    oid = np.zeros(max(cap, 1), np.int32)
    err = ctypes.create_string_buffer(512)
    k = L.ref_coverage_buffer(len(cols[0]), *[_p(a) for a in cols], _p(dom), float(spacing_fraction), float(factor),
                              *[_p(a) for a in o], oid.ctypes.data_as(_pi), cap, err, 512)
    if k < 0:
        raise ValueError(err.value.decode())
    return oid[:k].copy(), np.stack([a[:k] for a in o], 1)


def ref_load_snapshot(path, format="auto", hfact=1.2, cap=None, units=None, allow_missing_velocity=False):
    """The reference's own load_sph_snapshot through oracle/_ref/libkiwi_pts.so.  Returns a dict of
    arrays (id and the 9 columns); raises ValueError with the reference's message."""
    L = ctypes.CDLL(PTS_SO)
    L.ref_load_snapshot.restype = _ll
    L.ref_load_snapshot.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _d, _pd, _i, _pd, _pi, _ll, ctypes.c_char_p, _i]
    if cap is None:
        cap = (os.path.getsize(path) if os.path.exists(path) else 0) // 16 + 16  # a data row takes >= 16 bytes
    cols = np.zeros((9, cap))
    ids = np.zeros(cap, np.int32)
    err = ctypes.create_string_buffer(4096)
    u = None if units is None else np.ascontiguousarray(units, dtype=np.float64)
    n = L.ref_load_snapshot(os.fsencode(path), format.encode(), float(hfact), None if u is None else _p(u),
                            int(bool(allow_missing_velocity)), _p(cols), ids.ctypes.data_as(_pi), cap, err, 4096)
    if n < 0:
        raise ValueError(err.value.decode(errors="replace"))
    assert n <= cap
    out = {"id": ids[:n].copy()}
    for k, f in enumerate(("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")):
        out[f] = cols[k, :n].copy()
    return out


def ref_detect_format(path):
    L = ctypes.CDLL(PTS_SO)
    L.ref_detect_format.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _i]
    buf = ctypes.create_string_buffer(4096)
    if L.ref_detect_format(os.fsencode(path), buf, 4096) != 0:
        raise ValueError(buf.value.decode(errors="replace"))
    return buf.value.decode()


def nearest_site(sites, q, chunk=4096):
    """Index of the nearest of sites (n, 3) to every point of q (m, 3), brute force: (s - q)^2 summed x,
    y, z in double with no fused multiply-add, as voro++ forms it (v_compute.cc); the first minimum
    (lowest index) on exact ties."""
    best = np.empty(len(q), np.int64)
    for c0 in range(0, len(q), chunk):
        qq = q[c0:c0 + chunk]
        dx = sites[None, :, 0] - qq[:, None, 0]
        dy = sites[None, :, 1] - qq[:, None, 1]
        dz = sites[None, :, 2] - qq[:, None, 2]
        best[c0:c0 + chunk] = np.argmin((dx * dx + dy * dy) + dz * dz, axis=1)
    return best


def cartesian_map(sites, ids, rows, domain, res, mapped=None, chunk=4096):
    """The cartesian-map resample (src/cartesian_map.cpp:150-183) by brute force: every cell centre
    x_min + (i + 0.5) * (extent / res) (:157-164), in (ix, iy, iz) order with z fastest (:164-165), takes
    the row of its nearest site, i.e. of the Voronoi cell voro++'s find_voronoi_cell reports.  Distances
    are (s - q)^2 summed x, y, z in double with no fused multiply-add (voro++ v_compute.cc); on an exact
    tie the lowest site index wins.  Sites outside [min, max] on an axis are dropped, as voro++'s
    pre_container::put drops them (:67-69).  sites: (n, 3); rows: (n, 4) density, vx, vy, vz; mapped:
    optional (n,) bool, False = the site's id has no mapping.dat row (:175-178: id kept, row zero).
    Returns (rows (res^3, 4), cell_id (res^3,), unresolved, unmapped)."""
    d = np.asarray(domain, dtype=np.float64)
    sites = np.asarray(sites, dtype=np.float64).reshape(-1, 3)
    keep = np.all((sites >= d[0::2]) & (sites <= d[1::2]), axis=1)
    ids, rows = np.asarray(ids)[keep], np.asarray(rows, dtype=np.float64).reshape(-1, 4)[keep]
    mapped = None if mapped is None else np.asarray(mapped, bool)[keep]
    sites = sites[keep]
    ax = [d[2 * a] + (np.arange(res) + 0.5) * ((d[2 * a + 1] - d[2 * a]) / res) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    q = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    nc = len(q)
    out = np.zeros((nc, 4))
    cid = np.zeros(nc, np.int32)
    if len(sites) == 0:
        return out, cid, nc, 0
    best = nearest_site(sites, q, chunk)
    cid[:] = ids[best]
    ok = np.ones(nc, bool) if mapped is None else mapped[best]
    out[ok] = rows[best[ok]]
    return out, cid, 0, int((~ok).sum())


def inverse_mapping(parts, pid, home, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr, cell_mass, k=None):
    """The inverse-mapping test (src/mapping.cpp:540-701) per particle: (m_recovered, coverage, counts),
    with cell_mass the forward map's density x volume per cell."""
    k = k or kernel()
    n = len(parts["x"])
    cols = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in ("x", "y", "z", "m", "h")]
    pid = np.ascontiguousarray(pid, dtype=np.int32)
    home = np.ascontiguousarray(home, dtype=np.int32)
    cfp = np.ascontiguousarray(cell_face_ptr, dtype=np.int64)
    fvp = np.ascontiguousarray(face_vert_ptr, dtype=np.int64)
    fxyz = np.ascontiguousarray(face_xyz, dtype=np.float64)
    fnb = np.ascontiguousarray(face_nbr, dtype=np.int32)
    cm = np.ascontiguousarray(cell_mass, dtype=np.float64)
    rec, cov = np.zeros(n), np.zeros(n)
    counts = np.zeros(6, dtype=np.int64)
    rc = lib().oracle_inverse_mapping(n, *[_p(c) for c in cols], pid.ctypes.data_as(_pi), home.ctypes.data_as(_pi),
                                      len(cfp) - 1, cfp.ctypes.data_as(_pll), fvp.ctypes.data_as(_pll), _p(fxyz),
                                      fnb.ctypes.data_as(_pi), ctypes.byref(k), _p(cm), _p(rec), _p(cov),
                                      counts.ctypes.data_as(_pll))
    if rc != 0:
        raise ValueError(f"oracle_inverse_mapping failed ({rc})")
    return rec, cov, counts


def voronoi_map_cells_quad(parts, pid, home, cell_face_ptr, face_vert_ptr, face_xyz, face_nbr, cells, k=None,
                           share_lower=True):
    """The Voronoi map's (mass, px, py, pz) for the listed cells with the face integrals in binary128 (the
    visits decided by the double-precision BFS).  share_lower: True (1) -- a face shared with a visited
    lower-index neighbour takes the negated integral of the neighbour's copy (the GPU map's rule); False
    (0) -- every cell uses its own copies; 2 -- the reference's face cache (a neighbour the particle's BFS
    processed earlier lends its copy).  Returns an array of shape (len(cells), 4)."""
    k = k or kernel()
    n = len(parts["x"])
    cols = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in ("x", "y", "z", "vx", "vy", "vz", "m", "h")]
    pid = np.ascontiguousarray(pid, dtype=np.int32)
    home = np.ascontiguousarray(home, dtype=np.int32)
    cfp = np.ascontiguousarray(cell_face_ptr, dtype=np.int64)
    fvp = np.ascontiguousarray(face_vert_ptr, dtype=np.int64)
    fxyz = np.ascontiguousarray(face_xyz, dtype=np.float64)
    fnb = np.ascontiguousarray(face_nbr, dtype=np.int32)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    out = np.zeros((len(cells), 4))
    rc = lib().oracle_voronoi_map_cells_q(n, *[_p(c) for c in cols], pid.ctypes.data_as(_pi),
                                          home.ctypes.data_as(_pi), len(cfp) - 1, cfp.ctypes.data_as(_pll),
                                          fvp.ctypes.data_as(_pll), _p(fxyz), fnb.ctypes.data_as(_pi),
                                          ctypes.byref(k), len(cells), cells.ctypes.data_as(_pll),
                                          int(share_lower), _p(out))
    if rc != 0:
        raise ValueError("oracle_voronoi_map_cells_q failed")
    return out
