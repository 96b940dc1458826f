"""The GPU tessellation's per-cell code (kiwi_amd/csrc/voro_cell.h), built for the host with g++, against
voro++ (the reference's vendored copy, through oracle/_ref/libkiwi_voro.so): the tessellation the
reference builds in run_tessellation_stage / build_cell_geometry (src/tessellation.cpp:23-61,
src/mapping.cpp:222-269).  Same cell order (voro++'s blocks), same neighbours per cell, the same face
cycles up to their starting vertex (vertices within 1e-12 of the domain extent), and the lattice's exact
cubes (tools/test_lattice_grid.cpp:95-198).  CPU only."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from geometry_check import assert_same_cells, assert_same_cycles, volumes  # noqa: E402

_pd = ctypes.POINTER(ctypes.c_double)
_pll = ctypes.POINTER(ctypes.c_int64)
_pi = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("vc") / "libvc_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "kiwi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "voro_cell_host.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.vc_host_tessellate.restype = ctypes.c_int
    L.vc_host_tessellate.argtypes = [ctypes.c_int, _pd, _pd, _pd, _pd, _pll, ctypes.POINTER(ctypes.c_void_p)]
    L.vc_host_fetch.argtypes = [ctypes.c_void_p, _pi, _pll, _pll, _pd, _pi]
    L.vc_host_free.argtypes = [ctypes.c_void_p]
    return L


def _p(a):
    return a.ctypes.data_as(_pd)


def host_tessellate(L, pts, dom):
    P = [np.ascontiguousarray(pts[:, k], dtype=np.float64) for k in range(3)]
    dom = np.ascontiguousarray(dom, dtype=np.float64)
    sizes = np.zeros(3, np.int64)
    h = ctypes.c_void_p()
    rc = L.vc_host_tessellate(len(pts), *[_p(a) for a in P], _p(dom), sizes.ctypes.data_as(_pll), ctypes.byref(h))
    nc, nf, nv = (int(v) for v in sizes)
    g = {"cell_site": np.zeros(nc, np.int32), "cell_face_ptr": np.zeros(nc + 1, np.int64),
         "face_vert_ptr": np.zeros(nf + 1, np.int64), "vert_xyz": np.zeros((nv, 3)), "face_nbr": np.zeros(nf, np.int32)}
    L.vc_host_fetch(h, g["cell_site"].ctypes.data_as(_pi), g["cell_face_ptr"].ctypes.data_as(_pll),
                    g["face_vert_ptr"].ctypes.data_as(_pll), _p(g["vert_xyz"]), g["face_nbr"].ctypes.data_as(_pi))
    L.vc_host_free(h)
    assert rc == 0, f"cell computation failed ({rc})"
    return g


def run_case(L, pts, dom, check_cycles=True):
    g = host_tessellate(L, pts, dom)
    r = O.tessellate(np.arange(len(pts)), pts, dom, pts[:1])
    ext = max(abs(v) for v in dom)
    assert_same_cells(g, r)
    if check_cycles:
        assert_same_cycles(g, r, 1e-12 * ext)
    V = volumes(g, pts[g["cell_site"]])
    np.testing.assert_allclose(V, r["volume"], rtol=0, atol=1e-12 * ext ** 3)
    return g, r


def test_uniform_points(host_lib):
    rng = np.random.default_rng(1)
    run_case(host_lib, rng.random((2000, 3)), [0, 1, 0, 1, 0, 1])


def test_clustered_points_offset_box(host_lib):
    rng = np.random.default_rng(2)
    n = 3000
    pts = np.column_stack([rng.uniform(-3, 5, n), rng.uniform(0, 2, n), rng.uniform(1, 4, n)])
    pts[:400] = pts[:400] * 0.05 + [0.0, 1.0, 2.0]
    run_case(host_lib, pts, [-3, 5, 0, 2, 1, 4])


def test_lattice_cubes(host_lib):
    """The Voronoi cells of a cubic lattice are its cubes, with 6 faces (tools/test_lattice_grid.cpp)."""
    ax = (np.arange(6) + 0.5) / 6
    pts = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T.copy()
    g, _ = run_case(host_lib, pts, [0, 1, 0, 1, 0, 1])
    assert np.all(np.diff(g["cell_face_ptr"]) == 6)
    assert np.all(np.diff(g["face_vert_ptr"]) == 4)
    np.testing.assert_allclose(volumes(g, pts[g["cell_site"]]), 1 / 216, rtol=1e-13)


def test_many_faced_cell(host_lib):
    """A site ringed by 150 sites on a sphere: one cell of more than 64 faces."""
    k = np.arange(150) + 0.5
    phi = np.arccos(1 - 2 * k / 150)
    th = np.pi * (1 + 5 ** 0.5) * k
    ring = 0.5 + 0.3 * np.column_stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)])
    rng = np.random.default_rng(3)
    outer = rng.random((149, 3))
    outer = outer[np.linalg.norm(outer - 0.5, axis=1) > 0.45]
    pts = np.vstack([[[0.5, 0.5, 0.5]], ring, outer])
    g, _ = run_case(host_lib, pts, [0, 1, 0, 1, 0, 1])
    assert np.diff(g["cell_face_ptr"]).max() > 64


def test_dropped_and_boundary_sites(host_lib):
    """Sites outside the box are dropped, sites on its faces and corners kept or dropped as voro++'s
    container keeps them (its block index on a max face), and the cell order follows."""
    rng = np.random.default_rng(4)
    pts = rng.random((400, 3)) * 2.0
    pts[:20, 0] = 2.0
    pts[20:30, 1] = 0.0
    pts[30:35] = [[2.0, 2.0, 2.0], [0.0, 0.0, 0.0], [2.0, 0.0, 2.0], [0.0, 2.0, 0.0], [1.0, 2.0, 1.0]]
    pts[35:60] = rng.uniform(-1.0, 3.0, (25, 3))
    run_case(host_lib, pts, [0, 2, 0, 2, 0, 2])


def test_few_sites(host_lib):
    for n in (1, 2, 5):
        rng = np.random.default_rng(10 + n)
        run_case(host_lib, rng.random((n, 3)), [0, 1, 0, 1, 0, 1])
