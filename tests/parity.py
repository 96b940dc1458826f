"""Parity rules for FP64 cell sums against the reference's double-precision result.

Every field must agree with the reference to `tol` of the field's maximum.  Where it does not, the cell
is recomputed with the reference's own algorithm in binary128 (oracle.sph_to_grid_cells_quad): the HIP
value must then lie within `quad_tol` of that, and closer to it than the reference's double value.  The
reference's face walk takes acos of arguments within rounding of 1 for particles close to a grid plane
(src/kernel_integral.cpp:593-594), which costs it up to ~1e-9 of a cell's mass there.
"""
import numpy as np

from oracle import oracle as O


def assert_cells_match(gpu4, ref4, parts, domain, res, tol=1e-10, quad_tol=1e-11, max_disputed=5000, k=None):
    gpu4 = [np.asarray(a) for a in gpu4]
    ref4 = [np.asarray(a) for a in ref4]
    scale = [max(np.max(np.abs(r)), 1e-300) for r in ref4]
    bad = np.zeros(len(ref4[0]), dtype=bool)
    for g, r, s in zip(gpu4, ref4, scale):
        bad |= np.abs(g - r) > tol * s
    cells = np.nonzero(bad)[0]
    if len(cells) == 0:
        return 0
    assert len(cells) <= max_disputed, f"{len(cells)} cells outside {tol:g} of the reference"
    q = O.sph_to_grid_cells_quad(parts, domain, *res, cells, k=k)
    for f in range(4):
        g = gpu4[f][cells]
        r = ref4[f][cells]
        eg = np.abs(g - q[:, f])
        er = np.abs(r - q[:, f])
        assert np.all(eg <= quad_tol * scale[f]), (f, cells[np.argmax(eg)], eg.max() / scale[f])
        worse = (eg > er) & (np.abs(g - r) > tol * scale[f])
        assert not np.any(worse), (f, cells[worse][:5])
    return len(cells)


def assert_voronoi_cells_match(gpu4, ref4, parts, pid, home, cells, tol=1e-12, quad_tol=1e-11, max_disputed=5000,
                               k=None):
    """The rule for the Voronoi map: every field within `tol` of the field maximum of the reference's
    double-precision result, or else, for that cell, within `quad_tol` of the cell recomputed with the
    face integrals in binary128 (oracle.voronoi_map_cells_quad, closing each cell with the same copy of a
    shared face as the GPU does).

    Two effects separate the double results from binary128.  (1) The reference's analytic cubic vertex
    integral sums an I-series (src/kernel_integral.cpp) whose double value is off by up to ~1e-10 of a
    cell's mass for larger smoothing lengths; in binary128 it agrees with the series-free generic path
    to ~1e-15 (test_oracle.py).  (2) A face shared by two visited cells is integrated once and enters
    the other cell negated (the reference's face cache, src/mapping.cpp:87-159; the GPU does the same
    with the lower cell index as owner).  voro++ computes each cell's vertices on its own, so the two
    copies of a face differ by rounding and a cell closed with its neighbour's copy is off by ~5e-16 of
    a particle's mass per pair -- more where voro++'s plane tolerance, which scales with the domain's
    extent, snaps a vertex differently in the two cells.  Both the reference and the GPU carry (2), each
    on the faces its owner rule assigns, so the HIP value is not required to be closer to binary128 than
    the reference's; the binary128 evaluation closes cells with the GPU's copies."""
    gpu4 = [np.asarray(a) for a in gpu4]
    ref4 = [np.asarray(a) for a in ref4]
    scale = [max(np.max(np.abs(r)), 1e-300) for r in ref4]
    bad = np.zeros(len(ref4[0]), dtype=bool)
    for g, r, s in zip(gpu4, ref4, scale):
        bad |= np.abs(g - r) > tol * s
    disputed = np.nonzero(bad)[0]
    if len(disputed) == 0:
        return 0
    assert len(disputed) <= max_disputed, f"{len(disputed)} cells outside {tol:g} of the reference"
    q = O.voronoi_map_cells_quad(parts, pid, home, cells["cell_face_ptr"], cells["face_vert_ptr"],
                                 cells["vert_xyz"], cells["face_nbr"], disputed, k=k)
    for f in range(4):
        eg = np.abs(gpu4[f][disputed] - q[:, f])
        assert np.all(eg <= quad_tol * scale[f]), (f, disputed[np.argmax(eg)], eg.max() / scale[f])
    return len(disputed)
