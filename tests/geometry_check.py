"""Comparisons of a flattened Voronoi tessellation (CSR: cell_face_ptr, face_vert_ptr, vert_xyz, face_nbr,
cell_site) with voro++'s (oracle.tessellate: cell_id, the same CSR arrays, volume)."""
import numpy as np


def volumes(g, sites):
    """Per cell: volume from the fan tetrahedra of the faces about the site (faces clockwise seen from
    outside, voro++'s face_vertices orientation, give positive volumes)."""
    cfp, fvp, xyz = (np.asarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "vert_xyz"))
    xyz = xyz.reshape(-1, 3)
    nc = len(cfp) - 1
    V = np.zeros(nc)
    for c in range(nc):
        for f in range(cfp[c], cfp[c + 1]):
            v = xyz[fvp[f]:fvp[f + 1]] - sites[c]
            V[c] -= np.einsum("ij,ij->i", np.broadcast_to(v[0], v[1:-1].shape), np.cross(v[1:-1], v[2:])).sum() / 6
    return V


def face_areas(g):
    """Area of every face: half the norm of the sum of v_k x v_(k+1) around its cycle."""
    fvp = np.asarray(g["face_vert_ptr"])
    xyz = np.asarray(g["vert_xyz"]).reshape(-1, 3)
    nf = len(fvp) - 1
    if nf == 0:
        return np.zeros(0)
    nxt = np.arange(len(xyz)) + 1
    nxt[fvp[1:] - 1] = fvp[:-1]  # the last vertex of a face wraps to its first
    cr = np.cross(xyz, xyz[nxt])
    s = np.add.reduceat(cr, fvp[:-1], axis=0)
    return 0.5 * np.linalg.norm(s, axis=1)


def neighbour_lists(g, min_area=None):
    """(cell of each face, neighbour of each face), sorted by cell then neighbour; only the faces of area
    above min_area if given."""
    cfp, nbr = np.asarray(g["cell_face_ptr"]), np.asarray(g["face_nbr"])
    cell = np.repeat(np.arange(len(cfp) - 1), np.diff(cfp))
    if min_area is not None:
        keep = face_areas(g) > min_area
        cell, nbr = cell[keep], nbr[keep]
    o = np.lexsort((nbr, cell))
    return cell[o], nbr[o]


def assert_same_cells(g, r, min_area=None):
    """Cell order and every cell's neighbour list (vectorised: any size).  With min_area, only the faces
    larger than it on either side must match: where four or more sites are cospherical within voro++'s
    plane tolerance, a face of area near 0 (a few 1e-14 of a unit box) exists or not by the rounding of
    the vertices and the order of the cuts, in voro++ as here."""
    np.testing.assert_array_equal(np.asarray(g["cell_site"]), r["cell_id"])
    if min_area is None:
        np.testing.assert_array_equal(np.diff(np.asarray(g["cell_face_ptr"])), np.diff(r["cell_face_ptr"]))
    a, b = neighbour_lists(g, min_area), neighbour_lists(r, min_area)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def assert_same_cycles(g, r, tol):
    """Every face's vertex cycle equals voro++'s up to its starting vertex, within tol per coordinate
    (faces matched by neighbour and vertex count)."""
    cfp, fvp, nbr = (np.asarray(g[k]) for k in ("cell_face_ptr", "face_vert_ptr", "face_nbr"))
    xyz = np.asarray(g["vert_xyz"]).reshape(-1, 3)
    rxyz = r["vert_xyz"].reshape(-1, 3)
    for c in range(len(cfp) - 1):
        fb = range(r["cell_face_ptr"][c], r["cell_face_ptr"][c + 1])
        B = [rxyz[r["face_vert_ptr"][f]:r["face_vert_ptr"][f + 1]] for f in fb]
        used = set()
        for f in range(cfp[c], cfp[c + 1]):
            A = xyz[fvp[f]:fvp[f + 1]]
            hit = False
            for j, f2 in enumerate(fb):
                if j in used or r["face_nbr"][f2] != nbr[f] or len(B[j]) != len(A):
                    continue
                if any(np.max(np.abs(np.roll(A, -s, axis=0) - B[j])) < tol for s in range(len(A))):
                    used.add(j)
                    hit = True
                    break
            assert hit, (c, f)


VORO_TOL = 10 * 2.220446049250313e-16 * 1e6  # voro++'s plane tolerance for default-constructed cells


def voro_tolerance_bounds(sites, r):
    """Per cell, how far voro++'s plane tolerance lets its geometry move: a vertex within VORO_TOL of
    2 p.n - |n|^2 = 0 counts as on the plane, i.e. up to delta = VORO_TOL / (2 |n|) from it, |n| at
    least the distance to the site's nearest neighbour.  Two correct cut sequences can then differ by
    about delta in the centroid and delta x surface area in the volume (for the smallest cells of a
    dense cluster, 1e-5 of the volume).  Returns (volume bound, centroid bound), twice those figures."""
    from scipy.spatial import cKDTree
    nn = cKDTree(sites).query(sites, k=2)[0][:, 1] if len(sites) > 1 else np.full(len(sites), np.inf)
    delta = VORO_TOL / (2 * nn)
    area = np.add.reduceat(face_areas(r), r["cell_face_ptr"][:-1]) if len(sites) else np.zeros(0)
    return 2 * delta * area, 2 * delta
