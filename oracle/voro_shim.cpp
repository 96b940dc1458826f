// C-ABI shim over the reference's vendored voro++ (third_party/voro), linked only into
// oracle/_ref/libkiwi_voro.so.  Test infrastructure: it tessellates a point set the way the reference's
// mapping stage does (non-periodic container over the domain box; src/mapping.cpp:222-269), flattens
// every cell's faces to plain CSR arrays with absolute vertex coordinates (f_coord + site, as
// compute_cell_integral forms them, src/mapping.cpp:168-175) and neighbour cell indices, and finds each
// particle's home cell with find_voronoi_cell (src/mapping.cpp:434-440); each cell's centroid gives
// the Lloyd step (lloyd_step, src/point_set.cpp:275-306).  The Voronoi golden fixture and the Voronoi
// test and measurement inputs are written from it.  Test infrastructure: the library is git-ignored but
// travels with the working tree to the GPU machine, where the Voronoi, inverse-mapping and cartesian-map
// GPU tests load it on the host to tessellate their inputs; libkiwi_hip never links or calls it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "voro++.hh"

namespace {

struct Geometry {
    std::vector<int> cell_id;
    std::vector<double> site, volume, centroid;
    std::vector<long long> cell_face_ptr{0}, face_vert_ptr{0};
    std::vector<double> vert_xyz;
    std::vector<int> face_nbr, home;
};

}  // namespace

extern "C" {

void* voro_build(int npts, const int* id, const double* x, const double* y, const double* z, const double* dom,
                 long long npart, const double* px, const double* py, const double* pz, long long* sizes) {
    using namespace voro;
    pre_container pcon(dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], false, false, false);
    for (int i = 0; i < npts; ++i) pcon.put(id[i], x[i], y[i], z[i]);
    // the reference's block counts: voro++'s guess scaled by the default grid_blocks.block_scale_factor
    // of 3 (guess_scaled_blocks, include/kiwi/voro_grid.hpp:18-26); the blocks set the cells' plane
    // tolerance, so near-degenerate vertices resolve as in the reference's own tessellation
    int nx, ny, nz;
    pcon.guess_optimal(nx, ny, nz);
    nx = std::max(1, static_cast<int>(std::round(nx * 3.0)));
    ny = std::max(1, static_cast<int>(std::round(ny * 3.0)));
    nz = std::max(1, static_cast<int>(std::round(nz * 3.0)));
    container con(dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], nx, ny, nz, false, false, false, 8);
    pcon.setup(con);

    Geometry* g = new Geometry();
    int max_id = 0;
    for (int i = 0; i < npts; ++i) max_id = std::max(max_id, id[i]);
    std::vector<int> reverse(static_cast<size_t>(max_id) + 1, -1);
    std::vector<std::vector<int>> f_vert_all, n_list_all;
    std::vector<std::vector<double>> f_coord_all;
    std::vector<double> sx, sy, sz;

    voronoicell_neighbor c;
    c_loop_all clo(con);
    double x0, y0, z0;
    if (clo.start()) do {
            if (con.compute_cell(c, clo)) {
                reverse[clo.pid()] = static_cast<int>(g->cell_id.size());
                g->cell_id.push_back(clo.pid());
                clo.pos(x0, y0, z0);
                sx.push_back(x0);
                sy.push_back(y0);
                sz.push_back(z0);
                std::vector<int> fv, nl;
                std::vector<double> fc;
                c.face_vertices(fv);
                c.vertices(fc);
                c.neighbors(nl);
                f_vert_all.push_back(fv);
                f_coord_all.push_back(fc);
                n_list_all.push_back(nl);
                g->volume.push_back(c.volume());
                double cx, cy, cz;
                c.centroid(cx, cy, cz);
                g->centroid.push_back(x0 + cx);
                g->centroid.push_back(y0 + cy);
                g->centroid.push_back(z0 + cz);
            }
        } while (clo.inc());

    for (size_t ci = 0; ci < g->cell_id.size(); ++ci) {
        g->site.push_back(sx[ci]);
        g->site.push_back(sy[ci]);
        g->site.push_back(sz[ci]);
        const std::vector<int>& fv = f_vert_all[ci];
        const std::vector<double>& fc = f_coord_all[ci];
        size_t it = 0;
        int face = 0;
        while (it < fv.size()) {
            const int cnt = fv[it];
            for (int k = 0; k < cnt; ++k) {
                const int vi = fv[it + 1 + k];
                g->vert_xyz.push_back(fc[3 * vi] + sx[ci]);
                g->vert_xyz.push_back(fc[3 * vi + 1] + sy[ci]);
                g->vert_xyz.push_back(fc[3 * vi + 2] + sz[ci]);
            }
            g->face_vert_ptr.push_back(g->face_vert_ptr.back() + cnt);
            const int nb = n_list_all[ci][face];
            g->face_nbr.push_back((nb >= 0 && nb <= max_id) ? reverse[nb] : -1);
            it += cnt + 1;
            ++face;
        }
        g->cell_face_ptr.push_back(static_cast<long long>(g->face_nbr.size()));
    }
    for (long long i = 0; i < npart; ++i) {
        double rx, ry, rz;
        int pid = -1;
        const bool found = con.find_voronoi_cell(px[i], py[i], pz[i], rx, ry, rz, pid);
        g->home.push_back((found && pid >= 0 && pid <= max_id) ? reverse[pid] : -1);
    }
    sizes[0] = static_cast<long long>(g->cell_id.size());
    sizes[1] = static_cast<long long>(g->face_nbr.size());
    sizes[2] = static_cast<long long>(g->vert_xyz.size() / 3);
    return g;
}

void voro_fetch(void* h, int* cell_id, double* site, double* volume, double* centroid, long long* cell_face_ptr,
                long long* face_vert_ptr, double* vert_xyz, int* face_nbr, int* home) {
    const Geometry* g = static_cast<const Geometry*>(h);
    std::memcpy(cell_id, g->cell_id.data(), sizeof(int) * g->cell_id.size());
    std::memcpy(site, g->site.data(), sizeof(double) * g->site.size());
    std::memcpy(volume, g->volume.data(), sizeof(double) * g->volume.size());
    std::memcpy(centroid, g->centroid.data(), sizeof(double) * g->centroid.size());
    std::memcpy(cell_face_ptr, g->cell_face_ptr.data(), sizeof(long long) * g->cell_face_ptr.size());
    std::memcpy(face_vert_ptr, g->face_vert_ptr.data(), sizeof(long long) * g->face_vert_ptr.size());
    std::memcpy(vert_xyz, g->vert_xyz.data(), sizeof(double) * g->vert_xyz.size());
    std::memcpy(face_nbr, g->face_nbr.data(), sizeof(int) * g->face_nbr.size());
    std::memcpy(home, g->home.data(), sizeof(int) * g->home.size());
}

void voro_free(void* h) { delete static_cast<Geometry*>(h); }

}  // extern "C"
