// Host build of the GPU tessellation's per-cell code (kiwi_amd/csrc/voro_cell.h) for the CPU tests: the
// same cell, bucket-grid and block-order functions the device kernels call, with the device's workspace
// layout (every cell in its own slot of one strided workspace: computed and saved, then reloaded and
// written out), driven by a plain loop, so that the geometry can be checked against voro++ without a
// GPU.  Test infrastructure, built by
// tests/test_voro_cell_host.py with g++ into a temporary directory.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "voro_cell.h"

using namespace kiwi_hip;

namespace {
struct Geometry {
    std::vector<int> cell_site;
    std::vector<int64_t> cfp{0}, fvp{0};
    std::vector<double> xyz;
    std::vector<int> nbr;
};
}  // namespace

extern "C" {

// returns 0, or the first failing cut's code; sizes: n_cells, n_faces, n_face_vertices
int vc_host_tessellate(int n, const double* x, const double* y, const double* z, const double* dom, long long* sizes,
                       void** handle) {
    Geometry* G = new Geometry();
    *handle = G;
    int64_t n_in = 0;
    for (int i = 0; i < n; ++i)
        n_in += x[i] >= dom[0] && x[i] <= dom[1] && y[i] >= dom[2] && y[i] <= dom[3] && z[i] >= dom[4] && z[i] <= dom[5];
    int nbk[3];
    vc_block_dims(dom, n_in, nbk);
    std::vector<int64_t> blk(n);
    for (int i = 0; i < n; ++i) blk[i] = vc_block_index(dom, nbk, x[i], y[i], z[i]);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const uint64_t ka = static_cast<uint64_t>(blk[a]), kb = static_cast<uint64_t>(blk[b]);  // -1 last
        return ka < kb;
    });
    std::vector<int> site_cell(n, -1);
    for (int i : order)
        if (blk[i] >= 0) {
            site_cell[i] = static_cast<int>(G->cell_site.size());
            G->cell_site.push_back(i);
        }
    const int nc = static_cast<int>(G->cell_site.size());
    const VcGrid g = vc_grid_make(dom, nc);
    const int64_t nb = static_cast<int64_t>(g.nb[0]) * g.nb[1] * g.nb[2];
    std::vector<int64_t> bkey(nc);
    for (int c = 0; c < nc; ++c) {
        const int i = G->cell_site[c];
        bkey[c] = (static_cast<int64_t>(vc_bucket_axis(g, 0, x[i])) * g.nb[1] + vc_bucket_axis(g, 1, y[i])) * g.nb[2] +
                  vc_bucket_axis(g, 2, z[i]);
    }
    std::vector<int> bo(nc);
    std::iota(bo.begin(), bo.end(), 0);
    std::stable_sort(bo.begin(), bo.end(), [&](int a, int b) { return bkey[a] < bkey[b]; });
    std::vector<VcSite> sites(nc);
    std::vector<int> range(2 * nb, 0);
    for (int t = 0; t < nc; ++t) {
        const int i = G->cell_site[bo[t]];
        sites[t] = VcSite{x[i], y[i], z[i], i, 0};
        const int64_t b = bkey[bo[t]];
        if (t == 0 || bkey[bo[t - 1]] != b) range[2 * b] = t;
        if (t == nc - 1 || bkey[bo[t + 1]] != b) range[2 * b + 1] = t + 1;
    }
    // the device layout: every cell in its own slot of one strided workspace, computed, saved, reloaded
    std::vector<char> store(static_cast<size_t>(VC_BYTES) * (nc > 0 ? nc : 1) + 64);
    VCell cell;
    int status = 0;
    for (int c = 0; c < nc; ++c) {
        const int i = G->cell_site[c];
        vc_bind(cell, store.data(), nc, c);
        const int rc = vc_compute(cell, dom, g, range.data(), sites.data(), i, x[i], y[i], z[i], VC_TOL_DEFAULT);
        if (rc != VC_CUT && !status) status = rc;
        vc_save(cell);
    }
    for (int c = 0; c < nc; ++c) {
        const int i = G->cell_site[c];
        vc_bind(cell, store.data(), nc, c);
        vc_load(cell);
        const int64_t f0 = G->cfp.back(), v0 = G->fvp.back();
        const int nfv = vc_face_vertices(cell);
        G->fvp.resize(f0 + cell.nf + 1);
        G->nbr.resize(f0 + cell.nf);
        G->xyz.resize(3 * (v0 + nfv));
        vc_write(cell, x[i], y[i], z[i], site_cell.data(), f0, v0, G->fvp.data(), G->xyz.data(), G->nbr.data());
        G->fvp[f0 + cell.nf] = v0 + nfv;
        G->cfp.push_back(f0 + cell.nf);
    }
    sizes[0] = nc;
    sizes[1] = static_cast<long long>(G->nbr.size());
    sizes[2] = static_cast<long long>(G->xyz.size() / 3);
    return status;
}

void vc_host_fetch(void* h, int* cell_site, int64_t* cfp, int64_t* fvp, double* xyz, int* nbr) {
    const Geometry* G = static_cast<const Geometry*>(h);
    std::copy(G->cell_site.begin(), G->cell_site.end(), cell_site);
    std::copy(G->cfp.begin(), G->cfp.end(), cfp);
    std::copy(G->fvp.begin(), G->fvp.end(), fvp);
    std::copy(G->xyz.begin(), G->xyz.end(), xyz);
    std::copy(G->nbr.begin(), G->nbr.end(), nbr);
}

void vc_host_free(void* h) { delete static_cast<Geometry*>(h); }

}  // extern "C"
