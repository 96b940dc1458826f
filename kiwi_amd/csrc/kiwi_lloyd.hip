// The Lloyd step of the point set on the GPU (kiwi_hip_voronoi_centroids): each Voronoi cell's
// centroid -- its site's next position (lloyd_step, src/point_set.cpp:275-302) -- and its volume, from
// the flattened tessellation the map consumes.
//
// The reference asks voro++ for every cell's centroid (voronoicell_base::centroid, third_party/voro/src/
// cell.cc:1861-1901), which cuts the cell into tetrahedra from its vertex 0 over a fan of every face.
// Here one thread per cell cuts it into tetrahedra from the site (inside its own cell), over the fan of
// every face from the face's first vertex: for a face with vertices v_0 .. v_{k-1}, the tetrahedra
// (s, v_0, v_j, v_{j+1}), j = 1 .. k-2, with 6 x signed volume t = (v_0 - s) . ((v_j - s) x (v_{j+1} - s))
// and centroid s + (u + v + w) / 4.  In exact arithmetic both decompositions give the same volume and
// centroid; in double they differ by rounding (about 1e-16 of the cell's extent).  The faces' winding is
// the same over the cell, so the sum of t has one sign; the volume is |sum t| / 6 and the centroid
// s + sum (u + v + w) t / (4 sum t).  A cell of zero volume keeps its site, as voro++ returns the
// cell-relative centroid (0, 0, 0) below its volume tolerance.
//
// FP64, fixed summation order: bit-reproducible.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"

namespace {

__global__ __launch_bounds__(THREADS) void k_centroids(int64_t nc, int64_t nf, int64_t nv,
                                                       const int64_t* __restrict__ cfp,
                                                       const int64_t* __restrict__ fvp,
                                                       const double* __restrict__ vxyz,
                                                       const double* __restrict__ site,
                                                       double* __restrict__ cen, double* __restrict__ vol,
                                                       unsigned long long* bad) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (c >= nc) return;
    const double sx = site[3 * c], sy = site[3 * c + 1], sz = site[3 * c + 2];
    const int64_t f0 = cfp[c], f1 = cfp[c + 1];
    double T = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    if (!(f0 >= 0 && f0 <= f1 && f1 <= nf)) {
        atomicAdd(bad, 1ull);
    } else {
        for (int64_t f = f0; f < f1; ++f) {
            const int64_t v0 = fvp[f], v1 = fvp[f + 1];
            if (!(v0 >= 0 && v0 <= v1 && v1 <= nv)) {
                atomicAdd(bad, 1ull);
                break;
            }
            if (v1 - v0 < 3) continue;
            const double ux = vxyz[3 * v0] - sx, uy = vxyz[3 * v0 + 1] - sy, uz = vxyz[3 * v0 + 2] - sz;
            double vx = vxyz[3 * v0 + 3] - sx, vy = vxyz[3 * v0 + 4] - sy, vz = vxyz[3 * v0 + 5] - sz;
            for (int64_t j = v0 + 2; j < v1; ++j) {
                const double wx = vxyz[3 * j] - sx, wy = vxyz[3 * j + 1] - sy, wz = vxyz[3 * j + 2] - sz;
                const double t = ux * (vy * wz - vz * wy) + uy * (vz * wx - vx * wz) + uz * (vx * wy - vy * wx);
                T += t;
                cx += (ux + vx + wx) * t;
                cy += (uy + vy + wy) * t;
                cz += (uz + vz + wz) * t;
                vx = wx;
                vy = wy;
                vz = wz;
            }
        }
    }
    if (T != 0.0) {
        const double inv = 0.25 / T;
        cen[3 * c] = sx + cx * inv;
        cen[3 * c + 1] = sy + cy * inv;
        cen[3 * c + 2] = sz + cz * inv;
    } else {
        cen[3 * c] = sx;
        cen[3 * c + 1] = sy;
        cen[3 * c + 2] = sz;
    }
    if (vol) vol[c] = fabs(T) * (1.0 / 6.0);
}

}  // namespace

extern "C" {

int kiwi_hip_voronoi_centroids(kiwi_hip_ctx* c, const kiwi_voronoi_cells* V, const double* site_xyz,
                               double* centroid_xyz, double* volume) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!V) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (V->n_cells < 0 || V->n_faces < 0 || V->n_face_vertices < 0)
        return fail(c, KIWI_ERR_INVALID_ARG, "negative tessellation size");
    const int64_t nc = V->n_cells, nf = V->n_faces, nv = V->n_face_vertices;
    if (nc == 0) return KIWI_OK;
    if (!V->cell_face_ptr || !site_xyz || !centroid_xyz || (nf > 0 && !V->face_vert_ptr) ||
        (nv > 0 && !V->vert_xyz))
        return fail(c, KIWI_ERR_INVALID_ARG, "null tessellation, site or centroid array");
    if (nc > 0x7fffffffLL || nf > 0x7fffffffLL)
        return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 cells or faces in one call");
    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    if (V->on_device) {  // every array a kernel reads or writes: device memory, as on_device says
        const void* need[6] = {V->cell_face_ptr, nf > 0 ? static_cast<const void*>(V->face_vert_ptr) : site_xyz,
                               nv > 0 ? V->vert_xyz : site_xyz, site_xyz, centroid_xyz,
                               volume ? volume : centroid_xyz};
        for (const void* p : need)
            if (!gpu_accessible(p))
                return fail(c, KIWI_ERR_INVALID_ARG, "cells->on_device is set but an array is host memory");
    }
    hipStream_t s = c->stream;
    const int64_t* cfp = V->cell_face_ptr;
    const int64_t* fvp = V->face_vert_ptr;
    const double* vxyz = V->vert_xyz;
    const double* site = site_xyz;
    double* cen = centroid_xyz;
    double* vol = volume;
    if (!V->on_device) {  // stage the geometry and the sites; results come back to the caller's arrays
        const size_t bytes = sizeof(int64_t) * (nc + 1) + sizeof(int64_t) * (nf + 1) + sizeof(double) * 3 * nv +
                             sizeof(double) * (3 * nc + 3 * nc + nc);
        KIWI_HIP_CHECK(c, reserve(c->vor[0], bytes));
        char* b = static_cast<char*>(c->vor[0].p);
        int64_t* d_cfp = reinterpret_cast<int64_t*>(b);
        int64_t* d_fvp = d_cfp + (nc + 1);
        double* d_v = reinterpret_cast<double*>(d_fvp + (nf + 1));
        double* d_site = d_v + 3 * nv;
        double* d_cen = d_site + 3 * nc;
        double* d_vol = d_cen + 3 * nc;
        KIWI_HIP_CHECK(c, hipMemcpyAsync(d_cfp, cfp, sizeof(int64_t) * (nc + 1), hipMemcpyHostToDevice, s));
        if (nf > 0) KIWI_HIP_CHECK(c, hipMemcpyAsync(d_fvp, fvp, sizeof(int64_t) * (nf + 1), hipMemcpyHostToDevice, s));
        if (nv > 0) KIWI_HIP_CHECK(c, hipMemcpyAsync(d_v, vxyz, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, s));
        KIWI_HIP_CHECK(c, hipMemcpyAsync(d_site, site, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, s));
        cfp = d_cfp;
        fvp = nf > 0 ? d_fvp : nullptr;
        vxyz = nv > 0 ? d_v : nullptr;
        site = d_site;
        cen = d_cen;
        vol = d_vol;
    }
    KIWI_HIP_CHECK(c, reserve(c->stats, sizeof(unsigned long long)));
    unsigned long long* bad = static_cast<unsigned long long*>(c->stats.p);
    KIWI_HIP_CHECK(c, hipMemsetAsync(bad, 0, sizeof(unsigned long long), s));
    k_centroids<<<grid1d(nc), THREADS, 0, s>>>(nc, nf, nv, cfp, fvp, vxyz, site, cen, vol, bad);
    KIWI_HIP_CHECK(c, hipGetLastError());
    unsigned long long hbad = 0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s));
    if (!V->on_device) {
        KIWI_HIP_CHECK(c, hipMemcpyAsync(centroid_xyz, cen, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, s));
        if (volume) KIWI_HIP_CHECK(c, hipMemcpyAsync(volume, vol, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
    }
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (hbad) return fail(c, KIWI_ERR_INVALID_ARG, "malformed tessellation: face or vertex offsets out of range");
    return KIWI_OK;
}

}  // extern "C"
