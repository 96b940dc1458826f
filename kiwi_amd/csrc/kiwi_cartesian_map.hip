// libkiwi_hip: the cartesian-map resample (run_cartesian_map_stage, src/cartesian_map.cpp:84-201) on
// MI355X.  Every cell centre of a res^3 grid over the domain takes the mapping.dat row of the Voronoi
// cell it lies in, i.e. of its nearest site (voro++ find_voronoi_cell: minimum squared distance,
// src/cartesian_map.cpp:162-178; voro++ v_compute.cc scan_all).
//
//   1. k_cm_keys       per site its bucket in a uniform grid over the domain (about 1 site a bucket);
//                      a site outside the domain (voro++ pre_container::put drops it) gets the key past
//                      the last bucket and takes no part
//   2. sort            stable radix sort of the (bucket, site) keys on the bucket bits only: each bucket's
//                      sites contiguous, in site order
//   3. k_cm_gather     the sites' coordinates in bucket order (structure of arrays), bucket ranges
//   4. k_cm_query      one thread per Cartesian cell, a workgroup per 4 x 8 x 8 (x, y, z) brick of
//                      cells, a wavefront per 8 x 8 (y, z) plane of it, so that a wavefront's lanes
//                      start in the same few buckets and walk them together: the own bucket, then
//                      shells of buckets of growing Chebyshev radius
//                      until the best squared distance is below the distance to the unsearched region;
//                      then the site's id and row (or zeros and an "unmapped" count)
//
// Distances are formed as voro++ forms them, (s - q)^2 summed x, y, z with no fused multiply-add, so
// every comparison rounds as the reference's.  Of sites at exactly equal distance the lowest index
// wins; voro++ keeps the first its block search meets, which depends on its block layout (exact ties
// need lattice-aligned sites and cell centres).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "common.h"

namespace {

constexpr double CM_SITES_PER_BUCKET = 1.0;  // swept 0.25-8 on 2e5 clustered sites: 1 is fastest
constexpr int CM_MAX_BUCKET_AXIS = 1024;
constexpr int CM_BRICK_Z = 8, CM_BRICK_Y = 8, CM_BRICK_X = 4;  // CM_BRICK_X * CM_BRICK_Y * CM_BRICK_Z == THREADS
static_assert(CM_BRICK_X * CM_BRICK_Y * CM_BRICK_Z == THREADS, "one cell a thread");

struct CmGrid {
    double lo[3], hi[3], edge[3];
    int nb[3];
};

__device__ __forceinline__ int cm_bucket_axis(const CmGrid& g, int a, double v) {
    const double f = floor((v - g.lo[a]) / g.edge[a]);
    return static_cast<int>(fmin(fmax(f, 0.0), static_cast<double>(g.nb[a] - 1)));
}

__global__ void k_cm_keys(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, CmGrid g, uint64_t* __restrict__ keys,
                          unsigned long long* counts) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {x[i], y[i], z[i]};
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) atomicAdd(&counts[0], 1ull);
    bool inside = true;
    for (int a = 0; a < 3; ++a) inside = inside && p[a] >= g.lo[a] && p[a] <= g.hi[a];
    const int bx = cm_bucket_axis(g, 0, p[0]), by = cm_bucket_axis(g, 1, p[1]), bz = cm_bucket_axis(g, 2, p[2]);
    const uint64_t b = inside ? (static_cast<uint64_t>(bx) * g.nb[1] + by) * g.nb[2] + bz
                              : static_cast<uint64_t>(g.nb[0]) * g.nb[1] * g.nb[2];
    if (!inside) atomicAdd(&counts[1], 1ull);
    keys[i] = (b << 32) | static_cast<uint64_t>(i);
}

// a site in bucket order: position and index in one 32-byte record, read by a wavefront as one
// uniform (scalar) load
struct __align__(32) CmSite {
    double x, y, z;
    int idx, pad;
};

// sites in bucket order; [beg, end) of every bucket (buckets without sites keep beg = end = 0); the
// dropped sites, sorted past the last bucket, are written to no range
__global__ void k_cm_gather(int64_t n, uint32_t nb, const uint64_t* __restrict__ keys, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z, CmSite* __restrict__ sites,
                            int2* __restrict__ range) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= n) return;
    const uint64_t k = keys[t];
    const uint32_t b = static_cast<uint32_t>(k >> 32);
    const int i = static_cast<int>(k & 0xffffffffull);
    sites[t] = CmSite{x[i], y[i], z[i], i, 0};
    if (b >= nb) return;
    if (t == 0 || static_cast<uint32_t>(keys[t - 1] >> 32) != b) range[b].x = static_cast<int>(t);
    if (t == n - 1 || static_cast<uint32_t>(keys[t + 1] >> 32) != b) range[b].y = static_cast<int>(t + 1);
}

// every lane's nearest site among the bucket's sites: squared distance as voro++ forms it, lowest index on
// ties; the bucket and its sites are wave-uniform
__device__ __forceinline__ void cm_scan(int2 r, double qx, double qy, double qz, const CmSite* __restrict__ sites,
                                        double& best, int& bi) {
    const int t0 = __builtin_amdgcn_readfirstlane(r.x), t1 = __builtin_amdgcn_readfirstlane(r.y);
    for (int t = t0; t < t1; ++t) {
        const CmSite st = sites[t];
        const double dx = sub_sep(st.x, qx), dy = sub_sep(st.y, qy), dz = sub_sep(st.z, qz);
        const double rs = add_sep(add_sep(mul_sep(dx, dx), mul_sep(dy, dy)), mul_sep(dz, dz));
        const bool take = rs < best || (rs == best && st.idx < bi);
        best = take ? rs : best;
        bi = take ? st.idx : bi;
    }
}

// min + (i + 0.5) d, the product and the sum rounded on their own as the reference's host code does
__device__ __forceinline__ double cm_centre(double lo, int i, double d) { return add_sep(lo, mul_sep(i + 0.5, d)); }

struct CmRows {
    const int* id;
    const double *density, *vx, *vy, *vz;
    const int* mapped;
};

// One wavefront resolves an 8 x 8 (y, z) plane of cells together.  Its lanes' buckets span the box
// [blo, bhi]; the wave scans the box, then shells of buckets of growing Chebyshev radius r around it,
// every lane evaluating every site of every scanned bucket.  Before shell r every lane knows that the
// sites it has not seen lie outside the box grown by r - 1, so at least its distance to that box's
// outside, less tol (a bound on where rounding puts the bucket edges), from its centre; the wave stops
// once every lane's best squared distance is below that square, or the box covers the grid.
__global__ __launch_bounds__(THREADS) void k_cm_query(int res, double x_min, double y_min, double z_min, double dpx,
                                                      double dpy, double dpz, CmGrid g, double tol,
                                                      const int2* __restrict__ range, const CmSite* __restrict__ sites,
                                                      CmRows rows, double* __restrict__ od, double* __restrict__ ovx,
                                                      double* __restrict__ ovy, double* __restrict__ ovz,
                                                      int* __restrict__ oid, unsigned long long* unmapped) {
    const int tiles_z = (res + CM_BRICK_Z - 1) / CM_BRICK_Z, tiles_y = (res + CM_BRICK_Y - 1) / CM_BRICK_Y;
    const int64_t blk = blockIdx.x;
    const int z0 = static_cast<int>(blk % tiles_z) * CM_BRICK_Z, y0 = static_cast<int>((blk / tiles_z) % tiles_y) * CM_BRICK_Y;
    const int ix = static_cast<int>(blk / (static_cast<int64_t>(tiles_z) * tiles_y)) * CM_BRICK_X +
                   static_cast<int>(threadIdx.x / (CM_BRICK_Z * CM_BRICK_Y));  // one plane a wavefront
    const int iz = z0 + static_cast<int>(threadIdx.x % CM_BRICK_Z);
    const int iy = y0 + static_cast<int>((threadIdx.x / CM_BRICK_Z) % CM_BRICK_Y);
    if (ix >= res) return;  // a whole wavefront
    const bool live = iy < res && iz < res;
    // cell centre as src/cartesian_map.cpp:157-164 forms it (lanes past the grid edge take the last cell)
    const int cy = min(iy, res - 1), cz = min(iz, res - 1);
    const double qx = cm_centre(x_min, ix, dpx), qy = cm_centre(y_min, cy, dpy), qz = cm_centre(z_min, cz, dpz);
    const double qv[3] = {qx, qy, qz};
    // the wave's bucket box: the buckets of its lowest and highest centres (the bucket map is monotone)
    const int y1 = min(y0 + CM_BRICK_Y - 1, res - 1), z1 = min(z0 + CM_BRICK_Z - 1, res - 1);
    const int blo[3] = {cm_bucket_axis(g, 0, qx), cm_bucket_axis(g, 1, cm_centre(y_min, y0, dpy)),
                        cm_bucket_axis(g, 2, cm_centre(z_min, z0, dpz))};
    const int bhi[3] = {blo[0], cm_bucket_axis(g, 1, cm_centre(y_min, y1, dpy)),
                        cm_bucket_axis(g, 2, cm_centre(z_min, z1, dpz))};
    double best = INFINITY;
    int bi = 0x7fffffff;
    for (int r = 0;; ++r) {
        if (r > 0) {
            bool covered = true;
            double dmin = INFINITY;
            for (int a = 0; a < 3; ++a) {
                if (blo[a] - (r - 1) > 0) {
                    covered = false;
                    dmin = fmin(dmin, qv[a] - (g.lo[a] + (blo[a] - (r - 1)) * g.edge[a]));
                }
                if (bhi[a] + (r - 1) < g.nb[a] - 1) {
                    covered = false;
                    dmin = fmin(dmin, g.lo[a] + (bhi[a] + r) * g.edge[a] - qv[a]);
                }
            }
            if (covered) break;  // every bucket scanned
            const double dm = dmin - tol;
            if (__all(dm > 0.0 && best < dm * dm)) break;
        }
        // the shell of radius r around the box, clipped to the grid
        const int x0 = max(blo[0] - r, 0), x1 = min(bhi[0] + r, g.nb[0] - 1);
        const int ya = max(blo[1] - r, 0), yb = min(bhi[1] + r, g.nb[1] - 1);
        const int za = max(blo[2] - r, 0), zb = min(bhi[2] + r, g.nb[2] - 1);
        for (int bx = x0; bx <= x1; ++bx)
            for (int by = ya; by <= yb; ++by) {
                const int64_t row = (static_cast<int64_t>(bx) * g.nb[1] + by) * g.nb[2];
                const bool inner = r > 0 && bx > blo[0] - r && bx < bhi[0] + r && by > blo[1] - r && by < bhi[1] + r;
                if (!inner) {
                    for (int bz = za; bz <= zb; ++bz) cm_scan(range[row + bz], qx, qy, qz, sites, best, bi);
                } else {
                    if (blo[2] - r >= 0) cm_scan(range[row + blo[2] - r], qx, qy, qz, sites, best, bi);
                    if (bhi[2] + r < g.nb[2]) cm_scan(range[row + bhi[2] + r], qx, qy, qz, sites, best, bi);
                }
            }
    }
    if (!live) return;
    const int64_t c = (static_cast<int64_t>(ix) * res + iy) * res + iz;
    double d = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    int id = 0;
    if (bi != 0x7fffffff) {
        id = rows.id[bi];
        if (!rows.mapped || rows.mapped[bi]) {
            d = rows.density[bi];
            vx = rows.vx[bi];
            vy = rows.vy[bi];
            vz = rows.vz[bi];
        } else {
            atomicAdd(unmapped, 1ull);
        }
    }
    od[c] = d;
    ovx[c] = vx;
    ovy[c] = vy;
    ovz[c] = vz;
    oid[c] = id;
}

}  // namespace

extern "C" {

int kiwi_hip_cartesian_map(kiwi_hip_ctx* c, const kiwi_sites* S, const double domain[6], int res,
                           const kiwi_cartesian_map_out* out, kiwi_cartesian_map_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!S || !domain || !out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (res <= 0) return fail(c, KIWI_ERR_INVALID_ARG, "cartesian_map.resolution must be > 0");  // :112
    // 2048^3 cells of 36 output bytes fill 309 GB, past one GPU's memory; the query grid stays < 2^31 blocks
    if (res > 2048) return fail(c, KIWI_ERR_LIMIT, "more than 2048 cells along a grid axis");
    if (S->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative site count");
    if (S->n > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 sites in one call");
    if (S->n > 0 && (!S->x || !S->y || !S->z || !S->id || !S->density || !S->vx || !S->vy || !S->vz))
        return fail(c, KIWI_ERR_INVALID_ARG, "null site array");
    if (!out->density || !out->vx || !out->vy || !out->vz || !out->cell_id)
        return fail(c, KIWI_ERR_INVALID_ARG, "null output array");
    for (int a = 0; a < 3; ++a)
        if (!(domain[2 * a] < domain[2 * a + 1]) || !std::isfinite(domain[2 * a]) || !std::isfinite(domain[2 * a + 1]))
            return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");

    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = S->n;
    const int64_t ncell = static_cast<int64_t>(res) * res * res;

    // sites (and their rows) on the device
    const double* sp[7] = {S->x, S->y, S->z, S->density, S->vx, S->vy, S->vz};
    const double* dsp[7];
    const int* did = S->id;
    const int* dmapped = S->mapped;
    if (S->on_device) {
        for (int k = 0; k < 7; ++k) dsp[k] = sp[k];
    } else {
        KIWI_HIP_CHECK(c, reserve(c->parts, (7 * sizeof(double) + 2 * sizeof(int)) * std::max<int64_t>(n, 1)));
        for (int k = 0; k < 7; ++k) {
            double* d = static_cast<double*>(c->parts.p) + k * n;
            if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(d, sp[k], n * sizeof(double), hipMemcpyHostToDevice, s));
            dsp[k] = d;
        }
        int* di = reinterpret_cast<int*>(static_cast<double*>(c->parts.p) + 7 * n);
        if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(di, S->id, n * sizeof(int), hipMemcpyHostToDevice, s));
        did = di;
        if (S->mapped) {
            if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(di + n, S->mapped, n * sizeof(int), hipMemcpyHostToDevice, s));
            dmapped = di + n;
        }
    }
    double* od[4];
    int* oid;
    if (out->on_device) {
        od[0] = out->density;
        od[1] = out->vx;
        od[2] = out->vy;
        od[3] = out->vz;
        oid = out->cell_id;
    } else {
        KIWI_HIP_CHECK(c, reserve(c->outbuf, (4 * sizeof(double) + sizeof(int)) * ncell));
        for (int k = 0; k < 4; ++k) od[k] = static_cast<double*>(c->outbuf.p) + k * ncell;
        oid = reinterpret_cast<int*>(static_cast<double*>(c->outbuf.p) + 4 * ncell);
    }

    // bucket grid: about CM_SITES_PER_BUCKET sites a bucket, at most CM_MAX_BUCKET_AXIS per axis
    CmGrid g;
    {
        const double L[3] = {domain[1] - domain[0], domain[3] - domain[2], domain[5] - domain[4]};
        const double e = std::cbrt(L[0] * L[1] * L[2] * CM_SITES_PER_BUCKET / static_cast<double>(std::max<int64_t>(n, 1)));
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = domain[2 * a];
            g.hi[a] = domain[2 * a + 1];
            g.nb[a] = static_cast<int>(std::min<double>(std::max(1.0, std::ceil(L[a] / e)), CM_MAX_BUCKET_AXIS));
            g.edge[a] = L[a] / g.nb[a];
        }
    }
    const int64_t nb = static_cast<int64_t>(g.nb[0]) * g.nb[1] * g.nb[2];
    KIWI_HIP_CHECK(c, reserve(c->keys_a, 2 * sizeof(uint64_t) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve(c->ranges, sizeof(int2) * nb));
    KIWI_HIP_CHECK(c, reserve(c->slab_sorted, sizeof(CmSite) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve(c->stats, 3 * sizeof(unsigned long long)));
    uint64_t* keys_in = static_cast<uint64_t*>(c->keys_a.p);
    uint64_t* keys = keys_in + std::max<int64_t>(n, 1);
    int2* range = static_cast<int2*>(c->ranges.p);
    CmSite* sorted = static_cast<CmSite*>(c->slab_sorted.p);
    // [0] unmapped centres, [1] non-finite sites, [2] sites outside the domain
    unsigned long long* dcount = static_cast<unsigned long long*>(c->stats.p);

    KIWI_HIP_CHECK(c, hipMemsetAsync(dcount, 0, 3 * sizeof(unsigned long long), s));
    KIWI_HIP_CHECK(c, hipMemsetAsync(range, 0, sizeof(int2) * nb, s));
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    unsigned long long hcount[2] = {0, 0};  // non-finite sites, sites outside the domain
    if (n > 0) {
        k_cm_keys<<<grid1d(n), THREADS, 0, s>>>(n, dsp[0], dsp[1], dsp[2], g, keys_in, dcount + 1);
        KIWI_HIP_CHECK(c, hipGetLastError());
        // the keys arrive in site order and the sort is stable: the bucket bits [32, hi_bit) suffice
        // (bucket values 0..nb, nb marking the dropped sites)
        int hi_bit = 33;
        while ((1LL << (hi_bit - 32)) <= nb) ++hi_bit;
        size_t tb = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys_in, keys, static_cast<int>(n), 32,
                                                            hi_bit, s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tb));
        tb = c->cub_tmp.cap;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(c->cub_tmp.p, tb, keys_in, keys, static_cast<int>(n), 32,
                                                            hi_bit, s));
        k_cm_gather<<<grid1d(n), THREADS, 0, s>>>(n, static_cast<uint32_t>(nb), keys, dsp[0], dsp[1], dsp[2], sorted,
                                                  range);
        KIWI_HIP_CHECK(c, hipGetLastError());
        KIWI_HIP_CHECK(c, hipMemcpyAsync(hcount, dcount + 1, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
        if (hcount[0]) return fail(c, KIWI_ERR_INVALID_ARG, "non-finite site coordinate");
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
    const CmRows rows{did, dsp[3], dsp[4], dsp[5], dsp[6], dmapped};
    const double dpx = (domain[1] - domain[0]) / res, dpy = (domain[3] - domain[2]) / res,
                 dpz = (domain[5] - domain[4]) / res;
    const int64_t bricks = static_cast<int64_t>((res + CM_BRICK_X - 1) / CM_BRICK_X) * ((res + CM_BRICK_Y - 1) / CM_BRICK_Y) *
                           ((res + CM_BRICK_Z - 1) / CM_BRICK_Z);
    // where floor() puts a coordinate relative to the bucket edges g.lo + k edge agrees with those edges to
    // a few units in the last place of the coordinates' magnitude; tol bounds that by a factor > 1000
    double scale = 0.0;
    for (int a = 0; a < 6; ++a) scale = std::max(scale, std::fabs(domain[a]));
    const double tol = 1e-12 * scale;
    k_cm_query<<<static_cast<int>(bricks), THREADS, 0, s>>>(res, domain[0], domain[2], domain[4], dpx, dpy, dpz, g, tol,
                                                           range, sorted, rows, od[0], od[1], od[2], od[3], oid, dcount);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));
    if (!out->on_device) {
        double* ho[4] = {out->density, out->vx, out->vy, out->vz};
        for (int k = 0; k < 4; ++k)
            KIWI_HIP_CHECK(c, hipMemcpyAsync(ho[k], od[k], ncell * sizeof(double), hipMemcpyDeviceToHost, s));
        KIWI_HIP_CHECK(c, hipMemcpyAsync(out->cell_id, oid, ncell * sizeof(int), hipMemcpyDeviceToHost, s));
    }
    unsigned long long unm = 0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&unm, dcount, sizeof(unm), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        // a non-empty point set gives every centre a nearest site
        stats->unresolved = n > static_cast<int64_t>(hcount[1]) ? 0 : ncell;
        stats->unmapped = static_cast<int64_t>(unm);
        float ms_sort = 0, ms_query = 0;
        (void)hipEventElapsedTime(&ms_sort, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_query, c->ev[1], c->ev[2]);
        stats->device_ms_prep = ms_sort;
        stats->device_ms_query = ms_query;
    }
    return KIWI_OK;
}

}  // extern "C"
