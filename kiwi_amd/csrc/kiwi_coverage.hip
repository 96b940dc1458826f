// libkiwi_hip: the coverage-grid buffer (add_coverage_grid_buffer, src/point_set.cpp:208-273) on MI355X.
// A candidate of the cell-centred lattice over the domain (generate_cartesian_lattice, :166-195) is
// covered when a particle lies within factor x h of it; the uncovered candidates become buffer sites.
//
// The reference tests every candidate against the particles of the 27 hash bins around it (bin =
// factor x h_max).  Here every particle marks the candidates it covers instead, so the work is the
// covered (particle, candidate) pairs rather than candidates x bin occupancy:
//
//   1. k_cov_scan      h_max (std::max from 0, NaN ignored) and a count of non-finite coordinates
//   2. k_cov_columns   per particle the lattice columns (x, y) of the box [p - r, p + r], r = factor h
//   3. scan            column offsets
//   4. k_cov_mark      one thread per (particle, column): the column's candidates in [z - r, z + r]
//                      with the reference's predicate, exactly: the particle's hash key is one of the
//                      27 around the candidate's, and (dx^2 + dy^2) + dz^2 <= r * r in double with no
//                      fused multiply-add; a covered candidate's keep flag is cleared (a plain store)
//   5. scan + k_cov_emit   the kept candidates in lattice order, ids n + 1, n + 2, ...
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "common.h"

namespace {

struct CovLattice {
    double lo[3], spacing;
    int n[3];
};

// std::round of the host: halves away from zero
double round_half_away(double v) {
    const double t = std::trunc(v);
    return std::fabs(v - t) >= 0.5 ? t + (v > 0 ? 1.0 : -1.0) : t;
}

// generate_cartesian_lattice's sizes (src/point_set.cpp:167-181)
bool cov_lattice(const double* dom, double fraction, CovLattice* L) {
    const double ext[3] = {dom[1] - dom[0], dom[3] - dom[2], dom[5] - dom[4]};
    L->spacing = fraction * std::min({ext[0], ext[1], ext[2]});
    for (int a = 0; a < 3; ++a) {
        L->lo[a] = dom[2 * a];
        const double v = round_half_away(ext[a] / L->spacing);
        if (!(v < 2147483647.0)) return false;
        L->n[a] = std::max(1, static_cast<int>(v));
    }
    return true;
}

__device__ __forceinline__ double cov_coord(const CovLattice& L, int a, int i) {
    return add_sep(L.lo[a], mul_sep(L.spacing, static_cast<double>(i) + 0.5));
}

// the hash key of one axis: floor(v / bin) as an int, + 2^19, clamped to [0, 2^20) (bin_of and bin_key,
// src/point_set.cpp:29-35, 217-222); the floor is clamped to +-2^30 first, where the reference's int
// conversion would overflow, which the key's own clamp makes immaterial
__device__ __forceinline__ int cov_bin(double v, double bin) {
    return static_cast<int>(fmin(fmax(floor(v / bin), -1073741824.0), 1073741824.0));
}
__device__ __forceinline__ int cov_key(int b) { return min(max(b + (1 << 19), 0), (1 << 20) - 1); }
__device__ __forceinline__ bool cov_near(int bp_key, int bc) {
    return bp_key == cov_key(bc - 1) || bp_key == cov_key(bc) || bp_key == cov_key(bc + 1);
}

// index range of the candidates whose coordinate lies within [p - ar, p + ar] on one axis, widened by one
// on each side against rounding (the exact test decides), clamped to the lattice; lo > hi when empty
__device__ __forceinline__ void cov_range(const CovLattice& L, int a, double p, double ar, int& lo, int& hi) {
    const double f0 = ceil((p - ar - L.lo[a]) / L.spacing - 0.5) - 1.0;
    const double f1 = floor((p + ar - L.lo[a]) / L.spacing - 0.5) + 1.0;
    lo = static_cast<int>(fmin(fmax(f0, 0.0), static_cast<double>(L.n[a])));
    hi = static_cast<int>(fmin(fmax(f1, -1.0), static_cast<double>(L.n[a] - 1)));
}

// a fixed grid strides over the particles; each block reduces in LDS and adds one atomic per counter
constexpr int COV_SCAN_BLOCKS = 1024;

__global__ __launch_bounds__(THREADS) void k_cov_scan(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ z,
                                                      const double* __restrict__ h,
                                                      unsigned long long* __restrict__ out) {
    __shared__ double s_h[THREADS / 64];
    __shared__ unsigned long long s_bad[THREADS / 64];
    double hv = 0.0;
    unsigned long long nbad = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * THREADS) {
        const double v = h[i];
        hv = v > hv ? v : hv;  // NaN and h <= 0 never raise the maximum from 0
        nbad += !(isfinite(x[i]) && isfinite(y[i]) && isfinite(z[i]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        hv = fmax(hv, __shfl_xor(hv, o));
        nbad += __shfl_xor(nbad, o);
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        s_h[w] = hv;
        s_bad[w] = nbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < THREADS / 64; ++k) {
            hv = fmax(hv, s_h[k]);
            nbad += s_bad[k];
        }
        // for doubles >= 0 the bit patterns order as the values
        if (hv > 0.0) atomicMax(&out[0], static_cast<unsigned long long>(__double_as_longlong(hv)));
        if (nbad) atomicAdd(&out[1], nbad);
    }
}

__global__ void k_cov_columns(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ h, double factor, CovLattice L, int64_t* __restrict__ cols) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i > n) return;
    int64_t c = 0;
    if (i < n) {
        const double r = mul_sep(factor, h[i]);
        if (mul_sep(r, r) >= 0.0) {  // a NaN radius covers nothing
            const double ar = fabs(r);
            int x0, x1, y0, y1;
            cov_range(L, 0, x[i], ar, x0, x1);
            cov_range(L, 1, y[i], ar, y0, y1);
            if (x0 <= x1 && y0 <= y1) c = static_cast<int64_t>(x1 - x0 + 1) * (y1 - y0 + 1);
        }
    }
    cols[i] = c;  // cols[n] = 0 closes the scan
}

__global__ void k_cov_fill(int64_t m, int* __restrict__ keep) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (k < m) keep[k] = k + 1 < m ? 1 : 0;  // the last entry closes the scan
}

__global__ void k_cov_mark(int64_t total, int64_t n, const int64_t* __restrict__ off, const double* __restrict__ x,
                           const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ h,
                           double factor, double bin, CovLattice L, int* __restrict__ keep) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * THREADS;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x; t < total; t += stride) {
        // the particle whose column range holds t: the last i with off[i] <= t
        int64_t a = 0, b = n - 1;
        while (a < b) {
            const int64_t m = (a + b + 1) >> 1;
            if (off[m] <= t) a = m;
            else b = m - 1;
        }
        const int64_t i = a;
        const double px = x[i], py = y[i], pz = z[i];
        const double r = mul_sep(factor, h[i]), r2 = mul_sep(r, r), ar = fabs(r);
        int x0, x1, y0, y1, z0, z1;
        cov_range(L, 0, px, ar, x0, x1);
        cov_range(L, 1, py, ar, y0, y1);
        cov_range(L, 2, pz, ar, z0, z1);
        const int64_t local = t - off[i];
        const int ny = y1 - y0 + 1;
        const int ix = x0 + static_cast<int>(local / ny), iy = y0 + static_cast<int>(local % ny);
        const double cx = cov_coord(L, 0, ix), cy = cov_coord(L, 1, iy);
        const int kx = cov_key(cov_bin(px, bin)), ky = cov_key(cov_bin(py, bin)), kz = cov_key(cov_bin(pz, bin));
        if (!cov_near(kx, cov_bin(cx, bin)) || !cov_near(ky, cov_bin(cy, bin))) continue;
        const double dx = sub_sep(px, cx), dy = sub_sep(py, cy);
        const double dxy = add_sep(mul_sep(dx, dx), mul_sep(dy, dy));
        const int64_t row = (static_cast<int64_t>(ix) * L.n[1] + iy) * L.n[2];
        for (int iz = z0; iz <= z1; ++iz) {
            const double cz = cov_coord(L, 2, iz);
            const double dz = sub_sep(pz, cz);
            if (add_sep(dxy, mul_sep(dz, dz)) <= r2 && cov_near(kz, cov_bin(cz, bin))) keep[row + iz] = 0;
        }
    }
}

__global__ void k_cov_emit(int64_t m, const int* __restrict__ keep, const int* __restrict__ pos, CovLattice L,
                           int64_t n_particles, double* __restrict__ ox, double* __restrict__ oy,
                           double* __restrict__ oz, int* __restrict__ oid) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (k >= m || !keep[k]) return;
    const int iz = static_cast<int>(k % L.n[2]), iy = static_cast<int>((k / L.n[2]) % L.n[1]),
              ix = static_cast<int>(k / (static_cast<int64_t>(L.n[1]) * L.n[2]));
    const int j = pos[k];
    ox[j] = cov_coord(L, 0, ix);
    oy[j] = cov_coord(L, 1, iy);
    oz[j] = cov_coord(L, 2, iz);
    oid[j] = static_cast<int>(n_particles + 1 + j);
}

}  // namespace

extern "C" {

int kiwi_hip_cartesian_lattice(const double domain[6], double spacing_fraction, kiwi_lattice* lattice) {
    if (!domain || !lattice) return KIWI_ERR_INVALID_ARG;
    if (!(spacing_fraction > 0.0)) return KIWI_ERR_INVALID_ARG;
    for (int a = 0; a < 3; ++a)
        if (!(domain[2 * a] < domain[2 * a + 1]) || !std::isfinite(domain[2 * a]) || !std::isfinite(domain[2 * a + 1]))
            return KIWI_ERR_INVALID_ARG;
    CovLattice L;
    if (!cov_lattice(domain, spacing_fraction, &L)) return KIWI_ERR_LIMIT;
    lattice->nx = L.n[0];
    lattice->ny = L.n[1];
    lattice->nz = L.n[2];
    lattice->spacing = L.spacing;
    return KIWI_OK;
}

int kiwi_hip_coverage_buffer(kiwi_hip_ctx* c, const kiwi_particles* P, const double domain[6], double spacing_fraction,
                             double coverage_radius_factor, const kiwi_buffer_sites* out, kiwi_coverage_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!P || !domain || !out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    // the configuration checks of src/config.cpp:109-112
    if (!(spacing_fraction > 0.0)) return fail(c, KIWI_ERR_INVALID_ARG, "buffer.coverage_grid.spacing_fraction must be > 0");
    if (!(coverage_radius_factor > 0.0))
        return fail(c, KIWI_ERR_INVALID_ARG, "buffer.coverage_grid.coverage_radius_factor must be > 0");
    for (int a = 0; a < 3; ++a)
        if (!(domain[2 * a] < domain[2 * a + 1]) || !std::isfinite(domain[2 * a]) || !std::isfinite(domain[2 * a + 1]))
            return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");
    if (P->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative particle count");
    if (P->n > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 particles in one call");
    if (P->n > 0 && (!P->x || !P->y || !P->z || !P->h)) return fail(c, KIWI_ERR_INVALID_ARG, "null particle array");
    if (out->capacity < 0 || (out->capacity > 0 && (!out->x || !out->y || !out->z || !out->id)))
        return fail(c, KIWI_ERR_INVALID_ARG, "null output array");
    CovLattice L;
    if (!cov_lattice(domain, spacing_fraction, &L)) return fail(c, KIWI_ERR_LIMIT, "candidate lattice too large");
    const int64_t m = static_cast<int64_t>(L.n[0]) * L.n[1] * L.n[2];
    if (m >= 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-2 lattice candidates");
    const int64_t n = P->n;
    if (n == 0) return fail(c, KIWI_ERR_INVALID_ARG, "add_coverage_grid_buffer: all smoothing lengths are <= 0");

    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const double* col[4] = {P->x, P->y, P->z, P->h};
    const double* d[4];
    if (P->on_device) {
        for (int k = 0; k < 4; ++k) d[k] = col[k];
    } else {
        KIWI_HIP_CHECK(c, reserve(c->parts, 4 * sizeof(double) * n));
        for (int k = 0; k < 4; ++k) {
            double* p = static_cast<double*>(c->parts.p) + k * n;
            KIWI_HIP_CHECK(c, hipMemcpyAsync(p, col[k], n * sizeof(double), hipMemcpyHostToDevice, s));
            d[k] = p;
        }
    }
    KIWI_HIP_CHECK(c, reserve(c->stats, 2 * sizeof(unsigned long long)));
    KIWI_HIP_CHECK(c, reserve(c->sizes, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->offsets, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->work, sizeof(int) * (m + 1)));
    KIWI_HIP_CHECK(c, reserve(c->pos_a, sizeof(int) * (m + 1)));
    unsigned long long* dst = static_cast<unsigned long long*>(c->stats.p);
    int64_t* cols = static_cast<int64_t*>(c->sizes.p);
    int64_t* off = static_cast<int64_t*>(c->offsets.p);
    int* keep = static_cast<int*>(c->work.p);
    int* pos = static_cast<int*>(c->pos_a.p);

    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    KIWI_HIP_CHECK(c, hipMemsetAsync(dst, 0, 2 * sizeof(unsigned long long), s));
    k_cov_scan<<<static_cast<int>(std::min<int64_t>(grid1d(n), COV_SCAN_BLOCKS)), THREADS, 0, s>>>(n, d[0], d[1], d[2],
                                                                                                 d[3], dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    unsigned long long hs[2] = {0, 0};
    KIWI_HIP_CHECK(c, hipMemcpyAsync(hs, dst, sizeof(hs), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (hs[1]) return fail(c, KIWI_ERR_INVALID_ARG, "non-finite particle coordinate");
    double h_max = 0.0;
    std::memcpy(&h_max, &hs[0], sizeof(double));
    if (!(h_max > 0.0)) return fail(c, KIWI_ERR_INVALID_ARG, "add_coverage_grid_buffer: all smoothing lengths are <= 0");
    const double bin = coverage_radius_factor * h_max;

    k_cov_columns<<<grid1d(n + 1), THREADS, 0, s>>>(n, d[0], d[1], d[3], coverage_radius_factor, L, cols);
    KIWI_HIP_CHECK(c, hipGetLastError());
    size_t tb = 0;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cols, off, static_cast<int>(n + 1), s));
    KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tb));
    tb = c->cub_tmp.cap;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tb, cols, off, static_cast<int>(n + 1), s));
    int64_t total = 0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&total, off + n, sizeof(total), hipMemcpyDeviceToHost, s));
    k_cov_fill<<<grid1d(m + 1), THREADS, 0, s>>>(m + 1, keep);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (total > 0) {
        const int blocks = static_cast<int>(std::min<int64_t>((total + THREADS - 1) / THREADS, 64LL * std::max(c->n_cu, 1)));
        k_cov_mark<<<blocks, THREADS, 0, s>>>(total, n, off, d[0], d[1], d[2], d[3], coverage_radius_factor, bin, L,
                                              keep);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    tb = 0;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, keep, pos, static_cast<int>(m + 1), s));
    KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tb));
    tb = c->cub_tmp.cap;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tb, keep, pos, static_cast<int>(m + 1), s));
    int kept = 0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&kept, pos + m, sizeof(kept), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        stats->candidates = m;
        stats->kept = kept;
        stats->h_max = h_max;
        stats->particle_columns = total;
        stats->device_ms = 0.0;
    }
    if (kept > out->capacity) return fail(c, KIWI_ERR_LIMIT, "buffer output capacity below the kept candidates");
    if (n + kept > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "buffer site ids past 2^31-1");
    if (kept > 0) {
        double *ox = out->x, *oy = out->y, *oz = out->z;
        int* oid = out->id;
        if (!out->on_device) {
            KIWI_HIP_CHECK(c, reserve(c->outbuf, (3 * sizeof(double) + sizeof(int)) * kept));
            ox = static_cast<double*>(c->outbuf.p);
            oy = ox + kept;
            oz = oy + kept;
            oid = reinterpret_cast<int*>(oz + kept);
        }
        k_cov_emit<<<grid1d(m), THREADS, 0, s>>>(m, keep, pos, L, n, ox, oy, oz, oid);
        KIWI_HIP_CHECK(c, hipGetLastError());
        if (!out->on_device) {
            KIWI_HIP_CHECK(c, hipMemcpyAsync(out->x, ox, kept * sizeof(double), hipMemcpyDeviceToHost, s));
            KIWI_HIP_CHECK(c, hipMemcpyAsync(out->y, oy, kept * sizeof(double), hipMemcpyDeviceToHost, s));
            KIWI_HIP_CHECK(c, hipMemcpyAsync(out->z, oz, kept * sizeof(double), hipMemcpyDeviceToHost, s));
            KIWI_HIP_CHECK(c, hipMemcpyAsync(out->id, oid, kept * sizeof(int), hipMemcpyDeviceToHost, s));
        }
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        stats->device_ms = ms;
    }
    return KIWI_OK;
}

}  // extern "C"
