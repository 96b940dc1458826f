// libkiwi_hip: the C ABI of include/kiwi_hip.h on MI355X (gfx950).
//
// sph-to-grid pipeline (one call of kiwi_hip_sph_to_grid, all on the context's stream):
//   1. k_prep      one thread per particle: the reference's bounding box of cells
//                  (src/cartesian_grid.cpp:209-238, FP64 floor, clamped), the cell blocks it overlaps,
//                  skipped / pair counters.
//   2. scan        exclusive prefix sum of blocks-per-particle (hipcub).
//   3. k_emit      one 64-bit key (block << 32 | particle) per (particle, block) work item.
//   4. sort        radix sort of the keys: every block's particle list comes out in particle order,
//                  so each cell sums its particles in the reference's (single-thread) order.
//   5. k_ranges    [begin, end) of every block's list.
//   6. k_gather    one workgroup per 8x8x8 cell block: for each particle of its list, the plane and
//                  line constants and the lattice values W (m4_lattice.h) are built in LDS, then every
//                  thread folds the 8-corner stencil into the cells it owns.  No float atomics: each
//                  cell is owned by exactly one thread of one workgroup, so results are bit-reproducible.
//   7. k_vsum      deterministic mean velocity over all particles (the empty-cell fallback).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/kiwi_hip.h"
#include "m4_lattice.h"

using namespace kiwi_hip;

namespace {

constexpr int BLK = 8;                       // cells per block edge
constexpr int BLK_CELLS = BLK * BLK * BLK;   // 512
constexpr int NL = BLK + 1;                  // lattice points per block edge
constexpr int THREADS = 256;
constexpr int CELLS_PER_THREAD = BLK_CELLS / THREADS;

struct DevStats {
    unsigned long long skipped_outside, n_negative, n_negative_warn, pairs_bbox, pairs_sphere;
    unsigned long long n_warn;  // number of warn records claimed
    unsigned long long pts[3], lines[3], planes;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS];
};

struct GridGeom {
    double x_min, y_min, z_min, dx, dy, dz;
    int nx, ny, nz;
    int nbx, nby, nbz;
};

// ------------------------------------------------------------------------------------------------
// 1. per-particle bounding box
// ------------------------------------------------------------------------------------------------
// floor(v) as a cell index, saturated to [-1, n] before the conversion so that huge supports and
// non-finite coordinates cannot overflow it (NaN lands on n, outside the grid).  Within range this is
// exactly the reference's static_cast<int>(std::floor(v)) (src/cartesian_grid.cpp:220-225).
__device__ __forceinline__ int cell_floor(double v, int n) {
    const double f = floor(v);
    if (f >= -1.0 && f <= static_cast<double>(n)) return static_cast<int>(f);
    return (f < -1.0) ? -1 : n;
}

__global__ void k_prep(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const double* __restrict__ h, double support_in_h,
                       GridGeom g, int* __restrict__ bbox, int64_t* __restrict__ nblk, DevStats* st) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long skipped = 0, pairs = 0;
    if (i == n) nblk[n] = 0;  // trailing zero of the exclusive scan
    if (i < n) {
        int64_t nb = 0;
        int* bb = bbox + 6 * i;
        bb[0] = 1;
        bb[1] = 0;  // empty unless set below
        const double hh = h[i];
        const double x0 = x[i], y0 = y[i], z0 = z[i];
        if (hh > 0.0) {
            const double rad = support_in_h * hh;
            int ix0 = cell_floor((x0 - rad - g.x_min) / g.dx, g.nx);
            int ix1 = cell_floor((x0 + rad - g.x_min) / g.dx, g.nx);
            int iy0 = cell_floor((y0 - rad - g.y_min) / g.dy, g.ny);
            int iy1 = cell_floor((y0 + rad - g.y_min) / g.dy, g.ny);
            int iz0 = cell_floor((z0 - rad - g.z_min) / g.dz, g.nz);
            int iz1 = cell_floor((z0 + rad - g.z_min) / g.dz, g.nz);
            ix0 = max(ix0, 0);
            ix1 = min(ix1, g.nx - 1);
            iy0 = max(iy0, 0);
            iy1 = min(iy1, g.ny - 1);
            iz0 = max(iz0, 0);
            iz1 = min(iz1, g.nz - 1);
            if (ix0 > ix1 || iy0 > iy1 || iz0 > iz1) {
                skipped = 1;
            } else {
                bb[0] = ix0;
                bb[1] = ix1;
                bb[2] = iy0;
                bb[3] = iy1;
                bb[4] = iz0;
                bb[5] = iz1;
                pairs = static_cast<unsigned long long>(ix1 - ix0 + 1) * (iy1 - iy0 + 1) * (iz1 - iz0 + 1);
                nb = static_cast<int64_t>(ix1 / BLK - ix0 / BLK + 1) * (iy1 / BLK - iy0 / BLK + 1) *
                     (iz1 / BLK - iz0 / BLK + 1);
            }
        }
        nblk[i] = nb;
    }
    // one atomic per wave for the counters
    for (int off = 32; off > 0; off >>= 1) {
        skipped += __shfl_down(skipped, off);
        pairs += __shfl_down(pairs, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (skipped) atomicAdd(&st->skipped_outside, skipped);
        if (pairs) atomicAdd(&st->pairs_bbox, pairs);
    }
}

// ------------------------------------------------------------------------------------------------
// 3. keys
// ------------------------------------------------------------------------------------------------
__global__ void k_emit(int64_t n, const int* __restrict__ bbox, const int64_t* __restrict__ offs, GridGeom g,
                       uint64_t* __restrict__ keys) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int* bb = bbox + 6 * i;
    if (bb[0] > bb[1]) return;
    int64_t o = offs[i];
    for (int bx = bb[0] / BLK; bx <= bb[1] / BLK; ++bx)
        for (int by = bb[2] / BLK; by <= bb[3] / BLK; ++by)
            for (int bz = bb[4] / BLK; bz <= bb[5] / BLK; ++bz) {
                const uint64_t b = (static_cast<uint64_t>(bx) * g.nby + by) * g.nbz + bz;
                keys[o++] = (b << 32) | static_cast<uint64_t>(i);
            }
}

// ------------------------------------------------------------------------------------------------
// 5. block ranges in the sorted key array
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_int(int* p, int64_t n, int v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_ranges(int64_t npairs, const uint64_t* __restrict__ keys, int* __restrict__ begin,
                         int* __restrict__ end) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const uint32_t b = static_cast<uint32_t>(keys[i] >> 32);
    if (i == 0 || static_cast<uint32_t>(keys[i - 1] >> 32) != b) begin[b] = static_cast<int>(i);
    if (i == npairs - 1 || static_cast<uint32_t>(keys[i + 1] >> 32) != b) end[b] = static_cast<int>(i + 1);
}

// ------------------------------------------------------------------------------------------------
// 6. the gather: one workgroup per cell block
// ------------------------------------------------------------------------------------------------
// ordered (wall axis a, edge-normal axis b) pairs -> line table index
__device__ __forceinline__ int pair_id(int a, int b) { return 2 * a + (b > a ? b - 1 : b); }

template <typename R>
struct GatherLds {
    R sd[3][NL];           // signed offsets lattice - particle, units of h
    R pc[3][NL][3];        // plane constants c1, c2, c3 of |sd|
    R D2[6][NL * NL];      // line constants per ordered pair, [ia * NL + ib]
    R D3[6][NL * NL];
    R U[NL * NL * NL];     // signed lattice values s * W
};

template <typename R>
__global__ __launch_bounds__(THREADS) void k_gather_m4(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, const double* __restrict__ h, const int* __restrict__ bbox,
    const uint64_t* __restrict__ keys, const int* __restrict__ blk_begin, const int* __restrict__ blk_end,
    GridGeom g, double* __restrict__ out_mass, double* __restrict__ out_px, double* __restrict__ out_py,
    double* __restrict__ out_pz, DevStats* st, int clamp_pairs) {
    __shared__ GatherLds<R> L;
    const int b = blockIdx.x;
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int ox = bx * BLK, oy = by * BLK, oz = bz * BLK;
    const int tid = threadIdx.x;

    double acc[CELLS_PER_THREAD][4];
    int cl[CELLS_PER_THREAD][3];
    for (int s = 0; s < CELLS_PER_THREAD; ++s) {
        const int l = tid + THREADS * s;
        cl[s][0] = l / (BLK * BLK);
        cl[s][1] = (l / BLK) % BLK;
        cl[s][2] = l % BLK;
        acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;
    }
    unsigned long long n_neg = 0, n_warn = 0, n_sphere = 0;
    unsigned long long c_pts[3] = {0, 0, 0}, c_lines[3] = {0, 0, 0}, c_planes = 0;

    const int q0 = blk_begin[b], q1 = blk_end[b];
    for (int q = q0; q < q1; ++q) {
        const int p = static_cast<int>(keys[q] & 0xffffffffull);
        const int* bb = bbox + 6 * p;
        const int lx0 = max(bb[0] - ox, 0), lx1 = min(bb[1] - ox, BLK - 1);
        const int ly0 = max(bb[2] - oy, 0), ly1 = min(bb[3] - oy, BLK - 1);
        const int lz0 = max(bb[4] - oz, 0), lz1 = min(bb[5] - oz, BLK - 1);
        const int nlx = lx1 - lx0 + 2, nly = ly1 - ly0 + 2, nlz = lz1 - lz0 + 2;
        const double px = x[p], py = y[p], pz = z[p], hh = h[p];
        const double inv_h = 1.0 / hh;

        // ---- phase A: plane and line constants ----
        const int nplanes = nlx + nly + nlz;
        const int nlines = 2 * (nlx * nly + nlx * nlz + nly * nlz);
        for (int t = tid; t < nplanes + nlines; t += THREADS) {
            if (t < nplanes) {
                int a, i;
                if (t < nlx) {
                    a = 0;
                    i = t;
                } else if (t < nlx + nly) {
                    a = 1;
                    i = t - nlx;
                } else {
                    a = 2;
                    i = t - nlx - nly;
                }
                double d;
                if (a == 0) d = (g.x_min + (ox + lx0 + i) * g.dx - px) * inv_h;
                else if (a == 1) d = (g.y_min + (oy + ly0 + i) * g.dy - py) * inv_h;
                else d = (g.z_min + (oz + lz0 + i) * g.dz - pz) * inv_h;
                ++c_planes;
                L.sd[a][i] = static_cast<R>(d);
                const M4Plane<R> pl = m4_plane<R>(static_cast<R>(fabs(d)));
                L.pc[a][i][0] = pl.c1;
                L.pc[a][i][1] = pl.c2;
                L.pc[a][i][2] = pl.c3;
            } else {
                // decode (ordered pair, ia, ib)
                int u = t - nplanes;
                const int na[3] = {nlx, nly, nlz};
                int a = 0, bq = 1;
                for (int pp = 0; pp < 6; ++pp) {
                    const int aa = pp >> 1;
                    const int bb2 = (pp & 1) ? (aa == 2 ? 1 : 2) : (aa == 0 ? 1 : 0);
                    const int cnt = na[aa] * na[bb2];
                    if (u < cnt) {
                        a = aa;
                        bq = bb2;
                        break;
                    }
                    u -= cnt;
                }
                const int ia = u / na[bq], ib = u % na[bq];
                const int ba[3] = {ox + lx0, oy + ly0, oz + lz0};
                const double mn[3] = {g.x_min, g.y_min, g.z_min};
                const double dd[3] = {g.dx, g.dy, g.dz};
                const double pp3[3] = {px, py, pz};
                const double da = fabs((mn[a] + (ba[a] + ia) * dd[a] - pp3[a]) * inv_h);
                const double db = fabs((mn[bq] + (ba[bq] + ib) * dd[bq] - pp3[bq]) * inv_h);
                R D2, D3;
                m4_line<R>(static_cast<R>(da), static_cast<R>(db), D2, D3);
                {
                    const double l2 = da * da + db * db;
                    ++c_lines[(l2 <= 1.0) ? 2 : ((l2 <= 4.0) ? 1 : 0)];
                }
                const int pid = pair_id(a, bq);
                L.D2[pid][ia * NL + ib] = D2;
                L.D3[pid][ia * NL + ib] = D3;
            }
        }
        __syncthreads();

        // ---- phase B: lattice values ----
        const int npts = nlx * nly * nlz;
        for (int t = tid; t < npts; t += THREADS) {
            const int i = t / (nly * nlz);
            const int j = (t / nlz) % nly;
            const int k = t % nlz;
            const R dx = L.sd[0][i], dy = L.sd[1][j], dz = L.sd[2][k];
            R U = R(0);
            if (dx != R(0) && dy != R(0) && dz != R(0)) {
                const R ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
                const R r2 = ax * ax + ay * ay + az * az;
                const R r = KMath<R>::sqrt_(r2);
                const int reg = m4_region(r2);
                ++c_pts[reg - 1];
                const R csum = L.pc[0][i][reg - 1] + L.pc[1][j][reg - 1] + L.pc[2][k][reg - 1];
                R dsum = R(0);
                if (reg == 2) {
                    dsum = L.D2[0][i * NL + j] + L.D2[1][i * NL + k] + L.D2[2][j * NL + i] + L.D2[3][j * NL + k] +
                           L.D2[4][k * NL + i] + L.D2[5][k * NL + j];
                } else if (reg == 3) {
                    dsum = L.D3[0][i * NL + j] + L.D3[1][i * NL + k] + L.D3[2][j * NL + i] + L.D3[3][j * NL + k] +
                           L.D3[4][k * NL + i] + L.D3[5][k * NL + j];
                }
                const R W = m4_lattice_free<R>(ax, ay, az, r, reg) * M4Consts<R>::inv_pi + R(0.5) * csum + dsum;
                const bool neg = (dx < R(0)) != ((dy < R(0)) != (dz < R(0)));
                U = neg ? -W : W;
            }
            L.U[(i * NL + j) * NL + k] = U;
        }
        __syncthreads();

        // ---- phase C: the 8-corner stencil into the owned cells ----
        const double mm = m[p], vxp = vx[p], vyp = vy[p], vzp = vz[p];
        for (int s = 0; s < CELLS_PER_THREAD; ++s) {
            const int cx = cl[s][0], cy = cl[s][1], cz = cl[s][2];
            if (cx < lx0 || cx > lx1 || cy < ly0 || cy > ly1 || cz < lz0 || cz > lz1) continue;
            const int i = cx - lx0, j = cy - ly0, k = cz - lz0;
            // exact box-sphere test: cells the support sphere (radius 2h) does not reach hold no mass
            R dist2 = R(0);
            {
                const R a0 = L.sd[0][i], a1 = L.sd[0][i + 1];
                const R b0 = L.sd[1][j], b1 = L.sd[1][j + 1];
                const R c0 = L.sd[2][k], c1 = L.sd[2][k + 1];
                const R ex = (a0 > R(0)) ? a0 : ((a1 < R(0)) ? -a1 : R(0));
                const R ey = (b0 > R(0)) ? b0 : ((b1 < R(0)) ? -b1 : R(0));
                const R ez = (c0 > R(0)) ? c0 : ((c1 < R(0)) ? -c1 : R(0));
                dist2 = ex * ex + ey * ey + ez * ez;
            }
            if (!(dist2 < R(4))) continue;
            ++n_sphere;
            const R* U0 = &L.U[(i * NL + j) * NL + k];
            double Ms = static_cast<double>(U0[NL * NL]) + static_cast<double>(U0[NL]) + static_cast<double>(U0[1]) +
                        static_cast<double>(U0[NL * NL + NL + 1]) - static_cast<double>(U0[0]) -
                        static_cast<double>(U0[NL * NL + NL]) - static_cast<double>(U0[NL * NL + 1]) -
                        static_cast<double>(U0[NL + 1]);
            if (clamp_pairs && Ms < 0.0) {
                ++n_neg;
                if (Ms <= -0.01) {
                    ++n_warn;
                    const unsigned long long slot = atomicAdd(&st->n_warn, 1ull);
                    if (slot < KIWI_MAX_WARN_RECORDS) {
                        st->warn[slot].particle = p;
                        st->warn[slot].cell =
                            (static_cast<int64_t>(ox + cx) * g.ny + (oy + cy)) * g.nz + (oz + cz);
                        st->warn[slot].value = Ms;
                    }
                }
                Ms = 0.0;
            }
            if (Ms == 0.0) continue;
            const double t = Ms * mm;
            acc[s][0] += t;
            acc[s][1] += t * vxp;
            acc[s][2] += t * vyp;
            acc[s][3] += t * vzp;
        }
        __syncthreads();
    }

    for (int s = 0; s < CELLS_PER_THREAD; ++s) {
        const int ix = ox + cl[s][0], iy = oy + cl[s][1], iz = oz + cl[s][2];
        if (ix >= g.nx || iy >= g.ny || iz >= g.nz) continue;
        const int64_t c = (static_cast<int64_t>(ix) * g.ny + iy) * g.nz + iz;
        out_mass[c] = acc[s][0];
        out_px[c] = acc[s][1];
        out_py[c] = acc[s][2];
        out_pz[c] = acc[s][3];
    }
    for (int off = 32; off > 0; off >>= 1) {
        n_neg += __shfl_down(n_neg, off);
        n_warn += __shfl_down(n_warn, off);
        n_sphere += __shfl_down(n_sphere, off);
        c_planes += __shfl_down(c_planes, off);
        for (int r = 0; r < 3; ++r) {
            c_pts[r] += __shfl_down(c_pts[r], off);
            c_lines[r] += __shfl_down(c_lines[r], off);
        }
    }
    if ((tid & 63) == 0) {
        if (n_neg) atomicAdd(&st->n_negative, n_neg);
        if (n_warn) atomicAdd(&st->n_negative_warn, n_warn);
        if (n_sphere) atomicAdd(&st->pairs_sphere, n_sphere);
        if (c_planes) atomicAdd(&st->planes, c_planes);
        for (int r = 0; r < 3; ++r) {
            if (c_pts[r]) atomicAdd(&st->pts[r], c_pts[r]);
            if (c_lines[r]) atomicAdd(&st->lines[r], c_lines[r]);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// 7. deterministic velocity sums: fixed per-thread strides, fixed tree order
// ------------------------------------------------------------------------------------------------
constexpr int VSUM_BLOCKS = 256;

__global__ __launch_bounds__(THREADS) void k_vsum_partial(int64_t n, const double* __restrict__ vx,
                                                          const double* __restrict__ vy,
                                                          const double* __restrict__ vz, double* partial) {
    __shared__ double s[3][THREADS];
    double a = 0, b = 0, c = 0;
    const int64_t stride = static_cast<int64_t>(VSUM_BLOCKS) * THREADS;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x; i < n; i += stride) {
        a += vx[i];
        b += vy[i];
        c += vz[i];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partial[3 * blockIdx.x + k] = s[k][0];
}

__global__ void k_vsum_final(const double* partial, double* out) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; ++k) {
        double t = 0;
        for (int i = 0; i < VSUM_BLOCKS; ++i) t += partial[3 * i + k];
        out[k] = t;
    }
}

}  // namespace

// ================================================================================================
// context
// ================================================================================================

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct kiwi_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    DevBuf parts, bbox, nblk, offs, keys_a, keys_b, cub_tmp, ranges, stats, outbuf, vsum;
};

namespace {

int fail(kiwi_hip_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(kiwi_hip_ctx* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? KIWI_ERR_OUT_OF_MEMORY : KIWI_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define KIWI_HIP_CHECK(ctx, expr)                                \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
    } while (0)

hipError_t reserve(DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess) return e;
        b.p = nullptr;
        b.cap = 0;
    }
    const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) return e;
    b.cap = want;
    return hipSuccess;
}

}  // namespace

extern "C" {

int kiwi_hip_abi_version(void) { return KIWI_HIP_ABI_VERSION; }

int kiwi_hip_create(kiwi_hip_ctx** out, int device_ordinal) {
    if (!out) return KIWI_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KIWI_ERR_DEVICE;
    if (device_ordinal < 0 || device_ordinal >= ndev) return KIWI_ERR_INVALID_ARG;
    kiwi_hip_ctx* c = new kiwi_hip_ctx();
    c->device = device_ordinal;
    if (hipSetDevice(device_ordinal) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KIWI_ERR_DEVICE;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            kiwi_hip_destroy(c);
            return KIWI_ERR_DEVICE;
        }
    }
    *out = c;
    return KIWI_OK;
}

void kiwi_hip_destroy(kiwi_hip_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : {&c->parts, &c->bbox, &c->nblk, &c->offs, &c->keys_a, &c->keys_b, &c->cub_tmp, &c->ranges,
                      &c->stats, &c->outbuf, &c->vsum})
        if (b->p) (void)hipFree(b->p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* kiwi_hip_last_error(const kiwi_hip_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* kiwi_hip_stream(kiwi_hip_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int kiwi_hip_sph_to_grid(kiwi_hip_ctx* c, const kiwi_particles* P, const kiwi_grid* G, const kiwi_kernel* K,
                         int precision, const kiwi_cell_sums* out, kiwi_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!P || !G || !K || !out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (P->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative particle count");
    if (P->n > 0 && (!P->x || !P->y || !P->z || !P->vx || !P->vy || !P->vz || !P->m || !P->h))
        return fail(c, KIWI_ERR_INVALID_ARG, "null particle array");
    if (!out->mass || !out->px || !out->py || !out->pz) return fail(c, KIWI_ERR_INVALID_ARG, "null output array");
    if (G->nx <= 0 || G->ny <= 0 || G->nz <= 0)
        return fail(c, KIWI_ERR_INVALID_ARG, "cartesian_grid resolution (resolved nx/ny/nz) must be > 0");
    if (!(G->x_min < G->x_max) || !(G->y_min < G->y_max) || !(G->z_min < G->z_max))
        return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");
    if (precision != KIWI_FP64 && precision != KIWI_MIXED) return fail(c, KIWI_ERR_INVALID_ARG, "unknown precision");
    if (K->kernel_id < 0 || K->kernel_id > KIWI_KERNEL_WC6) return fail(c, KIWI_ERR_INVALID_ARG, "unknown kernel id");
    const double sup = K->support_in_h > 0 ? K->support_in_h : (K->kernel_id == KIWI_KERNEL_M4 ? 2.0 : 0.0);
    if (K->kernel_id != KIWI_KERNEL_M4 || K->generic_path || sup != 2.0 || K->tabulated)
        return fail(c, KIWI_ERR_UNSUPPORTED, "this build implements the analytic cubic-spline (M4, support 2h) path only");
    if (P->n > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 particles in one call");

    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = P->n;
    GridGeom g;
    g.x_min = G->x_min;
    g.y_min = G->y_min;
    g.z_min = G->z_min;
    g.dx = (G->x_max - G->x_min) / G->nx;
    g.dy = (G->y_max - G->y_min) / G->ny;
    g.dz = (G->z_max - G->z_min) / G->nz;
    g.nx = G->nx;
    g.ny = G->ny;
    g.nz = G->nz;
    g.nbx = (G->nx + BLK - 1) / BLK;
    g.nby = (G->ny + BLK - 1) / BLK;
    g.nbz = (G->nz + BLK - 1) / BLK;
    const int64_t nblocks = static_cast<int64_t>(g.nbx) * g.nby * g.nbz;
    const int64_t ncells = static_cast<int64_t>(G->nx) * G->ny * G->nz;
    if (nblocks > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "grid too large");

    // particle arrays on the device
    const double* dp[8];
    const double* hp[8] = {P->x, P->y, P->z, P->vx, P->vy, P->vz, P->m, P->h};
    if (P->on_device) {
        for (int k = 0; k < 8; ++k) dp[k] = hp[k];
    } else {
        KIWI_HIP_CHECK(c, reserve(c->parts, 8 * sizeof(double) * std::max<int64_t>(n, 1)));
        for (int k = 0; k < 8; ++k) {
            double* d = static_cast<double*>(c->parts.p) + k * n;
            if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(d, hp[k], n * sizeof(double), hipMemcpyHostToDevice, s));
            dp[k] = d;
        }
    }
    double* om[4];
    if (out->on_device) {
        om[0] = out->mass;
        om[1] = out->px;
        om[2] = out->py;
        om[3] = out->pz;
    } else {
        KIWI_HIP_CHECK(c, reserve(c->outbuf, 4 * sizeof(double) * ncells));
        for (int k = 0; k < 4; ++k) om[k] = static_cast<double*>(c->outbuf.p) + k * ncells;
    }

    KIWI_HIP_CHECK(c, reserve(c->stats, sizeof(DevStats)));
    DevStats* dst = static_cast<DevStats*>(c->stats.p);
    KIWI_HIP_CHECK(c, reserve(c->bbox, 6 * sizeof(int) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve(c->nblk, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->offs, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->ranges, 2 * sizeof(int) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->vsum, sizeof(double) * (3 * VSUM_BLOCKS + 3)));
    int* bbox = static_cast<int*>(c->bbox.p);
    int64_t* nblk = static_cast<int64_t*>(c->nblk.p);
    int64_t* offs = static_cast<int64_t*>(c->offs.p);
    int* blk_begin = static_cast<int*>(c->ranges.p);
    int* blk_end = blk_begin + nblocks;
    double* vpart = static_cast<double*>(c->vsum.p);

    KIWI_HIP_CHECK(c, hipMemsetAsync(dst, 0, sizeof(DevStats), s));
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    const double support = 2.0;
    // prep covers n + 1 threads: thread n writes the scan's trailing zero
    const int pgrid = static_cast<int>((n + 1 + THREADS - 1) / THREADS);
    k_prep<<<pgrid, THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], support, g, bbox, nblk, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    size_t tmp_bytes = 0;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, nblk, offs, static_cast<int>(n + 1), s));
    KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tmp_bytes));
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tmp_bytes, nblk, offs, static_cast<int>(n + 1), s));
    int64_t npairs = 0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&npairs, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (npairs > 0x7fffffffLL)
        return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 (particle, cell-block) work items in one call");

    KIWI_HIP_CHECK(c, reserve(c->keys_a, sizeof(uint64_t) * std::max<int64_t>(npairs, 1)));
    KIWI_HIP_CHECK(c, reserve(c->keys_b, sizeof(uint64_t) * std::max<int64_t>(npairs, 1)));
    uint64_t* keys_a = static_cast<uint64_t*>(c->keys_a.p);
    uint64_t* keys_b = static_cast<uint64_t*>(c->keys_b.p);
    if (n > 0) {
        k_emit<<<static_cast<int>((n + THREADS - 1) / THREADS), THREADS, 0, s>>>(n, bbox, offs, g, keys_a);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    int end_bit = 32;
    while ((1LL << (end_bit - 32)) < nblocks) ++end_bit;
    if (npairs > 0) {
        size_t sort_bytes = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys_a, keys_b,
                                                             static_cast<int>(npairs), 0, end_bit, s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, sort_bytes));
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(c->cub_tmp.p, sort_bytes, keys_a, keys_b,
                                                             static_cast<int>(npairs), 0, end_bit, s));
    }
    const int rgrid = static_cast<int>((2 * nblocks + THREADS - 1) / THREADS);
    k_fill_int<<<rgrid, THREADS, 0, s>>>(blk_begin, 2 * nblocks, 0);
    KIWI_HIP_CHECK(c, hipGetLastError());
    if (npairs > 0) {
        k_ranges<<<static_cast<int>((npairs + THREADS - 1) / THREADS), THREADS, 0, s>>>(npairs, keys_b, blk_begin,
                                                                                         blk_end);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
    if (precision == KIWI_FP64) {
        k_gather_m4<double><<<static_cast<int>(nblocks), THREADS, 0, s>>>(
            dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], bbox, keys_b, blk_begin, blk_end, g, om[0], om[1],
            om[2], om[3], dst, 1);
    } else {
        k_gather_m4<float><<<static_cast<int>(nblocks), THREADS, 0, s>>>(
            dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], bbox, keys_b, blk_begin, blk_end, g, om[0], om[1],
            om[2], om[3], dst, 0);
    }
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));
    k_vsum_partial<<<VSUM_BLOCKS, THREADS, 0, s>>>(n, dp[3], dp[4], dp[5], vpart);
    k_vsum_final<<<1, 64, 0, s>>>(vpart, vpart + 3 * VSUM_BLOCKS);
    KIWI_HIP_CHECK(c, hipGetLastError());

    if (!out->on_device) {
        double* ho[4] = {out->mass, out->px, out->py, out->pz};
        for (int k = 0; k < 4; ++k)
            KIWI_HIP_CHECK(c, hipMemcpyAsync(ho[k], om[k], ncells * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    DevStats hs;
    double vs[3];
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&hs, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipMemcpyAsync(vs, vpart + 3 * VSUM_BLOCKS, sizeof(vs), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->skipped_outside = static_cast<int64_t>(hs.skipped_outside);
        stats->n_negative = static_cast<int64_t>(hs.n_negative);
        stats->n_negative_warn = static_cast<int64_t>(hs.n_negative_warn);
        stats->pairs_bbox = static_cast<int64_t>(hs.pairs_bbox);
        stats->pairs_sphere = static_cast<int64_t>(hs.pairs_sphere);
        stats->particle_blocks = npairs;
        for (int k = 0; k < 3; ++k) stats->v_mean[k] = n > 0 ? vs[k] / static_cast<double>(n) : 0.0;
        float ms_prep = 0, ms_gather = 0;
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_gather, c->ev[1], c->ev[2]);
        stats->device_ms_prep = ms_prep;
        stats->device_ms_gather = ms_gather;
        for (int r = 0; r < 3; ++r) {
            stats->lattice_points[r] = static_cast<int64_t>(hs.pts[r]);
            stats->line_evals[r] = static_cast<int64_t>(hs.lines[r]);
        }
        stats->plane_evals = static_cast<int64_t>(hs.planes);
        const int nw = static_cast<int>(std::min<unsigned long long>(hs.n_warn, KIWI_MAX_WARN_RECORDS));
        stats->n_warn_records = nw;
        for (int k = 0; k < nw; ++k) stats->warn[k] = hs.warn[k];
    }
    return KIWI_OK;
}

}  // extern "C"
