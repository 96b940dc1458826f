// libkiwi_hip: the C ABI of include/kiwi_hip.h on MI355X (gfx950).
//
// sph-to-grid pipeline (one call of kiwi_hip_sph_to_grid, all on the context's stream):
//   1. k_prep        one thread per particle: the reference's bounding box of cells
//                    (src/cartesian_grid.cpp:209-238, FP64 floor, clamped), the number of cell blocks it
//                    overlaps and the size of its edge-line table; skipped / pair counters.
//   2. scans         exclusive prefix sums of both (hipcub).
//   3. k_emit        one 64-bit key (block << 32 | particle) per (particle, cell block) work item.
//   4. sort          radix sort of the keys: every block's particle list comes out in particle order,
//                    so each cell adds its particles in the reference's single-thread order.
//   5. k_ranges      [begin, end) of every block's list; k_block_work + a second sort give the
//                    blocks in order of decreasing work (longest-processing-time-first).
//   6. k_lines       one wavefront per particle: the region-matching constants (D2, D3) of every
//                    edge line of the particle's lattice, computed once per particle into HBM.
//   7. k_gather      one workgroup per cell block, blocks taken in work order.  Particles of the
//                    block's list are processed in batches: the batch's wall-plane constants and its
//                    slice of the line tables are staged in LDS, the lattice values W (m4_lattice.h)
//                    of all batch particles are evaluated as one flat task list, then every thread
//                    folds the 8-corner stencil of each batch particle into the cells it owns.  Each
//                    cell is owned by exactly one thread of one workgroup: no float atomics, and the
//                    results are bit-reproducible.
//   8. k_vsum        deterministic mean velocity over all particles (the empty-cell fallback).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/kiwi_hip.h"
#include "m4_lattice.h"

using namespace kiwi_hip;

namespace {

constexpr int THREADS = 256;

struct DevStats {
    unsigned long long skipped_outside, n_negative, n_negative_warn, pairs_bbox, pairs_sphere;
    unsigned long long n_warn;  // number of warn records claimed
    unsigned long long pts[3], lines[3], planes;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS];
};

struct GridGeom {
    double x_min, y_min, z_min, dx, dy, dz;
    int nx, ny, nz;
    int bsx, bsy, bsz;  // cell-block edge lengths
    int nbx, nby, nbz;  // blocks per axis
};

template <typename R>
struct Pair2 {
    R a, b;
};

// select one of three by a runtime axis index (no indexed private arrays: those would live in scratch)
template <typename T>
__device__ __forceinline__ T sel3(int a, T v0, T v1, T v2) { return a == 0 ? v0 : (a == 1 ? v1 : v2); }
__device__ __forceinline__ double gmin(const GridGeom& g, int a) { return sel3(a, g.x_min, g.y_min, g.z_min); }
__device__ __forceinline__ double gstep(const GridGeom& g, int a) { return sel3(a, g.dx, g.dy, g.dz); }

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
                       GridGeom g, int* __restrict__ bbox, int64_t* __restrict__ nblk, int64_t* __restrict__ nlines,
                       DevStats* st) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long skipped = 0, pairs = 0;
    if (i == n) {  // trailing zeros of the exclusive scans
        nblk[n] = 0;
        nlines[n] = 0;
    }
    if (i < n) {
        int64_t nb = 0, nl = 0;
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
                nb = static_cast<int64_t>(ix1 / g.bsx - ix0 / g.bsx + 1) * (iy1 / g.bsy - iy0 / g.bsy + 1) *
                     (iz1 / g.bsz - iz0 / g.bsz + 1);
                const int64_t Lx = ix1 - ix0 + 2, Ly = iy1 - iy0 + 2, Lz = iz1 - iz0 + 2;
                nl = 2 * (Lx * Ly + Lx * Lz + Ly * Lz);
            }
        }
        nblk[i] = nb;
        nlines[i] = nl;
    }
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
    for (int bx = bb[0] / g.bsx; bx <= bb[1] / g.bsx; ++bx)
        for (int by = bb[2] / g.bsy; by <= bb[3] / g.bsy; ++by)
            for (int bz = bb[4] / g.bsz; bz <= bb[5] / g.bsz; ++bz) {
                const uint64_t b = (static_cast<uint64_t>(bx) * g.nby + by) * g.nbz + bz;
                keys[o++] = (b << 32) | static_cast<uint64_t>(i);
            }
}

// ------------------------------------------------------------------------------------------------
// 5. block ranges in the sorted key array; work per block for the processing order
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_int(int* p, int64_t n, int v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_iota(int* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = static_cast<int>(i);
}

__global__ void k_ranges(int64_t npairs, const uint64_t* __restrict__ keys, int* __restrict__ begin,
                         int* __restrict__ end) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const uint32_t b = static_cast<uint32_t>(keys[i] >> 32);
    if (i == 0 || static_cast<uint32_t>(keys[i - 1] >> 32) != b) begin[b] = static_cast<int>(i);
    if (i == npairs - 1 || static_cast<uint32_t>(keys[i + 1] >> 32) != b) end[b] = static_cast<int>(i + 1);
}

// lattice points of a (particle, block) work item: the work estimate for the block order.  The keys
// are sorted by block, so equal blocks form runs of lanes; each run adds its total with one atomic.
__global__ void k_block_work(int64_t npairs, const uint64_t* __restrict__ keys, const int* __restrict__ bbox,
                             GridGeom g, unsigned long long* __restrict__ work) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < npairs;
    uint32_t b = 0xffffffffu;
    unsigned long long w = 0;
    if (valid) {
        b = static_cast<uint32_t>(keys[i] >> 32);
        const int p = static_cast<int>(keys[i] & 0xffffffffull);
        const int bz = b % g.nbz, by = (b / g.nbz) % g.nby, bx = b / (g.nbz * g.nby);
        const int* bb = bbox + 6 * p;
        const int nx = min(bb[1], bx * g.bsx + g.bsx - 1) - max(bb[0], bx * g.bsx) + 2;
        const int ny = min(bb[3], by * g.bsy + g.bsy - 1) - max(bb[2], by * g.bsy) + 2;
        const int nz = min(bb[5], bz * g.bsz + g.bsz - 1) - max(bb[4], bz * g.bsz) + 2;
        w = static_cast<unsigned long long>(nx * ny * nz + 64);
    }
    // inclusive segmented scan (segments = runs of equal b) towards higher lanes
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long wv = __shfl_up(w, off);
        const uint32_t bv = __shfl_up(b, off);
        if (lane >= off && bv == b) w += wv;
    }
    const uint32_t bnext = __shfl_down(b, 1);
    const bool last = (lane == 63) || (bnext != b);
    if (valid && last) atomicAdd(&work[b], w);
}

// ------------------------------------------------------------------------------------------------
// 6. edge-line constants, once per particle
// ------------------------------------------------------------------------------------------------
// Table layout per particle (lattice extents L = bbox cells + 1 per axis), six ordered (wall a,
// edge-normal b) blocks in the order (x,y) (x,z) (y,x) (y,z) (z,x) (z,y), each La x Lb with the
// a-index major: entry (ia, ib) holds (D2, D3) of the line at lattice coordinates (lo_a + ia, lo_b + ib).
template <typename R>
__global__ __launch_bounds__(THREADS) void k_lines(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const double* __restrict__ h,
                                                   const int* __restrict__ bbox, const int64_t* __restrict__ loffs,
                                                   GridGeom g, Pair2<R>* __restrict__ table, DevStats* st) {
    // grid-stride over particles, one wavefront per particle at a time; counters stay in registers
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (THREADS / 64);
    unsigned long long cnt0 = 0, cnt1 = 0, cnt2 = 0;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * (THREADS / 64) + (threadIdx.x >> 6); p < n; p += nwaves) {
        const int* bb = bbox + 6 * p;
        if (bb[0] > bb[1]) continue;
        const int lo0 = bb[0], lo1 = bb[2], lo2 = bb[4];
        const int L0 = bb[1] - bb[0] + 2, L1 = bb[3] - bb[2] + 2, L2 = bb[5] - bb[4] + 2;
        const int c01 = L0 * L1, c02 = L0 * L2, c12 = L1 * L2;
        const int nl = 2 * (c01 + c02 + c12);
        const double inv_h = 1.0 / h[p];
        const double p0 = x[p], p1 = y[p], p2 = z[p];
        Pair2<R>* out = table + loffs[p];
        for (int t = lane; t < nl; t += 64) {
            // pair blocks (x,y) (x,z) (y,x) (y,z) (z,x) (z,y) of sizes c01 c02 c01 c12 c02 c12
            int u = t, a, b;
            if (u < c01) { a = 0; b = 1; }
            else if ((u -= c01) < c02) { a = 0; b = 2; }
            else if ((u -= c02) < c01) { a = 1; b = 0; }
            else if ((u -= c01) < c12) { a = 1; b = 2; }
            else if ((u -= c12) < c02) { a = 2; b = 0; }
            else { u -= c02; a = 2; b = 1; }
            const int Lb = sel3(b, L0, L1, L2);
            const int ia = u / Lb, ib = u - ia * Lb;
            const double da = fabs((gmin(g, a) + (sel3(a, lo0, lo1, lo2) + ia) * gstep(g, a) - sel3(a, p0, p1, p2)) * inv_h);
            const double db = fabs((gmin(g, b) + (sel3(b, lo0, lo1, lo2) + ib) * gstep(g, b) - sel3(b, p0, p1, p2)) * inv_h);
            R D2, D3;
            m4_line<R>(static_cast<R>(da), static_cast<R>(db), D2, D3);
            const double l2 = da * da + db * db;
            if (l2 <= 1.0) ++cnt2;
            else if (l2 <= 4.0) ++cnt1;
            else ++cnt0;
            out[t] = Pair2<R>{D2, D3};
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cnt0 += __shfl_down(cnt0, off);
        cnt1 += __shfl_down(cnt1, off);
        cnt2 += __shfl_down(cnt2, off);
    }
    if (lane == 0) {
        if (cnt0) atomicAdd(&st->lines[0], cnt0);
        if (cnt1) atomicAdd(&st->lines[1], cnt1);
        if (cnt2) atomicAdd(&st->lines[2], cnt2);
    }
}

// ------------------------------------------------------------------------------------------------
// 7. the gather: one workgroup per cell block
// ------------------------------------------------------------------------------------------------
template <int BX, int BY, int BZ>
struct Shape {
    static constexpr int NX = BX + 1, NY = BY + 1, NZ = BZ + 1;      // lattice points per block edge
    static constexpr int CELLS = BX * BY * BZ;
    static constexpr int CPT = CELLS / THREADS;                      // cells per thread
    static constexpr int COLS = BY * BZ;                             // (y, z) columns of the block
    static constexpr int TPC = THREADS / COLS;                       // threads per column
    static constexpr int QMAX = 16;                                  // particles per batch
    static constexpr int PMAX = NX * NY * NZ > 1024 ? NX * NY * NZ : 1024;            // lattice values per batch
    static constexpr int LMAX_ONE = 2 * (NX * NY + NX * NZ + NY * NZ);
    static constexpr int LNMAX = LMAX_ONE > 1024 ? LMAX_ONE : 1024;                  // line entries per batch
    static constexpr int PLMAX = QMAX * (NX + NY + NZ);                              // plane entries per batch
    static_assert(CELLS % THREADS == 0, "block cells must be a multiple of the workgroup size");
    static_assert(THREADS % COLS == 0 && TPC * CPT == BX, "each thread owns a run of cells along x");
};

template <typename R, class S>
struct GatherLds {
    R U[S::PMAX];              // signed lattice values s * W, per batch particle, (i, j, k) k fastest
    R sd[S::PLMAX];            // signed offsets (lattice - particle) / h, per batch particle: x..., y..., z...
    union {
        struct {
            R pc[3][S::PLMAX];     // plane constants c1, c2, c3 of |sd|        (phases A, B)
            Pair2<R> ln[S::LNMAX]; // the batch's slices of the line tables     (phases A, B)
        } ab;
        double Ms[S::PMAX];        // Msum of every (batch particle, intersection cell)   (phases C1, C2)
    } u;
    int q_p[S::QMAX];
    int q_lo[3][S::QMAX];      // intersection low cell, block-local
    int q_n[3][S::QMAX];       // intersection lattice points per axis (cells + 1)
    int q_g[3][S::QMAX];       // intersection low lattice coordinate relative to the particle's bbox
    int q_L[3][S::QMAX];       // particle bbox lattice extents
    int q_pt[S::QMAX + 1], q_pl[S::QMAX + 1], q_ln[S::QMAX + 1], q_cl[S::QMAX + 1];
    float q_inv_nyz[S::QMAX], q_inv_nz[S::QMAX], q_inv_cyz[S::QMAX], q_inv_cz[S::QMAX];
    long long q_lbase[S::QMAX];
    double q_m[S::QMAX], q_v[3][S::QMAX], q_pos[3][S::QMAX], q_ih[S::QMAX];
    int s_nq;
};

// index of the batch particle owning flat task t, given the batch prefix offsets off[0..nq]
template <int QMAX>
__device__ __forceinline__ int owner(const int* off, int nq, int t) {
    int q = 0;
#pragma unroll
    for (int s = QMAX / 2; s > 0; s >>= 1)
        if (q + s < nq && off[q + s] <= t) q += s;
    return q;
}

// u / d for small non-negative integers through a float reciprocal, corrected to be exact
__device__ __forceinline__ int small_div(int u, int d, float inv_d) {
    int q = static_cast<int>((static_cast<float>(u) + 0.5f) * inv_d);
    const int r = u - q * d;
    q += (r >= d) ? 1 : 0;
    q -= (r < 0) ? 1 : 0;
    return q;
}

// distance component of the interval [lo, hi] (units of h) from the particle
template <typename R>
__device__ __forceinline__ R gap(R lo, R hi) {
    return (lo > R(0)) ? lo : ((hi < R(0)) ? -hi : R(0));
}

template <typename R, int BX, int BY, int BZ>
__global__ __launch_bounds__(THREADS) void k_gather(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, const double* __restrict__ h, const int* __restrict__ bbox,
    const uint64_t* __restrict__ keys, const int* __restrict__ blk_begin, const int* __restrict__ blk_end,
    const int* __restrict__ order, const int64_t* __restrict__ loffs, const Pair2<R>* __restrict__ lines, GridGeom g,
    double* __restrict__ out_mass, double* __restrict__ out_px, double* __restrict__ out_py,
    double* __restrict__ out_pz, DevStats* st, int clamp_pairs) {
    using S = Shape<BX, BY, BZ>;
    __shared__ GatherLds<R, S> L;
    const int b = order[blockIdx.x];
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int o[3] = {bx * BX, by * BY, bz * BZ};
    const int tid = threadIdx.x;
    // owned cells: the column (cy, cz) and a run of CPT cells along x starting at cx0
    const int col = tid % S::COLS;
    const int own_cy = col / BZ, own_cz = col % BZ;
    const int own_cx0 = (tid / S::COLS) * S::CPT;

    double acc[S::CPT][4];
#pragma unroll
    for (int s = 0; s < S::CPT; ++s) acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;
    unsigned long long n_neg = 0, n_warn = 0, n_sphere = 0;
    unsigned long long c_pts[3] = {0, 0, 0}, c_planes = 0;

    const int q1 = blk_end[b];
    int cursor = blk_begin[b];
    while (cursor < q1) {
        // ---- phase 0: the next batch (wave 0) ----
        if (tid < 64) {
            const int lane = tid;
            const int idx = cursor + lane;
            const bool valid = lane < S::QMAX && idx < q1;
            constexpr int BIG = 1 << 20;
            int npt = BIG, nln = BIG, npl = 0, ncl = 0, p = 0;
            int lo[3] = {0, 0, 0}, nn[3] = {0, 0, 0}, gg[3] = {0, 0, 0}, LL[3] = {0, 0, 0};
            if (valid) {
                p = static_cast<int>(keys[idx] & 0xffffffffull);
                const int* bb = bbox + 6 * p;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int blo = bb[2 * a], bhi = bb[2 * a + 1];
                    const int l0 = max(blo - o[a], 0), l1 = min(bhi - o[a], sel3(a, BX, BY, BZ) - 1);
                    lo[a] = l0;
                    nn[a] = l1 - l0 + 2;
                    gg[a] = o[a] + l0 - blo;
                    LL[a] = bhi - blo + 2;
                }
                npt = nn[0] * nn[1] * nn[2];
                nln = 2 * (nn[0] * nn[1] + nn[0] * nn[2] + nn[1] * nn[2]);
                npl = nn[0] + nn[1] + nn[2];
                ncl = (nn[0] - 1) * (nn[1] - 1) * (nn[2] - 1);
            }
            int ipt = npt, iln = nln, ipl = npl, icl = ncl;
            for (int off = 1; off < 64; off <<= 1) {
                const int a = __shfl_up(ipt, off), c = __shfl_up(iln, off), d = __shfl_up(ipl, off),
                          e = __shfl_up(icl, off);
                if (lane >= off) {
                    ipt += a;
                    iln += c;
                    ipl += d;
                    icl += e;
                }
            }
            const bool take = valid && ipt <= S::PMAX && iln <= S::LNMAX;
            const int nq = __popcll(__ballot(take));
            if (take) {
                L.q_p[lane] = p;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    L.q_lo[a][lane] = lo[a];
                    L.q_n[a][lane] = nn[a];
                    L.q_g[a][lane] = gg[a];
                    L.q_L[a][lane] = LL[a];
                }
                L.q_pt[lane] = ipt - npt;
                L.q_ln[lane] = iln - nln;
                L.q_pl[lane] = ipl - npl;
                L.q_cl[lane] = icl - ncl;
                L.q_inv_nyz[lane] = 1.0f / static_cast<float>(nn[1] * nn[2]);
                L.q_inv_nz[lane] = 1.0f / static_cast<float>(nn[2]);
                L.q_inv_cyz[lane] = 1.0f / static_cast<float>((nn[1] - 1) * (nn[2] - 1));
                L.q_inv_cz[lane] = 1.0f / static_cast<float>(nn[2] - 1);
                L.q_lbase[lane] = loffs[p];
                L.q_m[lane] = m[p];
                L.q_v[0][lane] = vx[p];
                L.q_v[1][lane] = vy[p];
                L.q_v[2][lane] = vz[p];
                L.q_pos[0][lane] = x[p];
                L.q_pos[1][lane] = y[p];
                L.q_pos[2][lane] = z[p];
                L.q_ih[lane] = 1.0 / h[p];
                if (lane == nq - 1) {
                    L.q_pt[nq] = ipt;
                    L.q_ln[nq] = iln;
                    L.q_pl[nq] = ipl;
                    L.q_cl[nq] = icl;
                }
            }
            if (lane == 0) L.s_nq = nq;
        }
        __syncthreads();
        const int nq = L.s_nq;
        cursor += nq;

        // ---- phase A: plane constants and the line-table slices ----
        const int npl_all = L.q_pl[nq];
        const int nln_all = L.q_ln[nq];
        for (int t = tid; t < npl_all + nln_all; t += THREADS) {
            if (t < npl_all) {
                const int q = owner<S::QMAX>(L.q_pl, nq, t);
                int u = t - L.q_pl[q], a = 0;
                if (u >= L.q_n[0][q]) {
                    u -= L.q_n[0][q];
                    a = 1;
                    if (u >= L.q_n[1][q]) {
                        u -= L.q_n[1][q];
                        a = 2;
                    }
                }
                const int gi = sel3(a, o[0], o[1], o[2]) + L.q_lo[a][q] + u;
                const double d = (gmin(g, a) + gi * gstep(g, a) - L.q_pos[a][q]) * L.q_ih[q];
                const R dr = static_cast<R>(d);
                const M4Plane<R> pl = m4_plane<R>(dr < R(0) ? -dr : dr);
                L.sd[t] = dr;
                L.u.ab.pc[0][t] = pl.c1;
                L.u.ab.pc[1][t] = pl.c2;
                L.u.ab.pc[2][t] = pl.c3;
                ++c_planes;
            } else {
                const int tl = t - npl_all;
                const int q = owner<S::QMAX>(L.q_ln, nq, tl);
                int u = tl - L.q_ln[q];
                const int n0 = L.q_n[0][q], n1 = L.q_n[1][q], n2 = L.q_n[2][q];
                const long long L0 = L.q_L[0][q], L1 = L.q_L[1][q], L2 = L.q_L[2][q];
                // pair blocks (x,y) (x,z) (y,x) (y,z) (z,x) (z,y): slice sizes n_a n_b, full-table sizes L_a L_b
                int a, bq;
                long long gbase;
                if (u < n0 * n1) { a = 0; bq = 1; gbase = 0; }
                else if ((u -= n0 * n1) < n0 * n2) { a = 0; bq = 2; gbase = L0 * L1; }
                else if ((u -= n0 * n2) < n1 * n0) { a = 1; bq = 0; gbase = L0 * L1 + L0 * L2; }
                else if ((u -= n1 * n0) < n1 * n2) { a = 1; bq = 2; gbase = 2 * L0 * L1 + L0 * L2; }
                else if ((u -= n1 * n2) < n2 * n0) { a = 2; bq = 0; gbase = 2 * L0 * L1 + L0 * L2 + L1 * L2; }
                else { u -= n2 * n0; a = 2; bq = 1; gbase = 2 * L0 * L1 + 2 * L0 * L2 + L1 * L2; }
                const int nb_ = sel3(bq, n0, n1, n2);
                const int ia = u / nb_, ib = u - ia * nb_;
                const long long gidx = L.q_lbase[q] + gbase +
                                       static_cast<long long>(L.q_g[a][q] + ia) * sel3(bq, L0, L1, L2) +
                                       (L.q_g[bq][q] + ib);
                L.u.ab.ln[tl] = lines[gidx];
            }
        }
        __syncthreads();

        // ---- phase B: lattice values of all batch particles ----
        const int npt_all = L.q_pt[nq];
        for (int t = tid; t < npt_all; t += THREADS) {
            const int q = owner<S::QMAX>(L.q_pt, nq, t);
            const int u = t - L.q_pt[q];
            const int nx = L.q_n[0][q], ny = L.q_n[1][q], nz = L.q_n[2][q];
            const int i = small_div(u, ny * nz, L.q_inv_nyz[q]);
            const int rem = u - i * (ny * nz);
            const int j = small_div(rem, nz, L.q_inv_nz[q]);
            const int k = rem - j * nz;
            const int pb = L.q_pl[q];
            const R dx = L.sd[pb + i], dy = L.sd[pb + nx + j], dz = L.sd[pb + nx + ny + k];
            R U = R(0);
            if (dx != R(0) && dy != R(0) && dz != R(0)) {
                // skip points that no cell of this intersection reaching the support sphere uses
                const R ex = gap(L.sd[pb + max(i - 1, 0)], L.sd[pb + min(i + 1, nx - 1)]);
                const R ey = gap(L.sd[pb + nx + max(j - 1, 0)], L.sd[pb + nx + min(j + 1, ny - 1)]);
                const R ez = gap(L.sd[pb + nx + ny + max(k - 1, 0)], L.sd[pb + nx + ny + min(k + 1, nz - 1)]);
                if (ex * ex + ey * ey + ez * ez < R(4)) {
                    const R ax = dx < R(0) ? -dx : dx, ay = dy < R(0) ? -dy : dy, az = dz < R(0) ? -dz : dz;
                    const R r2 = ax * ax + ay * ay + az * az;
                    const R r = KMath<R>::sqrt_(r2);
                    const int reg = m4_region(r2);
                    ++c_pts[reg - 1];
                    const R csum = L.u.ab.pc[reg - 1][pb + i] + L.u.ab.pc[reg - 1][pb + nx + j] +
                                   L.u.ab.pc[reg - 1][pb + nx + ny + k];
                    R dsum = R(0);
                    if (reg != 1) {
                        const int lb = L.q_ln[q];
                        const int pxy = lb, pxz = pxy + nx * ny, pyx = pxz + nx * nz, pyz = pyx + ny * nx,
                                  pzx = pyz + ny * nz, pzy = pzx + nz * nx;
                        const Pair2<R> l0 = L.u.ab.ln[pxy + i * ny + j], l1 = L.u.ab.ln[pxz + i * nz + k];
                        const Pair2<R> l2 = L.u.ab.ln[pyx + j * nx + i], l3 = L.u.ab.ln[pyz + j * nz + k];
                        const Pair2<R> l4 = L.u.ab.ln[pzx + k * nx + i], l5 = L.u.ab.ln[pzy + k * ny + j];
                        dsum = (reg == 2) ? (((l0.a + l1.a) + (l2.a + l3.a)) + (l4.a + l5.a))
                                          : (((l0.b + l1.b) + (l2.b + l3.b)) + (l4.b + l5.b));
                    }
                    const R W = m4_lattice_free<R>(ax, ay, az, r, reg) * M4Consts<R>::inv_pi + R(0.5) * csum + dsum;
                    const bool neg = (dx < R(0)) != ((dy < R(0)) != (dz < R(0)));
                    U = neg ? -W : W;
                }
            }
            L.U[t] = U;
        }
        __syncthreads();

        // ---- phase C1: Msum of every (batch particle, intersection cell), a flat task list ----
        const int ncl_all = L.q_cl[nq];
        for (int t = tid; t < ncl_all; t += THREADS) {
            const int q = owner<S::QMAX>(L.q_cl, nq, t);
            const int u = t - L.q_cl[q];
            const int nx = L.q_n[0][q], ny = L.q_n[1][q], nz = L.q_n[2][q];
            const int cyz = (ny - 1) * (nz - 1);
            const int i = small_div(u, cyz, L.q_inv_cyz[q]);
            const int rem = u - i * cyz;
            const int j = small_div(rem, nz - 1, L.q_inv_cz[q]);
            const int k = rem - j * (nz - 1);
            const int pb = L.q_pl[q];
            const R ex = gap(L.sd[pb + i], L.sd[pb + i + 1]);
            const R ey = gap(L.sd[pb + nx + j], L.sd[pb + nx + j + 1]);
            const R ez = gap(L.sd[pb + nx + ny + k], L.sd[pb + nx + ny + k + 1]);
            double Ms = 0.0;
            if (ex * ex + ey * ey + ez * ez < R(4)) {
                ++n_sphere;
                const R* U0 = &L.U[L.q_pt[q] + (i * ny + j) * nz + k];
                const int sx = ny * nz, sy = nz;
                Ms = static_cast<double>(U0[sx]) + static_cast<double>(U0[sy]) + static_cast<double>(U0[1]) +
                     static_cast<double>(U0[sx + sy + 1]) - static_cast<double>(U0[0]) -
                     static_cast<double>(U0[sx + sy]) - static_cast<double>(U0[sx + 1]) -
                     static_cast<double>(U0[sy + 1]);
                if (clamp_pairs && Ms < 0.0) {
                    ++n_neg;
                    if (Ms <= -0.01) {
                        ++n_warn;
                        const unsigned long long slot = atomicAdd(&st->n_warn, 1ull);
                        if (slot < KIWI_MAX_WARN_RECORDS) {
                            st->warn[slot].particle = L.q_p[q];
                            st->warn[slot].cell =
                                (static_cast<int64_t>(o[0] + L.q_lo[0][q] + i) * g.ny + (o[1] + L.q_lo[1][q] + j)) * g.nz +
                                (o[2] + L.q_lo[2][q] + k);
                            st->warn[slot].value = Ms;
                        }
                    }
                    Ms = 0.0;
                }
            }
            L.u.Ms[t] = Ms;
        }
        __syncthreads();

        // ---- phase C2: every thread adds its owned cells' Msum, batch particles in list order ----
        for (int q = 0; q < nq; ++q) {
            const int ly0 = L.q_lo[1][q], lz0 = L.q_lo[2][q];
            const int jy = own_cy - ly0, kz = own_cz - lz0;
            const int ny1 = L.q_n[1][q] - 1, nz1 = L.q_n[2][q] - 1;
            if (jy < 0 || jy >= ny1 || kz < 0 || kz >= nz1) continue;
            const int ix0 = own_cx0 - L.q_lo[0][q];
            const int nx1 = L.q_n[0][q] - 1;
            const double* Mq = &L.u.Ms[L.q_cl[q] + jy * nz1 + kz];
            const int sxc = ny1 * nz1;
            const double mm = L.q_m[q], v0 = L.q_v[0][q], v1 = L.q_v[1][q], v2 = L.q_v[2][q];
#pragma unroll
            for (int s = 0; s < S::CPT; ++s) {
                const int i = ix0 + s;
                if (i < 0 || i >= nx1) continue;
                const double Ms = Mq[i * sxc];
                if (Ms == 0.0) continue;
                const double t = Ms * mm;
                acc[s][0] += t;
                acc[s][1] += t * v0;
                acc[s][2] += t * v1;
                acc[s][3] += t * v2;
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < S::CPT; ++s) {
        const int ix = o[0] + own_cx0 + s, iy = o[1] + own_cy, iz = o[2] + own_cz;
        if (ix >= g.nx || iy >= g.ny || iz >= g.nz) continue;
        const int64_t c = (static_cast<int64_t>(ix) * g.ny + iy) * g.nz + iz;
        double mass = acc[s][0], p0 = acc[s][1], p1 = acc[s][2], p2 = acc[s][3];
        if (!clamp_pairs && mass < 0.0) {
            // mixed mode clamps the cell's total instead of every pair: a negative total is rounding
            // noise of an (almost) empty cell, which then takes the mean-velocity fallback like one
            mass = p0 = p1 = p2 = 0.0;
        }
        out_mass[c] = mass;
        out_px[c] = p0;
        out_py[c] = p1;
        out_pz[c] = p2;
    }
    for (int off = 32; off > 0; off >>= 1) {
        n_neg += __shfl_down(n_neg, off);
        n_warn += __shfl_down(n_warn, off);
        n_sphere += __shfl_down(n_sphere, off);
        c_planes += __shfl_down(c_planes, off);
        for (int r = 0; r < 3; ++r) c_pts[r] += __shfl_down(c_pts[r], off);
    }
    if ((tid & 63) == 0) {
        if (n_neg) atomicAdd(&st->n_negative, n_neg);
        if (n_warn) atomicAdd(&st->n_negative_warn, n_warn);
        if (n_sphere) atomicAdd(&st->pairs_sphere, n_sphere);
        if (c_planes) atomicAdd(&st->planes, c_planes);
        for (int r = 0; r < 3; ++r)
            if (c_pts[r]) atomicAdd(&st->pts[r], c_pts[r]);
    }
}

// ------------------------------------------------------------------------------------------------
// 8. deterministic velocity sums: fixed per-thread strides, fixed tree order
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

inline int grid1d(int64_t n) { return static_cast<int>((n + THREADS - 1) / THREADS); }

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
    int block_shape = 1;  // 0: 8x8x8 cells, 1: 16x16x8 cells (KIWI_HIP_BLOCK=8 or 16 overrides)
    DevBuf parts, bbox, nblk, nlin, offs, loffs, keys_a, keys_b, cub_tmp, ranges, work, work_sorted, order_in,
        order, lines, stats, outbuf, vsum;
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
    const size_t want = std::max<size_t>(bytes + bytes / 8, 256);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) return e;
    b.cap = want;
    return hipSuccess;
}

template <typename R, int BX, int BY, int BZ>
void launch_gather(int nblocks, hipStream_t s, const double* const* dp, const int* bbox, const uint64_t* keys,
                   const int* blk_begin, const int* blk_end, const int* order, const int64_t* loffs, const void* lines,
                   const GridGeom& g, double* const* om, DevStats* dst, int clamp) {
    k_gather<R, BX, BY, BZ><<<nblocks, THREADS, 0, s>>>(dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], bbox,
                                                         keys, blk_begin, blk_end, order, loffs,
                                                         static_cast<const Pair2<R>*>(lines), g, om[0], om[1], om[2],
                                                         om[3], dst, clamp);
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
    if (const char* e = std::getenv("KIWI_HIP_BLOCK")) c->block_shape = (std::atoi(e) == 8) ? 0 : 1;
    if (hipSetDevice(device_ordinal) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
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
    for (DevBuf* b : {&c->parts, &c->bbox, &c->nblk, &c->nlin, &c->offs, &c->loffs, &c->keys_a, &c->keys_b,
                      &c->cub_tmp, &c->ranges, &c->work, &c->work_sorted, &c->order_in, &c->order, &c->lines,
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
    const int bsx = c->block_shape ? 16 : 8, bsy = c->block_shape ? 16 : 8, bsz = 8;
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
    g.bsx = bsx;
    g.bsy = bsy;
    g.bsz = bsz;
    g.nbx = (G->nx + bsx - 1) / bsx;
    g.nby = (G->ny + bsy - 1) / bsy;
    g.nbz = (G->nz + bsz - 1) / bsz;
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
    KIWI_HIP_CHECK(c, reserve(c->nlin, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->offs, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->loffs, sizeof(int64_t) * (n + 1)));
    KIWI_HIP_CHECK(c, reserve(c->ranges, 2 * sizeof(int) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->work, sizeof(unsigned long long) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->work_sorted, sizeof(unsigned long long) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->order_in, sizeof(int) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->order, sizeof(int) * nblocks));
    KIWI_HIP_CHECK(c, reserve(c->vsum, sizeof(double) * (3 * VSUM_BLOCKS + 3)));
    int* bbox = static_cast<int*>(c->bbox.p);
    int64_t* nblk = static_cast<int64_t*>(c->nblk.p);
    int64_t* nlin = static_cast<int64_t*>(c->nlin.p);
    int64_t* offs = static_cast<int64_t*>(c->offs.p);
    int64_t* loffs = static_cast<int64_t*>(c->loffs.p);
    int* blk_begin = static_cast<int*>(c->ranges.p);
    int* blk_end = blk_begin + nblocks;
    unsigned long long* work = static_cast<unsigned long long*>(c->work.p);
    unsigned long long* work_sorted = static_cast<unsigned long long*>(c->work_sorted.p);
    int* order_in = static_cast<int*>(c->order_in.p);
    int* order = static_cast<int*>(c->order.p);
    double* vpart = static_cast<double*>(c->vsum.p);

    KIWI_HIP_CHECK(c, hipMemsetAsync(dst, 0, sizeof(DevStats), s));
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    const double support = 2.0;
    k_prep<<<grid1d(n + 1), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], support, g, bbox, nblk, nlin, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    size_t tmp_bytes = 0, tmp2 = 0;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, nblk, offs, static_cast<int>(n + 1), s));
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, nlin, loffs, static_cast<int>(n + 1), s));
    KIWI_HIP_CHECK(c, reserve(c->cub_tmp, std::max(tmp_bytes, tmp2)));
    tmp_bytes = c->cub_tmp.cap;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tmp_bytes, nblk, offs, static_cast<int>(n + 1), s));
    tmp_bytes = c->cub_tmp.cap;
    KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tmp_bytes, nlin, loffs, static_cast<int>(n + 1), s));
    int64_t totals[2] = {0, 0};
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&totals[0], offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&totals[1], loffs + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    const int64_t npairs = totals[0], nlines_total = totals[1];
    if (npairs > 0x7fffffffLL)
        return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 (particle, cell-block) work items in one call");

    const size_t line_bytes = (precision == KIWI_FP64 ? sizeof(Pair2<double>) : sizeof(Pair2<float>)) *
                              static_cast<size_t>(std::max<int64_t>(nlines_total, 1));
    KIWI_HIP_CHECK(c, reserve(c->lines, line_bytes));
    KIWI_HIP_CHECK(c, reserve(c->keys_a, sizeof(uint64_t) * std::max<int64_t>(npairs, 1)));
    KIWI_HIP_CHECK(c, reserve(c->keys_b, sizeof(uint64_t) * std::max<int64_t>(npairs, 1)));
    uint64_t* keys_a = static_cast<uint64_t*>(c->keys_a.p);
    uint64_t* keys_b = static_cast<uint64_t*>(c->keys_b.p);
    if (n > 0) {
        k_emit<<<grid1d(n), THREADS, 0, s>>>(n, bbox, offs, g, keys_a);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    int end_bit = 32;
    while ((1LL << (end_bit - 32)) < nblocks) ++end_bit;
    if (npairs > 0) {
        size_t sort_bytes = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys_a, keys_b,
                                                             static_cast<int>(npairs), 0, end_bit, s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, sort_bytes));
        sort_bytes = c->cub_tmp.cap;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(c->cub_tmp.p, sort_bytes, keys_a, keys_b,
                                                             static_cast<int>(npairs), 0, end_bit, s));
    }
    k_fill_int<<<grid1d(2 * nblocks), THREADS, 0, s>>>(blk_begin, 2 * nblocks, 0);
    k_fill_u64<<<grid1d(nblocks), THREADS, 0, s>>>(work, nblocks);
    k_iota<<<grid1d(nblocks), THREADS, 0, s>>>(order_in, nblocks);
    KIWI_HIP_CHECK(c, hipGetLastError());
    if (npairs > 0) {
        k_ranges<<<grid1d(npairs), THREADS, 0, s>>>(npairs, keys_b, blk_begin, blk_end);
        k_block_work<<<grid1d(npairs), THREADS, 0, s>>>(npairs, keys_b, bbox, g, work);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    {
        size_t ob = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, ob, work, work_sorted, order_in, order,
                                                                        static_cast<int>(nblocks), 0, 64, s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, ob));
        ob = c->cub_tmp.cap;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(c->cub_tmp.p, ob, work, work_sorted, order_in,
                                                                        order, static_cast<int>(nblocks), 0, 64, s));
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
    const int lgrid = static_cast<int>(std::min<int64_t>((n + 3) / 4, 2048));
    if (n > 0) {
        if (precision == KIWI_FP64)
            k_lines<double><<<lgrid, THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], bbox, loffs, g,
                                                      static_cast<Pair2<double>*>(c->lines.p), dst);
        else
            k_lines<float><<<lgrid, THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], bbox, loffs, g,
                                                     static_cast<Pair2<float>*>(c->lines.p), dst);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));
    const int nb = static_cast<int>(nblocks);
    const int clamp = (precision == KIWI_FP64) ? 1 : 0;
    if (precision == KIWI_FP64) {
        if (c->block_shape)
            launch_gather<double, 16, 16, 8>(nb, s, dp, bbox, keys_b, blk_begin, blk_end, order, loffs, c->lines.p, g,
                                              om, dst, clamp);
        else
            launch_gather<double, 8, 8, 8>(nb, s, dp, bbox, keys_b, blk_begin, blk_end, order, loffs, c->lines.p, g,
                                            om, dst, clamp);
    } else {
        if (c->block_shape)
            launch_gather<float, 16, 16, 8>(nb, s, dp, bbox, keys_b, blk_begin, blk_end, order, loffs, c->lines.p, g,
                                             om, dst, clamp);
        else
            launch_gather<float, 8, 8, 8>(nb, s, dp, bbox, keys_b, blk_begin, blk_end, order, loffs, c->lines.p, g,
                                           om, dst, clamp);
    }
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[3], s));
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
        float ms_prep = 0, ms_lines = 0, ms_gather = 0;
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_lines, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&ms_gather, c->ev[2], c->ev[3]);
        stats->device_ms_prep = ms_prep;
        stats->device_ms_lines = ms_lines;
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
