// libkiwi_hip: the C ABI of include/kiwi_hip.h on MI355X (gfx950).
//
// sph-to-grid pipeline (one call of kiwi_hip_sph_to_grid, all on the context's stream):
//   1. k_prep        one thread per particle: the reference's bounding box of cells
//                    (src/cartesian_grid.cpp:209-238, FP64 floor, clamped), the number of 8^3 cell
//                    blocks it overlaps and of edge lines of its lattice; counters.
//   2. scans         exclusive prefix sums of those sizes (hipcub).
//   3. k_emit        one item per (particle, cell block) pair, as the 64-bit key block << 32 | particle,
//                    written in particle order.
//   4. sort          radix sort of the keys by block (hipcub, stable): every block's item list comes
//                    out in particle order, so each cell adds its particles in the reference's
//                    single-thread order.  The sort carries the item's particle-order position along.
//   5. k_item_cells  cells of every item (its box = particle box intersected with the block); their
//                    prefix sum in block order places each item's cell sums in one contiguous slab per
//                    block; k_scatter hands every item its slab offset in particle order.
//      k_ranges, k_block_work + sort: [begin, end) of every block's items and the blocks in order of
//                    decreasing work (longest-processing-time-first) for the accumulation.
//   6. k_lines       one wavefront per particle: the region-matching constants (D2, D3) of every edge
//                    line of the particle's lattice, summed over the two lines through each pair of
//                    lattice coordinates, once per particle, into HBM.
//   7. k_particles   wavefronts take items in particle order from a shared cursor; per item: wall-plane
//                    constants and per-axis distance gaps in the wave's LDS, a compacted list of the
//                    lattice points any in-sphere cell needs, their lattice values W (m4_lattice.h),
//                    then the 8-corner stencil gives Msum of every cell, stored to the item's slab.
//                    Everything a wave touches in LDS is its own: no workgroup barriers.
//   8. k_accum       one workgroup per 8^3 cell block, blocks in work order: every thread owns two
//                    cells and adds Msum m and Msum m v of the block's items, read from the block's
//                    contiguous slab, in list order.  Each cell is owned by exactly one thread: no
//                    float atomics, and the results are bit-reproducible.
//   9. k_vsum        deterministic mean velocity over all particles (the empty-cell fallback; common.h).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "gen_lattice.h"
#include "m4_lattice.h"
#include "m4_table.h"
#include "radix_sort.h"

using namespace kiwi_hip;

// The allocation-free entry point's scratch (kiwi_hip_sph_to_grid_noalloc): a fixed device area that
// every buffer of its pipeline run is carved from.  Each device holds its own copy from the library's
// load until the process ends; every call on a device shares it.
static __device__ __align__(256) unsigned char scratch[KIWI_NOALLOC_SCRATCH_BYTES];

namespace {

// scratch's device address on each device (host code never passes the array's name itself)
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};
bool g_scratch_done[MAX_DEVICES] = {};

// called with the device current (the context's creation); false if the address cannot be had
bool scratch_address(int device) {
    if (device < 0 || device >= MAX_DEVICES) return false;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return false;
        g_scratch_ptr[device] = p;
        g_scratch_done[device] = true;
    }
    return g_scratch_ptr[device] != nullptr;
}

constexpr int BLK_PTS = (BLK + 1) * (BLK + 1) * (BLK + 1);
constexpr int AX_STRIDE = BLK + 1;  // LDS entries per axis in the cell kernel
constexpr int LINE_LDS_AX = 128;    // coordinates per axis the line kernel keeps in LDS
constexpr int MAX_AXIS = 8192;                // grid cells per axis
// The cell kernel's items are split into `cursors` contiguous slices (the context's launch shape,
// default 8), each with its own work cursor, from which a wavefront takes `item_grab` items at a time
// (default 16): one cursor serialises its atomics (about 12 ns apart), which capped a single cursor at
// ~80M grabs/s.  A workgroup starts on slice blockIdx % cursors (workgroups are dispatched round-robin
// over the 8 XCDs, so with 8 slices a slice's items, and the line tables of their particles, stay in
// one XCD's L2) and then takes from the other slices until all are drained.
constexpr int CURSOR_STRIDE = 544;  // 64-bit words between cursors (4352 B: separate lines)
// two sets: the FP32 items and the FP64 fallback's items of mixed precision
constexpr size_t CURSOR_BYTES = 2 * sizeof(unsigned long long) * KIWI_MAX_CURSORS * CURSOR_STRIDE;

struct DevStats {  // one per problem
    unsigned long long skipped_outside, n_negative, n_negative_warn, pairs_bbox, pairs_sphere;
    unsigned long long n_warn;  // number of warn records claimed
    unsigned long long pts[3], lines[3], planes;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS];
};

// One problem of a call (a single sph-to-grid call is a batch of one).  Particles of all problems are
// concatenated, and so are their cell blocks: problem q owns particles [part_off, part_off + n_parts)
// and blocks [block_off, block_off + nbx nby nbz).
struct Prob {
    GridGeom g;
    int64_t part_off, n_parts;
    int64_t block_off;
    double* out[4];         // device outputs: mass, px, py, pz
    const double* src[8];   // the caller's device columns x y z vx vy vz m h (null: host columns)
};

// A batch's device-resident particle columns concatenated into the pipeline's arrays (8 columns of n
// each) by one launch instead of 8 copies per problem: blockIdx.y is the problem.
__global__ __launch_bounds__(THREADS) void k_gather_parts(const Prob* __restrict__ pr, int64_t n,
                                                          double* __restrict__ dst) {
    const Prob& P = pr[blockIdx.y];
    if (!P.src[0]) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x; i < P.n_parts;
         i += static_cast<int64_t>(gridDim.x) * THREADS) {
#pragma unroll
        for (int k = 0; k < 8; ++k) dst[k * n + P.part_off + i] = P.src[k][i];
    }
}

// the problem that owns global block b (nprob is small; the search is wave-uniform)
__device__ __forceinline__ int block_problem(const Prob* __restrict__ pr, int nprob, int64_t b) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pr[mid].block_off <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// add a counter of each lane to its problem's slot: one atomic per wave when the wave's lanes share a
// problem (the common case: particles are grouped by problem), per lane otherwise
__device__ __forceinline__ void add_by_problem(unsigned long long v, int q, DevStats* st,
                                               unsigned long long DevStats::*field) {
    const int q0 = __shfl(q, 0);
    if (__all(q == q0)) {
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&(st[q0].*field), v);
    } else if (v && q >= 0) {
        atomicAdd(&(st[q].*field), v);
    }
}

// ------------------------------------------------------------------------------------------------
// 1. per-particle bounding box and work sizes
// ------------------------------------------------------------------------------------------------
// The accumulation's view of a particle: mass, velocity and cell box in one 64-byte record.  k_accum
// reads one per item in block order, that is at scattered particle positions: one cache line instead
// of one for each of five separate arrays.
struct alignas(64) AccPart {
    double m, v[3];
    int bb[6];
    int pad[2];
};

// Squared distance (units of h) between the particle's coordinate p and the cell range [c0, c1] along
// axis a, with the range's faces formed as the cell kernel forms its lattice coordinates, lo + c step.
// Every operation is rounded on its own, so that k_prep and k_emit take identical decisions.
__device__ __forceinline__ double range_gap2(const GridGeom& g, int a, int c0, int c1, double p, double inv_h) {
    const double lo = mul_sep(sub_sep(add_sep(gmin(g, a), mul_sep(static_cast<double>(c0), gstep(g, a))), p), inv_h);
    const double hi =
        mul_sep(sub_sep(add_sep(gmin(g, a), mul_sep(static_cast<double>(c1 + 1), gstep(g, a))), p), inv_h);
    const double d = gap(lo, hi);
    return mul_sep(d, d);
}

// Calls f(bx, by, bz) for every 8^3 cell block the particle's box [bb] overlaps that holds a cell its
// support sphere (radius sup h) reaches; with skip off, for every block of the box.  A block is left out
// only when the nearest point of its part of the box lies beyond the support by a relative margin of
// 1e-5 (squared), far above the rounding of the cell kernel's own in-sphere test: none of its cells
// passes that test, so every one of them would hold an exact zero and no counter would count it.
template <class F>
__device__ __forceinline__ void for_each_block(const int* bb, const GridGeom& g, double px, double py, double pz,
                                               double hh, double sup, bool skip, F f) {
    const double inv_h = 1.0 / hh;
    const double lim = mul_sep(mul_sep(sup, sup), 1.00001);
    for (int bx = bb[0] / BLK; bx <= bb[1] / BLK; ++bx) {
        const double gx = skip ? range_gap2(g, 0, max(bb[0], bx * BLK), min(bb[1], bx * BLK + BLK - 1), px, inv_h) : 0.0;
        if (gx >= lim) continue;
        for (int by = bb[2] / BLK; by <= bb[3] / BLK; ++by) {
            const double gxy =
                skip ? add_sep(gx, range_gap2(g, 1, max(bb[2], by * BLK), min(bb[3], by * BLK + BLK - 1), py, inv_h))
                     : 0.0;
            if (gxy >= lim) continue;
            for (int bz = bb[4] / BLK; bz <= bb[5] / BLK; ++bz) {
                if (skip &&
                    add_sep(gxy, range_gap2(g, 2, max(bb[4], bz * BLK), min(bb[5], bz * BLK + BLK - 1), pz, inv_h)) >=
                        lim)
                    continue;
                f(bx, by, bz);
            }
        }
    }
}

// Boxes of up to LANE_BLOCKS cell blocks are walked by their own lane; larger ones (a support of tens of
// cells spans thousands of blocks) by the whole wavefront, so that one large particle does not hold its
// wavefront for thousands of serial steps.
constexpr int LANE_BLOCKS = 8;

__device__ __forceinline__ int box_blocks(const int* bb) {
    return (bb[1] / BLK - bb[0] / BLK + 1) * (bb[3] / BLK - bb[2] / BLK + 1) * (bb[5] / BLK - bb[4] / BLK + 1);
}

// for_each_block for one particle by the whole wavefront (every argument wave-uniform, every lane
// calling): the lanes test 64 consecutive blocks of the box at a time, in for_each_block's order (x, then
// y, then z) and with its decisions; f(bx, by, bz, rank) runs for every block taken, rank = the number of
// blocks taken before it.  Returns the number of blocks taken.
template <class F>
__device__ __forceinline__ int for_each_block_wave(const int* bb, const GridGeom& g, double px, double py,
                                                   double pz, double hh, double sup, bool skip, F f) {
    const double inv_h = 1.0 / hh;
    const double lim = mul_sep(mul_sep(sup, sup), 1.00001);
    const int lane = static_cast<int>(threadIdx.x & 63);
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    const int bx0 = bb[0] / BLK, by0 = bb[2] / BLK, bz0 = bb[4] / BLK;
    const int nby = bb[3] / BLK - by0 + 1, nbz = bb[5] / BLK - bz0 + 1;
    const int nyz = nby * nbz;
    const int total = (bb[1] / BLK - bx0 + 1) * nyz;  // at most (8192 / 8)^3 blocks
    int taken = 0;
    for (int t0 = 0; t0 < total; t0 += 64) {
        const int t = t0 + lane;
        bool take = false;
        int bx = 0, by = 0, bz = 0;
        if (t < total) {
            const int tx = t / nyz, r = t - tx * nyz;
            const int ty = r / nbz;
            bx = bx0 + tx;
            by = by0 + ty;
            bz = bz0 + (r - ty * nbz);
            take = true;
            if (skip) {
                const double gx = range_gap2(g, 0, max(bb[0], bx * BLK), min(bb[1], bx * BLK + BLK - 1), px, inv_h);
                const double gxy =
                    add_sep(gx, range_gap2(g, 1, max(bb[2], by * BLK), min(bb[3], by * BLK + BLK - 1), py, inv_h));
                const double gxyz =
                    add_sep(gxy, range_gap2(g, 2, max(bb[4], bz * BLK), min(bb[5], bz * BLK + BLK - 1), pz, inv_h));
                take = !(gx >= lim) && !(gxy >= lim) && !(gxyz >= lim);
            }
        }
        const unsigned long long bal = __ballot(take);
        if (take) f(bx, by, bz, taken + __popcll(bal & lane_lt));
        taken += __popcll(bal);
    }
    return taken;
}

// the wavefront's lanes in `todo` one after the other: bb, the position and h of the lane's particle
// broadcast to every lane, then g(src, bb, px, py, pz, hh)
template <class G>
__device__ __forceinline__ void each_wave_particle(unsigned long long todo, const int* bb, double px, double py,
                                                   double pz, double hh, G g) {
    while (todo) {
        const int src = __ffsll(static_cast<long long>(todo)) - 1;
        todo &= todo - 1ull;
        int wb[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) wb[k] = __shfl(bb[k], src);
        g(src, wb, __shfl(px, src), __shfl(py, src), __shfl(pz, src), __shfl(hh, src));
    }
}

// Mixed precision's FP64 fallback: a particle whose support radius spans at least fb_cells of the
// smallest cell edge of its grid has its lattice values computed in FP64.  A cell's sum is a difference
// of O(1) lattice values while the cell holds O((cell / support)^3) of the particle, so the FP32 error of
// a cell sum, relative to the cell's share, grows with the cube of the support in cells.
__device__ __forceinline__ bool fp64_fallback(const GridGeom& g, double support_in_h, double h, double fb_cells) {
    return support_in_h * h >= fb_cells * fmin(g.dx, fmin(g.dy, g.dz));
}

__global__ void k_prep(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const double* __restrict__ h, const double* __restrict__ m,
                       const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
                       double support_in_h, int skip_outside, int axis_units, double fb_cells,
                       const Prob* __restrict__ pr, const int* __restrict__ part_prob, int* __restrict__ bbox,
                       AccPart* __restrict__ ap, int64_t* __restrict__ nblk, int64_t* __restrict__ nlines,
                       int64_t* __restrict__ nblk_fb, DevStats* st) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long skipped = 0, pairs = 0;
    const int q = (i < n) ? part_prob[i] : -1;
    if (i == n) {  // trailing zeros of the exclusive scans
        nblk[n] = 0;
        nlines[n] = 0;
        if (fb_cells > 0.0) nblk_fb[n] = 0;
    }
    int64_t nb = 0, nl = 0;
    bool fb = false, wide = false;
    int qb[6] = {1, 0, 0, 0, 0, 0};  // empty unless set below
    double hh = 0.0, x0 = 0.0, y0 = 0.0, z0 = 0.0;
    if (i < n) {
        hh = h[i];
        x0 = x[i];
        y0 = y[i];
        z0 = z[i];
        if (hh > 0.0) {
            int tb[6];
            if (!particle_bbox(pr[q].g, support_in_h, x0, y0, z0, hh, tb)) {
                skipped = 1;
            } else {
                for (int k = 0; k < 6; ++k) qb[k] = tb[k];
                const int64_t Nx = qb[1] - qb[0] + 1, Ny = qb[3] - qb[2] + 1, Nz = qb[5] - qb[4] + 1;
                pairs = static_cast<unsigned long long>(Nx * Ny * Nz);
                wide = box_blocks(qb) > LANE_BLOCKS;
                if (!wide)
                    for_each_block(qb, pr[q].g, x0, y0, z0, hh, support_in_h, skip_outside != 0,
                                   [&](int, int, int) { ++nb; });
                nl = (Nx + 1) * (Ny + 1) + (Nx + 1) * (Nz + 1) + (Ny + 1) * (Nz + 1);
                if (axis_units > 0) nl += axis_units - 1 + static_cast<int64_t>(axis_units) * (Nx + Ny + Nz + 3);
                if (fb_cells > 0.0) {
                    // with the FP64 fallback on, sizes count FP32 entries: a fallback particle's FP64 entries
                    // take two each, and every size is even, so that its region starts at an FP64 entry
                    fb = fp64_fallback(pr[q].g, support_in_h, hh, fb_cells);
                    if (fb) nl *= 2;
                    nl = (nl + 1) & ~static_cast<int64_t>(1);
                }
            }
        }
    }
    // the wide boxes' blocks, counted by the whole wavefront (every lane takes part)
    const int lane = static_cast<int>(threadIdx.x & 63);
    each_wave_particle(__ballot(wide), qb, x0, y0, z0, hh,
                       [&](int src, const int* wb, double wx, double wy, double wz, double wh) {
                           const int wq = __shfl(q, src);
                           const int cnt = for_each_block_wave(wb, pr[wq].g, wx, wy, wz, wh, support_in_h,
                                                               skip_outside != 0, [](int, int, int, int) {});
                           if (lane == src) nb = cnt;
                       });
    if (i < n) {
        int* bb = bbox + 6 * i;
        AccPart a;
        a.m = m[i];
        a.v[0] = vx[i];
        a.v[1] = vy[i];
        a.v[2] = vz[i];
        for (int k = 0; k < 6; ++k) bb[k] = a.bb[k] = qb[k];
        a.pad[0] = a.pad[1] = 0;
        ap[i] = a;
        nblk[i] = fb ? 0 : nb;
        if (fb_cells > 0.0) nblk_fb[i] = fb ? nb : 0;
        nlines[i] = nl;
    }
    add_by_problem(skipped, q, st, &DevStats::skipped_outside);
    add_by_problem(pairs, q, st, &DevStats::pairs_bbox);
}

// ------------------------------------------------------------------------------------------------
// 3. items: (particle, cell block) keys in particle order
// ------------------------------------------------------------------------------------------------
// the blocks k_prep counted, in the same order of decisions
__global__ void k_emit(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const double* __restrict__ h, double support_in_h,
                       int skip_outside, const int* __restrict__ bbox, const int64_t* __restrict__ offs,
                       const int64_t* __restrict__ offs_fb, const Prob* __restrict__ pr,
                       const int* __restrict__ part_prob, uint64_t* __restrict__ keys) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int qb[6] = {1, 0, 0, 0, 0, 0};
    int q = 0;
    int64_t o = 0, o_end = 0;
    double px = 0.0, py = 0.0, pz = 0.0, hh = 0.0;
    bool wide = false;
    if (i < n) {
        const int* bb = bbox + 6 * i;
        for (int k = 0; k < 6; ++k) qb[k] = bb[k];
        if (qb[0] <= qb[1]) {
            q = part_prob[i];
            px = x[i];
            py = y[i];
            pz = z[i];
            hh = h[i];
            // the FP64 fallback's items follow all the others (offs[n] of them), in particle order
            o = offs[i];
            o_end = offs[i + 1];
            if (offs_fb && offs_fb[i + 1] > offs_fb[i]) {
                o = offs[n] + offs_fb[i];
                o_end = offs[n] + offs_fb[i + 1];
            }
            wide = box_blocks(qb) > LANE_BLOCKS;
            if (!wide) {
                const Prob& P = pr[q];
                const GridGeom& g = P.g;
                for_each_block(qb, g, px, py, pz, hh, support_in_h, skip_outside != 0, [&](int bx, int by, int bz) {
                    if (o < o_end) {
                        const uint64_t b =
                            static_cast<uint64_t>(P.block_off) + (static_cast<uint64_t>(bx) * g.nby + by) * g.nbz + bz;
                        keys[o++] = (b << 32) | static_cast<uint64_t>(i);
                    }
                });
            }
        }
    }
    // the wide boxes by the whole wavefront: consecutive lanes write consecutive keys
    const int64_t i0 = i - static_cast<int64_t>(threadIdx.x & 63);  // the wavefront's first particle
    each_wave_particle(__ballot(wide), qb, px, py, pz, hh,
                       [&](int src, const int* wb, double wx, double wy, double wz, double wh) {
                           const Prob& P = pr[__shfl(q, src)];
                           const GridGeom& g = P.g;
                           const int64_t wo = __shfl(o, src), wend = __shfl(o_end, src);
                           const uint64_t wi = static_cast<uint64_t>(i0 + src);
                           for_each_block_wave(wb, g, wx, wy, wz, wh, support_in_h, skip_outside != 0,
                                               [&](int bx, int by, int bz, int rank) {
                                                   if (wo + rank < wend) {
                                                       const uint64_t b = static_cast<uint64_t>(P.block_off) +
                                                                          (static_cast<uint64_t>(bx) * g.nby + by) * g.nbz + bz;
                                                       keys[wo + rank] = (b << 32) | wi;
                                                   }
                                               });
                       });
}

// ------------------------------------------------------------------------------------------------
// 5. item boxes, slab offsets, block ranges, work per block
// ------------------------------------------------------------------------------------------------
struct ItemBox {
    int x0, y0, z0;  // first cell of the item (grid coordinates)
    int mx, my, mz;  // cells per axis, 1..8
};

// the item's cells: the particle's box intersected with the block
__device__ __forceinline__ ItemBox item_box(const int* bb, uint32_t b, const GridGeom& g) {
    const int bz = static_cast<int>(b % static_cast<uint32_t>(g.nbz));
    const uint32_t bxy = b / static_cast<uint32_t>(g.nbz);
    const int by = static_cast<int>(bxy % static_cast<uint32_t>(g.nby));
    const int bx = static_cast<int>(bxy / static_cast<uint32_t>(g.nby));
    ItemBox r;
    r.x0 = max(bb[0], bx * BLK);
    r.y0 = max(bb[2], by * BLK);
    r.z0 = max(bb[4], bz * BLK);
    r.mx = min(bb[1], bx * BLK + BLK - 1) - r.x0 + 1;
    r.my = min(bb[3], by * BLK + BLK - 1) - r.y0 + 1;
    r.mz = min(bb[5], bz * BLK + BLK - 1) - r.z0 + 1;
    return r;
}

// cells of every item in block order (the sizes whose prefix sum gives the slab offsets)
__global__ void k_item_cells(int64_t npairs, const uint64_t* __restrict__ keys, const int* __restrict__ bbox,
                             const Prob* __restrict__ pr, const int* __restrict__ part_prob,
                             int64_t* __restrict__ cells) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == npairs) cells[npairs] = 0;
    if (i >= npairs) return;
    const uint64_t key = keys[i];
    const uint32_t p = static_cast<uint32_t>(key & 0xffffffffull);
    const Prob& P = pr[part_prob[p]];
    const ItemBox r = item_box(bbox + 6 * p, static_cast<uint32_t>((key >> 32) - P.block_off), P.g);
    cells[i] = static_cast<int64_t>(r.mx * r.my * r.mz);
}

// slab offset of every item at its particle-order position
__global__ void k_scatter(int64_t npairs, const int* __restrict__ pos, const int64_t* __restrict__ off_sorted,
                          int64_t* __restrict__ off_by_particle) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < npairs) off_by_particle[pos[i]] = off_sorted[i];
}

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

// cells of an item plus a fixed per-item cost: the work estimate for the block order.  The keys are
// sorted by block, so equal blocks form runs of lanes; each run adds its total with one atomic.
__global__ void k_block_work(int64_t npairs, const uint64_t* __restrict__ keys, const int64_t* __restrict__ cells,
                             unsigned long long* __restrict__ work) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < npairs;
    uint32_t b = 0xffffffffu;
    unsigned long long w = 0;
    if (valid) {
        b = static_cast<uint32_t>(keys[i] >> 32);
        w = static_cast<unsigned long long>(cells[i] + 32);
    }
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
// Kernel policies: what the line and cell kernels need to know about the kernel function
// ------------------------------------------------------------------------------------------------
// M4 analytic (m4_lattice.h): lattice offsets in units of h, support 2; regions inside h, inside 2h,
// outside; plane constants c1..c3; the line table holds the region-2 and region-3 offsets (region 1
// has none).  R = double is the parity mode, R = float the mixed mode.
template <typename R>
struct PolicyM4 {
    using Real = R;
    static constexpr int WAVES_PER_EU = sizeof(R) == 4 ? 8 : 1;  // the cell kernel's occupancy request
    static constexpr int NPL = 3;  // plane values per lattice coordinate
    static constexpr int NLN = 2;  // line-table values per axis pair
    static constexpr bool HAS_LINES = true;
    static constexpr bool SPHERE_SKIP = true;  // cells and points beyond the support are exact zeros
    struct Ctx {
        double support_in_h;  // support / h = 2
    };
    static __device__ __forceinline__ double unit_in_h(const Ctx&) { return 1.0; }  // lattice unit: h
    static __device__ __forceinline__ R support2() { return R(4); }
    static __device__ __forceinline__ int region(const Ctx&, R r2) { return m4_region(r2); }
    static __device__ __forceinline__ void plane(const Ctx&, R ad, R* v) {
        const M4Plane<R> pl = m4_plane<R>(ad);
        v[0] = pl.c1;
        v[1] = pl.c2;
        v[2] = pl.c3;
    }
    // line-table slot of region k, -1: none
    static __device__ __forceinline__ int line_slot(int k) { return k - 2; }
    static __device__ __forceinline__ void line(const Ctx&, R da, R db, R* v) { m4_line<R>(da, db, v[0], v[1]); }
    // the table entry of a pair of coordinates from their |d| and plane values: the two lines' offsets
    // summed, shared terms formed once (m4_line_pair)
    static constexpr bool FUSED_PAIRS = true;
    static __device__ __forceinline__ void line_pair(const Ctx&, R da, const R* pa, R db, const R* pb, R* v) {
        m4_line_pair<R>(da, db, pa[0] - pa[1], pa[1] - pa[2], pb[0] - pb[1], pb[1] - pb[2], v[0], v[1]);
    }
    static __device__ __forceinline__ int line_kind(double l2) { return l2 < 1.0 ? 2 : (l2 < 4.0 ? 1 : 0); }
    // the lattice value without its plane and line terms, and the weight of the plane sum
    static __device__ __forceinline__ R free_part(const Ctx&, R ax, R ay, R az, R r, int k) {
        if (k == 3) return R(-0.25);  // -(pi/4) / pi: outside the support the free part is constant
        return m4_lattice_free<R>(ax, ay, az, r, k) * M4Consts<R>::inv_pi;
    }
    static __device__ __forceinline__ R plane_weight() { return R(0.5); }
};

// Generic tabulated-coefficient kernels (gen_lattice.h): lattice offsets in units of the support H,
// support 1; regions of the kernel table; plane values Phi_k; the line table holds the offsets of
// every region.  R = double is the parity mode; R = float (mixed) is used for the kernels whose
// single-precision lattice values are as accurate as the cubic spline's (gen_mixed_ok) and, with the FP64
// fallback on, for M5 below its scaled fallback threshold (gen_forward_fp32), with the plane
// values still formed in double and rounded.  KID >= 0: kernel KID's table compiled into the code
// (gen_fixed_tab); KID = -1: the context's table, read at run time.
template <typename R, int KID = -1>
struct PolicyGen {
    using Real = R;
    static constexpr int WAVES_PER_EU = 1;
    static constexpr int NPL = GEN_NREG;
    static constexpr int NLN = GEN_NREG;
    static constexpr bool HAS_LINES = true;
    static constexpr bool SPHERE_SKIP = true;
    struct Ctx {
        double support_in_h;  // H / h
        const GenTab* tab;
    };
    static __device__ __forceinline__ const GenTab& tab(const Ctx& c) {
        if constexpr (KID >= 0)
            return gen_fixed_tab<KID>();
        else
            return *c.tab;
    }
    static __device__ __forceinline__ double unit_in_h(const Ctx& c) { return c.support_in_h; }  // lattice unit: H
    static __device__ __forceinline__ R support2() { return R(1); }
    static __device__ __forceinline__ int region(const Ctx& c, R r2) {
        return gen_region_of(tab(c), static_cast<double>(KMath<R>::sqrt_(r2)));
    }
    static __device__ __forceinline__ void plane(const Ctx& c, R ad, R* v) {
        double Phi[GEN_NREG];
        gen_plane_phi(tab(c), static_cast<double>(ad), Phi);
#pragma unroll
        for (int k = 0; k < GEN_NREG; ++k) v[k] = static_cast<R>(Phi[k]);
    }
    static __device__ __forceinline__ int line_slot(int k) { return k - 1; }
    static __device__ __forceinline__ void line(const Ctx& c, R da, R db, R* v) { gen_line<R>(tab(c), da, db, v); }
    static constexpr bool FUSED_PAIRS = false;
    static __device__ __forceinline__ void line_pair(const Ctx&, R, const R*, R, const R*, R*) {}
    static __device__ __forceinline__ int line_kind(double l2) { return l2 < 1.0 ? 1 : 0; }
    static __device__ __forceinline__ R free_part(const Ctx& c, R ax, R ay, R az, R r, int k) {
        if constexpr (KID >= 0)
            return gen_lattice_free_fixed<R, KID>(ax, ay, az, r, k);
        else
            return gen_lattice_free<R>(gen_regions<R>(*c.tab)[k - 1], ax, ay, az, r);
    }
    static __device__ __forceinline__ R plane_weight() { return R(0.5 * GEN_PI); }
};

// Calls f(P{}) with the generic policy P of kernel kid in precision R, its table compiled in.  MIXED:
// only the kernels that mixed precision computes in FP32 (gen_forward_fp32), whose FP64 fallback is the
// double-precision policy; single precision implies it.
template <typename R, bool MIXED = std::is_same<R, float>::value, class F>
hipError_t with_gen_policy(int kid, F&& f) {
    switch (kid) {
        case KIWI_KERNEL_M4:
            return f(PolicyGen<R, KIWI_KERNEL_M4>{});
        case KIWI_KERNEL_M5:
            return f(PolicyGen<R, KIWI_KERNEL_M5>{});
        case KIWI_KERNEL_WC2:
            return f(PolicyGen<R, KIWI_KERNEL_WC2>{});
        default:
            break;
    }
    if constexpr (!MIXED) {
        switch (kid) {
            case KIWI_KERNEL_M6:
                return f(PolicyGen<R, KIWI_KERNEL_M6>{});
            case KIWI_KERNEL_WC4:
                return f(PolicyGen<R, KIWI_KERNEL_WC4>{});
            case KIWI_KERNEL_WC6:
                return f(PolicyGen<R, KIWI_KERNEL_WC6>{});
            default:
                break;
        }
    }
    return hipErrorInvalidValue;
}

// the FP64 fallback's policy: with_gen_policy for the kernels mixed precision computes in FP32
template <class F>
hipError_t with_fallback_gen_policy(int kid, F&& f) {
    return with_gen_policy<double, true>(kid, std::forward<F>(f));
}

// Tabulated cubic spline (m4_table.h): the lattice value is the plain sum of the 6 tabulated vertex
// integrals anchored at the point (no plane or line terms).  The table's interpolation error does not
// cancel outside the support, so, as in the reference, every bounding-box cell is evaluated.
struct PolicyTab {
    using Real = double;
    static constexpr int WAVES_PER_EU = 1;
    static constexpr int NPL = 1;  // unused
    static constexpr int NLN = 1;  // unused: the tabulated path has no line tables
    static constexpr bool HAS_LINES = false;
    static constexpr bool SPHERE_SKIP = false;
    struct Ctx {
        double support_in_h;
        const double* table;
    };
    static __device__ __forceinline__ double unit_in_h(const Ctx&) { return 1.0; }
    static __device__ __forceinline__ double support2() { return 4.0; }
    static __device__ __forceinline__ int region(const Ctx&, double r2) { return m4_region(r2); }
    static __device__ __forceinline__ void plane(const Ctx&, double, double* v) { v[0] = 0.0; }
    static __device__ __forceinline__ int line_slot(int) { return -1; }
    static __device__ __forceinline__ void line(const Ctx&, double, double, double* v) { v[0] = 0.0; }
    static __device__ __forceinline__ int line_kind(double) { return 0; }
    static __device__ __forceinline__ double free_part(const Ctx& c, double ax, double ay, double az, double, int) {
        // pyramid (wall a, edge normal b, along c): r0 = d_a, R0 = d_b, cos(phi) = d_b / sqrt(d_b^2 + d_c^2)
        const double rxy = sqrt(ax * ax + ay * ay), rxz = sqrt(ax * ax + az * az), ryz = sqrt(ay * ay + az * az);
        const double* T = c.table;
        return ((m4_gridded_value(T, ax, ay, ay / ryz) + m4_gridded_value(T, ax, az, az / ryz)) +
                (m4_gridded_value(T, ay, ax, ax / rxz) + m4_gridded_value(T, ay, az, az / rxz))) +
               (m4_gridded_value(T, az, ax, ax / rxy) + m4_gridded_value(T, az, ay, ay / rxy));
    }
    static __device__ __forceinline__ double plane_weight() { return 0.0; }
};

// the lookup table (src/kernel_integral.cpp:57-119): entry (i, j, k), i = 1..nr0, j < nR0, k < nph
__global__ void k_table_build(double* __restrict__ tab) {
    using G = M4TableGeom;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = static_cast<int64_t>(G::nr0) * G::nR0 * G::nph;
    if (t >= n) return;
    const int k = static_cast<int>(t % G::nph);
    const int j = static_cast<int>((t / G::nph) % G::nR0);
    const int i = static_cast<int>(t / (static_cast<int64_t>(G::nph) * G::nR0)) + 1;
    const double r0 = G::r0_of(i), mu0 = G::mu0_of(j);
    const double R0 = r0 * sqrt(1.0 - mu0 * mu0) / mu0;  // j = 0: +inf, on purpose
    tab[G::at(i, j, k)] = m4_vertex_value(r0, R0, acos(G::cphi_of(k)));
}

// row nr0 + 1 repeats row nr0 (the interpolation reads i + 1 of the last row)
__global__ void k_table_pad(double* __restrict__ tab) {
    using G = M4TableGeom;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G::nR0 * G::nph) return;
    const int j = t / G::nph, k = t % G::nph;
    tab[G::at(G::nr0 + 1, j, k)] = tab[G::at(G::nr0, j, k)];
}

template <typename R, int N>
struct LineVals {
    R v[N];
};

// Per-axis values of a particle's lattice coordinates: per axis a the L_a = box cells + 1 coordinates,
// axes x, y, z one after the other.  k_lines writes them after the particle's line-pair table (at the
// next multiple of the entry size); the cell kernel copies each item's window into its LDS.
template <typename R, int NPL>
struct alignas(8 * sizeof(R)) AxisVals {
    R d;        // signed offset (lattice - particle) / scale
    R cg2;      // squared gap between the particle and the cell [u, u + 1] (unused for the last coordinate)
    R pv[NPL];  // plane values of |d| (M4: c1..c3; generic: Phi_1..Phi_4)
};

// line-table entries per axis entry (0: the path has no line tables)
template <class P>
constexpr int axis_units() {
    if constexpr (P::HAS_LINES)
        return static_cast<int>(sizeof(AxisVals<typename P::Real, P::NPL>) /
                                sizeof(LineVals<typename P::Real, P::NLN>));
    else
        return 0;
}

// A particle's region of the line table: its line pairs (La Lb entries for each axis pair), then up to
// units - 1 entries of padding, then its axis table.  Returns the axis table's first entry (in line-
// table entries from the start of the table).
__device__ __forceinline__ int64_t axis_table_start(int64_t loff, int npairs_p, int units) {
    return (loff + npairs_p + units - 1) / units * units;
}

// ------------------------------------------------------------------------------------------------
// 6. edge-line constants, once per particle
// ------------------------------------------------------------------------------------------------
// Table layout per particle (lattice extents L = bbox cells + 1 per axis): three unordered axis pairs
// (x,y) (x,z) (y,z), each La x Lb with the a-index major.  Entry (ia, ib) of pair {a, b} holds, per
// line-table slot, the sum of the region offsets of the two edge lines through lattice coordinates
// (lo_a + ia, lo_b + ib): wall a with edge normal b, and wall b with edge normal a.  A lattice point in
// region k adds slot k of its three axis pairs.
template <class P>
__global__ __launch_bounds__(THREADS) void k_lines(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const double* __restrict__ h,
                                                   const int* __restrict__ bbox, const int64_t* __restrict__ loffs,
                                                   const int64_t* __restrict__ sel, int loff_shift,
                                                   const Prob* __restrict__ pr, const int* __restrict__ part_prob,
                                                   typename P::Ctx ctx,
                                                   LineVals<typename P::Real, P::NLN>* __restrict__ table,
                                                   DevStats* st) {
    using R = typename P::Real;
    // grid-stride over particles, one wavefront per particle at a time; counters stay in registers and
    // go to the problem's statistics when the wavefront moves on to another problem
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * WAVES;
    constexpr int UNITS = axis_units<P>();
    struct LAx {
        R ad;          // |d|
        R pv[P::NPL];  // plane values of |d|
    };
    __shared__ LAx LDS_AX[WAVES][P::FUSED_PAIRS ? 3 * LINE_LDS_AX : 1];
    LAx* const LA = LDS_AX[wave_uniform(threadIdx.x >> 6)];
    unsigned long long cnt0 = 0, cnt1 = 0, cnt2 = 0;
    unsigned long long c_planes = 0;  // wave total (uniform)
    int cur_q = -1;
    auto flush = [&]() {
        for (int off = 32; off > 0; off >>= 1) {
            cnt0 += __shfl_down(cnt0, off);
            cnt1 += __shfl_down(cnt1, off);
            cnt2 += __shfl_down(cnt2, off);
        }
        if (lane == 0 && cur_q >= 0) {
            if (cnt0) atomicAdd(&st[cur_q].lines[0], cnt0);
            if (cnt1) atomicAdd(&st[cur_q].lines[1], cnt1);
            if (cnt2) atomicAdd(&st[cur_q].lines[2], cnt2);
            if (c_planes) atomicAdd(&st[cur_q].planes, c_planes);
        }
        cnt0 = cnt1 = cnt2 = c_planes = 0;
    };
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * WAVES + wave_uniform(threadIdx.x >> 6); p < n; p += nwaves) {
        const int* bb = bbox + 6 * p;
        if (bb[0] > bb[1]) continue;
        // with the FP64 fallback on, each precision's launch takes only the particles with items of its own
        // (sel: their item counts)
        if (sel && sel[p] == 0) continue;
        const int q = static_cast<int>(wave_uniform(static_cast<uint32_t>(part_prob[p])));
        if (q != cur_q) {
            flush();
            cur_q = q;
        }
        const GridGeom& g = pr[q].g;
        const int64_t loff = loffs[p] >> loff_shift;  // in this precision's entries
        const int L0 = bb[1] - bb[0] + 2, L1 = bb[3] - bb[2] + 2, L2 = bb[5] - bb[4] + 2;
        const int c01 = L0 * L1, c02 = L0 * L2, c12 = L1 * L2;
        const int nl = c01 + c02 + c12;
        const float inv1 = __builtin_amdgcn_rcpf(static_cast<float>(L1)), inv2 = __builtin_amdgcn_rcpf(static_cast<float>(L2));
        const double inv_s = 1.0 / (P::unit_in_h(ctx) * h[p]);  // lattice units
        const double p0 = x[p], p1 = y[p], p2 = z[p];
        LineVals<R, P::NLN>* out = table + loff;
        // fused pairs take |d| and the plane values of both coordinates from the wave's LDS copy of the
        // axis table (particles with at most LINE_LDS_AX coordinates per axis)
        const bool use_lds = P::FUSED_PAIRS && L0 <= LINE_LDS_AX && L1 <= LINE_LDS_AX && L2 <= LINE_LDS_AX;

        // the axis table: offsets formed as the reference forms vertex coordinates, (x_min + i dx) - x_p
        AxisVals<R, P::NPL>* at = reinterpret_cast<AxisVals<R, P::NPL>*>(table + axis_table_start(loff, nl, UNITS));
        const int na = L0 + L1 + L2;
        for (int t = lane; t < na; t += 64) {
            int a = 0, u = t;
            if (u >= L0) {
                u -= L0;
                a = 1;
                if (u >= L1) {
                    u -= L1;
                    a = 2;
                }
            }
            const int g0 = sel3(a, bb[0], bb[2], bb[4]);
            const double lo = gmin(g, a), pa = sel3(a, p0, p1, p2), st_a = gstep(g, a);
            const double d = (lo + (g0 + u) * st_a - pa) * inv_s;
            const double dn = (lo + (g0 + u + 1) * st_a - pa) * inv_s;
            AxisVals<R, P::NPL> e;
            e.d = static_cast<R>(d);
            const R gc = static_cast<R>(gap(d, dn));
            e.cg2 = gc * gc;
            const R ad = e.d < R(0) ? -e.d : e.d;
            P::plane(ctx, ad, e.pv);
            at[t] = e;
            if (use_lds) {
                LAx& la = LA[a * LINE_LDS_AX + u];
                la.ad = ad;
#pragma unroll
                for (int k = 0; k < P::NPL; ++k) la.pv[k] = e.pv[k];
            }
        }
        c_planes += static_cast<unsigned>(na);
        if (use_lds) wave_sync();

        for (int t = lane; t < nl; t += 64) {
            // pair blocks (x,y) (x,z) (y,z) of sizes c01 c02 c12
            int u = t, a = 0, b = 1, Lb = L1;
            float inv_b = inv1;
            if (u >= c01) {
                u -= c01;
                b = 2;
                Lb = L2;
                inv_b = inv2;
                if (u >= c02) {
                    u -= c02;
                    a = 1;
                }
            }
            const int ia = small_div(u, Lb, inv_b), ib = u - ia * Lb;
            LineVals<R, P::NLN> e;
            int kind;
            if (use_lds) {
                const LAx& A = LA[a * LINE_LDS_AX + ia];
                const LAx& B = LA[b * LINE_LDS_AX + ib];
                P::line_pair(ctx, A.ad, A.pv, B.ad, B.pv, e.v);
                kind = P::line_kind(static_cast<double>(A.ad) * A.ad + static_cast<double>(B.ad) * B.ad);
            } else {
                const double da = fabs((gmin(g, a) + (sel3(a, bb[0], bb[2], bb[4]) + ia) * gstep(g, a) -
                                        sel3(a, p0, p1, p2)) * inv_s);
                const double db = fabs((gmin(g, b) + (sel3(b, bb[0], bb[2], bb[4]) + ib) * gstep(g, b) -
                                        sel3(b, p0, p1, p2)) * inv_s);
                R vab[P::NLN], vba[P::NLN];
                P::line(ctx, static_cast<R>(da), static_cast<R>(db), vab);
                P::line(ctx, static_cast<R>(db), static_cast<R>(da), vba);
#pragma unroll
                for (int k = 0; k < P::NLN; ++k) e.v[k] = vab[k] + vba[k];
                kind = P::line_kind(da * da + db * db);
            }
            cnt2 += (kind == 2) ? 2 : 0;
            cnt1 += (kind == 1) ? 2 : 0;
            cnt0 += (kind == 0) ? 2 : 0;
            out[t] = e;
        }
        if (use_lds) wave_sync();  // the LDS copy is rewritten for the next particle
    }
    flush();
}

// ------------------------------------------------------------------------------------------------
// 7. lattice values and cell sums, one wavefront per item
// ------------------------------------------------------------------------------------------------
// Per-wave LDS.  Axis arrays hold the item's lattice coordinates x..., y..., z... (n_x + n_y + n_z
// <= 27 entries, n = cells + 1 per axis).
template <typename R, int NPL>
struct WaveLds {
    // Axis arrays at a fixed stride: x coordinates at [0, 9), y at [9, 18), z at [18, 27) (an item has at
    // most 9 lattice coordinates per axis), so that every read of them takes an immediate offset.
    Pair2<R> sg[3 * AX_STRIDE];  // a: signed offset (lattice - particle) / scale; b: squared gap to the
                                 // cells on either side of the lattice plane within the item (point test)
    R pc[NPL][3 * AX_STRIDE];    // plane values of |offset| (M4: c1..c3; generic: Phi_1..Phi_4)
    R cg2[3 * AX_STRIDE];        // squared gap to the cell slab [u, u+1] (cell test; last entry unused)
    uint16_t list[BLK_PTS];      // needed points: i | j << 4 | k << 8; inside the support from the
                                 // front, outside from the back
    R U[BLK_PTS];                // signed lattice values s * W, (i, j, k) k fastest
};

// ceil(2^16 / d) for d = 1..81.  With m = ceil(2^16 / d) = (2^16 + e) / d, 0 <= e < d:
// (u m) >> 16 = floor(u / d + u e / (d 2^16)), which is floor(u / d) when (u mod d) + u e / 2^16 < d;
// for 0 <= u <= 64, u e < 64 * 81 < 2^16 makes that hold.  tests/test_cell_kernel_tables.py checks the
// table and the identity for every (u, d) the walks use.
__constant__ uint32_t kMagic16[82] = {
    0,     65536, 32768, 21846, 16384, 13108, 10923, 9363, 8192, 7282, 6554, 5958, 5462, 5042,
    4682,  4370,  4096,  3856,  3641,  3450,  3277,  3121, 2979, 2850, 2731, 2622, 2521, 2428,
    2341,  2260,  2185,  2115,  2048,  1986,  1928,  1873, 1821, 1772, 1725, 1681, 1639, 1599,
    1561,  1525,  1490,  1457,  1425,  1395,  1366,  1338, 1311, 1286, 1261, 1237, 1214, 1192,
    1171,  1150,  1130,  1111,  1093,  1075,  1058,  1041, 1024, 1009, 993,  979,  964,  950,
    937,   924,   911,   898,   886,   874,   863,   852,  841,  830,  820,  810};

// Lane-strided walk over an n0 x n1 x n2 box (index u = (i n1 + j) n2 + k): the lane's first point
// and the (i, j, k) step of 64 points, so that each iteration advances with adds and compares.
// n1 n2 <= 81; the divisions are multiplications by kMagic16.
struct BoxWalk {
    int i, j, k;
    int di, dj, dk;
    __device__ __forceinline__ BoxWalk(int lane, int n1, int n2) {
        const int n12 = n1 * n2;
        const uint32_t m12 = kMagic16[n12], m2 = kMagic16[n2];
        i = static_cast<int>((static_cast<uint32_t>(lane) * m12) >> 16);
        const int r = lane - i * n12;
        j = static_cast<int>((static_cast<uint32_t>(r) * m2) >> 16);
        k = r - j * n2;
        di = static_cast<int>((64u * m12) >> 16);
        const int s = 64 - di * n12;
        dj = static_cast<int>((static_cast<uint32_t>(s) * m2) >> 16);
        dk = s - dj * n2;
    }
    __device__ __forceinline__ void next(int n1, int n2) {
        k += dk;
        j += dj;
        if (k >= n2) {
            k -= n2;
            ++j;
        }
        i += di;
        if (j >= n1) {
            j -= n1;
            ++i;
        }
    }
};

// the value of lane - 1 within each row of 16 lanes (DPP row_shr:1); the first lane of a row keeps its own
template <typename R>
__device__ __forceinline__ R row_prev(R v) {
    if constexpr (sizeof(R) == 4) {
        const int b = __builtin_bit_cast(int, v);
        return __builtin_bit_cast(R, __builtin_amdgcn_update_dpp(b, b, 0x111, 0xf, 0xf, false));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = static_cast<int>(b & 0xffffffffll), hi = static_cast<int>(b >> 32);
        const int plo = __builtin_amdgcn_update_dpp(lo, lo, 0x111, 0xf, 0xf, false);
        const int phi = __builtin_amdgcn_update_dpp(hi, hi, 0x111, 0xf, 0xf, false);
        return __builtin_bit_cast(R, (static_cast<long long>(phi) << 32) | static_cast<uint32_t>(plo));
    }
}

// occupancy: the mixed cubic path is fitted to 8 wavefronts per SIMD (64 VGPRs, 20 KB of LDS per
// workgroup); the cell kernel waits on memory for about a third of its wavefront cycles, and more
// resident wavefronts hide it
// Items [first, first + npairs) of the particle-order item list; cell sums are stored as OutR (the
// FP64 fallback of mixed precision computes in double and stores float, like the FP32 items).
template <class P, typename OutR>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(P::WAVES_PER_EU))) void k_particles(
    uint32_t first, uint32_t npairs, const uint64_t* __restrict__ keys, const int64_t* __restrict__ slab,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ h, const int* __restrict__ bbox, const int64_t* __restrict__ loffs, int loff_shift,
    const Prob* __restrict__ pr, const int* __restrict__ part_prob, typename P::Ctx ctx,
    const LineVals<typename P::Real, P::NLN>* __restrict__ lines, OutR* __restrict__ ms, DevStats* st,
    unsigned long long* cursor, int grab, int ncurs) {
    using R = typename P::Real;
    // the reference clamps every (particle, cell) sum at zero (src/cartesian_grid.cpp:264-269): FP64 mode
    // does the same; mixed precision (float per-pair sums) clamps each cell's total instead (k_accum)
    constexpr bool CLAMP = std::is_same<OutR, double>::value;
    constexpr int NLN = P::NLN;
    constexpr int UNITS = axis_units<P>();
    __shared__ WaveLds<R, P::NPL> WL[WAVES];
    const int wave = static_cast<int>(wave_uniform(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    WaveLds<R, P::NPL>& L = WL[wave];
    const R S2 = P::support2();
    // per-lane counters (clamped negatives, rare) and wavefront totals (counted from ballots; lane 0 takes
    // part in every pass of the loops that count them, so lane 0's copies are the wave's totals)
    unsigned long long n_neg = 0, n_warn = 0;
    unsigned long long n_sphere = 0, c_planes = 0;
    unsigned long long c_pts[3] = {0, 0, 0};
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    // counters go to the problem's statistics when the wavefront moves on to another problem
    int cur_q = -1;
    auto flush = [&]() {
        for (int off = 32; off > 0; off >>= 1) {
            n_neg += __shfl_down(n_neg, off);
            n_warn += __shfl_down(n_warn, off);
        }
        if (lane == 0 && cur_q >= 0) {
            DevStats* sq = st + cur_q;
            if (n_neg) atomicAdd(&sq->n_negative, n_neg);
            if (n_warn) atomicAdd(&sq->n_negative_warn, n_warn);
            if (n_sphere) atomicAdd(&sq->pairs_sphere, n_sphere);
            if (c_planes) atomicAdd(&sq->planes, c_planes);
            c_pts[1] -= c_pts[0] + c_pts[2];  // c_pts[1] counted every evaluated point
            for (int r = 0; r < 3; ++r)
                if (c_pts[r]) atomicAdd(&sq->pts[r], c_pts[r]);
        }
        n_neg = n_warn = n_sphere = c_planes = 0;
        c_pts[0] = c_pts[1] = c_pts[2] = 0;
    };
    // the lane's axis entry in the set-up of an item: axis lane / 16, coordinate lane % 16
    const int ax = lane >> 4, au = lane & 15;

    // the current particle's values (a wavefront's items come in particle order: most share them)
    uint32_t cur_p = 0xffffffffu;
    int bb[6] = {0, 0, 0, 0, 0, 0};
    int Lx = 0, Ly = 0, Lz = 0;
    double px = 0.0, py = 0.0, pz = 0.0, inv_s = 0.0;
    const R* etab = nullptr;
    const AxisVals<R, P::NPL>* atab = nullptr;

    const int home = static_cast<int>(blockIdx.x % static_cast<unsigned>(ncurs));
    const unsigned long long ugrab = static_cast<unsigned long long>(grab);
    for (int d = 0; d < ncurs; ++d) {
        const int sl = (home + d) % ncurs;
        const uint64_t lo = first + static_cast<uint64_t>(npairs) * sl / ncurs;
        const uint64_t hi = first + static_cast<uint64_t>(npairs) * (sl + 1) / ncurs;
        unsigned long long* cs = cursor + sl * CURSOR_STRIDE;
        uint64_t cur = 0;
        if (lane == 0) cur = lo + atomicAdd(cs, ugrab);
        cur = wave_uniform64(cur);
        while (cur < hi) {
            uint64_t nxt = 0;
            if (lane == 0) nxt = lo + atomicAdd(cs, ugrab);
            const uint32_t tend = static_cast<uint32_t>(min(cur + ugrab, hi));
            for (uint32_t t = static_cast<uint32_t>(cur); t < tend; ++t) {
                const uint64_t key = keys[t];
                const uint32_t p = static_cast<uint32_t>(key & 0xffffffffull);
                const int q = static_cast<int>(wave_uniform(static_cast<uint32_t>(part_prob[p])));
                if (q != cur_q) {
                    flush();
                    cur_q = q;
                }
                const GridGeom& g = pr[q].g;
                if (p != cur_p) {
                    cur_p = p;
                    const int* pb = bbox + 6 * static_cast<size_t>(p);
                    for (int k = 0; k < 6; ++k) bb[k] = pb[k];
                    Lx = bb[1] - bb[0] + 2;  // particle lattice
                    Ly = bb[3] - bb[2] + 2;
                    Lz = bb[5] - bb[4] + 2;
                    px = x[p];
                    py = y[p];
                    pz = z[p];
                    inv_s = 1.0 / (P::unit_in_h(ctx) * h[p]);  // lattice units
                    if constexpr (P::HAS_LINES) {
                        const int npairs_p = Lx * Ly + Lx * Lz + Ly * Lz;
                        const int64_t lo_p = loffs[p] >> loff_shift;  // in this precision's entries
                        etab = reinterpret_cast<const R*>(lines + lo_p);
                        atab = reinterpret_cast<const AxisVals<R, P::NPL>*>(lines + axis_table_start(lo_p, npairs_p, UNITS));
                    }
                }
                const ItemBox ib = item_box(bb, static_cast<uint32_t>((key >> 32) - pr[q].block_off), g);
                const int mx = ib.mx, my = ib.my, mz = ib.mz;
                const int nx = mx + 1, ny = my + 1, nz = mz + 1;  // item lattice
                const int ox = ib.x0 - bb[0], oy = ib.y0 - bb[2], oz = ib.z0 - bb[4];
                OutR* out = ms + slab[t];

                // ---- per-axis constants of the item's lattice planes ----
                {
                    const int n_a = (ax == 0) ? nx : ((ax == 1) ? ny : nz);
                    if (ax < 3 && au < n_a) {
                        R dr, cg2v, pv[P::NPL];
                        if constexpr (P::HAS_LINES) {
                            // the item's window of the particle's axis table
                            const int o_a = (ax == 0) ? ox : ((ax == 1) ? oy : oz);
                            const int base = (ax == 0) ? 0 : ((ax == 1) ? Lx : Lx + Ly);
                            const AxisVals<R, P::NPL> e = atab[base + o_a + au];
                            dr = e.d;
                            cg2v = e.cg2;
#pragma unroll
                            for (int k = 0; k < P::NPL; ++k) pv[k] = e.pv[k];
                        } else {
                            // offsets formed as the reference forms vertex coordinates: (x_min + i dx) - x_p
                            const int g0 = sel3(ax, ib.x0, ib.y0, ib.z0);
                            const double lo_a = gmin(g, ax), pa = sel3(ax, px, py, pz), st_a = gstep(g, ax);
                            const double dd = (lo_a + (g0 + au) * st_a - pa) * inv_s;
                            const double dn = (lo_a + (g0 + au + 1) * st_a - pa) * inv_s;
                            dr = static_cast<R>(dd);
                            const R gc = static_cast<R>(gap(dd, dn));
                            cg2v = gc * gc;
                            P::plane(ctx, dr < R(0) ? -dr : dr, pv);
                        }
                        // point test: the gap of the cells on either side of the plane within the item,
                        // min(gap of cell u - 1, gap of cell u), clamped at the item's first and last plane
                        const R prev = row_prev(cg2v);  // cell u - 1 (the first plane keeps its own)
                        const R gp2 = (au == n_a - 1) ? prev : (prev < cg2v ? prev : cg2v);
                        const int slot = ax * AX_STRIDE + au;
                        L.sg[slot] = Pair2<R>{dr, gp2};
#pragma unroll
                        for (int k = 0; k < P::NPL; ++k) L.pc[k][slot] = pv[k];
                        L.cg2[slot] = cg2v;
                    }
                }
                if constexpr (!P::HAS_LINES) c_planes += static_cast<unsigned>(nx + ny + nz);
                wave_sync();

                // ---- classify the lattice points: not needed, inside the support, outside ----
                // (points in a plane through the particle are listed like the others; their value is 0)
                const int nyz = ny * nz, npts = nx * nyz;
                uint32_t nA = 0, nB = 0;
                BoxWalk w(lane, ny, nz);
                for (int u0 = 0; u0 < npts; u0 += 64, w.next(ny, nz)) {
                    const int u = u0 + lane;
                    const Pair2<R> ex = L.sg[w.i], ey = L.sg[AX_STRIDE + w.j], ez = L.sg[2 * AX_STRIDE + w.k];
                    const bool needed = u < npts && (!P::SPHERE_SKIP || ex.b + ey.b + ez.b < S2);
                    const R r2 = ex.a * ex.a + ey.a * ey.a + ez.a * ez.a;
                    const bool isA = needed && r2 < S2, isB = needed && !(r2 < S2);
                    const unsigned long long bA = __ballot(isA), bB = __ballot(isB);
                    const uint16_t code = static_cast<uint16_t>(w.i | (w.j << 4) | (w.k << 8));
                    if (isA) L.list[nA + __popcll(bA & lane_lt)] = code;
                    if (isB) L.list[BLK_PTS - 1 - (nB + __popcll(bB & lane_lt))] = code;
                    nA += __popcll(bA);
                    nB += __popcll(bB);
                }
                wave_sync();

                // ---- lattice values of the needed points (inside the support first, then outside) ----
                // line-table blocks (x,y) (x,z) (y,z) of the particle's table
                const int exz = Lx * Ly, eyz = exz + Lx * Lz;
                const uint32_t ntot = nA + nB;
                for (uint32_t v = lane; v < ntot; v += 64) {
                    const uint32_t code = (v < nA) ? L.list[v] : L.list[BLK_PTS - 1 - (v - nA)];
                    const int i = static_cast<int>(code & 15u), j = static_cast<int>((code >> 4) & 15u),
                              k = static_cast<int>(code >> 8);
                    const int u = (i * ny + j) * nz + k;
                    const R dx = L.sg[i].a, dy = L.sg[AX_STRIDE + j].a, dz = L.sg[2 * AX_STRIDE + k].a;
                    const R ax_ = dx < R(0) ? -dx : dx, ay = dy < R(0) ? -dy : dy, az = dz < R(0) ? -dz : dz;
                    const R r2 = ax_ * ax_ + ay * ay + az * az;
                    const R r = KMath<R>::sqrt_(r2);
                    const int reg = P::region(ctx, r2);
                    // lane 0 takes part in every pass of this loop, so its counters hold the wave's totals
                    // (region 2 is the rest: c_pts[1] gets the item's total below)
                    c_pts[0] += __popcll(__ballot(reg == 1));
                    c_pts[2] += __popcll(__ballot(reg >= 3));
                    // plane and line terms: the analytic paths only (the tabulated path keeps one unused plane
                    // row, so its regions must not index the rows)
                    R csum = R(0);
                    if constexpr (P::HAS_LINES)
                        csum = L.pc[reg - 1][i] + L.pc[reg - 1][AX_STRIDE + j] + L.pc[reg - 1][2 * AX_STRIDE + k];
                    R dsum = R(0);
                    const int slot = P::line_slot(reg);
                    if (slot >= 0) {
                        const int gi = ox + i, gj = oy + j, gk = oz + k;  // particle-lattice coordinates
                        const uint32_t c = static_cast<uint32_t>(slot);
                        dsum = (etab[NLN * static_cast<uint32_t>(gi * Ly + gj) + c] +
                                etab[NLN * static_cast<uint32_t>(exz + gi * Lz + gk) + c]) +
                               etab[NLN * static_cast<uint32_t>(eyz + gj * Lz + gk) + c];
                    }
                    const R W = P::free_part(ctx, ax_, ay, az, r, reg) + P::plane_weight() * csum + dsum;
                    const bool neg = (dx < R(0)) != ((dy < R(0)) != (dz < R(0)));
                    // every pyramid anchored in a plane through the particle is flat
                    const bool flat = dx == R(0) || dy == R(0) || dz == R(0);
                    L.U[u] = flat ? R(0) : (neg ? -W : W);
                }
                c_pts[1] += ntot;  // less the region-1 and region-3 points counted above (flush)
                wave_sync();

                // ---- cell sums: the 8-corner stencil, stored to the item's slab ----
                // the reference's per-pair clamp at zero, with its warning records (FP64 mode)
                auto clamp_cell = [&](double& Ms, int i, int j, int k) {
                    if (CLAMP && Ms < 0.0) {
                        ++n_neg;
                        if (Ms <= -0.01) {
                            ++n_warn;
                            DevStats* sq = st + cur_q;
                            const unsigned long long wslot = atomicAdd(&sq->n_warn, 1ull);
                            if (wslot < KIWI_MAX_WARN_RECORDS) {
                                sq->warn[wslot].particle = static_cast<int64_t>(p) - pr[cur_q].part_off;
                                sq->warn[wslot].cell =
                                    (static_cast<int64_t>(ib.x0 + i) * g.ny + (ib.y0 + j)) * g.nz + (ib.z0 + k);
                                sq->warn[wslot].value = Ms;
                            }
                        }
                        Ms = 0.0;
                    }
                };
                const int cyz = my * mz;
                if (cyz > 32) {
                    // x-walk: lane (j, k) of the y-z face walks its row of cells along x.  A cell's stencil is
                    // the difference of the face sums E(i + 1) - E(i), E(i) = U[i][j+1][k+1] - U[i][j+1][k]
                    // - U[i][j][k+1] + U[i][j][k] (sums of four floats, exact in double), so every face sum
                    // is formed once instead of twice.  Stores stay contiguous: cells (i, j, k) k fastest.
                    const bool act = lane < cyz;
                    const int j = act ? static_cast<int>((static_cast<uint32_t>(lane) * kMagic16[mz]) >> 16) : 0;
                    const int k = act ? lane - j * mz : 0;
                    const R cgy = L.cg2[AX_STRIDE + j], cgz = L.cg2[2 * AX_STRIDE + k];
                    const R* Uc = &L.U[j * nz + k];
                    auto face = [&](int i) {
                        const R* q = Uc + i * nyz;
                        return (static_cast<double>(q[nz + 1]) + static_cast<double>(q[0])) -
                               static_cast<double>(q[nz]) - static_cast<double>(q[1]);
                    };
                    double Ep = face(0);
                    for (int i = 0; i < mx; ++i) {
                        const double En = face(i + 1);
                        const bool in_sphere = act && (!P::SPHERE_SKIP || L.cg2[i] + cgy + cgz < S2);
                        n_sphere += __popcll(__ballot(in_sphere));  // wave total, held by lane 0 as above
                        double Ms = 0.0;
                        if (in_sphere) {
                            Ms = En - Ep;
                            clamp_cell(Ms, i, j, k);
                        }
                        if (act) out[i * cyz + lane] = static_cast<OutR>(Ms);
                        Ep = En;
                    }
                } else {
                    const int ncells = mx * cyz;
                    BoxWalk wc(lane, my, mz);
                    for (int u = lane; u < ncells; u += 64, wc.next(my, mz)) {
                        const int i = wc.i, j = wc.j, k = wc.k;
                        double Ms = 0.0;
                        const bool in_sphere =
                            !P::SPHERE_SKIP || L.cg2[i] + L.cg2[AX_STRIDE + j] + L.cg2[2 * AX_STRIDE + k] < S2;
                        n_sphere += __popcll(__ballot(in_sphere));  // wave total, held by lane 0 as above
                        if (in_sphere) {
                            const R* U0 = &L.U[(i * ny + j) * nz + k];
                            Ms = static_cast<double>(U0[nyz]) + static_cast<double>(U0[nz]) +
                                 static_cast<double>(U0[1]) + static_cast<double>(U0[nyz + nz + 1]) -
                                 static_cast<double>(U0[0]) - static_cast<double>(U0[nyz + nz]) -
                                 static_cast<double>(U0[nyz + 1]) - static_cast<double>(U0[nz + 1]);
                            clamp_cell(Ms, i, j, k);
                        }
                        out[u] = static_cast<OutR>(Ms);
                    }
                }
                wave_sync();
            }
            cur = wave_uniform64(nxt);
        }
    }
    flush();
}

// ------------------------------------------------------------------------------------------------
// 8. accumulation: one workgroup per 8^3 cell block
// ------------------------------------------------------------------------------------------------
constexpr int ACC_CPT = BLK * BLK * BLK / THREADS;  // 2 cells per thread, along x
constexpr int ACC_GROUP = 8;                         // items whose sums are loaded before adding
// A block with more than ACC_CHUNK items is added up in chunks of ACC_CHUNK items, one workgroup per
// chunk, each writing FP64 partial sums; k_acc_combine then adds a block's partials in chunk order.  A
// single workgroup per block left the densest blocks (5.7e4 items in a dense cutout) holding the launch.
// The split depends only on the block's own item count, so the sums stay deterministic and a batch
// still matches its single calls bit for bit.
constexpr int ACC_CHUNK = 2048;
constexpr int BLK_CELLS = BLK * BLK * BLK;

// one accumulation unit: LPT position k of its block, chunk index, partial slot (-1: writes the grid)
struct AccUnit {
    int k, chunk, slot;
};

// units per block in LPT order, and partial slots for split blocks; entry nblocks is 0 (scan totals)
__global__ void k_acc_count(int nblocks, const int* __restrict__ order, const int* __restrict__ beg,
                            const int* __restrict__ end, int* __restrict__ nunits, int* __restrict__ nsplit) {
    const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k > nblocks) return;
    int u = 0, sp = 0;
    if (k < nblocks) {
        const int b = order[k];
        const int n = end[b] - beg[b];
        u = n > ACC_CHUNK ? (n + ACC_CHUNK - 1) / ACC_CHUNK : 1;
        sp = n > ACC_CHUNK ? u : 0;
    }
    nunits[k] = u;
    nsplit[k] = sp;
}

__global__ void k_acc_fill(int nblocks, const int* __restrict__ unit_off, const int* __restrict__ split_off,
                           AccUnit* __restrict__ units) {
    const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nblocks) return;
    const int u0 = unit_off[k], nu = unit_off[k + 1] - u0;
    const bool split = split_off[k + 1] > split_off[k];
    for (int c = 0; c < nu; ++c) units[u0 + c] = AccUnit{k, c, split ? split_off[k] + c : -1};
}

// a cell's final sums into its problem's grid; mixed mode clamps the cell's total instead of every pair:
// a negative total is rounding noise of an (almost) empty cell, which then takes the mean-velocity
// fallback like one
__device__ __forceinline__ void acc_store(const Prob& P, int64_t c, double mass, double p0, double p1, double p2,
                                          int clamp_pairs) {
    if (!clamp_pairs && mass < 0.0) mass = p0 = p1 = p2 = 0.0;
    P.out[0][c] = mass;
    P.out[1][c] = p0;
    P.out[2][c] = p1;
    P.out[3][c] = p2;
}

struct AccLds {
    long long off[THREADS];  // item slab offsets
    uint32_t box[THREADS];   // block-local origin and extents, 4 bits each: ox oy oz mx my mz
    double m[THREADS], v[3][THREADS];
    uint16_t list[WAVES][THREADS];  // per wavefront: the staged items whose x-range meets its slices
};

template <typename R>
__global__ __launch_bounds__(THREADS) void k_accum(
    const AccPart* __restrict__ ap, const uint64_t* __restrict__ keys,
    const int64_t* __restrict__ slab, const int* __restrict__ blk_begin, const int* __restrict__ blk_end,
    const int* __restrict__ order, const R* __restrict__ ms, const Prob* __restrict__ pr, int nprob,
    int clamp_pairs, int nblocks, const AccUnit* __restrict__ units, const int* __restrict__ unit_off,
    double* __restrict__ partial) {
    __shared__ AccLds A;
    if (static_cast<int>(blockIdx.x) >= unit_off[nblocks]) return;  // the grid is an upper bound on the units
    const AccUnit un = units[blockIdx.x];
    const int b = order[un.k];  // global block
    const int qb = block_problem(pr, nprob, b);
    const GridGeom& g = pr[qb].g;
    const int bl = static_cast<int>(b - pr[qb].block_off);  // block within its problem's grid
    const int bz = bl % g.nbz;
    const int by = (bl / g.nbz) % g.nby;
    const int bx = bl / (g.nbz * g.nby);
    const int tid = threadIdx.x;
    const int wave = static_cast<int>(wave_uniform(threadIdx.x >> 6));
    // owned cells (block-local): x = 2 * wave + s, y = (tid % 64) / 8, z = tid % 8
    const int cy = (tid & 63) >> 3, cz = tid & 7;
    const int cx0 = wave * ACC_CPT;

    double acc[ACC_CPT][4];
#pragma unroll
    for (int s = 0; s < ACC_CPT; ++s) acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;

    const int q0 = blk_begin[b] + un.chunk * ACC_CHUNK, q1 = min(blk_end[b], q0 + ACC_CHUNK);
    const int lane = tid & 63;
    const unsigned long long lane_lt = (1ull << lane) - 1ull;
    uint16_t* const wlist = A.list[wave];
    for (int base = q0; base < q1; base += THREADS) {
        const int cnt = min(THREADS, q1 - base);
        __syncthreads();
        if (tid < cnt) {
            const uint64_t key = keys[base + tid];
            const AccPart& a = ap[key & 0xffffffffull];
            const ItemBox ib = item_box(a.bb, static_cast<uint32_t>(bl), g);
            const uint32_t ox = ib.x0 - bx * BLK, oy = ib.y0 - by * BLK, oz = ib.z0 - bz * BLK;
            A.box[tid] = ox | (oy << 4) | (oz << 8) | (static_cast<uint32_t>(ib.mx) << 12) |
                         (static_cast<uint32_t>(ib.my) << 16) | (static_cast<uint32_t>(ib.mz) << 20);
            A.off[tid] = slab[base + tid];
            A.m[tid] = a.m;
            A.v[0][tid] = a.v[0];
            A.v[1][tid] = a.v[1];
            A.v[2][tid] = a.v[2];
        }
        __syncthreads();
        // the wavefront's own list: staged items whose x-range meets its slices [cx0, cx0 + ACC_CPT), in
        // list order (the others add nothing to its cells)
        int nw = 0;
        for (int c0 = 0; c0 < cnt; c0 += 64) {
            const int t = c0 + lane;
            bool ov = false;
            if (t < cnt) {
                const uint32_t bq = A.box[t];
                const int ox = bq & 15, mx = (bq >> 12) & 15;
                ov = ox < cx0 + ACC_CPT && ox + mx > cx0;
            }
            const unsigned long long bal = __ballot(ov);
            if (ov) wlist[nw + __popcll(bal & lane_lt)] = static_cast<uint16_t>(t);
            nw += __popcll(bal);
        }
        wave_sync();
        for (int g0 = 0; g0 < nw; g0 += ACC_GROUP) {
            // issue the loads of a group of items first (independent loads in flight) ...
            // Every load is issued unconditionally from an in-bounds address (coordinates outside the
            // item are clamped into it) and masked afterwards: no branch separates a load from its use,
            // so the group's loads stay in flight together.  An item's slab is (x my + y) mz + z.
            R raw[ACC_GROUP][ACC_CPT];
            bool ok[ACC_GROUP][ACC_CPT];
#pragma unroll
            for (int u = 0; u < ACC_GROUP; ++u) {
                const int q = wave_uniform(wlist[min(g0 + u, nw - 1)]);
                // item geometry is the same for every lane: scalar registers
                const uint32_t bq = wave_uniform(A.box[q]);
                const int ox = bq & 15, oy = (bq >> 4) & 15, oz = (bq >> 8) & 15;
                const int mx = (bq >> 12) & 15, my = (bq >> 16) & 15, mz = (bq >> 20) & 15;
                const int ly = cy - oy, lz = cz - oz;
                const bool in_yz = (g0 + u < nw) && ly >= 0 && ly < my && lz >= 0 && lz < mz;
                const R* row = ms + A.off[q] + (min(max(ly, 0), my - 1) * mz + min(max(lz, 0), mz - 1));
#pragma unroll
                for (int s = 0; s < ACC_CPT; ++s) {
                    const int lx = cx0 + s - ox;
                    ok[u][s] = in_yz && lx >= 0 && lx < mx;
                    raw[u][s] = row[min(max(lx, 0), mx - 1) * my * mz];
                }
            }
            // ... then add them in list order
#pragma unroll
            for (int u = 0; u < ACC_GROUP; ++u) {
                if (g0 + u >= nw) break;
                const int q = wave_uniform(wlist[g0 + u]);
                const double mm = A.m[q], v0 = A.v[0][q], v1 = A.v[1][q], v2 = A.v[2][q];
#pragma unroll
                for (int s = 0; s < ACC_CPT; ++s) {
                    const double Ms = ok[u][s] ? static_cast<double>(raw[u][s]) : 0.0;
                    if (Ms == 0.0) continue;
                    const double t = Ms * mm;
                    acc[s][0] += t;
                    acc[s][1] += t * v0;
                    acc[s][2] += t * v1;
                    acc[s][3] += t * v2;
                }
            }
        }
    }

    if (un.slot >= 0) {  // a chunk of a split block: FP64 partial sums, [slot][field][block-local cell]
        double* const part = partial + static_cast<int64_t>(un.slot) * 4 * BLK_CELLS;
#pragma unroll
        for (int s = 0; s < ACC_CPT; ++s) {
            const int local = (cx0 + s) * BLK * BLK + (tid & 63);
#pragma unroll
            for (int f = 0; f < 4; ++f) part[f * BLK_CELLS + local] = acc[s][f];
        }
        return;
    }
    const int iy = by * BLK + cy, iz = bz * BLK + cz;
#pragma unroll
    for (int s = 0; s < ACC_CPT; ++s) {
        const int ix = bx * BLK + cx0 + s;
        if (ix >= g.nx || iy >= g.ny || iz >= g.nz) continue;
        const int64_t c = (static_cast<int64_t>(ix) * g.ny + iy) * g.nz + iz;
        acc_store(pr[qb], c, acc[s][0], acc[s][1], acc[s][2], acc[s][3], clamp_pairs);
    }
}

// the cells of every split block: its chunks' partial sums added in chunk order
__global__ __launch_bounds__(THREADS) void k_acc_combine(int nblocks, const int* __restrict__ order,
                                                         const int* __restrict__ split_off,
                                                         const double* __restrict__ partial,
                                                         const Prob* __restrict__ pr, int nprob, int clamp_pairs) {
    const int k = blockIdx.x;
    const int s0 = split_off[k], s1 = split_off[k + 1];
    if (s0 == s1) return;
    const int b = order[k];
    const int qb = block_problem(pr, nprob, b);
    const GridGeom& g = pr[qb].g;
    const int bl = static_cast<int>(b - pr[qb].block_off);
    const int bz = bl % g.nbz;
    const int by = (bl / g.nbz) % g.nby;
    const int bx = bl / (g.nbz * g.nby);
    for (int local = threadIdx.x; local < BLK_CELLS; local += THREADS) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int sl = s0; sl < s1; ++sl) {
            const double* part = partial + static_cast<int64_t>(sl) * 4 * BLK_CELLS;
#pragma unroll
            for (int f = 0; f < 4; ++f) v[f] += part[f * BLK_CELLS + local];
        }
        const int ix = bx * BLK + local / (BLK * BLK), iy = by * BLK + (local / BLK) % BLK, iz = bz * BLK + local % BLK;
        if (ix >= g.nx || iy >= g.ny || iz >= g.nz) continue;
        const int64_t c = (static_cast<int64_t>(ix) * g.ny + iy) * g.nz + iz;
        acc_store(pr[qb], c, v[0], v[1], v[2], v[3], clamp_pairs);
    }
}

// the problem of every particle (binary search over the problems' first particles)
__global__ void k_part_prob(int64_t n, const Prob* __restrict__ pr, int nprob, int* __restrict__ part_prob) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pr[mid].part_off <= i) lo = mid;
        else hi = mid - 1;
    }
    part_prob[i] = lo;
}

// velocity sums per problem: VSUM_BLOCKS workgroups per problem with fixed strides, then a fixed-order
// sum of their partials (deterministic)
__global__ __launch_bounds__(THREADS) void k_vsum_problems(const Prob* __restrict__ pr,
                                                           const double* __restrict__ vx,
                                                           const double* __restrict__ vy,
                                                           const double* __restrict__ vz, double* __restrict__ part) {
    __shared__ double sh[3][THREADS];
    const int q = blockIdx.x / VSUM_BLOCKS, j = blockIdx.x % VSUM_BLOCKS;
    const Prob& P = pr[q];
    double a = 0, b = 0, c = 0;
    for (int64_t i = static_cast<int64_t>(j) * THREADS + threadIdx.x; i < P.n_parts;
         i += static_cast<int64_t>(VSUM_BLOCKS) * THREADS) {
        a += vx[P.part_off + i];
        b += vy[P.part_off + i];
        c += vz[P.part_off + i];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[3 * static_cast<int64_t>(blockIdx.x) + k] = sh[k][0];
}

__global__ void k_vsum_problems_final(int nprob, const double* __restrict__ part, double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nprob) return;
    for (int k = 0; k < 3; ++k) {
        double t = 0;
        for (int j = 0; j < VSUM_BLOCKS; ++j) t += part[3 * (static_cast<int64_t>(q) * VSUM_BLOCKS + j) + k];
        out[3 * q + k] = t;
    }
}

}  // namespace

namespace {

// the cell kernel of a kernel path over items [first, first + count), cell sums stored as OutR
template <class P, typename OutR>
hipError_t launch_particles(const kiwi_hip_ctx* c, hipStream_t s, typename P::Ctx pctx, int64_t first, int64_t count,
                            const double* const* dp, const int* bbox, const int64_t* loffs, int loff_shift,
                            const uint64_t* keys_particle_order, const int64_t* slab, const Prob* pr,
                            const int* part_prob, void* lines, void* ms, DevStats* dst, unsigned long long* cursor,
                            int clamp) {
    using Line = LineVals<typename P::Real, P::NLN>;
    if ((clamp != 0) != std::is_same<OutR, double>::value) return hipErrorInvalidValue;  // the kernel's CLAMP
    if (count > 0)
        k_particles<P, OutR><<<resident_grid(c, k_particles<P, OutR>, c->tune.item_grab, count), THREADS, 0, s>>>(
            static_cast<uint32_t>(first), static_cast<uint32_t>(count), keys_particle_order, slab, dp[0], dp[1], dp[2],
            dp[7], bbox, loffs, loff_shift, pr, part_prob, pctx, static_cast<const Line*>(lines),
            static_cast<OutR*>(ms), dst, cursor, c->tune.item_grab, c->tune.cursors);
    return hipGetLastError();
}

// the accumulation of the per-pair cell sums (type R) into the grids
template <typename R>
hipError_t launch_accum(hipStream_t s, const AccPart* ap, const uint64_t* keys_block_order,
                        const int64_t* slab_sorted, const int* blk_begin, const int* blk_end, const int* order,
                        int nblocks, const Prob* pr, int nprob, const void* ms, int clamp, const AccUnit* units,
                        const int* unit_off, const int* split_off, int64_t max_units, double* partial) {
    if (nblocks > 0) {
        k_accum<R><<<static_cast<unsigned>(max_units), THREADS, 0, s>>>(
            ap, keys_block_order, slab_sorted, blk_begin, blk_end, order, static_cast<const R*>(ms), pr, nprob, clamp,
            nblocks, units, unit_off, partial);
        k_acc_combine<<<nblocks, THREADS, 0, s>>>(nblocks, order, split_off, partial, pr, nprob, clamp);
    }
    return hipGetLastError();
}

// the per-particle line tables (k_lines) of a kernel path, on stream s
// (sel: with the FP64 fallback on, the item counts of this precision's particles; loff_shift: 1 for the
// fallback's FP64 tables, whose offsets count FP32 entries)
template <class P>
hipError_t launch_lines(const kiwi_hip_ctx* c, hipStream_t s, typename P::Ctx pctx, int64_t n, const double* const* dp,
                        const int* bbox, const int64_t* loffs, const int64_t* sel, int loff_shift, const Prob* pr,
                        const int* part_prob, void* lines, DevStats* dst) {
    using Line = LineVals<typename P::Real, P::NLN>;
    // The line kernel runs beside the item layout, whose short kernels need free slots while it runs (a
    // grid of the full resident size holds every slot until it ends): twice the resident size, so that its
    // workgroups retire and the layout's kernels are dispatched between them.  Measured on 1M log-normal
    // particles -> 256^3 against the resident grid: FP64 53.1 -> 51.9 ms, mixed 24.0 -> 23.9 ms, Wendland
    // C2 mixed 31.1 -> 31.3 ms (4x and 8x: FP64 51.6, mixed 24.1-24.2; half the resident grid: mixed 23.5,
    // FP64 58.0 ms).
    const int64_t resident = resident_grid(c, k_lines<P>, 1, n);
    const int64_t grid = std::min<int64_t>((n + WAVES - 1) / WAVES, 2 * resident);
    if (n > 0 && P::HAS_LINES)
        k_lines<P><<<static_cast<unsigned>(grid), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], bbox, loffs, sel,
                                                                   loff_shift, pr, part_prob, pctx,
                                                                   static_cast<Line*>(lines), dst);
    return hipGetLastError();
}

// The FP64 fallback's tables sit in the FP32 table at even FP32-entry offsets: an FP64 entry is two FP32
// entries and an axis-table entry spans the same number of line entries in both precisions.
template <class P32, class P64>
constexpr bool fallback_layout_ok() {
    return sizeof(LineVals<typename P64::Real, P64::NLN>) == 2 * sizeof(LineVals<typename P32::Real, P32::NLN>) &&
           axis_units<P32>() == axis_units<P64>() && P32::NLN == P64::NLN && P32::NPL == P64::NPL;
}
static_assert(fallback_layout_ok<PolicyM4<float>, PolicyM4<double>>(), "M4 fallback table layout");
static_assert(fallback_layout_ok<PolicyGen<float>, PolicyGen<double>>(), "generic fallback table layout");

}  // namespace

// the tabulated cubic spline's lookup table, built once per context: 22.4M analytic vertex integrals, 182 MB
hipError_t kiwi_build_m4_table(kiwi_hip_ctx* c, hipStream_t s) {
    if (c->m4table_ready) return hipSuccess;
    using G = M4TableGeom;
    hipError_t e = reserve(c->m4table, sizeof(double) * static_cast<size_t>(G::size));
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(c->m4table.p, 0, sizeof(double) * static_cast<size_t>(G::size), s)) != hipSuccess) return e;
    k_table_build<<<grid1d(static_cast<int64_t>(G::nr0) * G::nR0 * G::nph), THREADS, 0, s>>>(
        static_cast<double*>(c->m4table.p));
    k_table_pad<<<grid1d(G::nR0 * G::nph), THREADS, 0, s>>>(static_cast<double*>(c->m4table.p));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    c->m4table_ready = true;
    return hipSuccess;
}

namespace {

struct HostProblem {
    const kiwi_particles* P;
    const kiwi_grid* G;
    const kiwi_cell_sums* out;
    kiwi_stats* stats;
};

// the argument rules of one problem (the reference's config and stage checks, src/config.cpp:59-206)
int check_problem(kiwi_hip_ctx* c, const kiwi_particles* P, const kiwi_grid* G, const kiwi_cell_sums* out) {
    if (P->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative particle count");
    if (P->n > 0 && (!P->x || !P->y || !P->z || !P->vx || !P->vy || !P->vz || !P->m || !P->h))
        return fail(c, KIWI_ERR_INVALID_ARG, "null particle array");
    if (!out->mass || !out->px || !out->py || !out->pz) return fail(c, KIWI_ERR_INVALID_ARG, "null output array");
    if (G->nx <= 0 || G->ny <= 0 || G->nz <= 0)
        return fail(c, KIWI_ERR_INVALID_ARG, "cartesian_grid resolution (resolved nx/ny/nz) must be > 0");
    if (!(G->x_min < G->x_max) || !(G->y_min < G->y_max) || !(G->z_min < G->z_max))
        return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");
    if (P->n > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 particles in one call");
    if (G->nx > MAX_AXIS || G->ny > MAX_AXIS || G->nz > MAX_AXIS)
        return fail(c, KIWI_ERR_LIMIT, "more than 8192 cells along a grid axis");
    return KIWI_OK;
}

// problems that can share one pipeline run: the same kernel evaluation and arithmetic
bool same_path(const kiwi_batch_problem& a, const KernelChoice& ka, const kiwi_batch_problem& b, const KernelChoice& kb,
               double fb_cells) {
    const bool fa = ka.tabulated || a.precision == KIWI_FP64 ||
                    (ka.generic && !gen_forward_fp32(a.kernel.kernel_id, fb_cells));
    const bool fb = kb.tabulated || b.precision == KIWI_FP64 ||
                    (kb.generic && !gen_forward_fp32(b.kernel.kernel_id, fb_cells));
    return ka.generic == kb.generic && ka.tabulated == kb.tabulated && ka.support_in_h == kb.support_in_h &&
           (!ka.generic || a.kernel.kernel_id == b.kernel.kernel_id) && fa == fb;
}

// one pipeline run over a group of problems with the same kernel path (a single call is a group of one)
int run_group(kiwi_hip_ctx* c, const HostProblem* hp, int nprob, const kiwi_kernel* K, const KernelChoice& kc,
              int precision, const PipeBufs& B, bool allow_graph) {
    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    // a line-table pass or fallback kernel of an earlier call that returned early (an error after its
    // launch) reads buffers this call rewrites: the main stream waits for them first
    KIWI_HIP_CHECK(c, hipStreamWaitEvent(s, c->side_done, 0));
    KIWI_HIP_CHECK(c, hipStreamWaitEvent(s, c->side2_done, 0));
    const double sup = kc.support_in_h;
    const bool generic = kc.generic;

    // problem table: grids, particle and block offsets, outputs
    std::vector<Prob> probs(static_cast<size_t>(nprob));
    int64_t n = 0, nblocks = 0, host_cells = 0;
    for (int q = 0; q < nprob; ++q) {
        const kiwi_grid* G = hp[q].G;
        GridGeom& g = probs[q].g;
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
        probs[q].part_off = n;
        probs[q].n_parts = hp[q].P->n;
        probs[q].block_off = nblocks;
        {
            const kiwi_particles* P = hp[q].P;
            const double* src[8] = {P->x, P->y, P->z, P->vx, P->vy, P->vz, P->m, P->h};
            for (int k = 0; k < 8; ++k) probs[q].src[k] = P->on_device ? src[k] : nullptr;
        }
        n += hp[q].P->n;
        nblocks += static_cast<int64_t>(g.nbx) * g.nby * g.nbz;
        if (!hp[q].out->on_device) host_cells += static_cast<int64_t>(G->nx) * G->ny * G->nz;
    }
    if (n > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 particles in one call");
    if (nblocks > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 cell blocks in one call");

    // outputs: the caller's device arrays, or staging for host arrays
    KIWI_HIP_CHECK(c, reserve((*B.outbuf), 4 * sizeof(double) * std::max<int64_t>(host_cells, 1)));
    {
        int64_t o = 0;
        for (int q = 0; q < nprob; ++q) {
            const kiwi_cell_sums* out = hp[q].out;
            const int64_t nc = static_cast<int64_t>(hp[q].G->nx) * hp[q].G->ny * hp[q].G->nz;
            if (out->on_device) {
                probs[q].out[0] = out->mass;
                probs[q].out[1] = out->px;
                probs[q].out[2] = out->py;
                probs[q].out[3] = out->pz;
            } else {
                for (int k = 0; k < 4; ++k) probs[q].out[k] = static_cast<double*>((*B.outbuf).p) + 4 * o + k * nc;
                o += nc;
            }
        }
    }
    KIWI_HIP_CHECK(c, reserve((*B.probs), sizeof(Prob) * nprob));
    Prob* dpr = static_cast<Prob*>((*B.probs).p);
    KIWI_HIP_CHECK(c, hipMemcpyAsync(dpr, probs.data(), sizeof(Prob) * nprob, hipMemcpyHostToDevice, s));
    // particle arrays on the device: the caller's own for a single device-resident problem, else the
    // problems' columns concatenated (device columns by k_gather_parts, host columns copied)
    const double* dp[8];
    if (nprob == 1 && hp[0].P->on_device) {
        for (int k = 0; k < 8; ++k) dp[k] = probs[0].src[k];
    } else {
        KIWI_HIP_CHECK(c, reserve((*B.parts), 8 * sizeof(double) * std::max<int64_t>(n, 1)));
        for (int k = 0; k < 8; ++k) dp[k] = static_cast<double*>((*B.parts).p) + k * n;
        int64_t dev_max = 0;
        for (int q = 0; q < nprob; ++q) {
            const kiwi_particles* P = hp[q].P;
            if (P->n == 0) continue;
            if (P->on_device) {
                dev_max = std::max<int64_t>(dev_max, P->n);
                continue;
            }
            const double* src[8] = {P->x, P->y, P->z, P->vx, P->vy, P->vz, P->m, P->h};
            for (int k = 0; k < 8; ++k)
                KIWI_HIP_CHECK(c, hipMemcpyAsync(const_cast<double*>(dp[k]) + probs[q].part_off, src[k],
                                                 P->n * sizeof(double), hipMemcpyHostToDevice, s));
        }
        if (dev_max > 0) {
            const unsigned gx = static_cast<unsigned>(std::min<int64_t>((dev_max + THREADS - 1) / THREADS, 1024));
            k_gather_parts<<<dim3(gx, static_cast<unsigned>(nprob)), THREADS, 0, s>>>(
                dpr, n, static_cast<double*>((*B.parts).p));
            KIWI_HIP_CHECK(c, hipGetLastError());
        }
    }
    KIWI_HIP_CHECK(c, reserve((*B.part_prob), sizeof(int) * std::max<int64_t>(n, 1)));
    int* part_prob = static_cast<int*>((*B.part_prob).p);

    // per-particle sizes (items, lines) and their exclusive scans, (n + 1) each
    const int64_t n1 = n + 1;
    // two statistics sets: everything, and the FP64 fallback's cell kernel (merged into the first at the end)
    const size_t stats_bytes = (sizeof(DevStats) * 2 * nprob + 255) / 256 * 256;
    KIWI_HIP_CHECK(c, reserve((*B.stats), stats_bytes + CURSOR_BYTES));
    DevStats* dst = static_cast<DevStats*>((*B.stats).p);
    DevStats* dst_fb = dst + nprob;
    unsigned long long* cursor = reinterpret_cast<unsigned long long*>(static_cast<char*>((*B.stats).p) + stats_bytes);
    KIWI_HIP_CHECK(c, reserve((*B.bbox), 6 * sizeof(int) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.acc_parts), sizeof(AccPart) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.sizes), 3 * sizeof(int64_t) * n1));
    KIWI_HIP_CHECK(c, reserve((*B.offsets), 3 * sizeof(int64_t) * n1));
    KIWI_HIP_CHECK(c, reserve((*B.ranges), 2 * sizeof(int) * std::max<int64_t>(nblocks, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.work), sizeof(unsigned long long) * std::max<int64_t>(nblocks, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.work_sorted), sizeof(unsigned long long) * std::max<int64_t>(nblocks, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.order_in), sizeof(int) * std::max<int64_t>(nblocks, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.order), sizeof(int) * std::max<int64_t>(nblocks, 1)));
    KIWI_HIP_CHECK(c, reserve((*B.vsum), sizeof(double) * 3 * nprob * (VSUM_BLOCKS + 1)));
    int* bbox = static_cast<int*>((*B.bbox).p);
    AccPart* ap = static_cast<AccPart*>((*B.acc_parts).p);
    int64_t* sz = static_cast<int64_t*>((*B.sizes).p);
    int64_t* of = static_cast<int64_t*>((*B.offsets).p);
    int64_t *nblk = sz, *nlin = sz + n1, *nblk_fb = sz + 2 * n1;
    int64_t *offs = of, *loffs = of + n1, *offs_fb = of + 2 * n1;
    int* blk_begin = static_cast<int*>((*B.ranges).p);
    int* blk_end = blk_begin + nblocks;
    unsigned long long* work = static_cast<unsigned long long*>((*B.work).p);
    unsigned long long* work_sorted = static_cast<unsigned long long*>((*B.work_sorted).p);
    int* order_in = static_cast<int*>((*B.order_in).p);
    int* order = static_cast<int*>((*B.order).p);
    double* vsums = static_cast<double*>((*B.vsum).p);

    {  // counters and work cursors start at zero (a fill kernel: no memset node in a captured stream)
        const int64_t words = static_cast<int64_t>((stats_bytes + CURSOR_BYTES) / sizeof(unsigned long long));
        k_fill_u64<<<grid1d(words), THREADS, 0, s>>>(reinterpret_cast<unsigned long long*>(dst), words);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    if (n > 0) k_part_prob<<<grid1d(n), THREADS, 0, s>>>(n, dpr, nprob, part_prob);
    // blocks wholly outside the support sphere hold only exact zeros and are left out, except on the
    // tabulated path, whose table values do not vanish there
    const int skip_outside = kc.tabulated ? 0 : 1;
    // the arithmetic of the path: mixed precision applies to the cubic path and to the generic kernels
    // (those beyond gen_mixed_ok only with the FP64 fallback on: gen_forward_fp32); the tabulated path
    // always computes in double
    const bool tabulated = kc.tabulated;
    const bool fp64 = tabulated || precision == KIWI_FP64 || (generic && !gen_forward_fp32(K->kernel_id, c->fb_cells));
    const int ax_units = tabulated ? 0
                         : generic ? (fp64 ? axis_units<PolicyGen<double>>() : axis_units<PolicyGen<float>>())
                                   : (fp64 ? axis_units<PolicyM4<double>>() : axis_units<PolicyM4<float>>());
    // mixed precision's FP64 fallback for particles whose support spans many cells (fp64_fallback): their
    // items follow the FP32 items and run through the FP64 cell kernel, which stores float cell sums; the
    // higher-order generic kernels' threshold is scaled down (gen_fallback_scale)
    const double fb_cells = fp64 ? 0.0 : c->fb_cells * (generic ? gen_fallback_scale(K->kernel_id) : 1.0);
    const bool split = fb_cells > 0.0;
    const int nscan = split ? 3 : 2;
    k_prep<<<grid1d(n1), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], dp[6], dp[3], dp[4], dp[5], sup, skip_outside,
                                          ax_units, fb_cells, dpr, part_prob, bbox, ap, nblk, nlin, nblk_fb, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    {
        size_t tmp_bytes = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, nblk, offs, static_cast<int>(n1), s));
        KIWI_HIP_CHECK(c, reserve((*B.cub_tmp), tmp_bytes));
        for (int k = 0; k < nscan; ++k) {
            size_t tb = (*B.cub_tmp).cap;
            KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum((*B.cub_tmp).p, tb, sz + k * n1, of + k * n1,
                                                               static_cast<int>(n1), s));
        }
    }
    int64_t totals[3] = {0, 0, 0};
    std::vector<DevStats> hs(static_cast<size_t>(nprob));
    for (int k = 0; k < nscan; ++k)
        KIWI_HIP_CHECK(c, hipMemcpyAsync(&totals[k], of + k * n1 + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipMemcpyAsync(hs.data(), dst, sizeof(DevStats) * nprob, hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    // items: the FP32 ones [0, n32), then the FP64 fallback's [n32, npairs)
    const int64_t n32 = totals[0], npairs = totals[0] + totals[2], nlines_total = totals[1];
    unsigned long long pairs_bbox = 0;
    for (int q = 0; q < nprob; ++q) pairs_bbox += hs[q].pairs_bbox;
    if (npairs > 0x7fffffffLL)
        return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 (particle, cell-block) items in one call");

    const size_t rsz = fp64 ? sizeof(double) : sizeof(float);
    const size_t line_entry = tabulated ? 0
                              : generic ? (fp64 ? sizeof(LineVals<double, GEN_NREG>) : sizeof(LineVals<float, GEN_NREG>))
                                        : (fp64 ? sizeof(LineVals<double, 2>) : sizeof(LineVals<float, 2>));
    const int64_t np1 = npairs + 1;
    KIWI_HIP_CHECK(c, reserve((*B.lines), line_entry * static_cast<size_t>(std::max<int64_t>(nlines_total, 1))));
    KIWI_HIP_CHECK(c, reserve((*B.ms), rsz * static_cast<size_t>(std::max<unsigned long long>(pairs_bbox, 1))));
    KIWI_HIP_CHECK(c, reserve((*B.keys_a), sizeof(uint64_t) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.keys_b), sizeof(uint64_t) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.pos_a), sizeof(int) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.pos_b), sizeof(int) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.cells), sizeof(int64_t) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.slab_sorted), sizeof(int64_t) * np1));
    KIWI_HIP_CHECK(c, reserve((*B.slab), sizeof(int64_t) * np1));
    uint64_t* keys_a = static_cast<uint64_t*>((*B.keys_a).p);  // particle order
    uint64_t* keys_b = static_cast<uint64_t*>((*B.keys_b).p);  // block order
    int* pos_a = static_cast<int*>((*B.pos_a).p);
    int* pos_b = static_cast<int*>((*B.pos_b).p);  // particle-order position of each block-order item
    int64_t* cells = static_cast<int64_t*>((*B.cells).p);
    int64_t* slab_sorted = static_cast<int64_t*>((*B.slab_sorted).p);
    int64_t* slab = static_cast<int64_t*>((*B.slab).p);
    const bool side_lines = n > 0 && !tabulated;
    int end_bit = 33;  // at least one key bit above the particle index
    while ((1LL << (end_bit - 32)) < nblocks) ++end_bit;
    // accumulation units (blocks split into chunks of ACC_CHUNK items) in LPT order; the unit count is
    // bounded by nblocks + 2 npairs / ACC_CHUNK and the split units by 2 npairs / ACC_CHUNK, so the grid
    // and the partial sums are sized without reading the counts back
    const int64_t max_split = 2 * (npairs / ACC_CHUNK) + 1;
    const int64_t max_units = nblocks + max_split;
    if (max_units > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "too many cell blocks in one call");
    KIWI_HIP_CHECK(c, reserve((*B.acc_units), sizeof(AccUnit) * max_units + 4 * sizeof(int) * (nblocks + 1) + 256));
    KIWI_HIP_CHECK(c, reserve((*B.acc_partial), sizeof(double) * 4 * BLK_CELLS * max_split));
    AccUnit* units = static_cast<AccUnit*>((*B.acc_units).p);
    int* nunits = reinterpret_cast<int*>(units + max_units);
    int* nsplit = nunits + (nblocks + 1);
    int* unit_off = nsplit + (nblocks + 1);
    int* split_off = unit_off + (nblocks + 1);
    double* partial = static_cast<double*>((*B.acc_partial).p);
    if (tabulated) KIWI_HIP_CHECK(c, kiwi_build_m4_table(c, s));  // once per context, outside any capture

    // Graph mode (kiwi_hip_set_graphs): everything from here to the statistics' copies is captured into a
    // HIP graph once per signature (sizes, launch shape, every buffer address) and replayed afterwards.
    // It needs device outputs (host copies would bake the caller's host addresses into the graph) and
    // sorts with the library's own radix sort, which records no memset nodes.
    bool dev_out = true;
    for (int q = 0; q < nprob; ++q) dev_out = dev_out && hp[q].out->on_device;
    const bool graph = allow_graph && c->graphs && dev_out && !tabulated && npairs > 0;
    // scan and sort storage, reserved before any launch of the rest (a capture must not allocate)
    int wbits = 1;  // a block's work (its items' cells) is at most pairs_bbox
    while (wbits < 64 && (pairs_bbox >> wbits) != 0) ++wbits;
    {
        size_t need = 0, tb = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cells, slab_sorted, static_cast<int>(np1), s));
        need = std::max(need, tb);
        KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, nunits, unit_off, static_cast<int>(nblocks + 1),
                                                           s));
        need = std::max(need, tb);
        if (graph) {
            const int64_t m = std::max<int64_t>(npairs, nblocks);
            KIWI_HIP_CHECK(c, rs_scan_bytes(m, &tb));
            need = std::max(need, tb);
            const size_t nh = sizeof(int64_t) * static_cast<size_t>(RS_BINS * rs_chunks(m) + 1);
            KIWI_HIP_CHECK(c, reserve(c->g_ktmp, sizeof(uint64_t) * static_cast<size_t>(m + 1)));
            KIWI_HIP_CHECK(c, reserve(c->g_vtmp, sizeof(int) * static_cast<size_t>(m + 1)));
            KIWI_HIP_CHECK(c, reserve(c->g_hist, nh));
            KIWI_HIP_CHECK(c, reserve(c->g_hoff, nh));
            const size_t hb = sizeof(DevStats) * 2 * nprob + sizeof(double) * 3 * nprob;
            if (hb > c->g_host_cap) {
                if (c->g_host) KIWI_HIP_CHECK(c, hipHostFree(c->g_host));
                c->g_host = nullptr;
                c->g_host_cap = 0;
                KIWI_HIP_CHECK(c, hipHostMalloc(&c->g_host, hb, hipHostMallocDefault));
                c->g_host_cap = hb;
            }
        } else {
            if (npairs > 0) {
                KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys_a, keys_b, pos_a, pos_b,
                                                                      static_cast<int>(npairs), 32, end_bit, s));
                need = std::max(need, tb);
            }
            if (nblocks > 0) {
                KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, work, work_sorted, order_in,
                                                                                order, static_cast<int>(nblocks), 0,
                                                                                64, s));
                need = std::max(need, tb);
            }
        }
        KIWI_HIP_CHECK(c, reserve((*B.cub_tmp), need));
    }
    void* cub = (*B.cub_tmp).p;
    const size_t cub_cap = (*B.cub_tmp).cap;
    const int clamp = fp64 ? 1 : 0;
    const int nb = static_cast<int>(nblocks);
    unsigned long long* cursor_fb = cursor + KIWI_MAX_CURSORS * CURSOR_STRIDE;  // the fallback's work cursors
    void* lt = (*B.lines).p;
    void* msb = (*B.ms).p;
    const GenTab* tabs = static_cast<const GenTab*>(c->tables.p) + K->kernel_id;
    const int kid = K->kernel_id;
    // the FP64 fallback's cell kernel runs on the third stream beside the FP32 one (disjoint items,
    // slab ranges, counters and work cursors); the accumulation waits for both
    const bool fb_on_side = split && npairs > n32 && !tabulated;
    std::vector<double> vs(static_cast<size_t>(3 * nprob));
    std::vector<DevStats> hs_fb(static_cast<size_t>(nprob));
    DevStats* g_stats = static_cast<DevStats*>(c->g_host);
    double* g_vs = reinterpret_cast<double*>(g_stats + (graph ? 2 * nprob : 0));

    // The rest of the pipeline.  Captured (graph mode): no timing events, the library's radix sort, and
    // the statistics copied to pinned host memory.
    auto enqueue = [&](bool capturing) -> int {
        // The line tables need only the boxes and line offsets: they are computed on the second stream
        // while the items are sorted and laid out on this one.  Mixed precision's FP64 fallback computes its
        // own line tables on the third stream, where its cell kernel follows them as soon as the items are
        // laid out: it is the longest of the cell kernels, and starts before the FP32 one.
        if (side_lines) {
            KIWI_HIP_CHECK(c, hipEventRecord(c->fork, s));
            KIWI_HIP_CHECK(c, hipStreamWaitEvent(c->side, c->fork, 0));
            if (fb_on_side) {
                KIWI_HIP_CHECK(c, hipStreamWaitEvent(c->side2, c->fork, 0));
                if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[4], c->side2));
                if (generic)
                    KIWI_HIP_CHECK(c, with_fallback_gen_policy(kid, [&](auto pol) {
                                       using P = decltype(pol);
                                       return launch_lines<P>(c, c->side2, typename P::Ctx{sup, tabs}, n, dp, bbox,
                                                              loffs, nblk_fb, 1, dpr, part_prob, lt, dst);
                                   }));
                else
                    KIWI_HIP_CHECK(c, launch_lines<PolicyM4<double>>(c, c->side2, {sup}, n, dp, bbox, loffs, nblk_fb, 1,
                                                                     dpr, part_prob, lt, dst));
                if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[5], c->side2));
            }
            if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[0], c->side));
            const int64_t* sel32 = split ? nblk : nullptr;
            if (generic && fp64) {
                KIWI_HIP_CHECK(c, with_gen_policy<double>(kid, [&](auto pol) {
                                   using P = decltype(pol);
                                   return launch_lines<P>(c, c->side, typename P::Ctx{sup, tabs}, n, dp, bbox, loffs,
                                                          nullptr, 0, dpr, part_prob, lt, dst);
                               }));
            } else if (generic) {
                KIWI_HIP_CHECK(c, with_gen_policy<float>(kid, [&](auto pol) {
                                   using P = decltype(pol);
                                   return launch_lines<P>(c, c->side, typename P::Ctx{sup, tabs}, n, dp, bbox, loffs,
                                                          sel32, 0, dpr, part_prob, lt, dst);
                               }));
            } else if (fp64) {
                KIWI_HIP_CHECK(c, launch_lines<PolicyM4<double>>(c, c->side, {sup}, n, dp, bbox, loffs, nullptr, 0, dpr,
                                                                 part_prob, lt, dst));
            } else {
                KIWI_HIP_CHECK(c, launch_lines<PolicyM4<float>>(c, c->side, {sup}, n, dp, bbox, loffs, sel32, 0, dpr,
                                                                part_prob, lt, dst));
            }
            if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[1], c->side));
            KIWI_HIP_CHECK(c, hipEventRecord(c->side_done, c->side));
        }
        if (n > 0) {
            k_emit<<<grid1d(n), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], sup, skip_outside, bbox, offs,
                                                 split ? offs_fb : nullptr, dpr, part_prob, keys_a);
            KIWI_HIP_CHECK(c, hipGetLastError());
        }
        if (npairs > 0) {
            k_iota<<<grid1d(npairs), THREADS, 0, s>>>(pos_a, npairs);
            if (capturing) {
                KIWI_HIP_CHECK(c, rs_sort_pairs(keys_a, pos_a, keys_b, pos_b, static_cast<uint64_t*>(c->g_ktmp.p),
                                                static_cast<int*>(c->g_vtmp.p), npairs, 32, end_bit, false,
                                                static_cast<int64_t*>(c->g_hist.p), static_cast<int64_t*>(c->g_hoff.p),
                                                cub, cub_cap, s));
            } else {
                size_t tb = cub_cap;
                KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairs(cub, tb, keys_a, keys_b, pos_a, pos_b,
                                                                      static_cast<int>(npairs), 32, end_bit, s));
            }
        }
        k_item_cells<<<grid1d(np1), THREADS, 0, s>>>(npairs, keys_b, bbox, dpr, part_prob, cells);
        {
            size_t tb = cub_cap;
            KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(cub, tb, cells, slab_sorted, static_cast<int>(np1), s));
        }
        if (nblocks > 0) {
            k_fill_int<<<grid1d(2 * nblocks), THREADS, 0, s>>>(blk_begin, 2 * nblocks, 0);
            k_fill_u64<<<grid1d(nblocks), THREADS, 0, s>>>(work, nblocks);
            k_iota<<<grid1d(nblocks), THREADS, 0, s>>>(order_in, nblocks);
        }
        KIWI_HIP_CHECK(c, hipGetLastError());
        if (npairs > 0) {
            k_scatter<<<grid1d(npairs), THREADS, 0, s>>>(npairs, pos_b, slab_sorted, slab);
            k_ranges<<<grid1d(npairs), THREADS, 0, s>>>(npairs, keys_b, blk_begin, blk_end);
            k_block_work<<<grid1d(npairs), THREADS, 0, s>>>(npairs, keys_b, cells, work);
            KIWI_HIP_CHECK(c, hipGetLastError());
        }
        if (nblocks > 0) {
            if (capturing) {
                KIWI_HIP_CHECK(c, rs_sort_pairs(reinterpret_cast<const uint64_t*>(work), order_in,
                                                reinterpret_cast<uint64_t*>(work_sorted), order,
                                                static_cast<uint64_t*>(c->g_ktmp.p),
                                                static_cast<int*>(c->g_vtmp.p), nblocks, 0, wbits, true,
                                                static_cast<int64_t*>(c->g_hist.p), static_cast<int64_t*>(c->g_hoff.p),
                                                cub, cub_cap, s));
            } else {
                size_t tb = cub_cap;
                KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(cub, tb, work, work_sorted, order_in,
                                                                                order, static_cast<int>(nblocks), 0,
                                                                                64, s));
            }
            k_acc_count<<<grid1d(nblocks + 1), THREADS, 0, s>>>(static_cast<int>(nblocks), order, blk_begin, blk_end,
                                                                 nunits, nsplit);
            for (int k = 0; k < 2; ++k) {
                size_t tb = cub_cap;
                KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(cub, tb, k ? nsplit : nunits, k ? split_off : unit_off,
                                                                   static_cast<int>(nblocks + 1), s));
            }
            k_acc_fill<<<grid1d(nblocks), THREADS, 0, s>>>(static_cast<int>(nblocks), unit_off, split_off, units);
            KIWI_HIP_CHECK(c, hipGetLastError());
        }
        if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
        if (fb_on_side) {
            // the fallback's cell kernel on the third stream, after its line tables there and the item layout
            // here (both cell kernels are VALU-issue bound: sharing the CUs does not shorten their sum, and a
            // smaller fallback grid only lengthens its tail; measured 16.3 / 17.0 / 19.7 ms for 1, 1/2, 1/4 of
            // its grid)
            KIWI_HIP_CHECK(c, hipEventRecord(c->fork2, s));
            KIWI_HIP_CHECK(c, hipStreamWaitEvent(c->side2, c->fork2, 0));
            if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[2], c->side2));
            if (generic) {
                KIWI_HIP_CHECK(c, with_fallback_gen_policy(kid, [&](auto pol) {
                                   using P = decltype(pol);
                                   return launch_particles<P, float>(c, c->side2, typename P::Ctx{sup, tabs}, n32,
                                                                     npairs - n32, dp, bbox, loffs, 1, keys_a, slab,
                                                                     dpr, part_prob, lt, msb, dst_fb, cursor_fb, clamp);
                               }));
            } else {
                KIWI_HIP_CHECK(c, (launch_particles<PolicyM4<double>, float>(c, c->side2, {sup}, n32, npairs - n32, dp,
                                                                             bbox, loffs, 1, keys_a, slab, dpr,
                                                                             part_prob, lt, msb, dst_fb, cursor_fb,
                                                                             clamp)));
            }
            if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev_side[3], c->side2));
            KIWI_HIP_CHECK(c, hipEventRecord(c->side2_done, c->side2));
        }
        if (side_lines) KIWI_HIP_CHECK(c, hipStreamWaitEvent(s, c->side_done, 0));  // the line tables are in
        if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));
        if (tabulated) {
            const PolicyTab::Ctx pc{sup, static_cast<const double*>(c->m4table.p)};
            KIWI_HIP_CHECK(c, (launch_particles<PolicyTab, double>(c, s, pc, 0, npairs, dp, bbox, loffs, 0, keys_a, slab,
                                                                   dpr, part_prob, lt, msb, dst, cursor, clamp)));
        } else if (generic) {
            if (fp64) {
                KIWI_HIP_CHECK(c, with_gen_policy<double>(kid, [&](auto pol) {
                                   using P = decltype(pol);
                                   return launch_particles<P, double>(c, s, typename P::Ctx{sup, tabs}, 0, npairs, dp,
                                                                      bbox, loffs, 0, keys_a, slab, dpr, part_prob, lt,
                                                                      msb, dst, cursor, clamp);
                               }));
            } else {
                KIWI_HIP_CHECK(c, with_gen_policy<float>(kid, [&](auto pol) {
                                   using P = decltype(pol);
                                   return launch_particles<P, float>(c, s, typename P::Ctx{sup, tabs}, 0, n32, dp, bbox,
                                                                     loffs, 0, keys_a, slab, dpr, part_prob, lt, msb,
                                                                     dst, cursor, clamp);
                               }));
            }
        } else if (fp64) {
            KIWI_HIP_CHECK(c, (launch_particles<PolicyM4<double>, double>(c, s, {sup}, 0, npairs, dp, bbox, loffs, 0,
                                                                          keys_a, slab, dpr, part_prob, lt, msb, dst,
                                                                          cursor, clamp)));
        } else {
            KIWI_HIP_CHECK(c, (launch_particles<PolicyM4<float>, float>(c, s, {sup}, 0, n32, dp, bbox, loffs, 0, keys_a,
                                                                        slab, dpr, part_prob, lt, msb, dst, cursor,
                                                                        clamp)));
        }
        if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev[5], s));  // the main cell kernel's end
        if (fb_on_side) KIWI_HIP_CHECK(c, hipStreamWaitEvent(s, c->side2_done, 0));
        if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev[3], s));
        if (fp64)
            KIWI_HIP_CHECK(c, launch_accum<double>(s, ap, keys_b, slab_sorted, blk_begin, blk_end, order, nb, dpr, nprob,
                                                   msb, clamp, units, unit_off, split_off, max_units, partial));
        else
            KIWI_HIP_CHECK(c, launch_accum<float>(s, ap, keys_b, slab_sorted, blk_begin, blk_end, order, nb, dpr, nprob,
                                                  msb, clamp, units, unit_off, split_off, max_units, partial));
        if (!capturing) KIWI_HIP_CHECK(c, hipEventRecord(c->ev[4], s));
        k_vsum_problems<<<nprob * VSUM_BLOCKS, THREADS, 0, s>>>(dpr, dp[3], dp[4], dp[5], vsums + 3 * nprob);
        k_vsum_problems_final<<<(nprob + 63) / 64, 64, 0, s>>>(nprob, vsums + 3 * nprob, vsums);
        KIWI_HIP_CHECK(c, hipGetLastError());
        if (capturing) {
            KIWI_HIP_CHECK(c, hipMemcpyAsync(g_stats, dst, sizeof(DevStats) * 2 * nprob, hipMemcpyDeviceToHost, s));
            KIWI_HIP_CHECK(c, hipMemcpyAsync(g_vs, vsums, sizeof(double) * 3 * nprob, hipMemcpyDeviceToHost, s));
            return KIWI_OK;
        }
        for (int q = 0; q < nprob; ++q) {
            const kiwi_cell_sums* out = hp[q].out;
            if (out->on_device) continue;
            const int64_t nc = static_cast<int64_t>(hp[q].G->nx) * hp[q].G->ny * hp[q].G->nz;
            double* ho[4] = {out->mass, out->px, out->py, out->pz};
            for (int k = 0; k < 4; ++k)
                KIWI_HIP_CHECK(c, hipMemcpyAsync(ho[k], probs[q].out[k], nc * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        KIWI_HIP_CHECK(c, hipMemcpyAsync(hs.data(), dst, sizeof(DevStats) * nprob, hipMemcpyDeviceToHost, s));
        KIWI_HIP_CHECK(c, hipMemcpyAsync(hs_fb.data(), dst_fb, sizeof(DevStats) * nprob, hipMemcpyDeviceToHost, s));
        KIWI_HIP_CHECK(c, hipMemcpyAsync(vs.data(), vsums, sizeof(double) * 3 * nprob, hipMemcpyDeviceToHost, s));
        return KIWI_OK;
    };

    if (!graph) {
        if (const int rc = enqueue(false)) return rc;
    } else {
        // what the captured launches hold: every size and switch they were enqueued with, the launch shape,
        // and the address of every buffer (a reallocated buffer changes the signature)
        std::vector<uint64_t> sig = {static_cast<uint64_t>(nprob), static_cast<uint64_t>(n),
                                     static_cast<uint64_t>(nblocks), static_cast<uint64_t>(npairs),
                                     static_cast<uint64_t>(n32), static_cast<uint64_t>(nlines_total),
                                     static_cast<uint64_t>(pairs_bbox), static_cast<uint64_t>(end_bit),
                                     static_cast<uint64_t>(wbits), static_cast<uint64_t>(generic),
                                     static_cast<uint64_t>(fp64), static_cast<uint64_t>(split),
                                     static_cast<uint64_t>(skip_outside), static_cast<uint64_t>(K->kernel_id),
                                     __builtin_bit_cast(uint64_t, sup), static_cast<uint64_t>(c->tune.item_grab),
                                     static_cast<uint64_t>(c->tune.cursors), static_cast<uint64_t>(cub_cap),
                                     reinterpret_cast<uint64_t>(cub), reinterpret_cast<uint64_t>(c->g_host)};
        for (int k = 0; k < 8; ++k) sig.push_back(reinterpret_cast<uint64_t>(dp[k]));
#define KIWI_PB_SIG(nm) sig.push_back(reinterpret_cast<uint64_t>((*B.nm).p));
        KIWI_PIPE_BUFS(KIWI_PB_SIG)
#undef KIWI_PB_SIG
        for (const DevBuf* b : {&c->g_ktmp, &c->g_vtmp, &c->g_hist, &c->g_hoff, &c->tables})
            sig.push_back(reinterpret_cast<uint64_t>(b->p));
        auto& cache = c->graph_cache;
        size_t hit = cache.size();
        for (size_t i = 0; i < cache.size(); ++i)
            if (cache[i].sig == sig) hit = i;
        KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));  // the sizing pass ends here
        if (hit == cache.size()) {
            KIWI_HIP_CHECK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(true);
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &gr);
            if (rc != KIWI_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (ec != hipSuccess) return hip_fail(c, ec, "hipStreamEndCapture");
            // a memset node would come from a library call: such graphs have given wrong results on replays
            // after the first, so the capture is refused
            size_t nnodes = 0;
            (void)hipGraphGetNodes(gr, nullptr, &nnodes);
            std::vector<hipGraphNode_t> nodes(nnodes);
            bool has_memset = false;
            if (nnodes > 0 && hipGraphGetNodes(gr, nodes.data(), &nnodes) == hipSuccess) {
                for (hipGraphNode_t nd : nodes) {
                    hipGraphNodeType ty;
                    if (hipGraphNodeGetType(nd, &ty) == hipSuccess && ty == hipGraphNodeTypeMemset) has_memset = true;
                }
            }
            if (has_memset) {
                (void)hipGraphDestroy(gr);
                return fail(c, KIWI_ERR_DEVICE, "the captured pipeline holds a memset node; graph mode refused");
            }
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) return hip_fail(c, ei, "hipGraphInstantiate");
            if (cache.size() >= KIWI_GRAPH_CACHE) {
                (void)hipGraphExecDestroy(cache.front().exec);
                cache.erase(cache.begin());
            }
            cache.push_back(GraphEntry{std::move(sig), ex});
        } else if (hit + 1 != cache.size()) {  // most recently used last
            GraphEntry e = std::move(cache[hit]);
            cache.erase(cache.begin() + static_cast<std::ptrdiff_t>(hit));
            cache.push_back(std::move(e));
        }
        KIWI_HIP_CHECK(c, hipGraphLaunch(cache.back().exec, s));
        KIWI_HIP_CHECK(c, hipEventRecord(c->ev[4], s));
    }
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (graph) {
        std::memcpy(hs.data(), g_stats, sizeof(DevStats) * nprob);
        std::memcpy(hs_fb.data(), g_stats + nprob, sizeof(DevStats) * nprob);
        std::memcpy(vs.data(), g_vs, sizeof(double) * 3 * nprob);
    }
    // device times; a replayed graph is timed as a whole (device_ms_gather), after the sizing pass
    float ms_prep = 0, ms_lines = 0, ms_cells = 0, ms_accum = 0;
    float ms_fb = 0, ms_cells_kernel = 0, ms_lines_kernel = 0;
    if (graph) {
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_cells, c->ev[1], c->ev[4]);
    } else {
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_lines, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&ms_cells, c->ev[2], c->ev[3]);
        (void)hipEventElapsedTime(&ms_accum, c->ev[3], c->ev[4]);
        (void)hipEventElapsedTime(&ms_cells_kernel, c->ev[2], c->ev[5]);
        if (fb_on_side) (void)hipEventElapsedTime(&ms_fb, c->ev_side[2], c->ev_side[3]);
        if (side_lines) {
            // from the line tables' start to the end of the last of them (the fallback's on the third stream)
            (void)hipEventElapsedTime(&ms_lines_kernel, c->ev_side[0], c->ev_side[1]);
            float ms_fb_lines = 0;
            if (fb_on_side && hipEventElapsedTime(&ms_fb_lines, c->ev_side[0], c->ev_side[5]) == hipSuccess)
                ms_lines_kernel = std::max(ms_lines_kernel, ms_fb_lines);
        }
    }
    for (int q = 0; q < nprob; ++q) {
        kiwi_stats* stats = hp[q].stats;
        if (!stats) continue;
        // the fallback's cell-kernel counters join the totals (its warning records stay apart: mixed
        // precision does not clamp pairs, so it records none)
        DevStats d = hs[q];
        const DevStats& f = hs_fb[q];
        d.pairs_sphere += f.pairs_sphere;
        d.n_negative += f.n_negative;
        d.n_negative_warn += f.n_negative_warn;
        d.planes += f.planes;
        for (int r = 0; r < 3; ++r) d.pts[r] += f.pts[r];
        std::memset(stats, 0, sizeof(*stats));
        stats->skipped_outside = static_cast<int64_t>(d.skipped_outside);
        stats->n_negative = static_cast<int64_t>(d.n_negative);
        stats->n_negative_warn = static_cast<int64_t>(d.n_negative_warn);
        stats->pairs_bbox = static_cast<int64_t>(d.pairs_bbox);
        stats->pairs_sphere = static_cast<int64_t>(d.pairs_sphere);
        stats->particle_blocks = nprob == 1 ? npairs : -1;
        const int64_t nq = hp[q].P->n;
        for (int k = 0; k < 3; ++k) stats->v_mean[k] = nq > 0 ? vs[3 * q + k] / static_cast<double>(nq) : 0.0;
        // device times cover the whole group's run
        stats->device_ms_prep = ms_prep;
        stats->device_ms_lines = ms_lines;
        stats->device_ms_gather = ms_cells;
        stats->device_ms_accum = ms_accum;
        stats->device_ms_lines_kernel = ms_lines_kernel;
        for (int r = 0; r < 3; ++r) {
            stats->lattice_points[r] = static_cast<int64_t>(d.pts[r]);
            stats->line_evals[r] = static_cast<int64_t>(d.lines[r]);
        }
        stats->plane_evals = static_cast<int64_t>(d.planes);
        stats->fp64_fallback_blocks = nprob == 1 ? npairs - n32 : -1;
        stats->fp64_fallback_pairs = static_cast<int64_t>(f.pairs_sphere);
        for (int r = 0; r < 3; ++r) stats->fp64_fallback_points[r] = static_cast<int64_t>(f.pts[r]);
        stats->device_ms_gather_fallback = ms_fb;
        stats->device_ms_gather_kernel = ms_cells_kernel;
        const int nw = static_cast<int>(std::min<unsigned long long>(d.n_warn, KIWI_MAX_WARN_RECORDS));
        stats->n_warn_records = nw;
        for (int k = 0; k < nw; ++k) stats->warn[k] = d.warn[k];
    }
    return KIWI_OK;
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
    if (hipSetDevice(device_ordinal) != hipSuccess ||
        hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device_ordinal) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KIWI_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side2_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork2, hipEventDisableTiming) != hipSuccess) {
        kiwi_hip_destroy(c);
        return KIWI_ERR_DEVICE;
    }
    for (auto& e : c->ev_side) {
        if (hipEventCreate(&e) != hipSuccess) {
            kiwi_hip_destroy(c);
            return KIWI_ERR_DEVICE;
        }
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            kiwi_hip_destroy(c);
            return KIWI_ERR_DEVICE;
        }
    }
    // the allocation-free scratch's address on this device (the device is current here)
    (void)scratch_address(device_ordinal);
    // the generic kernels' coefficient tables, resident for the context's lifetime
    GenTab tabs[KIWI_TAB_NKERNELS];
    for (int k = 0; k < KIWI_TAB_NKERNELS; ++k) gen_tab_init(tabs[k], kHostTables[k]);
    if (reserve(c->tables, sizeof(tabs)) != hipSuccess ||
        hipMemcpy(c->tables.p, tabs, sizeof(tabs), hipMemcpyHostToDevice) != hipSuccess) {
        kiwi_hip_destroy(c);
        return KIWI_ERR_DEVICE;
    }
    *out = c;
    return KIWI_OK;
}

void kiwi_hip_destroy(kiwi_hip_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->side2) (void)hipStreamSynchronize(c->side2);
    for (GraphEntry& g : c->graph_cache)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->graph_cache.clear();
    if (c->g_host) (void)hipHostFree(c->g_host);
    c->for_each_buf([](DevBuf& b) {
        if (b.p) (void)hipFree(b.p);
    });
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ev_side)
        if (e) (void)hipEventDestroy(e);
    if (c->side_done) (void)hipEventDestroy(c->side_done);
    if (c->side2_done) (void)hipEventDestroy(c->side2_done);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->fork2) (void)hipEventDestroy(c->fork2);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->side2) (void)hipStreamDestroy(c->side2);
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
    if (const int rc = check_problem(c, P, G, out)) return rc;
    if (precision != KIWI_FP64 && precision != KIWI_MIXED) return fail(c, KIWI_ERR_INVALID_ARG, "unknown precision");
    KernelChoice kc;
    if (const int rc = resolve_kernel(c, K, &kc)) return rc;
    const HostProblem hp{P, G, out, stats};
    return run_group(c, &hp, 1, K, kc, precision, c->own_bufs(), true);
}

int kiwi_hip_sph_to_grid_noalloc(kiwi_hip_ctx* c, const kiwi_particles* P, const kiwi_grid* G, const kiwi_kernel* K,
                                 int precision, const kiwi_cell_sums* out, kiwi_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!P || !G || !K || !out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (const int rc = check_problem(c, P, G, out)) return rc;
    if (precision != KIWI_FP64 && precision != KIWI_MIXED) return fail(c, KIWI_ERR_INVALID_ARG, "unknown precision");
    if ((P->n > 0 && !P->on_device) || !out->on_device)
        return fail(c, KIWI_ERR_INVALID_ARG, "the allocation-free entry point takes device particles and outputs");
    KernelChoice kc;
    if (const int rc = resolve_kernel(c, K, &kc)) return rc;
    if (kc.tabulated)
        return fail(c, KIWI_ERR_UNSUPPORTED, "the allocation-free entry point has no tabulated integral (its table "
                                             "is allocated on first use)");
    if (!scratch_address(c->device) || !g_scratch_ptr[c->device])
        return fail(c, KIWI_ERR_DEVICE, "the allocation-free scratch has no device address");
    // every buffer of the run is carved afresh from the start of this device's scratch
    c->noalloc_arena = ScratchArena{static_cast<char*>(g_scratch_ptr[c->device]), KIWI_NOALLOC_SCRATCH_BYTES, 0};
    PipeBufs B;
    int slot = 0;
#define KIWI_PB_SHADOW(n)                                       \
    c->noalloc_bufs[slot] = DevBuf{nullptr, 0, &c->noalloc_arena}; \
    B.n = &c->noalloc_bufs[slot++];
    KIWI_PIPE_BUFS(KIWI_PB_SHADOW)
#undef KIWI_PB_SHADOW
    const HostProblem hp{P, G, out, stats};
    const int rc = run_group(c, &hp, 1, K, kc, precision, B, false);
    if (rc == KIWI_ERR_OUT_OF_MEMORY) {
        return fail(c, KIWI_ERR_LIMIT, "the problem needs more device scratch than the allocation-free entry point "
                                       "holds (KIWI_NOALLOC_SCRATCH_BYTES = " +
                                           std::to_string(KIWI_NOALLOC_SCRATCH_BYTES >> 20) + " MiB): " + c->err);
    }
    return rc;
}

int kiwi_hip_sph_to_grid_batch(kiwi_hip_ctx* c, int n_problems, const kiwi_batch_problem* problems) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (n_problems < 0 || (n_problems > 0 && !problems)) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (n_problems > KIWI_MAX_BATCH) return fail(c, KIWI_ERR_LIMIT, "more problems than KIWI_MAX_BATCH in one batch");
    std::vector<KernelChoice> kcs(static_cast<size_t>(n_problems));
    for (int i = 0; i < n_problems; ++i) {
        const kiwi_batch_problem& q = problems[i];
        if (const int rc = check_problem(c, &q.particles, &q.grid, &q.out)) {
            c->err = "problem " + std::to_string(i) + ": " + c->err;
            return rc;
        }
        if (q.precision != KIWI_FP64 && q.precision != KIWI_MIXED) {
            return fail(c, KIWI_ERR_INVALID_ARG, "problem " + std::to_string(i) + ": unknown precision");
        }
        if (const int rc = resolve_kernel(c, &q.kernel, &kcs[i])) {
            c->err = "problem " + std::to_string(i) + ": " + c->err;
            return rc;
        }
    }
    // one pipeline run per kernel path: problems with the same kernel, support and arithmetic
    std::vector<char> done(static_cast<size_t>(n_problems), 0);
    for (int i = 0; i < n_problems; ++i) {
        if (done[i]) continue;
        std::vector<HostProblem> group;
        for (int j = i; j < n_problems; ++j) {
            if (done[j]) continue;
            if (same_path(problems[i], kcs[i], problems[j], kcs[j], c->fb_cells)) {
                group.push_back(HostProblem{&problems[j].particles, &problems[j].grid, &problems[j].out,
                                            problems[j].stats});
                done[j] = 1;
            }
        }
        if (const int rc = run_group(c, group.data(), static_cast<int>(group.size()), &problems[i].kernel, kcs[i],
                                     problems[i].precision, c->own_bufs(), true))
            return rc;
    }
    return KIWI_OK;
}

int kiwi_hip_set_tuning(kiwi_hip_ctx* c, const kiwi_tuning* t) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    if (!t) {
        c->tune = kiwi_tuning{16, 8};
        return KIWI_OK;
    }
    if (t->item_grab < 1 || t->item_grab > 256) return fail(c, KIWI_ERR_INVALID_ARG, "item_grab must be in 1..256");
    if (t->cursors < 1 || t->cursors > KIWI_MAX_CURSORS)
        return fail(c, KIWI_ERR_INVALID_ARG, "cursors must be in 1..KIWI_MAX_CURSORS");
    c->tune = *t;
    return KIWI_OK;
}

int kiwi_hip_set_graphs(kiwi_hip_ctx* c, int on) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    if (on != 0 && on != 1) return fail(c, KIWI_ERR_INVALID_ARG, "graph mode must be 0 or 1");
    c->graphs = on;
    if (!on) {  // the captured pipelines are released with the mode
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (GraphEntry& g : c->graph_cache)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        c->graph_cache.clear();
    }
    return KIWI_OK;
}

int kiwi_hip_get_graphs(const kiwi_hip_ctx* c, int* on) {
    if (!c || !on) return KIWI_ERR_INVALID_ARG;
    *on = c->graphs;
    return KIWI_OK;
}

int kiwi_hip_set_fp64_fallback(kiwi_hip_ctx* c, double support_cells) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    if (std::isnan(support_cells)) return fail(c, KIWI_ERR_INVALID_ARG, "fp64 fallback threshold is NaN");
    c->fb_cells = support_cells;
    return KIWI_OK;
}

int kiwi_hip_get_fp64_fallback(const kiwi_hip_ctx* c, double* support_cells) {
    if (!c || !support_cells) return KIWI_ERR_INVALID_ARG;
    *support_cells = c->fb_cells;
    return KIWI_OK;
}

int kiwi_hip_get_tuning(const kiwi_hip_ctx* c, kiwi_tuning* t) {
    if (!c || !t) return KIWI_ERR_INVALID_ARG;
    *t = c->tune;
    return KIWI_OK;
}

// The launch-shape autotuner: every candidate runs the caller's batch twice (the calls are synchronous,
// so the host clock brackets the whole call); the fastest candidate stays on the context.
int kiwi_hip_autotune(kiwi_hip_ctx* c, int n_problems, const kiwi_batch_problem* problems, kiwi_tuning* best,
                      double* best_ms) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    if (!best) return fail(c, KIWI_ERR_INVALID_ARG, "null tuning result");
    const kiwi_tuning saved = c->tune;
    // warm-up with the current shape: buffers reach their size, tables are built
    if (const int rc = kiwi_hip_sph_to_grid_batch(c, n_problems, problems)) return rc;
    static const int kGrab[] = {8, 16, 32, 64};
    static const int kCursors[] = {1, 4, 8, 16};
    kiwi_tuning bt = saved;
    double bms = INFINITY;
    for (const int g : kGrab)
        for (const int k : kCursors) {
            c->tune = kiwi_tuning{g, k};
            double t = INFINITY;
            for (int rep = 0; rep < 2; ++rep) {
                const auto t0 = std::chrono::steady_clock::now();
                if (const int rc = kiwi_hip_sph_to_grid_batch(c, n_problems, problems)) {
                    c->tune = saved;
                    return rc;
                }
                t = std::min(t, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            }
            if (t < bms) {
                bms = t;
                bt = c->tune;
            }
        }
    c->tune = bt;
    *best = bt;
    if (best_ms) *best_ms = bms;
    return KIWI_OK;
}

}  // extern "C"
