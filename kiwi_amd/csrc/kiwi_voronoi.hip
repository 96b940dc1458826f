// libkiwi_hip -- the Voronoi map: per-cell mass and momentum of SPH particles over a Voronoi
// tessellation by exact face integrals, the GPU form of run_mapping_stage's particle loop
// (src/mapping.cpp:400-506) with compute_cell_integral (src/mapping.cpp:105-206).
//
// The reference walks, per particle, a BFS outward from the particle's home cell: a cell is visited
// when it is the home cell or when a visited neighbour has a face, shared with it, that comes closer to
// the particle than the support radius.  Within the support ball (convex, inside the convex domain) the
// cells met by the ball are connected through faces met by the ball, so the visited set is exactly
// {home} + {cells with a face closer than the support}: a property of each (particle, cell) pair alone.
// The GPU therefore enumerates candidate pairs with a bucket grid, decides visits per pair, and keeps
// the reference's face cache as a structure: a face shared by two visited cells of one particle is
// integrated once, by the lower cell index, and enters the other cell negated (the reference integrates
// it from whichever side its BFS reaches first; the two sides agree to rounding).
//
//   1. k_vor_cells,     per-cell bounding boxes and the face -> cell map (validating the CSR geometry),
//      k_vor_twins      then each face's twin across its neighbour
//   2. bucket grid      cells binned by bounding box (count, scan, emit, sort, ranges)
//   3. candidates       per particle, the cells whose boxes meet the support ball (count, scan, fill)
//   4. k_vor_reach      visited flag per candidate pair; compaction; (particle, cell) sort
//   5. owned faces      per visited pair, which of its faces it integrates and where mirrored values
//                       come from; compaction to a flat list of (pair, face) items, grouped by vertex
//                       count (FP64 face walks) or sorted by face (FP32 face walks)
//   6. k_vor_faces      the face integrals, one thread per (particle, face) -- the hot kernel
//   7. k_vor_pairsum    per-pair cell integral in the cell's face order, clamped at 0
//   8. accumulation     pairs sorted by (cell, particle); one owner per cell sums in particle order
//
// Everything is FP64 except, for a mixed-precision request with the analytic cubic spline, the face
// integrals (single precision on vertex coordinates relative to the particle); the visit test, the
// per-pair and per-cell sums stay FP64, and the sums are deterministic (no floating-point atomics).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "common.h"
#include "face_walk.h"

using namespace kiwi_hip;

namespace {

constexpr double BUCKET_CELLS = 2.0;  // mean cells per bucket of the candidate grid
constexpr int MAX_BUCKET_AXIS = 1024;
constexpr double SUPPORT_MARGIN = 1.0 + 1e-7;  // candidate test slack over the exact face test
// Lanes per (particle, cell) pair in the per-pair kernels (a quarter wavefront): the pair's faces are
// taken 16 at a time, so that the face arrays and the pair's slots are read and written contiguously.
constexpr int PAIR_LANES = 16;

struct MapDevStats {
    unsigned long long unresolved, unmapped, n_negative, n_negative_warn;
    unsigned long long bad_geometry, bad_particle, bad_home, fill_mismatch;
    unsigned long long n_warn;
    kiwi_warn_record warn[KIWI_MAX_WARN_RECORDS];
};

struct BucketGrid {
    double lo0, lo1, lo2;
    double inv0, inv1, inv2;  // buckets per unit length
    int g0, g1, g2;
};

// bucket index along axis a of coordinate x, clamped into the grid (NaN -> 0)
__device__ __forceinline__ int bucket_axis(const BucketGrid& B, int a, double x) {
    const double lo = sel3(a, B.lo0, B.lo1, B.lo2), inv = sel3(a, B.inv0, B.inv1, B.inv2);
    const int g = sel3(a, B.g0, B.g1, B.g2);
    const double t = floor((x - lo) * inv);
    if (!(t >= 0.0)) return 0;
    if (t >= static_cast<double>(g - 1)) return g - 1;
    return static_cast<int>(t);
}

// the reachability-only face walk: no vertex integrals
struct EvNone {
    __device__ void wall(double) {}
    __device__ void edge(double) {}
    __device__ double at(double, double) const { return 0.0; }
};

__global__ void k_fill_i32(int* p, int64_t n, int v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// 1. geometry: cell boxes, face -> cell, twins, validation
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void k_vor_cells(int64_t nc, int64_t nf, int64_t nv,
                                                       const int64_t* __restrict__ cfp,
                                                       const int64_t* __restrict__ fvp,
                                                       const double* __restrict__ vxyz,
                                                       const int* __restrict__ fnb, double* __restrict__ box,
                                                       int* __restrict__ face_cell, MapDevStats* st) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (c >= nc) return;
    const int64_t f0 = cfp[c], f1 = cfp[c + 1];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool bad = !(f0 >= 0 && f0 <= f1 && f1 <= nf);
    if (!bad) {
        for (int64_t f = f0; f < f1; ++f) {
            face_cell[f] = static_cast<int>(c);
            const int64_t v0 = fvp[f], v1 = fvp[f + 1];
            const int nb = fnb[f];
            if (!(v0 >= 0 && v0 <= v1 && v1 <= nv) || nb >= nc || nb < -1 || nb == c) {
                bad = true;
                break;
            }
            for (int64_t v = v0; v < v1; ++v) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double x = vxyz[3 * v + a];
                    lo[a] = fmin(lo[a], x);
                    hi[a] = fmax(hi[a], x);
                }
            }
        }
    }
    if (c == 0 && (cfp[0] != 0 || cfp[nc] != nf)) bad = true;
    if (bad) {
        atomicAdd(&st->bad_geometry, 1ull);
        lo[0] = lo[1] = lo[2] = INFINITY;
        hi[0] = hi[1] = hi[2] = -INFINITY;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        box[6 * c + a] = lo[a];
        box[6 * c + 3 + a] = hi[a];
    }
}

// twin[f]: the face of the neighbouring cell that points back, when each of the two cells has exactly
// one face towards the other (always so for a non-periodic tessellation); -1 otherwise
__global__ __launch_bounds__(THREADS) void k_vor_twins(int64_t nf, const int64_t* __restrict__ cfp,
                                                       const int* __restrict__ fnb, const int* __restrict__ face_cell,
                                                       int* __restrict__ twin) {
    const int64_t f = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (f >= nf) return;
    const int nb = fnb[f];
    int t = -1;
    if (nb >= 0) {
        const int c = face_cell[f];
        int own = 0;
        for (int64_t g = cfp[c]; g < cfp[c + 1]; ++g) own += (fnb[g] == nb) ? 1 : 0;
        int back = 0;
        for (int64_t g = cfp[nb]; g < cfp[nb + 1]; ++g)
            if (fnb[g] == c) {
                ++back;
                t = static_cast<int>(g);
            }
        if (own != 1 || back != 1) t = -1;
    }
    twin[f] = t;
}

// Per-face bounds for the visit test, one cache line per face: the unit normal and offset of the plane
// through the first three vertices (n . x + off is the signed plane distance; all zero for a face without
// a plane), and a bounding sphere of the vertices (their centroid and largest distance from it).
struct alignas(64) FaceGeo {
    double nx, ny, nz, off;
    double cx, cy, cz, rad;
};

__global__ __launch_bounds__(THREADS) void k_vor_face_geo(int64_t nf, const int64_t* __restrict__ fvp,
                                                          const double* __restrict__ vxyz, FaceGeo* __restrict__ fg) {
    const int64_t f = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (f >= nf) return;
    const int64_t v0 = fvp[f], n = fvp[f + 1] - v0;
    const double* v = vxyz + 3 * v0;
    FaceGeo g = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, INFINITY};
    if (n >= 3) {
        const double ax = v[3] - v[0], ay = v[4] - v[1], az = v[5] - v[2];
        const double bx = v[6] - v[0], by = v[7] - v[1], bz = v[8] - v[2];
        const double A = ay * bz - az * by, B = az * bx - ax * bz, C = ax * by - ay * bx;
        const double nrm = sqrt(A * A + B * B + C * C);
        if (nrm > 0.0) {
            g.nx = A / nrm;
            g.ny = B / nrm;
            g.nz = C / nrm;
            g.off = -(g.nx * v[0] + g.ny * v[1] + g.nz * v[2]);
        }
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int64_t k = 0; k < n; ++k) {
            sx += v[3 * k];
            sy += v[3 * k + 1];
            sz += v[3 * k + 2];
        }
        g.cx = sx / static_cast<double>(n);
        g.cy = sy / static_cast<double>(n);
        g.cz = sz / static_cast<double>(n);
        double r2 = 0.0;
        for (int64_t k = 0; k < n; ++k) {
            const double dx = v[3 * k] - g.cx, dy = v[3 * k + 1] - g.cy, dz = v[3 * k + 2] - g.cz;
            r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
        }
        g.rad = sqrt(r2);
    }
    fg[f] = g;
}

// ------------------------------------------------------------------------------------------------
// 2. bucket grid over the cell boxes
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool cell_bucket_range(const BucketGrid& B, const double* __restrict__ box, int64_t c,
                                                  int* blo, int* bhi) {
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double lo = box[6 * c + a], hi = box[6 * c + 3 + a];
        empty |= !(lo <= hi);
        blo[a] = bucket_axis(B, a, lo);
        bhi[a] = bucket_axis(B, a, hi);
    }
    return !empty;
}

// entries per cell, and the cell's lowest bucket per axis packed 10 bits each (x | y << 10 | z << 20)
// What the candidate test reads of a cell, in one cache line: its box and lowest bucket per axis.
struct alignas(64) CellRec {
    double lo[3], hi[3];
    int blo;  // x | y << 10 | z << 20
};

__global__ __launch_bounds__(THREADS) void k_vor_bucket_count(int64_t nc, BucketGrid B,
                                                              const double* __restrict__ box,
                                                              int64_t* __restrict__ cnt, CellRec* __restrict__ crec) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (c > nc) return;
    int64_t k = 0;
    int blo[3], bhi[3];
    if (c < nc) {
        if (cell_bucket_range(B, box, c, blo, bhi))
            k = static_cast<int64_t>(bhi[0] - blo[0] + 1) * (bhi[1] - blo[1] + 1) * (bhi[2] - blo[2] + 1);
        CellRec r;
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = box[6 * c + a];
            r.hi[a] = box[6 * c + 3 + a];
        }
        r.blo = blo[0] | (blo[1] << 10) | (blo[2] << 20);
        crec[c] = r;
    }
    cnt[c] = k;  // cnt[nc] = 0 closes the scan
}

__global__ __launch_bounds__(THREADS) void k_vor_bucket_emit(int64_t nc, BucketGrid B,
                                                             const double* __restrict__ box,
                                                             const int64_t* __restrict__ off,
                                                             uint64_t* __restrict__ keys) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (c >= nc) return;
    int blo[3], bhi[3];
    if (!cell_bucket_range(B, box, c, blo, bhi)) return;
    int64_t o = off[c];
    for (int x = blo[0]; x <= bhi[0]; ++x)
        for (int y = blo[1]; y <= bhi[1]; ++y)
            for (int z = blo[2]; z <= bhi[2]; ++z) {
                const uint64_t b = (static_cast<uint64_t>(x) * B.g1 + y) * B.g2 + z;
                keys[o++] = (b << 32) | static_cast<uint32_t>(c);
            }
}

// boff[b] = first sorted entry of bucket b, boff[nbuckets] = ne: entries of bucket b are
// [boff[b], boff[b+1]), and a run of consecutive buckets is one contiguous range
__global__ __launch_bounds__(THREADS) void k_vor_bucket_offsets(int64_t ne, int64_t nbuckets,
                                                                const uint64_t* __restrict__ keys,
                                                                int* __restrict__ boff) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= ne) return;
    const int64_t b = static_cast<int64_t>(keys[t] >> 32);
    const int64_t prev = (t == 0) ? -1 : static_cast<int64_t>(keys[t - 1] >> 32);
    for (int64_t bb = prev + 1; bb <= b; ++bb) boff[bb] = static_cast<int>(t);
    if (t == ne - 1)
        for (int64_t bb = b + 1; bb <= nbuckets; ++bb) boff[bb] = static_cast<int>(ne);
}

// ------------------------------------------------------------------------------------------------
// 3. candidate (particle, cell) pairs: cell boxes that meet the support ball, plus the home cell
// ------------------------------------------------------------------------------------------------
// One wavefront per particle.  The query box's buckets are walked as runs along z, each one contiguous
// range of sorted entries that the 64 lanes test in parallel; accepted cells are compacted in lane
// order with a ballot, so the count and fill passes produce the same sequence.  Active particles have
// h > 0 and a home cell (src/mapping.cpp:425, 436-440).  Each cell is taken once per particle: in the
// lowest bucket, per axis, that both its box and the query box cover.
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

template <bool FILL>
__global__ __launch_bounds__(THREADS) void k_vor_candidates(
    int64_t n, const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ ph, const int* __restrict__ home, int64_t nc, double sup, BucketGrid B,
    const CellRec* __restrict__ crec, const uint64_t* __restrict__ bkeys,
    const int* __restrict__ boff, int64_t* __restrict__ cnt, const int64_t* __restrict__ off,
    int* __restrict__ cand_cell, int* __restrict__ cand_part, MapDevStats* st) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int64_t i = (static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x) >> 6;  // uniform per wave
    if (i > n) return;
    if (i == n) {
        if (!FILL && lane == 0) cnt[n] = 0;
        return;
    }
    const double h = ph[i];
    const int hc = home[i];
    int64_t k = 0;
    int64_t o = FILL ? off[i] : 0;
    const int64_t o_end = FILL ? off[i + 1] : 0;
    const bool finite_h = (h <= 0.0) || (h < INFINITY);  // NaN and +inf are rejected by the caller
    if (finite_h && h > 0.0 && hc >= 0 && hc < nc) {
        const double H = sup * h;
        const double p0 = px[i], p1 = py[i], p2 = pz[i];
        const double lim = H * SUPPORT_MARGIN;
        const double lim2 = lim * lim;
        const int q0 = bucket_axis(B, 0, p0 - lim), q0e = bucket_axis(B, 0, p0 + lim);
        const int q1 = bucket_axis(B, 1, p1 - lim), q1e = bucket_axis(B, 1, p1 + lim);
        const int q2 = bucket_axis(B, 2, p2 - lim), q2e = bucket_axis(B, 2, p2 + lim);
        const uint64_t below = lanes_below(lane);
        bool home_seen = false;
        for (int x = q0; x <= q0e; ++x)
            for (int y = q1; y <= q1e; ++y) {
                const int row = (x * B.g1 + y) * B.g2;
                const int e0 = boff[row + q2], e1 = boff[row + q2e + 1];
                for (int base = e0; base < e1; base += 64) {
                    const int e = base + lane;
                    bool acc = false, is_home = false;
                    int c = 0;
                    if (e < e1) {
                        const uint64_t key = bkeys[e];
                        c = static_cast<int>(key & 0xffffffffu);
                        const int z = static_cast<int>(key >> 32) - row;
                        const CellRec& cr = crec[c];
                        const int cb = cr.blo;
                        // lowest shared bucket per axis
                        if (max(cb & 1023, q0) == x && max((cb >> 10) & 1023, q1) == y && max(cb >> 20, q2) == z) {
                            // rounded operations: the count and fill passes take identical decisions
                            const double g0 = gap(cr.lo[0] - p0, cr.hi[0] - p0);
                            const double g1 = gap(cr.lo[1] - p1, cr.hi[1] - p1);
                            const double g2 = gap(cr.lo[2] - p2, cr.hi[2] - p2);
                            const double d2 = add_sep(add_sep(mul_sep(g0, g0), mul_sep(g1, g1)),
                                                        mul_sep(g2, g2));
                            is_home = (c == hc);
                            acc = (d2 < lim2) || is_home;
                        }
                    }
                    const uint64_t m = __ballot(acc);
                    home_seen |= (__ballot(acc && is_home) != 0ull);
                    if (FILL && acc) {
                        const int64_t pos = o + __popcll(m & below);
                        if (pos < o_end) {
                            cand_cell[pos] = c;
                            cand_part[pos] = static_cast<int>(i);
                        }
                    }
                    if (FILL) o += __popcll(m);
                    k += __popcll(m);
                }
            }
        if (!home_seen) {
            if (FILL && lane == 0 && o < o_end) {
                cand_cell[o] = hc;
                cand_part[o] = static_cast<int>(i);
            }
            if (FILL) ++o;
            ++k;
        }
        if (FILL && k != o_end - off[i]) {
            // cannot happen with identical decisions; keep every slot valid and report it
            if (lane == 0) {
                for (int64_t q = off[i] + (k < o_end - off[i] ? k : o_end - off[i]); q < o_end; ++q) {
                    cand_cell[q] = hc;
                    cand_part[q] = static_cast<int>(i);
                }
                atomicAdd(&st->fill_mismatch, 1ull);
            }
        }
    } else if (!FILL && lane == 0) {
        if (!finite_h) atomicAdd(&st->bad_particle, 1ull);
        else if (h > 0.0 && hc >= nc) atomicAdd(&st->bad_home, 1ull);
        else if (h > 0.0) atomicAdd(&st->unresolved, 1ull);  // h <= 0 is skipped silently (:425)
    }
    if (!FILL && lane == 0) cnt[i] = k;
}

// ------------------------------------------------------------------------------------------------
// 4. visits
// ------------------------------------------------------------------------------------------------
// the face reachability test of the reference's face walk (out_vert != n, src/kernel_integral.cpp:
// 478-481, 642-645): a face with fewer than three vertices counts as reachable unless it has none.  The
// face's distance is at least its plane's, so |r0| >= H rejects it before the edge walk.
__device__ __forceinline__ bool face_reachable(double x0, double y0, double z0, double H, const double* v, int nvf) {
    if (nvf < 3) return nvf != 0;
    const double x1 = v[0], y1 = v[1], z1 = v[2];
    const double ax = v[3] - x1, ay = v[4] - y1, az = v[5] - z1;
    const double bx = v[6] - x1, by = v[7] - y1, bz = v[8] - z1;
    const double A = ay * bz - by * az, B = az * bx - bz * ax, C = ax * by - bx * ay;
    const double nrm = sqrt(A * A + B * B + C * C);
    const double r0 = (A * (x0 - x1) + B * (y0 - y1) + C * (z0 - z1)) / nrm;
    if (fabs(r0) >= H) return false;
    EvNone ev;
    bool r = false;
    (void)face_walk_integral(x0, y0, z0, H, v, nvf, ev, &r);
    return r;
}

// The cheap part of a face's test: a face whose plane or bounding sphere lies beyond the support is
// rejected without its vertices (the bound is widened by 1e-9 H plus 1e-12 of the coordinates' magnitude,
// far above the rounding of either distance, so a rejected face is one the exact test rejects too); one of
// its first three vertices inside the support by a relative margin of 1e-6 on the squared distance makes
// it reachable (its distance is at most that vertex's); a face of fewer than three vertices is decided by
// face_reachable at once.  FACE_EXACT: the rest, which take the exact test.
constexpr int FACE_OUT = 0, FACE_IN = 1, FACE_EXACT = 2;

__device__ __forceinline__ int face_quick(double x0, double y0, double z0, double H, double slack, const FaceGeo& g,
                                          const double* v, int nvf) {
    if (fabs(g.nx * x0 + g.ny * y0 + g.nz * z0 + g.off) >= H + slack + 1e-12 * fabs(g.off)) return FACE_OUT;
    const double dx = x0 - g.cx, dy = y0 - g.cy, dz = z0 - g.cz;
    if (sqrt(dx * dx + dy * dy + dz * dz) - g.rad >= H + slack + 1e-12 * g.rad) return FACE_OUT;
    if (nvf < 3) return nvf != 0 ? FACE_IN : FACE_OUT;
    const double in2 = H * H * (1.0 - 1e-6);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double ex = v[3 * k] - x0, ey = v[3 * k + 1] - y0, ez = v[3 * k + 2] - z0;
        if (ex * ex + ey * ey + ez * ez < in2) return FACE_IN;
    }
    return FACE_EXACT;
}

// The visit test of a candidate pair is the OR over its cell's faces of face_quick, with face_reachable
// for the faces face_quick leaves open.  Those exact tests are few and long; tested where they arise, a
// wavefront's other lanes would wait on them, so they run in a pass of their own, one lane per (pair,
// face).  First pass: the home cell and the faces' cheap tests; the pair's visit flag, or the number of
// its faces left open (nexact) and which (exmask, bit j for the cell's face j < 64).  Entry ncand of
// flag and nexact is 0 (the scans' totals).
__global__ __launch_bounds__(THREADS) void k_vor_reach(int64_t ncand, const int* __restrict__ cand_cell,
                                                       const int* __restrict__ cand_part,
                                                       const double* __restrict__ px,
                                                       const double* __restrict__ py,
                                                       const double* __restrict__ pz,
                                                       const double* __restrict__ ph, const int* __restrict__ home,
                                                       double sup, const int64_t* __restrict__ cfp,
                                                       const int64_t* __restrict__ fvp,
                                                       const double* __restrict__ vxyz,
                                                       const FaceGeo* __restrict__ fg, int64_t* __restrict__ flag,
                                                       int64_t* __restrict__ nexact, uint64_t* __restrict__ exmask) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t > ncand) return;
    if (t == ncand) {
        flag[t] = 0;
        nexact[t] = 0;
        return;
    }
    const int c = cand_cell[t], i = cand_part[t];
    bool vis = (home[i] == c);
    int64_t nex = 0;
    uint64_t mask = 0;
    if (!vis) {
        const double x0 = px[i], y0 = py[i], z0 = pz[i], H = sup * ph[i];
        const double slack = 1e-9 * H + 1e-12 * (fabs(x0) + fabs(y0) + fabs(z0));
        const int64_t f0 = cfp[c];
        for (int64_t f = f0; f < cfp[c + 1] && !vis; ++f) {
            const int64_t v0 = fvp[f];
            const int q = face_quick(x0, y0, z0, H, slack, fg[f], vxyz + 3 * v0, static_cast<int>(fvp[f + 1] - v0));
            vis = q == FACE_IN;
            if (q == FACE_EXACT) {
                ++nex;
                if (f - f0 < 64) mask |= 1ull << (f - f0);
            }
        }
    }
    flag[t] = vis ? 1 : 0;
    nexact[t] = vis ? 0 : nex;
    exmask[t] = mask;
}

// Second pass: the (pair, face) entries of the faces left open, pair t's at ex_off[t], in face order:
// read from exmask, or, for a cell of more than 64 faces, decided again by the same cheap tests.
__global__ __launch_bounds__(THREADS) void k_vor_reach_list(int64_t ncand, const int* __restrict__ cand_cell,
                                                            const int* __restrict__ cand_part,
                                                            const double* __restrict__ px,
                                                            const double* __restrict__ py,
                                                            const double* __restrict__ pz,
                                                            const double* __restrict__ ph, double sup,
                                                            const int64_t* __restrict__ cfp,
                                                            const int64_t* __restrict__ fvp,
                                                            const double* __restrict__ vxyz,
                                                            const FaceGeo* __restrict__ fg,
                                                            const int64_t* __restrict__ nexact,
                                                            const uint64_t* __restrict__ exmask,
                                                            const int64_t* __restrict__ ex_off,
                                                            uint64_t* __restrict__ ex) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= ncand || nexact[t] == 0) return;
    const int c = cand_cell[t], i = cand_part[t];
    if (cfp[c + 1] - cfp[c] <= 64) {
        int64_t o = ex_off[t];
        for (uint64_t m = exmask[t]; m; m &= m - 1ull)
            ex[o++] = (static_cast<uint64_t>(t) << 32) | static_cast<uint64_t>(__ffsll(static_cast<long long>(m)) - 1);
        return;
    }
    const double x0 = px[i], y0 = py[i], z0 = pz[i], H = sup * ph[i];
    const double slack = 1e-9 * H + 1e-12 * (fabs(x0) + fabs(y0) + fabs(z0));
    int64_t o = ex_off[t];
    const int64_t o_end = o + nexact[t];
    for (int64_t f = cfp[c]; f < cfp[c + 1] && o < o_end; ++f) {
        const int64_t v0 = fvp[f];
        if (face_quick(x0, y0, z0, H, slack, fg[f], vxyz + 3 * v0, static_cast<int>(fvp[f + 1] - v0)) == FACE_EXACT)
            ex[o++] = (static_cast<uint64_t>(t) << 32) | static_cast<uint64_t>(f - cfp[c]);
    }
}

// Third pass: one exact test per listed (pair, face); a reachable face sets its pair's flag (every
// writer writes 1, so the order of the writes does not matter).
__global__ __launch_bounds__(THREADS) void k_vor_reach_exact(int64_t nex, const uint64_t* __restrict__ ex,
                                                             const int* __restrict__ cand_cell,
                                                             const int* __restrict__ cand_part,
                                                             const double* __restrict__ px,
                                                             const double* __restrict__ py,
                                                             const double* __restrict__ pz,
                                                             const double* __restrict__ ph, double sup,
                                                             const int64_t* __restrict__ cfp,
                                                             const int64_t* __restrict__ fvp,
                                                             const double* __restrict__ vxyz,
                                                             int64_t* __restrict__ flag) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (q >= nex) return;
    const uint64_t e = ex[q];
    const int64_t t = static_cast<int64_t>(e >> 32);
    const int c = cand_cell[t], i = cand_part[t];
    const int64_t f = cfp[c] + static_cast<int64_t>(e & 0xffffffffull);
    const int64_t v0 = fvp[f];
    if (face_reachable(px[i], py[i], pz[i], sup * ph[i], vxyz + 3 * v0, static_cast<int>(fvp[f + 1] - v0))) flag[t] = 1;
}

// visited pairs of particle i: [vis_off[i], vis_off[i+1]); particles that reached only their home cell
__global__ __launch_bounds__(THREADS) void k_vor_particle_visits(int64_t n, const int64_t* __restrict__ cand_off,
                                                                 const int64_t* __restrict__ flag_off,
                                                                 int64_t* __restrict__ vis_off, MapDevStats* st) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i > n) return;
    const int64_t v = flag_off[cand_off[i]];
    vis_off[i] = v;
    if (i < n && flag_off[cand_off[i + 1]] - v == 1) atomicAdd(&st->unmapped, 1ull);
}

__global__ __launch_bounds__(THREADS) void k_vor_compact(int64_t ncand, const int64_t* __restrict__ flag,
                                                         const int64_t* __restrict__ flag_off,
                                                         const int* __restrict__ cand_cell,
                                                         const int* __restrict__ cand_part, int cbits,
                                                         uint64_t* __restrict__ keys) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= ncand || !flag[t]) return;
    keys[flag_off[t]] = (static_cast<uint64_t>(cand_part[t]) << cbits) | static_cast<uint32_t>(cand_cell[t]);
}

// ------------------------------------------------------------------------------------------------
// 5. owned faces
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void k_vor_pair_faces(int64_t nvis, const uint64_t* __restrict__ keys,
                                                            uint64_t cmask, const int64_t* __restrict__ cfp,
                                                            int64_t* __restrict__ nface) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (q > nvis) return;
    if (q == nvis) {
        nface[q] = 0;
        return;
    }
    const int64_t c = static_cast<int64_t>(keys[q] & cmask);
    nface[q] = cfp[c + 1] - cfp[c];
}

// position of key in the sorted range [lo, hi), or -1
__device__ __forceinline__ int64_t find_key(const uint64_t* __restrict__ keys, int64_t lo, int64_t e, uint64_t key) {
    int64_t hi = e;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const uint64_t k = keys[mid];
        if (k < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < e && keys[lo] == key) ? lo : -1;
}

// The pairs a block's searches can reach, [first pair of the block's first particle, end of the block),
// are staged in LDS when they fit OWN_LDS keys; the binary searches then read LDS.
constexpr int OWN_LDS = 2048;

__device__ __forceinline__ int64_t find_key_lds(const uint64_t* __restrict__ sk, int64_t base, int64_t lo, int64_t e,
                                                uint64_t key) {
    int64_t l = lo - base, h = e - base;
    const int64_t end = h;
    while (l < h) {
        const int64_t mid = (l + h) >> 1;
        if (sk[mid] < key) l = mid + 1;
        else h = mid;
    }
    return (l < end && sk[l] == key) ? l + base : -1;
}

// sum over the 16 lanes of a quarter wavefront
__device__ __forceinline__ int64_t row_sum(int64_t v) {
    for (int m = 1; m < PAIR_LANES; m <<= 1) v += __shfl_xor(v, m, PAIR_LANES);
    return v;
}

// src[slot]: the slot itself for a face this pair integrates, ~(owner's slot) for a face shared with a
// lower-index visited cell of the same particle (integrated there, entered here negated)
__global__ __launch_bounds__(THREADS) void k_vor_owned(int64_t nvis, const uint64_t* __restrict__ keys, int cbits,
                                                       const int64_t* __restrict__ vis_off,
                                                       const int64_t* __restrict__ slot_off,
                                                       const int64_t* __restrict__ cfp, const int* __restrict__ fnb,
                                                       const int* __restrict__ twin, int64_t* __restrict__ src,
                                                       int64_t* __restrict__ nown) {
    constexpr int PAIRS = THREADS / PAIR_LANES;
    __shared__ uint64_t sk[OWN_LDS];
    const int64_t q0 = static_cast<int64_t>(blockIdx.x) * PAIRS;
    const int64_t q = q0 + threadIdx.x / PAIR_LANES;
    const int r = static_cast<int>(threadIdx.x % PAIR_LANES);
    // the block's search range (uniform): every search of pair q lies in [vis_off[i(q)], q)
    const int64_t qe = min(q0 + PAIRS, nvis);
    int64_t base = 0;
    bool staged = false;
    if (q0 < nvis) {
        base = vis_off[static_cast<int64_t>(keys[q0] >> cbits)];
        staged = qe - base <= OWN_LDS;
        if (staged)
            for (int64_t k = base + threadIdx.x; k < qe; k += THREADS) sk[k - base] = keys[k];
    }
    __syncthreads();
    if (q > nvis) return;
    if (q == nvis) {
        if (r == 0) nown[q] = 0;
        return;
    }
    const uint64_t cmask = (1ull << cbits) - 1;
    const uint64_t key = keys[q];
    const int c = static_cast<int>(key & cmask);
    const uint64_t pbase = key & ~cmask;
    const int64_t i = static_cast<int64_t>(key >> cbits);
    const int64_t lo = vis_off[i];
    const int64_t f0 = cfp[c], f1 = cfp[c + 1];
    const int64_t slot0 = slot_off[q];
    int64_t owned = 0;
    for (int64_t f = f0 + r; f < f1; f += PAIR_LANES) {
        const int64_t slot = slot0 + (f - f0);
        const int nb = fnb[f], tw = twin[f];
        int64_t s = slot;
        if (nb >= 0 && nb < c && tw >= 0) {
            // the neighbour's pair sorts before this one
            const uint64_t k2 = pbase | static_cast<uint32_t>(nb);
            const int64_t q2 = staged ? find_key_lds(sk, base, lo, q, k2) : find_key(keys, lo, q, k2);
            if (q2 >= 0) s = ~(slot_off[q2] + (tw - cfp[nb]));
        }
        src[slot] = s;
        owned += (s >= 0) ? 1 : 0;
    }
    owned = row_sum(owned);
    if (r == 0) nown[q] = owned;
}

struct FaceItem {
    int pair;  // visited-pair index
    int face;  // global face index
};

// the pair's owned faces in face order, 16 at a time: each lane's rank among the owned faces of its round
// comes from the quarter wavefront's ballot bits
__global__ __launch_bounds__(THREADS) void k_vor_items(int64_t nvis, const uint64_t* __restrict__ keys,
                                                       uint64_t cmask, const int64_t* __restrict__ slot_off,
                                                       const int64_t* __restrict__ own_off,
                                                       const int64_t* __restrict__ cfp,
                                                       const int64_t* __restrict__ src, FaceItem* __restrict__ items) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    const int64_t q = gid / PAIR_LANES;
    const int r = static_cast<int>(gid % PAIR_LANES);
    const int shift = static_cast<int>(threadIdx.x & 63) & ~(PAIR_LANES - 1);  // the row's first lane
    const bool live = q < nvis;
    int64_t f0 = 0, f1 = 0, slot0 = 0, o = 0;
    if (live) {
        const int64_t c = static_cast<int64_t>(keys[q] & cmask);
        f0 = cfp[c];
        f1 = cfp[c + 1];
        slot0 = slot_off[q];
        o = own_off[q];
    }
    // every lane of the wavefront takes part in every round (the ballots), live or not
    int64_t nf = f1 - f0;
    for (int off = 1; off < 64; off <<= 1) nf = max(nf, static_cast<int64_t>(__shfl_xor(nf, off)));
    for (int64_t k = 0; k < nf; k += PAIR_LANES) {
        const int64_t f = f0 + k + r;
        const bool own = live && f < f1 && src[slot0 + k + r] >= 0;
        const uint64_t b = __ballot(own);
        const uint32_t rowbits = static_cast<uint32_t>((b >> shift) & 0xffffull);
        if (own) items[o + __popc(rowbits & ((1u << r) - 1u))] = FaceItem{static_cast<int>(q), static_cast<int>(f)};
        o += __popc(rowbits);
    }
}

// Face items in the order the face kernel reads them best.  Every item writes its own slot, so the order
// changes no result.
//  - FP32 face walks are close to memory-bound: the items are sorted by face (a stable radix sort of the
//    items read as 64-bit keys, face in the high word), so the lanes of a wavefront mostly walk the same
//    face for different particles and its vertices are read once per wavefront, not once per lane
//    (2e5-site map: the FP32 face kernel 29.3 -> 13.3 ms).
//  - FP64 face walks are VALU-bound and branch on the particle's distances: they keep the particle order
//    and group the items by vertex count (a stable counting sort), so that the lanes of a wavefront walk
//    faces with the same number of edges for one particle (sorted by face they ran 52.3 -> 58.7 ms).
static_assert(sizeof(FaceItem) == 8 && offsetof(FaceItem, face) == 4, "FaceItem read as a 64-bit sort key");
//
// Vertex-count groups: one wavefront per chunk of BIN_CHUNK items; hist[bin][chunk] counts, and after an
// exclusive scan in (bin, chunk) order gives each chunk's first slot per bin.
constexpr int NV_BINS = 14;  // 3 .. 15 vertices, and 16 or more
constexpr int BIN_CHUNK = 512;

__device__ __forceinline__ int nv_bin(const int64_t* __restrict__ fvp, int f) {
    const int64_t nv = fvp[f + 1] - fvp[f];
    return nv < 3 ? 0 : (nv > 16 ? NV_BINS - 1 : static_cast<int>(nv) - 3);
}

template <bool SCATTER>
__global__ __launch_bounds__(THREADS) void k_vor_item_bins(int64_t nitems, int64_t nchunks,
                                                           const FaceItem* __restrict__ items,
                                                           const int64_t* __restrict__ fvp,
                                                           int64_t* __restrict__ hist, FaceItem* __restrict__ out) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int64_t chunk = (static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x) >> 6;
    if (chunk >= nchunks) return;
    const uint64_t below = lanes_below(lane);
    // per-bin running position (counts in the first pass), one bin per lane 0..NV_BINS-1
    int64_t run = SCATTER && lane < NV_BINS ? hist[static_cast<int64_t>(lane) * nchunks + chunk] : 0;
    const int64_t t0 = chunk * BIN_CHUNK;
    for (int r = 0; r < BIN_CHUNK / 64; ++r) {
        const int64_t t = t0 + r * 64 + lane;
        const bool ok = t < nitems;
        FaceItem it{0, 0};
        int b = -1;
        if (ok) {
            it = items[t];
            b = nv_bin(fvp, it.face);
        }
        int64_t pos = 0;
        for (int q = 0; q < NV_BINS; ++q) {
            const uint64_t m = __ballot(b == q);
            const int64_t base = __shfl(run, q);
            if (b == q) pos = base + __popcll(m & below);
            if (lane == q) run += __popcll(m);
        }
        if (SCATTER && ok) out[pos] = it;
    }
    if (!SCATTER && lane < NV_BINS) hist[static_cast<int64_t>(lane) * nchunks + chunk] = run;
}

// ------------------------------------------------------------------------------------------------
// 6. face integrals
// ------------------------------------------------------------------------------------------------
// SINGLE: the face walk runs in single precision on vertex coordinates relative to the particle
struct EvFactoryM4 {
    static constexpr bool SINGLE = false;
    double sup;
    __device__ EvM4 make(double h) const { return EvM4{1.0 / h}; }
};
struct EvFactoryM4F {
    static constexpr bool SINGLE = true;
    double sup;
    __device__ EvM4T<float> make(double h) const { return EvM4T<float>{static_cast<float>(1.0 / h)}; }
};
struct EvFactoryGen {
    static constexpr bool SINGLE = false;
    double sup;
    const GenTab* tab;
    __device__ EvGen<GenTab> make(double h) const { return EvGen<GenTab>{tab, 1.0 / (sup * h)}; }
};
struct EvFactoryTab {
    static constexpr bool SINGLE = false;
    double sup;
    const double* table;
    __device__ EvTab make(double h) const { return EvTab{table, 1.0 / h}; }
};

template <class F>
__global__ __launch_bounds__(THREADS) void k_vor_faces(int64_t nitems, const FaceItem* __restrict__ items,
                                                       const uint64_t* __restrict__ keys, int cbits,
                                                       const int64_t* __restrict__ slot_off,
                                                       const double* __restrict__ px, const double* __restrict__ py,
                                                       const double* __restrict__ pz, const double* __restrict__ ph,
                                                       F fac, const int64_t* __restrict__ cfp,
                                                       const int64_t* __restrict__ fvp,
                                                       const double* __restrict__ vxyz, double* __restrict__ val) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= nitems) return;
    const FaceItem it = items[t];
    const uint64_t key = keys[it.pair];
    const int64_t i = static_cast<int64_t>(key >> cbits);
    const int64_t c = static_cast<int64_t>(key & ((1ull << cbits) - 1));
    const double h = ph[i];
    auto ev = fac.make(h);
    const int64_t v0 = fvp[it.face];
    const int nv = static_cast<int>(fvp[it.face + 1] - v0);
    bool r;
    double m;
    if constexpr (F::SINGLE) {
        const RelVerts<float> vs{vxyz + 3 * v0, px[i], py[i], pz[i]};
        m = static_cast<double>(face_walk_t<float>(0.0f, 0.0f, 0.0f, static_cast<float>(fac.sup * h), vs, nv, ev, &r));
    } else {
        m = face_walk_integral(px[i], py[i], pz[i], fac.sup * h, vxyz + 3 * v0, nv, ev, &r);
    }
    val[slot_off[it.pair] + (it.face - cfp[c])] = m;
}

// ------------------------------------------------------------------------------------------------
// 7. per-pair cell integrals
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void k_vor_pairsum(int64_t nvis, const uint64_t* __restrict__ keys, int cbits,
                                                         int pbits, const int64_t* __restrict__ slot_off,
                                                         const int64_t* __restrict__ src,
                                                         const double* __restrict__ val, double* __restrict__ ms,
                                                         uint64_t* __restrict__ ckeys, int* __restrict__ cidx,
                                                         MapDevStats* st, int clamp_pairs) {
    // 16 lanes per pair: lane r adds slots r, r + 16, ... of the pair, then the quarter wavefront adds its
    // lanes in a fixed butterfly order (deterministic)
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    const int64_t q = gid / PAIR_LANES;
    const int r = static_cast<int>(gid % PAIR_LANES);
    if (q >= nvis) return;  // whole quarter wavefronts leave together
    double m = 0.0;
    for (int64_t s = slot_off[q] + r; s < slot_off[q + 1]; s += PAIR_LANES) {
        const int64_t k = src[s];
        m += (k >= 0) ? val[k] : -val[~k];
    }
    for (int msk = 1; msk < PAIR_LANES; msk <<= 1) m += __shfl_xor(m, msk, PAIR_LANES);
    if (r != 0) return;
    const uint64_t key = keys[q];
    const uint64_t i = key >> cbits, c = key & ((1ull << cbits) - 1);
    // src/mapping.cpp:198-203; mixed precision clamps cell totals instead (k_vor_accum): a per-pair clamp
    // of single-precision rounding noise around 0 would bias the total mass upward
    if (clamp_pairs && m < 0.0) {
        atomicAdd(&st->n_negative, 1ull);
        if (m <= -0.01) {
            atomicAdd(&st->n_negative_warn, 1ull);
            const unsigned long long w = atomicAdd(&st->n_warn, 1ull);
            if (w < KIWI_MAX_WARN_RECORDS) {
                st->warn[w].particle = static_cast<int64_t>(i);
                st->warn[w].cell = static_cast<int64_t>(c);
                st->warn[w].value = m;
            }
        }
        m = 0.0;
    }
    ms[q] = m;
    ckeys[q] = (c << pbits) | i;
    cidx[q] = static_cast<int>(q);
}

// ------------------------------------------------------------------------------------------------
// 8. per-cell accumulation in particle order
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void k_vor_cell_ranges(int64_t nvis, const uint64_t* __restrict__ ckeys,
                                                             int pbits, int* __restrict__ beg,
                                                             int* __restrict__ end) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (t >= nvis) return;
    const uint64_t c = ckeys[t] >> pbits;
    if (t == 0 || (ckeys[t - 1] >> pbits) != c) beg[c] = static_cast<int>(t);
    if (t == nvis - 1 || (ckeys[t + 1] >> pbits) != c) end[c] = static_cast<int>(t + 1);
}

// the inverse-mapping test per particle (src/mapping.cpp:640-667): its visited pairs are contiguous and in
// cell order; cells without forward mass drop out; m_recovered = sum (Ms m / M_c) M_c, coverage = m_rec / m
__global__ __launch_bounds__(THREADS) void k_vor_inverse(int64_t n, const int64_t* __restrict__ vis_off,
                                                         const uint64_t* __restrict__ keys, int cbits,
                                                         const double* __restrict__ ms, const double* __restrict__ pm,
                                                         const double* __restrict__ cell_mass,
                                                         double* __restrict__ rec, double* __restrict__ cov) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i >= n) return;
    const uint64_t cmask = (1ull << cbits) - 1;
    const double m = pm[i];
    double r = 0.0;
    for (int64_t q = vis_off[i]; q < vis_off[i + 1]; ++q) {
        const double mc = cell_mass[keys[q] & cmask];
        if (mc > 0.0) r += (ms[q] * m / mc) * mc;
    }
    rec[i] = r;
    cov[i] = (m > 0.0) ? r / m : 0.0;
}

__global__ __launch_bounds__(THREADS) void k_vor_accum(int64_t nc, const uint64_t* __restrict__ ckeys, int pbits,
                                                       const int* __restrict__ cidx, const int* __restrict__ beg,
                                                       const int* __restrict__ end, const double* __restrict__ ms,
                                                       const double* __restrict__ pm, const double* __restrict__ vx,
                                                       const double* __restrict__ vy, const double* __restrict__ vz,
                                                       double* __restrict__ om, double* __restrict__ ox,
                                                       double* __restrict__ oy, double* __restrict__ oz,
                                                       int clamp_totals) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (c >= nc) return;
    const uint64_t pmask = (1ull << pbits) - 1;
    double a = 0.0, b = 0.0, d = 0.0, e = 0.0;
    for (int t = beg[c]; t < end[c]; ++t) {
        const int64_t i = static_cast<int64_t>(ckeys[t] & pmask);
        const double w = ms[cidx[t]] * pm[i];  // src/mapping.cpp:452-455
        a += w;
        b += w * vx[i];
        d += w * vy[i];
        e += w * vz[i];
    }
    if (clamp_totals && a < 0.0) a = b = d = e = 0.0;  // rounding noise of an (almost) empty cell
    om[c] = a;
    ox[c] = b;
    oy[c] = d;
    oz[c] = e;
}

int bits_for(int64_t count) {  // bits to hold 0 .. count-1 (at least 1)
    int b = 1;
    while ((1LL << b) < count) ++b;
    return b;
}

template <class T>
T* at_offset(void* base, size_t& off, int64_t count) {
    off = (off + 255) & ~static_cast<size_t>(255);
    T* p = reinterpret_cast<T*>(static_cast<char*>(base) + off);
    off += sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 0));
    return p;
}

size_t bytes_of(std::initializer_list<size_t> parts) {
    size_t t = 0;
    for (size_t p : parts) t = ((t + 255) & ~static_cast<size_t>(255)) + p;
    return t + 256;
}

hipError_t scan_i64(kiwi_hip_ctx* c, const int64_t* in, int64_t* out, int64_t count, hipStream_t s) {
    size_t tb = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, static_cast<int>(count), s);
    if (e != hipSuccess) return e;
    if ((e = reserve(c->cub_tmp, tb)) != hipSuccess) return e;
    tb = c->cub_tmp.cap;
    return hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tb, in, out, static_cast<int>(count), s);
}

hipError_t sort_keys(kiwi_hip_ctx* c, const uint64_t* in, uint64_t* out, int64_t count, int b0, int b1,
                     hipStream_t s) {
    size_t tb = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, tb, in, out, static_cast<int>(count), b0, b1, s);
    if (e != hipSuccess) return e;
    if ((e = reserve(c->cub_tmp, tb)) != hipSuccess) return e;
    tb = c->cub_tmp.cap;
    return hipcub::DeviceRadixSort::SortKeys(c->cub_tmp.p, tb, in, out, static_cast<int>(count), b0, b1, s);
}

hipError_t sort_pairs(kiwi_hip_ctx* c, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, int64_t count,
                      int b0, int b1, hipStream_t s) {
    size_t tb = 0;
    hipError_t e =
        hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, static_cast<int>(count), b0, b1, s);
    if (e != hipSuccess) return e;
    if ((e = reserve(c->cub_tmp, tb)) != hipSuccess) return e;
    tb = c->cub_tmp.cap;
    return hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, tb, kin, kout, vin, vout, static_cast<int>(count), b0, b1,
                                              s);
}

template <class T>
hipError_t read_back(T* dst, const T* src, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, sizeof(T), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// The map's pipeline; with inv set, the inverse-mapping test's per-particle sums replace the per-cell
// accumulation (and its outputs).
int voronoi_run(kiwi_hip_ctx* c, const kiwi_particles* P, const int* home_cell, const kiwi_voronoi_cells* V,
                const kiwi_kernel* K, int precision, const kiwi_cell_sums* out, const double* cell_mass,
                const kiwi_inverse_rows* inv, kiwi_map_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!P || !V || !K || (!out && !inv)) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (P->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative particle count");
    if (P->n > 0 && (!P->x || !P->y || !P->z || !P->vx || !P->vy || !P->vz || !P->m || !P->h || !home_cell))
        return fail(c, KIWI_ERR_INVALID_ARG, "null particle array");
    if (V->n_cells < 0 || V->n_faces < 0 || V->n_face_vertices < 0)
        return fail(c, KIWI_ERR_INVALID_ARG, "negative tessellation size");
    if (!V->cell_face_ptr || (V->n_faces > 0 && (!V->face_vert_ptr || !V->face_neighbor)) ||
        (V->n_face_vertices > 0 && !V->vert_xyz) || (V->n_faces > 0 && V->n_cells == 0))
        return fail(c, KIWI_ERR_INVALID_ARG, "null tessellation array");
    if (out && V->n_cells > 0 && (!out->mass || !out->px || !out->py || !out->pz))
        return fail(c, KIWI_ERR_INVALID_ARG, "null output array");
    if (inv && ((V->n_cells > 0 && !cell_mass) || (P->n > 0 && (!inv->m_recovered || !inv->coverage_fraction))))
        return fail(c, KIWI_ERR_INVALID_ARG, "null cell-mass or per-particle output array");
    for (int a = 0; a < 3; ++a)
        if (!(V->domain[2 * a] < V->domain[2 * a + 1]) || !std::isfinite(V->domain[2 * a]) ||
            !std::isfinite(V->domain[2 * a + 1]))
            return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");
    if (precision != KIWI_FP64 && precision != KIWI_MIXED) return fail(c, KIWI_ERR_INVALID_ARG, "unknown precision");
    KernelChoice kc;
    if (const int rc = resolve_kernel(c, K, &kc)) return rc;
    // single-precision face integrals: a mixed request with the analytic cubic spline
    const bool single = precision == KIWI_MIXED && !kc.generic && !kc.tabulated;
    // per-pair clamp (the reference's compute_cell_integral) in FP64; in single precision a per-pair
    // clamp of rounding noise around 0 would bias the sums upward, so the map clamps cell totals instead
    // and the inverse test keeps the noise in the particle's sum
    const bool clamp_pairs = !single;
    if (P->n > 0x7fffffffLL || V->n_cells > 0x7fffffffLL || V->n_faces > 0x7fffffffLL)
        return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 particles, cells or faces in one call");

    KIWI_HIP_CHECK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = P->n, nc = V->n_cells, nf = V->n_faces, nv = V->n_face_vertices;

    // inputs on the device
    const double* dp[8];
    const double* hp[8] = {P->x, P->y, P->z, P->vx, P->vy, P->vz, P->m, P->h};
    const int* home = home_cell;
    if (P->on_device) {
        for (int k = 0; k < 8; ++k) dp[k] = hp[k];
    } else {
        KIWI_HIP_CHECK(c, reserve(c->parts, 8 * sizeof(double) * std::max<int64_t>(n, 1)));
        for (int k = 0; k < 8; ++k) {
            double* d = static_cast<double*>(c->parts.p) + k * n;
            if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(d, hp[k], n * sizeof(double), hipMemcpyHostToDevice, s));
            dp[k] = d;
        }
        KIWI_HIP_CHECK(c, reserve(c->vor[1], sizeof(int) * std::max<int64_t>(n, 1)));
        if (n) KIWI_HIP_CHECK(c, hipMemcpyAsync(c->vor[1].p, home_cell, n * sizeof(int), hipMemcpyHostToDevice, s));
        home = static_cast<const int*>(c->vor[1].p);
    }
    const int64_t* cfp = V->cell_face_ptr;
    const int64_t* fvp = V->face_vert_ptr;
    const double* vxyz = V->vert_xyz;
    const int* fnb = V->face_neighbor;
    if (!V->on_device) {
        const size_t need = bytes_of({sizeof(int64_t) * (nc + 1), sizeof(int64_t) * (nf + 1),
                                      3 * sizeof(double) * nv, sizeof(int) * nf});
        KIWI_HIP_CHECK(c, reserve(c->vor[0], need));
        size_t o = 0;
        int64_t* d_cfp = at_offset<int64_t>(c->vor[0].p, o, nc + 1);
        int64_t* d_fvp = at_offset<int64_t>(c->vor[0].p, o, nf + 1);
        double* d_v = at_offset<double>(c->vor[0].p, o, 3 * nv);
        int* d_nb = at_offset<int>(c->vor[0].p, o, nf);
        KIWI_HIP_CHECK(c, hipMemcpyAsync(d_cfp, cfp, sizeof(int64_t) * (nc + 1), hipMemcpyHostToDevice, s));
        if (nf > 0) {
            KIWI_HIP_CHECK(c, hipMemcpyAsync(d_fvp, fvp, sizeof(int64_t) * (nf + 1), hipMemcpyHostToDevice, s));
            KIWI_HIP_CHECK(c, hipMemcpyAsync(d_nb, fnb, sizeof(int) * nf, hipMemcpyHostToDevice, s));
        }
        if (nv > 0) KIWI_HIP_CHECK(c, hipMemcpyAsync(d_v, vxyz, 3 * sizeof(double) * nv, hipMemcpyHostToDevice, s));
        cfp = d_cfp;
        fvp = d_fvp;
        vxyz = d_v;
        fnb = d_nb;
    }
    double* om[4] = {nullptr, nullptr, nullptr, nullptr};
    const double* cmass = cell_mass;  // inverse test: forward cell masses on the device
    double *rec = nullptr, *cov = nullptr;
    if (out && out->on_device) {
        om[0] = out->mass;
        om[1] = out->px;
        om[2] = out->py;
        om[3] = out->pz;
    } else if (out) {
        KIWI_HIP_CHECK(c, reserve(c->outbuf, 4 * sizeof(double) * std::max<int64_t>(nc, 1)));
        for (int k = 0; k < 4; ++k) om[k] = static_cast<double*>(c->outbuf.p) + k * nc;
    } else if (inv->on_device) {
        rec = inv->m_recovered;
        cov = inv->coverage_fraction;
    } else {
        KIWI_HIP_CHECK(c, reserve(c->outbuf, sizeof(double) * (2 * std::max<int64_t>(n, 1) + std::max<int64_t>(nc, 1))));
        rec = static_cast<double*>(c->outbuf.p);
        cov = rec + std::max<int64_t>(n, 1);
        double* dcm = cov + std::max<int64_t>(n, 1);
        if (nc) KIWI_HIP_CHECK(c, hipMemcpyAsync(dcm, cell_mass, nc * sizeof(double), hipMemcpyHostToDevice, s));
        cmass = dcm;
    }

    KIWI_HIP_CHECK(c, reserve(c->stats, sizeof(MapDevStats)));
    MapDevStats* dst = static_cast<MapDevStats*>(c->stats.p);
    KIWI_HIP_CHECK(c, hipMemsetAsync(dst, 0, sizeof(MapDevStats), s));
    KIWI_HIP_CHECK(c, reserve(c->vsum, sizeof(double) * (3 * VSUM_BLOCKS + 3)));
    double* vpart = static_cast<double*>(c->vsum.p);
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));

    // 1. geometry
    KIWI_HIP_CHECK(c, reserve(c->vor[2], bytes_of({6 * sizeof(double) * nc, sizeof(int) * nf, sizeof(int) * nf,
                                                   sizeof(FaceGeo) * nf})));
    size_t o2 = 0;
    double* box = at_offset<double>(c->vor[2].p, o2, 6 * nc);
    int* face_cell = at_offset<int>(c->vor[2].p, o2, nf);
    int* twin = at_offset<int>(c->vor[2].p, o2, nf);
    FaceGeo* fgeo = at_offset<FaceGeo>(c->vor[2].p, o2, nf);
    if (nc > 0) k_vor_cells<<<grid1d(nc), THREADS, 0, s>>>(nc, nf, nv, cfp, fvp, vxyz, fnb, box, face_cell, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    {
        unsigned long long bad = 0;
        KIWI_HIP_CHECK(c, read_back(&bad, &dst->bad_geometry, s));
        if (bad) return fail(c, KIWI_ERR_INVALID_ARG, "malformed tessellation (face or vertex ranges, neighbours)");
    }
    if (nf > 0) {
        k_vor_twins<<<grid1d(nf), THREADS, 0, s>>>(nf, cfp, fnb, face_cell, twin);
        k_vor_face_geo<<<grid1d(nf), THREADS, 0, s>>>(nf, fvp, vxyz, fgeo);
    }
    KIWI_HIP_CHECK(c, hipGetLastError());

    // 2. bucket grid over the domain
    BucketGrid B;
    {
        const double L[3] = {V->domain[1] - V->domain[0], V->domain[3] - V->domain[2], V->domain[5] - V->domain[4]};
        const double edge = std::cbrt(L[0] * L[1] * L[2] * BUCKET_CELLS / static_cast<double>(std::max<int64_t>(nc, 1)));
        int g[3];
        for (int a = 0; a < 3; ++a)
            g[a] = static_cast<int>(std::min<double>(MAX_BUCKET_AXIS, std::max(1.0, std::ceil(L[a] / edge))));
        B = BucketGrid{V->domain[0], V->domain[2], V->domain[4], g[0] / L[0], g[1] / L[1], g[2] / L[2],
                       g[0], g[1], g[2]};
    }
    const int64_t nbuckets = static_cast<int64_t>(B.g0) * B.g1 * B.g2;
    const int bbits = bits_for(nbuckets);
    if (nbuckets > 0x7ffffffeLL) return fail(c, KIWI_ERR_LIMIT, "bucket grid too large");
    KIWI_HIP_CHECK(c, reserve(c->vor[3], bytes_of({sizeof(int64_t) * (nc + 1), sizeof(int64_t) * (nc + 1),
                                                   sizeof(int) * (nbuckets + 1), sizeof(CellRec) * nc})));
    size_t o3 = 0;
    int64_t* bcnt = at_offset<int64_t>(c->vor[3].p, o3, nc + 1);
    int64_t* boff = at_offset<int64_t>(c->vor[3].p, o3, nc + 1);
    int* bucket_off = at_offset<int>(c->vor[3].p, o3, nbuckets + 1);
    CellRec* crec = at_offset<CellRec>(c->vor[3].p, o3, nc);
    k_vor_bucket_count<<<grid1d(nc + 1), THREADS, 0, s>>>(nc, B, box, bcnt, crec);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, scan_i64(c, bcnt, boff, nc + 1, s));
    int64_t nentries = 0;
    KIWI_HIP_CHECK(c, read_back(&nentries, boff + nc, s));
    if (nentries > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 cell-bucket entries");
    KIWI_HIP_CHECK(c, reserve(c->vor[4], bytes_of({sizeof(uint64_t) * nentries, sizeof(uint64_t) * nentries})));
    size_t o4 = 0;
    uint64_t* bkeys_in = at_offset<uint64_t>(c->vor[4].p, o4, nentries);
    uint64_t* bkeys = at_offset<uint64_t>(c->vor[4].p, o4, nentries);
    k_fill_i32<<<grid1d(nbuckets + 1), THREADS, 0, s>>>(bucket_off, nbuckets + 1, 0);
    if (nc > 0) k_vor_bucket_emit<<<grid1d(nc), THREADS, 0, s>>>(nc, B, box, boff, bkeys_in);
    KIWI_HIP_CHECK(c, hipGetLastError());
    if (nentries > 0) {
        KIWI_HIP_CHECK(c, sort_keys(c, bkeys_in, bkeys, nentries, 32, 32 + bbits, s));
        k_vor_bucket_offsets<<<grid1d(nentries), THREADS, 0, s>>>(nentries, nbuckets, bkeys, bucket_off);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }

    // 3. candidates
    const double sup = kc.support_in_h;
    KIWI_HIP_CHECK(c, reserve(c->vor[5], bytes_of({sizeof(int64_t) * (n + 1), sizeof(int64_t) * (n + 1), sizeof(int64_t) * (n + 1)})));
    size_t o5 = 0;
    int64_t* ccnt = at_offset<int64_t>(c->vor[5].p, o5, n + 1);
    int64_t* coff = at_offset<int64_t>(c->vor[5].p, o5, n + 1);
    int64_t* vis_off = at_offset<int64_t>(c->vor[5].p, o5, n + 1);
    k_vor_candidates<false><<<grid1d(64 * (n + 1)), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], home, nc, sup, B,
                                                                      crec, bkeys, bucket_off, ccnt, nullptr,
                                                                      nullptr, nullptr, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, scan_i64(c, ccnt, coff, n + 1, s));
    int64_t ncand = 0;
    unsigned long long bad[2] = {0, 0};
    KIWI_HIP_CHECK(c, hipMemcpyAsync(bad, &dst->bad_particle, sizeof(bad), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, read_back(&ncand, coff + n, s));
    if (bad[0]) return fail(c, KIWI_ERR_INVALID_ARG, "non-finite smoothing length");
    if (bad[1]) return fail(c, KIWI_ERR_INVALID_ARG, "home cell index out of range");
    if (ncand > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 candidate (particle, cell) pairs");
    KIWI_HIP_CHECK(c, reserve(c->vor[6], bytes_of({sizeof(int) * ncand, sizeof(int) * ncand,
                                                   sizeof(int64_t) * (ncand + 1), sizeof(int64_t) * (ncand + 1)})));
    size_t o6 = 0;
    int* cand_cell = at_offset<int>(c->vor[6].p, o6, ncand);
    int* cand_part = at_offset<int>(c->vor[6].p, o6, ncand);
    int64_t* flag = at_offset<int64_t>(c->vor[6].p, o6, ncand + 1);
    int64_t* flag_off = at_offset<int64_t>(c->vor[6].p, o6, ncand + 1);
    if (n > 0)
        k_vor_candidates<true><<<grid1d(64 * n), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], home, nc, sup, B, crec,
                                                                  bkeys, bucket_off, nullptr, coff, cand_cell,
                                                                  cand_part, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    {
        unsigned long long mism = 0;
        KIWI_HIP_CHECK(c, read_back(&mism, &dst->fill_mismatch, s));
        if (mism) return fail(c, KIWI_ERR_DEVICE, "internal error: candidate count and fill passes disagree");
    }

    // 4. visits: cheap tests per pair, then the exact tests they leave open, one lane each
    KIWI_HIP_CHECK(c, reserve(c->vor[11], bytes_of({sizeof(int64_t) * (ncand + 1), sizeof(int64_t) * (ncand + 1),
                                                    sizeof(uint64_t) * (ncand + 1)})));
    size_t o11 = 0;
    int64_t* nexact = at_offset<int64_t>(c->vor[11].p, o11, ncand + 1);
    int64_t* ex_off = at_offset<int64_t>(c->vor[11].p, o11, ncand + 1);
    uint64_t* exmask = at_offset<uint64_t>(c->vor[11].p, o11, ncand + 1);
    k_vor_reach<<<grid1d(ncand + 1), THREADS, 0, s>>>(ncand, cand_cell, cand_part, dp[0], dp[1], dp[2], dp[7], home,
                                                      sup, cfp, fvp, vxyz, fgeo, flag, nexact, exmask);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, scan_i64(c, nexact, ex_off, ncand + 1, s));
    int64_t nex = 0;
    KIWI_HIP_CHECK(c, read_back(&nex, ex_off + ncand, s));
    if (nex > 0) {
        KIWI_HIP_CHECK(c, reserve(c->vor[12], sizeof(uint64_t) * nex));
        uint64_t* ex = static_cast<uint64_t*>(c->vor[12].p);
        k_vor_reach_list<<<grid1d(ncand), THREADS, 0, s>>>(ncand, cand_cell, cand_part, dp[0], dp[1], dp[2], dp[7],
                                                           sup, cfp, fvp, vxyz, fgeo, nexact, exmask, ex_off, ex);
        k_vor_reach_exact<<<grid1d(nex), THREADS, 0, s>>>(nex, ex, cand_cell, cand_part, dp[0], dp[1], dp[2], dp[7],
                                                          sup, cfp, fvp, vxyz, flag);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    KIWI_HIP_CHECK(c, scan_i64(c, flag, flag_off, ncand + 1, s));
    k_vor_particle_visits<<<grid1d(n + 1), THREADS, 0, s>>>(n, coff, flag_off, vis_off, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    int64_t nvis = 0;
    KIWI_HIP_CHECK(c, read_back(&nvis, flag_off + ncand, s));
    const int cbits = bits_for(nc), pbits = bits_for(n);
    KIWI_HIP_CHECK(c, reserve(c->vor[7], bytes_of({sizeof(uint64_t) * nvis, sizeof(uint64_t) * nvis, sizeof(int64_t) * (nvis + 1),
                                                   sizeof(int64_t) * (nvis + 1), sizeof(int64_t) * (nvis + 1),
                                                   sizeof(int64_t) * (nvis + 1)})));
    size_t o7 = 0;
    uint64_t* vkeys_in = at_offset<uint64_t>(c->vor[7].p, o7, nvis);
    uint64_t* vkeys = at_offset<uint64_t>(c->vor[7].p, o7, nvis);
    int64_t* nface = at_offset<int64_t>(c->vor[7].p, o7, nvis + 1);
    int64_t* slot_off = at_offset<int64_t>(c->vor[7].p, o7, nvis + 1);
    int64_t* nown = at_offset<int64_t>(c->vor[7].p, o7, nvis + 1);
    int64_t* own_off = at_offset<int64_t>(c->vor[7].p, o7, nvis + 1);
    if (ncand > 0) k_vor_compact<<<grid1d(ncand), THREADS, 0, s>>>(ncand, flag, flag_off, cand_cell, cand_part, cbits,
                                                                   vkeys_in);
    KIWI_HIP_CHECK(c, hipGetLastError());
    if (nvis > 0) KIWI_HIP_CHECK(c, sort_keys(c, vkeys_in, vkeys, nvis, 0, cbits + pbits, s));

    // 5. owned faces
    const uint64_t cmask = (1ull << cbits) - 1;
    k_vor_pair_faces<<<grid1d(nvis + 1), THREADS, 0, s>>>(nvis, vkeys, cmask, cfp, nface);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, scan_i64(c, nface, slot_off, nvis + 1, s));
    int64_t nslots = 0;
    KIWI_HIP_CHECK(c, read_back(&nslots, slot_off + nvis, s));
    KIWI_HIP_CHECK(c, reserve(c->vor[8], bytes_of({sizeof(int64_t) * nslots, sizeof(double) * nslots})));
    size_t o8 = 0;
    int64_t* src = at_offset<int64_t>(c->vor[8].p, o8, nslots);
    double* val = at_offset<double>(c->vor[8].p, o8, nslots);
    k_vor_owned<<<grid1d(PAIR_LANES * (nvis + 1)), THREADS, 0, s>>>(nvis, vkeys, cbits, vis_off, slot_off, cfp, fnb,
                                                                     twin, src, nown);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, scan_i64(c, nown, own_off, nvis + 1, s));
    int64_t nitems = 0;
    KIWI_HIP_CHECK(c, read_back(&nitems, own_off + nvis, s));
    if (nitems > 0x7fffffffLL) return fail(c, KIWI_ERR_LIMIT, "more than 2^31-1 face integrals in one call");
    const int64_t nchunks = (nitems + BIN_CHUNK - 1) / BIN_CHUNK;
    KIWI_HIP_CHECK(c, reserve(c->vor[9], bytes_of({sizeof(FaceItem) * nitems, sizeof(FaceItem) * nitems,
                                                   sizeof(int64_t) * (NV_BINS * nchunks + 1),
                                                   sizeof(int64_t) * (NV_BINS * nchunks + 1)})));
    size_t o9 = 0;
    FaceItem* items_pair = at_offset<FaceItem>(c->vor[9].p, o9, nitems);
    FaceItem* items = at_offset<FaceItem>(c->vor[9].p, o9, nitems);
    int64_t* bin_cnt = at_offset<int64_t>(c->vor[9].p, o9, NV_BINS * nchunks + 1);
    int64_t* bin_off = at_offset<int64_t>(c->vor[9].p, o9, NV_BINS * nchunks + 1);
    if (nvis > 0)
        k_vor_items<<<grid1d(PAIR_LANES * nvis), THREADS, 0, s>>>(nvis, vkeys, cmask, slot_off, own_off, cfp, src,
                                                                  items_pair);
    KIWI_HIP_CHECK(c, hipGetLastError());
    if (nitems > 0 && single) {
        KIWI_HIP_CHECK(c, sort_keys(c, reinterpret_cast<const uint64_t*>(items_pair), reinterpret_cast<uint64_t*>(items),
                                    nitems, 32, 32 + bits_for(nf), s));
    } else if (nitems > 0) {
        k_fill_i64<<<1, 64, 0, s>>>(bin_cnt + NV_BINS * nchunks, 1, 0);
        k_vor_item_bins<false><<<grid1d(64 * nchunks), THREADS, 0, s>>>(nitems, nchunks, items_pair, fvp, bin_cnt,
                                                                       nullptr);
        KIWI_HIP_CHECK(c, hipGetLastError());
        KIWI_HIP_CHECK(c, scan_i64(c, bin_cnt, bin_off, NV_BINS * nchunks + 1, s));
        k_vor_item_bins<true><<<grid1d(64 * nchunks), THREADS, 0, s>>>(nitems, nchunks, items_pair, fvp, bin_off,
                                                                      items);
    }
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));

    // 6. face integrals
    if (nitems > 0) {
        if (kc.tabulated) {
            KIWI_HIP_CHECK(c, kiwi_build_m4_table(c, s));
            const EvFactoryTab fac{sup, static_cast<const double*>(c->m4table.p)};
            k_vor_faces<<<grid1d(nitems), THREADS, 0, s>>>(nitems, items, vkeys, cbits, slot_off, dp[0], dp[1], dp[2],
                                                           dp[7], fac, cfp, fvp, vxyz, val);
        } else if (kc.generic) {
            const EvFactoryGen fac{sup, static_cast<const GenTab*>(c->tables.p) + K->kernel_id};
            k_vor_faces<<<grid1d(nitems), THREADS, 0, s>>>(nitems, items, vkeys, cbits, slot_off, dp[0], dp[1], dp[2],
                                                           dp[7], fac, cfp, fvp, vxyz, val);
        } else if (single) {
            const EvFactoryM4F fac{2.0};
            k_vor_faces<<<grid1d(nitems), THREADS, 0, s>>>(nitems, items, vkeys, cbits, slot_off, dp[0], dp[1], dp[2],
                                                           dp[7], fac, cfp, fvp, vxyz, val);
        } else {
            const EvFactoryM4 fac{2.0};
            k_vor_faces<<<grid1d(nitems), THREADS, 0, s>>>(nitems, items, vkeys, cbits, slot_off, dp[0], dp[1], dp[2],
                                                           dp[7], fac, cfp, fvp, vxyz, val);
        }
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));

    // 7-8. per-pair sums, then per-cell accumulation
    KIWI_HIP_CHECK(c, reserve(c->vor[10], bytes_of({sizeof(double) * nvis, sizeof(uint64_t) * nvis, sizeof(uint64_t) * nvis,
                                                    sizeof(int) * nvis, sizeof(int) * nvis, 2 * sizeof(int) * nc})));
    size_t o10 = 0;
    double* ms = at_offset<double>(c->vor[10].p, o10, nvis);
    uint64_t* ckeys_in = at_offset<uint64_t>(c->vor[10].p, o10, nvis);
    uint64_t* ckeys = at_offset<uint64_t>(c->vor[10].p, o10, nvis);
    int* cidx_in = at_offset<int>(c->vor[10].p, o10, nvis);
    int* cidx = at_offset<int>(c->vor[10].p, o10, nvis);
    int* cbeg = at_offset<int>(c->vor[10].p, o10, 2 * nc);
    int* cend = cbeg + nc;
    if (nvis > 0) {
        k_vor_pairsum<<<grid1d(PAIR_LANES * nvis), THREADS, 0, s>>>(nvis, vkeys, cbits, pbits, slot_off, src, val, ms, ckeys_in,
                                                       cidx_in, dst, clamp_pairs ? 1 : 0);
        KIWI_HIP_CHECK(c, hipGetLastError());
        if (!inv) KIWI_HIP_CHECK(c, sort_pairs(c, ckeys_in, ckeys, cidx_in, cidx, nvis, 0, cbits + pbits, s));
    }
    if (inv) {
        if (n > 0) k_vor_inverse<<<grid1d(n), THREADS, 0, s>>>(n, vis_off, vkeys, cbits, ms, dp[6], cmass, rec, cov);
        KIWI_HIP_CHECK(c, hipGetLastError());
    } else if (nc > 0) {
        k_fill_i32<<<grid1d(2 * nc), THREADS, 0, s>>>(cbeg, 2 * nc, 0);
        if (nvis > 0) k_vor_cell_ranges<<<grid1d(nvis), THREADS, 0, s>>>(nvis, ckeys, pbits, cbeg, cend);
        k_vor_accum<<<grid1d(nc), THREADS, 0, s>>>(nc, ckeys, pbits, cidx, cbeg, cend, ms, dp[6], dp[3], dp[4], dp[5],
                                                   om[0], om[1], om[2], om[3], single ? 1 : 0);
        KIWI_HIP_CHECK(c, hipGetLastError());
    }
    k_vsum_partial<<<VSUM_BLOCKS, THREADS, 0, s>>>(n, dp[3], dp[4], dp[5], vpart);
    k_vsum_final<<<1, 64, 0, s>>>(vpart, vpart + 3 * VSUM_BLOCKS);
    KIWI_HIP_CHECK(c, hipGetLastError());
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[3], s));

    if (out && !out->on_device && nc > 0) {
        double* ho[4] = {out->mass, out->px, out->py, out->pz};
        for (int k = 0; k < 4; ++k)
            KIWI_HIP_CHECK(c, hipMemcpyAsync(ho[k], om[k], nc * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (inv && !inv->on_device && n > 0) {
        KIWI_HIP_CHECK(c, hipMemcpyAsync(inv->m_recovered, rec, n * sizeof(double), hipMemcpyDeviceToHost, s));
        KIWI_HIP_CHECK(c, hipMemcpyAsync(inv->coverage_fraction, cov, n * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    MapDevStats hs;
    double vs[3];
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&hs, dst, sizeof(MapDevStats), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipMemcpyAsync(vs, vpart + 3 * VSUM_BLOCKS, sizeof(vs), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->unresolved = static_cast<int64_t>(hs.unresolved);
        stats->unmapped = static_cast<int64_t>(hs.unmapped);
        stats->n_negative = static_cast<int64_t>(hs.n_negative);
        stats->n_negative_warn = static_cast<int64_t>(hs.n_negative_warn);
        stats->pairs_candidate = ncand;
        stats->pairs_visited = nvis;
        stats->face_evals = nitems;
        for (int k = 0; k < 3; ++k) stats->v_mean[k] = n > 0 ? vs[k] / static_cast<double>(n) : 0.0;
        float ms_prep = 0, ms_faces = 0, ms_accum = 0;
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_faces, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&ms_accum, c->ev[2], c->ev[3]);
        stats->device_ms_prep = ms_prep;
        stats->device_ms_faces = ms_faces;
        stats->device_ms_accum = ms_accum;
        const int nw = static_cast<int>(std::min<unsigned long long>(hs.n_warn, KIWI_MAX_WARN_RECORDS));
        stats->n_warn_records = nw;
        for (int k = 0; k < nw; ++k) stats->warn[k] = hs.warn[k];
    }
    return KIWI_OK;
}

}  // namespace

extern "C" {

int kiwi_hip_voronoi_map(kiwi_hip_ctx* c, const kiwi_particles* P, const int* home_cell,
                         const kiwi_voronoi_cells* V, const kiwi_kernel* K, int precision, const kiwi_cell_sums* out,
                         kiwi_map_stats* stats) {
    if (c && !out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    return voronoi_run(c, P, home_cell, V, K, precision, out, nullptr, nullptr, stats);
}

int kiwi_hip_inverse_mapping(kiwi_hip_ctx* c, const kiwi_particles* P, const int* home_cell,
                             const kiwi_voronoi_cells* V, const kiwi_kernel* K, int precision, const double* cell_mass,
                             const kiwi_inverse_rows* rows, kiwi_map_stats* stats) {
    if (c && !rows) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    return voronoi_run(c, P, home_cell, V, K, precision, nullptr, cell_mass, rows, stats);
}

}  // extern "C"
