// libkiwi_hip -- the adjoint of the Cartesian sph-to-grid for the analytic cubic spline and the generic
// kernels: given dL/d(cell mass, px, py, pz), the gradients dL/d(m, x, y, z, h, vx, vy, vz) per particle.  The reference
// has no adjoint; this differentiates the forward of kiwi_hip.hip (the lattice restatement of
// src/cartesian_grid.cpp:239-279), recomputing the kernel integrals instead of storing per-pair terms.
//
// With mass(c) = sum_i m_i Ms_i(c) and p(c) = sum_i m_i v_i Ms_i(c), and Ms_i(c) = sum over the cell's
// 8 corners e * U((corner - x_i) / h_i) (U = s W, m4_lattice.h):
//   dL/dm_i   = sum_c Ms_i(c) w_i(c),         w_i(c) = dL/dmass(c) + v_i . dL/dp(c)
//   dL/dv_i   = m_i sum_c Ms_i(c) dL/dp(c)
//   dL/dx_i   = -(m_i / h_i) sum_c w_i(c) sum_corners e dU/da
//   dL/dh_i   = -(m_i / h_i) sum_c w_i(c) sum_corners e (a dU/da + b dU/db + c dU/dc)
// where (a, b, c) are a corner's offsets in units of h (for a generic kernel, offsets are in units of the
// support H = s h and both chain factors carry 1/H instead of 1/h).  dU/da comes from m4_adjoint.h or
// gen_adjoint.h (dual numbers through the forward's closed form).  Cells beyond the support are exact zeros in the forward and
// are skipped here; in FP64 a pair the forward clamps at 0 (Ms < 0) has zero derivative.  Mixed mode
// clamps only negative cell totals, rounding noise of empty cells, and the adjoint ignores that clamp.
//
//   1. k_adj_prep       per particle: the forward's bounding box, line-table size, work estimate
//   2. k_adj_lines      one wavefront per particle: the edge-line constants with their derivatives
//   3. order            particles by work, largest first (hipcub radix sort)
//   4. k_adj_particles  one wavefront per particle: per 8^3 cell block, the lattice values and
//                       gradients into LDS, then the cell pass; register accumulators, a fixed-order
//                       wavefront reduction, one write per gradient.  Deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "common.h"
#include "gen_adjoint.h"
#include "m4_adjoint.h"

using namespace kiwi_hip;

namespace {

constexpr int ADJ_PTS = (BLK + 1) * (BLK + 1) * (BLK + 1);

struct AdjDevStats {
    unsigned long long skipped_outside, pairs_bbox, points, cells, active;
};

// The kernel path of the adjoint: the analytic cubic spline (offsets in units of h, support 2) or a generic
// kernel (offsets in units of the support H, support 1); the line-table entry; a lattice value with
// its gradient.
template <typename T>
struct AdjM4 {
    using Real = T;
    using Line = AdjLine<T>;
    struct Ctx {
        double support_in_h;
    };
    static __device__ __forceinline__ double unit_in_h(const Ctx&) { return 1.0; }
    static __device__ __forceinline__ T support2() { return T(4); }
    static __device__ __forceinline__ Line line(const Ctx&, T da, T db) { return m4_adj_line<T>(da, db); }
    static __device__ __forceinline__ void point(const Ctx&, T ax, T ay, T az, const Line& lxy, const Line& lxz,
                                                 const Line& lyz, T& W, T& Wx, T& Wy, T& Wz) {
        m4_adj_point<T>(ax, ay, az, lxy, lxz, lyz, W, Wx, Wy, Wz);
    }
};

// KID >= 0: kernel KID's table compiled into the code (gen_fixed_tab), as in the forward's PolicyGen
template <typename T, int KID = -1>
struct AdjGen {
    using Real = T;
    using Line = AdjLineGen<T>;
    struct Ctx {
        double support_in_h;
        const GenTab* tab;
    };
    static __device__ __forceinline__ const GenTab& tab(const Ctx& c) {
        if constexpr (KID >= 0)
            return gen_fixed_tab<KID>();
        else
            return *c.tab;
    }
    static __device__ __forceinline__ double unit_in_h(const Ctx& c) { return c.support_in_h; }
    static __device__ __forceinline__ T support2() { return T(1); }
    static __device__ __forceinline__ Line line(const Ctx& c, T da, T db) { return gen_adj_line<T>(tab(c), da, db); }
    static __device__ __forceinline__ void point(const Ctx& c, T ax, T ay, T az, const Line& lxy, const Line& lxz,
                                                 const Line& lyz, T& W, T& Wx, T& Wy, T& Wz) {
        // the forward's region decision (PolicyGen::region): the radius in T, compared in double
        const int k = gen_region_of(tab(c), static_cast<double>(KMath<T>::sqrt_(ax * ax + ay * ay + az * az)));
        gen_adj_point<T, KID>(tab(c), k, ax, ay, az, lxy, lxz, lyz, W, Wx, Wy, Wz);
    }
};

// Calls f(P{}) with the adjoint's generic policy of kernel kid in precision T (single precision: the
// kernels gen_mixed_ok names).
template <typename T, class F>
hipError_t with_adj_gen(int kid, F&& f) {
    switch (kid) {
        case KIWI_KERNEL_M4:
            return f(AdjGen<T, KIWI_KERNEL_M4>{});
        case KIWI_KERNEL_WC2:
            return f(AdjGen<T, KIWI_KERNEL_WC2>{});
        default:
            break;
    }
    if constexpr (std::is_same<T, double>::value) {
        switch (kid) {
            case KIWI_KERNEL_M5:
                return f(AdjGen<T, KIWI_KERNEL_M5>{});
            case KIWI_KERNEL_M6:
                return f(AdjGen<T, KIWI_KERNEL_M6>{});
            case KIWI_KERNEL_WC4:
                return f(AdjGen<T, KIWI_KERNEL_WC4>{});
            case KIWI_KERNEL_WC6:
                return f(AdjGen<T, KIWI_KERNEL_WC6>{});
            default:
                break;
        }
    }
    return hipErrorInvalidValue;
}

__global__ void k_adj_fill(double* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------------------------------------
// 1. bounding boxes, line-table sizes, work
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void k_adj_prep(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ z,
                                                      const double* __restrict__ h, GridGeom g, double sup,
                                                      int* __restrict__ bbox, int64_t* __restrict__ nlines,
                                                      uint32_t* __restrict__ work, int* __restrict__ order,
                                                      AdjDevStats* st) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x;
    if (i == n) nlines[n] = 0;
    unsigned long long pairs = 0, active = 0, skipped = 0;
    if (i < n) {  // lanes past the particles only join the counter reduction below, with zeros
        int* bb = bbox + 6 * i;
        bb[0] = 1;
        bb[1] = 0;
        int64_t nl = 0;
        uint32_t wk = 0;
        const double hh = h[i];
        if (hh > 0.0) {
            int q[6];
            if (particle_bbox(g, sup, x[i], y[i], z[i], hh, q)) {
                for (int k = 0; k < 6; ++k) bb[k] = q[k];
                const int64_t Nx = q[1] - q[0] + 1, Ny = q[3] - q[2] + 1, Nz = q[5] - q[4] + 1;
                nl = (Nx + 1) * (Ny + 1) + (Nx + 1) * (Nz + 1) + (Ny + 1) * (Nz + 1);
                const int64_t blocks = static_cast<int64_t>(q[1] / BLK - q[0] / BLK + 1) *
                                       (q[3] / BLK - q[2] / BLK + 1) * (q[5] / BLK - q[4] / BLK + 1);
                wk = static_cast<uint32_t>(std::min<int64_t>(blocks * ADJ_PTS + Nx * Ny * Nz, 0xffffffffLL));
                pairs = static_cast<unsigned long long>(Nx * Ny * Nz);
                active = 1;
            } else {
                skipped = 1;
            }
        }
        nlines[i] = nl;
        work[i] = wk;
        order[i] = static_cast<int>(i);
    }
    // one atomic per wavefront and counter
    for (int off = 32; off > 0; off >>= 1) {
        pairs += __shfl_down(pairs, off);
        active += __shfl_down(active, off);
        skipped += __shfl_down(skipped, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (pairs) atomicAdd(&st->pairs_bbox, pairs);
        if (active) atomicAdd(&st->active, active);
        if (skipped) atomicAdd(&st->skipped_outside, skipped);
    }
}

// ------------------------------------------------------------------------------------------------
// 2. edge-line constants with derivatives; table layout as the forward's k_lines: three unordered
//    axis pairs (x,y) (x,z) (y,z), each La x Lb, a-index major
// ------------------------------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(THREADS) void k_adj_lines(int64_t n, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       const double* __restrict__ h, const int* __restrict__ bbox,
                                                       const int64_t* __restrict__ loffs, GridGeom g,
                                                       typename P::Ctx ctx, typename P::Line* __restrict__ table) {
    using T = typename P::Real;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * WAVES;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * WAVES + wave_uniform(threadIdx.x >> 6); p < n; p += nwaves) {
        const int* bb = bbox + 6 * p;
        if (bb[0] > bb[1]) continue;
        const int L0 = bb[1] - bb[0] + 2, L1 = bb[3] - bb[2] + 2, L2 = bb[5] - bb[4] + 2;
        const int c01 = L0 * L1, c02 = L0 * L2, c12 = L1 * L2;
        const int nl = c01 + c02 + c12;
        const double inv_h = 1.0 / (P::unit_in_h(ctx) * h[p]);  // lattice units
        const double p0 = x[p], p1 = y[p], p2 = z[p];
        typename P::Line* out = table + loffs[p];
        for (int t = lane; t < nl; t += 64) {
            int u = t, a = 0, b = 1, Lb = L1;
            if (u >= c01) {
                u -= c01;
                b = 2;
                Lb = L2;
                if (u >= c02) {
                    u -= c02;
                    a = 1;
                }
            }
            const int ia = u / Lb, ib = u - ia * Lb;
            const double da = fabs((gmin(g, a) + (sel3(a, bb[0], bb[2], bb[4]) + ia) * gstep(g, a) -
                                    sel3(a, p0, p1, p2)) * inv_h);
            const double db = fabs((gmin(g, b) + (sel3(b, bb[0], bb[2], bb[4]) + ib) * gstep(g, b) -
                                    sel3(b, p0, p1, p2)) * inv_h);
            out[t] = P::line(ctx, static_cast<T>(da), static_cast<T>(db));
        }
    }
}

// ------------------------------------------------------------------------------------------------
// 4. per particle: lattice values and gradients, then the cell pass
// ------------------------------------------------------------------------------------------------
template <typename T>
struct AdjLds {
    T U[ADJ_PTS], Ux[ADJ_PTS], Uy[ADJ_PTS], Uz[ADJ_PTS], Uh[ADJ_PTS];  // Uh = a Ux + b Uy + c Uz
    T d[3 * (BLK + 1)];   // signed offset of each lattice coordinate (units of h), x..., y..., z...
    T gp[3 * (BLK + 1)];  // squared gap of the cells on either side of the lattice plane
    T cg[3 * BLK];        // squared gap of each cell layer
};

template <class P>
__global__ __launch_bounds__(64) void k_adj_particles(
    const int* __restrict__ order, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vz, const double* __restrict__ m, const double* __restrict__ h,
    const int* __restrict__ bbox, const int64_t* __restrict__ loffs, const typename P::Line* __restrict__ table,
    typename P::Ctx ctx, GridGeom g,
    const double* __restrict__ gm, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, int clamp_pairs, double* __restrict__ o_m, double* __restrict__ o_x,
    double* __restrict__ o_y, double* __restrict__ o_z, double* __restrict__ o_h, double* __restrict__ o_vx,
    double* __restrict__ o_vy, double* __restrict__ o_vz, AdjDevStats* st) {
    using T = typename P::Real;
    using Line = typename P::Line;
    __shared__ AdjLds<T> L;
    const int lane = threadIdx.x;
    const int p = order[blockIdx.x];
    const int* bb = bbox + 6 * static_cast<int64_t>(p);
    const double x0 = x[p], y0 = y[p], z0 = z[p], hh = h[p], mm = m[p];
    const double v0 = vx[p], v1 = vy[p], v2 = vz[p];
    const double inv_h = 1.0 / (P::unit_in_h(ctx) * hh);  // lattice units (h, or the support H)
    const T S2 = P::support2();                            // support^2 in lattice units
    const int Lx = bb[1] - bb[0] + 2, Ly = bb[3] - bb[2] + 2, Lz = bb[5] - bb[4] + 2;
    const Line* lt = table + loffs[p];
    const int exz = Lx * Ly, eyz = exz + Lx * Lz;
    double a_m = 0, a_vx = 0, a_vy = 0, a_vz = 0, a_x = 0, a_y = 0, a_z = 0, a_h = 0;
    unsigned long long n_pts = 0, n_cells = 0;
    for (int bxk = bb[0] / BLK; bxk <= bb[1] / BLK; ++bxk)
        for (int byk = bb[2] / BLK; byk <= bb[3] / BLK; ++byk)
            for (int bzk = bb[4] / BLK; bzk <= bb[5] / BLK; ++bzk) {
                const int cx0 = max(bxk * BLK, bb[0]), cx1 = min(bxk * BLK + BLK - 1, bb[1]);
                const int cy0 = max(byk * BLK, bb[2]), cy1 = min(byk * BLK + BLK - 1, bb[3]);
                const int cz0 = max(bzk * BLK, bb[4]), cz1 = min(bzk * BLK + BLK - 1, bb[5]);
                const int mx = cx1 - cx0 + 1, my = cy1 - cy0 + 1, mz = cz1 - cz0 + 1;  // cells
                const int nx = mx + 1, ny = my + 1, nz = mz + 1;                      // lattice points
                // ---- axis arrays: offsets, gaps of the neighbouring cells, cell-layer gaps ----
                if (lane < nx + ny + nz) {
                    int a = 0, u = lane, n_a = nx, c0 = cx0;
                    if (u >= nx) {
                        u -= nx;
                        a = 1;
                        n_a = ny;
                        c0 = cy0;
                        if (u >= ny) {
                            u -= ny;
                            a = 2;
                            n_a = nz;
                            c0 = cz0;
                        }
                    }
                    const double lo = gmin(g, a), pa = sel3(a, x0, y0, z0), st_a = gstep(g, a);
                    const double d = (lo + (c0 + u) * st_a - pa) * inv_h;
                    const double dm = (lo + (c0 + max(u - 1, 0)) * st_a - pa) * inv_h;
                    const double dp = (lo + (c0 + min(u + 1, n_a - 1)) * st_a - pa) * inv_h;
                    const double dn = (lo + (c0 + u + 1) * st_a - pa) * inv_h;
                    const T gpa = static_cast<T>(gap(dm, dp));
                    L.d[lane] = static_cast<T>(d);
                    L.gp[lane] = gpa * gpa;
                    if (u < n_a - 1) {
                        const T gc = static_cast<T>(gap(d, dn));
                        L.cg[sel3(a, 0, mx, mx + my) + u] = gc * gc;
                    }
                }
                __syncthreads();
                // ---- lattice values and gradients ----
                const int nyz = ny * nz, npts = nx * nyz;
                const int ox = cx0 - bb[0], oy = cy0 - bb[2], oz = cz0 - bb[4];  // particle-lattice offsets
                for (int t = lane; t < npts; t += 64) {
                    const int i = t / nyz, r = t - i * nyz, j = r / nz, k = r - j * nz;
                    const T dx = L.d[i], dy = L.d[nx + j], dz = L.d[nx + ny + k];
                    T U = T(0), Ux = T(0), Uy = T(0), Uz = T(0);
                    if (dx != T(0) && dy != T(0) && dz != T(0) && L.gp[i] + L.gp[nx + j] + L.gp[nx + ny + k] < S2) {
                        ++n_pts;
                        const int gi = ox + i, gj = oy + j, gk = oz + k;
                        const Line lxy = lt[gi * Ly + gj], lxz = lt[exz + gi * Lz + gk], lyz = lt[eyz + gj * Lz + gk];
                        T W, Wx, Wy, Wz;
                        const T ax = dx < T(0) ? -dx : dx, ay = dy < T(0) ? -dy : dy, az = dz < T(0) ? -dz : dz;
                        P::point(ctx, ax, ay, az, lxy, lxz, lyz, W, Wx, Wy, Wz);
                        const T sx = dx < T(0) ? T(-1) : T(1), sy = dy < T(0) ? T(-1) : T(1),
                                sz = dz < T(0) ? T(-1) : T(1);
                        U = sx * sy * sz * W;
                        Ux = sy * sz * Wx;
                        Uy = sx * sz * Wy;
                        Uz = sx * sy * Wz;
                    }
                    L.U[t] = U;
                    L.Ux[t] = Ux;
                    L.Uy[t] = Uy;
                    L.Uz[t] = Uz;
                    L.Uh[t] = dx * Ux + dy * Uy + dz * Uz;
                }
                __syncthreads();
                // ---- cells: the 8-corner stencil of the values and of the gradients ----
                const int mzy = my * mz, ncells = mx * mzy;
                for (int t = lane; t < ncells; t += 64) {
                    const int i = t / mzy, r = t - i * mzy, j = r / mz, k = r - j * mz;
                    if (!(L.cg[i] + L.cg[mx + j] + L.cg[mx + my + k] < S2)) continue;  // exact zero
                    ++n_cells;
                    const int b0 = (i * ny + j) * nz + k;
                    // corners with e = +1: (1,0,0) (0,1,0) (0,0,1) (1,1,1); e = -1: the other four
                    const int cp0 = b0 + nyz, cp1 = b0 + nz, cp2 = b0 + 1, cp3 = b0 + nyz + nz + 1;
                    const int cm0 = b0, cm1 = b0 + nyz + nz, cm2 = b0 + nyz + 1, cm3 = b0 + nz + 1;
                    auto st8 = [&](const T* A) {
                        return (static_cast<double>(A[cp0]) + static_cast<double>(A[cp1]) +
                                static_cast<double>(A[cp2]) + static_cast<double>(A[cp3])) -
                               (static_cast<double>(A[cm0]) + static_cast<double>(A[cm1]) +
                                static_cast<double>(A[cm2]) + static_cast<double>(A[cm3]));
                    };
                    double Ms = st8(L.U);
                    double dX = st8(L.Ux), dY = st8(L.Uy), dZ = st8(L.Uz), dH = st8(L.Uh);
                    if (clamp_pairs && Ms < 0.0) Ms = dX = dY = dZ = dH = 0.0;
                    const int64_t c = (static_cast<int64_t>(cx0 + i) * g.ny + (cy0 + j)) * g.nz + (cz0 + k);
                    const double Gm = gm[c], Gx = gx[c], Gy = gy[c], Gz = gz[c];
                    const double w = Gm + v0 * Gx + v1 * Gy + v2 * Gz;
                    a_m += Ms * w;
                    a_vx += Ms * Gx;
                    a_vy += Ms * Gy;
                    a_vz += Ms * Gz;
                    a_x += w * dX;
                    a_y += w * dY;
                    a_z += w * dZ;
                    a_h += w * dH;
                }
                __syncthreads();
            }
    // fixed-order wavefront reduction
    for (int off = 32; off > 0; off >>= 1) {
        a_m += __shfl_down(a_m, off);
        a_vx += __shfl_down(a_vx, off);
        a_vy += __shfl_down(a_vy, off);
        a_vz += __shfl_down(a_vz, off);
        a_x += __shfl_down(a_x, off);
        a_y += __shfl_down(a_y, off);
        a_z += __shfl_down(a_z, off);
        a_h += __shfl_down(a_h, off);
        n_pts += __shfl_down(n_pts, off);
        n_cells += __shfl_down(n_cells, off);
    }
    if (lane == 0) {
        // an offset is (corner - x) / (u h) in lattice units u h: d/dx = -1 / (u h), d/dh = -offset / h
        const double s = -mm * inv_h, sh = -mm / hh;
        o_m[p] = a_m;
        o_vx[p] = mm * a_vx;
        o_vy[p] = mm * a_vy;
        o_vz[p] = mm * a_vz;
        o_x[p] = s * a_x;
        o_y[p] = s * a_y;
        o_z[p] = s * a_z;
        o_h[p] = sh * a_h;
        atomicAdd(&st->points, n_pts);
        atomicAdd(&st->cells, n_cells);
    }
}

template <class T>
T* carve(void* base, size_t& off, int64_t count) {
    off = (off + 255) & ~static_cast<size_t>(255);
    T* q = reinterpret_cast<T*>(static_cast<char*>(base) + off);
    off += sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 0));
    return q;
}

size_t carve_bytes(std::initializer_list<size_t> parts) {
    size_t t = 0;
    for (size_t q : parts) t = ((t + 255) & ~static_cast<size_t>(255)) + q;
    return t + 256;
}

template <class P>
hipError_t run_adjoint(kiwi_hip_ctx* c, hipStream_t s, typename P::Ctx ctx, int64_t n, int64_t n_active,
                       const double* const* dp, const int* bbox, const int64_t* loffs, int64_t nlines, const int* order,
                       const GridGeom& g, const double* const* gin, int clamp, double* const* go, AdjDevStats* dst) {
    using Line = typename P::Line;
    hipError_t e = reserve(c->lines, sizeof(Line) * static_cast<size_t>(std::max<int64_t>(nlines, 1)));
    if (e != hipSuccess) return e;
    Line* table = static_cast<Line*>(c->lines.p);
    if (n > 0) {
        k_adj_lines<P><<<resident_grid(c, k_adj_lines<P>, 1, n), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], bbox,
                                                                                 loffs, g, ctx, table);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (n_active > 0) {
        k_adj_particles<P><<<static_cast<unsigned>(n_active), 64, 0, s>>>(
            order, dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], bbox, loffs, table, ctx, g, gin[0], gin[1],
            gin[2], gin[3], clamp, go[0], go[1], go[2], go[3], go[4], go[5], go[6], go[7], dst);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int kiwi_hip_sph_to_grid_adjoint(kiwi_hip_ctx* c, const kiwi_particles* P, const kiwi_grid* G, const kiwi_kernel* K,
                                 int precision, const kiwi_cell_sums* grad_cells, const kiwi_particle_grads* grad_out,
                                 kiwi_adjoint_stats* stats) {
    if (!c) return KIWI_ERR_INVALID_ARG;
    c->err.clear();
    if (!P || !G || !K || !grad_cells || !grad_out) return fail(c, KIWI_ERR_INVALID_ARG, "null argument");
    if (P->n < 0) return fail(c, KIWI_ERR_INVALID_ARG, "negative particle count");
    if (P->n > 0 && (!P->x || !P->y || !P->z || !P->vx || !P->vy || !P->vz || !P->m || !P->h))
        return fail(c, KIWI_ERR_INVALID_ARG, "null particle array");
    if (!grad_cells->mass || !grad_cells->px || !grad_cells->py || !grad_cells->pz)
        return fail(c, KIWI_ERR_INVALID_ARG, "null cell gradient array");
    if (P->n > 0 && (!grad_out->m || !grad_out->x || !grad_out->y || !grad_out->z || !grad_out->h || !grad_out->vx ||
                     !grad_out->vy || !grad_out->vz))
        return fail(c, KIWI_ERR_INVALID_ARG, "null particle gradient array");
    if (G->nx <= 0 || G->ny <= 0 || G->nz <= 0)
        return fail(c, KIWI_ERR_INVALID_ARG, "cartesian_grid resolution (resolved nx/ny/nz) must be > 0");
    if (!(G->x_min < G->x_max) || !(G->y_min < G->y_max) || !(G->z_min < G->z_max))
        return fail(c, KIWI_ERR_INVALID_ARG, "domain min must be < max on every axis");
    if (precision != KIWI_FP64 && precision != KIWI_MIXED) return fail(c, KIWI_ERR_INVALID_ARG, "unknown precision");
    KernelChoice kc;
    if (const int rc = resolve_kernel(c, K, &kc)) return rc;
    if (kc.tabulated)
        return fail(c, KIWI_ERR_UNSUPPORTED,
                    "the adjoint is available for the analytic and generic integrals (the tabulated integral's "
                    "trilinear lookup is not differentiated)");
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
    const int64_t ncells = static_cast<int64_t>(G->nx) * G->ny * G->nz;

    // inputs and outputs on the device
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
    const double* gin[4] = {grad_cells->mass, grad_cells->px, grad_cells->py, grad_cells->pz};
    if (!grad_cells->on_device) {
        KIWI_HIP_CHECK(c, reserve(c->outbuf, 4 * sizeof(double) * ncells));
        for (int k = 0; k < 4; ++k) {
            double* d = static_cast<double*>(c->outbuf.p) + k * ncells;
            KIWI_HIP_CHECK(c, hipMemcpyAsync(d, gin[k], ncells * sizeof(double), hipMemcpyHostToDevice, s));
            gin[k] = d;
        }
    }
    double* go[8] = {grad_out->m, grad_out->x, grad_out->y, grad_out->z,
                     grad_out->h, grad_out->vx, grad_out->vy, grad_out->vz};
    if (!grad_out->on_device) {
        KIWI_HIP_CHECK(c, reserve(c->ms, 8 * sizeof(double) * std::max<int64_t>(n, 1)));
        for (int k = 0; k < 8; ++k) go[k] = static_cast<double*>(c->ms.p) + k * n;
    }

    // 1. boxes, line-table sizes, work; 3. largest work first
    const int64_t n1 = n + 1;
    KIWI_HIP_CHECK(c, reserve(c->bbox, 6 * sizeof(int) * std::max<int64_t>(n, 1)));
    KIWI_HIP_CHECK(c, reserve(c->sizes, carve_bytes({sizeof(int64_t) * n1, sizeof(int64_t) * n1})));
    KIWI_HIP_CHECK(c, reserve(c->work, carve_bytes({sizeof(uint32_t) * n1, sizeof(uint32_t) * n1, sizeof(int) * n1,
                                                    sizeof(int) * n1})));
    KIWI_HIP_CHECK(c, reserve(c->stats, sizeof(AdjDevStats)));
    int* bbox = static_cast<int*>(c->bbox.p);
    size_t o1 = 0, o2 = 0;
    int64_t* nlin = carve<int64_t>(c->sizes.p, o1, n1);
    int64_t* loffs = carve<int64_t>(c->sizes.p, o1, n1);
    uint32_t* work = carve<uint32_t>(c->work.p, o2, n1);
    uint32_t* work_sorted = carve<uint32_t>(c->work.p, o2, n1);
    int* order_in = carve<int>(c->work.p, o2, n1);
    int* order = carve<int>(c->work.p, o2, n1);
    AdjDevStats* dst = static_cast<AdjDevStats*>(c->stats.p);
    KIWI_HIP_CHECK(c, hipMemsetAsync(dst, 0, sizeof(AdjDevStats), s));
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[0], s));
    for (int k = 0; k < 8; ++k)
        if (n) k_adj_fill<<<grid1d(n), THREADS, 0, s>>>(go[k], n);
    const double sup = kc.support_in_h;
    k_adj_prep<<<grid1d(n1), THREADS, 0, s>>>(n, dp[0], dp[1], dp[2], dp[7], g, sup, bbox, nlin, work, order_in, dst);
    KIWI_HIP_CHECK(c, hipGetLastError());
    {
        size_t tb = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, nlin, loffs, static_cast<int>(n1), s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tb));
        tb = c->cub_tmp.cap;
        KIWI_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, tb, nlin, loffs, static_cast<int>(n1), s));
    }
    if (n > 0) {
        size_t tb = 0;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, work, work_sorted, order_in, order,
                                                                        static_cast<int>(n), 0, 32, s));
        KIWI_HIP_CHECK(c, reserve(c->cub_tmp, tb));
        tb = c->cub_tmp.cap;
        KIWI_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairsDescending(c->cub_tmp.p, tb, work, work_sorted, order_in,
                                                                        order, static_cast<int>(n), 0, 32, s));
    }
    int64_t nlines = 0;
    AdjDevStats hs0;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&nlines, loffs + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&hs0, dst, sizeof(AdjDevStats), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    // active particles (non-zero work) come first in the order
    const int64_t n_active = static_cast<int64_t>(hs0.active);
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[1], s));
    const int clamp = precision == KIWI_FP64 ? 1 : 0;
    // the forward's arithmetic: mixed precision in FP32 for the cubic spline and the generic kernels that
    // allow it (gen_mixed_ok), FP64 otherwise
    const GenTab* tabs = static_cast<const GenTab*>(c->tables.p) + K->kernel_id;
    const bool single = precision == KIWI_MIXED && (!kc.generic || gen_mixed_ok(K->kernel_id));
    if (kc.generic && single)
        KIWI_HIP_CHECK(c, with_adj_gen<float>(K->kernel_id, [&](auto pol) {
                           using P = decltype(pol);
                           return run_adjoint<P>(c, s, typename P::Ctx{sup, tabs}, n, n_active, dp, bbox, loffs, nlines,
                                                 order, g, gin, clamp, go, dst);
                       }));
    else if (kc.generic)
        KIWI_HIP_CHECK(c, with_adj_gen<double>(K->kernel_id, [&](auto pol) {
                           using P = decltype(pol);
                           return run_adjoint<P>(c, s, typename P::Ctx{sup, tabs}, n, n_active, dp, bbox, loffs, nlines,
                                                 order, g, gin, 1, go, dst);  // FP64: per-pair clamp
                       }));
    else if (single)
        KIWI_HIP_CHECK(c, run_adjoint<AdjM4<float>>(c, s, {sup}, n, n_active, dp, bbox, loffs, nlines, order, g, gin,
                                                    clamp, go, dst));
    else
        KIWI_HIP_CHECK(c, run_adjoint<AdjM4<double>>(c, s, {sup}, n, n_active, dp, bbox, loffs, nlines, order, g, gin,
                                                     clamp, go, dst));
    KIWI_HIP_CHECK(c, hipEventRecord(c->ev[2], s));
    if (!grad_out->on_device && n > 0) {
        double* ho[8] = {grad_out->m, grad_out->x, grad_out->y, grad_out->z,
                         grad_out->h, grad_out->vx, grad_out->vy, grad_out->vz};
        for (int k = 0; k < 8; ++k)
            KIWI_HIP_CHECK(c, hipMemcpyAsync(ho[k], go[k], n * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    AdjDevStats hs;
    KIWI_HIP_CHECK(c, hipMemcpyAsync(&hs, dst, sizeof(AdjDevStats), hipMemcpyDeviceToHost, s));
    KIWI_HIP_CHECK(c, hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->skipped_outside = static_cast<int64_t>(hs.skipped_outside);
        stats->pairs_bbox = static_cast<int64_t>(hs.pairs_bbox);
        stats->lattice_points = static_cast<int64_t>(hs.points);
        stats->cells = static_cast<int64_t>(hs.cells);
        float ms_prep = 0, ms_main = 0;
        (void)hipEventElapsedTime(&ms_prep, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_main, c->ev[1], c->ev[2]);
        stats->device_ms_prep = ms_prep;
        stats->device_ms = ms_main;
    }
    return KIWI_OK;
}

}  // extern "C"
