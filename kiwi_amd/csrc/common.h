// Shared by the library's translation units: launch shape, small device helpers, and the context
// (device, stream, events, device buffers that grow and are reused across calls).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kiwi_hip.h"
#include "kernel_tables.h"

namespace {

constexpr int THREADS = 256;
constexpr int WAVES = THREADS / 64;

template <typename R>
struct Pair2 {
    R a, b;
};

// select one of three by a runtime axis index (no indexed private arrays: those would live in scratch)
template <typename T>
__device__ __forceinline__ T sel3(int a, T v0, T v1, T v2) { return a == 0 ? v0 : (a == 1 ? v1 : v2); }

// u / d for non-negative integers through a float reciprocal inv_d ~ 1/d (within a few ulp), corrected
// by one step either way to be exact; valid while u / d < 2^22
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

// a value that is the same in every lane, moved to a scalar register
// Double operations rounded on their own, never fused into a multiply-add: the comparisons that
// decide which site, cell or candidate a result uses follow the reference's separately rounded
// arithmetic.  (The compiler contracts a + b * c into one fma by default, also when written with
// __dadd_rn / __dmul_rn, which are plain + and * here.)
__device__ __forceinline__ double mul_sep(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_sep(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double sub_sep(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t wave_uniform64(uint64_t v) {
    return static_cast<uint64_t>(wave_uniform(static_cast<uint32_t>(v))) |
           (static_cast<uint64_t>(wave_uniform(static_cast<uint32_t>(v >> 32))) << 32);
}

// LDS writes of one lane made visible to the other lanes of the same wavefront
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

inline int grid1d(int64_t n) { return static_cast<int>((n + THREADS - 1) / THREADS); }

}  // namespace

// A fixed device area that buffers are carved from in place of hipMalloc (the allocation-free entry
// point): regions are handed out in order and never freed; a call starts again from offset 0.
struct ScratchArena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    ScratchArena* arena = nullptr;  // carve from this area instead of allocating
};

// One captured sph-to-grid pipeline (kiwi_hip_set_graphs) and what its launches hold: sizes, launch
// shape, and the address of every buffer they touch (a buffer that has been reallocated since the
// capture changes the signature, so a replay only ever uses live buffers).
struct GraphEntry {
    std::vector<uint64_t> sig;
    hipGraphExec_t exec = nullptr;
};
constexpr size_t KIWI_GRAPH_CACHE = 8;  // captured pipelines kept per context (least recently used out)

// The device buffers of one sph-to-grid pipeline run: the context's own, or shadows carved from the
// allocation-free scratch area.
#define KIWI_PIPE_BUFS(X)                                                                                      \
    X(parts) X(bbox) X(sizes) X(offsets) X(keys_a) X(keys_b) X(pos_a) X(pos_b) X(cells) X(slab_sorted) X(slab) \
    X(cub_tmp) X(ranges) X(work) X(work_sorted) X(order_in) X(order) X(lines) X(ms) X(stats) X(outbuf) X(vsum)  \
    X(probs) X(part_prob) X(acc_units) X(acc_partial) X(acc_parts)
#define KIWI_PB_COUNT(n) +1
constexpr int KIWI_N_PIPE_BUFS = 0 KIWI_PIPE_BUFS(KIWI_PB_COUNT);
#undef KIWI_PB_COUNT
struct PipeBufs {
#define KIWI_PB_MEMBER(n) DevBuf* n;
    KIWI_PIPE_BUFS(KIWI_PB_MEMBER)
#undef KIWI_PB_MEMBER
};

// The last tessellation's search structures (kiwi_hip_voronoi_tessellate), for the home-cell lookup:
// voro++'s block grid over the domain and the bucket grid of the kept sites.
struct TessGrid {
    double dom[6], lo[3], edge[3], tolb;
    int nbk[3], nb[3];
};

struct kiwi_hip_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;      // second stream (default priority): line tables beside the item sort
    hipEvent_t side_done = nullptr;  // recorded on side after its work; the main stream waits on it
    // third stream (default priority): mixed precision's FP64 fallback, its line tables and cell kernel
    hipStream_t side2 = nullptr;
    hipEvent_t side2_done = nullptr;  // recorded on side2 after its work; the main stream waits on it
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // [0, 1] bracket the line-table kernel on the second stream; on the third: [2, 3] the FP64 fallback's
    // cell kernel, [4, 5] its line tables
    hipEvent_t ev_side[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    std::string err;
    DevBuf parts, bbox, sizes, offsets, keys_a, keys_b, pos_a, pos_b, cells, slab_sorted, slab, cub_tmp, ranges, work,
        work_sorted, order_in, order, lines, ms, stats, outbuf, vsum, tables, m4table, probs, part_prob, acc_units,
        acc_partial, acc_parts;
    bool m4table_ready = false;
    kiwi_tuning tune{16, 8};  // the cell kernel's launch shape (kiwi_hip_set_tuning)
    // mixed precision's FP64 fallback: support radius in cells from which a particle computes in FP64
    // (kiwi_hip_set_fp64_fallback; <= 0: off)
    double fb_cells = KIWI_FP64_FALLBACK_CELLS;
    DevBuf vor[16];  // the Voronoi map's buffers
    // the Voronoi tessellation's buffers; [8] .. [11] hold the last tessellation (kiwi_hip_voronoi_tessellate)
    DevBuf tess[12];
    int64_t tess_sizes[3] = {-1, 0, 0};  // the last tessellation's cells, faces, face-vertex entries (-1: none)
    TessGrid tess_grid{};
    // sph-to-grid HIP graphs (kiwi_hip_set_graphs): the captured pipelines, most recently used last; the
    // dependency-only events that fork the second stream in a capture; the graph-mode radix sort's
    // buffers; pinned host memory that the replayed pipeline copies its statistics to
    int graphs = 0;
    std::vector<GraphEntry> graph_cache;
    hipEvent_t fork = nullptr, fork2 = nullptr;
    DevBuf g_ktmp, g_vtmp, g_hist, g_hoff;
    void* g_host = nullptr;
    size_t g_host_cap = 0;
    // the allocation-free entry point: its arena over the device's scratch area and shadow buffers
    ScratchArena noalloc_arena;
    DevBuf noalloc_bufs[KIWI_N_PIPE_BUFS];
    PipeBufs own_bufs() {
        PipeBufs b;
#define KIWI_PB_OWN(n) b.n = &n;
        KIWI_PIPE_BUFS(KIWI_PB_OWN)
#undef KIWI_PB_OWN
        return b;
    }
    template <class F>
    void for_each_buf(F f) {
        for (DevBuf* b : {&parts, &bbox, &sizes, &offsets, &keys_a, &keys_b, &pos_a, &pos_b, &cells, &slab_sorted,
                          &slab, &cub_tmp, &ranges, &work, &work_sorted, &order_in, &order, &lines, &ms, &stats,
                          &outbuf, &vsum, &tables, &m4table, &probs, &part_prob, &acc_units, &acc_partial,
                          &acc_parts, &g_ktmp, &g_vtmp, &g_hist, &g_hoff})
            f(*b);
        for (DevBuf& b : vor) f(b);
        for (DevBuf& b : tess) f(b);
    }
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

// Whether a kernel may dereference p: device, managed or pinned host memory (not pageable host memory,
// which a kernel faults on).  A failed query leaves no error behind for later checks.
inline bool gpu_accessible(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type != hipMemoryTypeUnregistered;
}

hipError_t reserve(DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return hipSuccess;
    if (b.arena) {  // a new region of the fixed area (the old one is not reused within the call)
        ScratchArena& a = *b.arena;
        const size_t off = (a.used + 255) & ~static_cast<size_t>(255);
        if (off > a.cap || bytes > a.cap - off) return hipErrorOutOfMemory;
        b.p = a.base + off;
        b.cap = bytes;
        a.used = off + bytes;
        return hipSuccess;
    }
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

// workgroups of a persistent kernel: as many as are resident on the device at once (every wave
// loops until the work runs out), fewer for small jobs
template <typename K>
int resident_grid(const kiwi_hip_ctx* c, K kernel, int64_t items_per_wave_round, int64_t items) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, THREADS, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int64_t full = static_cast<int64_t>(per_cu) * std::max(c->n_cu, 1);
    const int64_t need = (items + WAVES * items_per_wave_round - 1) / (WAVES * items_per_wave_round);
    return static_cast<int>(std::max<int64_t>(1, std::min(full, need)));
}

// ------------------------------------------------------------------------------------------------
// Cartesian grid geometry shared by the forward and the adjoint
// ------------------------------------------------------------------------------------------------
constexpr int BLK = 8;  // cell block edge (cells)

struct GridGeom {
    double x_min, y_min, z_min, dx, dy, dz;
    int nx, ny, nz;
    int nbx, nby, nbz;  // cell blocks per axis
};

__device__ __forceinline__ double gmin(const GridGeom& g, int a) { return sel3(a, g.x_min, g.y_min, g.z_min); }
__device__ __forceinline__ double gstep(const GridGeom& g, int a) { return sel3(a, g.dx, g.dy, g.dz); }

// floor(v) as a cell index, saturated to [-1, n] before the conversion so that huge supports and
// non-finite coordinates cannot overflow it (NaN lands on n, outside the grid).  Within range this is
// exactly the reference's static_cast<int>(std::floor(v)) (src/cartesian_grid.cpp:220-225).
__device__ __forceinline__ int cell_floor(double v, int n) {
    const double f = floor(v);
    if (f >= -1.0 && f <= static_cast<double>(n)) return static_cast<int>(f);
    return (f < -1.0) ? -1 : n;
}

// the particle's cell bounding box [q0, q1] x [q2, q3] x [q4, q5] clamped to the grid
// (src/cartesian_grid.cpp:216-238); false if it misses the grid.  Requires h > 0.
__device__ __forceinline__ bool particle_bbox(const GridGeom& g, double support_in_h, double x0, double y0, double z0,
                                              double hh, int* q) {
    const double rad = mul_sep(support_in_h, hh);  // x0 - rad as the reference rounds it, not one fma
    q[0] = max(cell_floor((x0 - rad - g.x_min) / g.dx, g.nx), 0);
    q[1] = min(cell_floor((x0 + rad - g.x_min) / g.dx, g.nx), g.nx - 1);
    q[2] = max(cell_floor((y0 - rad - g.y_min) / g.dy, g.ny), 0);
    q[3] = min(cell_floor((y0 + rad - g.y_min) / g.dy, g.ny), g.ny - 1);
    q[4] = max(cell_floor((z0 - rad - g.z_min) / g.dz, g.nz), 0);
    q[5] = min(cell_floor((z0 + rad - g.z_min) / g.dz, g.nz), g.nz - 1);
    return q[0] <= q[1] && q[2] <= q[3] && q[4] <= q[5];
}

// deterministic velocity sums over all particles: fixed per-thread strides, fixed tree order
constexpr int VSUM_BLOCKS = 256;

__global__ __launch_bounds__(THREADS) __attribute__((unused)) void k_vsum_partial(int64_t n, const double* __restrict__ vx,
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

__global__ __attribute__((unused)) void k_vsum_final(const double* partial, double* out) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; ++k) {
        double t = 0;
        for (int i = 0; i < VSUM_BLOCKS; ++i) t += partial[3 * i + k];
        out[k] = t;
    }
}

}  // namespace

namespace {

// the kernel tables on the host, in KIWI_KERNEL_M4 .. KIWI_KERNEL_WC6 order
const KiwiKernelTable kHostTables[KIWI_TAB_NKERNELS] = {KIWI_KERNEL_TABLE_DATA};

// Generic kernels whose mixed-precision requests run in single precision: the cubic spline and
// Wendland C2, whose FP32 lattice values carry the cubic path's error (a few 1e-7 of a particle's mass
// in total, growing with h in cells like the cubic path's).  The higher-order kernels (M5, M6, Wendland
// C4 and C6) lose accuracy in their basis recurrences in single precision and run in FP64 whatever the
// request.
inline bool gen_mixed_ok(int kernel_id) { return kernel_id == KIWI_KERNEL_M4 || kernel_id == KIWI_KERNEL_WC2; }

// The sph-to-grid forward in mixed precision also computes M5 in FP32 while the FP64 fallback is on.  Its
// monomial expansion cancels more in FP32 than the cubic spline's or Wendland C2's (lattice terms up to
// 6.5 against |W| <= 1/8, against 2-3.4), so its fallback starts at 0.75 of the threshold, where an
// isolated particle's FP32 cell error matches theirs at the full threshold (host measurement, mass-
// weighted L1 of the FP32 cell sums against FP64 at supports of 4 / 6 / 8 / 12 cells: WC2 1.1e-4 /
// 2.9e-4 / 5.6e-4 / 1.6e-3, M5 1.8e-4 / 5.2e-4 / 1.1e-3 / 3.3e-3).  M6 and Wendland C4 and C6 stay in
// FP64: their terms reach 13-150, and a particle cut by the domain boundary (whose mass inside is a
// difference of boundary lattice values, whatever its size) keeps a total-mass error of 1.6e-6 to 1.9e-5
// in FP32, above the 1e-6 that mixed precision keeps (500 particles -> 20^3 of the generic fixtures).
inline bool gen_forward_fp32(int kernel_id, double fb_cells) {
    return gen_mixed_ok(kernel_id) || (kernel_id == KIWI_KERNEL_M5 && fb_cells > 0.0);
}
inline double gen_fallback_scale(int kernel_id) { return kernel_id == KIWI_KERNEL_M5 ? 0.75 : 1.0; }

// Which integral a kernel request selects: make_kernel_integrator (include/kiwi/kernel_integrator.hpp:
// 14-46, src/config.cpp:269-277) -- the cubic analytic path iff M4 with support 2h and no generic request,
// the generic path otherwise; the tabulated integral only on the cubic path (src/config.cpp:174-178).
struct KernelChoice {
    double support_in_h;
    bool generic, tabulated;
};

inline int resolve_kernel(kiwi_hip_ctx* c, const kiwi_kernel* K, KernelChoice* out) {
    if (K->kernel_id < 0 || K->kernel_id > KIWI_KERNEL_WC6) return fail(c, KIWI_ERR_INVALID_ARG, "unknown kernel id");
    if (!(K->support_in_h >= 0.0) || !(K->support_in_h < INFINITY))  // src/config.cpp:169-170
        return fail(c, KIWI_ERR_INVALID_ARG, "kernel.support_radius_in_h must be > 0 (or 0 for the kernel's default)");
    out->support_in_h = K->support_in_h > 0 ? K->support_in_h : kHostTables[K->kernel_id].radkern;
    out->generic = K->generic_path || K->kernel_id != KIWI_KERNEL_M4 || out->support_in_h != 2.0;
    out->tabulated = K->tabulated != 0;
    if (out->tabulated && out->generic)
        return fail(c, KIWI_ERR_INVALID_ARG, "the tabulated integral is only available on the cubic-spline fast path");
    return KIWI_OK;
}

}  // namespace

// the tabulated cubic spline's lookup table of the context (built on first use; kiwi_hip.hip)
hipError_t kiwi_build_m4_table(kiwi_hip_ctx* c, hipStream_t s);
