// Stable LSD radix sort of 64-bit keys carrying 32-bit values, on the key bits [b0, b1), four bits per
// pass.  It exists for the sph-to-grid pipeline's captured HIP graphs: it launches only its own kernels
// and hipcub's exclusive scan, so a captured sort holds no memset or memcpy nodes (hipcub's radix sort
// clears its digit offsets and look-back states with hipMemsetAsync).
//
// Each pass: k_rs_hist counts the digits of every chunk of RS_CHUNK keys (one wavefront per chunk, one
// ballot per digit value), an exclusive scan over hist[digit][chunk] gives every chunk's first slot per
// digit, and k_rs_scatter writes each key at that slot plus its rank among the chunk's keys of the same
// digit: chunks in order, keys in order within a chunk, so the sort is stable.
#pragma once

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "common.h"

namespace kiwi_hip {

constexpr int RS_BITS = 4;
constexpr int RS_BINS = 1 << RS_BITS;
constexpr int RS_CHUNK = 1024;  // keys per wavefront and pass

inline int64_t rs_chunks(int64_t n) { return (n + RS_CHUNK - 1) / RS_CHUNK; }

// the digit of a key in the pass at bit `shift` with `nb` bits (descending: mirrored)
__device__ __forceinline__ int rs_digit(uint64_t key, int shift, int nb, bool desc) {
    const int mask = (1 << nb) - 1;
    const int d = static_cast<int>((key >> shift) & static_cast<uint64_t>(mask));
    return desc ? mask - d : d;
}

__global__ __launch_bounds__(THREADS) void k_rs_hist(int64_t n, int64_t nchunks, const uint64_t* __restrict__ keys,
                                                     int shift, int nb, int desc, int64_t* __restrict__ hist) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int64_t chunk = (static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x) >> 6;
    if (chunk >= nchunks) return;
    int64_t cnt = 0;  // lane q < RS_BINS: the chunk's keys with digit q
    const int64_t t0 = chunk * RS_CHUNK;
    for (int r = 0; r < RS_CHUNK / 64; ++r) {
        const int64_t t = t0 + r * 64 + lane;
        const int d = t < n ? rs_digit(keys[t], shift, nb, desc != 0) : -1;
#pragma unroll
        for (int q = 0; q < RS_BINS; ++q) {
            const uint64_t m = __ballot(d == q);
            if (lane == q) cnt += __popcll(m);
        }
    }
    if (lane < RS_BINS) hist[static_cast<int64_t>(lane) * nchunks + chunk] = cnt;
    if (chunk == 0 && lane == 0) hist[static_cast<int64_t>(RS_BINS) * nchunks] = 0;  // the scan's last entry
}

__global__ __launch_bounds__(THREADS) void k_rs_scatter(int64_t n, int64_t nchunks, const uint64_t* __restrict__ kin,
                                                        const int* __restrict__ vin, int shift, int nb, int desc,
                                                        const int64_t* __restrict__ off, uint64_t* __restrict__ kout,
                                                        int* __restrict__ vout) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int64_t chunk = (static_cast<int64_t>(blockIdx.x) * THREADS + threadIdx.x) >> 6;
    if (chunk >= nchunks) return;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t run = lane < RS_BINS ? off[static_cast<int64_t>(lane) * nchunks + chunk] : 0;
    const int64_t t0 = chunk * RS_CHUNK;
    for (int r = 0; r < RS_CHUNK / 64; ++r) {
        const int64_t t = t0 + r * 64 + lane;
        uint64_t key = 0;
        int val = 0, d = -1;
        if (t < n) {
            key = kin[t];
            val = vin[t];
            d = rs_digit(key, shift, nb, desc != 0);
        }
        int64_t pos = 0;
#pragma unroll
        for (int q = 0; q < RS_BINS; ++q) {
            const uint64_t m = __ballot(d == q);
            const int64_t base = __shfl(run, q);
            if (d == q) pos = base + __popcll(m & below);
            if (lane == q) run += __popcll(m);
        }
        if (t < n) {
            kout[pos] = key;
            vout[pos] = val;
        }
    }
}

// hipcub scan storage a sort of n keys needs (queried on the host; reserve it before a capture)
inline hipError_t rs_scan_bytes(int64_t n, size_t* bytes) {
    *bytes = 0;
    const int64_t m = RS_BINS * rs_chunks(n) + 1;
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, static_cast<const int64_t*>(nullptr),
                                            static_cast<int64_t*>(nullptr), static_cast<int>(m));
}

// Sorts (kin, vin) by key bits [b0, b1), ascending or descending, into (kout, vout); kin and vin are left
// unchanged.  ktmp/vtmp: scratch of n entries; hist and hoff: RS_BINS * rs_chunks(n) + 1 entries each;
// scan_tmp: rs_scan_bytes(n) bytes.
inline hipError_t rs_sort_pairs(const uint64_t* kin, const int* vin, uint64_t* kout, int* vout, uint64_t* ktmp,
                                int* vtmp, int64_t n, int b0, int b1, bool desc, int64_t* hist, int64_t* hoff,
                                void* scan_tmp, size_t scan_bytes, hipStream_t s) {
    if (n <= 0 || b1 <= b0) return hipSuccess;
    const int64_t nchunks = rs_chunks(n);
    const int64_t m = RS_BINS * nchunks + 1;
    const int passes = (b1 - b0 + RS_BITS - 1) / RS_BITS;
    // ping-pong so that the last pass writes (kout, vout)
    const uint64_t* ks = kin;
    const int* vs = vin;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* kd = to_out ? kout : ktmp;
        int* vd = to_out ? vout : vtmp;
        const int shift = b0 + p * RS_BITS;
        const int nb = b1 - shift < RS_BITS ? b1 - shift : RS_BITS;
        k_rs_hist<<<grid1d(64 * nchunks), THREADS, 0, s>>>(n, nchunks, ks, shift, nb, desc ? 1 : 0, hist);
        size_t tb = scan_bytes;
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, hist, hoff, static_cast<int>(m), s);
        if (e != hipSuccess) return e;
        k_rs_scatter<<<grid1d(64 * nchunks), THREADS, 0, s>>>(n, nchunks, ks, vs, shift, nb, desc ? 1 : 0, hoff, kd,
                                                              vd);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        ks = kd;
        vs = vd;
    }
    return hipSuccess;
}

}  // namespace kiwi_hip
