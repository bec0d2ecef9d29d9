// Device helpers of the serving loop.
//
// bigmul_combine: a product of operands up to 4096 limbs is computed as four
// 2048 x 2048-limb NTT products laid out per op as [P00 | P11 | P01 | P10]
// (4096 limbs each, 16384 per op). With X = 2^(64*2048),
//   a*b = P00 + (P01 + P10) X + P11 X^2,
// and P00 | P11 already is P00 + P11 X^2 (disjoint limbs), so the combine
// adds P01 and then P10 into limbs [2048, 8192): two chained carry-resolved
// additions on register-resident rows (6144 limbs = 16 waves x 3 rows).
//
// stage_in / stage_out: the batch's host<->HBM transfers (below). The
// operands of requests above the ragged size limit cross the bus packed
// (their limbs only) and are spread into their zero-padded class slots on
// the way in; their results are gathered from the padded outputs into
// their packed places on the way out.
#include <algorithm>

#include "add_core.hpp"
#include "midint_internal.h"

namespace midint {

__global__ __launch_bounds__(1024) void bigmul_combine_kernel(uint64_t* __restrict__ blocks) {
    constexpr int R = 3;
    constexpr uint32_t W = 6144;
    __shared__ uint32_t lds_gp[16];
    uint64_t* base = blocks + (size_t)blockIdx.x * 16384;
    uint64_t* dst = base + 2048;
    const uint64_t* p01 = base + 8192;
    const uint64_t* p10 = base + 12288;
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
    Rows<R> t;
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        t.x0[v] = k < W ? dst[k] : 0ull;
        t.x1[v] = k + 1 < W ? dst[k + 1] : 0ull;
        t.y0[v] = k < 4096 ? p01[k] : 0ull;
        t.y1[v] = k + 1 < 4096 ? p01[k + 1] : 0ull;
        t.s0[v] = (k == 0);
        t.s1[v] = false;
    }
    block_add_rows<R>(t, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        t.y0[v] = k < 4096 ? p10[k] : 0ull;
        t.y1[v] = k + 1 < 4096 ? p10[k + 1] : 0ull;
    }
    block_add_rows<R>(t, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        if (k < W) dst[k] = t.x0[v];
        if (k + 1 < W) dst[k + 1] = t.x1[v];
    }
}

hipError_t launch_bigmul_combine(uint32_t n, uint64_t* blocks16k, hipStream_t s) {
    if (n == 0) return hipSuccess;
    bigmul_combine_kernel<<<n, 1024, 0, s>>>(blocks16k);
    return hipGetLastError();
}

}  // namespace midint

namespace midint {

// The serving loop's transfers between its pinned staging buffer and HBM,
// on the batch's own stream in place of copy-engine transfers. Blocks [0,
// copy_blocks) move `n16` 16-byte words, then one trailing 8-byte word when
// `tail8` is set (each thread keeps 4 loads in flight so that reads over the
// bus overlap); each further block handles one entry of `slots` (inbound:
// packed limbs read from the staging buffer into a zero-padded device slot)
// or `items` (outbound: a result read from its padded device output into
// its packed place in the staging buffer). The tables are read where the
// host wrote them, in the staging buffer.
__device__ __forceinline__ void stage_copy_part(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16,
                                                uint32_t tail8, uint32_t copy_blocks) {
    constexpr int U = 4;
    const size_t stride = (size_t)copy_blocks * 256;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + (U - 1) * stride < n16; i += U * stride) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; u++) dst[i + u * stride] = v[u];
    }
    for (; i < n16; i += stride) dst[i] = src[i];
    if (tail8 && blockIdx.x == 0 && threadIdx.x == 0)
        reinterpret_cast<uint64_t*>(dst + n16)[0] = reinterpret_cast<const uint64_t*>(src + n16)[0];
}

__global__ __launch_bounds__(256) void stage_in_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       size_t n16, uint32_t tail8, uint32_t copy_blocks,
                                                       const ServeSlot* __restrict__ slots) {
    if (blockIdx.x < copy_blocks) {
        stage_copy_part(src, dst, n16, tail8, copy_blocks);
        return;
    }
    // dst[0, n) = src[0, n), dst[n, size) = 0, in chunks of 16 limbs per
    // thread whose reads over the bus are all in flight at once
    const ServeSlot sl = slots[blockIdx.x - copy_blocks];
    constexpr uint32_t U = 16;
    for (uint32_t base = 0; base < sl.size; base += U * 256) {
        uint64_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = base + u * 256 + threadIdx.x;
            v[u] = i < sl.n ? sl.src[i] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = base + u * 256 + threadIdx.x;
            if (i < sl.size) sl.dst[i] = v[u];
        }
    }
}

__global__ __launch_bounds__(256) void stage_out_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        size_t n16, uint32_t tail8, uint32_t copy_blocks,
                                                        const ServeGather* __restrict__ items) {
    if (blockIdx.x < copy_blocks) {
        stage_copy_part(src, dst, n16, tail8, copy_blocks);
        return;
    }
    // dst[0, n) = src[0, n), except that with a carry byte the top limb is
    // that byte; 16 limbs per thread at once
    const ServeGather it = items[blockIdx.x - copy_blocks];
    const uint32_t m = it.carry ? it.n - 1 : it.n;
    constexpr uint32_t U = 16;
    for (uint32_t base = 0; base < m; base += U * 256) {
        uint64_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = base + u * 256 + threadIdx.x;
            v[u] = i < m ? it.src[i] : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = base + u * 256 + threadIdx.x;
            if (i < m) it.dst[i] = v[u];
        }
    }
    if (it.carry && threadIdx.x == 0) it.dst[m] = *it.carry;
}

namespace {

hipError_t stage_geometry(const void* src, void* dst, size_t bytes, size_t* n16, uint32_t* blocks) {
    if (bytes % 8 || reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(dst) % 16)
        return hipErrorInvalidValue;
    *n16 = bytes / 16;
    *blocks = bytes == 0 ? 0u : (uint32_t)std::min<size_t>(1024, std::max<size_t>(1, (*n16 + 1023) / 1024));
    return hipSuccess;
}

}  // namespace

hipError_t launch_stage_in(const void* src, void* dst, size_t bytes, const ServeSlot* slots, uint32_t n_slots,
                           hipStream_t s) {
    size_t n16 = 0;
    uint32_t blocks = 0;
    hipError_t e = stage_geometry(src, dst, bytes, &n16, &blocks);
    if (e != hipSuccess || blocks + n_slots == 0) return e;
    stage_in_kernel<<<blocks + n_slots, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16,
                                                     (uint32_t)(bytes % 16 != 0), blocks, slots);
    return hipGetLastError();
}

hipError_t launch_stage_out(const void* src, void* dst, size_t bytes, const ServeGather* items, uint32_t n_items,
                            hipStream_t s) {
    size_t n16 = 0;
    uint32_t blocks = 0;
    hipError_t e = stage_geometry(src, dst, bytes, &n16, &blocks);
    if (e != hipSuccess || blocks + n_items == 0) return e;
    stage_out_kernel<<<blocks + n_items, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16,
                                                      (uint32_t)(bytes % 16 != 0), blocks, items);
    return hipGetLastError();
}

}  // namespace midint
