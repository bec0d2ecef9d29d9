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
// stage_copy: the batch's host<->HBM transfers (below).
//
// serve_expand / serve_gather: the operands of requests above the ragged
// size limit cross the bus packed (their limbs only) and are spread into
// their zero-padded class slots on the device; their results are gathered
// from the padded outputs into one packed area, which is all that goes
// back.
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

// Staging copy between pinned host memory and HBM for the serving loop:
// moves `n16` 16-byte words, then one trailing 8-byte word when `tail8` is
// set. It runs on the batch's own stream in place of a copy-engine transfer;
// each thread keeps 4 loads in flight so that host reads over the bus
// overlap.
__global__ __launch_bounds__(256) void stage_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         size_t n16, uint32_t tail8) {
    constexpr int U = 4;
    const size_t stride = (size_t)gridDim.x * 256;
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

hipError_t launch_stage_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if (bytes % 8 || reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(dst) % 16)
        return hipErrorInvalidValue;
    const size_t n16 = bytes / 16;
    const size_t blocks = std::min<size_t>(1024, std::max<size_t>(1, (n16 + 1023) / 1024));
    stage_copy_kernel<<<(uint32_t)blocks, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16,
                                                        (uint32_t)(bytes % 16 != 0));
    return hipGetLastError();
}


// One workgroup per slot: dst[0, n) = src[0, n), dst[n, size) = 0.
__global__ __launch_bounds__(256) void serve_expand_kernel(const ServeSlot* __restrict__ slots) {
    const ServeSlot sl = slots[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < sl.size; i += 256) sl.dst[i] = i < sl.n ? sl.src[i] : 0ull;
}

// One workgroup per result: dst[0, n) = src[0, n), except that with a carry
// byte the top limb dst[n - 1] is that byte.
__global__ __launch_bounds__(256) void serve_gather_kernel(const ServeGather* __restrict__ items) {
    const ServeGather it = items[blockIdx.x];
    const uint32_t m = it.carry ? it.n - 1 : it.n;
    for (uint32_t i = threadIdx.x; i < m; i += 256) it.dst[i] = it.src[i];
    if (it.carry && threadIdx.x == 0) it.dst[m] = *it.carry;
}

hipError_t launch_serve_expand(uint32_t n, const ServeSlot* slots, hipStream_t s) {
    if (n == 0) return hipSuccess;
    serve_expand_kernel<<<n, 256, 0, s>>>(slots);
    return hipGetLastError();
}

hipError_t launch_serve_gather(uint32_t n, const ServeGather* items, hipStream_t s) {
    if (n == 0) return hipSuccess;
    serve_gather_kernel<<<n, 256, 0, s>>>(items);
    return hipGetLastError();
}

}  // namespace midint
