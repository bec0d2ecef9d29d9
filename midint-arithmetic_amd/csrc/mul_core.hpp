// Schoolbook column-group convolution shared by the fixed-size and the
// ragged multiplication kernels.
//
// Operands are u32 word vectors in LDS: the shorter one S (ns words, zero
// padded by at least Q words) and the longer one W stored at W[(Q-1) + j]
// with Q-1 zeros before and >= Q zeros after its nl words. Column c of the
// product is sum_{i+j=c} S_i W_j.
//
// A lane owns Q consecutive columns at a time. Its loop runs NB blocks of Q
// a-words each: blocks [0, sw) accumulate the columns [k1, k1+Q) with
// i = 0, Q, 2Q...; at block sw the accumulators are parked in `snap` (3Q
// words) and blocks [sw, NB) continue on columns [k2, k2+Q) with i starting
// at i2. Keeping one accumulator set and parking it costs one masked LDS
// store per lane instead of per-product selects; the second group is the
// difference of the final and the parked values.
//
// Lane kinds for an ns x nl product (ns <= nl, both multiples of Q):
//   pair lane t < ns/Q  : k1 = Qt,       sw = t+1, k2 = nl + Qt, i2 = Qt
//   plateau lane p      : k1 = ns + Qp,  sw = NB (no switch)
// and NB = ns/Q + 1 for every lane, so all lanes of a wave run the same
// number of blocks (the reference's per-thread column pairing,
// cuda/ker-classic-mul.cu.h:406-423, balances work per thread but leaves
// lanes of a warp waiting on each other's longer loops).
#pragma once

#include "midint_dev.hpp"

namespace midint {

constexpr uint32_t kMulQ = 4;

// acc[q] += a * b_q (32x32->64) and ovf[q] += carry-out, for q = 0..3.
// One v_mad_u64_u32 per product writes the 64-bit sum and its carry-out lane
// mask; one v_addc_co_u32 folds the mask into the overflow count. Each mask
// is read three VALU instructions after it is written.
__device__ __forceinline__ void mac4(uint32_t a, uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3,
                                     uint64_t (&acc)[4], uint32_t (&ovf)[4]) {
    uint64_t k0, k1, k2, k3;
    asm volatile(
        "v_mad_u64_u32 %[c0], %[k0], %[a], %[b0], %[c0]\n\t"
        "v_mad_u64_u32 %[c1], %[k1], %[a], %[b1], %[c1]\n\t"
        "v_mad_u64_u32 %[c2], %[k2], %[a], %[b2], %[c2]\n\t"
        "v_mad_u64_u32 %[c3], %[k3], %[a], %[b3], %[c3]\n\t"
        "v_addc_co_u32_e64 %[o0], %[k0], %[o0], 0, %[k0]\n\t"
        "v_addc_co_u32_e64 %[o1], %[k1], %[o1], 0, %[k1]\n\t"
        "v_addc_co_u32_e64 %[o2], %[k2], %[o2], 0, %[k2]\n\t"
        "v_addc_co_u32_e64 %[o3], %[k3], %[o3], 0, %[k3]"
        : [c0] "+v"(acc[0]), [c1] "+v"(acc[1]), [c2] "+v"(acc[2]), [c3] "+v"(acc[3]), [o0] "+v"(ovf[0]),
          [o1] "+v"(ovf[1]), [o2] "+v"(ovf[2]), [o3] "+v"(ovf[3]), [k0] "=&s"(k0), [k1] "=&s"(k1),
          [k2] "=&s"(k2), [k3] "=&s"(k3)
        : [a] "v"(a), [b0] "v"(b0), [b1] "v"(b1), [b2] "v"(b2), [b3] "v"(b3));
}

struct ColGroups {
    uint64_t acc[kMulQ];  // after the loop: group 1 + group 2 (or the single group)
    uint32_t ovf[kMulQ];
    uint64_t sacc[kMulQ];  // parked group 1 (pair lanes)
    uint32_t sovf[kMulQ];
};

__device__ __forceinline__ void conv_lane(const uint32_t* S, const uint32_t* W, uint32_t k1, uint32_t k2,
                                          uint32_t i2, uint32_t sw, uint32_t NB, uint32_t* snap, ColGroups& g) {
    constexpr uint32_t Q = kMulQ;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        g.acc[q] = 0;
        g.ovf[q] = 0;
    }
    const uint32_t* pa = S;
    const uint32_t* pb = W + k1;
    for (uint32_t beta = 0; beta < NB; beta++) {
        if (beta == sw) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                snap[3 * q] = (uint32_t)g.acc[q];
                snap[3 * q + 1] = (uint32_t)(g.acc[q] >> 32);
                snap[3 * q + 2] = g.ovf[q];
            }
            pa = S + i2;
            pb = W + (k2 - i2);
        }
        const uint4 av = *reinterpret_cast<const uint4*>(pa);
        const uint4 bw0 = *reinterpret_cast<const uint4*>(pb);
        const uint4 bw1 = *reinterpret_cast<const uint4*>(pb + 4);
        // column q, a-word u uses window word q - u + Q - 1
        mac4(av.x, bw0.w, bw1.x, bw1.y, bw1.z, g.acc, g.ovf);
        mac4(av.y, bw0.z, bw0.w, bw1.x, bw1.y, g.acc, g.ovf);
        mac4(av.z, bw0.y, bw0.z, bw0.w, bw1.x, g.acc, g.ovf);
        mac4(av.w, bw0.x, bw0.y, bw0.z, bw0.w, g.acc, g.ovf);
        pa += Q;
        pb -= Q;
    }
    if (sw < NB) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            g.sacc[q] = (uint64_t)snap[3 * q] | ((uint64_t)snap[3 * q + 1] << 32);
            g.sovf[q] = snap[3 * q + 2];
        }
    }
}

// Splits the value sum_q (va[q] + vo[q] 2^64) 2^(32q) of a column group at
// word column kcol into Q/2 low limbs written to L[kcol/2 ...] and a high
// part of two limbs written to H[kcol/2 + Q/2 ...] (dropped at or beyond P
// limbs). H[0..Q/2) of the segment is cleared by the group at column 0.
__device__ __forceinline__ void group_store(const uint64_t (&va)[kMulQ], const uint32_t (&vo)[kMulQ],
                                            uint32_t kcol, uint64_t* L, uint64_t* H, uint32_t P) {
    constexpr uint32_t Q = kMulQ;
    uint64_t lo = 0, hi = 0;
    uint32_t w[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        uint64_t nlo = lo + va[q];
        hi += (uint64_t)vo[q] + (nlo < lo ? 1ull : 0ull);
        lo = nlo;
        w[q] = (uint32_t)lo;
        lo = (lo >> 32) | (hi << 32);
        hi >>= 32;
    }
    const uint32_t limb = kcol / 2;
#pragma unroll
    for (int i = 0; i < Q / 2; i++) L[limb + i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    if (limb + Q / 2 < P) H[limb + Q / 2] = lo;
    if (limb + Q / 2 + 1 < P) H[limb + Q / 2 + 1] = hi;
    if (kcol == 0) {
#pragma unroll
        for (int i = 0; i < Q / 2; i++) H[i] = 0;
    }
}

// Stores the one or two column groups of a lane after conv_lane.
__device__ __forceinline__ void lane_store(const ColGroups& g, bool two, uint32_t k1, uint32_t k2, uint64_t* L,
                                           uint64_t* H, uint32_t P) {
    constexpr uint32_t Q = kMulQ;
    if (two) {
        group_store(g.sacc, g.sovf, k1, L, H, P);
        uint64_t va[Q];
        uint32_t vo[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            va[q] = g.acc[q] - g.sacc[q];
            vo[q] = g.ovf[q] - g.sovf[q] - (g.acc[q] < g.sacc[q] ? 1u : 0u);
        }
        group_store(va, vo, k2, L, H, P);
    } else {
        group_store(g.acc, g.ovf, k1, L, H, P);
    }
}

}  // namespace midint
