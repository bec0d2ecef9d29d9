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
//
// With PF (prefetch) the LDS reads of block beta+1 are issued before the
// products of block beta, so their latency overlaps the VALU work.
#pragma once

#include "midint_dev.hpp"

namespace midint {

// c_q += a * b_q (32x32->64) and o_q += carry-out, q = 0..3. One
// v_mad_u64_u32 per product writes the 64-bit sum and its carry-out lane
// mask; one v_addc_co_u32 folds the mask into the overflow count. Each mask
// is read three VALU instructions after it is written.
__device__ __forceinline__ void mac4(uint32_t a, uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, uint64_t& c0,
                                     uint64_t& c1, uint64_t& c2, uint64_t& c3, uint32_t& o0, uint32_t& o1,
                                     uint32_t& o2, uint32_t& o3) {
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
        : [c0] "+v"(c0), [c1] "+v"(c1), [c2] "+v"(c2), [c3] "+v"(c3), [o0] "+v"(o0), [o1] "+v"(o1),
          [o2] "+v"(o2), [o3] "+v"(o3), [k0] "=&s"(k0), [k1] "=&s"(k1), [k2] "=&s"(k2), [k3] "=&s"(k3)
        : [a] "v"(a), [b0] "v"(b0), [b1] "v"(b1), [b2] "v"(b2), [b3] "v"(b3));
}

// Eight columns against one a-word: 8 x v_mad_u64_u32 (carry masks in eight
// SGPR pairs) followed by 8 x v_addc_co_u32, so every mask is read eight
// instructions after it is written (fewer stalls at low occupancy than the
// 4+4 order of mac4).
__device__ __forceinline__ void mac8(uint32_t a, const uint32_t* b, uint64_t* c, uint32_t* o) {
    uint64_t k0, k1, k2, k3, k4, k5, k6, k7;
    asm volatile(
        "v_mad_u64_u32 %[c0], %[k0], %[a], %[b0], %[c0]\n\t"
        "v_mad_u64_u32 %[c1], %[k1], %[a], %[b1], %[c1]\n\t"
        "v_mad_u64_u32 %[c2], %[k2], %[a], %[b2], %[c2]\n\t"
        "v_mad_u64_u32 %[c3], %[k3], %[a], %[b3], %[c3]\n\t"
        "v_mad_u64_u32 %[c4], %[k4], %[a], %[b4], %[c4]\n\t"
        "v_mad_u64_u32 %[c5], %[k5], %[a], %[b5], %[c5]\n\t"
        "v_mad_u64_u32 %[c6], %[k6], %[a], %[b6], %[c6]\n\t"
        "v_mad_u64_u32 %[c7], %[k7], %[a], %[b7], %[c7]"
        : [c0] "+v"(c[0]), [c1] "+v"(c[1]), [c2] "+v"(c[2]), [c3] "+v"(c[3]), [c4] "+v"(c[4]), [c5] "+v"(c[5]),
          [c6] "+v"(c[6]), [c7] "+v"(c[7]), [k0] "=&s"(k0), [k1] "=&s"(k1), [k2] "=&s"(k2), [k3] "=&s"(k3),
          [k4] "=&s"(k4), [k5] "=&s"(k5), [k6] "=&s"(k6), [k7] "=&s"(k7)
        : [a] "v"(a), [b0] "v"(b[0]), [b1] "v"(b[1]), [b2] "v"(b[2]), [b3] "v"(b[3]), [b4] "v"(b[4]),
          [b5] "v"(b[5]), [b6] "v"(b[6]), [b7] "v"(b[7]));
    asm volatile(
        "v_addc_co_u32_e64 %[o0], %[k0], %[o0], 0, %[k0]\n\t"
        "v_addc_co_u32_e64 %[o1], %[k1], %[o1], 0, %[k1]\n\t"
        "v_addc_co_u32_e64 %[o2], %[k2], %[o2], 0, %[k2]\n\t"
        "v_addc_co_u32_e64 %[o3], %[k3], %[o3], 0, %[k3]\n\t"
        "v_addc_co_u32_e64 %[o4], %[k4], %[o4], 0, %[k4]\n\t"
        "v_addc_co_u32_e64 %[o5], %[k5], %[o5], 0, %[k5]\n\t"
        "v_addc_co_u32_e64 %[o6], %[k6], %[o6], 0, %[k6]\n\t"
        "v_addc_co_u32_e64 %[o7], %[k7], %[o7], 0, %[k7]"
        : [o0] "+v"(o[0]), [o1] "+v"(o[1]), [o2] "+v"(o[2]), [o3] "+v"(o[3]), [o4] "+v"(o[4]), [o5] "+v"(o[5]),
          [o6] "+v"(o[6]), [o7] "+v"(o[7]), [k0] "+s"(k0), [k1] "+s"(k1), [k2] "+s"(k2), [k3] "+s"(k3),
          [k4] "+s"(k4), [k5] "+s"(k5), [k6] "+s"(k6), [k7] "+s"(k7));
}

template <int Q>
struct ColGroups {
    uint64_t acc[Q];  // after the loop: group 1 + group 2 (or the single group)
    uint32_t ovf[Q];
    uint64_t sacc[Q];  // parked group 1 (pair lanes)
    uint32_t sovf[Q];
};

template <int Q>
struct Operands {
    uint32_t a[Q];
    uint32_t b[2 * Q];
};

// Skewed layout of the sliding-window operand W (fixed-size kernels, Q >= 8):
// logical word y is stored at y + 4 floor(y / Q), so every Q-word segment
// is followed by 4 pad words. Lane windows start on segment boundaries, Q
// logical words apart, hence Q + 4 physical words apart: an odd number of
// 4-bank quads (3 for Q = 8, 5 for Q = 16), which puts the 16-byte reads of
// any 16 consecutive lanes on distinct bank quads. Unskewed, lanes 8 words
// apart share quads pairwise and every window read is a two-way bank
// conflict. A window (2Q logical words) is two whole segments: physical
// [p, p + Q) and [p + Q + 4, p + 2Q + 4), and moving it down by one block
// moves p down by Q + 4.
template <int Q>
__host__ __device__ constexpr uint32_t skew_idx(uint32_t y) {
    return y + 4 * (y / Q);
}
template <int Q>
__host__ __device__ constexpr uint32_t skew_len(uint32_t n) {
    return n + 4 * ((n + Q - 1) / Q);
}

template <int Q, bool SKEW = false>
__device__ __forceinline__ void load_block(const uint32_t* pa, const uint32_t* pb, Operands<Q>& o) {
#pragma unroll
    for (int i = 0; i < Q; i += 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(pa + i);
        o.a[i] = v.x;
        o.a[i + 1] = v.y;
        o.a[i + 2] = v.z;
        o.a[i + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < 2 * Q; i += 4) {
        const uint32_t off = SKEW && i >= Q ? i + 4 : i;
        const uint4 v = *reinterpret_cast<const uint4*>(pb + off);
        o.b[i] = v.x;
        o.b[i + 1] = v.y;
        o.b[i + 2] = v.z;
        o.b[i + 3] = v.w;
    }
}

// Digit mode: operands are d-bit digits (d <= 28), so every product is
// below 2^56 and a 64-bit accumulator absorbs 255 of them on top of a value
// below 2^40: one v_mad_u64_u32 per product and no carry handling. Every
// DigitFlush::blocks blocks the bits from 2^40 up move into the overflow
// word, which so counts units of 2^40 (a column sum stays below 2^70, so
// the count stays below 2^30): sum = ovf * 2^40 + acc.
constexpr uint32_t kDigitBits = 28;
template <int Q>
struct DigitFlush {
    static constexpr uint32_t blocks = 255 / Q;  // Q products per accumulator per block
};

constexpr uint32_t kDigitFlushBit = 40;

template <int Q>
__device__ __forceinline__ void flush(ColGroups<Q>& g) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
        g.ovf[q] += (uint32_t)(g.acc[q] >> kDigitFlushBit);
        g.acc[q] &= (1ull << kDigitFlushBit) - 1;
    }
}

// Column q, a-word u uses window word q - u + Q - 1.
template <int Q, bool DIG = false>
__device__ __forceinline__ void mac_block(const Operands<Q>& o, ColGroups<Q>& g) {
    if (DIG) {
#pragma unroll
        for (int u = 0; u < Q; u++)
#pragma unroll
            for (int q = 0; q < Q; q++) g.acc[q] += (uint64_t)o.a[u] * o.b[q - u + Q - 1];
        return;
    }
    if (Q % 8 == 0) {
#pragma unroll
        for (int u = 0; u < Q; u++)
#pragma unroll
            for (int q0 = 0; q0 < Q; q0 += 8) mac8(o.a[u], &o.b[q0 + Q - 1 - u], &g.acc[q0], &g.ovf[q0]);
        return;
    }
#pragma unroll
    for (int u = 0; u < Q; u++) {
#pragma unroll
        for (int q0 = 0; q0 < Q; q0 += 4) {
            const int x = q0 - u + Q - 1;
            mac4(o.a[u], o.b[x], o.b[x + 1], o.b[x + 2], o.b[x + 3], g.acc[q0], g.acc[q0 + 1], g.acc[q0 + 2],
                 g.acc[q0 + 3], g.ovf[q0], g.ovf[q0 + 1], g.ovf[q0 + 2], g.ovf[q0 + 3]);
        }
    }
}

template <int Q>
__device__ __forceinline__ void park(uint32_t* snap, const ColGroups<Q>& g) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
        snap[3 * q] = (uint32_t)g.acc[q];
        snap[3 * q + 1] = (uint32_t)(g.acc[q] >> 32);
        snap[3 * q + 2] = g.ovf[q];
    }
}

// Parks the running sums as the lane's first column group (registers).
template <int Q>
__device__ __forceinline__ void park_regs(ColGroups<Q>& g) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
        g.sacc[q] = g.acc[q];
        g.sovf[q] = g.ovf[q];
    }
}

// SKEW: W is in the skewed layout (skew_idx). RPARK: the first group is
// parked in registers (sacc/sovf) instead of the LDS snapshot, trading 3Q
// VGPRs and 3Q moves per block for the 3Q words per lane of LDS.
template <int Q, bool PF, bool DIG = false, bool SKEW = false, bool RPARK = false>
__device__ __forceinline__ void conv_lane(const uint32_t* S, const uint32_t* W, uint32_t k1, uint32_t k2,
                                          uint32_t i2, uint32_t sw, uint32_t NB, uint32_t* snap, ColGroups<Q>& g) {
    static_assert(!SKEW || Q >= 8, "the skewed layout needs Q >= 8");
    constexpr uint32_t F = DigitFlush<Q>::blocks;
    constexpr uint32_t step = SKEW ? Q + 4 : Q;  // physical move of the window per block
    auto wpos = [&](uint32_t y) { return W + (SKEW ? skew_idx<Q>(y) : y); };
#pragma unroll
    for (int q = 0; q < Q; q++) {
        g.acc[q] = 0;
        g.ovf[q] = 0;
    }
    if (PF) {
        // Two operand buffers in ping-pong, loop unrolled by two. The load
        // of block beta+1 is issued before the products of block beta. The
        // prefetch past the last block reads one block beyond the lane's
        // range, which stays inside the LDS staging area (pa <= S + ns + 2Q,
        // pb >= W - step >= S); its values are never used.
        const uint32_t* pa = S;
        const uint32_t* pb = wpos(k1);
        Operands<Q> x0, x1;
        load_block<Q, SKEW>(pa, pb, x0);
        for (uint32_t beta = 0; beta < NB; beta += 2) {
            if (beta + 1 == sw) {
                pa = S + i2;
                pb = wpos(k2 - i2);
            } else {
                pa += Q;
                pb -= step;
            }
            load_block<Q, SKEW>(pa, pb, x1);
            if (beta == sw) {
                if (RPARK)
                    park_regs<Q>(g);
                else
                    park<Q>(snap, g);
            }
            mac_block<Q, DIG>(x0, g);
            if (DIG && (beta + 1) % F == 0) flush<Q>(g);
            if (beta + 1 >= NB) break;
            if (beta + 2 == sw) {
                pa = S + i2;
                pb = wpos(k2 - i2);
            } else {
                pa += Q;
                pb -= step;
            }
            load_block<Q, SKEW>(pa, pb, x0);
            if (beta + 1 == sw) {
                if (RPARK)
                    park_regs<Q>(g);
                else
                    park<Q>(snap, g);
            }
            mac_block<Q, DIG>(x1, g);
            if (DIG && (beta + 2) % F == 0) flush<Q>(g);
        }
    } else {
        const uint32_t* pa = S;
        const uint32_t* pb = wpos(k1);
        for (uint32_t beta = 0; beta < NB; beta++) {
            if (beta == sw) {
                if (RPARK)
                    park_regs<Q>(g);
                else
                    park<Q>(snap, g);
                pa = S + i2;
                pb = wpos(k2 - i2);
            }
            Operands<Q> cur;
            load_block<Q, SKEW>(pa, pb, cur);
            mac_block<Q, DIG>(cur, g);
            if (DIG && (beta + 1) % F == 0) flush<Q>(g);
            pa += Q;
            pb -= step;
        }
    }
    if (!RPARK && sw < NB) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            g.sacc[q] = (uint64_t)snap[3 * q] | ((uint64_t)snap[3 * q + 1] << 32);
            g.sovf[q] = snap[3 * q + 2];
        }
    }
}

// Splits the value sum_q (va[q] + vo[q] 2^64) 2^(32q) of a column group at
// word column kcol into Q/2 low limbs written to L[kcol/2 ...] and a high
// part of at most two limbs written to the H slot H[kcol/2 + Q/2 ...] of
// Q/2 limbs (rest of the slot zeroed; limbs at or beyond P dropped).
// H[0..Q/2) of the segment is cleared by the group at column 0.
template <int Q>
__device__ __forceinline__ void group_store(const uint64_t (&va)[Q], const uint32_t (&vo)[Q], uint32_t kcol,
                                            uint64_t* L, uint64_t* H, uint32_t P) {
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
    // the rest of this group's H slot (Q > 4: the slot is Q/2 limbs wide)
#pragma unroll
    for (int i = 2; i < Q / 2; i++)
        if (limb + Q / 2 + i < P) H[limb + Q / 2 + i] = 0;
    if (kcol == 0) {
#pragma unroll
        for (int i = 0; i < Q / 2; i++) H[i] = 0;
    }
}

// Stores the one or two column groups of a lane after conv_lane.
template <int Q>
__device__ __forceinline__ void lane_store(const ColGroups<Q>& g, bool two, uint32_t k1, uint32_t k2, uint64_t* L,
                                           uint64_t* H, uint32_t P) {
    if (two) {
        group_store<Q>(g.sacc, g.sovf, k1, L, H, P);
        uint64_t va[Q];
        uint32_t vo[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            va[q] = g.acc[q] - g.sacc[q];
            vo[q] = g.ovf[q] - g.sovf[q] - (g.acc[q] < g.sacc[q] ? 1u : 0u);
        }
        group_store<Q>(va, vo, k2, L, H, P);
    } else {
        group_store<Q>(g.acc, g.ovf, k1, L, H, P);
    }
}

// Digit mode, Q = 8: the group of columns [kc, kc + 8) holds
// V = sum_q S_q 2^(28 q), S_q = s_hi 2^64 + s_lo (< 2^96), so V < 2^293,
// at bit 28 kc = 224 g (g = kc / 8). Even groups start on a limb (limb
// 3.5 g), odd groups 32 bits into limb 3.5 g - 0.5. Even groups go to
// vector E, odd ones to O; in each vector consecutive groups are 7 limbs
// apart and V spans at most 6 limbs, so every group writes its 7-limb slot
// (value and zero fill) and the slots tile the vector: the product is E + O.
// The writer of group 0 also clears O's first 3 limbs. Limbs >= P dropped.
__device__ __forceinline__ void digit_group_store(const uint64_t (&s_lo)[8], const uint64_t (&s_hi)[8],
                                                  uint32_t kc, uint64_t* E, uint64_t* O, uint32_t P) {
    static_assert(kDigitBits == 28, "slot geometry assumes 28-bit digits and 8 columns");
    uint64_t V[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int o = 28 * q;
        const int l = o >> 6, r = o & 63;
        const uint64_t p0 = s_lo[q] << r;
        const uint64_t p1 = r ? (s_hi[q] << r) | (s_lo[q] >> (64 - r)) : s_hi[q];
        const uint64_t p2 = r ? (s_hi[q] >> (64 - r)) : 0ull;
        // V[l..] += (p2:p1:p0) with carries
        uint64_t x = V[l] + p0;
        uint64_t c = x < p0;
        V[l] = x;
        x = V[l + 1] + p1;
        uint64_t c2 = x < p1;
        x += c;
        c2 += x < c;
        V[l + 1] = x;
        x = V[l + 2] + p2;
        uint64_t c3 = x < p2;
        x += c2;
        c3 += x < c2;
        V[l + 2] = x;
#pragma unroll
        for (int i = l + 3; i < 6; i++) {
            x = V[i] + c3;
            c3 = x < c3;
            V[i] = x;
        }
    }
    const uint32_t g = kc / 8;
    if (g & 1u) {
        const uint32_t base = 7 * (g >> 1) + 3;
        const uint64_t W[7] = {V[0] << 32,
                               (V[1] << 32) | (V[0] >> 32),
                               (V[2] << 32) | (V[1] >> 32),
                               (V[3] << 32) | (V[2] >> 32),
                               (V[4] << 32) | (V[3] >> 32),
                               (V[5] << 32) | (V[4] >> 32),
                               V[5] >> 32};
#pragma unroll
        for (int i = 0; i < 7; i++)
            if (base + i < P) O[base + i] = W[i];
    } else {
        const uint32_t base = 7 * (g >> 1);
#pragma unroll
        for (int i = 0; i < 7; i++)
            if (base + i < P) E[base + i] = i < 6 ? V[i] : 0ull;
        if (g == 0) {
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (i < (int)P) O[i] = 0;
        }
    }
}

// Column sum acc + ovf 2^40 (digit mode) as s_hi 2^64 + s_lo.
__device__ __forceinline__ void digit_canon(uint64_t acc, uint32_t ovf, uint64_t& s_lo, uint64_t& s_hi) {
    const uint64_t t = (acc >> kDigitFlushBit) + (uint64_t)ovf;  // < 2^33
    s_lo = (acc & ((1ull << kDigitFlushBit) - 1)) | (t << kDigitFlushBit);
    s_hi = t >> (64 - kDigitFlushBit);
}

// Stores the one or two digit column groups of a lane after conv_lane; the
// second group is the 96-bit difference of the final and the parked sums.
__device__ __forceinline__ void digit_lane_store(const ColGroups<8>& g, bool two, uint32_t k1, uint32_t k2,
                                                 uint64_t* E, uint64_t* O, uint32_t P) {
    uint64_t fl[8], fh[8];
#pragma unroll
    for (int q = 0; q < 8; q++) digit_canon(g.acc[q], g.ovf[q], fl[q], fh[q]);
    if (!two) {
        digit_group_store(fl, fh, k1, E, O, P);
        return;
    }
    uint64_t pl[8], ph[8];
#pragma unroll
    for (int q = 0; q < 8; q++) digit_canon(g.sacc[q], g.sovf[q], pl[q], ph[q]);
    digit_group_store(pl, ph, k1, E, O, P);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint64_t lo = fl[q] - pl[q];
        fh[q] = fh[q] - ph[q] - (fl[q] < pl[q] ? 1ull : 0ull);
        fl[q] = lo;
    }
    digit_group_store(fl, fh, k2, E, O, P);
}


}  // namespace midint
