// Register-pass Goldilocks NTT for transform lengths L = 2^10 .. 2^14 with
// T = L / V threads, V values per thread (used by ntt.hip).
//
// Stage s of the forward decimation-in-frequency transform pairs x[i] and
// x[i + h], h = L >> (s+1), (i & h) == 0:  x[i] = u + v,
// x[i + h] = (u - v) w^e with e = (i mod h) << s. The inverse is the
// decimation-in-time transform with w^-e, stages in reverse order
// (bit-reversed in, natural out), so no permutation pass exists.
//
// A pass applies R consecutive stages (h from H down to h_min = H >> (R-1))
// to groups of 2^R elements that those stages only mix among themselves:
// group (g_hi, g_lo), g_lo < h_min, holds i = g_hi * 2H + g_lo + k * h_min,
// k < 2^R. Each thread owns V / 2^R groups, g = tid + T * j. Between passes
// the values go through LDS with one padding slot every 32 values (index i
// lives at i + (i >> 5)), which keeps the strided group accesses of every
// pass free of (or at most 2-way) bank conflicts.
//
// The transform is split into passes of RM = log2(V) stages, the last one
// taking what is left: NttPlan<LG, V>.
#pragma once

#include "midint_dev.hpp"

namespace midint {

__device__ __forceinline__ uint32_t ntt_pad(uint32_t i) { return i + (i >> 5); }

// x, with the compiler unable to see its value: per-phase index arithmetic
// derived from it is recomputed instead of being kept live across phases.
__device__ __forceinline__ uint32_t opaque_u32(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Group j of thread tid in a pass with G groups per thread: wave-contiguous,
// g = wave 64 G + 64 j + lane, so that wave w's groups are [64 G w, 64 G (w+1))
// (the same stride between lanes as tid + T j). With this assignment every
// pass after the first touches only the elements whose top log2(T / 64)
// index bits are the wave's index (the first pass's stages span exactly
// those bits): wave w's block of L / (T / 64) consecutive elements. The LDS
// exchanges between those passes are wave-local and need no workgroup
// barrier.
template <int G>
__device__ __forceinline__ uint32_t ntt_group(uint32_t tid, int j) {
    if constexpr (G == 1) {
        return tid;
    } else {
        constexpr int lgG = G == 2 ? 1 : G == 4 ? 2 : G == 8 ? 3 : G == 16 ? 4 : 5;
        static_assert((1 << lgG) == G, "groups per thread");
        return ((tid >> 6) << (6 + lgG)) | ((uint32_t)j << 6) | (tid & 63u);
    }
}

// Orders a wave's LDS stores before its following LDS loads (LDS executes a
// wave's instructions in order; this keeps the compiler from moving loads of
// other lanes' data above the stores).
__device__ __forceinline__ void ntt_wave_sync() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

template <int LG, int V>
struct NttPlan {
    static constexpr int RM = V == 32 ? 5 : 4;  // stages per full pass
    static constexpr int NP = (LG + RM - 1) / RM;
    static constexpr int radix(int p) { return (p + 1) * RM <= LG ? RM : LG - p * RM; }
    static constexpr uint32_t T = (1u << LG) / V;
};

// Radix-4 units of the passes with table twiddles. Two consecutive stages s
// (local span J) and s + 1 (span J / 2) mix the four values at local offsets
// k, k + J/2, k + J, k + 3J/2 of a group. With W = w^E, E the stage-s
// exponent of offset k, the stage-s twiddle of offset k + J/2 is W w^(L/4),
// and w^(L/4) = 2^48 (the root has w^(L/64) = 8). So the forward pair of
// decimation-in-frequency stages is
//   b0 = x0 + x2, b2 = x0 - x2, b1 = x1 + x3, b3 = (x1 - x3) 2^48,
//   y0 = b0 + b1, y1 = (b0 - b1) W^2, y2 = (b2 + b3) W, y3 = (b2 - b3) W^3
// (slots k, k + J/2, k + J, k + 3J/2): three table twiddles and one shift per
// four values instead of four table twiddles. The inverse pair (twiddles
// w^-E, 2^-48 = -2^48) undoes it up to the factor 4 that L^-1 absorbs:
//   Y1 = y1 W^-2, Y2 = y2 W^-1, Y3 = y3 W^-3, c0 = y0 + Y1, c1 = y0 - Y1,
//   c2 = Y2 + Y3, c3 = (Y3 - Y2) 2^48,
//   x0 = c0 + c2, x2 = c0 - c2, x1 = c1 + c3, x3 = c1 - c3.
// The twiddle tables hold w^i and w^-i for i < L (3E can exceed L / 2).
template <bool INV, int LG, bool ZERO_HI, int FM = 0>
__device__ __forceinline__ void ntt_radix4(uint64_t& s0, uint64_t& s1, uint64_t& s2, uint64_t& s3, uint32_t e,
                                           const uint64_t* __restrict__ tw) {
    constexpr uint32_t kMask = (1u << LG) - 1;
    const uint64_t w1 = tw[e], w2 = tw[2 * e], w3 = tw[(3 * e) & kMask];
    if (!INV) {
        const uint64_t x0 = s0, x1 = s1;
        uint64_t b0, b1, b2, b3;
        if (ZERO_HI) {  // x2 = x3 = 0
            b0 = b2 = x0;
            b1 = x1;
            b3 = gl_mul_pow2(x1, 48);
        } else {
            const uint64_t x2 = s2, x3 = s3;
            b0 = gl_add(x0, x2);
            b2 = gl_sub(x0, x2);
            b1 = gl_add(x1, x3);
            b3 = gl_mul_pow2(gl_sub(x1, x3), 48);
        }
        s0 = gl_add(b0, b1);
        s1 = gl_mulv<FM>(gl_sub(b0, b1), w2);
        s2 = gl_mulv<FM>(gl_add(b2, b3), w1);
        s3 = gl_mulv<FM>(gl_sub(b2, b3), w3);
    } else {
        const uint64_t y0 = s0, Y1 = gl_mulv<FM>(s1, w2), Y2 = gl_mulv<FM>(s2, w1), Y3 = gl_mulv<FM>(s3, w3);
        const uint64_t c0 = gl_add(y0, Y1), c1 = gl_sub(y0, Y1);
        const uint64_t c2 = gl_add(Y2, Y3), c3 = gl_mul_pow2(gl_sub(Y3, Y2), 48);
        s0 = gl_add(c0, c2);
        s2 = gl_sub(c0, c2);
        s1 = gl_add(c1, c3);
        s3 = gl_sub(c1, c3);
    }
}

// The stages of one group (values v[base .. base + 2^R)) of a pass with table
// twiddles, as radix-4 units: forward, units at spans (2^jj, 2^(jj-1)) for
// jj = R-1, R-3, ..., then the span-1 stage alone if R is odd; the inverse
// runs the mirror order. U counts stage positions; the recursion keeps every
// index a compile-time constant.
template <int R, bool INV, int V, int LG, int S0, bool ZHI, int FM, int U>
struct NttUnits {
    static __device__ __forceinline__ void run(uint64_t (&v)[V], int base, uint32_t g_lo,
                                               const uint64_t* __restrict__ tw) {
        constexpr int K = 1 << R;
        constexpr int lgmin = (LG - 1 - S0) - (R - 1);
        constexpr bool odd = (R & 1) != 0;
        // forward position U covers jj = R-1-U; inverse position U covers jj = U
        constexpr bool single = odd && (INV ? U == 0 : U == R - 1);
        constexpr bool pair = INV ? ((odd ? (U & 1) == 1 : (U & 1) == 0) && U + 1 < R) : ((U & 1) == 0 && U + 1 < R);
        if constexpr (single) {
            constexpr int s = S0 + (R - 1);  // span h_min
            const uint64_t wt = tw[g_lo << s];
#pragma unroll
            for (int k = 0; k < K; k += 2) {
                uint64_t& a = v[base + k];
                uint64_t& b = v[base + k + 1];
                if constexpr (!INV) {
                    const uint64_t x = a, y = b;
                    a = gl_add(x, y);
                    b = gl_mulv<FM>(gl_sub(x, y), wt);
                } else {
                    const uint64_t x = a, y = gl_mulv<FM>(b, wt);
                    a = gl_add(x, y);
                    b = gl_sub(x, y);
                }
            }
        } else if constexpr (pair) {
            constexpr int big = INV ? U + 1 : R - 1 - U;  // spans 2^big and 2^(big-1)
            constexpr int s = S0 + (R - 1 - big);
            constexpr int J = 1 << big, J2 = J >> 1;
            constexpr bool zhi = !INV && ZHI && S0 == 0 && s == 0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (k & (J | J2)) continue;
                const uint32_t klo = (uint32_t)(k & (J2 - 1));
                const uint32_t e = (g_lo + (klo << lgmin)) << s;
                ntt_radix4<INV, LG, zhi, FM>(v[base + k], v[base + k + J2], v[base + k + J], v[base + k + J + J2], e, tw);
            }
        }
        NttUnits<R, INV, V, LG, S0, ZHI, FM, U + 1>::run(v, base, g_lo, tw);
    }
};
template <int R, bool INV, int V, int LG, int S0, bool ZHI, int FM>
struct NttUnits<R, INV, V, LG, S0, ZHI, FM, R> {
    static __device__ __forceinline__ void run(uint64_t (&)[V], int, uint32_t, const uint64_t* __restrict__) {}
};

// One pass of R stages starting at stage S0 on the thread's registers v[V]
// (V / 2^R groups of 2^R values, group j at v[j << R ...]). In the last
// pass h_min = 1, so g_lo = 0 and every twiddle index is a compile-time
// constant; the butterflies whose index is 0 (twiddle 1) skip the multiply.
// The other passes run their stages as radix-4 units (ntt_radix4): pairs
// from the largest span down in the forward direction, a leftover single
// stage of span 1 last; the inverse mirrors that order.
// ZHI (forward, first pass): the upper half of the input is zero, so the
// first stage is a = u, b = u w.
template <int R, bool INV, int V, int LG, int S0, bool ZHI = false, int FM = 0>
__device__ __forceinline__ void ntt_pass_regs(uint64_t (&v)[V], uint32_t tid, const uint64_t* __restrict__ tw) {
    constexpr int G = V >> R;  // groups per thread
    constexpr int K = 1 << R;  // values per group
    constexpr int lgH = LG - 1 - S0;      // H = 2^lgH
    constexpr int lgmin = lgH - (R - 1);  // h_min
    if constexpr (lgmin > 0) {
#pragma unroll
        for (int j = 0; j < G; j++) {
            const uint32_t g = ntt_group<G>(tid, j);
            const uint32_t g_lo = g & ((1u << lgmin) - 1);
            NttUnits<R, INV, V, LG, S0, ZHI, FM, 0>::run(v, j * K, g_lo, tw);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < G; j++) {
        const uint32_t g = ntt_group<G>(tid, j);
        const uint32_t g_lo = g & ((1u << lgmin) - 1);
#pragma unroll
        for (int st = 0; st < R; st++) {
            // forward: jj = R-1 .. 0 (h = H .. h_min); inverse: jj = 0 .. R-1
            const int jj = INV ? st : (R - 1 - st);
            const int s = S0 + (R - 1 - jj);  // stage index of h = h_min << jj
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (k & (1 << jj)) continue;
                const uint32_t klo = (uint32_t)(k & ((1 << jj) - 1));
                const bool unit = lgmin == 0 && klo == 0;  // w^0
                // last pass (h_min = 1): the twiddle of span h = 2^lgh is
                // w^(klo L / 2h) = 2^(sh), sh = (96 / h) klo (the root has
                // w^(L/64) = 8), a shift instead of a table multiply
                const int lgh = LG - 1 - s;
                const int sh = lgmin == 0 && lgh <= 5 ? (96 >> lgh) * (int)klo : -1;
                uint64_t& a = v[j * K + k];
                uint64_t& b = v[j * K + k + (1 << jj)];
                if (!INV) {
                    const uint64_t u = a, x = b;
                    const bool zero_x = ZHI && S0 == 0 && s == 0;
                    a = zero_x ? u : gl_add(u, x);
                    const uint64_t d = zero_x ? u : gl_sub(u, x);
                    b = unit           ? d
                        : sh > 0 && sh < 64 ? gl_mul_pow2(d, sh)
                                            : gl_mulv<FM>(d, tw[(g_lo + (klo << lgmin)) << s]);
                } else if (sh > 32) {
                    // w^-e = 2^(192 - sh) = -2^(96 - sh): swap the add and the sub
                    const uint64_t u = a, y = gl_mul_pow2(b, 96 - sh);
                    a = gl_sub(u, y);
                    b = gl_add(u, y);
                } else {
                    const uint64_t u = a, x = unit ? b : gl_mulv<FM>(b, tw[(g_lo + (klo << lgmin)) << s]);
                    a = gl_add(u, x);
                    b = gl_sub(u, x);
                }
            }
        }
    }
}

// Element index of value k of the thread's group j in the pass starting at
// stage s0 with R stages: B_j + k * h_min with B_j = g_hi * 2H + g_lo.
template <int R>
__device__ __forceinline__ uint32_t ntt_group_index(uint32_t lg, uint32_t s0, uint32_t tid, uint32_t T, int j,
                                                    int k) {
    const uint32_t lgH = lg - 1 - s0;
    const uint32_t lgmin = lgH - (R - 1);
    const uint32_t g = tid + T * j;
    const uint32_t g_lo = g & ((1u << lgmin) - 1);
    const uint32_t g_hi = g >> lgmin;
    return (g_hi << (lgH + 1)) + g_lo + ((uint32_t)k << lgmin);
}

// Padded LDS slot of that element. The padding is linear inside a group:
// pad(B + C) = pad(B) + C + (C >> 5) for C = k * h_min, because the low five
// bits of B and C never carry (g_lo < h_min; when h_min < 32 the group lies
// in an aligned 2H-block or 2H >= 32 and B is 32-aligned apart from g_lo).
// So each group needs one base register and the per-k part is a
// compile-time offset.
template <int R, int V, int LG, int S0>
__device__ __forceinline__ void ntt_load_pass(const uint64_t* x, uint64_t (&v)[V], uint32_t tid) {
    constexpr int lgH = LG - 1 - S0;
    constexpr int lgmin = lgH - (R - 1);
#pragma unroll
    for (int j = 0; j < (V >> R); j++) {
        const uint32_t g = ntt_group<(V >> R)>(tid, j);
        const uint64_t* xb = x + ntt_pad(((g >> lgmin) << (lgH + 1)) + (g & ((1u << lgmin) - 1)));
#pragma unroll
        for (int k = 0; k < (1 << R); k++) {
            const uint32_t c = (uint32_t)k << lgmin;
            v[(j << R) + k] = xb[c + (c >> 5)];
        }
    }
}

template <int R, int V, int LG, int S0>
__device__ __forceinline__ void ntt_store_pass(uint64_t* x, const uint64_t (&v)[V], uint32_t tid) {
    constexpr int lgH = LG - 1 - S0;
    constexpr int lgmin = lgH - (R - 1);
#pragma unroll
    for (int j = 0; j < (V >> R); j++) {
        const uint32_t g = ntt_group<(V >> R)>(tid, j);
        uint64_t* xb = x + ntt_pad(((g >> lgmin) << (lgH + 1)) + (g & ((1u << lgmin) - 1)));
#pragma unroll
        for (int k = 0; k < (1 << R); k++) {
            const uint32_t c = (uint32_t)k << lgmin;
            xb[c + (c >> 5)] = v[(j << R) + k];
        }
    }
}

// Forward passes p .. NP-1 starting from values of pass p's grouping in v;
// returns with the last pass's values in v.
template <int LG, int V, int P, bool ZHI = false, int FM = 0>
struct NttFwd {
    __device__ static __forceinline__ void run(uint64_t* x, uint64_t (&v)[V], uint32_t tid,
                                               const uint64_t* __restrict__ twf) {
        using Pl = NttPlan<LG, V>;
        constexpr int R = Pl::radix(P);
        ntt_pass_regs<R, false, V, LG, P * Pl::RM, ZHI && P == 0, FM>(v, tid, twf);
        if constexpr (P + 1 < Pl::NP) {
            constexpr int R1 = Pl::radix(P + 1);
            ntt_store_pass<R, V, LG, P * Pl::RM>(x, v, tid);
            if constexpr (P >= 1) {
                ntt_wave_sync();  // passes P and P + 1 both within the wave's block
            } else {
                __syncthreads();
            }
            ntt_load_pass<R1, V, LG, (P + 1) * Pl::RM>(x, v, tid);
            NttFwd<LG, V, P + 1, false, FM>::run(x, v, tid, twf);
        }
    }
};

// Inverse passes p .. 0 starting from values of pass p's grouping in v; ends
// with the natural-order result stored in LDS (padded) and a barrier. The
// passes after the first touch only the wave's own block (ntt_group), so
// the first store needs no barrier after a forward transform's last pass.
template <int LG, int V, int P, int FM = 0>
struct NttInv {
    __device__ static __forceinline__ void run(uint64_t* x, uint64_t (&v)[V], uint32_t tid,
                                               const uint64_t* __restrict__ twi) {
        using Pl = NttPlan<LG, V>;
        constexpr int R = Pl::radix(P);
        ntt_pass_regs<R, true, V, LG, P * Pl::RM, false, FM>(v, tid, twi);
        ntt_store_pass<R, V, LG, P * Pl::RM>(x, v, tid);
        if constexpr (P >= 2) {
            ntt_wave_sync();  // passes P and P - 1 both within the wave's block
        } else {
            __syncthreads();
        }
        if constexpr (P > 0) {
            constexpr int R1 = Pl::radix(P - 1);
            ntt_load_pass<R1, V, LG, (P - 1) * Pl::RM>(x, v, tid);
            NttInv<LG, V, P - 1, FM>::run(x, v, tid, twi);
        }
    }
};

}  // namespace midint
