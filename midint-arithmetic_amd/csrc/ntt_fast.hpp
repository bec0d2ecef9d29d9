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

template <int LG, int V>
struct NttPlan {
    static constexpr int RM = V == 32 ? 5 : 4;  // stages per full pass
    static constexpr int NP = (LG + RM - 1) / RM;
    static constexpr int radix(int p) { return (p + 1) * RM <= LG ? RM : LG - p * RM; }
    static constexpr uint32_t T = (1u << LG) / V;
};

// One pass of R stages starting at stage S0 on the thread's registers v[V]
// (V / 2^R groups of 2^R values, group j at v[j << R ...]). In the last
// pass h_min = 1, so g_lo = 0 and every twiddle index is a compile-time
// constant; the butterflies whose index is 0 (twiddle 1) skip the multiply.
// ZHI (forward, first pass): the upper half of the input is zero, so the
// first stage is a = u, b = u w.
template <int R, bool INV, int V, int LG, int S0, bool ZHI = false>
__device__ __forceinline__ void ntt_pass_regs(uint64_t (&v)[V], uint32_t tid, const uint64_t* __restrict__ tw) {
    constexpr int G = V >> R;  // groups per thread
    constexpr int K = 1 << R;  // values per group
    constexpr uint32_t T = (1u << LG) / V;
    constexpr int lgH = LG - 1 - S0;      // H = 2^lgH
    constexpr int lgmin = lgH - (R - 1);  // h_min
#pragma unroll
    for (int j = 0; j < G; j++) {
        const uint32_t g = tid + T * j;
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
                                            : gl_mul(d, tw[(g_lo + (klo << lgmin)) << s]);
                } else if (sh > 32) {
                    // w^-e = 2^(192 - sh) = -2^(96 - sh): swap the add and the sub
                    const uint64_t u = a, y = gl_mul_pow2(b, 96 - sh);
                    a = gl_sub(u, y);
                    b = gl_add(u, y);
                } else {
                    const uint64_t u = a, x = unit ? b : gl_mul(b, tw[(g_lo + (klo << lgmin)) << s]);
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
    constexpr uint32_t T = (1u << LG) / V;
    constexpr int lgH = LG - 1 - S0;
    constexpr int lgmin = lgH - (R - 1);
#pragma unroll
    for (int j = 0; j < (V >> R); j++) {
        const uint32_t g = tid + T * j;
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
    constexpr uint32_t T = (1u << LG) / V;
    constexpr int lgH = LG - 1 - S0;
    constexpr int lgmin = lgH - (R - 1);
#pragma unroll
    for (int j = 0; j < (V >> R); j++) {
        const uint32_t g = tid + T * j;
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
template <int LG, int V, int P, bool ZHI = false>
struct NttFwd {
    __device__ static __forceinline__ void run(uint64_t* x, uint64_t (&v)[V], uint32_t tid,
                                               const uint64_t* __restrict__ twf) {
        using Pl = NttPlan<LG, V>;
        constexpr int R = Pl::radix(P);
        ntt_pass_regs<R, false, V, LG, P * Pl::RM, ZHI && P == 0>(v, tid, twf);
        if constexpr (P + 1 < Pl::NP) {
            constexpr int R1 = Pl::radix(P + 1);
            ntt_store_pass<R, V, LG, P * Pl::RM>(x, v, tid);
            __syncthreads();
            ntt_load_pass<R1, V, LG, (P + 1) * Pl::RM>(x, v, tid);
            NttFwd<LG, V, P + 1>::run(x, v, tid, twf);
        }
    }
};

// Inverse passes p .. 0 starting from values of pass p's grouping in v; ends
// with the natural-order result stored in LDS (padded) and a barrier.
template <int LG, int V, int P>
struct NttInv {
    __device__ static __forceinline__ void run(uint64_t* x, uint64_t (&v)[V], uint32_t tid,
                                               const uint64_t* __restrict__ twi) {
        using Pl = NttPlan<LG, V>;
        constexpr int R = Pl::radix(P);
        ntt_pass_regs<R, true, V, LG, P * Pl::RM>(v, tid, twi);
        ntt_store_pass<R, V, LG, P * Pl::RM>(x, v, tid);
        __syncthreads();
        if constexpr (P > 0) {
            constexpr int R1 = Pl::radix(P - 1);
            ntt_load_pass<R1, V, LG, (P - 1) * Pl::RM>(x, v, tid);
            NttInv<LG, V, P - 1>::run(x, v, tid, twi);
        }
    }
};

}  // namespace midint
