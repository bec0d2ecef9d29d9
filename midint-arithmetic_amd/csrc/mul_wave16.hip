// Batched classical multiplication, one product per wave, on the int8
// matrix cores' 16 x 16 x 64 instruction, specialised at compile time for the
// operand size (M limbs). The reference's kernel for this operation is the
// column-sum schoolbook of cuda/ker-classic-mul.cu.h:299-463 (one block per
// instance, u32 or u64 words); digits here are balanced base 256
// (mfma_digits.hpp), D = 8M + 1 per operand.
//
// Tiling. Output digit p = 16I + r (block I of 16 digits, r < 16). With
// windows of 64 digits of a that start every 16 digits and an offset
// DL = 48,
//
//     c_{16I+r} = sum_k sum_{e<64} b_{64k+DL+r-e} a_{16(I-4k)-DL+e}:
//
// every term has i + j = 16I + r, and for fixed I the windows
// [16(I-4k) - DL, +64) over all k tile the integers, so each digit pair
// appears once. For a tile of 16 consecutive blocks I = 16t + n and a step k
// that is one v_mfma_i32_16x16x64_i8: A = T_k (16 x 64, T_k[r][e] =
// b_{64k+DL+r-e}, the same for every column), B = the 16 windows of the
// tile (64 x 16), D accumulates the tile's 16 x 16 column sums. The window
// of column n at step k starts at digit 16(n + 4u) - DL with u = 4t - k, so
// one B fragment serves every (t, k) with the same u. The (t, k) pairs with
// a nonzero term are found at compile time (w16_plan); for 8192-bit
// operands that is 84 instructions per full product (issued / useful digit
// products 1.31; 32 x 32 tiles aligned on 32-digit blocks need 66
// instructions of twice the size).
//
// Operand layout of a lane (l = n + 16 g, n < 16, g < 4). The hardware pairs
// byte e of the A operand of lane (r, g) with byte e of the B operand of
// lane (n, g) for the same k index, so any assignment of the 64 k indices to
// (g, e) that both operands share gives the same sum; here k = 16 g + e.
// A: the 16 bytes b_{64k+DL+r-16g-e}, e < 16, read from b's digits stored
// in reverse (br[y] = b_digit[X0 - y]) as five aligned dwords and four
// funnel shifts (the misalignment depends on r only). B: the 16 bytes
// a_{16(n+g+4u)-DL+e}, one aligned 16-byte LDS read. All 17 T_k of an
// 8192-bit product stay in registers; the 20 B fragments stream from LDS.
//
// Column sums to limbs. Lane (n, g) of tile t holds the sums of output
// digits 16I + 4g + v (v < 4), I = 16t + n: word w = 4I + g of the product,
// V_w = sum_v (c + 255 B) 2^(8v) < 2^55 (four v_mad_i64_i32; one B more in
// column 0), kept in registers: cross-lane swaps move the four words of
// lane l's two limbs into lane l (no LDS round trip). Limb L of the product
// is W_L = V_{2L} + V_{2L+1} 2^32 < 2^88: its low 64 bits, plus the bits
// above 64 of W_{L-1}, one pass of the row adder over the wave's rows.
//
// The same machinery runs the reference's polyKerQ composite (four low-half
// products and three additions per instance, poly_wave16_kernel) with the
// intermediates in registers.
//
// No barrier: every wave owns its LDS region (digits of a, reversed digits
// of b) and one product at a time; LDS operations of a wave execute in
// order.
#include "add_core.hpp"
#include "mfma_digits.hpp"
#include "midint_internal.h"

namespace midint {

namespace {

constexpr int kW16Delta = 48;    // DL
constexpr int kW16MaxTiles = 16;  // tiles of the largest specialisation

// Which steps k each tile needs.
struct W16Plan {
    int kmin[kW16MaxTiles], kmax[kW16MaxTiles];
    int khi;         // largest k of any tile (the smallest is 0)
    int umin, umax;  // range of u = 4t - k
    int mfma;        // instructions per product
};

// Tile t has a nonzero term at step k iff some column I = 16t + n has a
// window digit i in [0, D) and a digit j in [0, D) with i + j in
// [16I, 16I + 15]: with i in [ilo, ihi] (the window clipped to [0, D)),
// the sums i + j cover [ilo, ihi + D - 1].
constexpr W16Plan w16_plan(int M, bool full) {
    W16Plan P{};
    const int D = 8 * M + 1;
    const int nblk = full ? M : M / 2;
    const int T = nblk / 16;
    P.khi = 0;
    P.umin = 1 << 30;
    P.umax = -(1 << 30);
    P.mfma = 0;
    for (int t = 0; t < T; t++) {
        P.kmin[t] = 1 << 30;
        P.kmax[t] = -(1 << 30);
        for (int k = -8; k <= D / 64 + 8; k++) {
            bool live = false;
            for (int n = 0; n < 16 && !live; n++) {
                const int I = 16 * t + n;
                const int L = 16 * I - 64 * k - kW16Delta;
                const int ilo = L > 0 ? L : 0;
                const int ihi = L + 63 < D - 1 ? L + 63 : D - 1;
                live = ilo <= ihi && ilo <= 16 * I + 15 && ihi + D - 1 >= 16 * I;
            }
            if (live) {
                if (k < P.kmin[t]) P.kmin[t] = k;
                if (k > P.kmax[t]) P.kmax[t] = k;
            }
        }
        for (int k = P.kmin[t]; k <= P.kmax[t]; k++) {
            const int u = 4 * t - k;
            if (u < P.umin) P.umin = u;
            if (u > P.umax) P.umax = u;
        }
        if (P.kmax[t] > P.khi) P.khi = P.kmax[t];
        P.mfma += P.kmax[t] - P.kmin[t] + 1;
    }
    return P;
}

constexpr int w16_round16(int x) { return (x + 15) & ~15; }

// Per-wave LDS layout (bytes) of a specialisation.
template <int M, bool FULL>
struct W16 {
    static constexpr W16Plan P = w16_plan(M, FULL);
    static constexpr int NBLK = FULL ? M : M / 2;  // 16-digit output blocks
    static constexpr int PL = 2 * NBLK;            // output limbs
    static constexpr int T = NBLK / 16;            // tiles
    static constexpr int ROWS_IN = (M + 127) / 128;
    static constexpr int ROWS_OUT = (PL + 127) / 128;
    // a region: digit i at byte AOFF + i; the windows read bytes
    // [AOFF + 16 (n + g) + 64 u - DL, +16); zeros outside the digits
    static constexpr int AOFF = kW16Delta - 64 * P.umin;
    static constexpr int A_END0 = AOFF + 16 * 18 + 64 * P.umax - kW16Delta + 16;
    static constexpr int A_BYTES = w16_round16(A_END0 > AOFF + 8 * M + 8 ? A_END0 : AOFF + 8 * M + 8);
    // br region (first in the wave's region, so that the A fragments' reads
    // of step k sit within the 1020-byte offset fields of ds_read2_b32 from
    // one base register): b's digit j at byte X0 - j; the A fragments read
    // five dwords from dword (X0 - 64k - DL - r + 16g) / 4; X0 = 15 mod 16
    // puts b's limb pairs on 16-byte boundaries
    static constexpr int X0_MIN0 = 64 * P.khi + kW16Delta + 15;
    static constexpr int X0_MIN = X0_MIN0 > 8 * M + 7 ? X0_MIN0 : 8 * M + 7;
    static constexpr int X0 = ((X0_MIN - 15 + 15) & ~15) + 15;
    static constexpr int BR_BYTES = w16_round16((X0 & ~3) + 20);
    static constexpr int WAVE_BYTES = A_BYTES + BR_BYTES;
    static_assert(T >= 1 && T <= kW16MaxTiles, "tile count");
    static_assert(AOFF % 16 == 0 && AOFF >= 0, "a region alignment");
    static_assert(X0 - 64 * P.khi - kW16Delta - 15 >= 0, "br region start");
    static_assert(X0 % 16 == 15, "br alignment");
};

constexpr uint32_t kW16Waves = 4;  // waves (independent products) per workgroup
// Workgroups per CU the register allocation must allow: 4 waves per SIMD
// (at most 128 VGPRs; with fewer waves the SIMDs idle on LDS latency)
constexpr uint32_t kW16MinBlocks = 4;

}  // namespace

// One wave's product machinery: its LDS region (digits of a, reversed digits
// of b, the words V) and the lane constants of the fragments.
template <int M, bool FULL>
struct W16Wave {
    using G = W16<M, FULL>;
    static constexpr W16Plan P = G::P;
    uint8_t* areg;
    uint8_t* breg;
    const uint8_t* pa;  // A fragment of step k: dwords from pa + 64 (khi - k)
    const uint8_t* pb;  // B fragment u: 16 bytes at pb + 64 (u - umin)
    uint32_t lane, sh;
    MfPow pw;

    // The region at `base`: zeros around the digits and before the words
    // (written once: the products only overwrite digit bytes and words).
    __device__ __forceinline__ void init(uint8_t* base, uint32_t pow8) {
        lane = threadIdx.x & 63;
        breg = base;
        areg = base + G::BR_BYTES;
        for (uint32_t x = 16 * lane; x < (uint32_t)G::WAVE_BYTES; x += 1024)
            *reinterpret_cast<uint4*>(base + x) = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t ln = lane & 15, lg = lane >> 4;
        const uint32_t yb = (uint32_t)(G::X0 - 64 * P.khi - kW16Delta) - ln + 16 * lg;
        sh = 8 * (yb & 3u);
        pa = breg + (yb & ~3u);
        // AOFF + 16 (n + g) + 64 u - DL = 16 (n + g) + 64 (u - umin)
        pb = areg + 16 * (ln + lg);
        pw = mf_pow(pow8);
    }

    // Digits of the operands held as rows (lane: limbs 128 ro + 2 lane, +1;
    // limbs at or above M are ignored): x = limbs + 0x8080...80 row by row
    // (the carry between rows in a scalar), digit bytes x ^ 0x80; a in
    // order, b reversed with byte-swapped limbs; the carry out is digit 8M.
    __device__ __forceinline__ void stage(const uint64_t (&xa)[G::ROWS_IN][2], const uint64_t (&xb)[G::ROWS_IN][2]) {
        uint32_t ca = 0, cb = 0;
#pragma unroll
        for (int ro = 0; ro < G::ROWS_IN; ro++) {
            const uint32_t j0 = 128 * ro + 2 * lane;
            const bool in = j0 < (uint32_t)M;  // M even: both limbs or none
            const uint64_t h = in ? kMfHalf : 0ull;
            uint64_t w[2][2];
#pragma unroll
            for (int o = 0; o < 2; o++) {
                // (limbs at or above M only exist when M is not a multiple of 128)
                const uint64_t x0 = (M % 128 == 0 || in) ? (o ? xb[ro][0] : xa[ro][0]) : 0ull;
                const uint64_t x1 = (M % 128 == 0 || in) ? (o ? xb[ro][1] : xa[ro][1]) : 0ull;
                const RowState st = row_build(x0, h, x1, h, false, false);
                const uint64_t C = row_chain(st, o ? &cb : &ca);
                row_apply(st, C, &w[o][0], &w[o][1]);
                w[o][0] ^= h;
                w[o][1] ^= h;
            }
            if (in) {
                *reinterpret_cast<ulonglong2*>(areg + G::AOFF + 8 * j0) = make_ulonglong2(w[0][0], w[0][1]);
                *reinterpret_cast<ulonglong2*>(breg + G::X0 - 8 * j0 - 15) =
                    make_ulonglong2(__builtin_bswap64(w[1][1]), __builtin_bswap64(w[1][0]));
            } else if (j0 == (uint32_t)M) {
                // the limb after the operand holds its carry (the rows end
                // inside a row when M is not a multiple of 128)
                *reinterpret_cast<uint64_t*>(areg + G::AOFF + 8 * M) = w[0][0];
                *reinterpret_cast<uint64_t*>(breg + G::X0 - 8 * M - 7) = __builtin_bswap64(w[1][0]);
            }
        }
        if (M % 128 == 0 && lane == 0) {
            *reinterpret_cast<uint64_t*>(areg + G::AOFF + 8 * M) = (uint64_t)ca;
            *reinterpret_cast<uint64_t*>(breg + G::X0 - 8 * M - 7) = __builtin_bswap64((uint64_t)cb);
        }
        asm volatile("" ::: "memory");
    }

    // The matrix-core phase of the staged operands and the words V_w: lane
    // (n, g) of tile t ends with word w = 64 t + 4 n + g in V[t].
    __device__ __forceinline__ void product(int64_t (&Vt)[G::T]) {
        mf_v4i A[P.khi + 1];
#pragma unroll
        for (int k = 0; k <= P.khi; k++) {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(pa + 64 * (P.khi - k));
            const uint32_t q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
            A[k] = mf_v4i{(int)__builtin_amdgcn_alignbit(q1, q0, sh), (int)__builtin_amdgcn_alignbit(q2, q1, sh),
                          (int)__builtin_amdgcn_alignbit(q3, q2, sh), (int)__builtin_amdgcn_alignbit(q4, q3, sh)};
        }
        mf_v4i acc[G::T];
#pragma unroll
        for (int t = 0; t < G::T; t++) acc[t] = mf_v4i{0, 0, 0, 0};
        // B fragments two steps ahead of their MFMAs (a ring of three): the
        // LDS latency of a fragment hides behind two steps of MFMAs (left
        // to itself the compiler loads each fragment just before its MFMAs
        // and waits for it)
        mf_v4i bq[3];
        bq[0] = *reinterpret_cast<const mf_v4i*>(pb);
        bq[1] = *reinterpret_cast<const mf_v4i*>(pb + 64);
#pragma unroll
        for (int u = P.umin; u <= P.umax; u++) {
            const int i = u - P.umin;
            if (u + 2 <= P.umax) bq[(i + 2) % 3] = *reinterpret_cast<const mf_v4i*>(pb + 64 * (i + 2));
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < G::T; t++) {
                const int k = 4 * t - u;
                if (k >= P.kmin[t] && k <= P.kmax[t])
                    acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[k], bq[i % 3], acc[t], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // column sums to words V_w (the one more B of column 0 is added in
        // finish: a lane-dependent constant here stays live across the loop)
#pragma unroll
        for (int t = 0; t < G::T; t++) {
            int64_t V = (int64_t)kMfFoldBias;
            V += (int64_t)acc[t][0] * (int32_t)pw.p0;
            V += (int64_t)acc[t][1] * (int32_t)pw.p8;
            V += (int64_t)acc[t][2] * (int32_t)pw.p16;
            V += (int64_t)acc[t][3] * (int32_t)pw.p24;
            Vt[t] = V;
        }
    }

    // Limbs of the product from the words: x_L = W_L mod 2^64, y_L = W_{L-1}
    // >> 64 (W_L = V_{2L} + V_{2L+1} 2^32), x + y, as rows (x0, x1; limbs at
    // or above PL are not defined); row ro's words come from its four tiles
    // through cross-lane swaps (mf_row_words), the zero words V_-2, V_-1
    // before row 0.
    __device__ __forceinline__ void finish(const int64_t (&Vt)[G::T], Rows<G::ROWS_OUT>& rows) {
        uint32_t prev[4] = {0u, 0u, 0u, 0u};  // words V_-2, V_-1 before row 0
#pragma unroll
        for (int ro = 0; ro < G::ROWS_OUT; ro++) {
            uint64_t S[4], w[4], p[2];
#pragma unroll
            for (int q = 0; q < 4; q++) S[q] = 4 * ro + q < G::T ? (uint64_t)Vt[4 * ro + q] : 0ull;
            mf_row_words(S, w, p, prev);
            if (ro == 0 && lane == 0) w[0] += kMfBias;  // V_0: column 0 takes 256 B
            const uint64_t sp = p[1] + (p[0] >> 32);
            const uint64_t s0 = w[1] + (w[0] >> 32);
            const uint64_t s1 = w[3] + (w[2] >> 32);
            rows.x0[ro] = (w[0] & 0xffffffffull) | (s0 << 32);
            rows.x1[ro] = (w[2] & 0xffffffffull) | (s1 << 32);
            rows.y0[ro] = sp >> 32;
            rows.y1[ro] = s0 >> 32;
            rows.s0[ro] = false;
            rows.s1[ro] = false;
        }
        wave_add_rows<G::ROWS_OUT>(rows);
    }
};

// Limbs 128 ro + 2 lane, +1 of an M-limb operand (zeros at or above M).
template <int M, int ROWS>
__device__ __forceinline__ void w16_load(const uint64_t* __restrict__ g, uint64_t (&x)[ROWS][2]) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int ro = 0; ro < ROWS; ro++) {
        const uint32_t j0 = 128 * ro + 2 * lane;
        const bool v = j0 < (uint32_t)M;  // M even: both limbs or none
        x[ro][0] = v ? g[j0] : 0ull;
        x[ro][1] = v ? g[j0 + 1] : 0ull;
    }
}

template <int M, bool FULL>
__global__ __launch_bounds__(64 * kW16Waves, kW16MinBlocks) void mul_wave16_kernel(uint32_t n_inst, uint32_t pow8,
                                                                    const uint64_t* __restrict__ a, uint64_t sa,
                                                                    const uint64_t* __restrict__ b, uint64_t sb,
                                                                    uint64_t* __restrict__ r, uint64_t sr) {
    using G = W16<M, FULL>;
    extern __shared__ __align__(16) uint64_t lds64[];
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    W16Wave<M, FULL> w;
    w.init(reinterpret_cast<uint8_t*>(lds64) + (size_t)wv * G::WAVE_BYTES, pow8);
    const uint64_t stride = (uint64_t)gridDim.x * kW16Waves;
    for (uint64_t inst = (uint64_t)blockIdx.x * kW16Waves + wv; inst < n_inst; inst += stride) {
        uint64_t xa[G::ROWS_IN][2], xb[G::ROWS_IN][2];
        w16_load<M, G::ROWS_IN>(a + inst * sa, xa);
        w16_load<M, G::ROWS_IN>(b + inst * sb, xb);
        w.stage(xa, xb);
        int64_t V[G::T];
        w.product(V);
        Rows<G::ROWS_OUT> rows;
        w.finish(V, rows);
        uint64_t* gr = r + inst * sr;
#pragma unroll
        for (int ro = 0; ro < G::ROWS_OUT; ro++) {
            const uint32_t j0 = 128 * ro + 2 * w.lane;
            if (j0 < (uint32_t)G::PL) {
                gr[j0] = rows.x0[ro];
                gr[j0 + 1] = rows.x1[ro];
            }
        }
        asm volatile("" ::: "memory");
    }
}

// The reference's polyKerQ (cuda/ker-classic-mul.cu.h:500-575), one instance
// per wave: r = ((a^2 + b)(b^2 + b) + ab) mod B^M, every step mod B^M in the
// reference's order. The four low-half products run through the wave's LDS
// region one after another; the operands, the intermediates and the sums
// stay in registers as rows.
template <int M>
__global__ __launch_bounds__(64 * kW16Waves, kW16MinBlocks) void poly_wave16_kernel(uint32_t n_inst, uint32_t pow8,
                                                                     const uint64_t* __restrict__ a, uint64_t sa,
                                                                     const uint64_t* __restrict__ b, uint64_t sb,
                                                                     uint64_t* __restrict__ r, uint64_t sr) {
    using G = W16<M, false>;
    static_assert(G::ROWS_IN == G::ROWS_OUT, "low-half products: rows in = rows out");
    constexpr int R = G::ROWS_IN;
    extern __shared__ __align__(16) uint64_t lds64[];
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    W16Wave<M, false> w;
    w.init(reinterpret_cast<uint8_t*>(lds64) + (size_t)wv * G::WAVE_BYTES, pow8);
    // x + y mod B^M of two row sets (limbs at or above M are ignored by every
    // consumer, so the carry past limb M - 1 does not matter)
    auto add_rows = [](const uint64_t (&x)[R][2], const uint64_t (&y)[R][2], uint64_t (&z)[R][2]) {
        Rows<R> t;
#pragma unroll
        for (int ro = 0; ro < R; ro++) {
            t.x0[ro] = x[ro][0];
            t.x1[ro] = x[ro][1];
            t.y0[ro] = y[ro][0];
            t.y1[ro] = y[ro][1];
            t.s0[ro] = false;
            t.s1[ro] = false;
        }
        wave_add_rows<R>(t);
#pragma unroll
        for (int ro = 0; ro < R; ro++) {
            z[ro][0] = t.x0[ro];
            z[ro][1] = t.x1[ro];
        }
    };
    auto mul = [&](const uint64_t (&x)[R][2], const uint64_t (&y)[R][2], uint64_t (&z)[R][2]) {
        w.stage(x, y);
        int64_t V[G::T];
        w.product(V);
        Rows<R> t;
        w.finish(V, t);
#pragma unroll
        for (int ro = 0; ro < R; ro++) {
            z[ro][0] = t.x0[ro];
            z[ro][1] = t.x1[ro];
        }
    };
    const uint64_t stride = (uint64_t)gridDim.x * kW16Waves;
    for (uint64_t inst = (uint64_t)blockIdx.x * kW16Waves + wv; inst < n_inst; inst += stride) {
        uint64_t xa[R][2], xb[R][2], t1[R][2], t2[R][2], t3[R][2];
        w16_load<M, R>(a + inst * sa, xa);
        w16_load<M, R>(b + inst * sb, xb);
        mul(xa, xa, t1);       // a^2
        add_rows(t1, xb, t1);  // a^2 + b
        mul(xb, xb, t2);       // b^2
        add_rows(t2, xb, t2);  // b^2 + b
        mul(t1, t2, t1);       // (a^2 + b)(b^2 + b)
        mul(xa, xb, t3);       // a b
        add_rows(t1, t3, t1);  // + a b
        uint64_t* gr = r + inst * sr;
#pragma unroll
        for (int ro = 0; ro < R; ro++) {
            const uint32_t j0 = 128 * ro + 2 * w.lane;
            if (j0 < (uint32_t)M) {
                gr[j0] = t1[ro][0];
                gr[j0 + 1] = t1[ro][1];
            }
        }
        asm volatile("" ::: "memory");
    }
}

namespace {

template <int M, bool FULL>
hipError_t launch_w16_t(uint32_t n, const uint64_t* a, const uint64_t* b, uint64_t* r, const Strides& st,
                        hipStream_t s) {
    using G = W16<M, FULL>;
    const size_t lds = (size_t)kW16Waves * G::WAVE_BYTES;
    const void* fn = (const void*)mul_wave16_kernel<M, FULL>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + kW16Waves - 1) / kW16Waves);
    mul_wave16_kernel<M, FULL><<<blocks, 64 * kW16Waves, lds, s>>>(n, 256u, a, st.a ? st.a : M, b,
                                                                   st.b ? st.b : M, r, st.r ? st.r : G::PL);
    return hipGetLastError();
}

template <int M>
hipError_t launch_poly_w16_t(uint32_t n, const uint64_t* a, const uint64_t* b, uint64_t* r, const Strides& st,
                             hipStream_t s) {
    using G = W16<M, false>;
    const size_t lds = (size_t)kW16Waves * G::WAVE_BYTES;
    const uint32_t blocks = (uint32_t)(((uint64_t)n + kW16Waves - 1) / kW16Waves);
    poly_wave16_kernel<M><<<blocks, 64 * kW16Waves, lds, s>>>(n, 256u, a, st.a ? st.a : M, b, st.b ? st.b : M,
                                                              r, st.r ? st.r : M);
    return hipGetLastError();
}

}  // namespace

bool mul_wave16_supported(uint32_t m, bool full) {
    (void)full;
    return m == 64 || m == 128;
}

hipError_t launch_mul_wave16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                             hipStream_t s, Strides st) {
    if (n == 0) return mul_wave16_supported(m, full) ? hipSuccess : hipErrorInvalidValue;
    switch (m) {
        case 64: return full ? launch_w16_t<64, true>(n, a, b, r, st, s) : launch_w16_t<64, false>(n, a, b, r, st, s);
        case 128:
            return full ? launch_w16_t<128, true>(n, a, b, r, st, s) : launch_w16_t<128, false>(n, a, b, r, st, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_poly_wave16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                              hipStream_t s, Strides st) {
    if (n == 0) return mul_wave16_supported(m, false) ? hipSuccess : hipErrorInvalidValue;
    switch (m) {
        case 64: return launch_poly_w16_t<64>(n, a, b, r, st, s);
        case 128: return launch_poly_w16_t<128>(n, a, b, r, st, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace midint
