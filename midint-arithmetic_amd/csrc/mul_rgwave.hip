// Ragged operations one per wave: products whose shorter operand has at most
// 128 limbs on the int8 matrix cores' 16 x 16 x 64 instruction, and full
// additions. No reference counterpart: the reference serves mixed sizes with
// one compile-time specialisation and one launch per size
// (runAdditions / runNaiveMuls, cuda/main.cu:661-709); its column-sum
// schoolbook is cuda/ker-classic-mul.cu.h:299-463 and its carry scan
// cuda/ker-addition.cu.h:141-194.
//
// Products (mul_rgw_kernel). The digits, tiling and fold are those of the
// one-product-per-wave kernel of mul_wave16.hip (balanced base-256 digits,
// D = 8n + 1 per operand, output block I = 16t + n of tile t, step k with
// the Toeplitz fragment T_k of the shorter operand b and the 64-digit
// windows of a starting at 16(n + g + 4u) - DL, u = 4t - k, DL = 48). The
// sizes are run-time values here and the products lopsided, so:
//
//  * all K = floor((8 n_b + 15) / 64) + 1 fragments T_0 .. T_{K-1} stay in
//    registers (K <= 17 for n_b <= 128; digits of b beyond its length read
//    as zero, and T_k for k >= K would multiply zeros only); the kernel body
//    is compiled for each K and the wave branches to its product's K;
//  * the output is produced one row of 128 limbs (4 tiles, 1024 digits) at a
//    time. Row R needs the windows of u = 16R - K + 1 .. 16R + 12, i.e. a's
//    digits [1024R - 64K + 16, 1024R + 1024): row R of a and at most 1072
//    digits before it. a's digits are staged row by row into a ring of four
//    1024-byte slots (digit i at byte i mod 4096; every window read is
//    16-byte aligned, so none wraps), the staging carry (a + 0x80..80) in a
//    scalar from row to row; the slots start as zeros, which are the digits
//    below 0;
//  * row R's 4 K MFMAs take K + 12 window fragments from the ring (tile i at
//    fragment j uses k = 4i + K - 1 - j), the tiles are folded into the
//    row's 256 words V (in registers, moved to the row adder's lanes by
//    cross-lane swaps), and the row is normalised at once: limb L of the
//    product is (W_L mod 2^64) + (W_{L-1} >> 64), W_L = V_{2L} + V_{2L+1}
//    2^32, added by the wave's row adder with the carry from row R - 1 in a
//    scalar (the last two words of row R - 1 kept in scalars for lane 0 of
//    row R). The column bias telescopes as in mul_wave16.hip: 255 B in every
//    column of the rows computed, one B more in column 0, B 2^(64 P) = 0
//    modulo the output length for any P >= n_a + n_b limbs.
//
// Per wave 5216 bytes of LDS: b's digits reversed (1120) and the ring
// (4096). Each wave owns its region and one op at a time, so waves never
// synchronise.
//
// Additions (add_rgw_kernel): one full sum (max(n_a, n_b) + 1 limbs) per
// wave, rows of 128 limbs with the carry from row to row in a scalar, the
// next row's loads in flight while a row is added.
#include "add_core.hpp"
#include "mfma_digits.hpp"
#include "midint_internal.h"

namespace midint {

namespace {

constexpr int kRgwDL = 48;        // DL
constexpr uint32_t kRgwWaves = 4;  // waves (independent ops) per workgroup
constexpr int kRgwSmallK = 9;      // the small-fragment kernel: K <= 9 (n_b <= 70)

// Per-wave LDS layout of the kernel for fragment counts K <= KMAX: b's
// digits reversed (digit j at byte X0 - j, X0 = 15 mod 16, bytes
// [X0 - 64 KMAX + 1, X0 + 20) read), the ring of a's digit rows (row R's
// windows reach back at most 64 KMAX - 16 digits: two slots for KMAX <= 9,
// four up to 17).
template <int KMAX>
struct RgwCfg {
    static constexpr uint32_t RING = KMAX <= 9 ? 2048 : 4096;
    static constexpr int32_t X0 = 64 * KMAX - 1;
    static constexpr uint32_t BR = ((uint32_t)X0 + 21 + 15) / 16 * 16;
    static constexpr uint32_t WAVE_BYTES = RING + BR;
    static_assert(X0 % 16 == 15 && X0 >= 64 * KMAX - 1, "b region origin");
    static_assert(64 * KMAX - 16 <= (int)RING - 1024, "ring look-back");
};

// Limbs 128 ro + 2 lane, +1 of an n-limb operand (zeros at or above n).
__device__ __forceinline__ void rgw_load_row(const uint64_t* __restrict__ g, uint32_t n, uint32_t ro, uint32_t lane,
                                             uint64_t& x0, uint64_t& x1) {
    const uint32_t j0 = 128 * ro + 2 * lane;
    x0 = j0 < n ? g[j0] : 0ull;
    x1 = j0 + 1 < n ? g[j0 + 1] : 0ull;
}

// Digits of one staged row: x = limbs + 0x80..80 (limbs below n only) with
// the carry *c from the previous row, digit bytes x ^ 0x80; the carry out
// of limb n - 1 lands in limb n as digit 8n (0 or 1), limbs above stay 0.
__device__ __forceinline__ void rgw_digits(uint64_t x0, uint64_t x1, uint32_t n, uint32_t ro, uint32_t lane,
                                           uint32_t* c, uint64_t& w0, uint64_t& w1) {
    const uint32_t j0 = 128 * ro + 2 * lane;
    const uint64_t h0 = j0 < n ? kMfHalf : 0ull, h1 = j0 + 1 < n ? kMfHalf : 0ull;
    const RowState st = row_build(x0, h0, x1, h1, false, false);
    const uint64_t C = row_chain(st, c);
    row_apply(st, C, &w0, &w1);
    w0 ^= h0;
    w1 ^= h1;
}

// The window fragments j < K + 12 that a row's MFMAs use: tile i takes
// fragment j at step k = 4 i + K - 1 - j, so every j when K >= 4 and only
// the j with j mod 4 < K below (4, 8 or 12 of the K + 12).
struct RgwSteps {
    int n;
    int j[32];
};
template <int K>
constexpr RgwSteps rgw_steps() {
    RgwSteps st{};
    for (int j = 0; j < K + 12; j++) {
        bool used = false;
        for (int i = 0; i < 4; i++) {
            const int k = 4 * i + K - 1 - j;
            used = used || (k >= 0 && k < K);
        }
        if (used) st.j[st.n++] = j;
    }
    return st;
}

template <int K, int KMAX>
__device__ __forceinline__ void rgw_product(const MfRgItem& it, uint8_t* base, uint32_t lane, const MfPow& pw) {
    using G = RgwCfg<KMAX>;
    constexpr uint32_t kRing = G::RING;
    constexpr int32_t kX0 = G::X0;
    uint8_t* breg = base;  // first, so the fragment reads' offsets fit ds_read2_b32's offset fields
    uint8_t* ring = base + G::BR;
    const uint32_t na = it.n_a, nb = it.n_b, pl = na + nb;
    const uint32_t ln = lane & 15, lg = lane >> 4;

    // the first rows of both operands in flight before anything else
    uint64_t bx0, bx1, nx0, nx1;  // b's row 0; a's next row (loaded a row ahead)
    rgw_load_row(it.b, nb, 0, lane, bx0, bx1);
    rgw_load_row(it.a, na, 0, lane, nx0, nx1);

    // ---- zero the ring, b's region and the word prefix
#pragma unroll
    for (uint32_t x = 0; x < kRing; x += 1024)
        *reinterpret_cast<uint4*>(ring + x + 16 * lane) = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t x = 16 * lane; x < G::BR; x += 1024)
        *reinterpret_cast<uint4*>(breg + x) = make_uint4(0u, 0u, 0u, 0u);
    asm volatile("" ::: "memory");

    // ---- b's digits reversed: digit j at byte X0 - j (n_b + 1 limbs: one
    // or two rows)
    {
        uint32_t cb = 0;
        const uint32_t rows_b = (nb + 128) / 128;  // row 1 (n_b = 128) holds the carry digit only
        for (uint32_t ro = 0; ro < rows_b; ro++) {
            uint64_t w0, w1;
            rgw_digits(ro ? 0ull : bx0, ro ? 0ull : bx1, nb, ro, lane, &cb, w0, w1);
            const uint32_t j0 = 128 * ro + 2 * lane;
            if (j0 <= nb)
                *reinterpret_cast<ulonglong2*>(breg + kX0 - 8 * j0 - 15) =
                    make_ulonglong2(__builtin_bswap64(w1), __builtin_bswap64(w0));
        }
    }
    asm volatile("" ::: "memory");

    // ---- the K Toeplitz fragments of b: lane (r, g) holds the 16 bytes
    // b_{64k + DL + r - 16g - e}, e < 16, bytes [Y, Y + 16) with
    // Y = X0 - 64k - DL - r + 16g (the funnel shift depends on the lane only)
    mf_v4i A[K];
    {
        const uint32_t y0 = (uint32_t)(kX0 - 64 * (K - 1) - kRgwDL) - ln + 16 * lg;
        const uint32_t sh = 8 * (y0 & 3u);
        const uint8_t* pa = breg + (y0 & ~3u);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(pa + 64 * (K - 1 - k));
            const uint32_t q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
            A[k] = mf_v4i{(int)__builtin_amdgcn_alignbit(q1, q0, sh), (int)__builtin_amdgcn_alignbit(q2, q1, sh),
                          (int)__builtin_amdgcn_alignbit(q3, q2, sh), (int)__builtin_amdgcn_alignbit(q4, q3, sh)};
        }
    }

    // ---- rows of the product
    const uint32_t rows = (pl + 127) / 128;
    uint32_t ca = 0, cn = 0;
    uint32_t prev[4] = {0u, 0u, 0u, 0u};  // the previous row's words 254, 255 (V_-2, V_-1 before row 0)
    for (uint32_t R = 0; R < rows; R++) {
        // stage a's row R into slot R mod 4
        {
            uint64_t w0, w1;
            rgw_digits(nx0, nx1, na, R, lane, &ca, w0, w1);
            *reinterpret_cast<ulonglong2*>(ring + ((1024 * R + 16 * lane) & (kRing - 1))) =
                make_ulonglong2(w0, w1);
        }
        if (R + 1 < rows) rgw_load_row(it.a, na, R + 1, lane, nx0, nx1);
        asm volatile("" ::: "memory");

        // 4 K MFMAs over the K + 12 window fragments (those used), two
        // fragments ahead
        constexpr RgwSteps ST = rgw_steps<K>();
        const int32_t fb = (int32_t)(16 * (ln + lg)) - kRgwDL + 1024 * (int32_t)R - 64 * (K - 1);
        auto frag = [&](int j) {
            return *reinterpret_cast<const mf_v4i*>(ring + ((uint32_t)(fb + 64 * j) & (kRing - 1)));
        };
        mf_v4i acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = mf_v4i{0, 0, 0, 0};
        mf_v4i bq[3];
        bq[0] = frag(ST.j[0]);
        bq[1] = frag(ST.j[1]);
#pragma unroll
        for (int q = 0; q < ST.n; q++) {
            const int j = ST.j[q];
            if (q + 2 < ST.n) bq[(q + 2) % 3] = frag(ST.j[q + 2]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = 4 * i + K - 1 - j;
                if (k >= 0 && k < K) acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[k], bq[q % 3], acc[i], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }

        // fold the tiles' column sums into the row's words, moved to the row
        // adder's lanes through cross-lane swaps (mf_row_words)
        uint64_t S[4], w[4], p[2];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int64_t V = (int64_t)kMfFoldBias;
            V += (int64_t)acc[i][0] * (int32_t)pw.p0;
            V += (int64_t)acc[i][1] * (int32_t)pw.p8;
            V += (int64_t)acc[i][2] * (int32_t)pw.p16;
            V += (int64_t)acc[i][3] * (int32_t)pw.p24;
            S[i] = (uint64_t)V;
        }
        mf_row_words(S, w, p, prev);

        // limbs of row R: x_L = W_L mod 2^64, y_L = W_{L-1} >> 64, x + y
        if (R == 0 && lane == 0) w[0] += kMfBias;  // V_0: column 0 takes 256 B
        const uint64_t sp = p[1] + (p[0] >> 32);
        const uint64_t s0 = w[1] + (w[0] >> 32);
        const uint64_t s1 = w[3] + (w[2] >> 32);
        const RowState st = row_build((w[0] & 0xffffffffull) | (s0 << 32), sp >> 32,
                                      (w[2] & 0xffffffffull) | (s1 << 32), s0 >> 32, false, false);
        const uint64_t C = row_chain(st, &cn);
        uint64_t z0, z1;
        row_apply(st, C, &z0, &z1);
        const uint32_t j0 = 128 * R + 2 * lane;
        if (j0 < pl) it.r[j0] = z0;
        if (j0 + 1 < pl) it.r[j0 + 1] = z1;
        asm volatile("" ::: "memory");
    }
}

}  // namespace

// One product per wave: item blockIdx.x * 4 + wave (n_a >= n_b, fragment
// count K = floor((8 n_b + 15) / 64) + 1 in [KMIN, KMAX]; KMAX <= 17, i.e.
// n_b <= 128). Operands and results may be only 8-byte aligned (packed
// ragged buffers): global accesses are limb by limb.
template <int KMIN, int KMAX>
__global__ __launch_bounds__(64 * kRgwWaves) void mul_rgw_kernel(const MfRgItem* __restrict__ items, uint32_t n_items,
                                                                  uint32_t pow8) {
    extern __shared__ __align__(16) uint64_t lds64[];
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t idx = blockIdx.x * kRgwWaves + wv;
    if (idx >= n_items) return;
    const MfRgItem it = items[idx];
    uint8_t* base = reinterpret_cast<uint8_t*>(lds64) + (size_t)wv * RgwCfg<KMAX>::WAVE_BYTES;
    const MfPow pw = mf_pow(pow8);
    const int K = (int)((8 * it.n_b + 15) / 64 + 1);
#define MIDINT_RGW_CASE(k)                                                           \
    case k:                                                                          \
        if constexpr (k >= KMIN && k <= KMAX) rgw_product<k, KMAX>(it, base, lane, pw); \
        break;
    switch (K) {
        MIDINT_RGW_CASE(1)
        MIDINT_RGW_CASE(2)
        MIDINT_RGW_CASE(3)
        MIDINT_RGW_CASE(4)
        MIDINT_RGW_CASE(5)
        MIDINT_RGW_CASE(6)
        MIDINT_RGW_CASE(7)
        MIDINT_RGW_CASE(8)
        MIDINT_RGW_CASE(9)
        MIDINT_RGW_CASE(10)
        MIDINT_RGW_CASE(11)
        MIDINT_RGW_CASE(12)
        MIDINT_RGW_CASE(13)
        MIDINT_RGW_CASE(14)
        MIDINT_RGW_CASE(15)
        MIDINT_RGW_CASE(16)
        MIDINT_RGW_CASE(17)
        default: break;
    }
#undef MIDINT_RGW_CASE
}

// One full sum per wave: r = a + b, max(n_a, n_b) + 1 limbs.
__global__ __launch_bounds__(64 * kRgwWaves) void add_rgw_kernel(const MfRgItem* __restrict__ items, uint32_t n_items) {
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t idx = blockIdx.x * kRgwWaves + wv;
    if (idx >= n_items) return;
    const MfRgItem it = items[idx];
    const uint32_t na = it.n_a, nb = it.n_b;
    const uint32_t out = (na > nb ? na : nb) + 1;
    const uint32_t rows = (out + 127) / 128;
    uint64_t x0, x1, y0, y1;
    rgw_load_row(it.a, na, 0, lane, x0, x1);
    rgw_load_row(it.b, nb, 0, lane, y0, y1);
    uint32_t c = 0;
    for (uint32_t R = 0; R < rows; R++) {
        const RowState st = row_build(x0, y0, x1, y1, false, false);
        if (R + 1 < rows) {
            rgw_load_row(it.a, na, R + 1, lane, x0, x1);
            rgw_load_row(it.b, nb, R + 1, lane, y0, y1);
        }
        const uint64_t C = row_chain(st, &c);
        uint64_t z0, z1;
        row_apply(st, C, &z0, &z1);
        const uint32_t j0 = 128 * R + 2 * lane;
        if (j0 < out) it.r[j0] = z0;
        if (j0 + 1 < out) it.r[j0 + 1] = z1;
    }
}

uint32_t rgw_max_short() { return 128; }

uint32_t rgw_small_max_short() { return (64 * kRgwSmallK - 15 - 1) / 8; }

size_t rgw_lds_bytes(int part) {
    return (size_t)kRgwWaves * (part == 1 ? RgwCfg<kRgwSmallK>::WAVE_BYTES : RgwCfg<17>::WAVE_BYTES);
}

uint32_t rgw_waves() { return kRgwWaves; }

const void* rgw_mul_kernel_ptr(int part) {
    if (part == 1) return (const void*)mul_rgw_kernel<1, kRgwSmallK>;
    if (part == 2) return (const void*)mul_rgw_kernel<kRgwSmallK + 1, 17>;
    return (const void*)mul_rgw_kernel<1, 17>;
}

const void* rgw_add_kernel_ptr() { return (const void*)add_rgw_kernel; }

}  // namespace midint
