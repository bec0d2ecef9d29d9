// Balanced base-256 digits for the int8 matrix-core multiplications
// (mul_mfma.hip, mul_wave16.hip): constants shared by both kernels.
//
// An operand of m limbs is written as x = a + 0x8080...80 (128 added to each
// of its 8m bytes, carries propagated); digit i (i < 8m) is byte i of x minus
// 128, in [-128, 127], stored as the int8 byte (byte i of x) ^ 0x80, and digit
// 8m is the carry out of x (0 or 1). Then a = sum_i digit_i 256^i exactly.
//
// A column sum (sum of digit products with one output position) is signed;
// 255 B (B = 2^21) added to every column makes it non-negative and below
// 2^30 for operands of up to 4096 limbs. Over the N output columns of a
// product of PL limbs (N = 8 PL) the biases add 255 B (256^N - 1) / 255 =
// B 256^N - B; one more B in column 0 makes the total B 256^N, which is zero
// modulo 2^(64 PL).
#pragma once

#include <cstdint>

namespace midint {

typedef int mf_v4i __attribute__((ext_vector_type(4)));

constexpr uint64_t kMfHalf = 0x8080808080808080ull;  // 128 in every byte
constexpr uint32_t kMfBias = 1u << 21;               // B

// The bias of four consecutive columns as one value: 255 B in each column,
// the columns 8 bits apart.
constexpr uint64_t kMfFoldBias = (uint64_t)(255u * kMfBias) * 0x01010101ull;

// 1, 2^8, 2^16, 2^24 in SGPRs the compiler cannot see through (built from
// a kernel argument), so that c * 2^k + G stays one v_mad_i64_i32 instead
// of a 64-bit shift and add.
struct MfPow {
    uint32_t p0, p8, p16, p24;
};

__device__ __forceinline__ MfPow mf_pow(uint32_t pow8) {
    MfPow pw;
    pw.p0 = pow8 >> 8;
    pw.p8 = pow8;
    pw.p16 = pow8 * pow8;
    pw.p24 = pw.p16 * pow8;
    return pw;
}

// The words V of one output row of 128 limbs (256 words) to the row
// adder's lanes. The row's four 16 x 16 tiles hold word 64 q + 4 n + g in
// lane n + 16 g of tile q (S[q]); lane l gets words 4 l .. 4 l + 3 in w[0 ..
// 3] (its limbs 2 l, 2 l + 1 are w[0] + w[1] 2^32, w[2] + w[3] 2^32) and
// the two words before them in p[0], p[1]. Words 4 l + j sit in lane
// (l mod 16) + 16 j of tile l / 16: a 4 x 4 transpose of 16-lane rows across
// the four tiles (permlane32 swaps of tiles q and q + 2, then permlane16
// swaps of q and q + 1, on both 32-bit halves) brings them to lane l; the
// words before them are lane l - 1's w[2], w[3] (wave_shr:1), and lane 0's
// are the previous row's words 254, 255 (*prev, wave-uniform; zero before
// the first row), which *prev receives for the next row.
__device__ __forceinline__ void mf_row_words(const uint64_t (&S)[4], uint64_t (&w)[4], uint64_t (&p)[2],
                                             uint32_t (&prev)[4]) {
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        lo[q] = (uint32_t)S[q];
        hi[q] = (uint32_t)(S[q] >> 32);
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const auto a = __builtin_amdgcn_permlane32_swap(lo[q], lo[q + 2], false, false);
        lo[q] = a[0];
        lo[q + 2] = a[1];
        const auto b = __builtin_amdgcn_permlane32_swap(hi[q], hi[q + 2], false, false);
        hi[q] = b[0];
        hi[q + 2] = b[1];
    }
#pragma unroll
    for (int q = 0; q < 4; q += 2) {
        const auto a = __builtin_amdgcn_permlane16_swap(lo[q], lo[q + 1], false, false);
        lo[q] = a[0];
        lo[q + 1] = a[1];
        const auto b = __builtin_amdgcn_permlane16_swap(hi[q], hi[q + 1], false, false);
        hi[q] = b[0];
        hi[q + 1] = b[1];
    }
    // lane l - 1's words 4 l - 2, 4 l - 1 (wave_shr:1; lane 0 keeps the old value)
    const uint32_t m2l = (uint32_t)__builtin_amdgcn_update_dpp((int)prev[0], (int)lo[2], 0x138, 0xf, 0xf, false);
    const uint32_t m2h = (uint32_t)__builtin_amdgcn_update_dpp((int)prev[1], (int)hi[2], 0x138, 0xf, 0xf, false);
    const uint32_t m3l = (uint32_t)__builtin_amdgcn_update_dpp((int)prev[2], (int)lo[3], 0x138, 0xf, 0xf, false);
    const uint32_t m3h = (uint32_t)__builtin_amdgcn_update_dpp((int)prev[3], (int)hi[3], 0x138, 0xf, 0xf, false);
    prev[0] = __builtin_amdgcn_readlane(lo[2], 63);
    prev[1] = __builtin_amdgcn_readlane(hi[2], 63);
    prev[2] = __builtin_amdgcn_readlane(lo[3], 63);
    prev[3] = __builtin_amdgcn_readlane(hi[3], 63);
#pragma unroll
    for (int j = 0; j < 4; j++) w[j] = ((uint64_t)hi[j] << 32) | lo[j];
    p[0] = ((uint64_t)m2h << 32) | m2l;
    p[1] = ((uint64_t)m3h << 32) | m3l;
}

}  // namespace midint
