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

}  // namespace midint
