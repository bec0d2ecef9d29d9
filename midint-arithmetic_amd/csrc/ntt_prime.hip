// NTT over the reference's two primes and Goldilocks (gfx950): batched
// transforms and exact products.
//
// Reference: FftPrime32 / FftPrime64 and zmod_t (cuda/ker-fft-mul.cu.h:44-140),
// fft / ifft (:304-368): bit-reversal permutation (permuteIP, :179-189), then
// radix-2 decimation-in-time stages where stage q (half-width h = 2^(q-1))
// multiplies x[k+j+h] by omega^(j * N / 2^q); the inverse runs the same with
// omega^-1 and scales by N^-1.
//
// This file keeps that algorithm (so the transform is bit-compatible with the
// reference's, root included) on one instance per block in LDS, with:
//  * Montgomery multiplication for p32 = 3*2^30+1 (R = 2^32) and
//    p64 = 29*2^57+1 (R = 2^64): twiddles are stored as w*R mod p, so
//    mul(x, w*R) = x*w and data stay in standard form; a data x data product
//    carries one R^-1, folded into the final scale constant;
//  * Goldilocks with the special-form reduction of midint_dev.hpp (R = 1).
//
// Product mode: operands split into DB-bit coefficients (16 for p64 and
// Goldilocks, 8 for p32, so every coefficient of the product stays below p),
// cyclic convolution of length L >= 2 * coefficient count, exact recombination
// limb j = sum_s c_{j*64/DB + s} 2^(DB s) (< 2^128) split into a low limb and a
// high part one limb up, then the carry-propagating row adder.
#include "add_core.hpp"
#include "midint_internal.h"

namespace midint {

namespace {

constexpr uint32_t neg_inv32(uint32_t p) {
    uint32_t x = 1;  // Newton: x = p^-1 mod 2^k, doubling k each step
    for (int i = 0; i < 5; i++) x *= 2u - p * x;
    return 0u - x;
}
constexpr uint64_t neg_inv64(uint64_t p) {
    uint64_t x = 1;
    for (int i = 0; i < 6; i++) x *= 2ull - p * x;
    return 0ull - x;
}

}  // namespace

struct FieldP32 {
    static constexpr uint64_t P = 3221225473ull;  // 3 * 2^30 + 1
    static constexpr uint32_t PINV = neg_inv32((uint32_t)P);
    __device__ static __forceinline__ uint64_t add(uint64_t a, uint64_t b) {
        const uint64_t s = a + b;
        return s >= P ? s - P : s;
    }
    __device__ static __forceinline__ uint64_t sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }
    // a * b * 2^-32 mod p for a, b < p
    __device__ static __forceinline__ uint64_t mul(uint64_t a, uint64_t b) {
        const uint64_t x = (uint64_t)(uint32_t)a * (uint32_t)b;
        const uint32_t q = (uint32_t)x * PINV;
        const uint64_t qp = (uint64_t)q * (uint32_t)P;
        const uint64_t s = x + qp;  // low 32 bits are 0
        const uint64_t c = s < x ? 1ull : 0ull;
        const uint64_t t = (s >> 32) | (c << 32);  // < 2p
        return t >= P ? t - P : t;
    }
};

struct FieldP64 {
    static constexpr uint64_t P = 4179340454199820289ull;  // 29 * 2^57 + 1
    static constexpr uint64_t PINV = neg_inv64(P);
    __device__ static __forceinline__ uint64_t add(uint64_t a, uint64_t b) {
        const uint64_t s = a + b;  // p < 2^62: no overflow
        return s >= P ? s - P : s;
    }
    __device__ static __forceinline__ uint64_t sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }
    // a * b * 2^-64 mod p for a, b < p
    __device__ static __forceinline__ uint64_t mul(uint64_t a, uint64_t b) {
        const uint64_t lo = a * b, hi = __umul64hi(a, b);
        const uint64_t q = lo * PINV;
        const uint64_t qh = __umul64hi(q, P);  // q * p = qh * 2^64 + (2^64 - lo) when lo != 0
        const uint64_t t = hi + qh + (lo != 0 ? 1ull : 0ull);  // < 2p < 2^63
        return t >= P ? t - P : t;
    }
};

struct FieldGL {
    static constexpr uint64_t P = kGlP;
    __device__ static __forceinline__ uint64_t add(uint64_t a, uint64_t b) { return gl_add(a, b); }
    __device__ static __forceinline__ uint64_t sub(uint64_t a, uint64_t b) { return gl_sub(a, b); }
    __device__ static __forceinline__ uint64_t mul(uint64_t a, uint64_t b) { return gl_mul(a, b); }
};

// Decimation-in-time stages on bit-reversed input (the permutation of
// permuteIP is applied when the values are stored to LDS); tw[i] = omega^i
// in the field's multiplier form, i < L/2. Ends with a barrier.
template <class F>
__device__ __forceinline__ void np_dit(uint64_t* x, uint32_t lg, const uint64_t* __restrict__ tw) {
    const uint32_t half = 1u << (lg - 1);
    for (uint32_t q = 1; q <= lg; q++) {
        const uint32_t lgh = q - 1;
        const uint32_t h = 1u << lgh;
        const uint32_t sh = lg - q;  // twiddle index j * (L >> q)
        for (uint32_t bfl = threadIdx.x; bfl < half; bfl += blockDim.x) {
            const uint32_t j = bfl & (h - 1);
            const uint32_t i0 = ((bfl >> lgh) << (lgh + 1)) + j;
            const uint64_t u = x[i0];
            const uint64_t tau = F::mul(x[i0 + h], tw[j << sh]);
            x[i0] = F::add(u, tau);
            x[i0 + h] = F::sub(u, tau);
        }
        __syncthreads();
    }
}

constexpr uint32_t kNpMaxPerThread = 16;

// MODE 0: exact product of m-limb operands (r: 2m limbs per instance).
// MODE 1: forward transform of L values; MODE 2: inverse (omega^-1, then
// the scale constant, which is N^-1 in the field's standard form).
template <class F, int MODE, int DB>
__global__ __launch_bounds__(1024) void ntt_prime_kernel(uint32_t m, uint32_t lg, const uint64_t* __restrict__ a,
                                                         const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                         const uint64_t* __restrict__ twf,
                                                         const uint64_t* __restrict__ twi, uint64_t scale) {
    extern __shared__ __align__(16) uint64_t x[];
    __shared__ uint32_t lds_gp[16];
    const uint32_t L = 1u << lg;
    const uint32_t T = blockDim.x;
    const uint32_t tid = threadIdx.x;

    if (MODE != 0) {
        const uint64_t* g = a + (size_t)blockIdx.x * L;
        for (uint32_t k = tid; k < L; k += T) x[__brev(k) >> (32 - lg)] = g[k];
        __syncthreads();
        np_dit<F>(x, lg, MODE == 1 ? twf : twi);
        uint64_t* gr = r + (size_t)blockIdx.x * L;
        for (uint32_t k = tid; k < L; k += T) gr[k] = MODE == 2 ? F::mul(x[k], scale) : x[k];
        return;
    }

    constexpr uint32_t CPL = 64 / DB;  // coefficients per limb
    constexpr uint64_t DMASK = (1ull << DB) - 1;
    const uint32_t ncoef = CPL * m;
    const uint32_t E = L / T;
    const uint64_t* ga = a + (size_t)blockIdx.x * m;
    const uint64_t* gb = b + (size_t)blockIdx.x * m;

    // forward transform of a (natural order out), kept in registers
    for (uint32_t k = tid; k < L; k += T)
        x[__brev(k) >> (32 - lg)] = k < ncoef ? (ga[k / CPL] >> (DB * (k % CPL))) & DMASK : 0ull;
    __syncthreads();
    np_dit<F>(x, lg, twf);
    uint64_t ah[kNpMaxPerThread];
#pragma unroll
    for (uint32_t e = 0; e < kNpMaxPerThread; e++)
        if (e < E) ah[e] = x[tid + T * e];
    __syncthreads();

    // forward transform of b, pointwise product stored bit-reversed for the inverse
    for (uint32_t k = tid; k < L; k += T)
        x[__brev(k) >> (32 - lg)] = k < ncoef ? (gb[k / CPL] >> (DB * (k % CPL))) & DMASK : 0ull;
    __syncthreads();
    np_dit<F>(x, lg, twf);
    uint64_t z[kNpMaxPerThread];
#pragma unroll
    for (uint32_t e = 0; e < kNpMaxPerThread; e++)
        if (e < E) z[e] = F::mul(x[tid + T * e], ah[e]);
    __syncthreads();
#pragma unroll
    for (uint32_t e = 0; e < kNpMaxPerThread; e++)
        if (e < E) x[__brev(tid + T * e) >> (32 - lg)] = z[e];
    __syncthreads();
    np_dit<F>(x, lg, twi);

    // recombination
    const uint32_t outl = 2 * m;
    const uint32_t LPT = (outl + T - 1) / T;  // <= 4 (L >= 2 * ncoef, T >= L / 16)
    uint64_t lo_v[4], hi_v[4];
#pragma unroll
    for (uint32_t e = 0; e < 4; e++) {
        const uint32_t j = tid + T * e;
        uint64_t lo = 0, hi = 0;
        if (e < LPT && j < outl) {
            unsigned __int128 acc = 0;
#pragma unroll
            for (uint32_t s = 0; s < CPL; s++)
                acc += (unsigned __int128)F::mul(x[CPL * j + s], scale) << (DB * s);
            lo = (uint64_t)acc;
            hi = (uint64_t)(acc >> 64);
        }
        lo_v[e] = lo;
        hi_v[e] = hi;
    }
    __syncthreads();
    uint64_t* Lb = x;
    uint64_t* Hb = x + outl;
    if (tid == 0) Hb[0] = 0;
#pragma unroll
    for (uint32_t e = 0; e < 4; e++) {
        const uint32_t j = tid + T * e;
        if (e < LPT && j < outl) {
            Lb[j] = lo_v[e];
            if (j + 1 < outl) Hb[j + 1] = hi_v[e];
        }
    }
    __syncthreads();
    constexpr int R = 2;  // rows per wave: T/64 waves * 2 rows * 128 limbs = 4T >= 2m
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
    Rows<R> rows;
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        rows.x0[v] = k < outl ? Lb[k] : 0ull;
        rows.x1[v] = k + 1 < outl ? Lb[k + 1] : 0ull;
        rows.y0[v] = k < outl ? Hb[k] : 0ull;
        rows.y1[v] = k + 1 < outl ? Hb[k + 1] : 0ull;
        rows.s0[v] = (k == 0);
        rows.s1[v] = false;
    }
    block_add_rows<R>(rows, lds_gp);
    uint64_t* gr = r + (size_t)blockIdx.x * outl;
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        if (k < outl) gr[k] = rows.x0[v];
        if (k + 1 < outl) gr[k + 1] = rows.x1[v];
    }
}

// ---------------------------------------------------------------------------
// Host side.

namespace {

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((unsigned __int128)a * b) % p); }
uint64_t h_pow(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = h_mulmod(r, a, p);
        a = h_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

struct PrimeInfo {
    uint64_t p, g;
    uint32_t n;      // 2-adicity of p - 1
    uint32_t rbits;  // Montgomery R = 2^rbits (0: none)
};

PrimeInfo prime_info(uint32_t prime) {
    switch (prime) {
        case MIDINT_PRIME_P32: return {FieldP32::P, 13, 30, 32};
        case MIDINT_PRIME_P64: return {FieldP64::P, 21, 57, 64};
        default: return {kGlP, 7, 32, 0};
    }
}

// x * R mod p (the multiplier form of a twiddle or constant)
uint64_t to_mult(const PrimeInfo& pi, uint64_t v) {
    if (pi.rbits == 0) return v;
    return (uint64_t)(((unsigned __int128)v << pi.rbits) % pi.p);
}

}  // namespace

int ntt_prime_tables(uint32_t prime, uint32_t lg, bool ref_root, uint64_t* fwd, uint64_t* inv, uint64_t* scale_inv,
                     uint64_t* scale_prod) {
    const PrimeInfo pi = prime_info(prime);
    if (lg == 0 || lg > pi.n) return -1;
    const uint64_t L = 1ull << lg;
    // the reference's getOmega (cuda/ker-fft-mul.cu.h:191-199) raises g to
    // 2^(n - lg) (computed there as an int shift, which for p64 exceeds 31
    // bits; the intended power of two is used here); a primitive L-th root
    // is g^((p-1)/L)
    const uint64_t w = ref_root ? h_pow(pi.g, 1ull << (pi.n - lg), pi.p) : h_pow(pi.g, (pi.p - 1) / L, pi.p);
    const uint64_t wi = h_pow(w, pi.p - 2, pi.p);
    uint64_t f = 1, g = 1;
    for (uint64_t i = 0; i < L / 2; i++) {
        fwd[i] = to_mult(pi, f);
        inv[i] = to_mult(pi, g);
        f = h_mulmod(f, w, pi.p);
        g = h_mulmod(g, wi, pi.p);
    }
    const uint64_t linv = h_pow(L % pi.p, pi.p - 2, pi.p);
    *scale_inv = to_mult(pi, linv);                 // mul(y, .) = y * L^-1
    *scale_prod = to_mult(pi, to_mult(pi, linv));   // undoes the R^-1 of the pointwise product too
    return 0;
}

uint32_t ntt_prime_digit_bits(uint32_t prime) { return prime == MIDINT_PRIME_P32 ? 8 : 16; }

static uint32_t np_threads(uint32_t lg) {
    uint32_t T = (1u << lg) / 16;
    return T < 64 ? 64 : (T > 1024 ? 1024 : T);
}

template <class F, int MODE, int DB>
static hipError_t np_launch(uint32_t n, uint32_t m, uint32_t lg, const uint64_t* a, const uint64_t* b, uint64_t* r,
                            const uint64_t* twf, const uint64_t* twi, uint64_t scale, hipStream_t s) {
    const uint32_t T = np_threads(lg);
    const size_t lds = ((size_t)1 << lg) * 8;
    const void* fn = (const void*)ntt_prime_kernel<F, MODE, DB>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    ntt_prime_kernel<F, MODE, DB><<<n, T, lds, s>>>(m, lg, a, b, r, twf, twi, scale);
    return hipGetLastError();
}

hipError_t launch_ntt_prime(uint32_t prime, int mode, uint32_t n, uint32_t m, uint32_t lg, const uint64_t* a,
                            const uint64_t* b, uint64_t* r, const uint64_t* twf, const uint64_t* twi, uint64_t scale,
                            hipStream_t s) {
    if (lg < 1 || lg > kNttMaxLg || (mode == 0 && lg < 6)) return hipErrorInvalidValue;
#define NP_CASE(F, DB)                                                                        \
    switch (mode) {                                                                           \
        case 0: return np_launch<F, 0, DB>(n, m, lg, a, b, r, twf, twi, scale, s);            \
        case 1: return np_launch<F, 1, DB>(n, m, lg, a, b, r, twf, twi, scale, s);            \
        default: return np_launch<F, 2, DB>(n, m, lg, a, b, r, twf, twi, scale, s);           \
    }
    switch (prime) {
        case MIDINT_PRIME_P32: NP_CASE(FieldP32, 8)
        case MIDINT_PRIME_P64: NP_CASE(FieldP64, 16)
        default: NP_CASE(FieldGL, 16)
    }
#undef NP_CASE
}

}  // namespace midint
