// NTT multiplication over the Goldilocks prime p = 2^64 - 2^32 + 1 (gfx950).
//
// Reference: bmulFFT / bmulFftReg (cuda/ker-fft-mul.cu.h:375-454) multiply
// one integer per block with a radix-2 NTT over a 32-bit prime and base-2^15
// digits, then recombine with a carry pass (cuda/ker-fft-help.cu.h). That
// path is not exact for arbitrary 16-bit digits; this one is:
//
//  * each operand of m u64 limbs is split into 4m 16-bit coefficients;
//  * the linear convolution (length < 8m) is computed with an L-point NTT,
//    L = 2^lg >= 8m, over Goldilocks -- every coefficient of the product is
//    below 4m * 2^32 < p, so the field result is the integer result;
//  * the forward transform is decimation-in-frequency (natural order in,
//    bit-reversed out) and the inverse is decimation-in-time (bit-reversed
//    in, natural out), so no bit-reversal permutation pass is needed
//    (the reference permutes in place, permuteIP :179-189);
//  * coefficients are recombined exactly: limb j gets
//    c_{4j} + c_{4j+1} 2^16 + c_{4j+2} 2^32 + c_{4j+3} 2^48 (< 2^94), split
//    into a low limb and a high part one limb up, and the two vectors are
//    added with the wave-ballot row adder.
//
// One instance per block. Twiddles w^i and w^-i (i < L/2) come from a
// per-length table built once on the host. Two kernels:
//  * ntt_kernel (L = 2^6 .. 2^9): the whole transform in LDS (8L bytes),
//    one radix-2 stage per barrier;
//  * ntt_fast_kernel (L = 2^10 .. 2^14): register passes of 4-5 stages
//    (ntt_fast.hpp), one LDS round trip (padded layout) between passes,
//    a's transform kept in registers for the pointwise product.
#include "add_core.hpp"
#include "midint_internal.h"
#include "ntt_fast.hpp"

namespace midint {

__device__ __forceinline__ void ntt_dif(uint64_t* x, uint32_t lg, const uint64_t* __restrict__ tw) {
    const uint32_t half = 1u << (lg - 1);
    const uint32_t T = blockDim.x;
    for (uint32_t s = 0; s < lg; s++) {
        const uint32_t lgh = lg - 1 - s;
        const uint32_t h = 1u << lgh;
        for (uint32_t bfl = threadIdx.x; bfl < half; bfl += T) {
            const uint32_t j = bfl & (h - 1);
            const uint32_t blk = bfl >> lgh;
            const uint32_t i0 = (blk << (lgh + 1)) + j;
            const uint32_t i1 = i0 + h;
            const uint64_t u = x[i0], v = x[i1];
            x[i0] = gl_add(u, v);
            x[i1] = gl_mul(gl_sub(u, v), tw[j << s]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void ntt_dit(uint64_t* x, uint32_t lg, const uint64_t* __restrict__ tw) {
    const uint32_t half = 1u << (lg - 1);
    const uint32_t T = blockDim.x;
    for (uint32_t lgh = 0; lgh < lg; lgh++) {
        const uint32_t h = 1u << lgh;
        const uint32_t sh = lg - 1 - lgh;
        for (uint32_t bfl = threadIdx.x; bfl < half; bfl += T) {
            const uint32_t j = bfl & (h - 1);
            const uint32_t blk = bfl >> lgh;
            const uint32_t i0 = (blk << (lgh + 1)) + j;
            const uint32_t i1 = i0 + h;
            const uint64_t u = x[i0];
            const uint64_t v = gl_mul(x[i1], tw[j << sh]);
            x[i0] = gl_add(u, v);
            x[i1] = gl_sub(u, v);
        }
        __syncthreads();
    }
}

constexpr uint32_t kNttMaxPerThread = 16;  // L / blockDim for every supported L
constexpr uint32_t kNttMaxLimbsPerThread = 4;  // 2m / blockDim, blockDim >= m/2

// MODE 0: integer product (r has 2m limbs per instance).
// MODE 1: raw cyclic convolution of L values (reference bmulFFTvalid).
template <int MODE>
__global__ __launch_bounds__(1024) void ntt_kernel(uint32_t m, uint32_t rl, uint32_t lg, const uint64_t* __restrict__ a,
                                                   const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                   const uint64_t* __restrict__ twf,
                                                   const uint64_t* __restrict__ twi, uint64_t inv_len) {
    extern __shared__ __align__(16) uint64_t x[];
    __shared__ uint32_t lds_gp[16];
    const uint32_t L = 1u << lg;
    const uint32_t T = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t E = L / T;

    // ---- forward transform of a, kept in registers (positions tid + T*e)
    const uint64_t* ga = a + (size_t)blockIdx.x * (MODE == 0 ? m : L);
    const uint64_t* gb = b + (size_t)blockIdx.x * (MODE == 0 ? m : L);
    for (uint32_t k = tid; k < L; k += T) {
        uint64_t v;
        if (MODE == 0)
            v = (k < 4 * m) ? ((ga[k >> 2] >> (16 * (k & 3))) & 0xFFFFull) : 0ull;
        else
            v = ga[k];
        x[k] = v;
    }
    __syncthreads();
    ntt_dif(x, lg, twf);
    uint64_t ah[kNttMaxPerThread];
#pragma unroll
    for (uint32_t e = 0; e < kNttMaxPerThread; e++)
        if (e < E) ah[e] = x[tid + T * e];
    __syncthreads();

    // ---- forward transform of b, pointwise product (scaled by L^-1)
    for (uint32_t k = tid; k < L; k += T) {
        uint64_t v;
        if (MODE == 0)
            v = (k < 4 * m) ? ((gb[k >> 2] >> (16 * (k & 3))) & 0xFFFFull) : 0ull;
        else
            v = gb[k];
        x[k] = v;
    }
    __syncthreads();
    ntt_dif(x, lg, twf);
#pragma unroll
    for (uint32_t e = 0; e < kNttMaxPerThread; e++) {
        if (e < E) {
            const uint32_t pos = tid + T * e;
            x[pos] = gl_mul(gl_mul(x[pos], ah[e]), inv_len);
        }
    }
    __syncthreads();
    ntt_dit(x, lg, twi);

    if (MODE == 1) {
        uint64_t* gr = r + (size_t)blockIdx.x * L;
        for (uint32_t k = tid; k < L; k += T) gr[k] = x[k];
        return;
    }

    // ---- exact recombination: limb j = sum_s c_{4j+s} 2^(16 s), split low/high
    const uint32_t outl = 2 * m;
    uint64_t lo_v[kNttMaxLimbsPerThread], hi_v[kNttMaxLimbsPerThread];
#pragma unroll
    for (uint32_t e = 0; e < kNttMaxLimbsPerThread; e++) {
        const uint32_t j = tid + T * e;
        uint64_t lo = 0, hi = 0;
        if (j < outl) {
            const uint64_t c0 = x[4 * j], c1 = x[4 * j + 1], c2 = x[4 * j + 2], c3 = x[4 * j + 3];
            unsigned __int128 acc = (unsigned __int128)c0 + ((unsigned __int128)c1 << 16) +
                                    ((unsigned __int128)c2 << 32) + ((unsigned __int128)c3 << 48);
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
    for (uint32_t e = 0; e < kNttMaxLimbsPerThread; e++) {
        const uint32_t j = tid + T * e;
        if (j < outl) {
            Lb[j] = lo_v[e];
            if (j + 1 < outl) Hb[j + 1] = hi_v[e];
        }
    }
    __syncthreads();

    constexpr int R = 2;
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
    uint64_t* gr = r + (size_t)blockIdx.x * rl;  // rl = 2m (full) or m (low half)
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        if (k < rl) gr[k] = rows.x0[v];
        if (k + 1 < rl) gr[k + 1] = rows.x1[v];
    }
}

// ---------------------------------------------------------------------------
// Register-pass kernel for L = 2^LG, LG in [10, 14]: T = L / V threads with
// V values each (NttPlan). a's transform stays in registers; b's last
// forward pass, the pointwise product (with L^-1) and the first inverse pass
// run on the same registers, so each transform costs NP - 1 LDS round trips.

// Coefficient i of operand `g`: 16-bit pieces of the m-limb integer for
// products (MODE 0), raw values for the convolution mode (MODE 1).
template <int MODE>
__device__ __forceinline__ uint64_t ntt_coef(const uint64_t* g, uint32_t i, uint32_t m) {
    if (MODE == 0) return i < 4 * m ? (g[i >> 2] >> (16 * (i & 3))) & 0xFFFFull : 0ull;
    return g[i];
}

// Forward transform of one operand read straight from global memory (in the
// first pass's grouping); returns with the last pass's values in v.
// SCALE: multiply the input by `scale` first (the inverse's L^-1, applied to
// a's nonzero inputs instead of to all L transformed values).
// SYNC: a barrier between loading the inputs and the first pass's LDS
// stores (for an input that lives in the transform's own LDS area).
template <int LG, int V, int MODE, bool SCALE = false, bool SYNC = false, int FM = 0>
__device__ __forceinline__ void ntt_fwd(const uint64_t* g, uint32_t m, uint64_t* x, uint64_t (&v)[V],
                                        const uint64_t* __restrict__ twf, uint64_t scale = 0) {
    using Pl = NttPlan<LG, V>;
    const uint32_t tid = opaque_u32(threadIdx.x);
    constexpr int R0 = Pl::radix(0);
    // products: the 4m <= L/2 coefficients sit in the lower half, which is
    // where k < 2^(R0-1) of the first pass's groups point
    constexpr bool ZHI = MODE == 0;
#pragma unroll
    for (int j = 0; j < (V >> R0); j++)
#pragma unroll
        for (int k = 0; k < (1 << R0); k++) {
            if (ZHI && k >= (1 << (R0 - 1))) {
                v[(j << R0) + k] = 0ull;
            } else {
                const uint64_t c = ntt_coef<MODE>(g, ntt_group_index<R0>(LG, 0, tid, Pl::T, j, k), m);
                v[(j << R0) + k] = SCALE ? gl_mulv<FM>(c, scale) : c;
            }
        }
    if (SYNC) __syncthreads();
    NttFwd<LG, V, 0, ZHI, FM>::run(x, v, tid, twf);
}

// Exact recombination of the 16-bit-spaced coefficients of a product,
// natural order in x (padded, after NttInv): limb j is c[4j] + c[4j+1] 2^16 +
// c[4j+2] 2^32 + c[4j+3] 2^48 (< 2^112), split into L (low 64 bits, at limb
// j) and H (high bits, at limb j + 1) in x, then one carry-propagating row
// addition L + H: limbs [0, outl) of the product in rows (R = V / 8 rows of
// 128 limbs per wave over L / 4 >= outl limbs; limb k = wave R 128 + v 128 +
// 2 lane). The low outl limbs need nothing above them: carries only go up.
template <int LG, int V>
__device__ __forceinline__ void ntt_recombine_rows(uint64_t* x, uint32_t outl, Rows<V / 8>& rows, uint32_t* lds_gp) {
    constexpr uint32_t T = NttPlan<LG, V>::T;
    constexpr uint32_t LPT = V / 4;  // limbs per thread: outl <= L / 4 = LPT * T
    constexpr int R = V / 8;         // rows per wave: (T / 64) * R * 128 = L / 4 limbs
    const uint32_t tid = threadIdx.x;
    uint64_t lo_v[LPT], hi_v[LPT];
#pragma unroll
    for (uint32_t e = 0; e < LPT; e++) {
        const uint32_t j = tid + T * e;
        uint64_t lo = 0, hi = 0;
        if (j < outl) {
            const uint64_t c0 = x[ntt_pad(4 * j)], c1 = x[ntt_pad(4 * j + 1)], c2 = x[ntt_pad(4 * j + 2)],
                           c3 = x[ntt_pad(4 * j + 3)];
            unsigned __int128 acc = (unsigned __int128)c0 + ((unsigned __int128)c1 << 16) +
                                    ((unsigned __int128)c2 << 32) + ((unsigned __int128)c3 << 48);
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
    for (uint32_t e = 0; e < LPT; e++) {
        const uint32_t j = tid + T * e;
        if (j < outl) {
            Lb[j] = lo_v[e];
            if (j + 1 < outl) Hb[j + 1] = hi_v[e];
        }
    }
    __syncthreads();
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
#pragma unroll
    for (int vv = 0; vv < R; vv++) {
        const uint32_t k = k_first + vv * kRowLimbs;
        rows.x0[vv] = k < outl ? Lb[k] : 0ull;
        rows.x1[vv] = k + 1 < outl ? Lb[k + 1] : 0ull;
        rows.y0[vv] = k < outl ? Hb[k] : 0ull;
        rows.y1[vv] = k + 1 < outl ? Hb[k + 1] : 0ull;
        rows.s0[vv] = (k == 0);
        rows.s1[vv] = false;
    }
    block_add_rows<R>(rows, lds_gp);
}

// rows (as left by ntt_recombine_rows) += y[0, n) mod B^n, y in global
// memory or LDS. Every thread of the block must call it.
template <int R>
__device__ __forceinline__ void ntt_rows_add(Rows<R>& rows, const uint64_t* y, uint32_t n, uint32_t* lds_gp) {
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
#pragma unroll
    for (int vv = 0; vv < R; vv++) {
        const uint32_t k = k_first + vv * kRowLimbs;
        rows.y0[vv] = k < n ? y[k] : 0ull;
        rows.y1[vv] = k + 1 < n ? y[k + 1] : 0ull;
        rows.s0[vv] = (k == 0);
        rows.s1[vv] = false;
    }
    block_add_rows<R>(rows, lds_gp);
}

// Limbs [0, n) of rows to dst (global memory or LDS).
template <int R>
__device__ __forceinline__ void ntt_store_rows(const Rows<R>& rows, uint64_t* dst, uint32_t n) {
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
#pragma unroll
    for (int vv = 0; vv < R; vv++) {
        const uint32_t k = k_first + vv * kRowLimbs;
        if (k < n) dst[k] = rows.x0[vv];
        if (k + 1 < n) dst[k + 1] = rows.x1[vv];
    }
}

template <int LG, int V, int MODE>
__global__ __launch_bounds__((1u << LG) / V) void ntt_fast_kernel(uint32_t m, uint32_t rl,
                                                                  const uint64_t* __restrict__ a,
                                                                  const uint64_t* __restrict__ b,
                                                                  uint64_t* __restrict__ r,
                                                                  const uint64_t* __restrict__ twf,
                                                                  const uint64_t* __restrict__ twi,
                                                                  uint64_t inv_len) {
    using Pl = NttPlan<LG, V>;
    constexpr uint32_t L = 1u << LG;
    constexpr uint32_t T = Pl::T;
    extern __shared__ __align__(16) uint64_t x[];
    __shared__ uint32_t lds_gp[16];
    const uint32_t tid = threadIdx.x;
    const uint64_t* ga = a + (size_t)blockIdx.x * (MODE == 0 ? m : L);
    const uint64_t* gb = b + (size_t)blockIdx.x * (MODE == 0 ? m : L);

    uint64_t ah[V], v[V];
    ntt_fwd<LG, V, MODE, true>(ga, m, x, ah, twf, inv_len);  // transform of a / L
    __syncthreads();  // every thread is done reading LDS for a's last pass
    ntt_fwd<LG, V, MODE>(gb, m, x, v, twf);
#pragma unroll
    for (int k = 0; k < V; k++) v[k] = gl_mul(v[k], ah[k]);
    // b's last forward pass read only this wave's block of LDS, which is all
    // the inverse's first store writes
    ntt_wave_sync();
    NttInv<LG, V, Pl::NP - 1>::run(x, v, opaque_u32(tid), twi);

    if (MODE == 1) {
        uint64_t* gr = r + (size_t)blockIdx.x * L;
        for (uint32_t k = tid; k < L; k += T) gr[k] = x[ntt_pad(k)];
        return;
    }

    Rows<V / 8> rows;
    ntt_recombine_rows<LG, V>(x, 2 * m, rows, lds_gp);
    ntt_store_rows<V / 8>(rows, r + (size_t)blockIdx.x * rl, rl);  // rl = 2m (full) or m (low half)
}

// The polynomial kernel's field multiplications reduce through the
// multiply-add (gl_mulv<0>), except in the inverse transform of a^2, which
// runs while NTT(a) is held for ab: there the carry-chain reduction
// (gl_mulv<1>), with fewer live registers, keeps the kernel within 128
// VGPRs without spilling (126; the multiply-add form everywhere spills 6).
constexpr int kPolyFM = 0;
constexpr int kPolyHeldFM = 1;

// The reference's polyFttKer (cuda/ker-fft-mul.cu.h:456-509) in one kernel:
// r = ((a^2 + b)(b^2 + b) + ab) mod B^m per instance (one per workgroup),
// the steps mod B^m in the reference's order, with eight L-point
// transforms (L >= 8m) and at most two transformed operands in registers:
//   P = NTT(a) / L;  a^2 + b     -> s1 in LDS (after the transform area)
//   Q = NTT(b) / L;  ab          -> r (held there until the end)
//                    b^2 + b     -> s2 in the transform area
//   Q = NTT(s2) / L, P = NTT(s1) / L;  s1 s2 + ab -> r
// Every pointwise product is of two transforms of inputs scaled by 1 / L,
// times L (a shift by LG bits), so the unscaled inverse returns the exact
// product coefficients; the low m limbs of each product need nothing above
// them (carries only go up). No scratch memory: intermediates stay in LDS
// and the instance's own m output limbs.
template <int LG, int V>
__global__ __launch_bounds__((1u << LG) / V) void poly_ntt_kernel(uint32_t m, const uint64_t* __restrict__ a,
                                                                  const uint64_t* __restrict__ b,
                                                                  uint64_t* __restrict__ r,
                                                                  const uint64_t* __restrict__ twf,
                                                                  const uint64_t* __restrict__ twi, uint64_t inv_len) {
    using Pl = NttPlan<LG, V>;
    constexpr uint32_t L = 1u << LG;
    constexpr int R = V / 8;
    extern __shared__ __align__(16) uint64_t x[];  // transform area (L + L / 32 words), then s1 (m words)
    __shared__ uint32_t lds_gp[16];
    uint64_t* s1 = x + (L + L / 32);
    const uint64_t* ga = a + (size_t)blockIdx.x * m;
    const uint64_t* gb = b + (size_t)blockIdx.x * m;
    uint64_t* gr = r + (size_t)blockIdx.x * m;
    uint64_t P[V], Q[V];
    Rows<R> rows;

    // ---- s1 = a^2 + b
    ntt_fwd<LG, V, 0, true, false, kPolyFM>(ga, m, x, P, twf, inv_len);
#pragma unroll
    for (int k = 0; k < V; k++) Q[k] = gl_mul_pow2(gl_mulv<kPolyFM>(P[k], P[k]), LG);
    ntt_wave_sync();  // the last forward pass and the first inverse store: the wave's own block
    NttInv<LG, V, Pl::NP - 1, kPolyHeldFM>::run(x, Q, opaque_u32(threadIdx.x), twi);
    ntt_recombine_rows<LG, V>(x, m, rows, lds_gp);
    ntt_rows_add<R>(rows, gb, m, lds_gp);
    ntt_store_rows<R>(rows, s1, m);
    __syncthreads();

    // ---- ab (kept in r)
    ntt_fwd<LG, V, 0, true, false, kPolyFM>(gb, m, x, Q, twf, inv_len);
#pragma unroll
    for (int k = 0; k < V; k++) P[k] = gl_mul_pow2(gl_mulv<kPolyFM>(P[k], Q[k]), LG);
    ntt_wave_sync();
    NttInv<LG, V, Pl::NP - 1, kPolyFM>::run(x, P, opaque_u32(threadIdx.x), twi);
    ntt_recombine_rows<LG, V>(x, m, rows, lds_gp);
    ntt_store_rows<R>(rows, gr, m);
    __syncthreads();

    // ---- s2 = b^2 + b, into the transform area
#pragma unroll
    for (int k = 0; k < V; k++) Q[k] = gl_mul_pow2(gl_mulv<kPolyFM>(Q[k], Q[k]), LG);
    NttInv<LG, V, Pl::NP - 1, kPolyFM>::run(x, Q, opaque_u32(threadIdx.x), twi);
    ntt_recombine_rows<LG, V>(x, m, rows, lds_gp);
    ntt_rows_add<R>(rows, gb, m, lds_gp);
    ntt_store_rows<R>(rows, x, m);
    __syncthreads();

    // ---- r = s1 s2 + ab
    ntt_fwd<LG, V, 0, true, true, kPolyFM>(x, m, x, Q, twf, inv_len);
    __syncthreads();  // every thread is done reading LDS for s2's last pass
    ntt_fwd<LG, V, 0, true, false, kPolyFM>(s1, m, x, P, twf, inv_len);
#pragma unroll
    for (int k = 0; k < V; k++) P[k] = gl_mul_pow2(gl_mulv<kPolyFM>(P[k], Q[k]), LG);
    ntt_wave_sync();
    NttInv<LG, V, Pl::NP - 1, kPolyFM>::run(x, P, opaque_u32(threadIdx.x), twi);
    ntt_recombine_rows<LG, V>(x, m, rows, lds_gp);
    ntt_rows_add<R>(rows, gr, m, lds_gp);
    ntt_store_rows<R>(rows, gr, m);
}

template <int LG, int V>
static hipError_t launch_poly_ntt_lg_v(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                                       const NttTables& t, hipStream_t s) {
    constexpr uint32_t L = 1u << LG;
    static_assert(L / V >= 64 && L / V <= 1024, "one to sixteen waves per block");
    const size_t lds = (size_t)(L + L / 32 + m) * 8;
    const void* fn = (const void*)poly_ntt_kernel<LG, V>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    poly_ntt_kernel<LG, V><<<n, L / V, lds, s>>>(m, a, b, r, t.d_fwd, t.d_inv, t.inv_len);
    return hipGetLastError();
}

template <int LG, int V, int MODE>
static hipError_t launch_fast(uint32_t n, uint32_t m, uint32_t rl, const uint64_t* a, const uint64_t* b, uint64_t* r,
                              const NttTables& t, hipStream_t s) {
    constexpr uint32_t L = 1u << LG;
    static_assert(L / V >= 64 && L / V <= 1024, "one to sixteen waves per block");
    const size_t lds = (size_t)(L + L / 32) * 8;
    const void* fn = (const void*)ntt_fast_kernel<LG, V, MODE>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    ntt_fast_kernel<LG, V, MODE><<<n, L / V, lds, s>>>(m, rl, a, b, r, t.d_fwd, t.d_inv, t.inv_len);
    return hipGetLastError();
}

// Values per thread: tunable per transform length ("ntt.vals@L", 16 or 32;
// 0 for the default). 16 by default: T = L / 16 threads (up to 1024 at 2^14
// points, 4 waves per SIMD) hide latency better than 32 values per thread,
// and both transforms fit in registers either way.
static uint32_t ntt_vals(uint32_t lg) {
    const uint32_t def = 16;
    const int64_t v = tune_get_class("ntt.vals", 1u << lg, def);
    return (v == 16 || (v == 32 && lg >= 11)) ? (uint32_t)v : def;
}

template <int MODE>
static hipError_t launch_fast_lg(uint32_t lg, uint32_t n, uint32_t m, uint32_t rl, const uint64_t* a,
                                 const uint64_t* b, uint64_t* r, const NttTables& t, hipStream_t s) {
    const bool v32 = ntt_vals(lg) == 32;
    switch (lg) {
        case 10: return launch_fast<10, 16, MODE>(n, m, rl, a, b, r, t, s);
        case 11: return v32 ? launch_fast<11, 32, MODE>(n, m, rl, a, b, r, t, s) : launch_fast<11, 16, MODE>(n, m, rl, a, b, r, t, s);
        case 12: return v32 ? launch_fast<12, 32, MODE>(n, m, rl, a, b, r, t, s) : launch_fast<12, 16, MODE>(n, m, rl, a, b, r, t, s);
        case 13: return v32 ? launch_fast<13, 32, MODE>(n, m, rl, a, b, r, t, s) : launch_fast<13, 16, MODE>(n, m, rl, a, b, r, t, s);
        case 14: return v32 ? launch_fast<14, 32, MODE>(n, m, rl, a, b, r, t, s) : launch_fast<14, 16, MODE>(n, m, rl, a, b, r, t, s);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Host side.

static inline uint64_t h_mulmod(uint64_t a, uint64_t b) {
    unsigned __int128 t = (unsigned __int128)a * b;
    return (uint64_t)(t % kGlP);
}
static uint64_t h_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = h_mulmod(r, a);
        a = h_mulmod(a, a);
        e >>= 1;
    }
    return r;
}

void ntt_host_tables(uint32_t lg, uint64_t* fwd, uint64_t* inv, uint64_t* inv_len) {
    const uint64_t L = 1ull << lg;
    // 7 generates the multiplicative group; 7^((p-1)/64) = 8^13 and
    // 13 * 5 = 1 (mod 64), so w = 7^(5 (p-1) / L) has w^(L/64) = 8 and the
    // twiddles of spans h <= 32 are powers of two (2^96 = -1 mod p):
    // w^(j L / 2h) = 2^(96 j / h). The kernels' last passes rely on it.
    const uint64_t w = h_pow(h_pow(7, (kGlP - 1) / L), 5);
    const uint64_t wi = h_pow(w, kGlP - 2);
    uint64_t f = 1, g = 1;
    for (uint64_t i = 0; i < L; i++) {
        fwd[i] = f;
        inv[i] = g;
        f = h_mulmod(f, w);
        g = h_mulmod(g, wi);
    }
    *inv_len = h_pow(L, kGlP - 2);
}

uint32_t ntt_lg_for_limbs(uint32_t m) {
    uint32_t lg = 6;
    while ((1ull << lg) < 8ull * m) lg++;
    return lg;
}

static uint32_t ntt_threads(uint32_t lg) {
    uint32_t L = 1u << lg;
    uint32_t T = L / 16;
    if (T < 64) T = 64;
    if (T > 1024) T = 1024;
    return T;
}

hipError_t launch_mul_ntt(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                          const NttTables& t, hipStream_t s, bool lo) {
    const uint32_t rl = lo ? m : 2 * m;
    if (t.lg > kNttMaxLg || t.lg < 6) return hipErrorInvalidValue;
    const uint32_t T = ntt_threads(t.lg);
    const size_t lds = ((size_t)1 << t.lg) * 8;
    if ((1u << t.lg) / T > kNttMaxPerThread) return hipErrorInvalidValue;
    if (2ull * m > (uint64_t)kNttMaxLimbsPerThread * T) return hipErrorInvalidValue;
    if (t.lg >= 10 && tune_get("ntt.fast", 1)) return launch_fast_lg<0>(t.lg, n, m, rl, a, b, r, t, s);
    hipError_t e = hipFuncSetAttribute((const void*)ntt_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    ntt_kernel<0><<<n, T, lds, s>>>(m, rl, t.lg, a, b, r, t.d_fwd, t.d_inv, t.inv_len);
    return hipGetLastError();
}

hipError_t launch_ntt_conv(uint32_t n, const uint64_t* a, const uint64_t* b, uint64_t* r, const NttTables& t,
                           hipStream_t s) {
    if (t.lg > kNttMaxLg || t.lg < 6) return hipErrorInvalidValue;
    const uint32_t T = ntt_threads(t.lg);
    const size_t lds = ((size_t)1 << t.lg) * 8;
    if (t.lg >= 10 && tune_get("ntt.fast", 1)) return launch_fast_lg<1>(t.lg, n, 0, 0, a, b, r, t, s);
    hipError_t e = hipFuncSetAttribute((const void*)ntt_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    ntt_kernel<1><<<n, T, lds, s>>>(0, 0, t.lg, a, b, r, t.d_fwd, t.d_inv, t.inv_len);
    return hipGetLastError();
}

uint32_t poly_ntt_lg(uint32_t m) {
    const uint32_t lg = ntt_lg_for_limbs(m);
    return lg < 10 ? 10 : lg;  // the register-pass kernel's shortest transform
}

// r = ((a^2 + b)(b^2 + b) + ab) mod B^m with NTT products (reference
// polyFttKer, cuda/ker-fft-mul.cu.h:456-509: the same sequence of three
// additions and four products as polyKerQ, every step mod B^m): one kernel
// (poly_ntt_kernel) on tables of 2^poly_ntt_lg(m) points.
hipError_t launch_poly_ntt(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                           const NttTables& t, hipStream_t s) {
    if (t.lg != poly_ntt_lg(m) || t.lg > kNttMaxLg || 8ull * m > (1ull << t.lg)) return hipErrorInvalidValue;
    const bool v32 = ntt_vals(t.lg) == 32;
    switch (t.lg) {
        case 10: return launch_poly_ntt_lg_v<10, 16>(n, m, a, b, r, t, s);
        case 11: return v32 ? launch_poly_ntt_lg_v<11, 32>(n, m, a, b, r, t, s) : launch_poly_ntt_lg_v<11, 16>(n, m, a, b, r, t, s);
        case 12: return v32 ? launch_poly_ntt_lg_v<12, 32>(n, m, a, b, r, t, s) : launch_poly_ntt_lg_v<12, 16>(n, m, a, b, r, t, s);
        case 13: return v32 ? launch_poly_ntt_lg_v<13, 32>(n, m, a, b, r, t, s) : launch_poly_ntt_lg_v<13, 16>(n, m, a, b, r, t, s);
        case 14: return v32 ? launch_poly_ntt_lg_v<14, 32>(n, m, a, b, r, t, s) : launch_poly_ntt_lg_v<14, 16>(n, m, a, b, r, t, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace midint
