// Classical (schoolbook) batched multiplication for gfx950.
//
// mul_kernel<FULL=true>  : r[0..2m) = a*b             (full product)
// mul_kernel<FULL=false> : r[0..m)  = a*b mod B^m     (reference bmulKerQ,
//                          cuda/ker-classic-mul.cu.h:468-498)
//
// Operands are viewed as n = 2m u32 words (zero-padded to n_pad, a multiple
// of 2Q). Column c of the product is sum_{i+j=c} a_i b_j. As in the
// reference (wrapperConvQ, cuda/ker-classic-mul.cu.h:406-423) each lane owns
// two groups of Q consecutive columns chosen so that their term counts add up
// to the same total for every lane. Unlike the reference, the two groups are
// walked in ONE loop of equal trip count for all lanes: lane t works on its
// first group for blocks 0..t and on its second group afterwards, so no lane
// of a wavefront idles while another finishes a longer column range. At the
// switch the lane parks its accumulators (in registers with RPARK, else in
// an LDS snapshot area) and keeps accumulating; the second group is
// recovered by subtraction.
//
// Each loop block processes Q a-words x Q columns = Q^2 32x32->64 products
// (Q = 4, 8 or 16) from Q/4 16-byte broadcast LDS reads of a and Q/2
// 16-byte reads of a sliding 2Q-word window of b. For Q >= 8 the window
// operand is stored skewed (4 pad words per Q words, skew_idx) so that the
// lanes' window reads hit distinct LDS banks; unskewed, Q = 8 windows 32
// bytes apart conflict pairwise, which cost a fifth of the kernel's time.
// Column sums are 64-bit accumulators plus a 32-bit overflow count (the
// reference's computeIter64, :11-26, at u32 granularity).
//
// Column sums are split into low words L and high words H shifted by one
// group (reference from4Reg2ShmQ, :299-350) and normalised with one
// carry-propagating addition L + H (reference baddRegs, :462) done by the
// wave-ballot row adder.
#include "add_core.hpp"
#include "midint_internal.h"
#include "mul_core.hpp"

namespace midint {

// Dynamic LDS a block may request: the 160 KB of a CU minus room for the
// kernels' static LDS (the 64-byte row-adder scratch).
constexpr size_t kMulLdsMax = 160 * 1024 - 256;

struct MulLayout {
    uint32_t Q;       // columns per lane group
    uint32_t n_pad;   // padded operand length in u32 words (multiple of 2Q)
    uint32_t T;       // lanes per instance
    uint32_t P;       // product limbs computed per instance
    uint32_t words;   // LDS u32 words per instance slot (conv phase)
    uint32_t ipb;     // instances per block
    uint32_t threads; // block size
    bool rpark;       // first column group parked in registers (no LDS snapshot area)
};

// Words of the W region of a slot: 2Q-padded window operand, skewed for Q >= 8.
template <int Q>
__host__ __device__ constexpr uint32_t mul_w_len(uint32_t n_pad) {
    return Q >= 8 ? skew_len<Q>(n_pad + 2 * Q) : n_pad + 2 * Q;
}

template <bool FULL, int Q, bool PF, bool RPARK>
__global__ __launch_bounds__(Q == 16 ? 256 : 1024) void mul_kernel(uint32_t n_inst, uint32_t m, uint32_t n_pad, uint32_t T,
                                                   uint32_t P, uint32_t words, uint32_t ipb, uint32_t out_len,
                                                   int32_t m_shift, const uint64_t* __restrict__ a,
                                                   const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                   uint64_t sa, uint64_t sb, uint64_t sr) {
    // sa, sb, sr: instance strides (limbs) of a, b and r; m, m, out_len for
    // the batch-major layout
    extern __shared__ __align__(16) uint64_t lds64[];
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
    __shared__ uint32_t lds_gp[16];

    const uint32_t tid = threadIdx.x;
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);

    // ---- stage operands into LDS: A[n_pad], then the window operand W with
    // W[(Q-1) + j] = b_j, in the skewed layout for Q >= 8 (skew_idx). Only
    // the pad words are cleared: A[2m, n_pad), W[0, Q-1) and
    // W[Q-1+2m, n_pad+2Q) of every slot (the snapshot area needs no init).
    constexpr bool SKEW = Q >= 8;
    auto wi = [](uint32_t y) { return SKEW ? skew_idx<Q>(y) : y; };
    {
        const uint32_t padA = n_pad - 2 * m;
        const uint32_t padB = (Q - 1) + (n_pad + Q + 1 - 2 * m);
        const uint32_t per = padA + padB;
        for (uint32_t f = tid; f < ipb * per; f += blockDim.x) {
            const uint32_t s = f / per;
            const uint32_t x = f - s * per;
            uint32_t* A = lds + s * words;
            uint32_t* Wp = A + n_pad;
            if (x < padA)
                A[2 * m + x] = 0u;
            else if (x < padA + (Q - 1))
                Wp[wi(x - padA)] = 0u;
            else
                Wp[wi((Q - 1) + 2 * m + (x - padA - (Q - 1)))] = 0u;
        }
    }
    {
        const uint32_t total = ninst * m;
        const uint64_t* ga = a + inst_base * sa;
        const uint64_t* gb = b + inst_base * sb;
        for (uint32_t f = tid; f < total; f += blockDim.x) {
            const uint32_t s = m_shift >= 0 ? (f >> m_shift) : f / m;
            const uint32_t i = f - s * m;
            const uint64_t va = ga[s * sa + i], vb = gb[s * sb + i];
            uint32_t* A = lds + s * words;
            uint32_t* Wp = A + n_pad;
            *reinterpret_cast<uint64_t*>(A + 2 * i) = va;
            Wp[wi((Q - 1) + 2 * i)] = (uint32_t)vb;
            Wp[wi((Q - 1) + 2 * i + 1)] = (uint32_t)(vb >> 32);
        }
        // slots past the last instance of the grid hold zero operands
        for (uint32_t f = ninst * m + tid; f < ipb * m; f += blockDim.x) {
            const uint32_t s = m_shift >= 0 ? (f >> m_shift) : f / m;
            const uint32_t i = f - s * m;
            uint32_t* A = lds + s * words;
            uint32_t* Wp = A + n_pad;
            *reinterpret_cast<uint64_t*>(A + 2 * i) = 0ull;
            Wp[wi((Q - 1) + 2 * i)] = 0u;
            Wp[wi((Q - 1) + 2 * i + 1)] = 0u;
        }
    }
    __syncthreads();

    // ---- convolution
    const uint32_t slot = tid / T;
    const uint32_t t = tid - slot * T;
    const bool active = slot < ipb;
    uint32_t k1 = Q * t, k2, i2, NB;
    if (FULL) {
        k2 = n_pad + Q * t;
        i2 = Q * t;
        NB = T + 1;
    } else {
        k2 = n_pad - Q - Q * t;
        i2 = 0;
        NB = 2 * T + 1;
    }
    ColGroups<Q> g;
    if (active) {
        const uint32_t* A = lds + slot * words;
        uint32_t* snap = lds + slot * words + n_pad + mul_w_len<Q>(n_pad) + 3 * Q * t;
        conv_lane<Q, PF, false, SKEW, RPARK>(A, A + n_pad, k1, k2, i2, t + 1, NB, snap, g);
    }
    __syncthreads();  // everyone is done with A/W/snap; reuse LDS for L/H

    uint64_t* Lb = lds64;
    uint64_t* Hb = lds64 + ipb * P;
    if (active) lane_store<Q>(g, true, k1, k2, Lb + slot * P, Hb + slot * P, P);
    __syncthreads();

    // ---- normalisation: L + H over ipb segments of P limbs (Q/2 rows per wave)
    constexpr int R = Q / 2;
    const uint32_t Lw = ipb * P;
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
    Rows<R> rows;
    uint32_t off0[R], inst0[R];
    {
        SegPos sp = segpos_init(k_first, P);
        const uint32_t sq = kRowLimbs / P, sr = kRowLimbs % P;
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            rows.x0[v] = k < Lw ? Lb[k] : 0ull;
            rows.x1[v] = k + 1 < Lw ? Lb[k + 1] : 0ull;
            rows.y0[v] = k < Lw ? Hb[k] : 0ull;
            rows.y1[v] = k + 1 < Lw ? Hb[k + 1] : 0ull;
            rows.s0[v] = (sp.off == 0) && (k < Lw);
            rows.s1[v] = (sp.off + 1 == P) && (k + 1 < Lw);
            off0[v] = sp.off;
            inst0[v] = sp.inst;
            segpos_next(&sp, P, sq, sr);
        }
    }
    block_add_rows<R>(rows, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const uint32_t o0 = off0[v];
        const uint32_t i0 = inst0[v];
        if (k < Lw && i0 < ninst && o0 < out_len) r[(inst_base + i0) * sr + o0] = rows.x0[v];
        const bool wrap = (o0 + 1 == P);
        const uint32_t o1 = wrap ? 0u : o0 + 1;
        const uint32_t i1 = i0 + (wrap ? 1u : 0u);
        if (k + 1 < Lw && i1 < ninst && o1 < out_len) r[(inst_base + i1) * sr + o1] = rows.x1[v];
    }
}

// ---------------------------------------------------------------------------
// Digit kernel: the same column-group schoolbook on 28-bit digits (Q = 8).
// Every product is below 2^56, so a column accumulator takes one
// v_mad_u64_u32 per product and no carry instruction; 255 / 8 = 31 blocks
// after each flush the high half moves to the overflow word. For 8192-bit
// operands that is 304^2 (padded) mads instead of 2 x 256^2 mad + addc.
// Column groups are 224 bits wide and tile two limb vectors E (even groups)
// and O (odd groups); the product is E + O (digit_group_store).

// Digit j (< D) of an operand of w32 = 2m u32 words: bits [28j, 28j + 28),
// one funnel shift of the two words that hold them (the word past the end
// reads as zero).
__device__ __forceinline__ uint32_t digit_at(const uint32_t* __restrict__ g32, uint32_t w32, uint32_t j) {
    const uint32_t bit = kDigitBits * j;
    const uint32_t w = bit >> 5;
    const uint32_t lo = g32[w];
    const uint32_t hi = w + 1 < w32 ? g32[w + 1] : 0u;
    return __builtin_amdgcn_alignbit(hi, lo, bit & 31) & ((1u << kDigitBits) - 1);
}

template <bool FULL, bool PF, bool RPARK>
__global__ __launch_bounds__(1024) void mul_digit_kernel(uint32_t n_inst, uint32_t m, uint32_t D, uint32_t n_pad,
                                                         uint32_t T, uint32_t words, uint32_t ipb, uint32_t P,
                                                         const uint64_t* __restrict__ a,
                                                         const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                         uint64_t sa, uint64_t sb, uint64_t sr) {
    constexpr int Q = 8;
    extern __shared__ __align__(16) uint64_t lds64[];
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
    __shared__ uint32_t lds_gp[16];

    const uint32_t tid = threadIdx.x;
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);

    // ---- stage digits: A[0, n_pad) and W[y] = digit y - (Q-1) of b for
    // y < n_pad + 2Q in the skewed layout (skew_idx), zero outside the
    // operand (and for slots past the grid)
    for (uint32_t s = 0; s < ipb; s++) {
        const bool live = s < ninst;
        const uint32_t* ga = reinterpret_cast<const uint32_t*>(a + (inst_base + s) * sa);
        const uint32_t* gb = reinterpret_cast<const uint32_t*>(b + (inst_base + s) * sb);
        uint32_t* A = lds + s * words;
        uint32_t* Wp = A + n_pad;
        for (uint32_t x = tid; x < n_pad; x += blockDim.x) A[x] = live && x < D ? digit_at(ga, 2 * m, x) : 0u;
        for (uint32_t y = tid; y < n_pad + 2 * Q; y += blockDim.x) {
            const uint32_t j = y - (Q - 1);  // wraps for y < Q-1
            Wp[skew_idx<Q>(y)] = live && j < D ? digit_at(gb, 2 * m, j) : 0u;
        }
    }
    __syncthreads();

    // ---- convolution (same lane / group pairing as mul_kernel)
    const uint32_t slot = tid / T;
    const uint32_t t = tid - slot * T;
    const bool active = slot < ipb;
    uint32_t k1 = Q * t, k2, i2, NB;
    if (FULL) {
        k2 = n_pad + Q * t;
        i2 = Q * t;
        NB = T + 1;
    } else {
        k2 = n_pad - Q - Q * t;
        i2 = 0;
        NB = 2 * T + 1;
    }
    ColGroups<Q> g;
    if (active) {
        const uint32_t* A = lds + slot * words;
        uint32_t* snap = lds + slot * words + n_pad + mul_w_len<Q>(n_pad) + 3 * Q * t;
        conv_lane<Q, PF, true, true, RPARK>(A, A + n_pad, k1, k2, i2, t + 1, NB, snap, g);
    }
    __syncthreads();  // everyone is done with A/B/snap; reuse LDS for E/O

    uint64_t* Eb = lds64;
    uint64_t* Ob = lds64 + ipb * P;
    if (active) digit_lane_store(g, true, k1, k2, Eb + slot * P, Ob + slot * P, P);
    __syncthreads();

    // ---- normalisation: E + O over ipb segments of P limbs (4 rows per wave)
    constexpr int R = 4;
    const uint32_t Lw = ipb * P;
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
    Rows<R> rows;
    uint32_t off0[R], inst0[R];
    {
        SegPos sp = segpos_init(k_first, P);
        const uint32_t sq = kRowLimbs / P, srr = kRowLimbs % P;
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            rows.x0[v] = k < Lw ? Eb[k] : 0ull;
            rows.x1[v] = k + 1 < Lw ? Eb[k + 1] : 0ull;
            rows.y0[v] = k < Lw ? Ob[k] : 0ull;
            rows.y1[v] = k + 1 < Lw ? Ob[k + 1] : 0ull;
            rows.s0[v] = (sp.off == 0) && (k < Lw);
            rows.s1[v] = (sp.off + 1 == P) && (k + 1 < Lw);
            off0[v] = sp.off;
            inst0[v] = sp.inst;
            segpos_next(&sp, P, sq, srr);
        }
    }
    block_add_rows<R>(rows, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const uint32_t o0 = off0[v];
        const uint32_t i0 = inst0[v];
        if (k < Lw && i0 < ninst) r[(inst_base + i0) * sr + o0] = rows.x0[v];
        const bool wrap = (o0 + 1 == P);
        const uint32_t o1 = wrap ? 0u : o0 + 1;
        const uint32_t i1 = i0 + (wrap ? 1u : 0u);
        if (k + 1 < Lw && i1 < ninst) r[(inst_base + i1) * sr + o1] = rows.x1[v];
    }
}

struct DigitLayout {
    uint32_t D, n_pad, T, words, ipb, threads, P;
    bool rpark;
};

static bool digit_layout(uint32_t m, bool full, bool rpark, DigitLayout* L) {
    constexpr uint32_t Q = 8;
    if (m == 0 || m > kMulMaxLimbs) return false;
    L->D = (64 * m + kDigitBits - 1) / kDigitBits;
    L->n_pad = (L->D + 2 * Q - 1) / (2 * Q) * (2 * Q);
    L->T = full ? L->n_pad / Q : L->n_pad / (2 * Q);
    L->P = full ? 2 * m : m;
    L->rpark = rpark;
    L->words = L->n_pad + mul_w_len<Q>(L->n_pad) + (rpark ? 0 : 3 * Q * L->T);
    if (L->T <= 256) {  // pack instances so that few lanes of the last wave idle
        uint32_t best = 1;
        double best_u = 0;
        for (uint32_t k = 1; k * L->T <= 256; k++) {
            const uint32_t th = (k * L->T + 63) / 64 * 64;
            const double u = (double)(k * L->T) / th;
            if (u >= best_u) {
                best_u = u;
                best = k;
            }
        }
        L->ipb = best;
        L->threads = (best * L->T + 63) / 64 * 64;
    } else {
        L->ipb = 1;
        L->threads = (L->T + 63) / 64 * 64;
    }
    if (L->threads > 1024) return false;
    // the normalisation covers 4 rows of 128 limbs per wave
    if ((uint64_t)(L->threads / 64) * 4 * 128 < (uint64_t)L->ipb * L->P) return false;
    return true;
}

static size_t digit_lds_bytes(const DigitLayout& L) {
    const size_t conv = (size_t)L.ipb * L.words * 4;
    const size_t norm = (size_t)L.ipb * L.P * 16;
    return conv > norm ? conv : norm;
}

template <bool FULL, bool PF, bool RPARK>
static hipError_t launch_digit_t(const DigitLayout& L, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b,
                                 uint64_t* r, const Strides& st, hipStream_t s) {
    const size_t lds = digit_lds_bytes(L);
    const void* fn = (const void*)mul_digit_kernel<FULL, PF, RPARK>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + L.ipb - 1) / L.ipb);
    mul_digit_kernel<FULL, PF, RPARK><<<blocks, L.threads, lds, s>>>(n, m, L.D, L.n_pad, L.T, L.words, L.ipb, L.P, a, b, r,
                                                             st.a ? st.a : m, st.b ? st.b : m, st.r ? st.r : L.P);
    return hipGetLastError();
}

static bool mul_layout(uint32_t m, bool full, uint32_t Q, bool rpark, MulLayout* L) {
    if (m == 0 || m > kMulMaxLimbs) return false;
    if (Q != 4 && Q != 8 && Q != 16) return false;
    L->Q = Q;
    L->rpark = rpark && Q >= 8;
    const uint32_t n = 2 * m;
    L->n_pad = (n + 2 * Q - 1) / (2 * Q) * (2 * Q);
    L->T = full ? L->n_pad / Q : L->n_pad / (2 * Q);
    L->P = full ? L->n_pad : L->n_pad / 2;
    const uint32_t wlen = Q == 16 ? mul_w_len<16>(L->n_pad) : Q == 8 ? mul_w_len<8>(L->n_pad) : mul_w_len<4>(L->n_pad);
    L->words = L->n_pad + wlen + (L->rpark ? 0 : 3 * Q * L->T);
    if (L->T <= 256) {
        // block size: tunable per size class ("mul.threads"); smaller blocks
        // pack the per-wave LDS footprint more tightly into 160 KB
        uint32_t th = (uint32_t)tune_get_class("mul.threads", m, 256);
        if (th < 64 || th > 256 || th % 64 != 0 || th < L->T) th = 256;
        L->threads = th;
        L->ipb = th / L->T;
    } else {
        L->threads = (L->T + 63) / 64 * 64;
        L->ipb = 1;
    }
    if (L->threads > (Q == 16 ? 256u : 1024u)) return false;  // the kernel's launch bound
    // normalisation: Q/2 rows of 128 limbs per wave must cover ipb*P limbs
    if ((uint64_t)(L->threads / 64) * (Q / 2) * 128 < (uint64_t)L->ipb * L->P) return false;
    return true;
}

static size_t mul_lds_bytes(const MulLayout& L) {
    size_t conv = (size_t)L.ipb * L.words * 4;
    size_t norm = (size_t)L.ipb * L.P * 16;
    return conv > norm ? conv : norm;
}

template <bool FULL, int Q, bool PF, bool RPARK>
static hipError_t launch_mul_t(const MulLayout& L, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b,
                               uint64_t* r, const Strides& st, hipStream_t s) {
    const size_t lds = mul_lds_bytes(L);
    if (lds > kMulLdsMax) return hipErrorInvalidValue;
    const void* fn = (const void*)mul_kernel<FULL, Q, PF, RPARK>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + L.ipb - 1) / L.ipb);
    const uint32_t out_len = FULL ? 2 * m : m;
    const int32_t m_shift = (m & (m - 1)) == 0 ? __builtin_ctz(m) : -1;
    mul_kernel<FULL, Q, PF, RPARK><<<blocks, L.threads, lds, s>>>(n, m, L.n_pad, L.T, L.P, L.words, L.ipb,
                                                                  out_len, m_shift, a, b, r, st.a ? st.a : m,
                                                                  st.b ? st.b : m, st.r ? st.r : out_len);
    return hipGetLastError();
}

template <bool FULL>
static hipError_t launch_mul_full_or_lo(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                                        const Strides& st, hipStream_t s) {
    if (tune_get_class("mul.mfma", m, 0) != 0) {
        if (tune_get_class("mul.wave16", m, 1) != 0 && mul_wave16_supported(m, FULL))
            return launch_mul_wave16(n, m, a, b, r, FULL, s, st);
        if (tune_get_class("mul.m16", m, 1) != 0 && mul_mfma16_supported(m, FULL))
            return launch_mul_mfma16(n, m, a, b, r, FULL, s, st);
        if (mul_mfma_supported(m, FULL)) return launch_mul_mfma(n, m, a, b, r, FULL, s, st);
    }
    if (tune_get_class("mul.digits", m, 0) != 0) {
        DigitLayout D;
        const bool rp = tune_get_class("mul.rpark", m, 1) != 0;
        if (digit_layout(m, FULL, rp, &D) && digit_lds_bytes(D) <= kMulLdsMax) {
            const bool pf = tune_get_class("mul.prefetch", m, 0) != 0;
            if (rp)
                return pf ? launch_digit_t<FULL, true, true>(D, n, m, a, b, r, st, s)
                          : launch_digit_t<FULL, false, true>(D, n, m, a, b, r, st, s);
            return pf ? launch_digit_t<FULL, true, false>(D, n, m, a, b, r, st, s)
                      : launch_digit_t<FULL, false, false>(D, n, m, a, b, r, st, s);
        }
    }
    const int64_t tq = tune_get_class("mul.q", m, 8);
    uint32_t Q = tq == 16 ? 16 : tq == 4 ? 4 : 8;
    const bool pf = tune_get_class("mul.prefetch", m, 0) != 0;
    const bool rp = tune_get_class("mul.rpark", m, 1) != 0;
    MulLayout L;
    if (!mul_layout(m, FULL, Q, rp, &L) || mul_lds_bytes(L) > kMulLdsMax) {
        Q = 4;  // fall back to the smallest footprint
        if (!mul_layout(m, FULL, Q, false, &L)) return hipErrorInvalidValue;
    }
    if (Q == 16) {
        if (L.rpark)
            return pf ? launch_mul_t<FULL, 16, true, true>(L, n, m, a, b, r, st, s)
                      : launch_mul_t<FULL, 16, false, true>(L, n, m, a, b, r, st, s);
        return pf ? launch_mul_t<FULL, 16, true, false>(L, n, m, a, b, r, st, s)
                  : launch_mul_t<FULL, 16, false, false>(L, n, m, a, b, r, st, s);
    }
    if (Q == 8) {
        if (L.rpark)
            return pf ? launch_mul_t<FULL, 8, true, true>(L, n, m, a, b, r, st, s)
                      : launch_mul_t<FULL, 8, false, true>(L, n, m, a, b, r, st, s);
        return pf ? launch_mul_t<FULL, 8, true, false>(L, n, m, a, b, r, st, s)
                  : launch_mul_t<FULL, 8, false, false>(L, n, m, a, b, r, st, s);
    }
    return pf ? launch_mul_t<FULL, 4, true, false>(L, n, m, a, b, r, st, s)
              : launch_mul_t<FULL, 4, false, false>(L, n, m, a, b, r, st, s);
}

hipError_t launch_mul(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                      hipStream_t s, Strides st) {
    return full ? launch_mul_full_or_lo<true>(n, m, a, b, r, st, s)
                : launch_mul_full_or_lo<false>(n, m, a, b, r, st, s);
}

// Whether launch_poly runs the fused kernel (and needs no scratch).
bool poly_uses_fused(uint32_t m) { return tune_get_class("poly.fused", m, 1) != 0 && poly_fused_supported(m); }

// Reference polyKerQ (cuda/ker-classic-mul.cu.h:533-559): the fused kernel
// where it fits (scratch unused), else a chain of launches over
// caller-provided scratch of 4*n*m limbs.
hipError_t launch_poly(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                       uint64_t* scratch, hipStream_t s, Strides st) {
    if (poly_uses_fused(m)) {
        if (tune_get_class("mul.wave16", m, 1) != 0 && mul_wave16_supported(m, false))
            return launch_poly_wave16(n, m, a, b, r, s, st);
        return launch_poly_fused(n, m, a, b, r, s, st);
    }
    const size_t nm = (size_t)n * m;
    uint64_t* t1 = scratch;
    uint64_t* t2 = scratch + nm;
    uint64_t* t3 = scratch + 2 * nm;
    uint64_t* t4 = scratch + 3 * nm;
    // strided operands / result only at the ends of the chain; temporaries are
    // batch-major. A zero stride keeps an addition on its 16-byte vector path.
    const uint64_t sa = st.a, sb = st.b;
    hipError_t e;
    if ((e = launch_mul(n, m, a, a, t1, false, s, {sa, sa, 0})) != hipSuccess) return e;         // t1 = a^2
    if ((e = launch_add(n, m, t1, b, t2, nullptr, false, s, {0, sb, 0})) != hipSuccess) return e; // t2 = a^2 + b
    if ((e = launch_mul(n, m, b, b, t3, false, s, {sb, sb, 0})) != hipSuccess) return e;         // t3 = b^2
    if ((e = launch_add(n, m, t3, b, t4, nullptr, false, s, {0, sb, 0})) != hipSuccess) return e; // t4 = b^2 + b
    if ((e = launch_mul(n, m, t2, t4, t1, false, s)) != hipSuccess) return e;                    // t5 = t2 * t4
    if ((e = launch_mul(n, m, a, b, t3, false, s, {sa, sb, 0})) != hipSuccess) return e;         // t6 = a * b
    return launch_add(n, m, t1, t3, r, nullptr, false, s, {0, 0, st.r});                         // r = t5 + t6
}

}  // namespace midint
