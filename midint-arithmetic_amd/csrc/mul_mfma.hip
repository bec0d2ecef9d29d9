// Classical (schoolbook) batched multiplication on the int8 matrix cores.
//
// Operands are split into 7-bit digits, non-negative in the signed int8
// operands of v_mfma_i32_32x32x32_i8. With blocks of 32 digits
// (a = sum_J a_J X^J, X = 2^224, D digits, P = ceil(D / 32) blocks), output
// block I (digits 32I .. 32I+31) of the column sums is
//
//     c_I = sum_{d=0..P} T_d a_{I-d},     T_d[r][s] = b_{32d + r - s},
//
// a 32 x 32 Toeplitz matrix of b times a 32-digit block of a. One MFMA
// evaluates T_d [a_{I0-d} .. a_{I0+31-d}] for the 32 output blocks of a
// "column tile" I0 .. I0+31 and accumulates over d in its int32 C registers,
// so no partial sum leaves the matrix core. A column sum is below
// D * 127^2 < 2^29.2 for m <= 4096 limbs (D <= 37450), inside int32, and a
// sum times a power of two below 2^32 stays below 2^61.2.
//
// LDS: a's digits as 32-byte blocks of 32 digits, the instances of a block
// one after another with a zero block before, between and after them (B
// operand: one aligned 16-byte read per lane of block med3(J, -1, P), so a
// block index outside the instance reads zeros and its MFMA adds nothing;
// the loop then needs no per-tile range tests), and per instance b's
// digits reversed (A operand: row r of T_d is
// br[y0 .. y0+15], y0 = X0 - 32d - r + 16kg; the misalignment of y0 is fixed
// per lane, so each row is five aligned dword reads and four funnel shifts).
// Since 32 digits x 7 bits = 224 bits = 7 u32 words, block J of an operand
// is words 7J .. 7J+6 and its digits sit at fixed shifts; br is b's blocks
// in reverse order, each byte-reversed. The product is written to two limb
// vectors E and O shared by the block's instances.
//
// Column sums to limbs: lane (n, kg) of a tile holds the 16 sums of output
// block I = I0 + n at rows 8q + 4kg + t (q, t < 4). It folds them into
// 32-bit-word accumulators (one v_mad_u64_u32 by a power of two per sum),
// carries those into 32-bit words, shifts by 28 bits when kg = 1 and adds
// its partner lane (lane ^ 32): the block's value V < 2^252 at bit 224 I.
// Even blocks start on a limb (3.5 I) and go to E, odd blocks start 32 bits
// into limb 3.5 I - 0.5 and go to O; every block writes its whole 7-limb
// slot, the slots of each vector tile it, and the product is E + O: one
// carry-resolving pass of the row adder (the same split as the 28-bit digit
// kernel in mul.hip).
#include <algorithm>

#include "add_core.hpp"
#include "midint_internal.h"

namespace midint {

typedef int mf_v4i __attribute__((ext_vector_type(4)));
typedef int mf_v16i __attribute__((ext_vector_type(16)));

constexpr int kMfTiles = 2;      // column tiles (accumulator sets) per wave pass
constexpr int kMfRawUnroll = 4;  // operand limbs per thread per staging round

// Bytes of the a region of ipb instances (a zero block, then per instance
// P digit blocks and a zero block) and of one instance's br region (P + 3
// blocks, the last three zero).
__host__ __device__ constexpr uint32_t mf_a_bytes(uint32_t P, uint32_t ipb) { return (1 + ipb * (P + 1)) * 32; }
__host__ __device__ constexpr uint32_t mf_br_bytes(uint32_t P) { return (P + 3) * 32; }

// floor(t / d) from mg = ceil(2^32 / d) (mf_magic, on the host) whenever
// t * d < 2^32: t mg / 2^32 = t / d + t e / (d 2^32) with e = mg d - 2^32 < d,
// and t e < 2^32 keeps the excess below 1 / d, which cannot carry past the
// next multiple of 1 / d. mfma_layout checks the bound for every division.
__device__ __forceinline__ uint32_t mf_div(uint32_t t, uint32_t mg) { return __umulhi(t, mg); }
__host__ constexpr uint32_t mf_magic(uint32_t d) { return (uint32_t)((0xFFFFFFFFull + d) / d); }

// min(max(x, lo), hi) for lo <= hi in one instruction.
__device__ __forceinline__ uint32_t mf_med3(uint32_t x, uint32_t lo, uint32_t hi) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}

__device__ __forceinline__ uint32_t mf_ld(const uint32_t* __restrict__ g, uint32_t w, uint32_t W) {
    return w < W ? g[w] : 0u;
}

// Digits x0 .. x0+3 of an operand of W u32 words, one per byte (byte e =
// digit x0 + e); digits past the operand read as zero.
__device__ __forceinline__ uint32_t mf_digits4(const uint32_t* __restrict__ g, uint32_t W, uint32_t x0) {
    const uint32_t bit = 7 * x0;
    const uint32_t w = bit >> 5;
    const uint32_t v = __builtin_amdgcn_alignbit(mf_ld(g, w + 1, W), mf_ld(g, w, W), bit & 31);
    return (v & 0x7Fu) | ((v << 1) & 0x7F00u) | ((v << 2) & 0x7F0000u) | ((v << 3) & 0x7F000000u);
}

// Folds a finished tile into the E / O slots of its 32 output blocks.
__device__ __forceinline__ void mf_store_tile(const mf_v16i& c, uint32_t I, uint32_t kg, uint64_t* E, uint64_t* O,
                                              uint32_t PL) {
    // row 8q + t (t < 4) of the lane's half sits at bit 7 (8q + t) (+ 28 kg)
    uint64_t A[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int bit = 7 * (8 * (v / 4) + (v % 4));
        A[bit >> 5] += (uint64_t)(uint32_t)c[v] << (bit & 31);
    }
    uint32_t u[8];
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < 6; w++) {
        const uint64_t t = A[w] + carry;
        u[w] = (uint32_t)t;
        carry = t >> 32;
    }
    u[6] = (uint32_t)carry;
    u[7] = 0;
    uint32_t V[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const uint32_t sh = __builtin_amdgcn_alignbit(u[w], w ? u[w - 1] : 0u, 4);  // (U << 28) word w
        V[w] = kg ? sh : u[w];
    }
    uint32_t S[9];
    uint32_t cy = 0;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const uint32_t p = (uint32_t)__shfl_xor((int)V[w], 32);
        const uint64_t t = (uint64_t)V[w] + p + cy;
        S[w] = (uint32_t)t;
        cy = (uint32_t)(t >> 32);
    }
    S[8] = cy;
    uint64_t L[7];
    uint64_t* dst;
    uint32_t base;
    if ((I & 1u) == 0) {
        base = 7 * (I >> 1);
        dst = E;
#pragma unroll
        for (int i = 0; i < 4; i++) L[i] = (uint64_t)S[2 * i] | ((uint64_t)S[2 * i + 1] << 32);
        L[4] = S[8];
    } else {
        base = 7 * (I >> 1) + 3;
        dst = O;
        L[0] = (uint64_t)S[0] << 32;
#pragma unroll
        for (int i = 1; i < 5; i++) L[i] = (uint64_t)S[2 * i - 1] | ((uint64_t)S[2 * i] << 32);
    }
    L[5] = 0;
    L[6] = 0;
    // the lane pair splits the slot: kg = 0 limbs 0..3, kg = 1 limbs 4..6
#pragma unroll
    for (int i = 0; i < 7; i++)
        if ((i < 4) == (kg == 0) && base + i < PL) dst[base + i] = L[i];
    if (I == 0 && kg == 1) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            if ((uint32_t)i < PL) O[i] = 0;
    }
}

template <bool FULL, int R>
__global__ __launch_bounds__(1024) void mul_mfma_kernel(uint32_t n_inst, uint32_t m, uint32_t P, uint32_t NT,
                                                        uint32_t wpi, uint32_t ipb, uint32_t PL, uint32_t mg_2m,
                                                        uint32_t mg_p1, uint32_t mg_nbr,
                                                        const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                        uint64_t sa, uint64_t sb, uint64_t sr) {
    extern __shared__ __align__(16) uint64_t lds64[];
    uint8_t* lds8 = reinterpret_cast<uint8_t*>(lds64);
    __shared__ uint32_t lds_gp[16];

    const uint32_t tid = threadIdx.x;
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);
    const uint32_t W32 = 2 * m;
    const uint32_t X0 = 32 * P + 31;
    const uint32_t A_BYTES = mf_a_bytes(P, ipb), BR_BYTES = mf_br_bytes(P);
    uint64_t* Eb = lds64;
    uint64_t* Ob = lds64 + (size_t)ipb * PL;
    uint8_t* Areg = lds8 + (size_t)16 * ipb * PL;  // a blocks of all instances
    uint8_t* BRreg = Areg + A_BYTES;                // br blocks, P + 3 per instance

    // ---- stage digits, in two steps: (1) the operands' limbs, coalesced,
    // into the E / O area (free until the MFMA phase: 2m limbs per instance
    // fit in its 2 PL); (2) per (instance, 32-digit block) item, the block's
    // 7 words from there to packed digits: a (zero-padded blocks) and b
    // reversed.
    {
        // kMfRawUnroll limbs per thread per round, all loads issued before
        // the stores
        uint64_t* raw = lds64;
        const uint32_t nl = ipb * 2 * m;
        for (uint32_t t0 = tid; t0 < nl; t0 += kMfRawUnroll * blockDim.x) {
            uint64_t v[kMfRawUnroll];
#pragma unroll
            for (int u = 0; u < kMfRawUnroll; u++) {
                const uint32_t t = t0 + u * blockDim.x;
                const uint32_t si = mf_div(t, mg_2m), j = t - si * 2 * m;
                v[u] = 0;
                if (t < nl && si < ninst) v[u] = j < m ? a[(inst_base + si) * sa + j] : b[(inst_base + si) * sb + (j - m)];
            }
#pragma unroll
            for (int u = 0; u < kMfRawUnroll; u++) {
                const uint32_t t = t0 + u * blockDim.x;
                if (t < nl) raw[t] = v[u];
            }
        }
    }
    __syncthreads();
    {
        const uint32_t NA = A_BYTES / 32, NBR = P + 3, total = NA + ipb * NBR;
        for (uint32_t t = tid; t < total; t += blockDim.x) {
            int32_t J;
            uint32_t si;
            uint8_t* dst8;
            const bool br = t >= NA;
            const uint32_t* g;
            if (!br) {  // a-block t: instance (t-1) / (P+1), block (t-1) % (P+1)
                const int32_t q = (int32_t)t - 1;
                si = q < 0 ? 0u : mf_div((uint32_t)q, mg_p1);
                J = q < 0 ? -1 : q - (int32_t)(si * (P + 1));
                g = reinterpret_cast<const uint32_t*>(lds64 + (size_t)si * 2 * m);
                dst8 = Areg + t * 32;
            } else {  // br block Y of instance si holds b-block P - Y, byte-reversed
                const uint32_t u = t - NA;
                si = mf_div(u, mg_nbr);
                const uint32_t Y = u - si * NBR;
                J = (int32_t)P - (int32_t)Y;
                g = reinterpret_cast<const uint32_t*>(lds64 + (size_t)si * 2 * m) + W32;
                dst8 = BRreg + (si * NBR + Y) * 32;
            }
            if (si >= ipb) J = -1;
            uint32_t wd[8];
            const bool have = J >= 0 && J < (int32_t)P;
#pragma unroll
            for (int i = 0; i < 7; i++) wd[i] = have ? mf_ld(g, 7 * (uint32_t)J + i, W32) : 0u;
            wd[7] = 0;
            uint32_t o[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int bit = 28 * k;
                const uint32_t v = __builtin_amdgcn_alignbit(wd[(bit >> 5) + 1], wd[bit >> 5], bit & 31);
                o[k] = (v & 0x7Fu) | ((v << 1) & 0x7F00u) | ((v << 2) & 0x7F0000u) | ((v << 3) & 0x7F000000u);
            }
            uint4* dst = reinterpret_cast<uint4*>(dst8);
            if (br) {
                dst[0] = make_uint4(__builtin_bswap32(o[7]), __builtin_bswap32(o[6]), __builtin_bswap32(o[5]),
                                    __builtin_bswap32(o[4]));
                dst[1] = make_uint4(__builtin_bswap32(o[3]), __builtin_bswap32(o[2]), __builtin_bswap32(o[1]),
                                    __builtin_bswap32(o[0]));
            } else {
                dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
                dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
            }
        }
    }
    __syncthreads();

    // ---- MFMA phase: wave wi of instance s takes tiles wi, wi + wpi, ...
    // in passes of up to kMfTiles tiles over the union of their d ranges.
    // Wave and d indices are scalar; each d step loads the A operand (five
    // dwords of br) and every tile's B operand (block clamped into
    // [-1, P]: zeros outside the instance), and the loads of step d-1 are
    // issued before the MFMAs of step d.
    {
        const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        const uint32_t lane = tid & 63;
        const uint32_t s = wv / wpi, wi = wv - s * wpi;
        const uint32_t n = lane & 31, kg = lane >> 5;
        const uint8_t* Ab = Areg + (size_t)(1 + s * (P + 1)) * 32;  // block 0 of instance s
        // B-operand byte addresses (from the start of LDS) of block 0 and of
        // the zero blocks at -1 and P for this lane's K half
        const uint32_t b_base = (uint32_t)(Ab - lds8) + 16 * kg;
        const uint32_t b_lo = b_base - 32, b_hi = b_base + 32 * P;
        const uint32_t* BR = reinterpret_cast<const uint32_t*>(BRreg + (size_t)s * BR_BYTES);
        const uint32_t ybase = X0 - n + 16 * kg;  // y0 at d = 0
        const uint32_t sh = (ybase & 3) * 8;
        const uint32_t* Abase = BR + (ybase >> 2);  // dword y0 >> 2 at d = 0
        uint64_t* E = Eb + (size_t)s * PL;
        uint64_t* O = Ob + (size_t)s * PL;
        for (uint32_t c0 = wi; c0 < NT; c0 += kMfTiles * wpi) {
            // d range of the pass: union of its tiles' ranges
            int32_t dlo = (int32_t)P, dhi = 0;
            uint32_t bb[kMfTiles];  // LDS byte address of the lane's B operand at d = 0
#pragma unroll
            for (int k = 0; k < kMfTiles; k++) {
                const uint32_t ct = c0 + k * wpi;
                const int32_t I0 = 32 * (int32_t)ct;
                if (ct < NT) {
                    dlo = min(dlo, max(0, I0 - (int32_t)P + 1));
                    dhi = max(dhi, min((int32_t)P, I0 + 31));
                }
                bb[k] = b_base + (uint32_t)(32 * (I0 + (int32_t)n));
            }
            mf_v16i acc[kMfTiles];
#pragma unroll
            for (int k = 0; k < kMfTiles; k++) acc[k] = mf_v16i{};
            // two operand sets in ping-pong: the loads of step d-1 go to one
            // set while the MFMAs of step d read the other (no copies)
            uint32_t qa0[5], qa1[5];
            mf_v4i bq0[kMfTiles], bq1[kMfTiles];
            auto issue = [&](int32_t d, uint32_t (&qa)[5], mf_v4i (&bq)[kMfTiles]) {
                const uint32_t* pa = Abase - 8 * d;
#pragma unroll
                for (int i = 0; i < 5; i++) qa[i] = pa[i];
#pragma unroll
                for (int k = 0; k < kMfTiles; k++) {
                    // block J = J0 - d, clamped to the zero blocks at -1 and P
                    const uint32_t addr = mf_med3(bb[k] - 32u * (uint32_t)d, b_lo, b_hi);
                    bq[k] = *reinterpret_cast<const mf_v4i*>(lds8 + addr);
                }
            };
            auto mma = [&](const uint32_t (&qc)[5], const mf_v4i (&bc)[kMfTiles]) {
                const mf_v4i av = {(int)__builtin_amdgcn_alignbit(qc[1], qc[0], sh),
                                   (int)__builtin_amdgcn_alignbit(qc[2], qc[1], sh),
                                   (int)__builtin_amdgcn_alignbit(qc[3], qc[2], sh),
                                   (int)__builtin_amdgcn_alignbit(qc[4], qc[3], sh)};
#pragma unroll
                for (int k = 0; k < kMfTiles; k++) acc[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bc[k], acc[k], 0, 0, 0);
            };
            // steps dhi, dhi-1, ..., dlo: pairs, then a last single step
            const int32_t steps = dhi - dlo + 1;
            issue(dhi, qa0, bq0);
            int32_t d = dhi;
            for (int32_t i = 0; i + 1 < steps; i += 2, d -= 2) {
                issue(d - 1, qa1, bq1);
                mma(qa0, bq0);
                issue(max(d - 2, dlo), qa0, bq0);
                mma(qa1, bq1);
            }
            if (steps & 1) mma(qa0, bq0);
#pragma unroll
            for (int k = 0; k < kMfTiles; k++) {
                const uint32_t ct = c0 + k * wpi;
                if (ct < NT) mf_store_tile(acc[k], 32 * ct + n, kg, E, O, PL);
            }
        }
    }
    __syncthreads();

    // ---- normalisation: E + O over ipb segments of PL limbs (R rows per wave)
    const uint32_t Lw = ipb * PL;
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
    Rows<R> rows;
    uint32_t off0[R], instv[R];
    {
        SegPos sp = segpos_init(k_first, PL);
        const uint32_t sq = kRowLimbs / PL, srr = kRowLimbs % PL;
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            rows.x0[v] = k < Lw ? Eb[k] : 0ull;
            rows.x1[v] = k + 1 < Lw ? Eb[k + 1] : 0ull;
            rows.y0[v] = k < Lw ? Ob[k] : 0ull;
            rows.y1[v] = k + 1 < Lw ? Ob[k + 1] : 0ull;
            rows.s0[v] = (sp.off == 0) && (k < Lw);
            rows.s1[v] = (sp.off + 1 == PL) && (k + 1 < Lw);
            off0[v] = sp.off;
            instv[v] = sp.inst;
            segpos_next(&sp, PL, sq, srr);
        }
    }
    block_add_rows<R>(rows, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const uint32_t o0 = off0[v];
        const uint32_t i0 = instv[v];
        if (k < Lw && i0 < ninst) r[(inst_base + i0) * sr + o0] = rows.x0[v];
        const bool wrap = (o0 + 1 == PL);
        const uint32_t o1 = wrap ? 0u : o0 + 1;
        const uint32_t i1 = i0 + (wrap ? 1u : 0u);
        if (k + 1 < Lw && i1 < ninst) r[(inst_base + i1) * sr + o1] = rows.x1[v];
    }
}

namespace {

struct MfmaLayout {
    uint32_t D, P, NT, wpi, ipb, waves, R, PL;
    size_t lds;
};

bool mfma_layout(uint32_t m, bool full, MfmaLayout* L) {
    if (m == 0 || m > kMulMaxLimbs) return false;
    L->D = (64 * m + 6) / 7;
    L->P = (L->D + 31) / 32;
    L->PL = full ? 2 * m : m;
    const uint32_t blocks_needed = (64 * L->PL + 223) / 224;  // output blocks below limb PL
    L->NT = (blocks_needed + 31) / 32;
    int64_t w = tune_get_class("mul.mfma_wpi", m, 0);
    if (w <= 0) w = (L->NT + 2) / 4;  // measured best: about a quarter of the column tiles
    L->wpi = (uint32_t)std::max<int64_t>(1, std::min<int64_t>({w, 16, (int64_t)L->NT}));
    L->ipb = std::max(1u, 4 / L->wpi);
    L->waves = L->ipb * L->wpi;
    const uint64_t limbs = (uint64_t)L->ipb * L->PL;
    uint32_t R = 1;
    while ((uint64_t)L->waves * R * 128 < limbs) R *= 2;
    if (R > 8) return false;
    L->R = R;
    L->lds = (size_t)16 * L->ipb * L->PL + mf_a_bytes(L->P, L->ipb) + (size_t)L->ipb * mf_br_bytes(L->P);
    // mf_div bounds: limb index x 2m, staging item x (P + 1) and x (P + 3)
    const uint64_t limbs2 = (uint64_t)L->ipb * 2 * m;
    const uint64_t items = 1 + (uint64_t)L->ipb * (2 * L->P + 4);
    if (limbs2 * 2 * m >= (1ull << 32) || items * (L->P + 3) >= (1ull << 32)) return false;
    return L->lds <= 160 * 1024 - 256;
}

template <bool FULL, int R>
hipError_t launch_mfma_t(const MfmaLayout& L, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b,
                         uint64_t* r, const Strides& st, hipStream_t s) {
    const void* fn = (const void*)mul_mfma_kernel<FULL, R>;
    if (L.lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + L.ipb - 1) / L.ipb);
    mul_mfma_kernel<FULL, R><<<blocks, 64 * L.waves, L.lds, s>>>(n, m, L.P, L.NT, L.wpi, L.ipb, L.PL,
                                                                 mf_magic(2 * m), mf_magic(L.P + 1),
                                                                 mf_magic(L.P + 3), a, b, r,
                                                                 st.a ? st.a : m, st.b ? st.b : m,
                                                                 st.r ? st.r : L.PL);
    return hipGetLastError();
}

template <bool FULL>
hipError_t launch_mfma_full_or_lo(const MfmaLayout& L, uint32_t n, uint32_t m, const uint64_t* a,
                                  const uint64_t* b, uint64_t* r, const Strides& st, hipStream_t s) {
    switch (L.R) {
        case 1: return launch_mfma_t<FULL, 1>(L, n, m, a, b, r, st, s);
        case 2: return launch_mfma_t<FULL, 2>(L, n, m, a, b, r, st, s);
        case 4: return launch_mfma_t<FULL, 4>(L, n, m, a, b, r, st, s);
        default: return launch_mfma_t<FULL, 8>(L, n, m, a, b, r, st, s);
    }
}

}  // namespace

bool mul_mfma_supported(uint32_t m, bool full) {
    MfmaLayout L;
    return mfma_layout(m, full, &L);
}

hipError_t launch_mul_mfma(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                           hipStream_t s, Strides st) {
    MfmaLayout L;
    if (!mfma_layout(m, full, &L)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    return full ? launch_mfma_full_or_lo<true>(L, n, m, a, b, r, st, s)
                : launch_mfma_full_or_lo<false>(L, n, m, a, b, r, st, s);
}

}  // namespace midint
