// Classical (schoolbook) batched multiplication on the int8 matrix cores.
//
// Digits. An operand of m limbs is written in balanced base 256: with
// x = a + 0x8080...80 (128 added to each of its 8m bytes, carries
// propagated), digit i (i < 8m) is byte i of x minus 128, in [-128, 127],
// stored as the int8 byte (byte i of x) ^ 0x80, and digit 8m is the carry out
// of x (0 or 1). Then a = sum_i digit_i 256^i exactly, with D = 8m + 1
// digits.
//
// Blocks of 32 digits are 256 bits = 4 limbs (a = sum_J a_J X^J, X = 2^256,
// P = floor(m / 4) + 1 blocks). Output block I (digits 32I .. 32I+31) of the
// column sums is
//
//     c_I = sum_d T_d a_{I-d},     T_d[r][s] = b_{32d + r - s},
//
// a 32 x 32 Toeplitz matrix of b times a 32-digit block of a; T_d is zero
// for d > dmax = floor((D + 30) / 32). One v_mfma_i32_32x32x32_i8 evaluates
// T_d [a_{I0-d} .. a_{I0+31-d}] for the 32 output blocks of a "column tile"
// I0 .. I0+31 and accumulates over d in its int32 C registers, so no partial
// sum leaves the matrix core.
//
// Bias. A column sum is a sum of at most D digit products in
// [-16256, 16384]. The accumulators start at zero and hold signed sums; when
// a tile is folded into limbs, 255 B (B = 2^21) is added to every column
// sum, which makes it non-negative and below 2^30 for m <= 4096 limbs. That
// adds B (2^256 - 1) to every block value; over the N output blocks the
// total is B 2^(256 N) - B, and 2^(256 N) is zero modulo the output length
// (4N >= output limbs), so the lowest column of block 0 gets 256 B instead
// and the blocks sum to the product exactly.
//
// LDS: a's digits as blocks, the instances of a workgroup one after another
// with a zero block before, between and after them, each block's two 16-byte
// K halves in two separate arrays (B operand: one aligned 16-byte read per
// lane of block med3(J, -1, P), so a block index outside the instance reads
// zeros and its MFMA adds nothing and the loop needs no per-tile range
// tests; with the halves apart, 16 lanes read 256 contiguous bytes, free of
// bank conflicts), and per instance b's digits reversed, br[y] =
// b_digit[X0 - y] with X0 = 32P + 31 (A operand: row r of T_d is
// br[y0 .. y0+15], y0 = X0 - 32d - r + 16kg; the misalignment of y0 is fixed
// per lane, so each row is five aligned dword reads and four funnel shifts).
// Limb j of an operand is digits 8j .. 8j+7: one 8-byte word of the a region
// (half (j / 2) mod 2 of block j / 4) and one byte-swapped 8-byte word of br,
// so staging needs no bit shuffling.
//
// Column sums to limbs: lane (n, kg) of a tile holds the 16 sums of output
// block I = I0 + n at rows 8q + 4kg + t (q, t < 4), i.e. at bits
// 64q + 32kg + 8t of the block. G_q = sum_t (c_{8q+4kg+t} + 255 B) 2^(8t)
// < 2^55 (four v_mad_i64_i32 onto the bias of the four columns), so the
// lane's part is sum_q G_q 2^(64q) shifted up by 32 kg bits, with no
// carries inside the lane. The kg = 0 lane takes its partner's G_q
// (ds_bpermute) and adds them 32 bits up in one 9-word carry
// chain: the block value V < 2^279 (5 limbs). Even blocks write vector E,
// odd blocks vector O, each an 8-limb slot at limb 4I (the kg = 1 lane
// writes its 3 zero limbs); the slots of each vector tile it and the product
// is E + O, one carry-resolving pass of the row adder.
#include <algorithm>

#include "add_core.hpp"
#include "mfma_digits.hpp"
#include "midint_internal.h"

namespace midint {

typedef int mf_v16i __attribute__((ext_vector_type(16)));

constexpr int kMfTiles = 2;                          // column tiles (accumulator sets) per wave pass

// Bytes of the a region of ipb instances (a zero block, then per instance
// P digit blocks and a zero block) and of one instance's br region (P + 3
// blocks).
__host__ __device__ constexpr uint32_t mf_a_bytes(uint32_t P, uint32_t ipb) { return (1 + ipb * (P + 1)) * 32; }
__host__ __device__ constexpr uint32_t mf_br_bytes(uint32_t P) { return (P + 3) * 32; }

// floor(t / d) from mg = ceil(2^32 / d) (mf_magic, on the host) whenever
// t * d < 2^32: t mg / 2^32 = t / d + t e / (d 2^32) with e = mg d - 2^32 < d,
// and t e < 2^32 keeps the excess below 1 / d, which cannot carry past the
// next multiple of 1 / d. mfma_layout and mfma_rg_config check the bound.
// d = 1 (2^32 does not fit) is encoded as mg = 0.
__device__ __forceinline__ uint32_t mf_div(uint32_t t, uint32_t mg) { return mg ? __umulhi(t, mg) : t; }
__host__ constexpr uint32_t mf_magic(uint32_t d) { return d == 1 ? 0u : (uint32_t)((0xFFFFFFFFull + d) / d); }

__device__ __forceinline__ SegPos mf_segpos(uint32_t k, uint32_t seg, uint32_t mg) {
    SegPos s;
    s.inst = mf_div(k, mg);
    s.off = k - s.inst * seg;
    return s;
}

// min(max(x, lo), hi) for lo <= hi in one instruction.
__device__ __forceinline__ uint32_t mf_med3(uint32_t x, uint32_t lo, uint32_t hi) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}

// A finished tile's 16 signed column sums of the lane (the accumulators
// start at zero) as the words of sum_q G_q 2^(64q), G_q = sum_t (c_{4q+t} +
// 255 B) 2^(8t) < 2^55; bias0 is G_0's bias (one B more in the lowest
// column of block 0).
__device__ __forceinline__ void mf_fold(const mf_v16i& c, const MfPow& pw, uint64_t bias0, uint32_t (&g)[8]) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int64_t G = (int64_t)(q == 0 ? bias0 : kMfFoldBias);
        G += (int64_t)c[4 * q] * (int32_t)pw.p0;
        G += (int64_t)c[4 * q + 1] * (int32_t)pw.p8;
        G += (int64_t)c[4 * q + 2] * (int32_t)pw.p16;
        G += (int64_t)c[4 * q + 3] * (int32_t)pw.p24;
        g[2 * q] = (uint32_t)G;
        g[2 * q + 1] = (uint32_t)((uint64_t)G >> 32);
    }
}

// Writes block I's E / O slot (only blocks below N) from the lane's words g
// and its partner's words p (the kg = 1 rows, 32 bits up).
__device__ __forceinline__ void mf_store_block(const uint32_t (&g)[8], const uint32_t (&p)[8], uint32_t I,
                                               uint32_t kg, uint32_t N, uint64_t* E, uint64_t* O) {
    uint32_t w1, w2, w3, w4, w5, w6, w7, w8;
    uint64_t cy;
    asm("v_add_co_u32_e64 %[w1], %[c], %[g1], %[p0]\n\t"
        "v_addc_co_u32_e64 %[w2], %[c], %[g2], %[p1], %[c]\n\t"
        "v_addc_co_u32_e64 %[w3], %[c], %[g3], %[p2], %[c]\n\t"
        "v_addc_co_u32_e64 %[w4], %[c], %[g4], %[p3], %[c]\n\t"
        "v_addc_co_u32_e64 %[w5], %[c], %[g5], %[p4], %[c]\n\t"
        "v_addc_co_u32_e64 %[w6], %[c], %[g6], %[p5], %[c]\n\t"
        "v_addc_co_u32_e64 %[w7], %[c], %[g7], %[p6], %[c]\n\t"
        "v_addc_co_u32_e64 %[w8], %[c], %[p7], 0, %[c]"
        : [w1] "=&v"(w1), [w2] "=&v"(w2), [w3] "=&v"(w3), [w4] "=&v"(w4), [w5] "=&v"(w5), [w6] "=&v"(w6),
          [w7] "=&v"(w7), [w8] "=&v"(w8), [c] "=&s"(cy)
        : [g1] "v"(g[1]), [g2] "v"(g[2]), [g3] "v"(g[3]), [g4] "v"(g[4]), [g5] "v"(g[5]), [g6] "v"(g[6]),
          [g7] "v"(g[7]), [p0] "v"(p[0]), [p1] "v"(p[1]), [p2] "v"(p[2]), [p3] "v"(p[3]), [p4] "v"(p[4]),
          [p5] "v"(p[5]), [p6] "v"(p[6]), [p7] "v"(p[7]));
    if (I < N) {
        // slots start on 32-byte boundaries (E, O and the instance stride
        // are multiples of 4 limbs): 16-byte stores
        uint64_t* dst = ((I & 1u) ? O : E) + 4 * I;
        if (kg == 0) {
            *reinterpret_cast<uint4*>(dst) = make_uint4(g[0], w1, w2, w3);
            *reinterpret_cast<uint4*>(dst + 2) = make_uint4(w4, w5, w6, w7);
            dst[4] = w8;
        } else {
            // a zero made here: a hoisted constant would hold registers
            // through the whole MFMA phase
            uint32_t z;
            asm volatile("v_mov_b32 %0, 0" : "=v"(z));
            const uint4 z4 = make_uint4(z, z, z, z);
            dst[5] = ((uint64_t)z << 32) | z;
            *reinterpret_cast<uint4*>(dst + 6) = z4;
            if (I == 0) {
                *reinterpret_cast<uint4*>(O) = z4;
                *reinterpret_cast<uint4*>(O + 2) = z4;
            }
        }
    }
}

// Per-wave operand addressing of one instance in the MFMA phase.
struct MfLane {
    const uint8_t* lds8;
    const uint32_t* Abase;  // dword y0 >> 2 of the lane's A row at d = 0
    uint32_t sh;            // bit misalignment of that row
    uint32_t b_base;        // LDS byte address of the lane's B operand, block 0
    uint32_t b_lo, b_hi;    // the zero blocks at -1 and P
    uint32_t n, kg, lane, paddr;
    uint64_t *E, *O;
};

// One pass of NTL tiles c0, c0 + wpi, ... over the union of their d ranges.
template <int NTL>
__device__ __forceinline__ void mf_pass(const MfLane& L, uint32_t c0, uint32_t wpi, uint32_t P, uint32_t dmax,
                                        uint32_t N, const MfPow& pw) {
    int32_t dlo = (int32_t)dmax, dhi = 0;
    uint32_t bb[NTL];  // LDS byte address of the lane's B operand at d = 0 (16 bytes per block)
#pragma unroll
    for (int k = 0; k < NTL; k++) {
        const int32_t I0 = 32 * (int32_t)(c0 + k * wpi);
        dlo = min(dlo, max(0, I0 - (int32_t)P + 1));
        dhi = max(dhi, min((int32_t)dmax, I0 + 31));
        bb[k] = L.b_base + (uint32_t)(16 * (I0 + (int32_t)L.n));
    }
    dlo = __builtin_amdgcn_readfirstlane(dlo);  // wave-uniform: keep d scalar
    dhi = __builtin_amdgcn_readfirstlane(dhi);
    mf_v16i acc[NTL];

    // two operand sets in ping-pong: the loads of step d-1 go to one set
    // while the MFMAs of step d read the other (no copies)
    uint32_t qa0[5], qa1[5];
    mf_v4i bq0[NTL], bq1[NTL];
    auto issue = [&](int32_t d, uint32_t (&qa)[5], mf_v4i (&bq)[NTL]) {
        const uint32_t* pa = L.Abase - 8 * d;
#pragma unroll
        for (int i = 0; i < 5; i++) qa[i] = pa[i];
#pragma unroll
        for (int k = 0; k < NTL; k++) {
            // block J = J0 - d, clamped to the zero blocks at -1 and P
            const uint32_t addr = mf_med3(bb[k] - 16u * (uint32_t)d, L.b_lo, L.b_hi);
            bq[k] = *reinterpret_cast<const mf_v4i*>(L.lds8 + addr);
        }
    };
    // the row of T_d: the lane's 16 digits from five aligned dwords
    auto a_op = [&](const uint32_t (&qc)[5]) {
        return mf_v4i{(int)__builtin_amdgcn_alignbit(qc[1], qc[0], L.sh),
                      (int)__builtin_amdgcn_alignbit(qc[2], qc[1], L.sh),
                      (int)__builtin_amdgcn_alignbit(qc[3], qc[2], L.sh),
                      (int)__builtin_amdgcn_alignbit(qc[4], qc[3], L.sh)};
    };
    auto mma = [&](const uint32_t (&qc)[5], const mf_v4i (&bc)[NTL]) {
        const mf_v4i av = a_op(qc);
#pragma unroll
        for (int k = 0; k < NTL; k++) acc[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bc[k], acc[k], 0, 0, 0);
    };
    // steps dhi, dhi-1, ..., dlo: the first from zero accumulators, then
    // pairs, then a last single step
    const int32_t steps = dhi - dlo + 1;
    if (steps > 0) {
        issue(dhi, qa0, bq0);
        issue(max(dhi - 1, dlo), qa1, bq1);
        {
            const mf_v4i av = a_op(qa0);
#pragma unroll
            for (int k = 0; k < NTL; k++) acc[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bq0[k], mf_v16i{}, 0, 0, 0);
        }
        // set 1 holds step d; rest steps remain, d down to dlo
        int32_t d = dhi - 1;
        const int32_t rest = steps - 1;
        for (int32_t i = 0; i + 1 < rest; i += 2, d -= 2) {
            issue(d - 1, qa0, bq0);
            mma(qa1, bq1);
            issue(max(d - 2, dlo), qa1, bq1);
            mma(qa0, bq0);
        }
        if (rest & 1) mma(qa1, bq1);
    } else {
#pragma unroll
        for (int k = 0; k < NTL; k++) acc[k] = mf_v16i{};
    }
    // column sums to E / O: both tiles' partner exchanges in flight together
    uint32_t gw[NTL][8], pw8[NTL][8];
    // lowest column of block 0: 256 B (the bias correction)
    const uint64_t bias0 = kMfFoldBias + (c0 == 0 && L.lane == 0 ? (uint64_t)kMfBias : 0ull);
#pragma unroll
    for (int k = 0; k < NTL; k++) mf_fold(acc[k], pw, k == 0 ? bias0 : kMfFoldBias, gw[k]);
#pragma unroll
    for (int k = 0; k < NTL; k++) {
#pragma unroll
        for (int i = 0; i < 8; i++) pw8[k][i] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)L.paddr, (int)gw[k][i]);
    }
#pragma unroll
    for (int k = 0; k < NTL; k++) mf_store_block(gw[k], pw8[k], 32 * (c0 + k * wpi) + L.n, L.kg, N, L.E, L.O);
}

// Sizes of the products of one workgroup (wave-uniform). a is the block
// operand, b the Toeplitz operand (for unbalanced products the shorter one:
// a tile's d range is at most dmax + 1 steps).
struct MfGeom {
    uint32_t na, nb;   // operand limbs
    uint32_t Pa, Pb;   // digit blocks, floor(n / 4) + 1
    uint32_t dmax;     // last d with a nonzero T_d, floor((8 nb + 31) / 32)
    uint32_t N, NT;    // output blocks below limb PL, column tiles
    uint32_t wpi, ipb; // waves per instance, instances per workgroup
    uint32_t PL, PLp;  // output limbs, LDS stride of E and O per instance
    uint32_t mg_pl;    // mf_magic(PL)
};

// One workgroup: ninst (<= ipb) products, instance s of a at a + s sa, of b
// at b + s sb, of the result at r + s sr.
template <int R>
__device__ __forceinline__ void mf_block(const MfGeom& g, uint32_t ninst, const uint64_t* __restrict__ a,
                                         uint64_t sa, const uint64_t* __restrict__ b, uint64_t sb,
                                         uint64_t* __restrict__ r, uint64_t sr, uint32_t pow8, uint64_t* lds64,
                                         uint32_t* lds_gp) {
    uint8_t* lds8 = reinterpret_cast<uint8_t*>(lds64);
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63;
    const uint32_t ipb = g.ipb;
    const uint32_t X0 = 32 * g.Pb + 31;
    const uint32_t A_BYTES = mf_a_bytes(g.Pa, ipb), BR_BYTES = mf_br_bytes(g.Pb);
    const uint32_t SB = 4 * (g.Pb + 3);  // br words per instance
    const uint32_t AH = A_BYTES / 16;     // u64 words of each half array of the a region
    uint64_t* Eb = lds64;
    uint64_t* Ob = lds64 + (size_t)ipb * g.PLp;
    uint8_t* Areg = lds8 + (size_t)16 * ipb * g.PLp;  // a blocks of all instances
    uint8_t* BRreg = Areg + A_BYTES;                   // br blocks, Pb + 3 per instance
    uint64_t* A64 = reinterpret_cast<uint64_t*>(Areg);
    uint64_t* BR64 = reinterpret_cast<uint64_t*>(BRreg);
    const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane;

    // ---- stage digits: one wave per operand (operand o < ipb: a of instance
    // o, else b of instance o - ipb), rows of 128 limbs in order with the
    // carry between rows in a scalar: x = limbs + 0x8080...80, digits =
    // x ^ 0x80 per byte below limb n, the carry at limb n and zeros above, up
    // to limb 4 (P + 1) (the blocks and, for a, the trailing zero block). a's
    // words go to the a region (a_word), b's to br in reverse order (word
    // SB - 9 - j for limb j, after br's 8 top zero words). The next row's
    // loads are issued before the current row is processed.
    {
        const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), waves = blockDim.x >> 6;
        const uint32_t nops = 2 * ipb;
        const uint32_t nrow_a = (g.na + kRowLimbs - 1) / kRowLimbs, nrow_b = (g.nb + kRowLimbs - 1) / kRowLimbs;
        // u64 word of limb j of instance s in the a region (block 1 + s (Pa + 1)
        // + j / 4, half (j / 2) mod 2)
        auto a_word = [&](uint32_t s, uint32_t j) {
            return ((j >> 1) & 1u) * AH + 2 * (1 + s * (g.Pa + 1) + (j >> 2)) + (j & 1u);
        };
        if (tid < 4) A64[(tid >> 1) * AH + (tid & 1u)] = 0;  // the zero block before instance 0
        auto fetch = [&](uint32_t o, uint32_t row, uint64_t& x0, uint64_t& x1) {
            const bool isb = o >= ipb;
            const uint32_t s = isb ? o - ipb : o;
            const uint32_t n = isb ? g.nb : g.na;
            const uint64_t* __restrict__ src = isb ? b + s * sb : a + s * sa;
            const uint32_t j0 = row * kRowLimbs + 2 * lane;
            if (s < ninst && (row + 1) * kRowLimbs <= n) {  // a whole row (wave-uniform)
                x0 = src[j0];
                x1 = src[j0 + 1];
            } else {
                x0 = s < ninst && j0 < n ? src[j0] : 0ull;
                x1 = s < ninst && j0 + 1 < n ? src[j0 + 1] : 0ull;
            }
        };
        uint32_t o = wv, row = 0, cin = 0;
        uint64_t nx0 = 0, nx1 = 0;
        if (o < nops) fetch(o, 0, nx0, nx1);
        while (o < nops) {
            const uint64_t x0 = nx0, x1 = nx1;
            const bool isb = o >= ipb;
            const uint32_t nrow = isb ? nrow_b : nrow_a;
            uint32_t on = o, rn = row + 1;
            if (rn == nrow) {
                rn = 0;
                on = o + waves;
            }
            if (on < nops) fetch(on, rn, nx0, nx1);
            const uint32_t s = isb ? o - ipb : o;
            const uint32_t n = isb ? g.nb : g.na;
            const uint32_t S4 = 4 * ((isb ? g.Pb : g.Pa) + 1);  // limbs written
            const uint32_t j0 = row * kRowLimbs + 2 * lane;
            const uint64_t h0 = j0 < n ? kMfHalf : 0ull, h1 = j0 + 1 < n ? kMfHalf : 0ull;
            const RowState st = row_build(x0, h0, x1, h1, false, false);
            const uint64_t C = row_chain(st, &cin);
            uint64_t r0, r1;
            row_apply(st, C, &r0, &r1);
            const uint64_t w0 = r0 ^ h0, w1 = r1 ^ h1;
            if (j0 < S4) {  // S4 and j0 even: both limbs
                if (!isb)
                    *reinterpret_cast<ulonglong2*>(A64 + a_word(s, j0)) = make_ulonglong2(w0, w1);
                else
                    *reinterpret_cast<ulonglong2*>(BR64 + s * SB + SB - 10 - j0) =
                        make_ulonglong2(__builtin_bswap64(w1), __builtin_bswap64(w0));
            }
            if (rn == 0) {
                // limbs past the rows (at most 8): the carry when the rows
                // end at limb n, zeros after; and br's top zero words
                const uint32_t jt = nrow * kRowLimbs + lane;
                if (jt < S4) {
                    const uint64_t w = jt == n ? (uint64_t)cin : 0ull;
                    if (!isb)
                        A64[a_word(s, jt)] = w;
                    else
                        BR64[s * SB + SB - 9 - jt] = __builtin_bswap64(w);
                }
                if (isb && lane < 8) BR64[s * SB + SB - 8 + lane] = 0;
                cin = 0;
            }
            o = on;
            row = rn;
        }
    }
    __syncthreads();

    // ---- MFMA phase: wave wi of instance s takes tiles wi, wi + wpi, ...
    // in passes of up to kMfTiles tiles over the union of their d ranges.
    {
        const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        // instance and first tile of the wave, by scalar subtraction (wv <
        // 16 waves): a division would run on the vector unit and leave
        // every loop bound in vector registers
        uint32_t s = 0, wi = wv;
        while (wi >= g.wpi) {
            wi -= g.wpi;
            s++;
        }
        MfLane L;
        L.lds8 = lds8;
        L.lane = lane;
        L.n = lane & 31;
        L.kg = lane >> 5;
        L.paddr = (lane ^ 32u) << 2;
        // block 0 of instance s in the lane's half array
        L.b_base = (uint32_t)(Areg - lds8) + L.kg * (A_BYTES / 2) + 16 * (1 + s * (g.Pa + 1));
        L.b_lo = L.b_base - 16;
        L.b_hi = L.b_base + 16 * g.Pa;
        const uint32_t* BR = reinterpret_cast<const uint32_t*>(BRreg + (size_t)s * BR_BYTES);
        const uint32_t ybase = X0 - L.n + 16 * L.kg;  // y0 at d = 0
        L.sh = (ybase & 3) * 8;
        L.Abase = BR + (ybase >> 2);
        L.E = Eb + (size_t)s * g.PLp;
        L.O = Ob + (size_t)s * g.PLp;
        const MfPow pw = mf_pow(pow8);
        for (uint32_t c0 = wi; c0 < g.NT; c0 += kMfTiles * g.wpi) {
            if (c0 + g.wpi < g.NT)
                mf_pass<2>(L, c0, g.wpi, g.Pa, g.dmax, g.N, pw);
            else
                mf_pass<1>(L, c0, g.wpi, g.Pa, g.dmax, g.N, pw);
        }
    }
    __syncthreads();

    // ---- normalisation: E + O over ipb segments of PL limbs (instance
    // stride PLp in LDS), R rows per wave
    const uint32_t PL = g.PL, PLp = g.PLp;
    const uint32_t Lw = ipb * PL;
    Rows<R> rows;
    uint32_t off0[R], instv[R];
    // the rows cover exactly the workgroup's ninst = ipb products of even
    // length: every limb pair is in range and inside one product, so the
    // pairs load and store without checks (E and O pairs are 16-byte aligned)
    const bool full = (blockDim.x >> 6) * R * kRowLimbs == Lw && ninst == ipb && (PL & 1u) == 0;
    {
        SegPos sp = mf_segpos(k_first, PL, g.mg_pl);
        const uint32_t sq = mf_div(kRowLimbs, g.mg_pl), srr = kRowLimbs - sq * PL;
        if (full) {
#pragma unroll
            for (int v = 0; v < R; v++) {
                const uint32_t e0 = sp.inst * PLp + sp.off;
                const ulonglong2 xe = *reinterpret_cast<const ulonglong2*>(Eb + e0);
                const ulonglong2 yo = *reinterpret_cast<const ulonglong2*>(Ob + e0);
                rows.x0[v] = xe.x;
                rows.x1[v] = xe.y;
                rows.y0[v] = yo.x;
                rows.y1[v] = yo.y;
                rows.s0[v] = sp.off == 0;
                rows.s1[v] = false;
                off0[v] = sp.off;
                instv[v] = sp.inst;
                segpos_next(&sp, PL, sq, srr);
            }
        } else
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            const bool wrap = sp.off + 1 == PL;
            const uint32_t e0 = sp.inst * PLp + sp.off;
            const uint32_t e1 = e0 + 1 + (wrap ? PLp - PL : 0u);
            rows.x0[v] = k < Lw ? Eb[e0] : 0ull;
            rows.x1[v] = k + 1 < Lw ? Eb[e1] : 0ull;
            rows.y0[v] = k < Lw ? Ob[e0] : 0ull;
            rows.y1[v] = k + 1 < Lw ? Ob[e1] : 0ull;
            rows.s0[v] = (sp.off == 0) && (k < Lw);
            rows.s1[v] = wrap && (k + 1 < Lw);
            off0[v] = sp.off;
            instv[v] = sp.inst;
            segpos_next(&sp, PL, sq, srr);
        }
    }
    // R 128-limb rows a multiple of PL: each wave holds whole products
    if ((R * kRowLimbs) % PL == 0)
        wave_add_rows<R>(rows);
    else
        block_add_rows<R>(rows, lds_gp);
    if (full) {
#pragma unroll
        for (int v = 0; v < R; v++) {
            uint64_t* dst = r + instv[v] * sr + off0[v];
            dst[0] = rows.x0[v];
            dst[1] = rows.x1[v];
        }
        return;
    }
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const uint32_t o0 = off0[v];
        const uint32_t i0 = instv[v];
        if (k < Lw && i0 < ninst) r[i0 * sr + o0] = rows.x0[v];
        const bool wrap = (o0 + 1 == PL);
        const uint32_t o1 = wrap ? 0u : o0 + 1;
        const uint32_t i1 = i0 + (wrap ? 1u : 0u);
        if (k + 1 < Lw && i1 < ninst) r[i1 * sr + o1] = rows.x1[v];
    }
}

// Batches of equal-size products: ipb instances per workgroup.
template <int R>
__global__ __launch_bounds__(1024) void mul_mfma_kernel(uint32_t n_inst, uint32_t m, uint32_t P, uint32_t dmax,
                                                        uint32_t N, uint32_t NT, uint32_t wpi, uint32_t ipb,
                                                        uint32_t PL, uint32_t PLp, uint32_t mg_pl, uint32_t pow8,
                                                        const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                        uint64_t sa, uint64_t sb, uint64_t sr) {
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t lds_gp[16];
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);
    const MfGeom g = {m, m, P, P, dmax, N, NT, wpi, ipb, PL, PLp, mg_pl};
    mf_block<R>(g, ninst, a + inst_base * sa, sa, b + inst_base * sb, sb, r + inst_base * sr, sr, pow8, lds64,
                lds_gp);
}

// Ragged batches: one product per workgroup (items[blockIdx.x], n_a >= n_b),
// all WAVES waves on its column tiles.
template <int R, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void mul_mfma_rg_kernel(const MfRgItem* __restrict__ items,
                                                                 uint32_t pow8) {
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t lds_gp[16];
    const MfRgItem it = items[blockIdx.x];
    MfGeom g;
    g.na = it.n_a;
    g.nb = it.n_b;
    g.Pa = g.na / 4 + 1;
    g.Pb = g.nb / 4 + 1;
    g.dmax = (8 * g.nb + 31) / 32;
    g.PL = g.na + g.nb;
    g.N = (g.PL + 3) / 4;
    g.NT = (g.N + 31) / 32;
    g.PLp = 4 * g.N + 4;
    g.wpi = WAVES;
    g.ipb = 1;
    g.mg_pl = it.mg_pl;
    mf_block<R>(g, 1, it.a, 0, it.b, 0, it.r, 0, pow8, lds64, lds_gp);
}

// z <- x + y per instance (mod B^m) over the ipb instances of a workgroup:
// x, y, z at instance strides sx, sy, sz (LDS or global; only instances
// below ninst are read or written), R rows of 128 limbs per wave covering
// ipb m limbs. Every wave of the workgroup must call it.
template <int R>
__device__ __forceinline__ void mf_seg_add(const uint64_t* x, uint64_t sx, const uint64_t* y, uint64_t sy,
                                           uint64_t* z, uint64_t sz, uint32_t m, uint32_t mg_m, uint32_t ipb,
                                           uint32_t ninst, uint32_t* lds_gp) {
    const uint32_t Lw = ipb * m;
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * (threadIdx.x & 63);
    SegPos sp = mf_segpos(k_first, m, mg_m);
    const uint32_t sq = mf_div(kRowLimbs, mg_m), srr = kRowLimbs - sq * m;
    Rows<R> rows;
    uint32_t off0[R], inst0[R];
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const bool wrap = sp.off + 1 == m;
        const uint32_t i1 = sp.inst + (wrap ? 1u : 0u), o1 = wrap ? 0u : sp.off + 1;
        const bool v0 = k < Lw && sp.inst < ninst, v1 = k + 1 < Lw && i1 < ninst;
        rows.x0[v] = v0 ? x[sp.inst * sx + sp.off] : 0ull;
        rows.y0[v] = v0 ? y[sp.inst * sy + sp.off] : 0ull;
        rows.x1[v] = v1 ? x[i1 * sx + o1] : 0ull;
        rows.y1[v] = v1 ? y[i1 * sy + o1] : 0ull;
        rows.s0[v] = sp.off == 0 && k < Lw;
        rows.s1[v] = wrap && k + 1 < Lw;
        off0[v] = sp.off;
        inst0[v] = sp.inst;
        segpos_next(&sp, m, sq, srr);
    }
    block_add_rows<R>(rows, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        const bool wrap = off0[v] + 1 == m;
        const uint32_t i1 = inst0[v] + (wrap ? 1u : 0u), o1 = wrap ? 0u : off0[v] + 1;
        if (k < Lw && inst0[v] < ninst) z[inst0[v] * sz + off0[v]] = rows.x0[v];
        if (k + 1 < Lw && i1 < ninst) z[i1 * sz + o1] = rows.x1[v];
    }
}

// The reference's polyKerQ (cuda/ker-classic-mul.cu.h:500-575) in one
// kernel: r = ((a^2 + b)(b^2 + b) + ab) mod B^m, every step mod B^m in the
// reference's order. The three products of the operands come from global
// memory into LDS buffers (T1 = a^2, T2 = b^2, T3 = ab), the additions run
// in LDS (T1 += b, T2 += b), the fourth product reads T1, T2 from LDS into
// T4, and r = T4 + T3 is the only global store. The matrix-core phases are
// mf_block's, with low-half products (PL = m).
template <int R>
__global__ __launch_bounds__(1024) void poly_mfma_kernel(uint32_t n_inst, uint32_t m, uint32_t P, uint32_t dmax,
                                                         uint32_t N, uint32_t NT, uint32_t wpi, uint32_t ipb,
                                                         uint32_t PLp, uint32_t mg_m, uint32_t t_off, uint32_t pow8,
                                                         const uint64_t* __restrict__ a,
                                                         const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                         uint64_t sa, uint64_t sb, uint64_t sr) {
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t lds_gp[16];
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);
    const MfGeom g = {m, m, P, P, dmax, N, NT, wpi, ipb, m, PLp, mg_m};
    const uint64_t* ga = a + inst_base * sa;
    const uint64_t* gb = b + inst_base * sb;
    uint64_t* T1 = lds64 + t_off;  // t_off: u64 words of mf_block's region
    uint64_t* T2 = T1 + (size_t)ipb * m;
    uint64_t* T3 = T2 + (size_t)ipb * m;
    uint64_t* T4 = T3 + (size_t)ipb * m;
    mf_block<R>(g, ninst, ga, sa, ga, sa, T1, m, pow8, lds64, lds_gp);  // T1 = a^2
    __syncthreads();
    mf_block<R>(g, ninst, gb, sb, gb, sb, T2, m, pow8, lds64, lds_gp);  // T2 = b^2
    __syncthreads();
    mf_block<R>(g, ninst, ga, sa, gb, sb, T3, m, pow8, lds64, lds_gp);  // T3 = a b
    __syncthreads();
    mf_seg_add<R>(T1, m, gb, sb, T1, m, m, mg_m, ipb, ninst, lds_gp);   // T1 = a^2 + b
    mf_seg_add<R>(T2, m, gb, sb, T2, m, m, mg_m, ipb, ninst, lds_gp);   // T2 = b^2 + b
    __syncthreads();
    mf_block<R>(g, ninst, T1, m, T2, m, T4, m, pow8, lds64, lds_gp);    // T4 = T1 T2
    __syncthreads();
    mf_seg_add<R>(T4, m, T3, m, r + inst_base * sr, sr, m, mg_m, ipb, ninst, lds_gp);  // r = T4 + T3
}

namespace {

struct MfmaLayout {
    uint32_t P, dmax, N, NT, wpi, ipb, waves, R, PL, PLp;
    size_t lds;
};

bool mfma_layout(uint32_t m, bool full, MfmaLayout* L) {
    if (m == 0 || m > kMulMaxLimbs) return false;
    L->P = m / 4 + 1;
    L->dmax = (8 * m + 31) / 32;  // (D + 30) / 32, D = 8m + 1
    L->PL = full ? 2 * m : m;
    L->N = (L->PL + 3) / 4;
    L->NT = (L->N + 31) / 32;
    L->PLp = 4 * L->N + 4;  // the last block's 8-limb slot fits
    int64_t w = tune_get_class("mul.mfma_wpi", m, 0);
    if (w <= 0) w = (L->NT + 2) / 4;  // about a quarter of the column tiles
    L->wpi = (uint32_t)std::max<int64_t>(1, std::min<int64_t>({w, 16, (int64_t)L->NT}));
    const int64_t ipb = tune_get_class("mul.mfma_ipb", m, 0);  // instances per workgroup, 0: 4 / wpi
    L->ipb = ipb > 0 ? (uint32_t)std::min<int64_t>(ipb, 16) : std::max(1u, 4 / L->wpi);
    if (L->ipb * L->wpi > 16) return false;
    L->waves = L->ipb * L->wpi;
    const uint64_t limbs = (uint64_t)L->ipb * L->PL;  // the normalisation's row window
    uint32_t R = 1;
    while ((uint64_t)L->waves * R * kRowLimbs < limbs) R *= 2;
    if (R > 8) return false;
    L->R = R;
    L->lds = (size_t)16 * L->ipb * L->PLp + mf_a_bytes(L->P, L->ipb) + (size_t)L->ipb * mf_br_bytes(L->P);
    // mf_div bound: window limb index (below waves R 128) x segment length
    const uint64_t kmax = (uint64_t)L->waves * R * kRowLimbs + 2;
    if (kmax * L->PL >= (1ull << 32)) return false;
    return L->lds <= 160 * 1024 - 256;
}

template <int R>
hipError_t launch_mfma_t(const MfmaLayout& L, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b,
                         uint64_t* r, const Strides& st, hipStream_t s) {
    const void* fn = (const void*)mul_mfma_kernel<R>;
    if (L.lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + L.ipb - 1) / L.ipb);
    mul_mfma_kernel<R><<<blocks, 64 * L.waves, L.lds, s>>>(n, m, L.P, L.dmax, L.N, L.NT, L.wpi, L.ipb, L.PL, L.PLp,
                                                           mf_magic(L.PL), 256u, a, b, r, st.a ? st.a : m,
                                                           st.b ? st.b : m, st.r ? st.r : L.PL);
    return hipGetLastError();
}

// The fused poly kernel: mfma_layout's low-half layout with four m-limb LDS
// buffers per instance after mf_block's region, fewer instances per
// workgroup if they do not fit.
bool poly_layout(uint32_t m, MfmaLayout* L, uint32_t* t_off) {
    if (!mfma_layout(m, false, L)) return false;
    for (;;) {
        const size_t region = (L->lds + 15) & ~(size_t)15;
        const size_t total = region + (size_t)4 * L->ipb * m * 8;
        if (total <= 160 * 1024 - 256) {
            *t_off = (uint32_t)(region / 8);
            L->lds = total;
            return true;
        }
        if (L->ipb == 1) return false;
        L->ipb /= 2;
        L->waves = L->ipb * L->wpi;
        uint32_t R = 1;
        while ((uint64_t)L->waves * R * kRowLimbs < (uint64_t)L->ipb * L->PL) R *= 2;
        L->R = R;
        L->lds = (size_t)16 * L->ipb * L->PLp + mf_a_bytes(L->P, L->ipb) + (size_t)L->ipb * mf_br_bytes(L->P);
    }
}

template <int R>
hipError_t launch_poly_t(const MfmaLayout& L, uint32_t t_off, uint32_t n, uint32_t m, const uint64_t* a,
                         const uint64_t* b, uint64_t* r, const Strides& st, hipStream_t s) {
    const void* fn = (const void*)poly_mfma_kernel<R>;
    if (L.lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t blocks = (uint32_t)(((uint64_t)n + L.ipb - 1) / L.ipb);
    poly_mfma_kernel<R><<<blocks, 64 * L.waves, L.lds, s>>>(n, m, L.P, L.dmax, L.N, L.NT, L.wpi, L.ipb, L.PLp,
                                                            mf_magic(m), t_off, 256u, a, b, r, st.a ? st.a : m,
                                                            st.b ? st.b : m, st.r ? st.r : m);
    return hipGetLastError();
}

}  // namespace

bool poly_fused_supported(uint32_t m) {
    MfmaLayout L;
    uint32_t t_off;
    return poly_layout(m, &L, &t_off);
}

hipError_t launch_poly_fused(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                             hipStream_t s, Strides st) {
    MfmaLayout L;
    uint32_t t_off;
    if (!poly_layout(m, &L, &t_off)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    switch (L.R) {
        case 1: return launch_poly_t<1>(L, t_off, n, m, a, b, r, st, s);
        case 2: return launch_poly_t<2>(L, t_off, n, m, a, b, r, st, s);
        case 4: return launch_poly_t<4>(L, t_off, n, m, a, b, r, st, s);
        default: return launch_poly_t<8>(L, t_off, n, m, a, b, r, st, s);
    }
}

bool mul_mfma_supported(uint32_t m, bool full) {
    MfmaLayout L;
    return mfma_layout(m, full, &L);
}

hipError_t launch_mul_mfma(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                           hipStream_t s, Strides st) {
    MfmaLayout L;
    if (!mfma_layout(m, full, &L)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    switch (L.R) {
        case 1: return launch_mfma_t<1>(L, n, m, a, b, r, st, s);
        case 2: return launch_mfma_t<2>(L, n, m, a, b, r, st, s);
        case 4: return launch_mfma_t<4>(L, n, m, a, b, r, st, s);
        default: return launch_mfma_t<8>(L, n, m, a, b, r, st, s);
    }
}

MfRgItem mfma_rg_item(const uint64_t* a, const uint64_t* b, uint64_t* r, uint32_t n_a, uint32_t n_b) {
    MfRgItem it = {};
    it.a = a;
    it.b = b;
    it.r = r;
    it.n_a = n_a;
    it.n_b = n_b;
    it.mg_pl = mf_magic(n_a + n_b);
    return it;
}

template <int WAVES>
static const void* mfma_rg_func(uint32_t R) {
    switch (R) {
        case 1: return (const void*)mul_mfma_rg_kernel<1, WAVES>;
        case 2: return (const void*)mul_mfma_rg_kernel<2, WAVES>;
        case 4: return (const void*)mul_mfma_rg_kernel<4, WAVES>;
        default: return (const void*)mul_mfma_rg_kernel<8, WAVES>;
    }
}

bool mfma_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out, uint32_t waves) {
    if (tune_get_class("ragged.m16", max_na, 0) == 1 && mfma16_rg_config(max_na, max_nb, max_pl, out, waves))
        return true;
    return mfma32_rg_config(max_na, max_nb, max_pl, out, waves);
}

bool mfma32_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out, uint32_t waves) {
    if (max_na == 0 || max_nb == 0 || max_na > kMulMaxLimbs || max_nb > max_na || max_pl > max_na + max_nb)
        return false;
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return false;
    uint32_t R = 1;
    while (waves * R * kRowLimbs < max_pl) R *= 2;
    if (R > 8) return false;
    // mf_div bound of the normalisation: window limb index x PL < 2^32
    if (((uint64_t)waves * R * kRowLimbs + 2) * max_pl >= (1ull << 32)) return false;
    const uint32_t N = (max_pl + 3) / 4, Pa = max_na / 4 + 1, Pb = max_nb / 4 + 1;
    out->lds = (size_t)16 * (4 * N + 4) + mf_a_bytes(Pa, 1) + mf_br_bytes(Pb);
    if (out->lds > 160 * 1024 - 256) return false;
    out->threads = 64 * waves;
    out->pow8 = 256u;
    switch (waves) {
        case 1: out->func = mfma_rg_func<1>(R); break;
        case 2: out->func = mfma_rg_func<2>(R); break;
        case 4: out->func = mfma_rg_func<4>(R); break;
        default: out->func = mfma_rg_func<8>(R); break;
    }
    return true;
}

}  // namespace midint
