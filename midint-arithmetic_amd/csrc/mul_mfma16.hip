// Classical multiplication of operands of any size on the int8 matrix cores'
// 16 x 16 x 64 instruction: one product per workgroup of W waves, the tile
// plan computed at run time. Same digits, bias and tiling as mul_wave16.hip
// (balanced base-256 digits, mfma_digits.hpp; output blocks of 16 digits;
// 64-digit windows of a every 16 digits against Toeplitz fragments T_k of
// b, c_{16I+r} = sum_k sum_{e<64} b_{64k+DL+r-e} a_{16(I-4k)-DL+e}); the
// reference's kernel for this operation is cuda/ker-classic-mul.cu.h:299-463.
//
// The 17 fragments T_k of an 8192-bit product fit one wave's registers; a
// 65536-bit product has 129. Here a wave takes a "pass" of four consecutive
// tiles t0 .. t0+3 (t0 = 4p, passes dealt to the waves in turn) and walks
// u = 4t - k upwards: at step u the four tiles need T_k for k = 4 t0 + 4i -
// u (i < 4), a window of 13 consecutive k that moves down by one per step.
// Sixteen register slots (slot k mod 16) hold it, so each step builds one
// new T_k (five dwords from LDS, four funnel shifts) and reads one B
// fragment (the windows of all four tiles at step u start at 16(n + 4u) -
// DL) for up to four MFMAs. Both loads run one step ahead of their MFMAs.
// The step loop is unrolled by 16 so every slot index is a compile-time
// constant (t0 is a multiple of 4, so k mod 16 = (4i - u) mod 16).
//
// A pass runs every step of its union of ranges with all four MFMAs: a
// (tile, step) pair outside the tile's range multiplies zeros only (zero
// margins around the digits, T_k = 0 for k < 0 from zero bytes above b's
// digits), so the steps need no per-tile test; the 16 slots are numbered
// from the pass's first step, which keeps their indices compile-time
// constants in blocks of 16 steps.
//
// Tile t needs the steps k in [kmin(t), kmax(t)]: column I (16-digit block)
// has a term iff its window [L, L + 63], L = 16I - 64k - DL, meets
// [max(0, 16I - Db + 1), min(Da - 1, 16I + 15)] (the digits of a that have a
// partner in b for some output digit of the block); both bounds are
// nondecreasing in I, so the tile's range runs from its first column's
// lower bound to its last column's upper bound.
//
// After its passes every wave folds its tiles' column sums into the words
// V_w (as in mul_wave16.hip), the workgroup synchronises, and the waves turn
// the words into limbs with one block-wide pass of the row adder.
#include <algorithm>

#include "add_core.hpp"
#include "mfma_digits.hpp"
#include "midint_internal.h"

namespace midint {

namespace {

constexpr int kM16Delta = 48;  // DL

__host__ __device__ constexpr int32_t m16_floordiv64(int32_t x) { return x >= 0 ? x / 64 : -((-x + 63) / 64); }

// Sizes and LDS layout (bytes) of one product; wave-uniform.
struct M16Geom {
    uint32_t na, nb;     // limbs of a (windows) and b (Toeplitz fragments)
    uint32_t PL;         // output limbs
    uint32_t nblk, T;    // 16-digit output blocks (ceil(PL / 2)), tiles of 16 blocks
    int32_t Da, Db;      // digits, 8n + 1
    uint32_t aoff;       // byte of a's digit 0 in the a region
    uint32_t a_bytes;    // a region
    uint32_t x0;         // br region: b's digit j at byte x0 - j
    uint32_t br_bytes;   // br region
    uint32_t v_bytes;    // words V (rows of 128 limbs, two words per limb)
    uint32_t bytes;      // a + br + 16 (zero words V_-2, V_-1) + V
};

// Step range of tile t.
__host__ __device__ inline void m16_krange(const M16Geom& g, uint32_t t, int32_t* kmin, int32_t* kmax) {
    const int32_t I0 = 16 * (int32_t)t;
    const int32_t I1 = std::min<int32_t>(I0 + 15, (int32_t)g.nblk - 1);
    const int32_t ihi0 = std::min<int32_t>(g.Da - 1, 16 * I0 + 15);
    const int32_t ilo1 = std::max<int32_t>(0, 16 * I1 - g.Db + 1);
    *kmin = -m16_floordiv64(-(16 * I0 - kM16Delta - ihi0));  // ceil
    *kmax = m16_floordiv64(16 * I1 - kM16Delta + 63 - ilo1);
}

__host__ __device__ inline uint32_t m16_round16(uint32_t x) { return (x + 15u) & ~15u; }

// The last step with a nonzero T_k: 64k + DL - 63 <= Db - 1.
__host__ __device__ inline int32_t m16_kcap(const M16Geom& g) { return (g.Db + 62 - kM16Delta) / 64; }

__host__ __device__ inline M16Geom m16_geom(uint32_t na, uint32_t nb, uint32_t PL) {
    M16Geom g;
    g.na = na;
    g.nb = nb;
    g.PL = PL;
    g.nblk = (PL + 1) / 2;
    g.T = (g.nblk + 15) / 16;
    g.Da = 8 * (int32_t)na + 1;
    g.Db = 8 * (int32_t)nb + 1;
    // window bytes: AOFF + 16 (n + g) + 64 u - DL, u = 4t - k over the used
    // (t, k). Both ends of a tile's step range grow by at most 4 from one
    // tile to the next (its columns move on by 16 blocks, 256 digits, 4
    // steps, and the clamps to the operands only slow that), so 4t - kmax(t)
    // and 4t - kmin(t) never decrease: the first tile gives the smallest u
    // and the last tile the largest (every tile has a nonempty range). This
    // is O(1) work per product instead of a loop over every tile, which
    // made this the largest share of the kernel's scalar instructions for
    // ragged products (checked against that loop for every pair of operand
    // sizes up to 2048 limbs and every uniform size up to 4096).
    int32_t lo0, hi0, loL, hiL;
    m16_krange(g, 0, &lo0, &hi0);
    m16_krange(g, g.T - 1, &loL, &hiL);
    const int32_t umin = -hi0;
    const int32_t umax = 4 * (int32_t)(g.T - 1) - loL;
    (void)lo0;
    (void)hiL;
    g.aoff = (uint32_t)std::max<int32_t>(0, kM16Delta - 64 * umin);
    g.aoff = m16_round16(g.aoff);
    // (the window of step umax + 1 is read, not used: one more step of bytes)
    const uint32_t a_end = (uint32_t)((int32_t)g.aoff + 16 * 18 + 64 * (umax + 1) - kM16Delta + 16);
    g.a_bytes = m16_round16(std::max<uint32_t>(a_end, g.aoff + 8 * na + 8));
    // T_k (0 <= k <= kcap, the last k with a digit of b) reads five dwords
    // from byte (x0 - 64k - DL - r + 16g) & ~3, r < 16, g < 4; x0 = 15 mod 16
    // puts b's limb pairs on 16-byte boundaries
    const uint32_t x0min = std::max<uint32_t>((uint32_t)(64 * m16_kcap(g) + kM16Delta + 15), 8 * nb + 7);
    g.x0 = x0min + (15u + 16u - x0min % 16u) % 16u;
    // zeros for k down to -13 above (the passes' last steps read them)
    g.br_bytes = m16_round16((g.x0 & ~3u) + 20 + 13 * 64);
    g.v_bytes = (PL + kRowLimbs - 1) / kRowLimbs * kRowLimbs * 16;
    g.bytes = g.a_bytes + g.br_bytes + 16 + g.v_bytes;
    return g;
}

// One product per workgroup: waves W = blockDim.x / 64, R rows of 128 limbs
// per wave in the normalisation (W R 128 >= PL).
template <int R>
__device__ __forceinline__ void m16_block(const M16Geom& g, const uint64_t* __restrict__ a,
                                          const uint64_t* __restrict__ b, uint64_t* __restrict__ r, uint32_t pow8,
                                          uint8_t* lds, uint32_t* lds_gp) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), waves = blockDim.x >> 6;
    uint8_t* areg = lds;
    uint8_t* breg = lds + g.a_bytes;
    uint8_t* vreg = breg + g.br_bytes + 16;

    // ---- zeros around the digits and the words V_-2, V_-1
    for (uint32_t x = 16 * tid; x < g.a_bytes + g.br_bytes + 16; x += 16 * blockDim.x)
        *reinterpret_cast<uint4*>(lds + x) = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();

    // ---- digits: x = limbs + 0x8080...80 as one block-wide row addition per
    // operand (wave w holds rows w R .. w R + R - 1, every row's limbs load at
    // once, the carry crosses the waves through LDS), digit bytes x ^ 0x80;
    // the carry out is digit 8n: limb n of x (0 + 0 + carry) when it lies in
    // the rows, else the addition's carry out
    for (uint32_t o = 0; o < 2; o++) {
        const uint32_t n = o ? g.nb : g.na;
        const uint64_t* src = o ? b : a;
        Rows<R> t;
        uint64_t h0[R], h1[R];
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t j0 = (wv * R + v) * kRowLimbs + 2 * lane;
            t.x0[v] = j0 < n ? src[j0] : 0ull;
            t.x1[v] = j0 + 1 < n ? src[j0 + 1] : 0ull;
            h0[v] = j0 < n ? kMfHalf : 0ull;
            h1[v] = j0 + 1 < n ? kMfHalf : 0ull;
            t.y0[v] = h0[v];
            t.y1[v] = h1[v];
            t.s0[v] = false;
            t.s1[v] = false;
        }
        uint32_t c = 0;
        block_add_rows_carry<R>(t, lds_gp, &c);
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t j0 = (wv * R + v) * kRowLimbs + 2 * lane;
            const uint64_t w0 = t.x0[v] ^ h0[v], w1 = t.x1[v] ^ h1[v];
            if (o == 0) {
                if (j0 <= n) *reinterpret_cast<uint64_t*>(areg + g.aoff + 8 * j0) = w0;
                if (j0 + 1 <= n) *reinterpret_cast<uint64_t*>(areg + g.aoff + 8 * (j0 + 1)) = w1;
            } else {
                if (j0 <= n) *reinterpret_cast<uint64_t*>(breg + g.x0 - 8 * j0 - 7) = __builtin_bswap64(w0);
                if (j0 + 1 <= n) *reinterpret_cast<uint64_t*>(breg + g.x0 - 8 * (j0 + 1) - 7) = __builtin_bswap64(w1);
            }
        }
        if (n == waves * R * kRowLimbs && tid == 0) {
            if (o == 0)
                *reinterpret_cast<uint64_t*>(areg + g.aoff + 8 * n) = (uint64_t)c;
            else
                *reinterpret_cast<uint64_t*>(breg + g.x0 - 8 * n - 7) = __builtin_bswap64((uint64_t)c);
        }
    }
    __syncthreads();

    // ---- matrix-core passes
    const uint32_t ln = lane & 15, lg = lane >> 4;
    // T_k: five dwords from byte (x0 - 64k - DL - r + 16g) & ~3; per lane the
    // byte at k = 0 and its misalignment
    const int32_t y0 = (int32_t)g.x0 - kM16Delta - (int32_t)ln + 16 * (int32_t)lg;
    const uint32_t sh = 8 * ((uint32_t)y0 & 3u);
    const int32_t ya = y0 & ~3;
    // B_u: 16 bytes at aoff + 16 (n + g) + 64 u - DL
    const int32_t yb = (int32_t)g.aoff + 16 * (int32_t)(ln + lg) - kM16Delta;
    const MfPow pw = mf_pow(pow8);
    const uint32_t passes = (g.T + 3) / 4;
#pragma unroll 1
    for (uint32_t p = wv; p < passes; p += waves) {
        const int32_t t0 = 4 * (int32_t)p;
        int32_t kmin[4], kmax[4];
        int32_t umin = 1 << 30, umax = -(1 << 30);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (t0 + i < (int32_t)g.T) {
                m16_krange(g, (uint32_t)(t0 + i), &kmin[i], &kmax[i]);
            } else {
                kmin[i] = 1;
                kmax[i] = 0;
            }
            if (kmin[i] <= kmax[i]) {
                umin = min(umin, 4 * (t0 + i) - kmax[i]);
                umax = max(umax, 4 * (t0 + i) - kmin[i]);
            }
            kmin[i] = __builtin_amdgcn_readfirstlane(kmin[i]);
            kmax[i] = __builtin_amdgcn_readfirstlane(kmax[i]);
        }
        umin = __builtin_amdgcn_readfirstlane(umin);
        umax = __builtin_amdgcn_readfirstlane(umax);
        mf_v4i acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = mf_v4i{0, 0, 0, 0};
        if (umin <= umax) {
            // step s = u - umin; tile i uses k = kap(s) + 4i, kap(s) = 4 t0 -
            // umin - s, in slot (4i - s) mod 16 (slot of k: (k - kap(0)) mod 16)
            const int32_t kap0 = 4 * t0 - umin;
            const int32_t ns = umax - umin + 1;
            const int32_t kcap = m16_kcap(g);
            mf_v4i Tk[16];
            auto loadT = [&](int32_t k, uint32_t (&q)[5]) {
                const uint32_t* src = reinterpret_cast<const uint32_t*>(breg + (ya - 64 * k));
#pragma unroll
                for (int e = 0; e < 5; e++) q[e] = src[e];
            };
            auto finishT = [&](const uint32_t (&q)[5]) {
                return mf_v4i{(int)__builtin_amdgcn_alignbit(q[1], q[0], sh), (int)__builtin_amdgcn_alignbit(q[2], q[1], sh),
                              (int)__builtin_amdgcn_alignbit(q[3], q[2], sh), (int)__builtin_amdgcn_alignbit(q[4], q[3], sh)};
            };
            // slots of k = kap0 .. kap0 + 12 (slot (k - kap0) mod 16; k above
            // kcap is zero and not read)
#pragma unroll
            for (int d = 0; d <= 12; d++) {
                const int32_t k = kap0 + d;
                uint32_t q[5] = {0, 0, 0, 0, 0};
                if (k <= kcap) loadT(k, q);
                Tk[d] = finishT(q);
            }
            mf_v4i bcur = *reinterpret_cast<const mf_v4i*>(areg + (yb + 64 * umin));
            // one step: loads for step s + 1 (T_k, k = kap(s) - 1, into slot
            // (15 - j) mod 16, reading zeros for k >= -13 below 0; B for u +
            // 1), the four MFMAs of step s, then the new slot
            auto step = [&](int32_t sb, int j) {
                const int32_t s1 = sb + j + 1;
                uint32_t q[5];
                loadT(kap0 - s1, q);
                const mf_v4i bnext = *reinterpret_cast<const mf_v4i*>(areg + (yb + 64 * (umin + s1)));
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < 4; i++)
                    acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Tk[(4 * i - j + 16) & 15], bcur, acc[i], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                Tk[(15 - j) & 15] = finishT(q);
                bcur = bnext;
            };
            int32_t sb = 0;
#pragma unroll 1
            for (; sb + 16 <= ns; sb += 16) {
#pragma unroll
                for (int j = 0; j < 16; j++) step(sb, j);
            }
            if (sb < ns) {
#pragma unroll
                for (int j = 0; j < 16; j++)
                    if (sb + j < ns) step(sb, j);
            }
        }
        // column sums of the pass's tiles to words V_w, w = 64 t + 4 n + g
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t t = (uint32_t)(t0 + i);
            if (t >= g.T) break;
            int64_t V = (int64_t)kMfFoldBias;
            V += (int64_t)acc[i][0] * (int32_t)pw.p0;
            V += (int64_t)acc[i][1] * (int32_t)pw.p8;
            V += (int64_t)acc[i][2] * (int32_t)pw.p16;
            V += (int64_t)acc[i][3] * (int32_t)pw.p24;
            const uint32_t w = 64 * t + 4 * ln + lg;
            if (w < g.v_bytes / 8) *reinterpret_cast<int64_t*>(vreg + 8 * w) = V;
        }
    }
    __syncthreads();

    // ---- limbs: x_L = W_L mod 2^64, y_L = W_{L-1} >> 64 (W_L = V_{2L} +
    // V_{2L+1} 2^32), x + y across the workgroup's rows
    Rows<R> rows;
    const uint32_t k_first = wv * R * kRowLimbs + 2 * lane;
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t L0 = k_first + v * kRowLimbs;  // limbs L0, L0 + 1
        uint64_t x0 = 0, x1 = 0, y0v = 0, y1v = 0;
        if (L0 < g.PL) {
            const uint8_t* pv = vreg + 16 * L0;  // word 2 L0
            const ulonglong2 v01 = *reinterpret_cast<const ulonglong2*>(pv - 16);
            ulonglong2 v23 = *reinterpret_cast<const ulonglong2*>(pv);
            const ulonglong2 v45 = *reinterpret_cast<const ulonglong2*>(pv + 16);
            if (L0 == 0) v23.x += kMfBias;  // V_0: column 0 takes 256 B
            const uint64_t sp = v01.y + (v01.x >> 32);
            const uint64_t s0 = v23.y + (v23.x >> 32);
            const uint64_t s1 = v45.y + (v45.x >> 32);
            x0 = (v23.x & 0xffffffffull) | (s0 << 32);
            y0v = sp >> 32;
            if (L0 + 1 < g.PL) {
                x1 = (v45.x & 0xffffffffull) | (s1 << 32);
                y1v = s0 >> 32;
            }
        }
        rows.x0[v] = x0;
        rows.x1[v] = x1;
        rows.y0[v] = y0v;
        rows.y1[v] = y1v;
        rows.s0[v] = false;
        rows.s1[v] = false;
    }
    block_add_rows<R>(rows, lds_gp);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t L0 = k_first + v * kRowLimbs;
        if (L0 < g.PL) r[L0] = rows.x0[v];
        if (L0 + 1 < g.PL) r[L0 + 1] = rows.x1[v];
    }
}

}  // namespace

// Batches of equal-size products, one per workgroup.
template <int R>
__global__ __launch_bounds__(1024) void mul_mfma16_kernel(uint32_t n_inst, uint32_t m, uint32_t PL, uint32_t pow8,
                                                          const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                          uint64_t* __restrict__ r, uint64_t sa, uint64_t sb, uint64_t sr) {
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t lds_gp[16];
    const M16Geom g = m16_geom(m, m, PL);
    const uint64_t inst = blockIdx.x;
    if (inst >= n_inst) return;
    m16_block<R>(g, a + inst * sa, b + inst * sb, r + inst * sr, pow8, reinterpret_cast<uint8_t*>(lds64), lds_gp);
}

// Ragged batches: one product per workgroup (items[blockIdx.x], n_a >= n_b),
// all waves on its passes. The shorter operand gives the windows and the
// longer one the Toeplitz fragments: a tile's steps then advance with its
// columns, so the four tiles of a pass share one short run of steps (for
// 1024 x 32 limbs 2.15x the useful digit products issued instead of 4.36x).
template <int R>
__global__ __launch_bounds__(1024) void mul_mfma16_rg_kernel(const MfRgItem* __restrict__ items, uint32_t pow8) {
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t lds_gp[16];
    const MfRgItem it = items[blockIdx.x];
    const M16Geom g = m16_geom(it.n_b, it.n_a, it.n_a + it.n_b);
    m16_block<R>(g, it.b, it.a, it.r, pow8, reinterpret_cast<uint8_t*>(lds64), lds_gp);
}

namespace {

struct M16Layout {
    uint32_t waves, R, PL;
    size_t lds;
};

bool m16_layout(uint32_t m, bool full, M16Layout* L) {
    if (m == 0 || m > kMulMaxLimbs) return false;
    L->PL = full ? 2 * m : m;
    const M16Geom g = m16_geom(m, m, L->PL);
    const uint32_t passes = (g.T + 3) / 4;
    int64_t w = tune_get_class("mul.m16_waves", m, 0);
    // measured: two passes per wave up to 8 passes (512 limbs), four beyond
    if (w <= 0) w = passes >= 16 ? passes / 4 : std::max<uint32_t>(1, passes / 2);
    L->waves = (uint32_t)std::min<int64_t>(16, std::max<int64_t>(1, w));
    // at least 4 rows per wave (rows past the product are inert); with 1 or
    // 2 the compiler's schedule of the passes spills
    uint32_t R = 4;
    while ((uint64_t)L->waves * R * kRowLimbs < L->PL) R *= 2;
    if (R > 8) return false;
    L->R = R;
    L->lds = g.bytes;
    return L->lds <= 160 * 1024 - 256;
}

template <int R>
hipError_t launch_m16_t(const M16Layout& L, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                        const Strides& st, hipStream_t s) {
    const void* fn = (const void*)mul_mfma16_kernel<R>;
    if (L.lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    mul_mfma16_kernel<R><<<n, 64 * L.waves, L.lds, s>>>(n, m, L.PL, 256u, a, b, r, st.a ? st.a : m, st.b ? st.b : m,
                                                        st.r ? st.r : L.PL);
    return hipGetLastError();
}

}  // namespace

bool mul_mfma16_supported(uint32_t m, bool full) {
    M16Layout L;
    return m16_layout(m, full, &L);
}

hipError_t launch_mul_mfma16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                             hipStream_t s, Strides st) {
    M16Layout L;
    if (!m16_layout(m, full, &L)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    return L.R == 4 ? launch_m16_t<4>(L, n, m, a, b, r, st, s) : launch_m16_t<8>(L, n, m, a, b, r, st, s);
}

// The geometry's LDS bytes grow with each of n_a, n_b and the product
// length (the a region with the last tile's windows and n_a, the br region
// with n_b, the words with the length), so the largest item of a launch
// sizes it.
bool mfma16_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out, uint32_t waves) {
    if (max_na == 0 || max_nb == 0 || max_na > kMulMaxLimbs || max_nb > max_na || max_pl > max_na + max_nb)
        return false;
    if (waves == 0 || waves > 16) return false;
    uint32_t R = 4;
    while ((uint64_t)waves * R * kRowLimbs < max_pl) R *= 2;
    if (R > 8) return false;
    const M16Geom g = m16_geom(max_nb, max_na, max_pl);  // windows: the shorter operand
    if (g.bytes > 160 * 1024 - 256) return false;
    out->lds = g.bytes;
    out->threads = 64 * waves;
    out->pow8 = 256u;
    out->func = R == 4 ? (const void*)mul_mfma16_rg_kernel<4> : (const void*)mul_mfma16_rg_kernel<8>;
    return true;
}

}  // namespace midint
