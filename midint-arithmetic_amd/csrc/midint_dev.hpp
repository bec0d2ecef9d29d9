// Device-side building blocks shared by the midint kernels (gfx950, wave64).
//
// Carry resolution. The reference resolves carries with a block-wide
// Hillis-Steele scan of a 3-bit (carry, all-ones, segment-start) flag over
// 32-lane warps and shared memory (cuda/ker-addition.cu.h:50-134). Here a
// wave64 resolves 64 lane-carries at once from two ballots: with generate
// mask G (lane sum overflowed) and propagate mask P (lane sum is all-ones),
// the carry into every lane is the bit-carry vector of the 64-bit integer
// addition G + (G|P) + cin, i.e. (G + (G|P) + cin) ^ G ^ (G|P). Segment
// (instance) boundaries are cut by clearing the G and P bits of the lane just
// before a segment start. Across waves, each wave contributes one
// (generate, propagate) pair through LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace midint {

constexpr int kWave = 64;
constexpr int kRowLimbs = 128;  // one row = 64 lanes x 2 limbs (16 bytes per lane)

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Carries into each lane of a 64-lane row. X = generate, Y = generate|propagate
// (both already cut at segment boundaries). Returns the per-lane carry-in bit
// vector; *cout receives the carry out of lane 63.
__device__ __forceinline__ uint64_t row_carries(uint64_t X, uint64_t Y, uint32_t cin, uint32_t* cout) {
    uint64_t s1 = X + Y;
    uint32_t c1 = s1 < X;
    uint64_t s2 = s1 + (uint64_t)cin;
    uint32_t c2 = s2 < s1;
    *cout = c1 | c2;
    return s2 ^ X ^ Y;
}

// Row-wise sums and wave-uniform carry masks for one 128-limb row. Each lane
// holds limbs (2*lane, 2*lane+1) of the row.
struct RowState {
    uint64_t r0, r1;      // limb sums without incoming carries
    uint64_t G, P;        // lane-level generate / propagate (wave-uniform)
    uint64_t S;           // lanes whose limb 0 starts a segment (wave-uniform)
    uint64_t G0, P0, S1;  // limb-0 generate/propagate, limb-1 segment start
};

// Builds the row state from the two operands of each lane. s0/s1: limb starts
// a segment. Invalid limbs must be passed as x=y=0 with s=false and are then
// kill positions only if followed by a segment start (never, past the end).
__device__ __forceinline__ RowState row_build(uint64_t x0, uint64_t y0, uint64_t x1, uint64_t y1,
                                              bool s0, bool s1) {
    RowState st;
    st.r0 = x0 + y0;
    st.r1 = x1 + y1;
    bool g0 = st.r0 < x0, p0 = st.r0 == ~0ull;
    bool g1 = st.r1 < x1, p1 = st.r1 == ~0ull;
    // lane aggregate: carry-in enters limb 0, reaches limb 1 unless limb 1
    // starts a new segment.
    bool gl = g1 || (p1 && g0 && !s1);
    bool pl = p0 && p1 && !s1;
    st.G = ballot(gl);
    st.P = ballot(pl);
    st.S = ballot(s0);
    st.G0 = ballot(g0);
    st.P0 = ballot(p0);
    st.S1 = ballot(s1);
    return st;
}

// Propagates the carry chain through one row. cin: carry into lane 0 (ignored
// when lane 0 starts a segment). Returns lane carry-in mask, updates *c to the
// row carry-out.
__device__ __forceinline__ uint64_t row_chain(const RowState& st, uint32_t* c) {
    uint32_t cin = (st.S & 1ull) ? 0u : *c;
    uint64_t K = st.S >> 1;  // lane i kills its carry-out if lane i+1 starts a segment
    uint64_t X = st.G & ~K;
    uint64_t Y = (st.G | st.P) & ~K;
    uint32_t cout;
    uint64_t C = row_carries(X, Y, cin, &cout);
    *c = cout;
    return C;
}

// Applies the lane carry-in mask C to the lane's two limb sums.
__device__ __forceinline__ void row_apply(const RowState& st, uint64_t C, uint64_t* out0, uint64_t* out1) {
    uint32_t lane = lane_id();
    uint64_t c0 = (C >> lane) & 1ull;
    // carry into limb 1: limb 0's carry-out, unless limb 1 starts a segment
    uint64_t C1 = ~st.S1 & (st.G0 | (st.P0 & C));
    uint64_t c1 = (C1 >> lane) & 1ull;
    *out0 = st.r0 + c0;
    *out1 = st.r1 + c1;
}

// Carry out of a limb given its sum state and carry-in (segment carry-out).
__device__ __forceinline__ uint32_t limb_cout(uint64_t x, uint64_t r, uint32_t cin) {
    return (r < x) | ((r == ~0ull) & cin);
}

// ---------------------------------------------------------------------------
// Segment bookkeeping: position of a window limb inside its segment and the
// segment (instance) index, advanced row by row without divisions.
struct SegPos {
    uint32_t off;   // offset of the lane's limb 0 inside its segment
    uint32_t inst;  // segment index within the window
};

__device__ __forceinline__ SegPos segpos_init(uint32_t k, uint32_t seg) {
    SegPos s;
    s.inst = k / seg;
    s.off = k - s.inst * seg;
    return s;
}

// Advance by kRowLimbs limbs; step_q = 128 / seg, step_r = 128 % seg.
__device__ __forceinline__ void segpos_next(SegPos* s, uint32_t seg, uint32_t step_q, uint32_t step_r) {
    s->off += step_r;
    s->inst += step_q;
    if (s->off >= seg) {
        s->off -= seg;
        s->inst += 1;
    }
}

// ---------------------------------------------------------------------------
// Block-level carry exchange between waves. Each wave publishes the
// carry-out of its whole range for cin=0 (g) and whether cin=1 changes it
// (p); wave w's carry-in is the fold of waves 0..w-1.
__device__ __forceinline__ uint32_t block_wave_carry_in(uint32_t g, uint32_t p, uint32_t* lds_gp) {
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t nwaves = blockDim.x / kWave;
    if (nwaves == 1) return 0;
    if (lane_id() == 0) lds_gp[wave] = g | (p << 1);
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t u = 0; u < wave; u++) {
        uint32_t v = lds_gp[u];
        c = (v & 1u) | ((v >> 1) & c);
    }
    __syncthreads();
    return c;
}

// ---------------------------------------------------------------------------
// Goldilocks field p = 2^64 - 2^32 + 1. Values are kept canonical (< p).
constexpr uint64_t kGlP = 0xFFFFFFFF00000001ull;
constexpr uint64_t kGlEps = 0xFFFFFFFFull;  // 2^64 mod p

__device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    bool c = s < a;
    return (c || s >= kGlP) ? s - kGlP : s;
}

__device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    return (a < b) ? d + kGlP : d;
}

// x = hi * 2^64 + lo reduced with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
__device__ __forceinline__ uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
    uint64_t hh = hi >> 32;
    uint64_t hl = hi & 0xFFFFFFFFull;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= kGlEps;
    uint64_t t1 = (hl << 32) - hl;
    uint64_t t2 = t0 + t1;
    if (t2 < t1) t2 += kGlEps;
    return t2 >= kGlP ? t2 - kGlP : t2;
}

__device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b) {
    return gl_reduce128(a * b, __umul64hi(a, b));
}

}  // namespace midint
