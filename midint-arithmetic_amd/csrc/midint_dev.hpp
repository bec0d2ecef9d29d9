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

// Like block_wave_carry_in, with a carry `cin` into wave 0: returns the carry
// into this wave and sets *cblock to the carry out of the block's last wave.
__device__ __forceinline__ uint32_t block_wave_carry_chain(uint32_t g, uint32_t p, uint32_t cin, uint32_t* lds_gp,
                                                           uint32_t* cblock) {
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t nwaves = blockDim.x / kWave;
    if (nwaves == 1) {
        *cblock = g | (p & cin);
        return cin;
    }
    if (lane_id() == 0) lds_gp[wave] = g | (p << 1);
    __syncthreads();
    uint32_t c = cin, mine = cin;
    for (uint32_t u = 0; u < nwaves; u++) {
        if (u == wave) mine = c;
        const uint32_t v = lds_gp[u];
        c = (v & 1u) | ((v >> 1) & c);
    }
    __syncthreads();
    *cblock = c;
    return mine;
}

// ---------------------------------------------------------------------------
// Goldilocks field p = 2^64 - 2^32 + 1. Values are kept canonical (< p).
constexpr uint64_t kGlP = 0xFFFFFFFF00000001ull;
constexpr uint64_t kGlEps = 0xFFFFFFFFull;  // 2^64 mod p

// The add/sub/reduce steps are written with the 32-bit carry chain
// (v_add_co / v_addc / v_sub_co / v_subb) so that every carry and borrow
// comes from the flag the add itself produces; the compiler otherwise
// rebuilds them with 64-bit compares. All inputs must be canonical.
//
// Each step is ONE asm statement and its carry masks die inside it: the
// compiler cannot see the SGPR writes of an asm statement, so a mask handed
// to later code could be spilled (v_writelane) or read without the wait
// states that VALU-written SGPRs need.

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint64_t join32(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// a + b mod p for canonical a, b: b' = b + eps cannot overflow (b < p), and
// t = a + b' = a + b - p + 2^64 carries exactly when a + b >= p; then t is
// the result, otherwise t - eps (= a + b, no borrow since t >= eps). 7 VALU.
__device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b) {
    uint32_t t0, t1, u0, u1, e;
    uint64_t k;
    asm("v_add_co_u32_e64 %[u0], %[k], %[b0], -1\n\t"
        "v_addc_co_u32_e64 %[u1], %[k], %[b1], 0, %[k]\n\t"
        "v_add_co_u32_e64 %[t0], %[k], %[a0], %[u0]\n\t"
        "v_addc_co_u32_e64 %[t1], %[k], %[a1], %[u1], %[k]\n\t"
        "v_cndmask_b32_e64 %[e], -1, 0, %[k]\n\t"
        "v_sub_co_u32_e64 %[t0], %[k], %[t0], %[e]\n\t"
        "v_subb_co_u32_e64 %[t1], %[k], %[t1], 0, %[k]"
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [u0] "=&v"(u0), [u1] "=&v"(u1), [e] "=&v"(e), [k] "=&s"(k)
        : [a0] "v"(lo32(a)), [a1] "v"(hi32(a)), [b0] "v"(lo32(b)), [b1] "v"(hi32(b)));
    return join32(t0, t1);
}

// a - b mod p: d = a - b, and on borrow d - eps (= d + p mod 2^64), which
// cannot borrow again for canonical inputs.
__device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b) {
    uint32_t d0, d1, e;
    uint64_t k;
    asm("v_sub_co_u32_e64 %[d0], %[k], %[a0], %[b0]\n\t"
        "v_subb_co_u32_e64 %[d1], %[k], %[a1], %[b1], %[k]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]\n\t"
        "v_sub_co_u32_e64 %[d0], %[k], %[d0], %[e]\n\t"
        "v_subb_co_u32_e64 %[d1], %[k], %[d1], 0, %[k]"
        : [d0] "=&v"(d0), [d1] "=&v"(d1), [e] "=&v"(e), [k] "=&s"(k)
        : [a0] "v"(lo32(a)), [a1] "v"(hi32(a)), [b0] "v"(lo32(b)), [b1] "v"(hi32(b)));
    return join32(d0, d1);
}

// x = hi * 2^64 + lo (lo = l1:l0, hi = hh:hl) reduced with 2^64 = 2^32 - 1
// and 2^96 = -1 (mod p):  t0 = lo - hh (borrow: -eps), t2 = t0 + hl*eps
// (carry: +eps; it cannot carry twice), then one conditional subtraction of
// p, done as t2 + eps with the carry as the condition.
__device__ __forceinline__ uint64_t gl_reduce(uint32_t l0, uint32_t l1, uint32_t hl, uint32_t hh) {
    uint32_t x0, x1, y0, y1, e;
    uint64_t k;
    asm("v_sub_co_u32_e64 %[x0], %[k], %[l0], %[hh]\n\t"
        "v_subb_co_u32_e64 %[x1], %[k], %[l1], 0, %[k]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]\n\t"
        "v_sub_co_u32_e64 %[x0], %[k], %[x0], %[e]\n\t"
        "v_subb_co_u32_e64 %[x1], %[k], %[x1], 0, %[k]\n\t"
        "v_sub_co_u32_e64 %[y0], %[k], 0, %[hl]\n\t"
        "v_subb_co_u32_e64 %[y1], %[k], %[hl], 0, %[k]\n\t"
        "v_add_co_u32_e64 %[x0], %[k], %[x0], %[y0]\n\t"
        "v_addc_co_u32_e64 %[x1], %[k], %[x1], %[y1], %[k]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]\n\t"
        "v_add_co_u32_e64 %[x0], %[k], %[x0], %[e]\n\t"
        "v_addc_co_u32_e64 %[x1], %[k], %[x1], 0, %[k]\n\t"
        "v_add_co_u32_e64 %[y0], %[k], %[x0], -1\n\t"
        "v_addc_co_u32_e64 %[y1], %[k], %[x1], 0, %[k]\n\t"
        "v_cndmask_b32_e64 %[x0], %[x0], %[y0], %[k]\n\t"
        "v_cndmask_b32_e64 %[x1], %[x1], %[y1], %[k]"
        : [x0] "=&v"(x0), [x1] "=&v"(x1), [y0] "=&v"(y0), [y1] "=&v"(y1), [e] "=&v"(e), [k] "=&s"(k)
        : [l0] "v"(l0), [l1] "v"(l1), [hl] "v"(hl), [hh] "v"(hh));
    return join32(x0, x1);
}

// The same reduction with the low half as one 64-bit value L = l1:l0, the
// 2^64 term taken by one multiply-add: T = L + hl * (2^32 - 1) (v_mad_u64_u32,
// 65 bits, carry kT), x = T + kT * eps - hh. T + kT eps cannot overflow (when
// kT is set, T < hl eps <= 2^64 - 2^33 + 1); a borrow of the subtraction of
// hh means x + 2^64 was computed, fixed by - eps (= + p - 2^64), which cannot
// borrow again; then one conditional subtraction of p. 13 VALU against 16.
// The carry of the multiply-add becomes a VGPR value inside the first asm
// statement, so no mask crosses statements.
__device__ __forceinline__ uint64_t gl_reduce_l(uint64_t L, uint32_t hl, uint32_t hh) {
    // (the second statement works in place on T's words and e, with one
    // temporary: the reduction sits inside register-heavy transforms)
    uint64_t T, k1;
    uint32_t e;
    asm("v_mad_u64_u32 %[T], %[k], %[hl], -1, %[L]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]"
        : [T] "=v"(T), [e] "=v"(e), [k] "=&s"(k1)
        : [hl] "v"(hl), [L] "v"(L));
    uint32_t t0 = lo32(T), t1 = hi32(T), z1;
    uint64_t k;
    asm("v_add_co_u32_e64 %[t0], %[k], %[t0], %[e]\n\t"
        "v_addc_co_u32_e64 %[t1], %[k], %[t1], 0, %[k]\n\t"
        "v_sub_co_u32_e64 %[t0], %[k], %[t0], %[hh]\n\t"
        "v_subb_co_u32_e64 %[t1], %[k], %[t1], 0, %[k]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]\n\t"
        "v_sub_co_u32_e64 %[t0], %[k], %[t0], %[e]\n\t"
        "v_subb_co_u32_e64 %[t1], %[k], %[t1], 0, %[k]\n\t"
        "v_add_co_u32_e64 %[e], %[k], %[t0], -1\n\t"
        "v_addc_co_u32_e64 %[z1], %[k], %[t1], 0, %[k]\n\t"
        "v_cndmask_b32_e64 %[t0], %[t0], %[e], %[k]\n\t"
        "v_cndmask_b32_e64 %[t1], %[t1], %[z1], %[k]"
        : [t0] "+v"(t0), [t1] "+v"(t1), [e] "+v"(e), [z1] "=&v"(z1), [k] "=&s"(k)
        : [hh] "v"(hh));
    return join32(t0, t1);
}

// L + 2^64 hl reduced mod p (canonical): gl_reduce_l with hh = 0, where
// T + kT eps is the value (no borrow step). 8 VALU.
__device__ __forceinline__ uint64_t gl_reduce_l0(uint64_t L, uint32_t hl) {
    uint64_t T, k1;
    uint32_t e;
    asm("v_mad_u64_u32 %[T], %[k], %[hl], -1, %[L]\n\t"
        "v_cndmask_b32_e64 %[e], 0, -1, %[k]"
        : [T] "=v"(T), [e] "=v"(e), [k] "=&s"(k1)
        : [hl] "v"(hl), [L] "v"(L));
    uint32_t t0 = lo32(T), t1 = hi32(T), z1;
    uint64_t k;
    asm("v_add_co_u32_e64 %[t0], %[k], %[t0], %[e]\n\t"
        "v_addc_co_u32_e64 %[t1], %[k], %[t1], 0, %[k]\n\t"
        "v_add_co_u32_e64 %[e], %[k], %[t0], -1\n\t"
        "v_addc_co_u32_e64 %[z1], %[k], %[t1], 0, %[k]\n\t"
        "v_cndmask_b32_e64 %[t0], %[t0], %[e], %[k]\n\t"
        "v_cndmask_b32_e64 %[t1], %[t1], %[z1], %[k]"
        : [t0] "+v"(t0), [t1] "+v"(t1), [e] "+v"(e), [z1] "=&v"(z1), [k] "=&s"(k));
    return join32(t0, t1);
}

// x * 2^s for 0 < s < 64: the 128-bit shift, then the reduction. Meant for
// compile-time s (the shifts and word selects fold away).
__device__ __forceinline__ uint64_t gl_mul_pow2(uint64_t x, int s) {
    const uint32_t x0 = lo32(x), x1 = hi32(x);
    if (s < 32) {
        return gl_reduce_l0(x << s, x1 >> (32 - s));
    } else if (s == 32) {
        return gl_reduce_l0(x << 32, x1);
    }
    return gl_reduce_l(x << s, __builtin_amdgcn_alignbit(x1, x0, 64 - s), x1 >> (64 - s));
}

// 64x64 -> 128 from four v_mad_u64_u32 (no partial sum can overflow 64
// bits), then the reduction: FM = 0 through the multiply-add (gl_reduce_l,
// fewer instructions), FM = 1 on the carry chain alone (gl_reduce, fewer
// live registers: for kernels that hold two transforms in registers).
template <int FM = 0>
__device__ __forceinline__ uint64_t gl_mulv(uint64_t a, uint64_t b) {
    const uint32_t al = lo32(a), ah = hi32(a), bl = lo32(b), bh = hi32(b);
    const uint64_t p00 = (uint64_t)al * bl;
    const uint64_t t = (uint64_t)ah * bl + hi32(p00);
    const uint64_t u = (uint64_t)al * bh + lo32(t);
    const uint64_t h = (uint64_t)ah * bh + hi32(t) + hi32(u);
    if constexpr (FM == 1) {
        return gl_reduce(lo32(p00), lo32(u), lo32(h), hi32(h));
    } else {
        return gl_reduce_l(join32(lo32(p00), lo32(u)), lo32(h), hi32(h));
    }
}

__device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b) { return gl_mulv<0>(a, b); }

}  // namespace midint
