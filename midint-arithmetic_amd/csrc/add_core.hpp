// Block-cooperative multi-limb addition on register-resident rows.
//
// Layout: a block owns a window of whole integers (segments of `seg` limbs).
// Wave w of the block holds rows w*R .. w*R+R-1 of the window, a row being 128
// consecutive limbs with lane l holding limbs 2l and 2l+1. This is the
// reference's Q-limbs-per-thread design (cuda/ker-addition.cu.h:141-194,
// cuda/ker-helpers.cu.h:121-170) with Q = 2 limbs per row and R rows per
// wave, arranged so that every global access is a contiguous 16-byte-per-lane
// (1 KiB per wave) transaction and no LDS transposition is needed.
#pragma once

#include "midint_dev.hpp"

namespace midint {

template <int R>
struct Rows {
    uint64_t x0[R], x1[R];  // first operand (and result)
    uint64_t y0[R], y1[R];  // second operand
    bool s0[R], s1[R];      // limb 0 / limb 1 starts a segment
    bool c0[R], c1[R];      // (out) carry into limb 0 / limb 1
};

// x <- x + y per segment (mod B^seg), carries resolved across the block.
// Every wave of the block must call this (it synchronises the block when the
// block has more than one wave).
template <int R>
__device__ __forceinline__ void block_add_rows(Rows<R>& t, uint32_t* lds_gp) {
    RowState st[R];
#pragma unroll
    for (int v = 0; v < R; v++) st[v] = row_build(t.x0[v], t.y0[v], t.x1[v], t.y1[v], t.s0[v], t.s1[v]);

    // wave aggregate: carry-out for cin = 0 and cin = 1
    uint32_t cz = 0, co = 1;
#pragma unroll
    for (int v = 0; v < R; v++) {
        (void)row_chain(st[v], &cz);
        (void)row_chain(st[v], &co);
    }
    uint32_t c = block_wave_carry_in(cz, co & ~cz, lds_gp);

    const uint32_t lane = lane_id();
#pragma unroll
    for (int v = 0; v < R; v++) {
        uint64_t C = row_chain(st[v], &c);
        row_apply(st[v], C, &t.x0[v], &t.x1[v]);
        t.c0[v] = (C >> lane) & 1ull;
        uint64_t Cl1 = ~st[v].S1 & (st[v].G0 | (st[v].P0 & C));
        t.c1[v] = (Cl1 >> lane) & 1ull;
    }
}

// x <- x + y + *c over one chunk of a longer integer: no segment starts, the
// carry *c enters the block's first limb and *c receives the carry out of its
// last limb. Every wave of the block must call this.
template <int R>
__device__ __forceinline__ void block_add_rows_carry(Rows<R>& t, uint32_t* lds_gp, uint32_t* c) {
    RowState st[R];
#pragma unroll
    for (int v = 0; v < R; v++) st[v] = row_build(t.x0[v], t.y0[v], t.x1[v], t.y1[v], false, false);
    uint32_t cz = 0, co = 1;
#pragma unroll
    for (int v = 0; v < R; v++) {
        (void)row_chain(st[v], &cz);
        (void)row_chain(st[v], &co);
    }
    uint32_t cw = block_wave_carry_chain(cz, co & ~cz, *c, lds_gp, c);
    const uint32_t lane = lane_id();
#pragma unroll
    for (int v = 0; v < R; v++) {
        uint64_t C = row_chain(st[v], &cw);
        row_apply(st[v], C, &t.x0[v], &t.x1[v]);
        t.c0[v] = (C >> lane) & 1ull;
        uint64_t Cl1 = st[v].G0 | (st[v].P0 & C);
        t.c1[v] = (Cl1 >> lane) & 1ull;
    }
}

// x <- x + y per segment with carries inside the wave only: for windows
// where every wave's R rows hold whole segments (no barrier, no exchange).
template <int R>
__device__ __forceinline__ void wave_add_rows(Rows<R>& t) {
    RowState st[R];
#pragma unroll
    for (int v = 0; v < R; v++) st[v] = row_build(t.x0[v], t.y0[v], t.x1[v], t.y1[v], t.s0[v], t.s1[v]);
    uint32_t c = 0;
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint64_t C = row_chain(st[v], &c);
        row_apply(st[v], C, &t.x0[v], &t.x1[v]);
    }
}

}  // namespace midint
