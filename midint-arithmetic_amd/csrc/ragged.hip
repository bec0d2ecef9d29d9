// Ragged batch of additions and multiplications with independent operand
// sizes, executed by ONE persistent kernel launch (inside a hipGraph).
//
// The reference runs one compile-time specialisation per size and one launch
// per size (runAdditions / runNaiveMuls, cuda/main.cu:661-709); a ragged
// batch there means padding every operand to the next compiled size. Here:
//
//  * the host planner groups the ops into work items: a multiplication
//    (u32-word lengths ns <= nl after ordering the operands) needs nl/Q
//    "virtual lanes" that all run ns/Q + 1 loop blocks (see mul_core.hpp);
//    small multiplications with the same or similar ns are packed into one
//    item (at most 512 lanes, 2048 output limbs), additions are packed into
//    items of at most 2048 output limbs; items are ordered longest first;
//  * each block of the persistent grid pops items from an atomic counter
//    until the list is exhausted (every wave leaves the loop on the same
//    popped index, so the grid drains);
//  * products are normalised and additions carried by the wave-ballot row
//    adder over the item's output window, segment boundaries taken from a
//    per-limb owner map in LDS.
//
// Results are exact: full products (na + nb limbs) and full sums
// (max(na, nb) + 1 limbs).
#include "add_core.hpp"
#include "midint_internal.h"
#include "mul_core.hpp"

namespace midint {

template <int THREADS>
struct RgCfg {
    static constexpr int kWaves = THREADS / kWave;
    static constexpr int kRows = kRgWindow / (kWaves * kRowLimbs);
    static constexpr int kRounds = kRgMaxLanes / THREADS;
};

// Zeroes the n work counters of a plan (one per persistent launch).
__global__ void ragged_reset(uint32_t* counter, uint32_t n) {
    if (blockIdx.x == 0 && threadIdx.x < n) counter[threadIdx.x] = 0u;
}

// Q: u32 columns per lane group of the products (4 or 8), the same for every
// item of the launch (the planner packs the ops of a launch with its Q).
template <int THREADS, int Q>
__global__ __launch_bounds__(THREADS) void ragged_kernel(const RgItem* __restrict__ items,
                                                         const RgOp* __restrict__ ops, uint32_t n_items,
                                                         uint32_t* __restrict__ counter,
                                                         const uint64_t* __restrict__ A,
                                                         const uint64_t* __restrict__ B, uint64_t* __restrict__ Rr,
                                                         uint32_t conv_cap) {
    using C = RgCfg<THREADS>;
    constexpr int R = C::kRows;
    extern __shared__ __align__(16) uint64_t lds64[];
    __shared__ uint32_t s_item;
    __shared__ uint32_t lds_gp[16];

    // LDS: [conv area / L,H window] [owner: window bytes] [lane_op: lanes bytes]; the
    // first column group of a lane is parked in registers (no snapshot area)
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds64);
    const uint32_t area_words = conv_cap > 4u * kRgWindow ? conv_cap : 4u * kRgWindow;
    uint8_t* owner = reinterpret_cast<uint8_t*>(lds + area_words);
    uint8_t* lane_op = owner + kRgWindow;
    uint64_t* Lb = lds64;
    uint64_t* Hb = lds64 + kRgWindow;

    const uint32_t tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        if (it >= n_items) break;
        const RgItem item = items[it];
        const RgOp* iops = ops + item.first;

        // ---- owner map of the output window (+ lane -> op for products)
        for (uint32_t k = 0; k < item.count; k++) {
            const RgOp& o = iops[k];
            for (uint32_t x = tid; x < o.seg_len; x += THREADS) owner[o.seg_base + x] = (uint8_t)k;
            if (item.kind == kRgMul)
                for (uint32_t x = tid; x < o.lanes; x += THREADS) lane_op[o.lane_base + x] = (uint8_t)k;
        }
        if (item.kind == kRgMul)
            for (uint32_t f = tid; f < item.conv_words; f += THREADS) lds[f] = 0u;
        __syncthreads();

        Rows<R> rows;
        const uint32_t k_first = (tid / kWave) * R * kRowLimbs + 2 * lane_id();
        const uint32_t win = item.window;

        if (item.kind == kRgMul) {
            // ---- stage operands as u32 words: S[i], W[(Q-1) + j]
            for (uint32_t k = 0; k < item.count; k++) {
                const RgOp& o = iops[k];
                uint32_t* S = lds + o.lds_base;
                uint32_t* W = S + o.ns_w + Q + (Q - 1);
                const uint64_t* gs = (o.swap ? B : A) + o.off_s;
                const uint64_t* gl = (o.swap ? A : B) + o.off_l;
                for (uint32_t x = tid; x < o.ns + o.nl; x += THREADS) {
                    if (x < o.ns) {
                        const uint64_t v = gs[x];
                        S[2 * x] = (uint32_t)v;
                        S[2 * x + 1] = (uint32_t)(v >> 32);
                    } else {
                        const uint32_t y = x - o.ns;
                        const uint64_t v = gl[y];
                        W[2 * y] = (uint32_t)v;
                        W[2 * y + 1] = (uint32_t)(v >> 32);
                    }
                }
            }
            __syncthreads();

            // ---- column-group convolution, up to kRounds virtual lanes per thread
            ColGroups<Q> g[C::kRounds];
            uint32_t gk1[C::kRounds], gk2[C::kRounds], gop[C::kRounds];
            bool gtwo[C::kRounds], gon[C::kRounds];
#pragma unroll
            for (int rr = 0; rr < C::kRounds; rr++) {
                const uint32_t v = tid + rr * THREADS;
                gon[rr] = v < item.lanes;
                if (gon[rr]) {
                    const uint32_t k = lane_op[v];
                    const RgOp& o = iops[k];
                    const uint32_t t = v - o.lane_base;
                    const uint32_t tp = o.ns_w / Q;
                    const uint32_t NB = tp + 1;
                    uint32_t k1, k2 = 0, i2 = 0, sw;
                    bool two;
                    if (t < tp) {
                        k1 = Q * t;
                        sw = t + 1;
                        k2 = o.nl_w + Q * t;
                        i2 = Q * t;
                        two = true;
                    } else {
                        k1 = o.ns_w + Q * (t - tp);
                        sw = NB;
                        two = false;
                    }
                    const uint32_t* S = lds + o.lds_base;
                    conv_lane<Q, false, false, false, true>(S, S + o.ns_w + Q, k1, k2, i2, sw, NB, nullptr, g[rr]);
                    gk1[rr] = k1;
                    gk2[rr] = k2;
                    gop[rr] = k;
                    gtwo[rr] = two;
                }
            }
            __syncthreads();  // conv area becomes the L/H window
#pragma unroll
            for (int rr = 0; rr < C::kRounds; rr++) {
                if (gon[rr]) {
                    const RgOp& o = iops[gop[rr]];
                    lane_store<Q>(g[rr], gtwo[rr], gk1[rr], gk2[rr], Lb + o.seg_base, Hb + o.seg_base, o.seg_len);
                }
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < R; v++) {
                const uint32_t k = k_first + v * kRowLimbs;
                rows.x0[v] = k < win ? Lb[k] : 0ull;
                rows.x1[v] = k + 1 < win ? Lb[k + 1] : 0ull;
                rows.y0[v] = k < win ? Hb[k] : 0ull;
                rows.y1[v] = k + 1 < win ? Hb[k + 1] : 0ull;
            }
        } else {
#pragma unroll
            for (int v = 0; v < R; v++) {
                const uint32_t k = k_first + v * kRowLimbs;
                uint64_t x[2] = {0, 0}, y[2] = {0, 0};
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (k + h < win) {
                        const RgOp& o = iops[owner[k + h]];
                        const uint32_t off = k + h - o.seg_base;
                        x[h] = off < o.ns ? A[o.off_s + off] : 0ull;
                        y[h] = off < o.nl ? B[o.off_l + off] : 0ull;
                    }
                }
                rows.x0[v] = x[0];
                rows.x1[v] = x[1];
                rows.y0[v] = y[0];
                rows.y1[v] = y[1];
            }
        }

        // ---- segment starts from the owner map, carry resolution, store
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            rows.s0[v] = k < win && (k == 0 || owner[k] != owner[k - 1]);
            rows.s1[v] = k + 1 < win && owner[k + 1] != owner[k];
        }
        block_add_rows<R>(rows, lds_gp);
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (k + h < win) {
                    const RgOp& o = iops[owner[k + h]];
                    const uint32_t off = k + h - o.seg_base;
                    if (off < o.out_len) Rr[o.off_r + off] = h ? rows.x1[v] : rows.x0[v];
                }
            }
        }
        __syncthreads();  // LDS is rewritten by the next item
    }
}

template <int THREADS>
static size_t ragged_lds_bytes(uint32_t conv_cap) {
    const uint32_t area_words = conv_cap > 4u * kRgWindow ? conv_cap : 4u * kRgWindow;
    return (size_t)area_words * 4 + kRgWindow + kRgMaxLanes;
}

size_t ragged_lds_size(uint32_t threads, uint32_t conv_cap) {
    return threads == 512 ? ragged_lds_bytes<512>(conv_cap) : ragged_lds_bytes<256>(conv_cap);
}

const void* ragged_kernel_ptr(uint32_t threads, uint32_t q) {
    if (q == 8) return threads == 512 ? (const void*)ragged_kernel<512, 8> : (const void*)ragged_kernel<256, 8>;
    return threads == 512 ? (const void*)ragged_kernel<512, 4> : (const void*)ragged_kernel<256, 4>;
}

const void* ragged_reset_ptr() { return (const void*)ragged_reset; }

}  // namespace midint
