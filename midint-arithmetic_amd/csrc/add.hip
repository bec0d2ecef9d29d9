// Batched big-integer addition kernels for gfx950.
//
// add_kernel  : r = (a + b) mod B^m per instance    (reference baddKer,
//               cuda/ker-addition.cu.h:227-260)
// add6_kernel : r = (6a + 10b) mod B^m per instance (reference a6pb10Ker,
//               cuda/ker-addition.cu.h:266-328), the value of its chain of
//               six additions in one register-resident addition.
//
// A block owns `ipb` whole instances (a window of ipb*m limbs); the window is
// split into rows of 128 limbs, R rows per wave. Loads and stores are 16-byte
// per lane and contiguous per wave when the window is 16-byte aligned (VEC),
// 8-byte otherwise.
#include "add_core.hpp"
#include "midint_internal.h"

namespace midint {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// NT: streaming (non-temporal) 16-byte accesses for data touched once.
template <bool VEC, bool NT = false>
__device__ __forceinline__ void load_pair(const uint64_t* __restrict__ p, uint32_t k, uint32_t L,
                                          uint64_t& v0, uint64_t& v1) {
    if (VEC && k + 1 < L) {
        u64x2 t;
        if (NT)
            t = __builtin_nontemporal_load(reinterpret_cast<const u64x2*>(p + k));
        else
            t = *reinterpret_cast<const u64x2*>(p + k);
        v0 = t.x;
        v1 = t.y;
    } else {
        v0 = k < L ? p[k] : 0ull;
        v1 = k + 1 < L ? p[k + 1] : 0ull;
    }
}

template <bool VEC, bool NT = false>
__device__ __forceinline__ void store_pair(uint64_t* __restrict__ p, uint32_t k, uint32_t L, uint64_t v0,
                                           uint64_t v1) {
    if (VEC && k + 1 < L) {
        u64x2 t;
        t.x = v0;
        t.y = v1;
        if (NT)
            __builtin_nontemporal_store(t, reinterpret_cast<u64x2*>(p + k));
        else
            *reinterpret_cast<u64x2*>(p + k) = t;
    } else {
        if (k < L) p[k] = v0;
        if (k + 1 < L) p[k + 1] = v1;
    }
}

// Limb k of a strided window, given its (instance, offset) position.
template <bool NT>
__device__ __forceinline__ uint64_t load_strided(const uint64_t* __restrict__ p, uint32_t k, uint32_t L,
                                                 uint32_t inst, uint32_t off, uint64_t stride) {
    if (k >= L) return 0ull;
    const uint64_t* q = p + (uint64_t)inst * stride + off;
    return NT ? __builtin_nontemporal_load(q) : *q;
}

// Position of the limb after (inst, off) in instances of m limbs.
__device__ __forceinline__ void next_pos(uint32_t inst, uint32_t off, uint32_t m, uint32_t& inst1, uint32_t& off1) {
    const bool wrap = off + 1 == m;
    off1 = wrap ? 0u : off + 1;
    inst1 = inst + (wrap ? 1u : 0u);
}

// Loads both operands of the block window into rows and marks segment starts.
// STRIDED: instance i of a (b) starts at limb i * sa (i * sb) instead of i * m.
template <int R, bool VEC, bool NT = false, bool STRIDED = false>
__device__ __forceinline__ void load_window(Rows<R>& t, const uint64_t* __restrict__ a,
                                            const uint64_t* __restrict__ b, uint32_t L, uint32_t m,
                                            uint32_t step_q, uint32_t step_r, uint32_t* off0,
                                            uint32_t* inst0, uint64_t sa = 0, uint64_t sb = 0) {
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t k_first = wave * R * kRowLimbs + 2 * lane_id();
    SegPos sp = segpos_init(k_first, m);
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        if (STRIDED) {
            uint32_t i1, o1;
            next_pos(sp.inst, sp.off, m, i1, o1);
            t.x0[v] = load_strided<NT>(a, k, L, sp.inst, sp.off, sa);
            t.x1[v] = load_strided<NT>(a, k + 1, L, i1, o1, sa);
            t.y0[v] = load_strided<NT>(b, k, L, sp.inst, sp.off, sb);
            t.y1[v] = load_strided<NT>(b, k + 1, L, i1, o1, sb);
        } else {
            load_pair<VEC, NT>(a, k, L, t.x0[v], t.x1[v]);
            load_pair<VEC, NT>(b, k, L, t.y0[v], t.y1[v]);
        }
        t.s0[v] = (sp.off == 0) && (k < L);
        t.s1[v] = (sp.off + 1 == m) && (k + 1 < L);
        off0[v] = sp.off;
        inst0[v] = sp.inst;
        segpos_next(&sp, m, step_q, step_r);
    }
}

// Stores the rows of a window (strided or batch-major).
template <int R, bool VEC, bool NT, bool STRIDED>
__device__ __forceinline__ void store_window(uint64_t* __restrict__ r, const Rows<R>& t, uint32_t L, uint32_t m,
                                             const uint32_t* off0, const uint32_t* inst0, uint64_t sr) {
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t k = k_first + v * kRowLimbs;
        if (STRIDED) {
            uint32_t i1, o1;
            next_pos(inst0[v], off0[v], m, i1, o1);
            if (k < L) r[(uint64_t)inst0[v] * sr + off0[v]] = t.x0[v];
            if (k + 1 < L) r[(uint64_t)i1 * sr + o1] = t.x1[v];
        } else {
            store_pair<VEC, NT>(r, k, L, t.x0[v], t.x1[v]);
        }
    }
}

template <int R, bool VEC, bool NT, bool STRIDED>
__global__ __launch_bounds__(1024) void add_kernel(uint32_t n_inst, uint32_t m, uint32_t ipb, uint32_t step_q,
                                                   uint32_t step_r, const uint64_t* __restrict__ a,
                                                   const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                   uint8_t* __restrict__ carry, uint64_t sa, uint64_t sb,
                                                   uint64_t sr) {
    __shared__ uint32_t lds_gp[16];
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);
    const uint32_t L = ninst * m;
    const uint64_t base = inst_base * m;

    Rows<R> t;
    uint32_t off0[R], inst0[R];
    if (STRIDED)
        load_window<R, false, NT, true>(t, a + inst_base * sa, b + inst_base * sb, L, m, step_q, step_r, off0,
                                        inst0, sa, sb);
    else
        load_window<R, VEC, NT>(t, a + base, b + base, L, m, step_q, step_r, off0, inst0);

    uint64_t ax0[R], ax1[R];
    if (carry) {
#pragma unroll
        for (int v = 0; v < R; v++) {
            ax0[v] = t.x0[v];
            ax1[v] = t.x1[v];
        }
    }

    block_add_rows<R>(t, lds_gp);

    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
    store_window<R, VEC, NT, STRIDED>(STRIDED ? r + inst_base * sr : r + base, t, L, m, off0, inst0, sr);

    if (carry) {
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = k_first + v * kRowLimbs;
            if (k < L && off0[v] == m - 1) {
                uint64_t pre = t.x0[v] - (uint64_t)t.c0[v];
                carry[inst_base + inst0[v]] = (uint8_t)limb_cout(ax0[v], pre, t.c0[v]);
            }
            const uint32_t off1 = (off0[v] + 1 == m) ? 0u : off0[v] + 1;
            const uint32_t i1 = inst0[v] + ((off0[v] + 1 == m) ? 1u : 0u);
            if (k + 1 < L && off1 == m - 1) {
                uint64_t pre = t.x1[v] - (uint64_t)t.c1[v];
                carry[inst_base + i1] = (uint8_t)limb_cout(ax1[v], pre, t.c1[v]);
            }
        }
    }
}

// 6a + 10b of one limb pair: the low 64 bits, and the high part (< 16) in hi.
__device__ __forceinline__ uint64_t a6b10(uint64_t a, uint64_t b, uint32_t& hi) {
    const unsigned __int128 s = (unsigned __int128)a * 6u + (unsigned __int128)b * 10u;
    hi = (uint32_t)(s >> 64);
    return (uint64_t)s;
}

// r = (6a + 10b) mod B^m in one carry-resolved addition. The reference's
// chain of six additions, each mod B^m, ends at the same value, since the
// reduction mod B^m commutes with every step. Per limb 6a_i + 10b_i = lo_i
// + hi_i B (hi_i < 16), so 6a + 10b = L + H with L_i = lo_i and H_i =
// hi_{i-1}: the high part moves one limb up (lane l's first limb takes lane
// l - 1's, through a wave_shr DPP move; lane 0 takes the previous row's lane
// 63, or, for a wave's first row, the limb before it, read again from
// memory), and is dropped at every instance start.
template <int R, bool VEC, bool STRIDED, bool NT = false>
__global__ __launch_bounds__(1024) void add6_kernel(uint32_t n_inst, uint32_t m, uint32_t ipb, uint32_t step_q,
                                                    uint32_t step_r, const uint64_t* __restrict__ a,
                                                    const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                    uint64_t sa, uint64_t sb, uint64_t sr) {
    __shared__ uint32_t lds_gp[16];
    const uint64_t inst_base = (uint64_t)blockIdx.x * ipb;
    const uint32_t ninst = (uint32_t)min((uint64_t)ipb, (uint64_t)n_inst - inst_base);
    const uint32_t L = ninst * m;
    const uint64_t base = inst_base * m;

    Rows<R> t;
    uint32_t off0[R], inst0[R];
    if (STRIDED)
        load_window<R, false, false, true>(t, a + inst_base * sa, b + inst_base * sb, L, m, step_q, step_r, off0,
                                           inst0, sa, sb);
    else
        load_window<R, VEC, NT>(t, a + base, b + base, L, m, step_q, step_r, off0, inst0);
    const uint32_t lane = lane_id();
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane;
    uint32_t h0[R], h1[R];
#pragma unroll
    for (int v = 0; v < R; v++) {
        t.x0[v] = a6b10(t.x0[v], t.y0[v], h0[v]);
        t.x1[v] = a6b10(t.x1[v], t.y1[v], h1[v]);
    }
    // the high part of the limb before the wave's first one (in the same
    // instance: off0[0] > 0, so k_first > 0)
    uint32_t before = 0;
    if (lane == 0 && k_first < L && off0[0] != 0) {
        const uint64_t* pa = STRIDED ? a + (inst_base + inst0[0]) * sa + off0[0] - 1 : a + base + k_first - 1;
        const uint64_t* pb = STRIDED ? b + (inst_base + inst0[0]) * sb + off0[0] - 1 : b + base + k_first - 1;
        (void)a6b10(*pa, *pb, before);
    }
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint32_t in0 = (uint32_t)__builtin_amdgcn_update_dpp((int)before, (int)h1[v], 0x138, 0xf, 0xf, false);
        t.y0[v] = t.s0[v] ? 0ull : (uint64_t)in0;
        t.y1[v] = t.s1[v] ? 0ull : (uint64_t)h0[v];
        before = (uint32_t)__builtin_amdgcn_readlane((int)h1[v], 63);
    }
    block_add_rows<R>(t, lds_gp);
    store_window<R, VEC, NT, STRIDED>(STRIDED ? r + inst_base * sr : r + base, t, L, m, off0, inst0, sr);
}

// Instances longer than one workgroup's rows (m > kAddRowMax limbs): one
// instance per workgroup, walked in chunks of (waves * R * 128) limbs. Each
// addition of the chain keeps its carry between chunks in a scalar, so any
// length runs with the registers of R rows. SIX: the value of the
// reference's six-add chain, (6a + 10b) mod B^m, one addition per chunk.
template <int R, bool VEC, bool SIX>
__global__ __launch_bounds__(1024) void add_long_kernel(uint32_t m, const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b, uint64_t* __restrict__ r,
                                                        uint8_t* __restrict__ carry, uint64_t sa, uint64_t sb,
                                                        uint64_t sr) {
    __shared__ uint32_t lds_gp[16];
    const uint64_t inst = blockIdx.x;
    const uint64_t* pa = a + inst * sa;
    const uint64_t* pb = b + inst * sb;
    uint64_t* pr = r + inst * sr;
    const uint32_t chunk = (blockDim.x / kWave) * R * kRowLimbs;
    const uint32_t k_first = (threadIdx.x / kWave) * R * kRowLimbs + 2 * lane_id();
    __shared__ uint32_t lds_hi[2];
    uint32_t c[1] = {0u};
    for (uint32_t base = 0; base < m; base += chunk) {
        Rows<R> t;
#pragma unroll
        for (int v = 0; v < R; v++) {
            const uint32_t k = base + k_first + v * kRowLimbs;
            load_pair<VEC, true>(pa, k, m, t.x0[v], t.x1[v]);
            load_pair<VEC, true>(pb, k, m, t.y0[v], t.y1[v]);
        }
        if (!SIX) {
            uint64_t ax0[R], ax1[R];
#pragma unroll
            for (int v = 0; v < R; v++) {
                ax0[v] = t.x0[v];
                ax1[v] = t.x1[v];
            }
            block_add_rows_carry<R>(t, lds_gp, &c[0]);
#pragma unroll
            for (int v = 0; v < R; v++) {
                const uint32_t k = base + k_first + v * kRowLimbs;
                store_pair<VEC, true>(pr, k, m, t.x0[v], t.x1[v]);
                if (carry && k == m - 1) carry[inst] = (uint8_t)limb_cout(ax0[v], t.x0[v] - t.c0[v], t.c0[v]);
                if (carry && k + 1 == m - 1)
                    carry[inst] = (uint8_t)limb_cout(ax1[v], t.x1[v] - t.c1[v], t.c1[v]);
            }
        } else {
            // (6a + 10b) mod B^m in one addition per chunk (as add6_kernel):
            // each limb's high part moves one limb up. A wave's first limb
            // takes the high part of the limb before it, read again from
            // memory (same chunk: not yet stored), except in wave 0, where
            // that limb is the previous chunk's last, already stored (and
            // overwritten when r = a): the block's last wave hands its high
            // part over in an LDS slot per chunk parity (written before this
            // chunk's barrier, read after it, rewritten only after the next)
            const uint32_t lane = lane_id();
            const uint32_t wave = threadIdx.x / kWave;
            const uint32_t par = (base / chunk) & 1u;
            uint32_t h0[R], h1[R];
#pragma unroll
            for (int v = 0; v < R; v++) {
                t.x0[v] = a6b10(t.x0[v], t.y0[v], h0[v]);
                t.x1[v] = a6b10(t.x1[v], t.y1[v], h1[v]);
            }
            uint32_t before = 0;
            if (wave == 0) {
                before = base == 0 ? 0u : lds_hi[par ^ 1u];
            } else if (lane == 0 && base + k_first < m) {
                (void)a6b10(pa[base + k_first - 1], pb[base + k_first - 1], before);
            }
#pragma unroll
            for (int v = 0; v < R; v++) {
                const uint32_t in0 =
                    (uint32_t)__builtin_amdgcn_update_dpp((int)before, (int)h1[v], 0x138, 0xf, 0xf, false);
                t.y0[v] = (uint64_t)in0;
                t.y1[v] = (uint64_t)h0[v];
                before = (uint32_t)__builtin_amdgcn_readlane((int)h1[v], 63);
            }
            if (wave == blockDim.x / kWave - 1 && lane == 0) lds_hi[par] = before;
            block_add_rows_carry<R>(t, lds_gp, &c[0]);
#pragma unroll
            for (int v = 0; v < R; v++)
                store_pair<VEC, true>(pr, base + k_first + v * kRowLimbs, m, t.x0[v], t.x1[v]);
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launch geometry.

// Chooses rows per wave R and waves per block so that one block holds at
// least one whole instance; small instances are packed `ipb` per block.
AddGeom add_geometry(uint32_t m) {
    AddGeom g;
    if (m <= 256) {
        g.rows = (uint32_t)tune_get_class("add.rows", m, 2);
        g.threads = (uint32_t)tune_get_class("add.threads", m, 256);
        if ((g.rows != 1 && g.rows != 2 && g.rows != 4 && g.rows != 8) || g.threads < 64 || g.threads > 1024 ||
            g.threads % 64 != 0 || (g.threads / kWave) * g.rows * kRowLimbs < m) {
            g.rows = 2;
            g.threads = 256;
        }
    } else if (m <= kAddRowMax) {
        const uint32_t waves = (m + 2 * kRowLimbs - 1) / (2 * kRowLimbs);
        g.rows = 2;
        g.threads = (waves < 4 ? 4 : waves) * kWave;
    } else {
        g.rows = 0;  // add_long_kernel
        g.threads = 0;
        g.ipb = 0;
        return g;
    }
    uint32_t cap = (g.threads / kWave) * g.rows * kRowLimbs;
    g.ipb = cap / m;
    if ((m & 1u) && g.ipb > 1) g.ipb &= ~1u;  // keep windows 16-byte aligned
    return g;
}

template <int R>
static hipError_t launch_add_r(const AddGeom& g, bool vec, bool six, uint32_t n, uint32_t m, const uint64_t* a,
                               const uint64_t* b, uint64_t* r, uint8_t* carry, const Strides& st, hipStream_t s) {
    const bool nt = tune_get_class("add.nt", m, 1) != 0;
    const bool strided = st.a || st.b || st.r;
    const uint64_t sa = st.a ? st.a : m, sb = st.b ? st.b : m, sr_ = st.r ? st.r : m;
    const uint32_t blocks = (uint32_t)(((uint64_t)n + g.ipb - 1) / g.ipb);
    const uint32_t sq = kRowLimbs / m, sr = kRowLimbs % m;
    if (six) {
        if (strided)
            add6_kernel<R, false, true><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, sa, sb, sr_);
        else if (vec && nt)
            add6_kernel<R, true, false, true><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, sa, sb, sr_);
        else if (vec)
            add6_kernel<R, true, false><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, sa, sb, sr_);
        else
            add6_kernel<R, false, false><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, sa, sb, sr_);
    } else {
        if (strided)
            add_kernel<R, false, true, true><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, carry, sa,
                                                                          sb, sr_);
        else if (vec && nt)
            add_kernel<R, true, true, false><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, carry, sa,
                                                                          sb, sr_);
        else if (vec)
            add_kernel<R, true, false, false><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, carry, sa,
                                                                           sb, sr_);
        else
            add_kernel<R, false, false, false><<<blocks, g.threads, 0, s>>>(n, m, g.ipb, sq, sr, a, b, r, carry,
                                                                            sa, sb, sr_);
    }
    return hipGetLastError();
}

// One instance per workgroup of 16 waves, rows of 4 x 128 limbs per wave.
static hipError_t launch_add_long(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                                  uint8_t* carry, bool six, hipStream_t s, const Strides& st) {
    const uint64_t sa = st.a ? st.a : m, sb = st.b ? st.b : m, sr = st.r ? st.r : m;
    const bool vec = ((((uintptr_t)a) | ((uintptr_t)b) | ((uintptr_t)r)) & 15u) == 0 && ((sa | sb | sr) & 1u) == 0;
    constexpr uint32_t kThreads = 1024;
    if (six) {
        if (vec)
            add_long_kernel<4, true, true><<<n, kThreads, 0, s>>>(m, a, b, r, nullptr, sa, sb, sr);
        else
            add_long_kernel<4, false, true><<<n, kThreads, 0, s>>>(m, a, b, r, nullptr, sa, sb, sr);
    } else {
        if (vec)
            add_long_kernel<4, true, false><<<n, kThreads, 0, s>>>(m, a, b, r, carry, sa, sb, sr);
        else
            add_long_kernel<4, false, false><<<n, kThreads, 0, s>>>(m, a, b, r, carry, sa, sb, sr);
    }
    return hipGetLastError();
}

hipError_t launch_add(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, uint8_t* carry,
                      bool six, hipStream_t s, Strides st) {
    if (m > kAddRowMax) return launch_add_long(n, m, a, b, r, carry, six, s, st);
    AddGeom g = add_geometry(m);
    if (g.ipb == 0) return hipErrorInvalidValue;
    const bool aligned = ((((uintptr_t)a) | ((uintptr_t)b) | ((uintptr_t)r)) & 15u) == 0;
    const bool vec = aligned && ((uint64_t)g.ipb * m) % 2 == 0;
    switch (g.rows) {
        case 1: return launch_add_r<1>(g, vec, six, n, m, a, b, r, carry, st, s);
        case 2: return launch_add_r<2>(g, vec, six, n, m, a, b, r, carry, st, s);
        case 4: return launch_add_r<4>(g, vec, six, n, m, a, b, r, carry, st, s);
        case 8: return launch_add_r<8>(g, vec, six, n, m, a, b, r, carry, st, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace midint
