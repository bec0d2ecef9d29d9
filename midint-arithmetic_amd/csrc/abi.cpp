// C ABI of libmidint (declared in include/midint.h): argument checking,
// per-size dispatch to the kernels, the context (twiddle cache, scratch) and
// the host-pointer drop-ins that mirror the reference's gpuAdd / gpuMultiply /
// gpuMulFFT drivers (cuda/main.cu:39-579).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <vector>
#include <mutex>

#include "midint.h"
#include "midint_internal.h"

using namespace midint;

struct midint_ctx {
    int device;
    std::mutex mu;
    std::map<uint32_t, NttTables> ntt;  // keyed by log2 length
    std::map<uint32_t, PrimeTables> prime_ntt;  // keyed by (prime, lg, root kind)
    // Context scratch (poly / poly_ntt temporaries, u32 repacking). Its last
    // user recorded scratch_ev on its stream; the next user's stream waits on
    // it, so calls on different streams take turns on the device.
    uint64_t* scratch = nullptr;
    size_t scratch_limbs = 0;
    hipEvent_t scratch_ev = nullptr;
    bool scratch_ev_live = false;
};

static thread_local int g_last_hip = 0;

static midint_status hip_status(hipError_t e) {
    if (e == hipSuccess) return MIDINT_OK;
    g_last_hip = (int)e;
    return MIDINT_EHIP;
}

static midint_status bind_device(const midint_ctx* ctx) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return hip_status(hipGetLastError());
    if (cur != ctx->device) return hip_status(hipSetDevice(ctx->device));
    return MIDINT_OK;
}

static bool aligned8(const void* p) { return (((uintptr_t)p) & 7u) == 0; }

// Makes at least `limbs` of context scratch available to work enqueued next
// on `stream`: the stream first waits for the scratch's previous user. Call
// with ctx->mu held, then enqueue, then release_scratch.
static midint_status acquire_scratch(midint_ctx* ctx, size_t limbs, hipStream_t stream) {
    hipError_t e;
    if (!ctx->scratch_ev) {
        if ((e = hipEventCreateWithFlags(&ctx->scratch_ev, hipEventDisableTiming)) != hipSuccess) {
            ctx->scratch_ev = nullptr;
            return hip_status(e);
        }
    }
    if (ctx->scratch_limbs < limbs) {
        if (ctx->scratch) {
            // the old buffer may still be in use on another stream
            if (ctx->scratch_ev_live && (e = hipEventSynchronize(ctx->scratch_ev)) != hipSuccess) return hip_status(e);
            e = hipFree(ctx->scratch);
            ctx->scratch = nullptr;
            ctx->scratch_limbs = 0;
            if (e != hipSuccess) return hip_status(e);
        }
        e = hipMalloc((void**)&ctx->scratch, limbs * sizeof(uint64_t));
        if (e != hipSuccess) {
            ctx->scratch = nullptr;
            (void)hipGetLastError();
            return MIDINT_ENOMEM;
        }
        ctx->scratch_limbs = limbs;
    } else if (ctx->scratch_ev_live && (e = hipStreamWaitEvent(stream, ctx->scratch_ev, 0)) != hipSuccess) {
        return hip_status(e);
    }
    return MIDINT_OK;
}

// Marks the end of the work just enqueued on `stream` that uses the scratch.
static midint_status release_scratch(midint_ctx* ctx, hipStream_t stream) {
    hipError_t e = hipEventRecord(ctx->scratch_ev, stream);
    if (e != hipSuccess) return hip_status(e);
    ctx->scratch_ev_live = true;
    return MIDINT_OK;
}

// Runs `enqueue(scratch)` on `stream` with `limbs` of context scratch.
template <class F>
static midint_status with_scratch(midint_ctx* ctx, size_t limbs, hipStream_t stream, F enqueue) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    midint_status st = acquire_scratch(ctx, limbs, stream);
    if (st != MIDINT_OK) return st;
    st = hip_status(enqueue(ctx->scratch));
    if (st != MIDINT_OK) return st;
    return release_scratch(ctx, stream);
}

extern "C" {

const char* midint_strerror(midint_status s) {
    switch (s) {
        case MIDINT_OK: return "success";
        case MIDINT_EINVAL: return "invalid argument";
        case MIDINT_EUNSUPPORTED: return "operand size not supported";
        case MIDINT_EHIP: return "HIP runtime error";
        case MIDINT_ENOMEM: return "out of memory";
        case MIDINT_EBUSY: return "resource busy";
        case MIDINT_ENOTFOUND: return "not found";
        case MIDINT_ETOOBIG: return "request exceeds the fixed-size static scratch";
    }
    return "unknown status";
}

const char* midint_version(void) { return "midint-mi355x 0.1 (gfx950)"; }

int midint_last_hip_error(void) { return g_last_hip; }

midint_status midint_init(midint_ctx** out, int device) {
    if (!out) return MIDINT_EINVAL;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_status(e);
    if (device < 0 || device >= count) return MIDINT_EINVAL;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_status(e);
    midint_ctx* c = new (std::nothrow) midint_ctx;
    if (!c) return MIDINT_ENOMEM;
    c->device = device;
    // device address of the no-allocation entry point's static scratch, looked
    // up once per device with the device current; a failure is kept and
    // reported by that entry point only
    (void)noalloc_setup(device);
    *out = c;
    return MIDINT_OK;
}

midint_status midint_destroy(midint_ctx* ctx) {
    if (!ctx) return MIDINT_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto& kv : ctx->ntt) {
        (void)hipFree(kv.second.d_fwd);
        (void)hipFree(kv.second.d_inv);
    }
    for (auto& kv : ctx->prime_ntt) {
        (void)hipFree(kv.second.d_fwd);
        (void)hipFree(kv.second.d_inv);
    }
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->scratch_ev) (void)hipEventDestroy(ctx->scratch_ev);
    delete ctx;
    return MIDINT_OK;
}

uint32_t midint_max_limbs_add(void) { return kAddMaxLimbs; }
uint32_t midint_max_limbs_mul(void) { return 2048u; }
uint32_t midint_max_limbs_ntt(void) { return (1u << kNttMaxLg) / 8u; }

static midint_status check_binary(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b,
                                  const uint64_t* r) {
    if (!ctx || m == 0) return MIDINT_EINVAL;
    if (n == 0) return MIDINT_OK;
    if (!a || !b || !r) return MIDINT_EINVAL;
    if (!aligned8(a) || !aligned8(b) || !aligned8(r)) return MIDINT_EINVAL;
    return MIDINT_OK;
}

midint_status midint_add(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                         uint64_t* d_r, uint8_t* d_carry, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_add()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    return hip_status(launch_add(n, m, d_a, d_b, d_r, d_carry, false, (hipStream_t)stream));
}

midint_status midint_add6(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                          uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_add()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    return hip_status(launch_add(n, m, d_a, d_b, d_r, nullptr, true, (hipStream_t)stream));
}

midint_status midint_mul_lo(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                            uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_mul()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    return hip_status(launch_mul(n, m, d_a, d_b, d_r, false, (hipStream_t)stream));
}

midint_status midint_mul_full(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a,
                              const uint64_t* d_b, uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_mul()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    return hip_status(launch_mul(n, m, d_a, d_b, d_r, true, (hipStream_t)stream));
}

midint_status midint_poly(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                          uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_mul()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (poly_uses_fused(m)) return hip_status(launch_poly(n, m, d_a, d_b, d_r, nullptr, s));
    return with_scratch(ctx, 4ull * n * m, s,
                        [&](uint64_t* scr) { return launch_poly(n, m, d_a, d_b, d_r, scr, s); });
}

// Returns the (cached) twiddle tables of length 2^lg.
static midint_status get_tables(midint_ctx* ctx, uint32_t lg, NttTables* out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->ntt.find(lg);
    if (it != ctx->ntt.end()) {
        *out = it->second;
        return MIDINT_OK;
    }
    const size_t len = (size_t)1 << lg;  // w^i and w^-i for i < L
    uint64_t* hf = (uint64_t*)malloc(len * sizeof(uint64_t));
    uint64_t* hi = (uint64_t*)malloc(len * sizeof(uint64_t));
    if (!hf || !hi) {
        free(hf);
        free(hi);
        return MIDINT_ENOMEM;
    }
    NttTables t;
    t.lg = lg;
    ntt_host_tables(lg, hf, hi, &t.inv_len);
    hipError_t e = hipMalloc((void**)&t.d_fwd, len * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&t.d_inv, len * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(t.d_fwd, hf, len * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.d_inv, hi, len * sizeof(uint64_t), hipMemcpyHostToDevice);
    free(hf);
    free(hi);
    if (e != hipSuccess) return hip_status(e);
    ctx->ntt[lg] = t;
    *out = t;
    return MIDINT_OK;
}

}  // extern "C"

midint_status midint::ctx_ntt_tables(midint_ctx* ctx, uint32_t lg, NttTables* out) { return get_tables(ctx, lg, out); }

extern "C" {

midint_status midint_ntt_prepare(midint_ctx* ctx, uint32_t m) {
    if (!ctx || m == 0) return MIDINT_EINVAL;
    if (m > midint_max_limbs_ntt()) return MIDINT_EUNSUPPORTED;
    midint_status st = bind_device(ctx);
    if (st != MIDINT_OK) return st;
    NttTables t;
    return get_tables(ctx, ntt_lg_for_limbs(m), &t);
}

midint_status midint_mul_ntt(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a,
                             const uint64_t* d_b, uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_ntt()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    NttTables t;
    if ((st = get_tables(ctx, ntt_lg_for_limbs(m), &t)) != MIDINT_OK) return st;
    return hip_status(launch_mul_ntt(n, m, d_a, d_b, d_r, t, (hipStream_t)stream));
}

midint_status midint_ntt_conv(midint_ctx* ctx, uint32_t n, uint32_t lg, const uint64_t* d_a,
                              const uint64_t* d_b, uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, 1, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (lg < 6 || lg > kNttMaxLg) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    NttTables t;
    if ((st = get_tables(ctx, lg, &t)) != MIDINT_OK) return st;
    return hip_status(launch_ntt_conv(n, d_a, d_b, d_r, t, (hipStream_t)stream));
}

midint_status midint_poly_ntt(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                              uint64_t* d_r, void* stream) {
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    if (m > midint_max_limbs_ntt()) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    NttTables t;
    if ((st = get_tables(ctx, poly_ntt_lg(m), &t)) != MIDINT_OK) return st;
    return hip_status(launch_poly_ntt(n, m, d_a, d_b, d_r, t, (hipStream_t)stream));
}

// ------------------------------------------------- strided u64 layouts
enum class Op { Add, Add6, MulLo, MulFull, Poly };

static uint32_t op_out_limbs(Op op, uint32_t m) { return op == Op::MulFull ? 2 * m : m; }

// u64 operation on instances at limb strides sa, sb, sr (0 = batch-major).
static midint_status u64_op(midint_ctx* ctx, Op op, uint32_t n, uint32_t m, const uint64_t* a, uint64_t sa,
                            const uint64_t* b, uint64_t sb, uint64_t* r, uint64_t sr, uint8_t* carry,
                            void* stream) {
    midint_status st = check_binary(ctx, n, m, a, b, r);
    if (st != MIDINT_OK || n == 0) return st;
    const bool is_add = op == Op::Add || op == Op::Add6;
    if (m > (is_add ? midint_max_limbs_add() : midint_max_limbs_mul())) return MIDINT_EUNSUPPORTED;
    if ((sa && sa < m) || (sb && sb < m) || (sr && sr < op_out_limbs(op, m))) return MIDINT_EINVAL;
    // the batch-major strides keep the 16-byte vector path
    if (sa == m) sa = 0;
    if (sb == m) sb = 0;
    if (sr == op_out_limbs(op, m)) sr = 0;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const Strides str{sa, sb, sr};
    switch (op) {
        case Op::Add: return hip_status(launch_add(n, m, a, b, r, carry, false, s, str));
        case Op::Add6: return hip_status(launch_add(n, m, a, b, r, nullptr, true, s, str));
        case Op::MulLo: return hip_status(launch_mul(n, m, a, b, r, false, s, str));
        case Op::MulFull: return hip_status(launch_mul(n, m, a, b, r, true, s, str));
        case Op::Poly:
            if (poly_uses_fused(m)) return hip_status(launch_poly(n, m, a, b, r, nullptr, s, str));
            return with_scratch(ctx, 4ull * n * m, s,
                                [&](uint64_t* scr) { return launch_poly(n, m, a, b, r, scr, s, str); });
    }
    return MIDINT_EINVAL;
}

midint_status midint_add_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                 const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, uint8_t* d_carry,
                                 void* stream) {
    return u64_op(ctx, Op::Add, n, m, d_a, sa, d_b, sb, d_r, sr, d_carry, stream);
}
midint_status midint_add6_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                  const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream) {
    return u64_op(ctx, Op::Add6, n, m, d_a, sa, d_b, sb, d_r, sr, nullptr, stream);
}
midint_status midint_mul_lo_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                    const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream) {
    return u64_op(ctx, Op::MulLo, n, m, d_a, sa, d_b, sb, d_r, sr, nullptr, stream);
}
midint_status midint_mul_full_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                      const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr,
                                      void* stream) {
    return u64_op(ctx, Op::MulFull, n, m, d_a, sa, d_b, sb, d_r, sr, nullptr, stream);
}
midint_status midint_poly_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                  const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream) {
    return u64_op(ctx, Op::Poly, n, m, d_a, sa, d_b, sb, d_r, sr, nullptr, stream);
}

// ------------------------------------------------------------ u32 limbs
// u32 operation on instances at word strides sa, sb, sr (0 = batch-major).
static midint_status u32_op(midint_ctx* ctx, Op op, uint32_t n, uint32_t m32, const uint32_t* a, uint64_t sa,
                            const uint32_t* b, uint64_t sb, uint32_t* r, uint64_t sr, uint8_t* carry,
                            void* stream) {
    if (!ctx || m32 == 0) return MIDINT_EINVAL;
    if (n == 0) return MIDINT_OK;
    if (!a || !b || !r || ((((uintptr_t)a) | ((uintptr_t)b) | ((uintptr_t)r)) & 3u)) return MIDINT_EINVAL;
    const uint32_t m = (m32 + 1) / 2;
    const bool is_add = op == Op::Add || op == Op::Add6;
    if (m > (is_add ? midint_max_limbs_add() : midint_max_limbs_mul())) return MIDINT_EUNSUPPORTED;
    const uint32_t out_w32 = op == Op::MulFull ? 2 * m32 : m32;
    sa = sa ? sa : m32;
    sb = sb ? sb : m32;
    sr = sr ? sr : out_w32;
    if (sa < m32 || sb < m32 || sr < out_w32) return MIDINT_EINVAL;
    if (!(m32 & 1u) && !((sa | sb | sr) & 1u) && aligned8(a) && aligned8(b) && aligned8(r))
        // same bytes as the u64 layout at half the strides
        return u64_op(ctx, op, n, m, (const uint64_t*)a, sa / 2, (const uint64_t*)b, sb / 2, (uint64_t*)r, sr / 2,
                      carry, stream);
    midint_status st = bind_device(ctx);
    if (st != MIDINT_OK) return st;
    // zero-extend each instance to mp batch-major limbs, run the u64 kernel,
    // keep the low words: every operation here commutes with truncation mod
    // 2^(32 m32). An addition keeps one spare word above m32 for its carry.
    const uint32_t mp = op == Op::Add ? m32 / 2 + 1 : m;
    if (op == Op::Add && mp > midint_max_limbs_add()) return MIDINT_EUNSUPPORTED;
    const uint32_t out_limbs = op_out_limbs(op, mp);
    const size_t nm = (size_t)n * mp;
    hipStream_t s = (hipStream_t)stream;
    const bool poly_chain = op == Op::Poly && !poly_uses_fused(mp);
    return with_scratch(ctx, 2 * nm + (size_t)n * out_limbs + (poly_chain ? 4 * nm : 0), s,
                        [&](uint64_t* scr) {
                            uint64_t* pa = scr;
                            uint64_t* pb = pa + nm;
                            uint64_t* pr = pb + nm;
                            uint64_t* extra = pr + (size_t)n * out_limbs;
                            hipError_t e = launch_pack_u32(n, m32, mp, a, sa, pa, s);
                            if (e == hipSuccess) e = launch_pack_u32(n, m32, mp, b, sb, pb, s);
                            if (e == hipSuccess) {
                                switch (op) {
                                    case Op::Add: e = launch_add(n, mp, pa, pb, pr, nullptr, false, s); break;
                                    case Op::Add6: e = launch_add(n, mp, pa, pb, pr, nullptr, true, s); break;
                                    case Op::MulLo: e = launch_mul(n, mp, pa, pb, pr, false, s); break;
                                    case Op::MulFull: e = launch_mul(n, mp, pa, pb, pr, true, s); break;
                                    case Op::Poly: e = launch_poly(n, mp, pa, pb, pr, extra, s); break;
                                }
                            }
                            if (e == hipSuccess)
                                e = launch_unpack_u32(n, out_w32, out_limbs, pr, r, sr, op == Op::Add ? carry : nullptr,
                                                      s);
                            return e;
                        });
}

midint_status midint_add_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                             uint32_t* d_r, uint8_t* d_carry, void* stream) {
    return u32_op(ctx, Op::Add, n, m32, d_a, 0, d_b, 0, d_r, 0, d_carry, stream);
}
midint_status midint_add6_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                              uint32_t* d_r, void* stream) {
    return u32_op(ctx, Op::Add6, n, m32, d_a, 0, d_b, 0, d_r, 0, nullptr, stream);
}
midint_status midint_mul_lo_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                const uint32_t* d_b, uint32_t* d_r, void* stream) {
    return u32_op(ctx, Op::MulLo, n, m32, d_a, 0, d_b, 0, d_r, 0, nullptr, stream);
}
midint_status midint_mul_full_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                  const uint32_t* d_b, uint32_t* d_r, void* stream) {
    return u32_op(ctx, Op::MulFull, n, m32, d_a, 0, d_b, 0, d_r, 0, nullptr, stream);
}
midint_status midint_poly_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                              uint32_t* d_r, void* stream) {
    return u32_op(ctx, Op::Poly, n, m32, d_a, 0, d_b, 0, d_r, 0, nullptr, stream);
}

midint_status midint_mul_lo_u32_noalloc(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                        const uint32_t* d_b, uint32_t* d_r, void* stream) {
    if (!ctx || m32 == 0) return MIDINT_EINVAL;
    if (n == 0) return MIDINT_OK;
    if (!d_a || !d_b || !d_r || ((((uintptr_t)d_a) | ((uintptr_t)d_b) | ((uintptr_t)d_r)) & 3u)) return MIDINT_EINVAL;
    const uint32_t m = (m32 + 1) / 2;
    if (m > midint_max_limbs_mul()) return MIDINT_EUNSUPPORTED;
    midint_status st = bind_device(ctx);
    if (st != MIDINT_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (!(m32 & 1u) && aligned8(d_a) && aligned8(d_b) && aligned8(d_r))  // the u64 layout: no scratch
        return hip_status(launch_mul(n, m, (const uint64_t*)d_a, (const uint64_t*)d_b, (uint64_t*)d_r, false, s));
    const size_t nm = (size_t)n * m;
    if (3 * nm > kNoAllocScratchLimbs) return MIDINT_ETOOBIG;
    uint64_t* scr = nullptr;
    hipError_t e = noalloc_scratch(ctx->device, &scr);
    if (e != hipSuccess) return hip_status(e);
    uint64_t* pa = scr;
    uint64_t* pb = pa + nm;
    uint64_t* pr = pb + nm;
    e = launch_pack_u32(n, m32, m, d_a, m32, pa, s);
    if (e == hipSuccess) e = launch_pack_u32(n, m32, m, d_b, m32, pb, s);
    if (e == hipSuccess) e = launch_mul(n, m, pa, pb, pr, false, s);
    if (e == hipSuccess) e = launch_unpack_u32(n, m32, m, pr, d_r, m32, nullptr, s);
    return hip_status(e);
}

// ------------------------------------------------------------ CGBN layout
midint_status midint_cgbn(midint_ctx* ctx, uint32_t op, uint32_t n, uint32_t bits, void* d_instances,
                          void* stream) {
    if (!ctx || bits == 0 || bits % 32 != 0) return MIDINT_EINVAL;
    Op o;
    switch (op) {
        case MIDINT_CGBN_ADD: o = Op::Add; break;
        case MIDINT_CGBN_ADD6: o = Op::Add6; break;
        case MIDINT_CGBN_MUL: o = Op::MulLo; break;
        case MIDINT_CGBN_POLY: o = Op::Poly; break;
        default: return MIDINT_EINVAL;
    }
    const uint32_t w = bits / 32;  // words per field; instance = {a, b, sum}
    uint32_t* base = (uint32_t*)d_instances;
    return u32_op(ctx, o, n, w, base, 3ull * w, base + w, 3ull * w, base + 2 * w, 3ull * w, nullptr, stream);
}

// ------------------------------------------------- prime-field variants
static midint_status get_prime_tables(midint_ctx* ctx, uint32_t prime, uint32_t lg, bool ref_root,
                                      PrimeTables* out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    const uint32_t key = (prime << 8) | (lg << 1) | (ref_root ? 1u : 0u);
    auto it = ctx->prime_ntt.find(key);
    if (it != ctx->prime_ntt.end()) {
        *out = it->second;
        return MIDINT_OK;
    }
    const size_t half = (size_t)1 << (lg - 1);
    std::vector<uint64_t> hf(half), hi(half);
    PrimeTables t;
    if (ntt_prime_tables(prime, lg, ref_root, hf.data(), hi.data(), &t.scale_inv, &t.scale_prod) != 0)
        return MIDINT_EUNSUPPORTED;
    t.d_fwd = t.d_inv = nullptr;
    hipError_t e = hipMalloc((void**)&t.d_fwd, half * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&t.d_inv, half * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(t.d_fwd, hf.data(), half * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.d_inv, hi.data(), half * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (t.d_fwd) (void)hipFree(t.d_fwd);
        if (t.d_inv) (void)hipFree(t.d_inv);
        return hip_status(e);
    }
    ctx->prime_ntt[key] = t;
    *out = t;
    return MIDINT_OK;
}

static bool valid_prime(uint32_t prime) {
    return prime == MIDINT_PRIME_GOLDILOCKS || prime == MIDINT_PRIME_P32 || prime == MIDINT_PRIME_P64;
}

midint_status midint_ntt_transform(midint_ctx* ctx, uint32_t prime, uint32_t flags, uint32_t n, uint32_t lg,
                                   const uint64_t* d_x, uint64_t* d_y, void* stream) {
    if (!ctx || !valid_prime(prime) || (flags & ~3u)) return MIDINT_EINVAL;
    const bool ref_root = (flags & MIDINT_NTT_REFERENCE_ROOT) != 0;
    if (ref_root && prime == MIDINT_PRIME_GOLDILOCKS) return MIDINT_EINVAL;
    if (lg < 1 || lg > kNttMaxLg) return MIDINT_EUNSUPPORTED;
    if (n == 0) return MIDINT_OK;
    if (!d_x || !d_y) return MIDINT_EINVAL;
    midint_status st = bind_device(ctx);
    if (st != MIDINT_OK) return st;
    PrimeTables t;
    if ((st = get_prime_tables(ctx, prime, lg, ref_root, &t)) != MIDINT_OK) return st;
    const int mode = (flags & MIDINT_NTT_INVERSE) ? 2 : 1;
    return hip_status(launch_ntt_prime(prime, mode, n, 0, lg, d_x, nullptr, d_y, t.d_fwd, t.d_inv, t.scale_inv,
                                       (hipStream_t)stream));
}

midint_status midint_mul_ntt_prime(midint_ctx* ctx, uint32_t prime, uint32_t n, uint32_t m, const uint64_t* d_a,
                                   const uint64_t* d_b, uint64_t* d_r, void* stream) {
    if (ctx && prime == MIDINT_PRIME_GOLDILOCKS) return midint_mul_ntt(ctx, n, m, d_a, d_b, d_r, stream);
    if (!ctx || !valid_prime(prime)) return MIDINT_EINVAL;
    midint_status st = check_binary(ctx, n, m, d_a, d_b, d_r);
    if (st != MIDINT_OK || n == 0) return st;
    const uint32_t ncoef = (64 / ntt_prime_digit_bits(prime)) * m;  // coefficients per operand
    uint32_t lg = 6;
    while ((1u << lg) < 2 * ncoef) lg++;
    if (lg > kNttMaxLg) return MIDINT_EUNSUPPORTED;
    if ((st = bind_device(ctx)) != MIDINT_OK) return st;
    PrimeTables t;
    if ((st = get_prime_tables(ctx, prime, lg, false, &t)) != MIDINT_OK) return st;
    return hip_status(launch_ntt_prime(prime, 0, n, m, lg, d_a, d_b, d_r, t.d_fwd, t.d_inv, t.scale_prod,
                                       (hipStream_t)stream));
}

// ----------------------------------------------------------- host drop-ins
enum class HostOp { Add, MulLo, MulFull, MulNtt, Add6, Poly, PolyNtt };

static midint_status run_host(int device, HostOp op, uint32_t n, uint32_t m, const uint64_t* h_a,
                              const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec,
                              uint32_t prime = MIDINT_PRIME_GOLDILOCKS) {
    if (m == 0 || (n > 0 && (!h_a || !h_b || !h_r))) return MIDINT_EINVAL;
    if (n == 0) return MIDINT_OK;
    midint_ctx* ctx = nullptr;
    midint_status st = midint_init(&ctx, device);
    if (st != MIDINT_OK) return st;
    const size_t in_bytes = (size_t)n * m * sizeof(uint64_t);
    const size_t out_limbs = (op == HostOp::MulFull || op == HostOp::MulNtt) ? (size_t)2 * n * m : (size_t)n * m;
    uint64_t *d_a = nullptr, *d_b = nullptr, *d_r = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipMalloc((void**)&d_a, in_bytes);
    if (he == hipSuccess) he = hipMalloc((void**)&d_b, in_bytes);
    if (he == hipSuccess) he = hipMalloc((void**)&d_r, out_limbs * sizeof(uint64_t));
    if (he == hipSuccess) he = hipMemcpy(d_a, h_a, in_bytes, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(d_b, h_b, in_bytes, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) st = hip_status(he);

    auto once = [&]() -> midint_status {
        switch (op) {
            case HostOp::Add: return midint_add(ctx, n, m, d_a, d_b, d_r, nullptr, nullptr);
            case HostOp::MulLo: return midint_mul_lo(ctx, n, m, d_a, d_b, d_r, nullptr);
            case HostOp::MulFull: return midint_mul_full(ctx, n, m, d_a, d_b, d_r, nullptr);
            case HostOp::MulNtt: return midint_mul_ntt_prime(ctx, prime, n, m, d_a, d_b, d_r, nullptr);
            case HostOp::Add6: return midint_add6(ctx, n, m, d_a, d_b, d_r, nullptr);
            case HostOp::Poly: return midint_poly(ctx, n, m, d_a, d_b, d_r, nullptr);
            case HostOp::PolyNtt: return midint_poly_ntt(ctx, n, m, d_a, d_b, d_r, nullptr);
        }
        return MIDINT_EINVAL;
    };
    if (st == MIDINT_OK) st = once();  // dry run (cuda/main.cu:80-85)
    if (st == MIDINT_OK) st = hip_status(hipDeviceSynchronize());
    if (st == MIDINT_OK && runs > 0) {
        st = hip_status(hipEventRecord(e0, nullptr));
        for (uint32_t i = 0; st == MIDINT_OK && i < runs; i++) st = once();
        if (st == MIDINT_OK) st = hip_status(hipEventRecord(e1, nullptr));
        if (st == MIDINT_OK) st = hip_status(hipEventSynchronize(e1));
        float ms = 0.f;
        if (st == MIDINT_OK) st = hip_status(hipEventElapsedTime(&ms, e0, e1));
        if (st == MIDINT_OK && usec) *usec = 1000.0 * ms / runs;
    }
    if (st == MIDINT_OK) st = hip_status(hipMemcpy(h_r, d_r, out_limbs * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_a);
    (void)hipFree(d_b);
    (void)hipFree(d_r);
    midint_destroy(ctx);
    return st;
}

midint_status midint_gpu_add_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                  uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, HostOp::Add, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_mul_host(int device, uint32_t n, uint32_t m, int full, const uint64_t* h_a,
                                  const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, full ? HostOp::MulFull : HostOp::MulLo, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_mul_fft_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a,
                                      const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, HostOp::MulNtt, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_add6_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                   uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, HostOp::Add6, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_poly_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                   uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, HostOp::Poly, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_poly_fft_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                       uint64_t* h_r, uint32_t runs, double* usec) {
    return run_host(device, HostOp::PolyNtt, n, m, h_a, h_b, h_r, runs, usec);
}

midint_status midint_gpu_mul_fft_prime_host(int device, uint32_t prime, uint32_t n, uint32_t m, const uint64_t* h_a,
                                            const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec) {
    if (!valid_prime(prime)) return MIDINT_EINVAL;
    return run_host(device, HostOp::MulNtt, n, m, h_a, h_b, h_r, runs, usec, prime);
}

// ----------------------------------------------------------- generators
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void midint_gen_limbs(uint64_t seed, uint64_t* out, size_t count) {
    uint64_t sm = seed;
    uint64_t s0 = splitmix64(&sm), s1 = splitmix64(&sm), s2 = splitmix64(&sm), s3 = splitmix64(&sm);
    for (size_t i = 0; i < count; i++) {
        out[i] = rotl64(s1 * 5, 7) * 9;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
    }
}

void midint_gen_ref_rand(uint32_t seed, size_t num_instances, size_t m_u32, size_t nz, uint32_t* out) {
    srand(seed);
    for (size_t i = 0; i < num_instances; i++)
        for (size_t k = 0; k < m_u32; k++) {
            uint32_t v = 0;
            if (k < nz) {
                const uint32_t lo = (uint32_t)rand() * 2u;
                const uint32_t hi = (uint32_t)rand() * 2u;
                v = (hi << 16) + lo;
            }
            out[i * m_u32 + k] = v;
        }
}

}  // extern "C"
