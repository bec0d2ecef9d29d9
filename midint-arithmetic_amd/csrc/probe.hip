// VALU throughput probes (reference: the basic-op latency kernels
// additionKer / multiplyKer / testBasicOps, cuda/ker-helpers.cu.h:189-317).
//
// Each lane runs 8 independent dependency chains of one operation for
// `iters` iterations; enough waves are launched to fill every SIMD, so the
// measured rate is the issue throughput of that instruction on gfx950. The
// result is the price of one 32x32->64 partial product used as the
// multiplication roofline.
#include "midint.h"
#include "midint_internal.h"
#include "mul_core.hpp"

namespace midint {

constexpr int kChains = 8;

// 32x32 -> 64 multiply-add (one v_mad_u64_u32 per step and chain).
__global__ __launch_bounds__(256) void probe_mad64(uint32_t iters, uint32_t seed, uint64_t* out) {
    uint64_t acc[kChains];
    uint32_t x = seed ^ threadIdx.x;
    uint32_t y = seed * 2654435761u + blockIdx.x;
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = (uint64_t)(x + c);
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) acc[c] = (uint64_t)(uint32_t)acc[c] * (uint64_t)y + acc[c];
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) s ^= acc[c];
    if (s == 0x123456789ull) out[0] = s;  // keeps the chains live
}

// 32-bit integer xor + add (full-rate reference point: two VALU ops per step).
__global__ __launch_bounds__(256) void probe_add32(uint32_t iters, uint32_t seed, uint64_t* out) {
    uint32_t acc[kChains];
    uint32_t y = seed * 2654435761u + blockIdx.x;
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = seed ^ (threadIdx.x + c);
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) acc[c] = (acc[c] ^ y) + (uint32_t)c + 0x9E3779B9u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) s ^= acc[c];
    if (s == 0x12345678u) out[0] = s;
}

// The classical kernel's accumulate pattern (mac4: 4 x v_mad_u64_u32 with
// carry-out + 4 x v_addc_co_u32) on register operands: partial products per
// second without LDS traffic or loop bookkeeping.
__global__ __launch_bounds__(256) void probe_mac(uint32_t iters, uint32_t seed, uint64_t* out) {
    uint64_t acc[8];
    uint32_t ovf[8];
    uint32_t a = seed ^ threadIdx.x, b0 = seed * 3u + blockIdx.x, b1 = b0 ^ 0x55u, b2 = b0 + 7u, b3 = b0 * 5u;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        acc[c] = c;
        ovf[c] = 0;
    }
    for (uint32_t i = 0; i < iters; i++) {
        mac4(a, b0, b1, b2, b3, acc[0], acc[1], acc[2], acc[3], ovf[0], ovf[1], ovf[2], ovf[3]);
        mac4(a, b1, b2, b3, b0, acc[4], acc[5], acc[6], acc[7], ovf[4], ovf[5], ovf[6], ovf[7]);
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) s ^= acc[c] + ovf[c];
    if (s == 0x123456789ull) out[0] = s;
}

// HBM streaming ceiling of the addition's access pattern: r = a + b per
// u64 limb with no carries, in the add kernel's launch shape for short
// instances (blocks of 4 waves, each wave 2 rows of 128 limbs, lane l holding
// limbs 2l and 2l+1 of a row, 16-byte non-temporal loads and stores).
typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void probe_stream_add(uint64_t limbs, const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b, uint64_t* __restrict__ r) {
    constexpr int R = 2;
    const uint64_t base = (uint64_t)blockIdx.x * (256 / 64) * R * 128 + (threadIdx.x / 64) * R * 128 +
                          2 * (threadIdx.x % 64);
    u64x2_t x[R], y[R];
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint64_t k = base + v * 128;
        if (k + 1 < limbs) {
            x[v] = __builtin_nontemporal_load(reinterpret_cast<const u64x2_t*>(a + k));
            y[v] = __builtin_nontemporal_load(reinterpret_cast<const u64x2_t*>(b + k));
        }
    }
#pragma unroll
    for (int v = 0; v < R; v++) {
        const uint64_t k = base + v * 128;
        if (k + 1 < limbs) __builtin_nontemporal_store(x[v] + y[v], reinterpret_cast<u64x2_t*>(r + k));
    }
}

typedef int probe_v4i __attribute__((ext_vector_type(4)));

// int8 matrix-core issue throughput: 8 independent accumulator chains of
// v_mfma_i32_16x16x64_i8 per wave; the B operand changes every step (one
// xor, one add) so the operands toggle like real digit data.
__global__ __launch_bounds__(256) void probe_mfma16(uint32_t iters, uint32_t seed, int32_t* out) {
    const uint32_t t = threadIdx.x + 977u * blockIdx.x;
    probe_v4i av = {(int)(seed * 2654435761u ^ t), (int)(seed + 0x9E3779B9u * t), (int)(t * 40503u + seed),
                    (int)(~seed ^ (t << 7))};
    probe_v4i bv = {(int)(t * 2246822519u), (int)(seed ^ 0x85EBCA6Bu), (int)(t + 0xC2B2AE35u), (int)(seed * t)};
    const int kx = (int)(seed * 0x27D4EB2Fu + t);
    probe_v4i acc[8];
#pragma unroll
    for (int c = 0; c < 8; c++) acc[c] = probe_v4i{c, 0, 0, 0};
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[c], 0, 0, 0);
        bv.x ^= kx;
        bv.z += kx;
    }
    int s = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) s ^= acc[c].x ^ acc[c].y ^ acc[c].z ^ acc[c].w;
    if (s == 0x1234567) out[0] = s;  // keeps the chains live
}

}  // namespace midint

using namespace midint;

extern "C" midint_status midint_probe_stream(uint64_t limbs, const uint64_t* d_a, const uint64_t* d_b,
                                             uint64_t* d_r, void* stream) {
    if (limbs == 0) return MIDINT_OK;
    if (!d_a || !d_b || !d_r || (limbs & 1u) || ((((uintptr_t)d_a) | ((uintptr_t)d_b) | ((uintptr_t)d_r)) & 15u))
        return MIDINT_EINVAL;
    constexpr uint64_t kPerBlock = 4 * 2 * 128;
    const uint64_t blocks = (limbs + kPerBlock - 1) / kPerBlock;
    if (blocks > 0x7FFFFFFFull) return MIDINT_EUNSUPPORTED;
    probe_stream_add<<<(uint32_t)blocks, 256, 0, (hipStream_t)stream>>>(limbs, d_a, d_b, d_r);
    return hipGetLastError() == hipSuccess ? MIDINT_OK : MIDINT_EHIP;
}

extern "C" midint_status midint_probe_mac(int device, double* pp_per_s) {
    if (!pp_per_s) return MIDINT_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MIDINT_EHIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return MIDINT_EHIP;
    uint64_t* d_out = nullptr;
    if (hipMalloc((void**)&d_out, 64) != hipSuccess) return MIDINT_ENOMEM;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t blocks = (uint32_t)cus * 32, threads = 256, iters = 2048;
    double best = 0;
    for (int rep = 0; rep < 3; rep++) {
        (void)hipEventRecord(e0, nullptr);
        probe_mac<<<blocks, threads>>>(iters, 777u + rep, d_out);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double pp = (double)blocks * threads * iters * 8;
        if (pp / (ms * 1e-3) > best) best = pp / (ms * 1e-3);
    }
    hipError_t e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_out);
    if (e != hipSuccess) return MIDINT_EHIP;
    *pp_per_s = best;
    return MIDINT_OK;
}

extern "C" midint_status midint_probe_valu(int device, double* mad64_per_s, double* add32_per_s) {
    if (!mad64_per_s || !add32_per_s) return MIDINT_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MIDINT_EHIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return MIDINT_EHIP;
    uint64_t* d_out = nullptr;
    if (hipMalloc((void**)&d_out, 64) != hipSuccess) return MIDINT_ENOMEM;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t blocks = (uint32_t)cus * 32;  // 8 waves per SIMD
    const uint32_t threads = 256;
    const uint32_t iters = 4096;
    double rates[2] = {0, 0};
    for (int which = 0; which < 2; which++) {
        for (int rep = 0; rep < 3; rep++) {
            (void)hipEventRecord(e0, nullptr);
            if (which == 0)
                probe_mad64<<<blocks, threads>>>(iters, 12345u + rep, d_out);
            else
                probe_add32<<<blocks, threads>>>(iters, 12345u + rep, d_out);
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            // probe_add32 issues two VALU ops (xor, add) per chain step
            const double ops = (double)blocks * threads * iters * kChains * (which == 0 ? 1 : 2);
            const double r = ops / (ms * 1e-3);
            if (r > rates[which]) rates[which] = r;
        }
    }
    hipError_t e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_out);
    if (e != hipSuccess) return MIDINT_EHIP;
    *mad64_per_s = rates[0];
    *add32_per_s = rates[1];
    return MIDINT_OK;
}

extern "C" midint_status midint_probe_mfma(int device, double* burst_ops_per_s, double* sustained_ops_per_s) {
    if (!burst_ops_per_s || !sustained_ops_per_s) return MIDINT_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MIDINT_EHIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return MIDINT_EHIP;
    int32_t* d_out = nullptr;
    if (hipMalloc((void**)&d_out, 64) != hipSuccess) return MIDINT_ENOMEM;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // 4 waves per SIMD, about 0.5 ms per launch at the published rate
    const uint32_t blocks = (uint32_t)cus * 4, threads = 256, iters = 2048;
    const double ops = (double)blocks * (threads / 64) * iters * 8 * 2.0 * 16 * 16 * 64;
    double best = 0, sustained = 0;
    for (int rep = 0; rep < 3; rep++) {  // short runs after idle: the burst rate
        (void)hipDeviceSynchronize();
        (void)hipEventRecord(e0, nullptr);
        probe_mfma16<<<blocks, threads>>>(iters, 4242u + rep, d_out);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms > 0.f && ops / (ms * 1e-3) > best) best = ops / (ms * 1e-3);
    }
    // 400 back-to-back launches (about 0.2 s): the rate under sustained load
    constexpr int kLaunches = 400, kTail = 100;
    for (int i = 0; i < kLaunches; i++) {
        if (i == kLaunches - kTail) (void)hipEventRecord(e0, nullptr);
        probe_mfma16<<<blocks, threads>>>(iters, 77u + i, d_out);
    }
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms > 0.f) sustained = ops * kTail / (ms * 1e-3);
    hipError_t e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_out);
    if (e != hipSuccess) return MIDINT_EHIP;
    *burst_ops_per_s = best;
    *sustained_ops_per_s = sustained;
    return MIDINT_OK;
}
