// u32-limb layout support (the reference's U32bits base,
// cuda/ker-helpers.cu.h:21-28): instance i of n integers of m32 u32 words
// occupies words [i*m32, (i+1)*m32), word 0 least significant. For even m32
// that is byte-for-byte the u64 layout with m32/2 limbs; for odd m32 the
// instances are repacked into whole u64 limbs (top half zero) in scratch
// and back, which these two kernels do.
#include <mutex>

#include "midint_internal.h"

namespace midint {

// dst[i][0..m) (u64) = src[i*ss .. i*ss + m32) (u32), zero-extended; 2m >= m32.
__global__ void pack_u32_kernel(uint64_t n, uint32_t m32, uint32_t m, const uint32_t* __restrict__ src, uint64_t ss,
                                uint64_t* __restrict__ dst) {
    const uint64_t total = n * m;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = t / m;
        const uint32_t j = (uint32_t)(t - i * m);
        const uint32_t* s = src + i * ss;
        const uint32_t lo = 2 * j < m32 ? s[2 * j] : 0u;
        const uint32_t hi = 2 * j + 1 < m32 ? s[2 * j + 1] : 0u;
        dst[t] = ((uint64_t)hi << 32) | lo;
    }
}

// dst[i*ds .. i*ds + w32) (u32) = the low w32 words of src[i] (u64, stride
// ms limbs); carry[i] (optional) = bit 0 of word w32 of src[i].
__global__ void unpack_u32_kernel(uint64_t n, uint32_t w32, uint32_t ms, const uint64_t* __restrict__ src,
                                  uint32_t* __restrict__ dst, uint64_t ds, uint8_t* __restrict__ carry) {
    const uint64_t total = n * w32;
    const uint32_t* s32 = reinterpret_cast<const uint32_t*>(src);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = t / w32;
        const uint32_t j = (uint32_t)(t - i * w32);
        dst[i * ds + j] = s32[i * 2 * ms + j];
        if (carry && j == 0) carry[i] = (uint8_t)(s32[i * 2 * ms + w32] & 1u);
    }
}

static uint32_t grid_for(uint64_t total) {
    const uint64_t g = (total + 255) / 256;
    return (uint32_t)(g < 8192 ? (g ? g : 1) : 8192);
}

hipError_t launch_pack_u32(uint64_t n, uint32_t m32, uint32_t m, const uint32_t* src, uint64_t ss, uint64_t* dst,
                           hipStream_t s) {
    if (2ull * m < m32 || ss < m32) return hipErrorInvalidValue;
    pack_u32_kernel<<<grid_for(n * m), 256, 0, s>>>(n, m32, m, src, ss, dst);
    return hipGetLastError();
}

hipError_t launch_unpack_u32(uint64_t n, uint32_t w32, uint32_t ms, const uint64_t* src, uint32_t* dst,
                             uint64_t ds, uint8_t* carry, hipStream_t s) {
    if (2ull * ms < (uint64_t)w32 + (carry ? 1 : 0) || ds < w32) return hipErrorInvalidValue;  // carry word must exist
    unpack_u32_kernel<<<grid_for(n * w32), 256, 0, s>>>(n, w32, ms, src, dst, ds, carry);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Static scratch of midint_mul_lo_u32_noalloc: the repacked operands and the
// product (3 * n * ceil(m32 / 2) limbs) live in this fixed device array
// instead of context scratch, so that entry point never allocates. The
// array's name is a host-side stand-in in host code; its device address is
// looked up once per device (noalloc_setup, called by midint_init with the
// device current) and only that address is handed to kernels.
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchLimbs];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
uint64_t* g_noalloc_ptr[kMaxDevices];
hipError_t g_noalloc_err[kMaxDevices];
}  // namespace

hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_noalloc_err[device] = e;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_noalloc_ptr[device] = (uint64_t*)p;
    return hipSuccess;
}

hipError_t noalloc_scratch(int device, uint64_t** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    *out = g_noalloc_ptr[device];
    if (*out) return hipSuccess;
    return g_noalloc_err[device] != hipSuccess ? g_noalloc_err[device] : hipErrorNotInitialized;
}

}  // namespace midint
