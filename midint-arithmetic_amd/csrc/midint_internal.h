// Internal host-side interface between the C ABI (abi.cpp) and the kernel
// translation units. Not installed; the public interface is include/midint.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "midint.h"

namespace midint {

struct AddGeom {
    uint32_t rows;     // rows of 128 limbs per wave
    uint32_t threads;  // threads per block
    uint32_t ipb;      // instances per block (0: size unsupported)
};
AddGeom add_geometry(uint32_t m);
// Longest instance (limbs) that one workgroup holds in registers (16 waves of
// 2 rows); longer instances run on add_long_kernel, chunk by chunk.
constexpr uint32_t kAddRowMax = 16 * 2 * 128;
// Longest addition operand (limbs).
constexpr uint32_t kAddMaxLimbs = 1u << 24;

// Instance strides (limbs) of the operands and the result; 0 = the
// batch-major stride (m, or the output length).
struct Strides {
    uint64_t a = 0, b = 0, r = 0;
};

// r = (a + b) mod B^m (six == false) or (6a + 10b) mod B^m (six == true).
hipError_t launch_add(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                      uint8_t* carry, bool six, hipStream_t s, Strides st = {});

// Largest operand size (limbs) of the classical multiplication kernels.
constexpr uint32_t kMulMaxLimbs = 4096;
// full: r[0..2m) = a*b ; otherwise r[0..m) = a*b mod B^m.
hipError_t launch_mul(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                      hipStream_t s, Strides st = {});
// The int8 matrix-core classical kernel (mul_mfma.hip): whether its layout
// fits this size, and the launch (same contract as launch_mul).
bool mul_mfma_supported(uint32_t m, bool full);
hipError_t launch_mul_mfma(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                           hipStream_t s, Strides st = {});
// One product per wave on the 16 x 16 x 64 int8 instruction (mul_wave16.hip),
// compiled for 64- and 128-limb operands (same contract as launch_mul).
bool mul_wave16_supported(uint32_t m, bool full);
hipError_t launch_mul_wave16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                             hipStream_t s, Strides st = {});
// Any size up to the LDS limit on the 16 x 16 x 64 instruction, one product per
// workgroup, the tile plan at run time (mul_mfma16.hip).
bool mul_mfma16_supported(uint32_t m, bool full);
hipError_t launch_mul_mfma16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, bool full,
                             hipStream_t s, Strides st = {});
// polyKerQ with one instance per wave on the same machinery (64 and 128 limbs).
hipError_t launch_poly_wave16(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                              hipStream_t s, Strides st = {});
// Ragged multiplications on the matrix cores: one product (n_a >= n_b limbs,
// full n_a + n_b limb result) per workgroup of 1, 2, 4 or 8 waves (default
// kMfRgWaves).
constexpr uint32_t kMfRgWaves = 4;
struct MfRgItem {
    const uint64_t* a;  // n_a limbs (the longer operand)
    const uint64_t* b;  // n_b limbs
    uint64_t* r;        // n_a + n_b limbs
    uint32_t n_a, n_b;
    uint32_t mg_pl;  // reciprocal of n_a + n_b for the kernel's divisions
    uint32_t pad;
};
MfRgItem mfma_rg_item(const uint64_t* a, const uint64_t* b, uint64_t* r, uint32_t n_a, uint32_t n_b);
// Launch configuration for items with n_a <= max_na, n_b <= max_nb and
// n_a + n_b <= max_pl; the kernel's arguments are (const MfRgItem* items,
// uint32_t pow8), one workgroup per item. False when the sizes do not fit.
struct MfRgConfig {
    const void* func;
    uint32_t threads;
    size_t lds;
    uint32_t pow8;
};
// 32 x 32 x 32 tiles (mfma32_rg_config) or 16 x 16 x 64 tiles
// (mfma16_rg_config, mul_mfma16.hip); mfma_rg_config (the serving loop's
// choice) takes the second when the tuning key ragged.m16 is 1. Ragged
// plans choose per product (ragged_plan.cpp).
bool mfma_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out,
                    uint32_t waves = kMfRgWaves);
bool mfma32_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out, uint32_t waves);
bool mfma16_rg_config(uint32_t max_na, uint32_t max_nb, uint32_t max_pl, MfRgConfig* out, uint32_t waves);
// One ragged op per wave (mul_rgw_kernel, add_rgw_kernel in mul_rgwave.hip):
// products whose shorter operand (the item's b, n_a >= n_b) has at most
// rgw_max_short() limbs, on 16 x 16 x 64 tiles, and full additions (a, b in
// the op's order). Product kernels: part 0 any n_b up to rgw_max_short(),
// part 1 n_b <= rgw_small_max_short() (fewer registers and LDS), part 2 the
// n_b above that. Kernel arguments (const MfRgItem* items, uint32_t n_items,
// uint32_t pow8) for products, (items, n_items) for additions; grid
// ceil(n_items / rgw_waves()) workgroups of 64 * rgw_waves() threads,
// product workgroups with rgw_lds_bytes(part) of LDS.
uint32_t rgw_max_short();
uint32_t rgw_small_max_short();
size_t rgw_lds_bytes(int part);
uint32_t rgw_waves();
const void* rgw_mul_kernel_ptr(int part);
const void* rgw_add_kernel_ptr();
// r = ((a^2 + b)(b^2 + b) + ab) mod B^m (reference polyKerQ): the launch
// chain over 4*n*m limbs of scratch, and the fused one-kernel version
// (mul_mfma.hip; no scratch) where its LDS layout fits.
hipError_t launch_poly(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                       uint64_t* scratch, hipStream_t s, Strides st = {});
bool poly_fused_supported(uint32_t m);
bool poly_uses_fused(uint32_t m);
hipError_t launch_poly_fused(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                             hipStream_t s, Strides st = {});

// NTT multiplication.
struct NttTables {
    uint32_t lg;          // log2 of the transform length
    uint64_t* d_fwd;      // w^i, i < L
    uint64_t* d_inv;      // w^-i, i < L
    uint64_t inv_len;     // L^-1 mod p
};
constexpr uint32_t kNttMaxLg = 14;  // 16384-point transform: 128 KiB of LDS
uint32_t ntt_lg_for_limbs(uint32_t m);
// Host computation of twiddle tables w^i, w^-i for i < L (allocated by the caller).
void ntt_host_tables(uint32_t lg, uint64_t* fwd, uint64_t* inv, uint64_t* inv_len);
// lo: write only the low m limbs of each product (stride m), else all 2m.
hipError_t launch_mul_ntt(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                          const NttTables& t, hipStream_t s, bool lo = false);
// r = ((a^2 + b)(b^2 + b) + ab) mod B^m with NTT products, one kernel, no
// scratch; t must be the tables of 2^poly_ntt_lg(m) points.
uint32_t poly_ntt_lg(uint32_t m);
hipError_t launch_poly_ntt(uint32_t n, uint32_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                           const NttTables& t, hipStream_t s);
// Raw cyclic convolution of n instances of L = 2^lg values (< p each).
hipError_t launch_ntt_conv(uint32_t n, const uint64_t* a, const uint64_t* b, uint64_t* r, const NttTables& t,
                           hipStream_t s);

// u32-limb layouts: zero-extend m32 words per instance (instance stride ss
// words) to m batch-major limbs (2m >= m32), and copy the low w32 words of
// each ms-limb instance back to a u32 array of instance stride ds (carry:
// bit 0 of word w32, optional; needs 2 ms > w32).
hipError_t launch_pack_u32(uint64_t n, uint32_t m32, uint32_t m, const uint32_t* src, uint64_t ss, uint64_t* dst,
                           hipStream_t s);
hipError_t launch_unpack_u32(uint64_t n, uint32_t w32, uint32_t ms, const uint64_t* src, uint32_t* dst,
                             uint64_t ds, uint8_t* carry, hipStream_t s);
// Static scratch of midint_mul_lo_u32_noalloc (u32.hip): limbs, the lookup of
// its device address on the current device (once per device, at midint_init)
// and that address for a device.
constexpr size_t kNoAllocScratchLimbs = (size_t)6 << 20;  // 48 MiB per device
hipError_t noalloc_setup(int device);
hipError_t noalloc_scratch(int device, uint64_t** out);

// NTT over the reference's primes (p32, p64) and Goldilocks: tables in the
// field's multiplier form (Montgomery for p32/p64), and the launcher.
// mode 0: exact product (m limbs in, 2m out); 1: forward transform of L
// values; 2: inverse transform.
int ntt_prime_tables(uint32_t prime, uint32_t lg, bool ref_root, uint64_t* fwd, uint64_t* inv,
                     uint64_t* scale_inv, uint64_t* scale_prod);
uint32_t ntt_prime_digit_bits(uint32_t prime);
hipError_t launch_ntt_prime(uint32_t prime, int mode, uint32_t n, uint32_t m, uint32_t lg, const uint64_t* a,
                            const uint64_t* b, uint64_t* r, const uint64_t* twf, const uint64_t* twi,
                            uint64_t scale, hipStream_t s);
struct PrimeTables {
    uint64_t* d_fwd;
    uint64_t* d_inv;
    uint64_t scale_inv, scale_prod;
};

// Ragged batches.
constexpr uint32_t kMulQ = 4;          // default columns per lane group of ragged products
constexpr uint32_t kRgWindow = 2048;   // output limbs per work item
constexpr uint32_t kRgMaxLanes = 512;  // virtual lanes per multiplication item
constexpr uint32_t kRgMaxLimbs = 1024; // operand limbs
constexpr uint32_t kRgAdd = 0;
constexpr uint32_t kRgMul = 1;

struct RgOp {
    uint64_t off_s, off_l, off_r;  // limb offsets: shorter operand, longer operand, result
    uint32_t ns, nl;               // limbs (additions: a and b in order)
    uint32_t ns_w, nl_w;           // padded u32 word counts (multiplications)
    uint32_t lanes, lane_base;     // virtual lanes and first lane within the item
    uint32_t lds_base;             // u32 word offset of the operand staging area
    uint32_t seg_base, seg_len;    // output window segment within the item
    uint32_t out_len;              // limbs written to the result
    uint32_t swap;                 // 1: S comes from b, L from a
    uint32_t pad;
};

struct RgItem {
    uint32_t kind, first, count, lanes, window, conv_words;
};

size_t ragged_lds_size(uint32_t threads, uint32_t conv_cap);
// Packs ops (offsets relative to base_a / base_b / base_r) into work items
// for a persistent kernel whose products use q (4 or 8) columns per lane group.
midint_status ragged_pack(const midint_op* ops, size_t n_ops, uint64_t base_a, uint64_t base_b, uint64_t base_r,
                          std::vector<RgItem>& items, std::vector<RgOp>& rgops, uint32_t* conv_cap,
                          uint32_t q = kMulQ);
// The persistent kernel for `threads` (256 or 512) and q (4 or 8); its
// arguments are (items, ops, n_items, counter, A, B, R, conv_cap).
const void* ragged_kernel_ptr(uint32_t threads, uint32_t q = kMulQ);
// Zeroes counters (uint32_t*, uint32_t n): one block of 64 threads, n <= 64.
const void* ragged_reset_ptr();

// Serving: products of operands above 2048 limbs are split into four
// 2048 x 2048 NTT products [P00 | P11 | P01 | P10] (4096 limbs each); the
// combine kernel adds P01 + P10 into limbs [2048, 8192) of P00 | P11.
hipError_t launch_bigmul_combine(uint32_t n, uint64_t* blocks16k, hipStream_t s);
// Serving loop transfers between the pinned staging buffer and HBM, kernels
// on `s`: `bytes` (a multiple of 8; both pointers 16-byte aligned) copied,
// plus per entry of `slots` a zero-padded device slot filled from packed
// limbs (dst[0, n) = src[0, n), zeros up to size) on the way in, and per
// entry of `items` a result gathered from its padded output (carry: the top
// limb is that byte when set) on the way out; the tables are read in place
// (host-mapped staging).
struct ServeSlot {
    const uint64_t* src;
    uint64_t* dst;
    uint32_t n, size;
};
struct ServeGather {
    const uint64_t* src;
    uint64_t* dst;
    const uint8_t* carry;
    uint32_t n, pad;
};
hipError_t launch_stage_in(const void* src, void* dst, size_t bytes, const ServeSlot* slots, uint32_t n_slots,
                           hipStream_t s);
hipError_t launch_stage_out(const void* src, void* dst, size_t bytes, const ServeGather* items, uint32_t n_items,
                            hipStream_t s);

// Autotuned launch parameter (see tuning.cpp), or dflt for unknown keys.
int64_t tune_get(const char* key, int64_t dflt);
// Value for operand size `limbs` ("key@C" for its power-of-two class C), else tune_get.
int64_t tune_get_class(const char* key, uint32_t limbs, int64_t dflt);

// NTT twiddle tables of a context (cached).
midint_status ctx_ntt_tables(midint_ctx* ctx, uint32_t lg, NttTables* out);

}  // namespace midint
