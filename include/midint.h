/*
 * midint -- batched midsize big-integer arithmetic on AMD Instinct MI355X
 * (gfx950). Public C ABI of libmidint.so.
 *
 * Data layout (same contract as the reference, cuda/ker-helpers.cu.h:12-19,
 * cuda/helper.h:116): an integer of m limbs is m uint64_t, limb 0 least
 * significant (a limb is two little-endian u32 words); a batch of n integers
 * is n*m contiguous limbs, instance i at [i*m, (i+1)*m).
 *
 * The reference's interface is C++ templates with compile-time sizes:
 * kernels baddKer/a6pb10Ker/bmulKerQ/polyKerQ<Base,IPB,M,Q>(num_instances,
 * ass, bss, rss) and bmulFFT<P,M,Q>(...), launched by host drivers
 * gpuAdd/gpuMultiply/gpuMulFFT<Base,m>(num_instances, h_as, h_bs, h_rs)
 * (cuda/main.cu). Here sizes are runtime arguments and every entry point is
 * plain C. Each declaration names the reference interface it replaces.
 *
 * Device-pointer entry points take pointers to device memory (8-byte
 * aligned; 16-byte alignment enables vector loads) and a hipStream_t passed
 * as `void*` (NULL = default stream). They enqueue work and return without
 * synchronising. Inputs are read-only; the output must not overlap the
 * inputs unless stated. Errors are returned as midint_status; nothing is
 * printed.
 */
#ifndef MIDINT_H
#define MIDINT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MIDINT_OK = 0,
    MIDINT_EINVAL = 1,       /* bad argument (NULL pointer, m == 0, ...)        */
    MIDINT_EUNSUPPORTED = 2, /* size outside what the kernels support          */
    MIDINT_EHIP = 3,         /* HIP runtime error (launch, allocation, copy)   */
    MIDINT_ENOMEM = 4,       /* host or device allocation failed              */
    MIDINT_EBUSY = 5,        /* resource temporarily unavailable (admission)  */
    MIDINT_ENOTFOUND = 6,    /* unknown request / handle                       */
    MIDINT_ETOOBIG = 7       /* request exceeds a fixed-size static buffer     */
} midint_status;

typedef struct midint_ctx midint_ctx;

/* Human-readable text of a status code (static string). */
const char* midint_strerror(midint_status s);
/* Library version string. */
const char* midint_version(void);
/* HIP error code of the last MIDINT_EHIP returned on this thread. */
int midint_last_hip_error(void);

/* Context: device binding, NTT twiddle cache, scratch. Replaces the
 * reference's hard-coded cudaSetDevice(1) (cuda/main.cu:757) and the
 * per-call twiddle construction in gpuMulFFT (cuda/main.cu:458-470). */
midint_status midint_init(midint_ctx** ctx, int device);
midint_status midint_destroy(midint_ctx* ctx);

/* Largest m (limbs) accepted by each operation. */
uint32_t midint_max_limbs_add(void);
uint32_t midint_max_limbs_mul(void);
uint32_t midint_max_limbs_ntt(void);

/* ---------------------------------------------------------------- add */
/* r = (a + b) mod B^m per instance; carry (optional, n bytes) receives the
 * dropped carry-out. In-place (r == a or r == b) is allowed.
 * Replaces baddKer<Base,IPB,M,Q> (cuda/ker-addition.cu.h:227-260) and its
 * launch geometry in gpuAdd (cuda/main.cu:63-82). */
midint_status midint_add(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                         uint64_t* d_r, uint8_t* d_carry, void* stream);

/* r = (6a + 10b) mod B^m per instance.
 * Replaces a6pb10Ker (cuda/ker-addition.cu.h:266-328). */
midint_status midint_add6(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                          uint64_t* d_r, void* stream);

/* ---------------------------------------------------------------- mul */
/* r = (a * b) mod B^m per instance (m limbs out).
 * Replaces bmulKerQ<Base,IPB,M,Q> (cuda/ker-classic-mul.cu.h:468-498) and
 * the geometry in gpuMultiply (cuda/main.cu:243-261). */
midint_status midint_mul_lo(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                            uint64_t* d_r, void* stream);

/* r = a * b per instance (2m limbs out), classical algorithm. The reference
 * obtains a full product by calling bmulKerQ at 2m limbs with the top half
 * of each operand zero (cuda/helper.h:93-114, ourMkRandom<m,nz>). */
midint_status midint_mul_full(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a,
                              const uint64_t* d_b, uint64_t* d_r, void* stream);

/* r = ((a^2 + b)(b^2 + b) + ab) mod B^m per instance.
 * Replaces polyKerQ (cuda/ker-classic-mul.cu.h:500-575): one kernel per
 * call, the four products on the int8 matrix cores and the intermediates in
 * LDS, for every m up to midint_max_limbs_mul(). (With the tuning key
 * "poly.fused" = 0 it runs as a chain of launches through 4*n*m limbs of
 * context scratch instead.) Calls that use context scratch (that chain and
 * the repacking u32 path) on different streams of one context are ordered
 * on the device: each waits for the previous scratch user. */
midint_status midint_poly(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                          uint64_t* d_r, void* stream);

/* r = a * b per instance (2m limbs out) through a Goldilocks NTT on 16-bit
 * coefficients; exact. Replaces bmulFFT<P,M,Q> (cuda/ker-fft-mul.cu.h:
 * 427-454) and gpuMulFFT's twiddle setup (cuda/main.cu:444-476). */
midint_status midint_mul_ntt(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a,
                             const uint64_t* d_b, uint64_t* d_r, void* stream);

/* Raw cyclic convolution of n instances of 2^lg values (each < p =
 * 2^64 - 2^32 + 1): r = INTT(NTT(a) . NTT(b)). Replaces bmulFFTvalid
 * (cuda/ker-fft-mul.cu.h:512-554), the kernel the reference's FFT
 * known-answer test (cuda/main.cu:631-655) checks. */
midint_status midint_ntt_conv(midint_ctx* ctx, uint32_t n, uint32_t lg, const uint64_t* d_a,
                              const uint64_t* d_b, uint64_t* d_r, void* stream);

/* r = ((a^2 + b)(b^2 + b) + ab) mod B^m per instance with NTT products, the
 * composite of midint_poly. Replaces polyFttKer (cuda/ker-fft-mul.cu.h:
 * 456-509): one kernel per call, one instance per workgroup, eight
 * transforms with the intermediates in LDS, no scratch memory;
 * m <= midint_max_limbs_ntt(). */
midint_status midint_poly_ntt(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, const uint64_t* d_b,
                              uint64_t* d_r, void* stream);

/* Precomputes the NTT twiddle tables for operands of m limbs (otherwise done
 * on first use). */
midint_status midint_ntt_prepare(midint_ctx* ctx, uint32_t m);

/* ------------------------------------------------------ strided layouts */
/* The same operations on instances placed at arbitrary limb strides: a's
 * instance i starts at d_a + i*sa, b's at d_b + i*sb, r's at d_r + i*sr
 * (0 selects the batch-major stride: m, or 2m for mul_full). Strides must be
 * at least the instance length. This runs directly on interleaved
 * array-of-structs layouts (see midint_cgbn) without a re-layout pass.
 * The Futhark entry layout [m][ipb][4n]u64 (futhark-clean/big-add.fut:
 * 124-128, sqr-mul.fut) flattened row-major is the batch-major layout itself
 * (m*ipb instances of 4n limbs), so it needs no stride. */
midint_status midint_add_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                 const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, uint8_t* d_carry,
                                 void* stream);
midint_status midint_add6_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                  const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream);
midint_status midint_mul_lo_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                    const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream);
midint_status midint_mul_full_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                      const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr,
                                      void* stream);
/* poly_strided runs the fused kernel of midint_poly at these strides. */
midint_status midint_poly_strided(midint_ctx* ctx, uint32_t n, uint32_t m, const uint64_t* d_a, uint64_t sa,
                                  const uint64_t* d_b, uint64_t sb, uint64_t* d_r, uint64_t sr, void* stream);

/* CGBN's instance_t array (cuda/cgbn-tests/cgbn-kers.cu.h:5-9): per
 * instance three fields of bits/32 u32 words, {a, b, sum}. The op writes
 * sum = (a + b), (6a + 10b), (a * b) or ((a^2 + b)(b^2 + b) + ab), each mod
 * 2^bits, like kernel_add / kernel_6adds / kernel_mul / kernel_poly
 * (:19-110). bits must be a multiple of 32; a multiple of 64 with an
 * 8-byte aligned array runs in place on the strided u64 path, other sizes
 * go through the u32 repack of the next section. */
enum { MIDINT_CGBN_ADD = 0, MIDINT_CGBN_ADD6 = 1, MIDINT_CGBN_MUL = 2, MIDINT_CGBN_POLY = 3 };
midint_status midint_cgbn(midint_ctx* ctx, uint32_t op, uint32_t n, uint32_t bits, void* d_instances,
                          void* stream);

/* ------------------------------------------------------------ u32 limbs */
/* The same operations on integers of m32 u32 words (the reference's U32bits
 * base, cuda/ker-helpers.cu.h:21-28, run by runAdditions / runNaiveMuls<
 * U32bits> in cuda/main.cu:768-774): instance i at words [i*m32, (i+1)*m32),
 * word 0 least significant; r has m32 words (2*m32 for mul_full), carry n
 * bytes (optional). Pointers need 4-byte alignment. Even m32 with 8-byte
 * aligned pointers is the u64 layout and runs in place; otherwise each
 * instance is zero-extended to whole u64 limbs in context scratch
 * (2..7 * n * ceil(m32/2) limbs); such calls on different streams of one
 * context take turns on the device (see midint_poly). */
midint_status midint_add_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                             uint32_t* d_r, uint8_t* d_carry, void* stream);
midint_status midint_add6_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                              uint32_t* d_r, void* stream);
midint_status midint_mul_lo_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                const uint32_t* d_b, uint32_t* d_r, void* stream);
midint_status midint_mul_full_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                  const uint32_t* d_b, uint32_t* d_r, void* stream);
midint_status midint_poly_u32(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a, const uint32_t* d_b,
                              uint32_t* d_r, void* stream);

/* midint_mul_lo_u32 without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Odd word counts (or pointers not 8-byte aligned) are repacked
 * in a fixed per-device static array of 48 MiB instead of context scratch,
 * so 3 * n * ceil(m32 / 2) must not exceed 6 * 2^20 limbs; larger requests
 * return MIDINT_ETOOBIG before anything is enqueued. Even word counts with
 * 8-byte aligned pointers use no scratch at all. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (different streams,
 * contexts or threads) race and must be serialised by the caller. Results
 * equal midint_mul_lo_u32's. */
midint_status midint_mul_lo_u32_noalloc(midint_ctx* ctx, uint32_t n, uint32_t m32, const uint32_t* d_a,
                                        const uint32_t* d_b, uint32_t* d_r, void* stream);

/* ------------------------------------------------- prime-field variants */
/* The reference's FFT primes (FftPrime32 / FftPrime64,
 * cuda/ker-fft-mul.cu.h:44-72; selected in cuda/main.cu:765-766) and
 * Goldilocks. */
typedef enum {
    MIDINT_PRIME_GOLDILOCKS = 0, /* 2^64 - 2^32 + 1, generator 7        */
    MIDINT_PRIME_P32 = 1,        /* 3 * 2^30 + 1 = 3221225473, g = 13  */
    MIDINT_PRIME_P64 = 2         /* 29 * 2^57 + 1 = 4179340454199820289, g = 21 */
} midint_prime;

enum {
    MIDINT_NTT_INVERSE = 1,       /* transform with omega^-1, then scale by L^-1 */
    MIDINT_NTT_REFERENCE_ROOT = 2 /* omega = g^(2^(n - lg)) as the reference's
                                     getOmega (cuda/ker-fft-mul.cu.h:191-199);
                                     p32 / p64 only. Otherwise omega =
                                     g^((p-1)/L), a primitive L-th root. */
};

/* Batched length-2^lg transform (1 <= lg <= 14) of n instances of 2^lg
 * values each (u64, every value < p), natural order in and out: the
 * reference's fft / ifft (cuda/ker-fft-mul.cu.h:304-368), i.e. a bit-reversal
 * permutation then radix-2 decimation-in-time stages. With the reference's
 * root the p32 / p64 transforms are bit-compatible with the reference's (that
 * root has order k * L for p - 1 = k * 2^n, so the map is not a DFT and the
 * inverse does not undo it); with the primitive root they are true NTTs. */
midint_status midint_ntt_transform(midint_ctx* ctx, uint32_t prime, uint32_t flags, uint32_t n, uint32_t lg,
                                   const uint64_t* d_x, uint64_t* d_y, void* stream);

/* Exact full product (2m limbs) through an NTT over the chosen prime, the
 * reference's gpuMulFFT<FftPrime32 / FftPrime64, m> (cuda/main.cu:415-579)
 * made exact: coefficients of 8 bits (p32, m <= 1024) or 16 bits (p64,
 * Goldilocks, m <= 2048) keep every product coefficient below p. */
midint_status midint_mul_ntt_prime(midint_ctx* ctx, uint32_t prime, uint32_t n, uint32_t m, const uint64_t* d_a,
                                   const uint64_t* d_b, uint64_t* d_r, void* stream);

/* ------------------------------------------------------- ragged batches */
/* One op of a ragged batch. Operands live in two packed device buffers
 * (d_a, d_b) at limb offsets off_a / off_b; the result is written to d_r at
 * off_r: n_a + n_b limbs for a product, max(n_a, n_b) + 1 limbs for a sum
 * (both exact, no truncation). Sizes are independent per operand. */
typedef enum { MIDINT_OP_ADD = 0, MIDINT_OP_MUL = 1 } midint_op_kind;
typedef struct {
    uint32_t kind;     /* midint_op_kind */
    uint32_t n_a;      /* limbs of a, 1 .. midint_max_limbs_ragged() */
    uint32_t n_b;      /* limbs of b */
    uint32_t reserved; /* must be 0 */
    uint64_t off_a;
    uint64_t off_b;
    uint64_t off_r;
} midint_op;

/* Launch parameters of the persistent kernel (0 = library default / tuned). */
typedef struct {
    uint32_t threads;       /* 256 or 512 threads per block */
    uint32_t blocks_per_cu; /* persistent blocks per CU (0: occupancy limit) */
} midint_ragged_tuning;

typedef struct midint_ragged_plan midint_ragged_plan;

uint32_t midint_max_limbs_ragged(void);

/* Plans a ragged batch over fixed buffers: packs the ops into work items,
 * uploads them, and builds a hipGraph (counter reset + ONE persistent
 * kernel). No reference counterpart: the reference launches one
 * compile-time specialisation per size (cuda/main.cu:661-709). */
midint_status midint_ragged_plan_create(midint_ctx* ctx, const midint_op* ops, size_t n_ops, const uint64_t* d_a,
                                        const uint64_t* d_b, uint64_t* d_r, const midint_ragged_tuning* tuning,
                                        midint_ragged_plan** plan);
/* Replays the plan's graph on `stream` (one graph launch). */
midint_status midint_ragged_launch(midint_ragged_plan* plan, void* stream);
/* Work items, persistent grid size (blocks) and LDS bytes per block. */
midint_status midint_ragged_plan_stats(const midint_ragged_plan* plan, uint64_t* n_items, uint64_t* grid,
                                       uint64_t* lds_bytes);
midint_status midint_ragged_plan_destroy(midint_ragged_plan* plan);

/* --------------------------------------------------------- serving loop */
/* Continuous batching of single add / mul requests (operands 1..4096 limbs).
 * Requests queue at submit; midint_server_poll() (non-blocking) retires
 * finished micro-batches, admits queued requests into new micro-batches
 * while their device region fits the free part of the HBM pool (requests
 * that do not fit wait; none is dropped), and launches them on free streams,
 * so later requests overlap on the device with earlier ones. Results are
 * exact (full sums and products) and independent of batching. One thread
 * drives a server; inside a call that copies more than about 1 MB of
 * operands or results (a large micro-batch, submit_many, result_many) the
 * library splits the host copies over up to 8 short-lived threads. No
 * reference counterpart (the reference has no serving path). */
typedef struct {
    uint64_t hbm_budget_bytes; /* device pool for in-flight batches (0: 1 GiB) */
    uint64_t max_batch_bytes;  /* device bytes of one micro-batch (0: 64 MiB) */
    uint32_t streams;          /* micro-batches in flight, 1..8 (0: 4) */
    uint32_t max_batch_ops;    /* requests per micro-batch (0: 8192) */
} midint_server_config;

typedef struct {
    uint64_t submitted, completed, queued;
    uint64_t inflight_batches, batches_launched, ops_in_batches;
    uint64_t pool_bytes_in_use, pool_bytes_total;
    uint64_t deferred_admissions; /* admission attempts stopped by the budget */
} midint_server_counters;

typedef struct midint_server midint_server;

midint_status midint_server_create(midint_ctx* ctx, const midint_server_config* cfg, midint_server** server);
/* Copies the operands and queues the request; *id identifies it. */
midint_status midint_server_submit(midint_server* s, uint32_t kind, const uint64_t* a, uint32_t n_a,
                                   const uint64_t* b, uint32_t n_b, uint64_t* id);
/* Queues n requests in one call, as n calls of midint_server_submit would:
 * request i is kinds[i] on a[i] (n_a[i] limbs) and b[i] (n_b[i] limbs); ids[i]
 * receives its id. Every request is checked before any is queued: on an
 * error nothing is queued. No reference counterpart: the batched form of
 * midint_server_submit, for drivers that submit bursts (one call instead of
 * n from an interpreted or remote client). */
midint_status midint_server_submit_many(midint_server* s, size_t n, const uint32_t* kinds,
                                        const uint64_t* const* a, const uint32_t* n_a,
                                        const uint64_t* const* b, const uint32_t* n_b, uint64_t* ids);
/* Drives the scheduler once. Up to `cap` ids of requests completed in this
 * call go to done_ids; *n_done receives their number. */
midint_status midint_server_poll(midint_server* s, uint64_t* done_ids, size_t cap, size_t* n_done);
/* Copies a completed request's result (n_a + n_b limbs for mul,
 * max(n_a, n_b) + 1 for add) and releases it; *latency_s = completion time
 * minus submission time. MIDINT_EBUSY while it is still in flight. */
midint_status midint_server_result(midint_server* s, uint64_t id, uint64_t* r, size_t cap_limbs, size_t* n_limbs,
                                   double* latency_s);
/* Results of n completed requests in one call, as n calls of
 * midint_server_result would: request ids[i]'s limbs at r + offsets[i]
 * (offsets[i] plus its length at most cap_limbs), its length in n_limbs[i]
 * and its latency in latency_s[i] (either array may be NULL). Every id is
 * checked first (ENOTFOUND, EBUSY, EINVAL for a result that does not fit):
 * on an error nothing is copied or released. No reference counterpart: the
 * batched form of midint_server_result. */
midint_status midint_server_result_many(midint_server* s, size_t n, const uint64_t* ids, uint64_t* r,
                                        const uint64_t* offsets, size_t cap_limbs, size_t* n_limbs,
                                        double* latency_s);
midint_status midint_server_stats(const midint_server* s, midint_server_counters* out);
midint_status midint_server_destroy(midint_server* s);

/* ------------------------------------------------- host-pointer drop-ins */
/* Mirror the reference host drivers: allocate device buffers, copy inputs,
 * run the operation `runs` times back to back (after one untimed run),
 * copy the result back and free. *usec (optional) receives the mean time of
 * one run in microseconds, measured with HIP events.
 *   midint_gpu_add_host      <- gpuAdd<Base,m>      (cuda/main.cu:39-162)
 *   midint_gpu_mul_host      <- gpuMultiply<Base,m> (cuda/main.cu:221-356);
 *                               full != 0 selects the 2m-limb product
 *   midint_gpu_mul_fft_host  <- gpuMulFFT<P,m>      (cuda/main.cu:415-579),
 *                               2m-limb exact product */
midint_status midint_gpu_add_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                  uint64_t* h_r, uint32_t runs, double* usec);
midint_status midint_gpu_mul_host(int device, uint32_t n, uint32_t m, int full, const uint64_t* h_a,
                                  const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec);
midint_status midint_gpu_mul_fft_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a,
                                      const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec);
/* The composites each reference driver times after its base operation,
 * with the same allocate / dry run / `runs` timed launches / copy-back flow
 * (m limbs out):
 *   midint_gpu_add6_host     <- a6pb10Ker in gpuAdd (cuda/main.cu:121-157)
 *   midint_gpu_poly_host     <- polyKerQ in gpuMultiply (cuda/main.cu:317-351)
 *   midint_gpu_poly_fft_host <- polyFttKer in gpuMulFFT (cuda/main.cu:526-570) */
midint_status midint_gpu_add6_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                   uint64_t* h_r, uint32_t runs, double* usec);
midint_status midint_gpu_poly_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                   uint64_t* h_r, uint32_t runs, double* usec);
midint_status midint_gpu_poly_fft_host(int device, uint32_t n, uint32_t m, const uint64_t* h_a, const uint64_t* h_b,
                                       uint64_t* h_r, uint32_t runs, double* usec);
/* The same through the chosen prime (midint_mul_ntt_prime): the reference's
 * gpuMulFFT<FftPrime32 / FftPrime64, m> (cuda/main.cu:415-579, selected in
 * cuda/main.cu:765-766), exact. */
midint_status midint_gpu_mul_fft_prime_host(int device, uint32_t prime, uint32_t n, uint32_t m, const uint64_t* h_a,
                                            const uint64_t* h_b, uint64_t* h_r, uint32_t runs, double* usec);

/* ------------------------------------------------------------- tuning */
/* Process-wide launch parameters used by the kernels' launchers, set by the
 * autotuner (midint-arithmetic_amd/autotune.py) or by hand, for example
 * "mul.q" (4|8), "mul.mfma" (0|1), "add.rows", "add.threads", "add.nt" (0|1),
 * "ntt.vals", "ragged.threads" (the full list is in csrc/tuning.cpp).
 * "key@C" (C a power of two) sets the value
 * for operands of (C/2, C] limbs. Unknown keys: MIDINT_ENOTFOUND. Results
 * never depend on these values, only speed. No reference counterpart (the reference fixes
 * Q and block sizes at compile time, cuda/main.cu:64-78, 243-261). */
midint_status midint_tune_set(const char* key, int64_t value);
midint_status midint_tune_get(const char* key, int64_t* value);
/* Applies a tuning file: one "key value" pair per line, '#' comments
 * (written by the autotuner, e.g. midint-arithmetic_amd/tune/mi355x.txt). */
midint_status midint_tune_load(const char* path);

/* ------------------------------------------------------ microbenchmark */
/* Issue throughput of the VALU operations the kernels are built from,
 * measured on `device`: 32x32->64 multiply-adds per second (the price of
 * one partial product) and 32-bit adds per second. Replaces the reference's
 * basic-op latency kernels testBasicOps / multiplyKer
 * (cuda/ker-helpers.cu.h:189-317). Synchronous. */
midint_status midint_probe_valu(int device, double* mad64_per_s, double* add32_per_s);
/* Partial products per second of the classical kernel's accumulate pattern
 * (v_mad_u64_u32 with carry-out + v_addc_co_u32 per product) on register
 * operands: the ceiling of the multiplication loop without LDS traffic. */
midint_status midint_probe_mac(int device, double* pp_per_s);
/* int8 matrix-core ceiling: dense v_mfma_i32_16x16x64_i8 chains (operands
 * changing every step) on every SIMD, 2 x 16 x 16 x 64 ops per instruction.
 * `burst`: best of three short launches after idle; `sustained`: the last
 * quarter of 400 back-to-back launches (about 0.2 s of load, when the
 * clocks have settled under the power limit). Synchronous. No reference
 * counterpart. */
midint_status midint_probe_mfma(int device, double* burst_ops_per_s, double* sustained_ops_per_s);

/* HBM streaming ceiling of the addition's access pattern: d_r = d_a + d_b
 * per u64 limb, no carries, in the add kernel's launch shape (blocks of 4
 * waves, 2 rows of 128 limbs per wave, 16-byte non-temporal accesses) on the
 * current device. `limbs` even, pointers 16-byte aligned. Asynchronous on
 * `stream`. The add's bandwidth is reported against it. No reference
 * counterpart (the reference times only its own add, cuda/main.cu:91-115). */
midint_status midint_probe_stream(uint64_t limbs, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_r,
                                  void* stream);

/* ------------------------------------------------------ data generation */
/* xoshiro256** stream (state seeded by four splitmix64 draws of `seed`)
 * filling `count` limbs; the generator of the benchmark workloads. */
void midint_gen_limbs(uint64_t seed, uint64_t* out, size_t count);
/* The reference's ourMkRandom<m,nz> (cuda/helper.h:99-114) with glibc
 * srand(seed)/rand(): m_u32 words per instance, first nz random, rest 0. */
void midint_gen_ref_rand(uint32_t seed, size_t num_instances, size_t m_u32, size_t nz, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* MIDINT_H */
