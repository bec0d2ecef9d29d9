/*
 * midint oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement, on the CPU, of the reference's batched midsize
 * big-integer arithmetic (coancea/midint-arithmetic, cuda/ker-*.cu.h files). It is the
 * parity checker for the HIP library and nothing else: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libmidint) never links, loads or calls it.
 *
 * Data layout (reference cuda/ker-helpers.cu.h:12-19, cuda/helper.h:116):
 * an integer of m limbs is m uint64_t, limb 0 least significant; a batch of n
 * integers is n*m contiguous limbs, instance i at [i*m, (i+1)*m).
 *
 * Pinning (see tests/test_oracle.py):
 *   add / add6 / mul_lo / mul_full / poly   -> libgmp mpn_add_n / mpn_mul
 *                                              (the reference's own oracle),
 *                                              Python integers, closed forms
 *   ntt (generic prime)                     -> fft-mul.fut:356-378 forward KAT (p32)
 *   ntt_conv                                -> main.cu:637-652 convolution KAT,
 *                                              schoolbook convolution
 *   mul_ntt                                 -> libgmp mpn_mul, Python integers
 *   generators                              -> published xoshiro256** /
 *                                              splitmix64 reference outputs
 */
#ifndef MIDINT_ORACLE_H
#define MIDINT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- addition (reference cuda/ker-addition.cu.h:141-194, 227-260) ---- */
/* r = (a + b) mod B^m; returns the dropped carry-out. */
uint32_t orc_add_n(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m);
void orc_add_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b,
                   uint64_t* r, uint8_t* carry /* may be NULL */);
/* r = (6a + 10b) mod B^m via the six-add chain of a6pb10Ker
 * (reference cuda/ker-addition.cu.h:301-324). */
void orc_add6_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r);

/* ---- classical multiplication (reference cuda/ker-classic-mul.cu.h) ---- */
/* r[0..2m) = a * b (full product). */
void orc_mul_full(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m);
/* r[0..m) = (a * b) mod B^m (reference bmulKerQ semantics, :468-498). */
void orc_mul_lo(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m);
void orc_mul_full_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r);
void orc_mul_lo_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r);
/* r = ((a^2 + b)(b^2 + b) + ab) mod B^m (reference polyKerQ, :500-575). */
void orc_poly_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r);
/* Unequal sizes: r[0..na+nb) = a[0..na) * b[0..nb). */
void orc_mul_unbal(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb);
/* r[0..max(na,nb)+1) = a + b (full, with the carry limb). */
void orc_add_unbal(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb);

/* ---- number-theoretic transform (reference cuda/ker-fft-mul.cu.h) ---- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
/* In-place forward transform X_k = sum_j x_j w^(jk) mod p, natural order
 * in and out, n a power of two (reference fft, :304-348). */
void orc_ntt(uint64_t* x, size_t n, uint64_t w, uint64_t p);
/* Goldilocks p = 2^64 - 2^32 + 1, generator 7: the primitive n-th root. */
uint64_t orc_gl_root(size_t n);
/* Raw cyclic convolution over Goldilocks via forward NTT, pointwise
 * product, inverse NTT and n^-1 scaling (reference bmulFFTvalid, :512-554).
 * a, b, out hold n values each. */
void orc_ntt_conv(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
/* Full product r[0..2m) = a * b through 16-bit coefficients, a Goldilocks
 * NTT of length 8m and exact carry recombination. */
void orc_mul_ntt(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m);
void orc_mul_ntt_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r);

/* ---- data generators ---- */
uint64_t orc_splitmix64(uint64_t* state);
/* xoshiro256** stream seeded by splitmix64(seed); fills count limbs. */
void orc_gen_limbs(uint64_t seed, uint64_t* out, size_t count);
/* Reference generator ourMkRandom<m,nz> (cuda/helper.h:99-114) driven by
 * glibc srand(seed)/rand(): m u32 words per instance, first nz random. */
void orc_gen_ref_rand(uint32_t seed, size_t num_instances, size_t m_u32, size_t nz,
                      uint32_t* out);

/* ---- GMP (the reference's own oracle, cuda/helper.h:146-170) ---- */
/* Returns 1 when libgmp.so.10 and its documented mpn entry points load. */
int orc_gmp_available(void);
uint64_t orc_gmp_add_n(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m);
void orc_gmp_mul(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb);
/* Batched GMP calls with OpenMP over instances (cpu baseline timing). */
void orc_gmp_add_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, int threads);
void orc_gmp_mul_full_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, int threads);
/* A ragged batch op by op with GMP (mpn_mul / mpn_add), OpenMP over ops. */
void orc_gmp_ragged(size_t n_ops, const uint32_t* kind, const uint32_t* na, const uint32_t* nb,
                    const uint64_t* off_a, const uint64_t* off_b, const uint64_t* off_r, const uint64_t* a,
                    const uint64_t* b, uint64_t* r, int threads);
/* Batched restatement with OpenMP over instances (cpu baseline timing). */
void orc_mul_full_batch_mt(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r, int threads);

#ifdef __cplusplus
}
#endif
#endif
