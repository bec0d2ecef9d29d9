/*
 * midint oracle -- TEST INFRASTRUCTURE ONLY (see midint_oracle.h).
 *
 * CPU restatement of the reference's algorithms, used to check the HIP
 * library bit for bit. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never does.
 *
 * Every routine below is written for clarity, not speed: limbs are processed
 * one at a time with 128-bit intermediates, and NTT twiddles are recomputed
 * with generic modular arithmetic.
 */
#define _GNU_SOURCE
#include "midint_oracle.h"

#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* Addition                                                            */
/* ------------------------------------------------------------------ */

/* Reference baddRegs (cuda/ker-addition.cu.h:141-194): limbwise sum, then a
 * carry chain over the whole integer; the carry out of limb m-1 is dropped
 * (mod B^m). Restated here as the sequential carry chain it parallelises. */
uint32_t orc_add_n(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m) {
    uint64_t c = 0;
    for (size_t i = 0; i < m; i++) {
        u128 s = (u128)a[i] + b[i] + c;
        r[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return (uint32_t)c;
}

/* Reference baddKer (cuda/ker-addition.cu.h:227-260): one independent
 * addition per instance of the batch. */
void orc_add_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b,
                   uint64_t* r, uint8_t* carry) {
    for (size_t i = 0; i < n; i++) {
        uint32_t c = orc_add_n(r + i * m, a + i * m, b + i * m, m);
        if (carry) carry[i] = (uint8_t)c;
    }
}

/* Reference a6pb10Ker (cuda/ker-addition.cu.h:301-324): the same six chained
 * additions t1=a+b, t2=t1+t1, t3=t2+b, t4=t3+t3, t5=t4+t3, r=t5+b. */
void orc_add6_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r) {
    uint64_t* t = (uint64_t*)malloc(5 * m * sizeof(uint64_t));
    uint64_t *t1 = t, *t2 = t + m, *t3 = t + 2 * m, *t4 = t + 3 * m, *t5 = t + 4 * m;
    for (size_t i = 0; i < n; i++) {
        const uint64_t* ai = a + i * m;
        const uint64_t* bi = b + i * m;
        orc_add_n(t1, ai, bi, m);
        orc_add_n(t2, t1, t1, m);
        orc_add_n(t3, t2, bi, m);
        orc_add_n(t4, t3, t3, m);
        orc_add_n(t5, t4, t3, m);
        orc_add_n(r + i * m, t5, bi, m);
    }
    free(t);
}

/* a (na limbs) + b (nb limbs) -> r (max(na,nb)+1 limbs), full width. */
void orc_add_unbal(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb) {
    size_t n = na > nb ? na : nb;
    uint64_t c = 0;
    for (size_t i = 0; i < n; i++) {
        u128 s = (u128)(i < na ? a[i] : 0) + (i < nb ? b[i] : 0) + c;
        r[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    r[n] = c;
}

/* ------------------------------------------------------------------ */
/* Classical multiplication                                            */
/* ------------------------------------------------------------------ */

/* Operand-scanning schoolbook: r = a * b over na x nb limbs. The reference
 * computes the same column sums c_k = sum_{i+j=k} a_i b_j
 * (cuda/ker-classic-mul.cu.h:374-404) and normalises them with one carry
 * scan (:425-463); the value is the same integer. */
void orc_mul_unbal(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb) {
    memset(r, 0, (na + nb) * sizeof(uint64_t));
    for (size_t i = 0; i < na; i++) {
        uint64_t c = 0;
        for (size_t j = 0; j < nb; j++) {
            u128 t = (u128)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        r[i + nb] = c;
    }
}

void orc_mul_full(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m) {
    orc_mul_unbal(r, a, m, b, m);
}

/* Low half only (reference bmulKerQ: the convolution covers columns k < M,
 * cuda/ker-classic-mul.cu.h:406-423; GMP result truncated to m words,
 * cuda/helper.h:164-169). */
void orc_mul_lo(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m) {
    memset(r, 0, m * sizeof(uint64_t));
    for (size_t i = 0; i < m; i++) {
        uint64_t c = 0;
        for (size_t j = 0; i + j < m; j++) {
            u128 t = (u128)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
    }
}

void orc_mul_full_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r) {
    for (size_t i = 0; i < n; i++) orc_mul_full(r + 2 * i * m, a + i * m, b + i * m, m);
}

void orc_mul_full_batch_mt(size_t n, size_t m, const uint64_t* a, const uint64_t* b,
                           uint64_t* r, int threads) {
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t i = 0; i < n; i++) orc_mul_full(r + 2 * i * m, a + i * m, b + i * m, m);
}

void orc_mul_lo_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r) {
    for (size_t i = 0; i < n; i++) orc_mul_lo(r + i * m, a + i * m, b + i * m, m);
}

/* Reference polyKerQ (cuda/ker-classic-mul.cu.h:533-559): every product is
 * the low-half product and every sum is mod B^m. */
void orc_poly_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r) {
    uint64_t* t = (uint64_t*)malloc(6 * m * sizeof(uint64_t));
    uint64_t *t1 = t, *t2 = t + m, *t3 = t + 2 * m, *t4 = t + 3 * m, *t5 = t + 4 * m, *t6 = t + 5 * m;
    for (size_t i = 0; i < n; i++) {
        const uint64_t* ai = a + i * m;
        const uint64_t* bi = b + i * m;
        orc_mul_lo(t1, ai, ai, m);
        orc_add_n(t2, t1, bi, m);
        orc_mul_lo(t3, bi, bi, m);
        orc_add_n(t4, t3, bi, m);
        orc_mul_lo(t5, t2, t4, m);
        orc_mul_lo(t6, ai, bi, m);
        orc_add_n(r + i * m, t5, t6, m);
    }
    free(t);
}

/* ------------------------------------------------------------------ */
/* Number-theoretic transform                                          */
/* ------------------------------------------------------------------ */

/* Reference zmod_t::mul (cuda/ker-fft-mul.cu.h:107-110): double-width
 * product reduced with %. */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p) {
    return (uint64_t)(((u128)a * b) % p);
}

/* Reference zmod_t::pow (cuda/ker-fft-mul.cu.h:130-139), square-and-multiply. */
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = orc_mulmod(r, a, p);
        a = orc_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

static uint64_t addmod(uint64_t a, uint64_t b, uint64_t p) {
    u128 s = (u128)a + b;
    if (s >= p) s -= p;
    return (uint64_t)s;
}
static uint64_t submod(uint64_t a, uint64_t b, uint64_t p) {
    return a >= b ? a - b : (uint64_t)((u128)a + p - b);
}

static uint32_t bitrev(uint32_t j, uint32_t nb) {
    uint32_t r = 0;
    for (; nb--; j >>= 1) r = (r << 1) | (j & 1);
    return r;
}

/* Reference fft (cuda/ker-fft-mul.cu.h:304-348): in-place bit-reversal
 * permutation (permuteIP, :179-189), then log2 n radix-2 decimation-in-time
 * stages; stage q with half-length L/2 = 2^(q-1) multiplies the odd element
 * by w^(r*j), r = n >> q (:325-341). */
void orc_ntt(uint64_t* x, size_t n, uint64_t w, uint64_t p) {
    uint32_t lgn = 0;
    while (((size_t)1 << lgn) < n) lgn++;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t j = bitrev(k, lgn);
        if (j > k) {
            uint64_t t = x[k];
            x[k] = x[j];
            x[j] = t;
        }
    }
    uint64_t* om = (uint64_t*)malloc(n * sizeof(uint64_t));
    om[0] = 1 % p;
    for (size_t i = 1; i < n; i++) om[i] = orc_mulmod(om[i - 1], w, p);
    for (uint32_t q = 1; q <= lgn; q++) {
        size_t L = (size_t)1 << q, h = L >> 1, r = n >> q;
        for (size_t k = 0; k < n; k += L) {
            for (size_t j = 0; j < h; j++) {
                uint64_t tau = orc_mulmod(om[r * j], x[k + j + h], p);
                uint64_t u = x[k + j];
                x[k + j] = addmod(u, tau, p);
                x[k + j + h] = submod(u, tau, p);
            }
        }
    }
    free(om);
}

#define GL_P 0xFFFFFFFF00000001ull

uint64_t orc_gl_root(size_t n) {
    return orc_powmod(7, (GL_P - 1) / n, GL_P);
}

/* Reference bmulFFTvalid (cuda/ker-fft-mul.cu.h:512-554): NTT both inputs,
 * pointwise product, inverse NTT (forward with w^-1), scale by n^-1
 * (ifft, :350-368). */
void orc_ntt_conv(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    uint64_t* A = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* B = (uint64_t*)malloc(n * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) {
        A[i] = a[i] % GL_P;
        B[i] = b[i] % GL_P;
    }
    uint64_t w = orc_gl_root(n);
    orc_ntt(A, n, w, GL_P);
    orc_ntt(B, n, w, GL_P);
    for (size_t i = 0; i < n; i++) A[i] = orc_mulmod(A[i], B[i], GL_P);
    uint64_t winv = orc_powmod(w, GL_P - 2, GL_P);
    orc_ntt(A, n, winv, GL_P);
    uint64_t ninv = orc_powmod((uint64_t)n % GL_P, GL_P - 2, GL_P);
    for (size_t i = 0; i < n; i++) out[i] = orc_mulmod(A[i], ninv, GL_P);
    free(A);
    free(B);
}

/* Full product by NTT. The reference splits limbs into u16 digits
 * (cuda/main.cu:593-605, uhlf_t inputs) and recombines with a carry pass
 * (cuda/ker-fft-help.cu.h:3-95). Here: 4 coefficients of 16 bits per limb,
 * linear convolution through a Goldilocks NTT of length >= 8m (each
 * coefficient of the product is < 4m * 2^32 < p, so it is exact), then
 * r = sum_k c_k 2^(16k) with an exact carry chain. */
void orc_mul_ntt(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m) {
    size_t nc = 4 * m, L = 1;
    while (L < 2 * nc) L <<= 1;
    uint64_t* A = (uint64_t*)calloc(L, sizeof(uint64_t));
    uint64_t* B = (uint64_t*)calloc(L, sizeof(uint64_t));
    uint64_t* C = (uint64_t*)calloc(L, sizeof(uint64_t));
    for (size_t k = 0; k < nc; k++) {
        A[k] = (a[k / 4] >> (16 * (k % 4))) & 0xFFFF;
        B[k] = (b[k / 4] >> (16 * (k % 4))) & 0xFFFF;
    }
    orc_ntt_conv(A, B, C, L);
    u128 acc = 0;
    for (size_t j = 0; j < 2 * m; j++) {
        for (int s = 0; s < 4; s++) {
            size_t k = 4 * j + s;
            if (k < L) acc += (u128)C[k] << (16 * s);
        }
        r[j] = (uint64_t)acc;
        acc >>= 64;
    }
    free(A);
    free(B);
    free(C);
}

void orc_mul_ntt_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r) {
    for (size_t i = 0; i < n; i++) orc_mul_ntt(r + 2 * i * m, a + i * m, b + i * m, m);
}

/* ------------------------------------------------------------------ */
/* Generators                                                          */
/* ------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea, Flood; public-domain reference by Vigna). */
uint64_t orc_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256** (Blackman, Vigna 2018), state seeded by four splitmix64 draws. */
void orc_gen_limbs(uint64_t seed, uint64_t* out, size_t count) {
    uint64_t sm = seed, s[4];
    for (int i = 0; i < 4; i++) s[i] = orc_splitmix64(&sm);
    for (size_t i = 0; i < count; i++) {
        out[i] = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
    }
}

/* Reference ourMkRandom<m,nz> (cuda/helper.h:99-114): two rand() draws per
 * u32 word, each doubled, so bit 0 of every word is 0. */
void orc_gen_ref_rand(uint32_t seed, size_t num_instances, size_t m_u32, size_t nz,
                      uint32_t* out) {
    srand(seed);
    for (size_t i = 0; i < num_instances; i++) {
        for (size_t k = 0; k < m_u32; k++) {
            uint32_t v = 0;
            if (k < nz) {
                uint32_t low = (uint32_t)rand() * 2u;
                uint32_t high = (uint32_t)rand() * 2u;
                v = (high << 16) + low;
            }
            out[i * m_u32 + k] = v;
        }
    }
}

/* ------------------------------------------------------------------ */
/* GMP                                                                 */
/* ------------------------------------------------------------------ */

/* Documented GMP low-level functions (GMP manual, "Low-level Functions");
 * mp_limb_t is a 64-bit unsigned long and mp_size_t a long on this ABI. */
typedef unsigned long (*gmpn_add_n_fn)(unsigned long*, const unsigned long*,
                                       const unsigned long*, long);
typedef unsigned long (*gmpn_mul_fn)(unsigned long*, const unsigned long*, long,
                                     const unsigned long*, long);
typedef unsigned long (*gmpn_add_fn)(unsigned long*, const unsigned long*, long,
                                     const unsigned long*, long);
static gmpn_add_n_fn p_add_n;
static gmpn_mul_fn p_mul;
static gmpn_add_fn p_add;
static int gmp_state = -1;

int orc_gmp_available(void) {
    if (gmp_state >= 0) return gmp_state;
    void* h = dlopen("libgmp.so.10", RTLD_NOW | RTLD_LOCAL);
    if (!h) return gmp_state = 0;
    p_add_n = (gmpn_add_n_fn)dlsym(h, "__gmpn_add_n");
    p_mul = (gmpn_mul_fn)dlsym(h, "__gmpn_mul");
    p_add = (gmpn_add_fn)dlsym(h, "__gmpn_add");
    gmp_state = (p_add_n && p_mul && p_add) ? 1 : 0;
    return gmp_state;
}

uint64_t orc_gmp_add_n(uint64_t* r, const uint64_t* a, const uint64_t* b, size_t m) {
    if (!orc_gmp_available()) abort();
    return p_add_n((unsigned long*)r, (const unsigned long*)a, (const unsigned long*)b, (long)m);
}

/* mpn_mul requires na >= nb >= 1 and a result area of na + nb limbs. */
void orc_gmp_mul(uint64_t* r, const uint64_t* a, size_t na, const uint64_t* b, size_t nb) {
    if (!orc_gmp_available()) abort();
    if (na >= nb)
        p_mul((unsigned long*)r, (const unsigned long*)a, (long)na, (const unsigned long*)b, (long)nb);
    else
        p_mul((unsigned long*)r, (const unsigned long*)b, (long)nb, (const unsigned long*)a, (long)na);
}

void orc_gmp_add_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b, uint64_t* r,
                       int threads) {
    if (!orc_gmp_available()) abort();
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t i = 0; i < n; i++)
        p_add_n((unsigned long*)(r + i * m), (const unsigned long*)(a + i * m),
                (const unsigned long*)(b + i * m), (long)m);
}

void orc_gmp_mul_full_batch(size_t n, size_t m, const uint64_t* a, const uint64_t* b,
                            uint64_t* r, int threads) {
    if (!orc_gmp_available()) abort();
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t i = 0; i < n; i++)
        p_mul((unsigned long*)(r + 2 * i * m), (const unsigned long*)(a + i * m), (long)m,
              (const unsigned long*)(b + i * m), (long)m);
}

/* A ragged batch op by op with GMP (the CPU baseline of the ragged
 * workload): op i is a product (kind 1: na + nb limbs out, mpn_mul with the
 * longer operand first) or a sum (kind 0: max(na, nb) + 1 limbs out,
 * mpn_add with the longer operand first and the carry as the top limb). */
void orc_gmp_ragged(size_t n_ops, const uint32_t* kind, const uint32_t* na, const uint32_t* nb,
                    const uint64_t* off_a, const uint64_t* off_b, const uint64_t* off_r, const uint64_t* a,
                    const uint64_t* b, uint64_t* r, int threads) {
    if (!orc_gmp_available()) abort();
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (size_t i = 0; i < n_ops; i++) {
        const unsigned long* x = (const unsigned long*)(a + off_a[i]);
        const unsigned long* y = (const unsigned long*)(b + off_b[i]);
        unsigned long* z = (unsigned long*)(r + off_r[i]);
        long nx = (long)na[i], ny = (long)nb[i];
        if (nx < ny) {
            const unsigned long* t = x;
            x = y;
            y = t;
            nx = (long)nb[i];
            ny = (long)na[i];
        }
        if (kind[i] == 1)
            p_mul(z, x, nx, y, ny);
        else
            z[nx] = p_add(z, x, nx, y, ny);
    }
}
