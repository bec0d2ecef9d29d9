/*
 * midint-harness: C host harness over libmidint's C ABI, mirroring the
 * reference's benchmark driver (cuda/main.cu:661-776):
 *
 *   runAdditions  -> run_additions   (cuda/main.cu:661-681)
 *   runNaiveMuls  -> run_naive_muls  (cuda/main.cu:683-709)
 *   runFFTMuls    -> run_fft_muls    (cuda/main.cu:711-746), validated here;
 *                    Goldilocks, then FftPrime32 as cuda/main.cu:766 runs it
 *
 * After each base operation the composite its reference driver times comes
 * too, validated here as well: six additions after each addition
 * (a6pb10Ker, cuda/main.cu:121-157), the polynomial after each classical
 * product (polyKerQ, :317-351) and after each Goldilocks NTT product
 * (polyFttKer, :526-570).
 *
 * Inputs come from the reference's mkRandArrays<32,32> (cuda/main.cu:18-33):
 * at the start of each sweep, ourMkRandom (cuda/helper.h:99-114) fills the
 * operand arrays a, then b, from ONE glibc rand() stream that is never
 * reseeded (glibc's default seed 1) and continues from sweep to sweep, and
 * every size of the sweep reads the first n*m words of those same arrays.
 * For each size (in u32 words, as the reference counts them) the
 * operation runs through the host drop-ins (midint_gpu_*_host: dry run, then
 * `runs` timed launches), and the result is compared word for word with
 * GMP's mpn functions loaded from libgmp.so.10 (cuda/helper.h:146-170,
 * validateExact :71-81). Output: the reference's printf lines
 * (cuda/main.cu:112-115, 307-310, 513-519) followed by one JSON line each.
 *
 * usage: midint-harness [total_work_u32_words] [device] [tune_file | none]
 *        default total_work = 2^27 (cuda/Makefile:25); the launch-parameter
 *        table defaults to ../tune/mi355x.txt next to this binary (the
 *        autotuned choice of kernel per size class), "none" keeps the
 *        library's built-in defaults
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "midint.h"

typedef unsigned long (*gmpn_add_n_fn)(unsigned long*, const unsigned long*, const unsigned long*, long);
typedef unsigned long (*gmpn_mul_fn)(unsigned long*, const unsigned long*, long, const unsigned long*, long);
static gmpn_add_n_fn g_add_n;
static gmpn_mul_fn g_mul;

static int load_gmp(void) {
    void* h = dlopen("libgmp.so.10", RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    g_add_n = (gmpn_add_n_fn)dlsym(h, "__gmpn_add_n");
    g_mul = (gmpn_mul_fn)dlsym(h, "__gmpn_mul");
    return g_add_n && g_mul;
}

/* Reference numAd32OpsOfMultInst (cuda/helper.h:36-43): 300 * n * log2 n for
 * n = u32 words of one operand. */
static double ops_of_mul(uint32_t m_u64) {
    uint32_t n = 2 * m_u64, lg = 0;
    for (uint32_t x = n; x > 1; x >>= 1) lg++;
    return 300.0 * n * lg;
}

static int validate_exact(const uint64_t* want, const uint64_t* got, size_t count) {
    for (size_t i = 0; i < count; i++)
        if (want[i] != got[i]) {
            printf("INVALID RESULT at flat index %zu: %lu vs %lu\n", i, (unsigned long)want[i],
                   (unsigned long)got[i]);
            return 0;
        }
    printf("VALID RESULT!\n");
    return 1;
}

enum { OP_ADD, OP_MUL_LO, OP_MUL_FFT, OP_ADD6, OP_POLY, OP_POLY_FFT };

/* r = a + b mod B^m, r may alias a or b. */
static void ref_add(uint64_t* r, const uint64_t* a, const uint64_t* b, uint32_t m) {
    (void)g_add_n((unsigned long*)r, (const unsigned long*)a, (const unsigned long*)b, (long)m);
}

/* r = a * b mod B^m through GMP's full product (t: 2m limbs). */
static void ref_mul_lo(uint64_t* r, const uint64_t* a, const uint64_t* b, uint32_t m, uint64_t* t) {
    g_mul((unsigned long*)t, (const unsigned long*)a, (long)m, (const unsigned long*)b, (long)m);
    memcpy(r, t, (size_t)m * 8);
}

/* The reference's six-add chain (cuda/ker-addition.cu.h:301-324), every
 * step mod B^m: 6a + 10b. t: 2m limbs. */
static void ref_add6(uint64_t* r, const uint64_t* a, const uint64_t* b, uint32_t m, uint64_t* t) {
    uint64_t *t1 = t, *t3 = t + m;
    ref_add(t1, a, b, m);   /* a + b      */
    ref_add(t1, t1, t1, m); /* 2a + 2b    */
    ref_add(t3, t1, b, m);  /* 2a + 3b    */
    ref_add(t1, t3, t3, m); /* 4a + 6b    */
    ref_add(t1, t1, t3, m); /* 6a + 9b    */
    ref_add(r, t1, b, m);   /* 6a + 10b   */
}

/* The reference's polynomial (cuda/ker-classic-mul.cu.h:533-559), every step
 * mod B^m: (a^2 + b)(b^2 + b) + ab. t: 5m limbs. */
static void ref_poly(uint64_t* r, const uint64_t* a, const uint64_t* b, uint32_t m, uint64_t* t) {
    uint64_t *s1 = t, *s2 = t + m, *full = t + 2 * m, *ab = t + 4 * m;
    ref_mul_lo(s1, a, a, m, full);
    ref_add(s1, s1, b, m);
    ref_mul_lo(s2, b, b, m, full);
    ref_add(s2, s2, b, m);
    ref_mul_lo(ab, a, b, m, full);
    ref_mul_lo(r, s1, s2, m, full);
    ref_add(r, r, ab, m);
}

/* The reference's ourMkRandom<32,32> over `words` u32 words (every word
 * random): two rand() calls per word, continuing the process's glibc stream. */
static void mk_random(size_t words, uint32_t* out) {
    for (size_t k = 0; k < words; k++) {
        const uint32_t lo = (uint32_t)rand() * 2u;
        const uint32_t hi = (uint32_t)rand() * 2u;
        out[k] = (hi << 16) + lo;
    }
}

/* mkRandArrays<32,32>(num_instances): a then b, `drawn` words each, from the
 * continuing stream; only the first `keep` words of each are stored (the
 * sizes of the sweep use at most that many). */
static int mk_rand_arrays(size_t drawn, size_t keep, uint32_t** a, uint32_t** b) {
    *a = malloc(keep * 4);
    *b = malloc(keep * 4);
    uint32_t* spill = malloc(4096 * 4);
    if (!*a || !*b || !spill) {
        free(*a), free(*b), free(spill);
        return 0;
    }
    uint32_t* outs[2] = {*a, *b};
    for (int o = 0; o < 2; o++) {
        mk_random(keep, outs[o]);
        for (size_t done = keep; done < drawn;) { /* the reference draws these too */
            const size_t c = drawn - done < 4096 ? drawn - done : 4096;
            mk_random(c, spill);
            done += c;
        }
    }
    free(spill);
    return 1;
}

static int run_one(int op, int device, uint64_t total_work, uint32_t m_u32, const uint32_t* wa, const uint32_t* wb,
                   int gmp, uint32_t prime) {
    const uint32_t m = m_u32 / 2;                    /* u64 limbs */
    const uint32_t n = (uint32_t)(total_work / m_u32);
    const size_t in_limbs = (size_t)n * m;
    const size_t out_limbs = (op == OP_MUL_FFT) ? 2 * in_limbs : in_limbs;  /* composites: m limbs */
    const uint64_t* a = (const uint64_t*)wa; /* a limb is two little-endian u32 words */
    const uint64_t* b = (const uint64_t*)wb;
    uint64_t* r = malloc(out_limbs * 8);
    uint64_t* want = malloc(out_limbs * 8);
    if (!r || !want) {
        printf("host allocation failed\n");
        free(r), free(want);
        return 0;
    }

    double usec = 0.0;
    midint_status st;
    /* GPU_RUNS_ADD / GPU_RUNS_MUL (cuda/main.cu:11-12) */
    const uint32_t runs = (op == OP_ADD || op == OP_ADD6) ? 50 : 25;
    if (op == OP_ADD)
        st = midint_gpu_add_host(device, n, m, a, b, r, runs, &usec);
    else if (op == OP_ADD6)
        st = midint_gpu_add6_host(device, n, m, a, b, r, runs, &usec);
    else if (op == OP_MUL_LO)
        st = midint_gpu_mul_host(device, n, m, 0, a, b, r, runs, &usec);
    else if (op == OP_POLY)
        st = midint_gpu_poly_host(device, n, m, a, b, r, runs, &usec);
    else if (op == OP_POLY_FFT)
        st = midint_gpu_poly_fft_host(device, n, m, a, b, r, runs, &usec);
    else
        st = midint_gpu_mul_fft_prime_host(device, prime, n, m, a, b, r, runs, &usec);
    if (st != MIDINT_OK) {
        printf("midint error: %s (hip %d)\n", midint_strerror(st), midint_last_hip_error());
        free(r), free(want);
        return 0;
    }
    const unsigned bits = m_u32 * 32;
    const char* name = op == OP_ADD          ? "add"
                       : op == OP_ADD6       ? "add6"
                       : op == OP_MUL_LO     ? "mul_lo"
                       : op == OP_POLY       ? "poly"
                       : op == OP_POLY_FFT   ? "poly_ntt"
                       : prime == MIDINT_PRIME_P32 ? "mul_ntt_p32"
                                                   : "mul_ntt";
    double metric;
    if (op == OP_ADD || op == OP_ADD6) {
        metric = 3.0 * in_limbs * 8 / (usec * 1000.0);
        printf("Our %s of %u-bit Big-Numbers (base u64) runs %u instances in: %.0f microsecs, "
               "GB/sec: %.2f, Mil-Instances/sec: %.2f\n",
               op == OP_ADD ? "Addition" : "SIX Additions", bits, n, usec, metric, n / usec);
    } else if (op == OP_POLY || op == OP_POLY_FFT) {
        metric = 4.0 * n * ops_of_mul(m) / (usec * 1000.0);
        printf("Our %s of %u-bits Big-Numbers (in base = 64 bits) runs %u instances in: %.0f microsecs, "
               "Gu32ops/sec: %.2f, Mil-Instances/sec: %.2f\n",
               op == OP_POLY ? "Polynomial" : "Poly-FFT Computation", bits, n, usec, metric, n / usec);
    } else {
        metric = (double)n * ops_of_mul(m) / (usec * 1000.0);
        printf("%s Multiplication of %u-bits Big-Numbers (in base = 64 bits) runs %u instances in: %.0f "
               "microsecs, Gu32ops/sec: %.2f, Mil-Instances/sec: %.2f\n",
               op == OP_MUL_LO ? "N^2" : "Our FFT", bits, n, usec, metric, n / usec);
    }
    int ok = -1;
    if (gmp) {
        uint64_t* t = malloc(5 * (size_t)m * 8);
        if (!t) {
            printf("host allocation failed\n");
            free(r), free(want);
            return 0;
        }
        for (uint32_t i = 0; i < n; i++) {
            const uint64_t* x = a + (size_t)i * m;
            const uint64_t* y = b + (size_t)i * m;
            uint64_t* w = want + (size_t)i * (op == OP_MUL_FFT ? 2 * m : m);
            switch (op) {
                case OP_ADD: ref_add(w, x, y, m); break;
                case OP_ADD6: ref_add6(w, x, y, m, t); break;
                case OP_MUL_LO: ref_mul_lo(w, x, y, m, t); break;
                case OP_POLY:
                case OP_POLY_FFT: ref_poly(w, x, y, m, t); break;
                default:
                    g_mul((unsigned long*)w, (const unsigned long*)x, (long)m, (const unsigned long*)y, (long)m);
                    break;
            }
        }
        free(t);
        ok = validate_exact(want, r, out_limbs);
    }
    printf("{\"op\": \"%s\", \"bits\": %u, \"instances\": %u, \"usec\": %.3f, \"%s\": %.3f, \"valid\": %s}\n",
           name, bits, n, usec, (op == OP_ADD || op == OP_ADD6) ? "GB_per_s" : "Gu32ops_per_s", metric,
           ok < 0 ? "null" : ok ? "true" : "false");
    fflush(stdout);
    free(r), free(want);
    return ok != 0;
}

/* Loads the launch-parameter table: `arg`, or the one shipped beside the
 * executable (../tune/mi355x.txt); "none" keeps the built-in defaults. */
static void load_tuning(const char* arg) {
    char path[4096];
    if (arg && strcmp(arg, "none") == 0) {
        printf("tuning: built-in defaults\n");
        return;
    }
    if (arg) {
        snprintf(path, sizeof path, "%s", arg);
    } else {
        const ssize_t n = readlink("/proc/self/exe", path, sizeof path - 1);
        if (n <= 0) return;
        path[n] = 0;
        char* slash = strrchr(path, '/');
        if (!slash) return;
        snprintf(slash, sizeof path - (size_t)(slash - path), "/../tune/mi355x.txt");
    }
    const midint_status st = midint_tune_load(path);
    printf("tuning: %s (%s)\n", path, st == MIDINT_OK ? "loaded" : midint_strerror(st));
}

int main(int argc, char** argv) {
    uint64_t total_work = 1ull << 27;
    int device = 0;
    if (argc > 1) total_work = strtoull(argv[1], NULL, 10);
    if (argc > 2) device = atoi(argv[2]);
    load_tuning(argc > 3 ? argv[3] : NULL);
    const int gmp = load_gmp();
    if (!gmp) printf("libgmp.so.10 not found: results are not validated\n");
    const uint32_t sizes[] = {4096, 2048, 1024, 512, 256, 128, 64, 32, 16};
    const size_t nsz = sizeof(sizes) / sizeof(sizes[0]);
    int ok = 1;
    uint32_t *a = NULL, *b = NULL;
    /* runAdditions: mkRandArrays<32,32>(total_work/32) -> total_work words each */
    if (!mk_rand_arrays(total_work, total_work, &a, &b)) return 1;
    for (size_t i = 0; i < nsz; i++) {
        ok &= run_one(OP_ADD, device, total_work, sizes[i], a, b, gmp, 0);
        ok &= run_one(OP_ADD6, device, total_work, sizes[i], a, b, gmp, 0);
    }
    free(a), free(b);
    /* runNaiveMuls: the same, continuing the stream */
    if (!mk_rand_arrays(total_work, total_work, &a, &b)) return 1;
    for (size_t i = 0; i < nsz; i++) {
        ok &= run_one(OP_MUL_LO, device, total_work, sizes[i], a, b, gmp, 0);
        ok &= run_one(OP_POLY, device, total_work, sizes[i], a, b, gmp, 0);
    }
    free(a), free(b);
    /* runFFTMuls<FftPrime32>: mkRandArrays<32,32>(total_work/16) draws
     * 2*total_work words per operand; the FFT sizes run total_work/16 words.
     * The Goldilocks sweep (no reference counterpart) reuses those inputs. */
    const uint64_t fft_work = total_work / 16;
    if (!mk_rand_arrays(2 * total_work, fft_work, &a, &b)) return 1;
    /* p32 products take 8-bit coefficients: up to 2048 u32 words (1024 limbs) */
    for (size_t i = 1; i < nsz; i++)
        ok &= run_one(OP_MUL_FFT, device, fft_work, sizes[i], a, b, gmp, MIDINT_PRIME_P32);
    for (size_t i = 0; i < nsz; i++) {
        ok &= run_one(OP_MUL_FFT, device, fft_work, sizes[i], a, b, gmp, MIDINT_PRIME_GOLDILOCKS);
        ok &= run_one(OP_POLY_FFT, device, fft_work, sizes[i], a, b, gmp, 0);
    }
    free(a), free(b);
    return ok ? 0 : 1;
}
