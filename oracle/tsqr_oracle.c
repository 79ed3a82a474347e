/*
 * tsqr_oracle.c -- CPU restatement of the reference TSQR / recursive QR path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The HIP product library never links, calls or
 * falls back to it.
 *
 * Contents (each function cites the reference file:line it restates):
 *   oracle_gen_uniform_{s,d}  test_tsqr.cu:24-28 input stream
 *   oracle_gen_normal_{s,d}   N(0,1) input of the recursive-QR configuration
 *                             (std::mt19937_64(0) + Box-Muller; no reference
 *                             counterpart), oracle_mt64_raw its engine
 *   oracle_qr_tile_{s,d}      kernelQR.h:69-291
 *   oracle_tsqr_{s,d}         kernelQR.h:299-497 + TallShinnyQR.h:86-149
 *   oracle_qr_{s,d}           block_QR.h:72-95
 *
 * Parity pins (all in tests/test_oracle.py):
 *   - generator: first values and FNV-1a hash against libstdc++ itself
 *     (tests/golden/gen_uniform_golden.cpp output committed as a fixture);
 *   - tile reflector: closed-form KATs (x=(3,4) gives R00=-5, u=(1.26491,
 *     0.63246)), diagonal and n=1 KATs, and LAPACK dgeqrf (numpy) after
 *     diagonal-sign normalisation;
 *   - tree and chunking: R and Q against LAPACK after sign normalisation,
 *     QR = A and Q^T Q = I invariants, and the reference's own all_close
 *     criterion (utils.h:352-357);
 *   - recursion: R12 = Q1^T A2 against a float64 GEMM, R21 = 0.
 * Not pinned: the reference's exact rounding (its warp-butterfly summation
 * order, kernelQR.h:110-122), which needs an NVIDIA GPU to observe.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Input generator of test_tsqr.cu:24-28 / test_qr.cu:24-28:
 *   std::default_random_engine eng(0U)            -> minstd_rand0,
 *       x_{k+1} = 16807 x_k mod (2^31-1); seed 0 maps to state 1
 *   std::uniform_real_distribution<T>(-1, 1)     -> 2*generate_canonical - 1
 * libstdc++ generate_canonical<T, digits>: k = ceil(digits / 30) engine
 * draws (1 for float, 2 for double), sum_i (g_i - 1) * r^i with r =
 * 2147483646 held as long double and the running power rounded to T,
 * divided by r^k (rounded to T); a result >= 1 becomes nextafter(1, 0).
 * The stream is column-major fill order (std::generate over A).
 * Jump-ahead (a^k mod p) lets chunks be generated independently; the output
 * is identical to the sequential stream.
 * ------------------------------------------------------------------------ */
static const uint64_t MINSTD_A = 16807u, MINSTD_P = 2147483647u;

static uint64_t minstd_pow(uint64_t k) {
    uint64_t result = 1, base = MINSTD_A;
    while (k) {
        if (k & 1) result = (result * base) % MINSTD_P;
        base = (base * base) % MINSTD_P;
        k >>= 1;
    }
    return result;
}

/* state after `draws` engine steps from seed `seed` */
static uint64_t minstd_state(uint32_t seed, uint64_t draws) {
    uint64_t x = seed % MINSTD_P;
    if (x == 0) x = 1;
    return (x * minstd_pow(draws)) % MINSTD_P;
}

static inline float canon_float(uint64_t g) {
    const long double r = 2147483646.0L;
    float sum = (float)(g - 1) * 1.0f;
    float tmp = (float)(1.0f * r);
    float ret = sum / tmp;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret;
}

static inline double canon_double(uint64_t g1, uint64_t g2) {
    const long double r = 2147483646.0L;
    double sum = 0.0, tmp = 1.0;
    sum += (double)(g1 - 1) * tmp;
    tmp = (double)((long double)tmp * r);
    sum += (double)(g2 - 1) * tmp;
    tmp = (double)((long double)tmp * r);
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* out[i] = i-th value of the float stream (values first..first+count) */
void oracle_gen_uniform_s(float *out, int64_t first, int64_t count, uint32_t seed) {
    const int64_t chunk = 1 << 20;
#pragma omp parallel for schedule(static)
    for (int64_t c0 = 0; c0 < count; c0 += chunk) {
        uint64_t x = minstd_state(seed, (uint64_t)(first + c0));
        int64_t end = c0 + chunk < count ? c0 + chunk : count;
        for (int64_t i = c0; i < end; ++i) {
            x = (x * MINSTD_A) % MINSTD_P;
            out[i] = canon_float(x) * 2.0f + -1.0f;
        }
    }
}

void oracle_gen_uniform_d(double *out, int64_t first, int64_t count, uint32_t seed) {
    const int64_t chunk = 1 << 20;
#pragma omp parallel for schedule(static)
    for (int64_t c0 = 0; c0 < count; c0 += chunk) {
        uint64_t x = minstd_state(seed, 2u * (uint64_t)(first + c0));
        int64_t end = c0 + chunk < count ? c0 + chunk : count;
        for (int64_t i = c0; i < end; ++i) {
            uint64_t g1 = (x * MINSTD_A) % MINSTD_P;
            uint64_t g2 = (g1 * MINSTD_A) % MINSTD_P;
            x = g2;
            out[i] = canon_double(g1, g2) * 2.0 + -1.0;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Seeded N(0,1) input of the recursive-QR configuration (no reference
 * counterpart: the reference harness only draws uniforms).  The engine is
 * std::mt19937_64 restated from its published definition (Matsumoto and
 * Nishimura 2000; degree 312, middle word 156, twist constant
 * 0xB5026F5AA96619E9, the C++ standard's seeding and tempering); the normal
 * values are Box-Muller in FP64 over pairs of 53-bit uniforms
 * u = (x >> 11) 2^-53: r = sqrt(-2 log(1 - u1)), t = 2 pi u2, z0 = r cos t,
 * z1 = r sin t.  Pinned by tests/golden/normal_stream.json, written by
 * libstdc++ and glibc themselves (tests/golden/gen_normal_golden.cpp).
 * ------------------------------------------------------------------------ */
typedef struct {
    uint64_t s[312];
    int i;
} mt64;

static void mt64_seed(mt64 *g, uint64_t seed) {
    g->s[0] = seed;
    for (int k = 1; k < 312; ++k)
        g->s[k] = 6364136223846793005ull * (g->s[k - 1] ^ (g->s[k - 1] >> 62)) + (uint64_t)k;
    g->i = 312;
}

static uint64_t mt64_next(mt64 *g) {
    if (g->i >= 312) {
        const uint64_t upper = 0xFFFFFFFF80000000ull, lower = 0x7FFFFFFFull;
        for (int k = 0; k < 312; ++k) {
            const uint64_t y = (g->s[k] & upper) | (g->s[(k + 1) % 312] & lower);
            g->s[k] = g->s[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1u) ? 0xB5026F5AA96619E9ull : 0ull);
        }
        g->i = 0;
    }
    uint64_t x = g->s[g->i++];
    x ^= (x >> 29) & 0x5555555555555555ull;
    x ^= (x << 17) & 0x71D67FFFEDA60000ull;
    x ^= (x << 37) & 0xFFF7EEE000000000ull;
    x ^= x >> 43;
    return x;
}

/* FNV-1a-64 of `bytes` bytes (the fixtures' array hash). */
uint64_t oracle_fnv1a64(const void *data, int64_t bytes) {
    const unsigned char *p = (const unsigned char *)data;
    uint64_t h = 1469598103934665603ull;
    for (int64_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

/* out[0..count): the raw engine outputs (for the fixture check). */
void oracle_mt64_raw(uint64_t *out, int64_t count, uint64_t seed) {
    mt64 g;
    mt64_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = mt64_next(&g);
}

static void gen_normal(double *outd, float *outf, int64_t count, uint64_t seed) {
    mt64 g;
    mt64_seed(&g, seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < count; i += 2) {
        const double u1 = (double)(mt64_next(&g) >> 11) * 0x1.0p-53;
        const double u2 = (double)(mt64_next(&g) >> 11) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(1.0 - u1));
        const double t = two_pi * u2;
        const double z0 = r * cos(t), z1 = r * sin(t);
        if (outd) outd[i] = z0;
        if (outf) outf[i] = (float)z0;
        if (i + 1 < count) {
            if (outd) outd[i + 1] = z1;
            if (outf) outf[i + 1] = (float)z1;
        }
    }
}

void oracle_gen_normal_s(float *out, int64_t count, uint64_t seed) { gen_normal(NULL, out, count, seed); }
void oracle_gen_normal_d(double *out, int64_t count, uint64_t seed) { gen_normal(out, NULL, count, seed); }

#define REAL float
#define SUF s
#include "oracle_impl.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF d
#include "oracle_impl.h"
#undef REAL
#undef SUF
