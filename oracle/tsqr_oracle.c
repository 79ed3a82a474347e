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
