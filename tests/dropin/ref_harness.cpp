// ref_harness.cpp -- a caller written the way the reference's harnesses
// call the library (test_tsqr.cu:16-108, test_qr.cu:16-108), built against
// include/tsqr_compat.hpp and linked with -ltsqr_amd instead of compiling
// the reference's headers.  It exercises the drop-in path end to end:
//   - input: std::default_random_engine(0U) with
//     uniform_real_distribution<T>(-1, 1), filled column-major
//     (test_tsqr.cu:24-28);
//   - workspace: ldwork = 2 * NUM_SM * BLOCK_SIZE rows of n elements
//     (test_tsqr.cu:42), through the reference-shaped tsqr<T> / qr<T>;
//   - one checked call, then warm-up and timed calls with A re-uploaded
//     before each, timed with events as the reference does, and its report
//     line "<kind> Latency: %f ms" (test_tsqr.cu:90);
//   - the reference's accuracy line "|A-QR|/|A| = ..., |I-Q^TQ| / n = ..."
//     (utils.h:380-381, spectral norms) and the elementwise
//     all_close(QR, A) test (abs 1e-4, rel 1e-5, utils.h:352-357), here
//     computed on the host in FP64 from the n x n Gram matrices of A - QR
//     and Q, so that no m x m workspace is needed.
// Usage: ref_harness m n dtype [tsqr|qr] [warmup repeat]; dtype 1 float,
// 2 double (test_tsqr.cu:113-128).  Exit status 0 only if every check holds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "tsqr_compat.hpp"

#define HIP_OK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            std::exit(2);                                                                   \
        }                                                                                   \
    } while (0)

namespace {

// Largest |eigenvalue| of the symmetric n x n matrix S (row-major, destroyed)
// by cyclic Jacobi rotations.
double max_abs_eig(std::vector<double>& S, int n) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += S[i * n + j] * S[i * n + j];
                if (i != j) off += S[i * n + j] * S[i * n + j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = S[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (S[q * n + q] - S[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double a = S[k * n + p], b = S[k * n + q];
                    S[k * n + p] = c * a - s * b;
                    S[k * n + q] = s * a + c * b;
                }
                for (int k = 0; k < n; ++k) {
                    const double a = S[p * n + k], b = S[q * n + k];
                    S[p * n + k] = c * a - s * b;
                    S[q * n + k] = s * a + c * b;
                }
            }
    }
    double mx = 0.0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(S[i * n + i]));
    return mx;
}

template <typename T>
int run(int m, int n, const char* kind, int warmup, int repeat) {
    const bool is_qr = std::strcmp(kind, "qr") == 0;
    std::vector<T> A((size_t)m * n);
    std::default_random_engine eng(0U);
    std::uniform_real_distribution<T> dis(-1.0, 1.0);
    for (auto& a : A) a = dis(eng);

    const int lda = m, ldr = n;
    const int ldwork = 2 * NUM_SM * BLOCK_SIZE;
    T *dA = nullptr, *dR = nullptr, *dW = nullptr;
    HIP_OK(hipMalloc(&dA, sizeof(T) * A.size()));
    HIP_OK(hipMalloc(&dR, sizeof(T) * n * n));
    HIP_OK(hipMalloc(&dW, sizeof(T) * (size_t)ldwork * n));
    hipStream_t stream;
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));

    auto call = [&]() {
        if (is_qr)
            qr<T>(m, n, dA, lda, dR, ldr, dW, ldwork);
        else
            tsqr<T>(m, n, dA, lda, dR, ldr, dW, ldwork);
    };

    HIP_OK(hipMemcpy(dA, A.data(), sizeof(T) * A.size(), hipMemcpyHostToDevice));
    call();
    HIP_OK(hipDeviceSynchronize());
    std::vector<T> Q(A.size()), R((size_t)n * n);
    HIP_OK(hipMemcpy(Q.data(), dA, sizeof(T) * Q.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(R.data(), dR, sizeof(T) * R.size(), hipMemcpyDeviceToHost));

    // FP64 checks on the host: E = A - QR row by row, Gram matrices of E, A, Q
    std::vector<double> GE((size_t)n * n, 0.0), GA((size_t)n * n, 0.0), GQ((size_t)n * n, 0.0);
    std::vector<double> e(n), a(n), q(n);
    bool close = true, lower_zero = true;
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) lower_zero = lower_zero && R[i + (size_t)j * ldr] == T(0);
    for (int r = 0; r < m; ++r) {
        for (int j = 0; j < n; ++j) {
            q[j] = (double)Q[r + (size_t)j * lda];
            a[j] = (double)A[r + (size_t)j * lda];
        }
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k <= j; ++k) s += q[k] * (double)R[k + (size_t)j * ldr];
            e[j] = a[j] - s;
            if (!(std::fabs(e[j]) <= std::max(1e-4, 1e-5 * std::fabs(a[j])))) close = false;
        }
        for (int i = 0; i < n; ++i)
            for (int j = i; j < n; ++j) {
                GE[i * n + j] += e[i] * e[j];
                GA[i * n + j] += a[i] * a[j];
                GQ[i * n + j] += q[i] * q[j];
            }
    }
    double fe = 0.0, fa = 0.0, fq = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            GE[j * n + i] = GE[i * n + j];
            GA[j * n + i] = GA[i * n + j];
            GQ[j * n + i] = GQ[i * n + j];
        }
    for (int i = 0; i < n; ++i) {
        fe += GE[i * n + i];
        fa += GA[i * n + i];
        for (int j = 0; j < n; ++j) {
            const double d = GQ[i * n + j] - (i == j ? 1.0 : 0.0);
            fq += d * d;
            GQ[i * n + j] = d;
        }
    }
    const double back2 = std::sqrt(max_abs_eig(GE, n) / max_abs_eig(GA, n));
    const double orth2 = max_abs_eig(GQ, n);
    const double backf = std::sqrt(fe / fa), orthf = std::sqrt(fq);

    hipEvent_t start, stop;
    HIP_OK(hipEventCreate(&start));
    HIP_OK(hipEventCreate(&stop));
    for (int i = 0; i < warmup; ++i) {
        HIP_OK(hipMemcpy(dA, A.data(), sizeof(T) * A.size(), hipMemcpyHostToDevice));
        HIP_OK(hipDeviceSynchronize());
        call();
        HIP_OK(hipDeviceSynchronize());
    }
    float total = 0.f;
    for (int i = 0; i < repeat; ++i) {
        HIP_OK(hipMemcpy(dA, A.data(), sizeof(T) * A.size(), hipMemcpyHostToDevice));
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(start, stream));
        call();
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(stop, stream));
        HIP_OK(hipEventSynchronize(stop));
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, start, stop));
        total += t;
    }
    std::printf("%s Latency: %f ms\n", is_qr ? "qr" : "tsqr", repeat > 0 ? total / repeat : 0.f);
    std::printf("|A-QR|/|A| = %.17f, |I-Q^TQ| / n = %.17f\n", back2, orth2 / n);
    std::printf("frobenius: |A-QR|_F/|A|_F = %.3e, |I-Q^TQ|_F = %.3e, all_close = %d, R lower zero = %d\n", backf,
                orthf, close ? 1 : 0, lower_zero ? 1 : 0);
    std::printf("status: %d\n", tsqr_compat_last_status());

    HIP_OK(hipEventDestroy(start));
    HIP_OK(hipEventDestroy(stop));
    HIP_OK(hipStreamDestroy(stream));
    HIP_OK(hipFree(dA));
    HIP_OK(hipFree(dR));
    HIP_OK(hipFree(dW));
    const double u = sizeof(T) == 4 ? 5.960464477539063e-08 : 1.1102230246251565e-16;
    const double bound = 8.0 * n * u;
    const bool ok = close && lower_zero && backf <= bound && orthf <= bound;
    if (!ok) std::printf("FAILED: bound %.3e\n", bound);
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    int m = 13824, n = 32, dtype = 2;
    const char* kind = "tsqr";
    int warmup = 20, repeat = 50;  // test_tsqr.cu:13-14
    if (argc >= 4) {
        m = std::atoi(argv[1]);
        n = std::atoi(argv[2]);
        dtype = std::atoi(argv[3]);
    }
    if (argc >= 5) kind = argv[4];
    if (std::strcmp(kind, "qr") == 0) warmup = 2, repeat = 5;  // test_qr.cu:13-14
    if (argc >= 7) {
        warmup = std::atoi(argv[5]);
        repeat = std::atoi(argv[6]);
    }
    if (m < 1 || n < 1 || (std::strcmp(kind, "tsqr") != 0 && std::strcmp(kind, "qr") != 0)) {
        std::fprintf(stderr, "usage: %s m n dtype(1 float, 2 double) [tsqr|qr] [warmup repeat]\n", argv[0]);
        return 2;
    }
    if (dtype == 1) return run<float>(m, n, kind, warmup, repeat);
    if (dtype == 2) return run<double>(m, n, kind, warmup, repeat);
    std::fprintf(stderr, "dtype must be 1 (float) or 2 (double)\n");
    return 2;
}
