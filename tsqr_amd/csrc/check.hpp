// check.hpp -- device-side FP64 checker of a computed factorisation A = QR.
//
// The reference checks its results on the GPU with cuBLAS and cuSOLVER
// (check_QR_accuracy, utils.h:317-455): it forms QR, tests all_close(QR, A)
// and prints ||A - QR||_2 / ||A||_2 and ||I - Q^T Q||_2 / n from an m x m
// gesvd workspace, so it cannot run above m ~ 46k.  Here every quantity
// comes from three n x n FP64 Gram matrices formed by streaming reductions
// over the m rows:
//   G_A = A^T A,  G_Q = Q^T Q,  G_E = E^T E  with  E = A - Q R,
// E formed in FP64 on the matrix cores (products of working-precision
// values are exact in FP64, the sum is rounded once).  From them (on the
// host, n <= 32):
//   ||A||_F^2 = tr G_A, ||A||_2^2 = lambda_max(G_A), ||E|| likewise,
//   ||I - Q^T Q||_F and ||I - Q^T Q||_2 = max |lambda(G_Q - I)|.
// G_A and G_Q use the streaming Gram kernel (gram.hpp); G_E uses the
// residual kernel below, which has the same reduction tail.
#pragma once
#include "../../include/tsqr_amd.h"
#include "gram.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

// G_E = (A - QR)^T (A - QR) over all m rows, FP64, by a persistent grid of P
// workgroups of 4 waves; wave w takes the 64-row blocks w, w + 4P, ....
// Per block: P = Q R on v_mfma_f64_16x16x4f64 (Q rows as the A operand, R
// upper triangular from registers as the B operand, as in tsqr_gram_q);
// the output layout (lane l, register v of MFMA t and column tile jt: row
// 16v + 4(l/16) + t, column 16 jt + l%16) is also the operand layout of the
// Gram MFMAs, so E = A - P is formed in place (A loaded in that layout) and
// E^T E accumulates with no data movement.  Partial sums as in
// tsqr_gram_stream (stream_reduce).
//
// RESID = false computes the Gram matrix of the product alone, (Q R)^T (Q R)
// (A unused): with Q := A and R := X = R1^{-1} (FP64, row-major: rs_k = 32,
// rs_j = 1) this is the Gram matrix of A R1^{-1}, the second CholeskyQR pass
// of the R-only entry points, formed without writing the product.  R is
// read as R[k rs_k + j rs_j] (column-major: rs_k = 1, rs_j = ldr).
// If run_if is given and run_if[0] == 0 (the flat path was not taken, so R
// is not written), every workgroup returns at once and nothing is written.
template <typename T, typename TR, int NT, bool VEC, bool RESID>
__global__ __launch_bounds__(256) void tsqr_resid_stream(int m, int n, const T* __restrict__ A, int lda,
                                                          const T* __restrict__ Q, int ldq, const TR* __restrict__ R,
                                                          int rs_k, int rs_j, double* __restrict__ partial,
                                                          int* __restrict__ tk, const int* __restrict__ run_if) {
    __shared__ GramLds L;
    __shared__ int bcast;
    if (run_if && !run_if[0]) return;
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lq = lane >> 4;
    const int P = gridDim.x;
    const int W = 4 * P;
    const int w = 4 * blockIdx.x + (threadIdx.x >> 6);
    // B operand of k-group kg (k = 4 kg + l/16) and column tile jt: R[k][j]
    double mb[8][2];
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int k = 4 * kg + lq;
            const int j = 16 * jt + li;
            mb[kg][jt] = (k < n && j < n && k <= j) ? (double)R[(int64_t)k * rs_k + (int64_t)j * rs_j] : 0.0;
        }
    const int nk = (n + 3) / 4;
    const int nblk = (m + 63) / 64;
    v4d e00 = {0, 0, 0, 0}, e01 = {0, 0, 0, 0}, e11 = {0, 0, 0, 0};
#pragma unroll 1
    for (int b = w; b < nblk; b += W) {
        const int r0 = 64 * b;
        const int rows = m - r0 < 64 ? m - r0 : 64;
        const T* Qb = Q + r0;
        const T* Ab = RESID ? A + r0 : nullptr;
        v4d acc[4][2];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = v4d{0, 0, 0, 0};
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) {
            if (kg < nk) {
                const Raw4<T> q = load_raw4<T, VEC>(Qb, ldq, 4 * li, 4 * kg + lq, rows, n);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const double x = (double)q.v[t];
                    acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, mb[kg][0], acc[t][0], 0, 0, 0);
                    if constexpr (NT == 2)
                        acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, mb[kg][1], acc[t][1], 0, 0, 0);
                }
            }
        }
        // e[jt][v][t] = A - P at row 16v + 4 lq + t, column 16 jt + li; zero
        // past the last row and in the padding columns (both terms are zero)
        double e[2][4][4];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if constexpr (RESID) {
                    const Raw4<T> a = load_raw4<T, VEC>(Ab, lda, 16 * v + 4 * lq, 16 * jt + li, rows, n);
#pragma unroll
                    for (int t = 0; t < 4; ++t) e[jt][v][t] = (double)a.v[t] - acc[t][jt][v];
                } else {
#pragma unroll
                    for (int t = 0; t < 4; ++t) e[jt][v][t] = acc[t][jt][v];
                }
            }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                e00 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[0][v][t], e[0][v][t], e00, 0, 0, 0);
                if constexpr (NT == 2) {
                    e01 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[0][v][t], e[1][v][t], e01, 0, 0, 0);
                    e11 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[1][v][t], e[1][v][t], e11, 0, 0, 0);
                }
            }
    }
    stream_reduce<NT>(L, e00, e01, e11, partial, tk, &bcast);
}

// Enqueues the three Gram reductions on st.  On completion the group sums
// of G_A, G_Q and G_E (ngroups = stream_group_count(P) blocks of 32 x 32
// FP64 each, P = stream_groups(m)) are at the returned offsets of work; the
// caller adds each set in order.  n <= 32.
template <typename T, int NT, bool VEC>
int check_launch_n(int m, int n, const T* A, int lda, const T* Q, int ldq, const T* R, int ldr, void* work,
                   size_t* offsets, hipStream_t st) {
    const int P = stream_groups(m);
    const size_t part = stream_partials(m) * 1024;  // doubles per area
    double* base = reinterpret_cast<double*>(work);
    int* tk = reinterpret_cast<int*>(base + 3 * part);
    hipLaunchKernelGGL(zero_ints, dim3((3 * kCtlStreamTickets + 255) / 256), dim3(256), 0, st, tk,
                       3 * kCtlStreamTickets);
    TreeDesc d{};
    d.m = m;
    d.n = n;
    hipLaunchKernelGGL((tsqr_gram_stream<T, NT, VEC>), dim3(P), dim3(256), 0, st, d, A, lda, base, tk);
    hipLaunchKernelGGL((tsqr_gram_stream<T, NT, VEC>), dim3(P), dim3(256), 0, st, d, Q, ldq, base + part,
                       tk + kCtlStreamTickets);
    hipLaunchKernelGGL((tsqr_resid_stream<T, T, NT, VEC, true>), dim3(P), dim3(256), 0, st, m, n, A, lda, Q, ldq, R,
                       1, ldr, base + 2 * part, tk + 2 * kCtlStreamTickets, (const int*)nullptr);
    for (int k = 0; k < 3; ++k) offsets[k] = (k * part + (size_t)P * 1024) * sizeof(double);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

template <typename T>
int check_launch(int m, int n, const T* A, int lda, const T* Q, int ldq, const T* R, int ldr, void* work,
                 size_t* offsets, hipStream_t st) {
    const bool vec = aligned16(A, lda, sizeof(T)) && aligned16(Q, ldq, sizeof(T));
    if (n > 16) {
        if (vec) return check_launch_n<T, 2, true>(m, n, A, lda, Q, ldq, R, ldr, work, offsets, st);
        return check_launch_n<T, 2, false>(m, n, A, lda, Q, ldq, R, ldr, work, offsets, st);
    }
    if (vec) return check_launch_n<T, 1, true>(m, n, A, lda, Q, ldq, R, ldr, work, offsets, st);
    return check_launch_n<T, 1, false>(m, n, A, lda, Q, ldq, R, ldr, work, offsets, st);
}

}  // namespace tsqr
