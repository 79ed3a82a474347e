// gemm_valu.hpp -- the two products of one recursive QR level on VALU.
//
// The recursive QR (block_QR.h:72-95) couples its column halves by one pass
// of block classical Gram-Schmidt:
//   R12 = Q1^T A2   (n1 x n2, reduction over all m rows)   block_QR.h:89-90
//   A2 -= Q1 R12    (m x n2, inner dimension n1)           block_QR.h:91-92
// The reference runs each as one thread per output element with a serial
// k loop (block_QR.h:7-60).  These are LDS-tiled VALU kernels, kept as the
// fallback and cross-check of the MFMA kernels.
//
// gemm_tn_partial: grid (chunks, ceil(n1/64), ceil(n2/64)), 256 threads.
//   Block (k, bi, bj) sums rows [k*kc, (k+1)*kc) of the 64x64 output tile
//   (bi, bj) into part[k][i + j*n1]; gemm_tn_reduce adds the chunks in a
//   fixed order, so the result is deterministic.
// gemm_nn_update: grid (ceil(m/256), ceil(n2/64)), 256 threads, one row per
//   thread, R12 staged in LDS 32 rows at a time.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wave.hpp"

namespace tsqr {

template <typename T>
__global__ __launch_bounds__(256) void gemm_tn_partial(int m, int n1, int n2, const T* __restrict__ Q, int64_t ldq,
                                                       const T* __restrict__ A, int64_t lda, int kc,
                                                       T* __restrict__ part) {
    __shared__ T sQ[32][64 + 1];
    __shared__ T sA[32][64 + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.z * 64;
    const int64_t kbeg = (int64_t)blockIdx.x * kc;
    const int64_t kend = kbeg + kc < m ? kbeg + kc : m;
    T acc[4][4] = {};
    for (int64_t k0 = kbeg; k0 < kend; k0 += 32) {
        const int kk = tid & 31;
        const int64_t row = k0 + kk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int col = (tid >> 5) + 8 * q;
            const bool rv = row < kend;
            sQ[kk][col] = (rv && i0 + col < n1) ? Q[row + (int64_t)(i0 + col) * ldq] : T(0);
            sA[kk][col] = (rv && j0 + col < n2) ? A[row + (int64_t)(j0 + col) * lda] : T(0);
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < 32; ++k) {
            T qa[4], ab[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qa[t] = sQ[k][4 * ty + t];
                ab[t] = sA[k][4 * tx + t];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fmat(qa[a], ab[b], acc[a][b]);
        }
        __syncthreads();
    }
    T* out = part + (int64_t)blockIdx.x * n1 * n2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + 4 * ty + a, j = j0 + 4 * tx + b;
            if (i < n1 && j < n2) out[i + (int64_t)j * n1] = acc[a][b];
        }
}

// C[i + j*ldc] = sum over chunks of part[k][i + j*n1], chunks in order.
template <typename T>
__global__ void gemm_tn_reduce(int n1, int n2, int chunks, const T* __restrict__ part, T* __restrict__ C,
                               int64_t ldc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n1 * n2) return;
    T s = T(0);
    for (int k = 0; k < chunks; ++k) s += part[(int64_t)k * n1 * n2 + idx];
    const int i = idx % n1, j = idx / n1;
    C[i + (int64_t)j * ldc] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void gemm_nn_update(int m, int n1, int n2, const T* __restrict__ Q, int64_t ldq,
                                                      const T* __restrict__ Rm, int64_t ldr, T* __restrict__ A,
                                                      int64_t lda) {
    __shared__ T sR[32][64];
    const int tid = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 256 + tid;
    const int j0 = blockIdx.y * 64;
    const int nj = n2 - j0 < 64 ? n2 - j0 : 64;
    T acc[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) acc[j] = T(0);
    for (int k0 = 0; k0 < n1; k0 += 32) {
        for (int e = tid; e < 32 * 64; e += 256) {
            const int kk = e & 31, jj = e >> 5;
            sR[kk][jj] = (k0 + kk < n1 && jj < nj) ? Rm[(k0 + kk) + (int64_t)(j0 + jj) * ldr] : T(0);
        }
        __syncthreads();
        if (row < m) {
#pragma unroll 4
            for (int kk = 0; kk < 32; ++kk) {
                const T q = (k0 + kk < n1) ? Q[row + (int64_t)(k0 + kk) * ldq] : T(0);
#pragma unroll
                for (int j = 0; j < 64; ++j) acc[j] = fmat(q, sR[kk][j], acc[j]);
            }
        }
        __syncthreads();
    }
    if (row < m) {
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if (j < nj) A[row + (int64_t)(j0 + j) * lda] -= acc[j];
    }
}

// Zeroes rows [r0, r0+nr) x columns [c0, c0+nc) of a column-major matrix.
template <typename T>
__global__ void zero_block(T* __restrict__ M, int64_t ld, int r0, int nr, int c0, int nc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nr * nc) return;
    const int r = idx % nr, c = idx / nr;
    M[(r0 + r) + (int64_t)(c0 + c) * ld] = T(0);
}

}  // namespace tsqr
