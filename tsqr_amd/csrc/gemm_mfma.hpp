// gemm_mfma.hpp -- the two coupling products of one recursive QR level on the
// matrix cores (v_mfma_f32_16x16x4f32 for float, v_mfma_f64_16x16x4f64 for
// double; products and sums in the working precision, like the reference;
// BF16x3 split products on v_mfma_f32_16x16x32_bf16 for the mixed-precision
// entry points).
//
// The recursive QR (block_QR.h:72-95) couples its column halves by one pass
// of block classical Gram-Schmidt:
//   R12 = Q1^T A2   (n1 x n2, reduction over all m rows)   block_QR.h:89-90
//   A2 -= Q1 R12    (m x n2, inner dimension n1)           block_QR.h:91-92
// The reference runs both as one thread per output element with a serial k
// loop (block_QR.h:7-60).
//
// Operand layout (both kernels).  A lane loads four consecutive rows of one
// column with one 16-byte load; MFMA number t (t = 0..3) feeds row 4*q + t
// (q = lane / 16, or lane % 16 in gemm_nn) as its k index or its row index,
// so four MFMAs consume 16 rows and every load is a contiguous 16 bytes.
//
// gemm_tn_mfma: grid (chunks, ceil(n1/64), ceil(n2/64)), one wave.  Block
//   (k, bi, bj) sums rows [k*kc, (k+1)*kc) of the 64 x 64 output tile
//   (bi, bj) into part[k][i + j*n1]; gemm_tn_reduce (below) adds the chunks
//   in a fixed order, so the result is deterministic.
// gemm_nn_mfma: grid (ceil(m/64), ceil(n2/64)), one wave.  The 64 x 64 block
//   of C is loaded into the accumulators, B is streamed 16 rows at a time in
//   the B-operand layout, and C -= X B is written back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bf16x3.hpp"
#include "wave.hpp"

namespace tsqr {

template <typename T>
struct Mfma16;

template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator register v of lane (lq = lane/16) holds output row row(v, lq)
    static __device__ __forceinline__ int row(int v, int lq) { return 4 * lq + v; }
};

template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lq) { return 4 * v + lq; }
};

template <typename T>
struct Four {
    T v[4];
};

// Rows [r, r+4) of column c of X (column-major, ld), zero outside rows < m
// and columns < ncols; VEC: 16-byte aligned base and ld.
template <typename T, bool VEC>
__device__ __forceinline__ Four<T> load_four(const T* __restrict__ X, int64_t ld, int64_t r, int c, int64_t m,
                                             int ncols) {
    Four<T> o;
    if (c < ncols) {
        const T* p = X + (int64_t)c * ld + r;
        if (VEC && r + 4 <= m) {
            if constexpr (sizeof(T) == 4) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                const f4 q = *reinterpret_cast<const f4*>(p);
#pragma unroll
                for (int t = 0; t < 4; ++t) o.v[t] = q[t];
            } else {
                typedef double d2 __attribute__((ext_vector_type(2)));
                const d2 q0 = *reinterpret_cast<const d2*>(p);
                const d2 q1 = *reinterpret_cast<const d2*>(p + 2);
                o.v[0] = q0[0];
                o.v[1] = q0[1];
                o.v[2] = q1[0];
                o.v[3] = q1[1];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) o.v[t] = (r + t < m) ? p[t] : T(0);
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) o.v[t] = T(0);
    }
    return o;
}

template <typename T, bool VEC>
__device__ __forceinline__ void store_four(const Four<T>& o, T* __restrict__ X, int64_t ld, int64_t r, int c,
                                           int64_t m, int ncols) {
    if (c >= ncols) return;
    T* p = X + (int64_t)c * ld + r;
    if (VEC && r + 4 <= m) {
        if constexpr (sizeof(T) == 4) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            *reinterpret_cast<f4*>(p) = f4{o.v[0], o.v[1], o.v[2], o.v[3]};
        } else {
            typedef double d2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<d2*>(p) = d2{o.v[0], o.v[1]};
            *reinterpret_cast<d2*>(p + 2) = d2{o.v[2], o.v[3]};
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (r + t < m) p[t] = o.v[t];
    }
}

// A 64 x 64 FP32 block of C (rows r0.., columns j0..) moves between HBM and
// the accumulator layout of the update kernels through the wave's LDS tile
// `tw` (column c, row r at c * kCTilePitch + r): HBM sees 16-byte column runs
// (lane l: rows 4(l%16) .. +3 of column 4p + l/16, p = 0..15, 256 contiguous
// bytes per column and instruction); the accumulator register v of (t, jt)
// holds row 16(l/16) + 4v + t, column 16 jt + l%16, and each lane reads and
// writes the same tile positions.
constexpr int kCTilePitch = 66;  // floats per LDS column of the tile
constexpr int kCTileFloats = 64 * kCTilePitch;

template <bool VEC>
__device__ __forceinline__ void c_tile_load(f4v (&acc)[4][4], float* tw, const float* __restrict__ C, int64_t ldc,
                                            int64_t r0, int j0, int64_t m, int N, int lane) {
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const Four<float> c = load_four<float, VEC>(C, ldc, r0 + 4 * li, j0 + 4 * p + lq, m, N);
#pragma unroll
        for (int t = 0; t < 4; ++t) tw[(4 * p + lq) * kCTilePitch + 4 * li + t] = c.v[t];
    }
    wave_sync();
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t][jt][v] = tw[(16 * jt + li) * kCTilePitch + 16 * lq + 4 * v + t];
}

template <bool VEC>
__device__ __forceinline__ void c_tile_store(const f4v (&acc)[4][4], float* tw, float* C, int64_t ldc, int64_t r0,
                                             int j0, int64_t m, int N, int lane) {
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int t = 0; t < 4; ++t) tw[(16 * jt + li) * kCTilePitch + 16 * lq + 4 * v + t] = acc[t][jt][v];
    wave_sync();
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        Four<float> c;
#pragma unroll
        for (int t = 0; t < 4; ++t) c.v[t] = tw[(4 * p + lq) * kCTilePitch + 4 * li + t];
        store_four<float, VEC>(c, C, ldc, r0 + 4 * li, j0 + 4 * p + lq, m, N);
    }
}

// part[chunk] (n1 x n2, ld n1) for the 64 x 64 output tile (blockIdx.y,
// blockIdx.z) = Q[rows of chunk, cols i]^T A[rows of chunk, cols j].
template <typename T, bool VEC>
__global__ __launch_bounds__(64) void gemm_tn_mfma(int m, int n1, int n2, const T* __restrict__ Q, int64_t ldq,
                                                   const T* __restrict__ A, int64_t lda, int kc,
                                                   T* __restrict__ part) {
    using M = Mfma16<T>;
    const int lane = threadIdx.x;
    const int li = lane & 15, lq = lane >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.z * 64;
    const int64_t kbeg = (int64_t)blockIdx.x * kc;
    const int64_t kend = kbeg + kc < m ? kbeg + kc : m;
    typename M::v4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = typename M::v4{0, 0, 0, 0};
    // the next 16-row step's loads are issued before this step's MFMAs (a
    // launch has few waves per SIMD, so each wave hides its own latency)
    Four<T> q[4], x[4];
    if (kbeg < kend) {
#pragma unroll
        for (int a = 0; a < 4; ++a) q[a] = load_four<T, VEC>(Q, ldq, kbeg + 4 * lq, i0 + 16 * a + li, kend, n1);
#pragma unroll
        for (int b = 0; b < 4; ++b) x[b] = load_four<T, VEC>(A, lda, kbeg + 4 * lq, j0 + 16 * b + li, kend, n2);
    }
#pragma unroll 1
    for (int64_t r0 = kbeg; r0 < kend; r0 += 16) {
        const int64_t rn = r0 + 16 + 4 * lq;
        const bool more = r0 + 16 < kend;
        Four<T> qn[4], xn[4];
        if (more) {
#pragma unroll
            for (int a = 0; a < 4; ++a) qn[a] = load_four<T, VEC>(Q, ldq, rn, i0 + 16 * a + li, kend, n1);
#pragma unroll
            for (int b = 0; b < 4; ++b) xn[b] = load_four<T, VEC>(A, lda, rn, j0 + 16 * b + li, kend, n2);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = M::mma(q[a].v[t], x[b].v[t], acc[a][b]);
        if (!more) break;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            q[a] = qn[a];
            x[a] = xn[a];
        }
    }
    T* out = part + (int64_t)blockIdx.x * n1 * n2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + 16 * a + M::row(v, lq), j = j0 + 16 * b + li;
                if (i < n1 && j < n2) out[i + (int64_t)j * n1] = acc[a][b][v];
            }
}

// C[i + j*ldc] = (accumulate ? C : 0) + sum over chunks of
// part[k][i + j*n1], summed and written in TO (T, or double for the
// reorthogonalisation Gram matrix).  256 threads per 16 outputs: thread
// (o, g) adds the chunks k = g, g + 16, ... of output o in order, then the
// 16 group sums are added in order of g, so the result is deterministic and
// no thread walks more than chunks / 16 partials.  Eight partials are
// loaded before they are added (the same order), so that the loads overlap.
// sym: the partials hold only the 64 x 64 tiles on and above the diagonal
// (gemm_tn_bf16x3 with sym), and an entry of a tile below it is read from
// the transposed position.
template <typename T, typename TO = T>
__global__ __launch_bounds__(256) void gemm_tn_reduce(int n1, int n2, int chunks, const T* __restrict__ part,
                                                      TO* __restrict__ C, int64_t ldc, int accumulate, int sym = 0) {
    // sums in FP64 whenever the partials or the output are FP64
    typedef typename std::conditional<sizeof(T) == 8 || sizeof(TO) == 8, double, T>::type TS;
    __shared__ TS red[16][17];
    const int o = threadIdx.x & 15, g = threadIdx.x >> 4;
    const int out = blockIdx.x * 16 + o;
    const int total = n1 * n2;
    int idx = out;
    if (sym && out < total && (out % n1) / 64 > (out / n1) / 64) idx = (out / n1) + (out % n1) * n1;
    TS s = TS(0);
    if (idx < total) {
        // eight chunks in flight per step, the last step's too (its loads
        // past the end reread chunk g and add zero)
#pragma unroll 1
        for (int k = g; k < chunks; k += 8 * 16) {
            T v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int kq = k + 16 * q < chunks ? k + 16 * q : g;
                v[q] = part[(int64_t)kq * total + idx];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) s += k + 16 * q < chunks ? (TS)v[q] : TS(0);
        }
    }
    red[o][g] = s;
    __syncthreads();
    if (g != 0 || out >= total) return;
    TS sum = TS(0);
#pragma unroll
    for (int q = 0; q < 16; ++q) sum += red[o][q];
    const int i = out % n1, j = out / n1;
    TO* c = C + i + (int64_t)j * ldc;
    *c = (TO)(accumulate ? (TS)*c + sum : sum);
}

// C[rows, j0:j0+64] -= X[rows, 0:K] B[0:K, j0:j0+64] for the 64-row block
// blockIdx.x and the column group blockIdx.y.  X: m x K (ld ldx), B: K x N
// (ld ldb), C: m x N (ld ldc), all column-major.
template <typename T, bool VEC>
__global__ __launch_bounds__(64) void gemm_nn_mfma(int m, int K, int N, const T* __restrict__ X, int64_t ldx,
                                                   const T* __restrict__ B, int64_t ldb, T* __restrict__ C,
                                                   int64_t ldc) {
    using M = Mfma16<T>;
    __shared__ float tile[sizeof(T) == 4 ? kCTileFloats : 1];
    const int lane = threadIdx.x;
    const int li = lane & 15, lq = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int j0 = blockIdx.y * 64;
    // acc[t][jt] register v: row r0 + 4*row(v, lq) + t, column j0 + 16 jt + li
    typename M::v4 acc[4][4];
    if constexpr (sizeof(T) == 4) {
        c_tile_load<VEC>(acc, tile, C, ldc, r0, j0, m, N, lane);  // row(v, lq) = 4 lq + v
    } else {
#pragma unroll
        for (int jt = 0; jt < 4; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const Four<T> c = load_four<T, VEC>(C, ldc, r0 + 4 * M::row(v, lq), j0 + 16 * jt + li, m, N);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t][jt][v] = c.v[t];
            }
    }
#pragma unroll 1
    for (int k0 = 0; k0 < K; k0 += 16) {
        // A operand: lane feeds row 4*li + t of column k0 + 4*kg + lq
        Four<T> xa[4];
        T bb[4][4];  // B[k0 + 4 kg + lq][j0 + 16 jt + li], negated
#pragma unroll
        for (int kg = 0; kg < 4; ++kg) {
            const int k = k0 + 4 * kg + lq;
            xa[kg] = load_four<T, VEC>(X, ldx, r0 + 4 * li, k, m, K);
#pragma unroll
            for (int jt = 0; jt < 4; ++jt) {
                const int j = j0 + 16 * jt + li;
                bb[kg][jt] = (k < K && j < N) ? -B[k + (int64_t)j * ldb] : T(0);
            }
        }
#pragma unroll
        for (int kg = 0; kg < 4; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int jt = 0; jt < 4; ++jt) acc[t][jt] = M::mma(xa[kg].v[t], bb[kg][jt], acc[t][jt]);
    }
    if constexpr (sizeof(T) == 4) {
        c_tile_store<VEC>(acc, tile, C, ldc, r0, j0, m, N, lane);
    } else {
#pragma unroll
        for (int jt = 0; jt < 4; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Four<T> c;
#pragma unroll
                for (int t = 0; t < 4; ++t) c.v[t] = acc[t][jt][v];
                store_four<T, VEC>(c, C, ldc, r0 + 4 * M::row(v, lq), j0 + 16 * jt + li, m, N);
            }
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision coupling (FP32 data, the qr_mixed entry points): the same
// two products on the BF16 matrix cores with BF16x3 split operands
// (bf16x3.hpp), FP32 accumulation, 32 rows (tn) or 32 inner columns (nn)
// per v_mfma_f32_16x16x32_bf16.  Grids and outputs are those of
// gemm_tn_mfma / gemm_nn_mfma.

// Rows [r, r+8) of column c of X, zero outside rows < m and columns < ncols.
template <bool VEC>
__device__ __forceinline__ void load_eight(float (&x)[8], const float* __restrict__ X, int64_t ld, int64_t r, int c,
                                           int64_t m, int ncols) {
    const Four<float> a = load_four<float, VEC>(X, ld, r, c, m, ncols);
    const Four<float> b = load_four<float, VEC>(X, ld, r + 4, c, m, ncols);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        x[t] = a.v[t];
        x[4 + t] = b.v[t];
    }
}

// part[chunk] for the 64 x 64 tile (blockIdx.y, blockIdx.z) of Q^T A: lane
// (li, lq) feeds rows 8 lq .. 8 lq + 7 of each 32-row step, column li of the
// tile's 16-column groups, as both operands.
// sym (Q == A, n1 == n2, the Gram matrix Q^T Q): tiles below the diagonal
// are not computed (gemm_tn_reduce mirrors them) and a diagonal tile loads
// and splits its 64 columns once for both operands.
// TP = double (the reorthogonalisation Gram matrix): each 32-row step is
// summed in FP32 from zero and added to FP64 sums, and the partials are
// FP64.  Long FP32 sums of positive terms (the diagonal of a Gram matrix)
// drift low on the matrix cores -- about 5e-7 relative over 1024 rows was
// measured on MI355X -- which the per-step FP64 sums avoid.
template <bool VEC, typename TP = float>
__global__ __launch_bounds__(64) void gemm_tn_bf16x3(int m, int n1, int n2, const float* __restrict__ Q, int64_t ldq,
                                                     const float* __restrict__ A, int64_t lda, int kc,
                                                     TP* __restrict__ part, int sym = 0) {
    constexpr bool F64 = sizeof(TP) == 8;
    if (sym && blockIdx.y > blockIdx.z) return;
    const bool same = sym && blockIdx.y == blockIdx.z;
    const int lane = threadIdx.x;
    const int li = lane & 15, lq = lane >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.z * 64;
    const int64_t kbeg = (int64_t)blockIdx.x * kc;
    const int64_t kend = kbeg + kc < m ? kbeg + kc : m;
    f4v acc[4][4];
    double acc64[F64 ? 4 : 1][F64 ? 4 : 1][F64 ? 4 : 1];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4v{0, 0, 0, 0};
    if constexpr (F64) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc64[a][b][v] = 0.0;
    }
    // the next 32-row step's loads are issued before this step's splits and
    // MFMAs (few waves per SIMD: each wave hides its own latency)
    float vq[4][8], vx[4][8];
    if (kbeg < kend) {
#pragma unroll
        for (int a = 0; a < 4; ++a) load_eight<VEC>(vq[a], Q, ldq, kbeg + 8 * lq, i0 + 16 * a + li, kend, n1);
        if (!same) {
#pragma unroll
            for (int b = 0; b < 4; ++b) load_eight<VEC>(vx[b], A, lda, kbeg + 8 * lq, j0 + 16 * b + li, kend, n2);
        }
    }
#pragma unroll 1
    for (int64_t r0 = kbeg; r0 < kend; r0 += 32) {
        const int64_t rn = r0 + 32 + 8 * lq;
        const bool more = r0 + 32 < kend;
        Split3 sq[4], sx[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            sq[a] = split3(vq[a]);
            sx[a] = same ? sq[a] : split3(vx[a]);
        }
        if (more) {
#pragma unroll
            for (int a = 0; a < 4; ++a) load_eight<VEC>(vq[a], Q, ldq, rn, i0 + 16 * a + li, kend, n1);
            if (!same) {
#pragma unroll
                for (int b = 0; b < 4; ++b) load_eight<VEC>(vx[b], A, lda, rn, j0 + 16 * b + li, kend, n2);
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if constexpr (F64) {
                    const f4v step = mma_bf16x3(sq[a], sx[b], f4v{0, 0, 0, 0});
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc64[a][b][v] += (double)step[v];
                } else {
                    acc[a][b] = mma_bf16x3(sq[a], sx[b], acc[a][b]);
                }
            }
    }
    TP* out = part + (int64_t)blockIdx.x * n1 * n2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + 16 * a + 4 * lq + v, j = j0 + 16 * b + li;
                if (i < n1 && j < n2) {
                    if constexpr (F64)
                        out[i + (int64_t)j * n1] = acc64[a][b][v];
                    else
                        out[i + (int64_t)j * n1] = acc[a][b][v];
                }
            }
}

// C[rows, j0:j0+64] -= X B for the 64-row block blockIdx.x, as gemm_nn_mfma.
// MFMA t (t = 0..3) takes rows r0 + 4 i + t (i = 0..15) of X as its A
// operand: lane (li, lq) loads rows r0 + 4 li .. +3 of the eight columns
// k0 + 8 lq + s, s = 0..7, so every load is 16 contiguous bytes.  Its output
// register v is row r0 + 16 lq + 4 v + t, column j0 + 16 jt + li.
// The body takes plain pointers, so that X and C may be the same array (the
// in-place reorthogonalisation, reorth_apply): every load of the wave comes
// before its stores.
template <bool VEC>
__device__ __forceinline__ void nn_bf16x3_block(int m, int K, int N, const float* X, int64_t ldx, const float* B,
                                                int64_t ldb, float* C, int64_t ldc, int64_t r0, int j0, float* tw) {
    const int lane = threadIdx.x;
    const int li = lane & 15, lq = lane >> 4;
    f4v acc[4][4];  // [t][jt]
    c_tile_load<VEC>(acc, tw, C, ldc, r0, j0, m, N, lane);
#pragma unroll 1
    for (int k0 = 0; k0 < K; k0 += 32) {
        Split3 sx[4], sb[4];
        {
            float xv[4][8];  // [t][s]: X[r0 + 4 li + t][k0 + 8 lq + s]
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const Four<float> f = load_four<float, VEC>(X, ldx, r0 + 4 * li, k0 + 8 * lq + s, m, K);
#pragma unroll
                for (int t = 0; t < 4; ++t) xv[t][s] = f.v[t];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) sx[t] = split3(xv[t]);
        }
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) {
            const int j = j0 + 16 * jt + li;
            float bv[8];  // -B[k0 + 8 lq + s][j]
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int k = k0 + 8 * lq + s;
                bv[s] = (k < K && j < N) ? -B[k + (int64_t)j * ldb] : 0.f;
            }
            sb[jt] = split3(bv);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int jt = 0; jt < 4; ++jt) acc[t][jt] = mma_bf16x3(sx[t], sb[jt], acc[t][jt]);
    }
    c_tile_store<VEC>(acc, tw, C, ldc, r0, j0, m, N, lane);
}

template <bool VEC>
__global__ __launch_bounds__(64) void gemm_nn_bf16x3(int m, int K, int N, const float* __restrict__ X, int64_t ldx,
                                                     const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                     int64_t ldc) {
    __shared__ float tile[kCTileFloats];
    nn_bf16x3_block<VEC>(m, K, N, X, ldx, B, ldb, C, ldc, (int64_t)blockIdx.x * 64, blockIdx.y * 64, tile);
}

// ---------------------------------------------------------------------------
// Whole-matrix reorthogonalisation of the mixed-precision recursive QR (one
// CholeskyQR pass over the finished Q, n <= kReorthMax columns; no
// counterpart in the reference, whose qr couples once, block_QR.h:88-93).
// The coupled column blocks of the recursion lose orthogonality to each
// other in proportion to cond(R22) times the accuracy of the coupling
// products; one CholeskyQR pass over the nearly orthonormal Q restores it
// independently of cond(A):
//   G = Q^T Q (BF16x3 products, each 32-row step added to FP64 sums),
//   e = ||G - I||_F;
//   if e > tol: R_c = chol(G), Q <- Q R_c^{-1} = Q - Q Dneg with
//   Dneg = I - R_c^{-1}, and R <- R_c R.
// Dneg is small, so holding it in FP32 keeps Q's own bits in the sum
// (rounding R_c^{-1} itself to FP32 would perturb every column's norm by
// u32), and its BF16x3 products on the matrix cores are accurate far beyond
// what the sum keeps.  Kernels: reorth_chol (one workgroup: e, the
// decision, R_c in FP64), reorth_inv (one workgroup per column of Dneg),
// reorth_rupdate (one workgroup per column of R), reorth_apply (one wave per
// 64 rows of Q and 64 columns).  Each does nothing unless reorth_chol set
// st[0].
constexpr int kReorthMax = 128;
constexpr int kReorthThreads = 256;

// G: n x n FP64 (ld n).  On e > tol and a successful Cholesky: Rc (n x n
// FP64, ld n, zeros below the diagonal) = chol(G) and st[0] = 1; otherwise
// st[0] = 0 (st[1] = 2 if the Cholesky broke down).  *e_out = e.
template <int UNUSED = 0>
__global__ __launch_bounds__(kReorthThreads) void reorth_chol(int n, const double* __restrict__ G,
                                                             double* __restrict__ Rc, int* __restrict__ st,
                                                             double* __restrict__ e_out, double tol) {
    __shared__ double L[kReorthMax][kReorthMax + 1];
    __shared__ double red[kReorthThreads];
    __shared__ int ok;
    const int t = threadIdx.x;
    double s = 0.0;
    for (int e = t; e < n * n; e += kReorthThreads) {
        const int i = e % n, j = e / n;
        const double g = G[e];
        L[i][j] = g;
        const double d = g - (i == j ? 1.0 : 0.0);
        s += d * d;
    }
    red[t] = s;
    __syncthreads();
    for (int w = kReorthThreads / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double err = __builtin_sqrt(red[0]);
    if (t == 0) {
        *e_out = err;
        st[0] = 0;
        st[1] = 0;
        ok = 1;
    }
    if (!(err > tol)) return;  // orthonormal enough (or NaN): Q and R are kept
    __syncthreads();
    // right-looking Cholesky, upper factor in place: row j of R_c, then the
    // trailing upper triangle
    for (int j = 0; j < n; ++j) {
        if (t == 0) {
            const double p = L[j][j];
            if (!(p > 0.0) || !(p < __builtin_inf())) ok = 0;
            L[j][j] = ok ? __builtin_sqrt(p) : 1.0;
        }
        __syncthreads();
        const double dj = L[j][j];
        for (int c = j + 1 + t; c < n; c += kReorthThreads) L[j][c] /= dj;
        __syncthreads();
        // trailing update over a 16 x 16 grid of threads (no index division)
        for (int r = j + 1 + (t >> 4); r < n; r += 16) {
            const double ljr = L[j][r];
            for (int c = r + ((j + 1 + (t & 15) - r) & 15); c < n; c += 16) L[r][c] -= ljr * L[j][c];
        }
        __syncthreads();
    }
    if (!ok) {
        if (t == 0) st[1] = 2;
        return;
    }
    for (int e = t; e < n * n; e += kReorthThreads) {
        const int i = e % n, j = e / n;
        Rc[e] = i <= j ? L[i][j] : 0.0;
    }
    if (t == 0) st[0] = 1;
}

// Column c = blockIdx.x of Dneg = I - R_c^{-1} (FP32, n x n, ld n): R_c x = e_c
// by column-oriented back substitution (x_i /= R_ii, then x_k -= R_ki x_i
// for k < i, in parallel over k), in FP64.
template <int UNUSED = 0>
__global__ __launch_bounds__(kReorthMax) void reorth_inv(int n, const double* __restrict__ Rc, float* __restrict__ D,
                                                         const int* __restrict__ st) {
    __shared__ double x[kReorthMax];
    if (!st[0]) return;
    const int c = blockIdx.x, t = threadIdx.x;
    if (t < n) x[t] = t == c ? 1.0 : 0.0;
    __syncthreads();
    for (int i = c; i >= 0; --i) {
        if (t == 0) x[i] /= Rc[i + (int64_t)i * n];
        __syncthreads();
        const double xi = x[i];
        if (t < i) x[t] -= Rc[t + (int64_t)i * n] * xi;
        __syncthreads();
    }
    if (t < n) D[t + (int64_t)c * n] = t <= c ? (float)((t == c ? 1.0 : 0.0) - x[t]) : 0.f;
}

// Column j = blockIdx.x of R <- R_c R (R upper triangular, ld ldr), in FP64.
template <typename T>
__global__ __launch_bounds__(kReorthMax) void reorth_rupdate(int n, const double* __restrict__ Rc, T* __restrict__ R,
                                                             int64_t ldr, const int* __restrict__ st) {
    __shared__ double col[kReorthMax];
    if (!st[0]) return;
    const int j = blockIdx.x, t = threadIdx.x;
    if (t < n) col[t] = t <= j ? (double)R[t + (int64_t)j * ldr] : 0.0;
    __syncthreads();
    if (t <= j && t < n) {
        double acc = 0.0;
        for (int k = t; k <= j; ++k) acc += Rc[t + (int64_t)k * n] * col[k];
        R[t + (int64_t)j * ldr] = (T)acc;
    }
}

// Columns [c0, c0 + N) of Q (m x *, ld ldq) += Q[:, 0:K] D[0:K, c0:c0 + N] in
// place, if st[0] is set, one wave per 64 rows: the BF16x3 product of
// gemm_nn_bf16x3 with B = Dneg = I - R_c^{-1} (C -= X Dneg).  D is upper
// triangular, so output columns [c0, c0 + N) read input columns < c0 + N
// only: the column groups are launched from the last to the first, and no
// launch reads a column an earlier one wrote.
template <bool VEC>
__global__ __launch_bounds__(64) void reorth_apply(int m, int K, int N, float* Q, int64_t ldq, int c0,
                                                   const float* __restrict__ Dneg, int64_t ldd,
                                                   const int* __restrict__ st) {
    __shared__ float tile[kCTileFloats];
    if (!st[0]) return;
    nn_bf16x3_block<VEC>(m, K, N, Q, ldq, Dneg + (int64_t)c0 * ldd, ldd, Q + (int64_t)c0 * ldq, ldq,
                         (int64_t)blockIdx.x * 64, 0, tile);
}

// ---------------------------------------------------------------------------
// Small triangular solves of the least-squares entry points (n <= 32): for
// column j < k of C (n x k, ldc), y = R^{-T} C[:, j] (forward substitution
// with the upper triangular R, n x n, ldr) and, unless transpose_only,
// x = R^{-1} y (back substitution), in FP64, one thread per column.  The
// result is written to X[:, j] (ldx), or added to it if accumulate.
// R is staged in LDS; the loops run to 32 with guards so the vectors stay
// in registers.
template <typename T>
__global__ __launch_bounds__(256) void tri_normal_solve(int n, int k, const T* __restrict__ R, int64_t ldr,
                                                        const T* __restrict__ C, int64_t ldc, T* __restrict__ X,
                                                        int64_t ldx, int accumulate, int transpose_only) {
    __shared__ double Rs[32][33];
    for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        const int i = e & 31, j = e >> 5;
        Rs[i][j] = (i < n && j < n && i <= j) ? (double)R[i + (int64_t)j * ldr] : 0.0;
    }
    __syncthreads();
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= k) return;
    double y[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) y[i] = i < n ? (double)C[i + (int64_t)col * ldc] : 0.0;
    // R^T y' = y: row i of R^T is column i of R
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        if (i < n) {
            double s = y[i];
#pragma unroll
            for (int p = 0; p < i; ++p) s -= Rs[p][i] * y[p];
            y[i] = s / Rs[i][i];
        }
    }
    if (!transpose_only) {
#pragma unroll
        for (int i = 31; i >= 0; --i) {
            if (i < n) {
                double s = y[i];
#pragma unroll
                for (int p = i + 1; p < 32; ++p) s -= Rs[i][p] * y[p];
                y[i] = s / Rs[i][i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        if (i < n) {
            T* x = X + i + (int64_t)col * ldx;
            *x = accumulate ? (T)((double)*x + y[i]) : (T)y[i];
        }
    }
}

// Zeroes the entries below the diagonal of the n x n column-major M.
template <typename T>
__global__ void zero_lower(T* __restrict__ M, int64_t ld, int n) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int r = (int)(idx % n), c = (int)(idx / n);
    if (r > c) M[r + (int64_t)c * ld] = T(0);
}

// dst[r + c*ldd] += src[r + c*lds] for an nr x nc block.
template <typename T>
__global__ void add_block(const T* __restrict__ src, int lds, T* __restrict__ dst, int64_t ldd, int nr, int nc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nr * nc) return;
    const int r = idx % nr, c = idx / nr;
    dst[r + (int64_t)c * ldd] += src[r + (int64_t)c * lds];
}

}  // namespace tsqr
