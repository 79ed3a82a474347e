// mixed.hpp -- the mixed-precision TSQR path (FP32 in and out, n <= 32).
//
// The Gram matrices run on the BF16 matrix cores instead of FP64 ones.  An
// FP32 value x is split exactly into three BF16 pieces
//   x = h + m + l,  h = bf16(x), m = bf16(x - h), l = x - h - m
// (RNE each time; |m| <= 2^-9 |x|, |l| <= 2^-18 |x|, and l is exact in BF16
// because x - h - m has at most 8 significant bits).  A product x*y is then
// the six leading terms
//   hh + hm + mh + hl + lh + mm          (dropped: ml, lm, ll < 2^-26 |xy|)
// each a BF16 x BF16 product, exact in FP32.  v_mfma_f32_16x16x32_bf16 sums
// 32 rows of such products into FP32 accumulators; every 64 rows the FP32
// partials are added into FP64 sums (in LDS), so FP32 rounding acts only
// within 64-row partial sums.  The Gram matrices are therefore accurate to
// about u32 per 64-row partial, averaging down with the row count: the
// residual check reaches ~1e-7 for m >= ~10^4 rows but only ~5e-7 for a few
// hundred rows at n = 32.
//
// Method (one flat CholeskyQR with an optional second pass, CholeskyQR2):
//   pass 1  G1 = A^T A (BF16x3, above), streamed over the whole matrix
//           (tsqr_mixed_stream), the 64-row partials summed in FP64;
//           R1 = chol(G1) and kappa_est = ||R1||_F ||R1^{-1}||_F <= gate in
//           FP64 (tsqr_gram_flat); Q1 = A R1^{-1} streamed (tsqr_mixed_qs):
//           BF16x3 products summed in FP32 where kappa_est <= 2n (R1^{-1}
//           well conditioned), else FP64 MFMA products rounded once; and in
//           the same pass G2 = Q1^T Q1 of the stored Q1 (BF16x3 again).
//   check   e1 = ||G2 - I||_F (the orthogonality of the stored Q1).  If
//           e1 <= tol the result is Q1, R1 (tsqr_mixed_check).
//   pass 2  otherwise R2 = chol(G2), Q = Q1 R2^{-1} (the same kernel, on
//           BF16x3 where e1 <= kBf16QMaxResidual, with G3 = Q^T Q fused
//           again), R = R2 R1 in FP64; e2 = ||G3 - I||_F is recorded and
//           status 1 is set if e2 > tol.
// If the first Cholesky breaks down or kappa_est exceeds the gate, pass 1
// is redone on the FP64 Gram of the AUTO path (exact products: the flat
// R1 holds to kappa_est <= kMixedGram64Gate, with Q1 formed from FP64
// products and the second pass restoring the orthogonality).  Only where
// that fails too does the whole matrix take the Householder tree (the leaf
// and tree kernels, driven by the same ctl word), so no input is rejected.
//
// ctl words (int): ctl[0] = 1 flat Gram taken, ctl[1] = 1 second pass
// requested, ctl[2] = status (0 ok, 1 residual above tol after the second
// pass, 2 second Cholesky broke down), ctl[3] / ctl[4] = 1 where pass 1 /
// pass 2 forms Q on the BF16 matrix cores.  The FP64 side block T0+2 holds
// e1, e2.
#pragma once
#include "bf16x3.hpp"
#include "gram.hpp"

namespace tsqr {

// Symmetric 32 x 32 Gram accumulator of one wave: blocks (0,0), (0,1),
// (1,1) of 16 x 16.  FP32 partials c** live in registers; flush() adds them
// into FP64 sums kept in LDS (12 doubles per lane, `sums` points at the
// wave's 12 x 64 slab), which keeps the streaming kernels' register count
// at two waves per SIMD.
// f32 accumulator layout: lane l, register v holds C[4(l/16) + v][l%16].
struct GramAcc32 {
    f4v c00, c01, c11;
    double* sums;  // LDS, [12][64]
    __device__ __forceinline__ void zero(int lane) {
        c00 = c01 = c11 = f4v{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 12; ++r) sums[r * 64 + lane] = 0.0;
    }
    // x0: column l%16, x1: column 16 + l%16, both at this lane's 8 k-slots
    template <int NT>
    __device__ __forceinline__ void add(const float (&x0)[8], const float (&x1)[8]) {
        const Split3 a0 = split3(x0);
        c00 = mma_bf16x3(a0, a0, c00);
        if constexpr (NT == 2) {
            const Split3 a1 = split3(x1);
            c01 = mma_bf16x3(a0, a1, c01);
            c11 = mma_bf16x3(a1, a1, c11);
        }
    }
    template <int NT>
    __device__ __forceinline__ void flush(int lane) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            sums[v * 64 + lane] += (double)c00[v];
            if constexpr (NT == 2) {
                sums[(4 + v) * 64 + lane] += (double)c01[v];
                sums[(8 + v) * 64 + lane] += (double)c11[v];
            }
        }
        c00 = c01 = c11 = f4v{0, 0, 0, 0};
    }
    // G row-major 32 x 32 (both triangles).
    __device__ __forceinline__ void store(double* __restrict__ G, int lane) const {
        const int col = lane & 15;
        const int i0 = 4 * (lane >> 4);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + v;
            const double g01 = sums[(4 + v) * 64 + lane];
            G[i * 32 + col] = sums[v * 64 + lane];
            G[i * 32 + 16 + col] = g01;
            G[(16 + col) * 32 + i] = g01;
            G[(16 + i) * 32 + 16 + col] = sums[(8 + v) * 64 + lane];
        }
    }
};

// Eight consecutive rows [r0, r0+8) of column c (zero outside rows < rows and
// columns < ncols) as FP32.
template <bool VEC>
__device__ __forceinline__ void load8(float (&x)[8], const float* __restrict__ base, int64_t ld, int r0, int c,
                                      int rows, int ncols) {
    const Raw4<float> a = load_raw4<float, VEC>(base, ld, r0, c, rows, ncols);
    const Raw4<float> b = load_raw4<float, VEC>(base, ld, r0 + 4, c, rows, ncols);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        x[t] = a.v[t];
        x[4 + t] = b.v[t];
    }
}

// Pass 1a, one wave: the BF16x3 Gram matrix of row block `blk` of S of an
// m-row panel, left in the wave's FP64 slab `sums`.  The row blocks
// partition the panel's 64-row chunks by stride: block j takes chunks j,
// j + S, j + 2S, ..., so the S waves of a panel sweep it together over a
// window of S chunks (the Grams are only ever summed).  Lane l loads rows
// 8(l/16)..+7 of column l%16 (and 16 + l%16) of each 32-row step.  Full
// chunks go through two buffers in turn (ping-pong): unconditional loads
// issued before the previous chunk's products (a branch around a load made
// the compiler wait for every load in flight), padding columns (n < 32)
// loaded as column 0 and zeroed where they are used; the short last chunk
// is loaded with bounds checks.  Used by the batch (one row block per wave
// of an item) and by tsqr_mixed_stream (the waves of a persistent grid).
template <int NT, bool VEC>
__device__ __forceinline__ void mixed_gram_body(int m, int n, const float* __restrict__ A, int lda, double* sums,
                                                int blk, int S) {
    const int lane = threadIdx.x & 63;
    const int col = lane & 15;
    const int rq = 8 * (lane >> 4);
    const int chunks = (m + 63) / 64;
    constexpr int D = 2;  // 32-row steps per chunk
    GramAcc32 g;
    g.sums = sums;
    g.zero(lane);
    const bool in0 = col < n, in1 = col + 16 < n;
    const int c0col = in0 ? col : 0, c1col = in1 ? col + 16 : 0;
    const int nfull = m / 64;
    auto load_chunk = [&](int ch, float(&x0)[D][8], float(&x1)[D][8]) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const Raw4<float> a = load_col4<float, VEC>(A, lda, 64 * ch + 32 * q + rq + 4 * h, c0col);
#pragma unroll
                for (int t = 0; t < 4; ++t) x0[q][4 * h + t] = a.v[t];
                if constexpr (NT == 2) {
                    const Raw4<float> b = load_col4<float, VEC>(A, lda, 64 * ch + 32 * q + rq + 4 * h, c1col);
#pragma unroll
                    for (int t = 0; t < 4; ++t) x1[q][4 * h + t] = b.v[t];
                }
            }
        }
    };
    auto use_chunk = [&](float(&x0)[D][8], float(&x1)[D][8]) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                x0[q][k] = in0 ? x0[q][k] : 0.f;
                if constexpr (NT == 2) x1[q][k] = in1 ? x1[q][k] : 0.f;
            }
            g.add<NT>(x0[q], x1[q]);
        }
        g.flush<NT>(lane);
    };
    if (blk < nfull) {
        const int nch = (nfull - blk + S - 1) / S;
        float a0[D][8], a1[D][8], b0[D][8], b1[D][8];
        load_chunk(blk, a0, a1);
#pragma unroll 1
        for (int k = 0; k + 1 < nch; k += 2) {
            load_chunk(blk + (k + 1) * S, b0, b1);
            __builtin_amdgcn_sched_barrier(0);
            use_chunk(a0, a1);
            __builtin_amdgcn_sched_barrier(0);
            const int ka = k + 2 < nch ? k + 2 : nch - 1;  // the last pair reloads a chunk it has
            load_chunk(blk + ka * S, a0, a1);
            __builtin_amdgcn_sched_barrier(0);
            use_chunk(b0, b1);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (nch & 1) use_chunk(a0, a1);
    }
    if (chunks > nfull && blk == nfull % S) {  // the short last chunk, bounds-checked loads
        float c0[D][8], c1[D][8];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            load8<VEC>(c0[q], A, lda, 64 * nfull + 32 * q + rq, col, m, n);
            if constexpr (NT == 2) load8<VEC>(c1[q], A, lda, 64 * nfull + 32 * q + rq, col + 16, m, n);
            g.add<NT>(c0[q], c1[q]);
        }
        g.flush<NT>(lane);
    }
}

// G (FP64 row-major 32 x 32, both triangles) = the sum of a workgroup's
// four GramAcc32 slabs sums[0..3] in wave order, by its 256 threads (after
// a barrier): thread t takes register t/64 of lane t%64 in every block.
__device__ __forceinline__ void gram32_combine(const double (*sums)[12 * 64], double* __restrict__ G) {
    const int t = threadIdx.x;
    const int lane = t & 63, v = t >> 6;
    const int col = lane & 15, i = 4 * (lane >> 4) + v;
    double g00 = sums[0][v * 64 + lane], g01 = sums[0][(4 + v) * 64 + lane], g11 = sums[0][(8 + v) * 64 + lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
        g00 += sums[w][v * 64 + lane];
        g01 += sums[w][(4 + v) * 64 + lane];
        g11 += sums[w][(8 + v) * 64 + lane];
    }
    G[i * 32 + col] = g00;
    G[i * 32 + 16 + col] = g01;
    G[(16 + col) * 32 + i] = g01;
    G[(16 + i) * 32 + 16 + col] = g11;
}

// End of a streaming launch that sums a Gram matrix with GramAcc32 (256
// threads, P = gridDim.x workgroups): the four waves' FP64 slabs sums[wave]
// are added in wave order into L.G (accumulator layout: lane l, register v
// holds row 4(l/16) + v, column l%16 of a block), then stream_finish.
template <int NT>
__device__ __forceinline__ void gram32_finish(GramLds& L, const double (*sums)[12 * 64], double* __restrict__ partial,
                                              int* __restrict__ tk, int* bcast) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int col = lane & 15;
    const int i0 = 4 * (lane >> 4);
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        if (wv == s) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + v;
                const double g00 = sums[s][v * 64 + lane];
                const double g01 = NT == 2 ? sums[s][(4 + v) * 64 + lane] : 0.0;
                const double g11 = NT == 2 ? sums[s][(8 + v) * 64 + lane] : 0.0;
                if (s == 0) {
                    L.G[i][col] = g00;
                    L.G[i][16 + col] = g01;
                    L.G[16 + col][i] = g01;
                    L.G[16 + i][16 + col] = g11;
                } else {
                    L.G[i][col] += g00;
                    L.G[i][16 + col] += g01;
                    L.G[16 + col][i] += g01;
                    L.G[16 + i][16 + col] += g11;
                }
            }
        }
        __syncthreads();
    }
    stream_finish(L, partial, tk, bcast);
}

// Pass 1a of a single matrix, streamed: G = A^T A with the BF16x3 products
// of mixed_gram_body over the whole matrix by a persistent grid, as
// tsqr_gram_stream does with FP64 products.  Workgroup b (4 waves) of P:
// wave w = 4b + 0..3 takes the 64-row chunks w, w + 4P, ... (the last one
// bounds-checked), each summed in FP32 on the matrix cores and added to the
// wave's FP64 sums in LDS; the workgroup adds its waves in wave order into
// partial[b] and the last of each group (stream_group_size) sums the group
// (stream_finish).  Every sum has a fixed order.
template <int NT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void tsqr_mixed_stream(
    const TreeDesc d, const float* __restrict__ A, int lda, double* __restrict__ partial, int* __restrict__ tk) {
    __shared__ double sums[4][12 * 64];
    __shared__ GramLds L;
    __shared__ int bcast;
    const int wv = threadIdx.x >> 6;
    mixed_gram_body<NT, VEC>(d.m, d.n, A, lda, sums[wv], 4 * (int)blockIdx.x + wv, 4 * (int)gridDim.x);
    gram32_finish<NT>(L, sums, partial, tk, &bcast);
}

// ---------------------------------------------------------------------------
// Q passes: Q = A X in place (X upper triangular, FP64 row-major 32 x 32),
// and in the same pass the Gram matrix of the stored Q (BF16x3 for float,
// FP64 MFMA for double).  Pass 1 runs when the flat Gram path was taken
// (ctl[0]); pass 2 only if the check asked for it (ctl[1]).
//
// Rows go in chunks of 16 RPL (RPL = q_rpl<T>(): 4 rows of FP32, 2 of
// FP64, i.e. 16 bytes per lane and column).  A chunk is loaded as 16-byte
// column runs: lane l holds rows RPL(l%16) .. +RPL-1 of column 4p + l/16 in
// buf[p], p = 0..7 (zero past the rows and columns).  The products take one
// of two forms:
//
// FP64 (every double pass; float where X is not known to be well
// conditioned): 16x16x4 FP64 MFMA, element t of buf[kg] is the A operand
// with k = l/16 <-> column 4kg + l/16 and i = l%16 <-> row RPL(l%16) + t;
// products and sums in FP64, one rounding.  Output register v of (t, jt)
// is Q[r0 + RPL(4v + l/16) + t][16 jt + l%16].
//
// BF16x3 (float, flag ctl[kCtlBf16Q1] / ctl[kCtlBf16Q2]): BF16x3 products
// (bf16x3.hpp) summed in FP32 on v_mfma_f32_16x16x32_bf16 with k = the 32
// columns of A.  Those sums carry an error of about u32 |A| |X| per entry,
// i.e. u32 kappa(X) relative to Q, which is why the flag is set only for a
// well-conditioned X (flat_factor, mixed_check_body).  The A operand needs
// eight columns of one row per lane, so the chunk goes through the wave's
// LDS tile `tw` (column c, row r at c * kTrPitch + r) and comes back as row
// tile rt (16 rows): lane l takes A[r0 + 16 rt + l%16][8(l/16) + s], s =
// 0..7.  X is the B operand, X[8(l/16) + s][16 jt + l%16] (zero outside
// n x n).  Output register v of (rt, jt) is Q[r0 + 16 rt + 4(l/16) + v][16 jt
// + l%16].
//
// For float, either form leaves lane l with q[jt][u][t] = Q[r0 + 16u + 4(l/16)
// + t][16 jt + l%16]; the chunk goes back through the LDS tile so that it is
// stored in the load layout (16-byte runs, 256 contiguous bytes per column
// and instruction), and q feeds the Gram: the 32 rows r0 + 32h + [0, 32) are
// BF16 k-slot s = 4(u - 2h) + t of lane l, a bijection onto the 32 rows,
// identical for both operands.  For double, register (v, t) of lane l is
// the 16x16x4 FP64 Gram operand with k = l/16 <-> row RPL(4v + l/16) + t.
// Zero rows and columns of the operands make the padding of Q exactly zero,
// and padded columns are not stored.
// ---------------------------------------------------------------------------
constexpr int kTrPitch = 66;  // floats per LDS column (8-byte aligned, reads conflict-free per half-wave)

template <typename T>
constexpr int q_rpl() {
    return sizeof(T) == 4 ? 4 : 2;
}

// x[t] = rows r0 + t (t < RPL) of column c, zero outside rows and columns.
template <typename T, int RPL, bool VEC>
__device__ __forceinline__ void load_run(T (&x)[RPL], const T* __restrict__ base, int64_t ld, int r0, int c,
                                         int rows, int ncols) {
    using C = Chunk16<T, RPL>;
    static_assert(C::kChunks == 1, "one 16-byte run per lane");
    if (c < ncols) {
        const T* p = base + (int64_t)c * ld + r0;
        if (VEC && r0 + RPL <= rows) {
            const typename C::vec v = *reinterpret_cast<const typename C::vec*>(p);
#pragma unroll
            for (int t = 0; t < RPL; ++t) x[t] = v[t];
        } else {
#pragma unroll
            for (int t = 0; t < RPL; ++t) x[t] = (r0 + t < rows) ? p[t] : T(0);
        }
    } else {
#pragma unroll
        for (int t = 0; t < RPL; ++t) x[t] = T(0);
    }
}

template <typename T, bool VEC>
__device__ __forceinline__ void q_chunk_load(T (&buf)[8][q_rpl<T>()], const T* __restrict__ base, int lda, int r0,
                                             int rows, int n, int lane) {
    constexpr int RPL = q_rpl<T>();
#pragma unroll
    for (int p = 0; p < 8; ++p)
        load_run<T, RPL, VEC>(buf[p], base, lda, r0 + RPL * (lane & 15), 4 * p + (lane >> 4), rows, n);
}

// B operands of the FP64 form: mb[kg][jt] = X[4kg + l/16][16jt + l%16].
__device__ __forceinline__ void q_x_f64(double (&mb)[8][2], const double* __restrict__ X, int lane) {
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) mb[kg][jt] = X[(4 * kg + lq) * 32 + 16 * jt + li];
}

// B operands of the BF16 form, split: X[8(l/16) + s][16jt + l%16].
template <int NT>
__device__ __forceinline__ void q_x_bf16(Split3 (&xb)[NT], const double* __restrict__ X, int n, int lane) {
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) {
        double xs[8];
        const int j = 16 * jt + li;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 8 * lq + s;
            xs[s] = (k < n && j < n) ? X[k * 32 + j] : 0.0;
        }
        xb[jt] = split3(xs);
    }
}

// Gram accumulators of a Q pass: GramAcc32 for float, FP64 MFMA blocks
// (0,0), (0,1), (1,1) for double.
struct QGram {
    GramAcc32 g;
    v4d d00, d01, d11;
};

// The chunk in buf into the wave's LDS tile (column-major, pitch kTrPitch).
__device__ __forceinline__ void q_tile_put(float* tw, const float (&buf)[8][4], int lane) {
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int p = 0; p < 8; ++p)
#pragma unroll
        for (int t = 0; t < 4; ++t) tw[(4 * p + lq) * kTrPitch + 4 * li + t] = buf[p][t];
}

// FP32 Q chunk (q[jt][u][t], layout above) to A through the LDS tile, then
// its BF16x3 Gram into g.
template <int NT, bool VEC>
__device__ __forceinline__ void q_chunk_finish32(float* tw, const float (&q)[2][4][4], float* __restrict__ base,
                                                 int lda, int rows, int n, int r0, GramAcc32& g, int lane) {
    const int li = lane & 15, lq = lane >> 4;
    wave_sync();  // every lane's reads of the tile are done
#pragma unroll
    for (int jt = 0; jt < NT; ++jt)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < 4; ++t) tw[(16 * jt + li) * kTrPitch + 16 * u + 4 * lq + t] = q[jt][u][t];
    wave_sync();
#pragma unroll
    for (int p = 0; p < 4 * NT; ++p) {
        const int c = 4 * p + lq;
        if (c < n) {
            float v[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = tw[c * kTrPitch + 4 * li + t];
            store_column<float, 4, VEC>(v, base + (int64_t)c * lda + r0 + 4 * li, rows, r0 + 4 * li);
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float x0[8], x1[8];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            x0[t] = q[0][2 * h][t];
            x0[4 + t] = q[0][2 * h + 1][t];
            x1[t] = q[NT - 1][2 * h][t];
            x1[4 + t] = q[NT - 1][2 * h + 1][t];
        }
        g.add<NT>(x0, x1);
    }
    g.flush<NT>(lane);
}

// FP64 form of one chunk: a[kg][t] = buf[kg][t] converted (the caller may
// reload buf meanwhile); Q to base (rows rows), Gram into acc.  tw: the
// wave's LDS tile (float only).
template <typename T, int NT, bool VEC>
__device__ __forceinline__ void q_chunk_f64(const double (&a)[8][q_rpl<T>()], const double (&mb)[8][2], int nk,
                                            T* __restrict__ base, int lda, int rows, int n, int r0, QGram& acc,
                                            float* tw, int lane) {
    constexpr int RPL = q_rpl<T>();
    const int li = lane & 15, lq = lane >> 4;
    v4d c[RPL][2];
#pragma unroll
    for (int t = 0; t < RPL; ++t) c[t][0] = c[t][1] = v4d{0, 0, 0, 0};
    // X is upper triangular: rows 16..31 are zero in columns 0..15
#pragma unroll
    for (int kg = 0; kg < 8; ++kg) {
        if (kg < nk) {
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                if (kg < 4) c[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][0], c[t][0], 0, 0, 0);
                if constexpr (NT == 2)
                    c[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][1], c[t][1], 0, 0, 0);
            }
        }
    }
    if constexpr (sizeof(T) == 4) {
        float q[2][4][4];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const bool cin = 16 * jt + li < n;
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int t = 0; t < 4; ++t) q[jt][v][t] = cin ? (float)c[t][jt][v] : 0.f;
        }
        q_chunk_finish32<NT, VEC>(tw, q, base, lda, rows, n, r0, acc.g, lane);
    } else {
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) {
            const int cc = 16 * jt + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                T q[RPL];
#pragma unroll
                for (int t = 0; t < RPL; ++t) q[t] = cc < n ? (T)c[t][jt][v] : T(0);
                const int rr = r0 + RPL * (4 * v + lq);
                if (cc < n) store_column<T, RPL, VEC>(q, base + (int64_t)cc * lda + rr, rows, rr);
#pragma unroll
                for (int t = 0; t < RPL; ++t) c[t][jt][v] = q[t];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const double q0 = c[t][0][v];
                acc.d00 = __builtin_amdgcn_mfma_f64_16x16x4f64(q0, q0, acc.d00, 0, 0, 0);
                if constexpr (NT == 2) {
                    const double q1 = c[t][1][v];
                    acc.d01 = __builtin_amdgcn_mfma_f64_16x16x4f64(q0, q1, acc.d01, 0, 0, 0);
                    acc.d11 = __builtin_amdgcn_mfma_f64_16x16x4f64(q1, q1, acc.d11, 0, 0, 0);
                }
            }
    }
}

// BF16x3 form of one FP32 chunk, from the LDS tile: Q to base, Gram into g.
template <int NT, bool VEC>
__device__ __forceinline__ void q_chunk_bf16(float* tw, const Split3 (&xb)[NT], float* __restrict__ base, int lda,
                                             int rows, int n, int r0, GramAcc32& g, int lane) {
    const int li = lane & 15, lq = lane >> 4;
    f4v c[4][NT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        float a[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) a[s] = tw[(8 * lq + s) * kTrPitch + 16 * rt + li];
        const Split3 sa = split3(a);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) c[rt][jt] = mma_bf16x3(sa, xb[jt], f4v{0, 0, 0, 0});
    }
    float q[2][4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            q[0][rt][v] = c[rt][0][v];
            q[1][rt][v] = c[rt][NT - 1][v];  // unused when NT == 1
        }
    q_chunk_finish32<NT, VEC>(tw, q, base, lda, rows, n, r0, g, lane);
}

// Q pass over row block `blk` of S of an m-row FP32 panel by one wave (the
// batched path; the chunks j, j + S, ... as in mixed_gram_body), the next
// chunk's loads in flight while the current one is multiplied; the Gram
// matrix of the stored rows is left in the wave's FP64 slab `sums`.  `bf16`
// picks the form; tw is the wave's LDS tile.
template <int NT, bool VEC>
__device__ __forceinline__ void mixed_q_leaf(int m, int n, float* __restrict__ A, int lda,
                                             const double* __restrict__ X, double* sums, float* tw, int blk, int S,
                                             bool bf16) {
    const int lane = threadIdx.x & 63;
    const int chunks = (m + 63) / 64;
    QGram acc;
    acc.g.sums = sums;
    acc.g.zero(lane);
    float buf[8][4];
    if (blk < chunks) q_chunk_load<float, VEC>(buf, A + (int64_t)blk * 64, lda, 0, m - 64 * blk, n, lane);
    if (bf16) {
        Split3 xb[NT];
        q_x_bf16<NT>(xb, X, n, lane);
#pragma unroll 1
        for (int ch = blk; ch < chunks; ch += S) {
            wave_sync();  // the previous chunk's LDS reads are done
            q_tile_put(tw, buf, lane);
            wave_sync();
            const int nx = ch + S;
            if (nx < chunks) q_chunk_load<float, VEC>(buf, A + (int64_t)nx * 64, lda, 0, m - 64 * nx, n, lane);
            q_chunk_bf16<NT, VEC>(tw, xb, A + (int64_t)ch * 64, lda, m - 64 * ch, n, 0, acc.g, lane);
        }
    } else {
        double mb[8][2];
        q_x_f64(mb, X, lane);
        const int nk = (n + 3) / 4;
#pragma unroll 1
        for (int ch = blk; ch < chunks; ch += S) {
            double a[8][4];
#pragma unroll
            for (int kg = 0; kg < 8; ++kg)
#pragma unroll
                for (int t = 0; t < 4; ++t) a[kg][t] = (double)buf[kg][t];
            const int nx = ch + S;
            if (nx < chunks) q_chunk_load<float, VEC>(buf, A + (int64_t)nx * 64, lda, 0, m - 64 * nx, n, lane);
            q_chunk_f64<float, NT, VEC>(a, mb, nk, A + (int64_t)ch * 64, lda, m - 64 * ch, n, 0, acc, tw, lane);
        }
    }
}

// Q pass of a single matrix, streamed by a persistent grid (as
// tsqr_mixed_stream): workgroup b (4 waves) of P, wave w = 4b + 0..3 takes
// the chunks w, w + 4P, ...  Each wave consumes its chunk's loads (into its
// LDS tile or FP64 registers) before issuing the next chunk's.  The Gram
// matrix of the stored Q is summed per wave (GramAcc32 slabs in LDS for
// float, FP64 MFMA registers for double), then per workgroup in wave order
// into partial[b], then per group (stream_finish); every sum has a fixed
// order.  X: R^{-1} of pass 1, R2^{-1} of pass 2 (side block T0).
template <typename T, int NT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void tsqr_mixed_qs(
    int m, int n, T* __restrict__ A, int lda, const double* __restrict__ X, double* __restrict__ partial,
    int* __restrict__ tk, const int* __restrict__ ctl, int pass) {
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int RPL = q_rpl<T>();
    constexpr int CR = 16 * RPL;  // rows per chunk
    __shared__ double sums[F32 ? 4 : 1][12 * 64];
    __shared__ float tr[F32 ? 4 : 1][32 * kTrPitch];
    __shared__ GramLds L;
    __shared__ int bcast;
    if (!ctl[0] || (pass == 2 && !ctl[1])) return;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int W = 4 * (int)gridDim.x;
    const int w = 4 * blockIdx.x + wv;
    const int chunks = (m + CR - 1) / CR;
    float* tw = tr[F32 ? wv : 0];
    QGram acc;
    acc.d00 = acc.d01 = acc.d11 = v4d{0, 0, 0, 0};
    if constexpr (F32) {
        acc.g.sums = sums[wv];
        acc.g.zero(lane);
    }
    T buf[8][RPL];
    if (w < chunks) q_chunk_load<T, VEC>(buf, A, lda, CR * w, m, n, lane);
    bool bf16 = false;
    if constexpr (F32) bf16 = ctl[pass == 1 ? kCtlBf16Q1 : kCtlBf16Q2] != 0;
    if (bf16) {
        if constexpr (F32) {
            Split3 xb[NT];
            q_x_bf16<NT>(xb, X, n, lane);
#pragma unroll 1
            for (int ch = w; ch < chunks; ch += W) {
                wave_sync();  // the previous chunk's LDS reads are done
                q_tile_put(tw, buf, lane);
                wave_sync();
                if (ch + W < chunks) q_chunk_load<T, VEC>(buf, A, lda, CR * (ch + W), m, n, lane);
                q_chunk_bf16<NT, VEC>(tw, xb, A, lda, m, n, CR * ch, acc.g, lane);
            }
        }
    } else {
        double mb[8][2];
        q_x_f64(mb, X, lane);
        const int nk = (n + 3) / 4;
#pragma unroll 1
        for (int ch = w; ch < chunks; ch += W) {
            double a[8][RPL];
#pragma unroll
            for (int kg = 0; kg < 8; ++kg)
#pragma unroll
                for (int t4 = 0; t4 < RPL; ++t4) a[kg][t4] = (double)buf[kg][t4];
            if (ch + W < chunks) q_chunk_load<T, VEC>(buf, A, lda, CR * (ch + W), m, n, lane);
            q_chunk_f64<T, NT, VEC>(a, mb, nk, A, lda, m, n, CR * ch, acc, tw, lane);
        }
    }
    if constexpr (F32)
        gram32_finish<NT>(L, sums, partial, tk, &bcast);
    else
        stream_reduce<NT>(L, acc.d00, acc.d01, acc.d11, partial, tk, &bcast);
}

// Residual check after pass `pass`, by one wave (lanes 0..63), on the Gram
// matrix of the pass's output in L.G: e = ||G - I||_F over the leading
// n x n block goes to misc[pass - 1].  Pass 1: e <= tol finishes;
// otherwise R2 = chol(G), R = R2 R1 (R1 in FP64, row-major) to the caller,
// X = R2^{-1} (FP64, row-major) and ctl[1] = 1.  Pass 2: ctl[2] = (e > tol).
// A second Cholesky that breaks down sets ctl[2] = 2 and keeps Q1, R1.
template <typename T>
__device__ __forceinline__ void mixed_check_body(GramLds& L, int n, T* __restrict__ R, int ldr,
                                                 double* __restrict__ X, const double* __restrict__ R1,
                                                 double* __restrict__ misc, int* __restrict__ ctl, int pass,
                                                 double tol, int lane) {
    const int c = lane & 31;
    double e = 0.0;
    if (c < n) {
#pragma unroll 4
        for (int r = 0; r < n; ++r) {
            const double v = L.G[r][c] - (r == c ? 1.0 : 0.0);
            e = fmat(v, v, e);
        }
    }
    e = __builtin_sqrt(sum_half(e));
    if (pass == 2) {
        if (lane == 0) {
            misc[1] = e;
            ctl[2] = (e <= tol) ? 0 : 1;
        }
        return;
    }
    if (lane == 0) misc[0] = e;
    if (e <= tol) return;
    bool ok;
    {
        double gcol[32];
        ok = chol_wave(L, n, lane, gcol);  // both half-waves factor the same G
        r_to_lds(gcol, L, lane);           // L.G = R2 (row-major, upper)
    }
    if (!ok) {
        if (lane == 0) ctl[2] = 2;
        return;
    }
    // column c of R = R2 R1: sum over r <= k <= c of R2[r][k] R1[k][c]
    // (R1 read from memory in the loop: this runs once per matrix and the
    // kernels it is fused into keep their registers for the streaming part)
    if (lane < 32 && c < n) {
#pragma unroll 1
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
#pragma unroll 1
            for (int k = r; k <= c; ++k) s = fmat(L.G[r][k], R1[k * 32 + c], s);
            R[r + (int64_t)c * ldr] = (r <= c) ? (T)s : T(0);
        }
    }
    double x[32];
    tri_inv_wave(L, lane, x);
    if (lane < 32) {
#pragma unroll
        for (int r = 0; r < 32; ++r) X[r * 32 + c] = x[r];
        if (lane == 0) {
            ctl[1] = 1;
            ctl[kCtlBf16Q2] = (e <= kBf16QMaxResidual) ? 1 : 0;
        }
    }
}

// Check kernel (one workgroup of kSumThreads threads): G = sum of the nparts
// partials (in order), then mixed_check_body with X = side block T0, R1 in
// side block T0+1 and the residuals in side block T0+2.
template <typename T>
__global__ __launch_bounds__(kSumThreads) void tsqr_mixed_check(const TreeDesc d, int nparts,
                                                        const double* __restrict__ partial, T* __restrict__ R,
                                                        int ldr, double* __restrict__ side, int* __restrict__ ctl,
                                                        int pass, double tol) {
    __shared__ GramLds L;
    if (!ctl[0] || (pass == 2 && !ctl[1])) return;
    const int t = threadIdx.x;
    sum_partials(L, nparts, partial, t);
    __syncthreads();
    if (t >= 64) return;
    const int64_t T0 = d.side_leaves;
    mixed_check_body<T>(L, d.n, R, ldr, side + T0 * 1024, side + (T0 + 1) * 1024, side + (T0 + 2) * 1024, ctl,
                        pass, tol, t);
}

}  // namespace tsqr
