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
//           FP64 (tsqr_gram_flat); Q1 = A R1^{-1} in FP64 MFMA, rounded once
//           (tsqr_mixed_q), and in the same pass G2 = Q1^T Q1 of the rounded
//           Q1 (BF16x3 again).
//   check   e1 = ||G2 - I||_F (the orthogonality of the stored Q1).  If
//           e1 <= tol the result is Q1, R1 (tsqr_mixed_check).
//   pass 2  otherwise R2 = chol(G2), Q = Q1 R2^{-1} (the same kernel, with
//           G3 = Q^T Q fused again), R = R2 R1 in FP64; e2 = ||G3 - I||_F is
//           recorded and status 1 is set if e2 > tol.
// If the first Cholesky breaks down or kappa_est exceeds the gate, the whole
// matrix takes the Householder tree instead (the superleaf and tree kernels,
// driven by the same ctl word), so no input is rejected.
//
// ctl words (int): ctl[0] = 1 flat Gram taken, ctl[1] = 1 second pass
// requested, ctl[2] = status (0 ok, 1 residual above tol after the second
// pass, 2 second Cholesky broke down).  The FP64 side block T0+2 holds
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

// Pass 1a of the batch: G_b = A_b^T A_b of one superleaf (one wave) with
// BF16x3 products, written to G (FP64, row-major 32 x 32).  Lane l loads
// rows 8(l/16)..+7 of column l%16 (and 16 + l%16) of each 32-row step; the
// two steps of a 64-row batch are loaded together and the next batch is in
// flight while the current one is multiplied.
template <int NT, bool VEC>
__device__ __forceinline__ void mixed_gram_body(const LevelDesc& L0, int n, const float* __restrict__ A, int lda,
                                                double* __restrict__ G, double* sums, int leaf) {
    const int lane = threadIdx.x & 63;
    const int rows = leaf_rows_of(L0, leaf);
    const float* base = A + (int64_t)leaf * L0.tile_rows;
    const int col = lane & 15;
    const int rq = 8 * (lane >> 4);
    constexpr int D = 2;  // 32-row steps per batch (a batch and its successor in flight: 64 VGPRs per column half)
    GramAcc32 g;
    g.sums = sums;
    g.zero(lane);
    float c0[D][8], c1[D][8];
#pragma unroll
    for (int q = 0; q < D; ++q) {
        load8<VEC>(c0[q], base, lda, 32 * q + rq, col, rows, n);
        if constexpr (NT == 2) load8<VEC>(c1[q], base, lda, 32 * q + rq, col + 16, rows, n);
    }
#pragma unroll 1
    for (int s0 = 0; s0 < rows; s0 += 32 * D) {
        const int next = s0 + 32 * D;
        float p0[D][8], p1[D][8];
        if (next < rows) {
#pragma unroll
            for (int q = 0; q < D; ++q) {
                load8<VEC>(p0[q], base, lda, next + 32 * q + rq, col, rows, n);
                if constexpr (NT == 2) load8<VEC>(p1[q], base, lda, next + 32 * q + rq, col + 16, rows, n);
            }
        }
#pragma unroll
        for (int q = 0; q < D; ++q) {
            g.add<NT>(c0[q], c1[q]);
            if (q & 1) g.flush<NT>(lane);
        }
        if (next >= rows) break;
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                c0[q][t] = p0[q][t];
                if constexpr (NT == 2) c1[q][t] = p1[q][t];
            }
    }
    g.store(G, lane);
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
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int m = d.m, n = d.n;
    const int W = 4 * (int)gridDim.x;
    const int w = 4 * blockIdx.x + wv;
    const int chunks = (m + 63) / 64;
    const int col = lane & 15;
    const int rq = 8 * (lane >> 4);
    constexpr int D = 2;  // 32-row steps per chunk
    GramAcc32 g;
    g.sums = sums[wv];
    g.zero(lane);
    // Full 64-row chunks in two buffers in turn (ping-pong, as
    // tsqr_gram_stream): unconditional loads issued before the previous
    // chunk's products, padding columns (n < 32) loaded as column 0 and
    // zeroed only where they are used.
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
    if (w < nfull) {
        const int nch = (nfull - w + W - 1) / W;
        float a0[D][8], a1[D][8], b0[D][8], b1[D][8];
        load_chunk(w, a0, a1);
#pragma unroll 1
        for (int k = 0; k + 1 < nch; k += 2) {
            load_chunk(w + (k + 1) * W, b0, b1);
            __builtin_amdgcn_sched_barrier(0);
            use_chunk(a0, a1);
            __builtin_amdgcn_sched_barrier(0);
            const int ka = k + 2 < nch ? k + 2 : nch - 1;  // the last pair reloads a chunk it has
            load_chunk(w + ka * W, a0, a1);
            __builtin_amdgcn_sched_barrier(0);
            use_chunk(b0, b1);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (nch & 1) use_chunk(a0, a1);
    }
    if (chunks > nfull && w == nfull % W) {  // the short last chunk, bounds-checked loads
        float c0[D][8], c1[D][8];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            load8<VEC>(c0[q], A, lda, 64 * nfull + 32 * q + rq, col, m, n);
            if constexpr (NT == 2) load8<VEC>(c1[q], A, lda, 64 * nfull + 32 * q + rq, col + 16, m, n);
            g.add<NT>(c0[q], c1[q]);
        }
        g.flush<NT>(lane);
    }
    // the workgroup's sum in L.G, waves added in order (accumulator layout:
    // lane l, register v holds row 4(l/16) + v, column l%16 of a block)
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
    stream_finish(L, partial, tk, &bcast);
}

// Pass 1b / 2: Q = A X in place for every superleaf (X upper triangular,
// FP64 row-major 32 x 32 in side block T0), products and sums in FP64 on the
// matrix cores and one rounding, as in tsqr_gram_q's flat case; then the
// Gram matrix of the rounded rows goes to side[leaf]: BF16x3 for float, FP64
// MFMA for double.  Pass 1 runs when the flat Gram path was taken (ctl[0]);
// pass 2 only if the check asked for it (ctl[1]).
//
// Layout: rows in blocks of 64; the FP64 MFMA output register v of (t, jt) is
// row r0 + 16v + 4(l/16) + t, column 16 jt + l%16.
//   float:  the 32 rows r0 + 32h + [0, 32) feed BF16 k-slots
//           s = 4(v - 2h) + t of lane l, i.e. k = 8(l/16) + s <-> row
//           16(v - 2h) + 4(l/16) + t: a bijection onto the 32 rows,
//           identical for both operands;
//   double: register (v, t) of lane l is the 16x16x4 f64 operand with
//           k = l/16 <-> row 16v + 4(l/16) + t, again the same for both
//           operands.
template <typename T, int NT, bool VEC>
__device__ __forceinline__ void mixed_q_body(const LevelDesc& L0, int n, T* __restrict__ A, int lda,
                                             const double* __restrict__ X, double* __restrict__ Gout, double* sums,
                                             int leaf) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lq = lane >> 4;
    // B operand of k-group kg (k = 4 kg + l/16), column tile jt
    double mb[8][2];
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) mb[kg][jt] = X[(4 * kg + lq) * 32 + 16 * jt + li];
    const int rows = leaf_rows_of(L0, leaf);
    T* base = A + (int64_t)leaf * L0.tile_rows;
    const int nk = (n + 3) / 4;
    Raw4<T> raw[8];
#pragma unroll
    for (int kg = 0; kg < 8; ++kg) {
        const int k = 4 * kg + lq;
        if (kg < nk) raw[kg] = load_raw4<T, VEC>(base, lda, 4 * li, k < n ? k : n - 1, rows, n);
    }
    GramAcc32 g;  // float: BF16x3, FP64 sums in LDS
    v4d d00 = {0, 0, 0, 0}, d01 = {0, 0, 0, 0}, d11 = {0, 0, 0, 0};  // double: FP64 MFMA
    if constexpr (sizeof(T) == 4) {
        g.sums = sums;
        g.zero(lane);
    }
#pragma unroll 1
    for (int r0 = 0; r0 < rows; r0 += 64) {
        double a[8][4];
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t) a[kg][t] = (double)raw[kg].v[t];
        const int r1 = r0 + 64;
        if (r1 < rows) {
#pragma unroll
            for (int kg = 0; kg < 8; ++kg) {
                const int k = 4 * kg + lq;
                if (kg < nk) raw[kg] = load_raw4<T, VEC>(base, lda, r1 + 4 * li, k < n ? k : n - 1, rows, n);
            }
        }
        v4d acc[4][2];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = v4d{0, 0, 0, 0};
        // X is upper triangular: rows 16..31 are zero in columns 0..15
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) {
            if (kg < nk) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (kg < 4)
                        acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][0], acc[t][0], 0, 0, 0);
                    if constexpr (NT == 2)
                        acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][1], acc[t][1], 0, 0, 0);
                }
            }
        }
        // q[jt][v][t] = Q[r0 + 16v + 4 lq + t][16 jt + li], rounded; zero in
        // padded columns (rows past the end are zero already)
        T q[2][4][4];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int cc = 16 * jt + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
#pragma unroll
                for (int t = 0; t < 4; ++t) q[jt][v][t] = cc < n ? (T)acc[t][jt][v] : T(0);
                if (cc < n)
                    store_column<T, 4, VEC>(q[jt][v], base + (int64_t)cc * lda + r0 + 16 * v + 4 * lq, rows,
                                            r0 + 16 * v + 4 * lq);
            }
        }
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float x0[8], x1[8];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    x0[t] = q[0][2 * h][t];
                    x0[4 + t] = q[0][2 * h + 1][t];
                    x1[t] = q[1][2 * h][t];
                    x1[4 + t] = q[1][2 * h + 1][t];
                }
                g.add<NT>(x0, x1);
            }
            g.flush<NT>(lane);
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    d00 = __builtin_amdgcn_mfma_f64_16x16x4f64(q[0][v][t], q[0][v][t], d00, 0, 0, 0);
                    if constexpr (NT == 2) {
                        d01 = __builtin_amdgcn_mfma_f64_16x16x4f64(q[0][v][t], q[1][v][t], d01, 0, 0, 0);
                        d11 = __builtin_amdgcn_mfma_f64_16x16x4f64(q[1][v][t], q[1][v][t], d11, 0, 0, 0);
                    }
                }
        }
    }
    if constexpr (sizeof(T) == 4) {
        g.store(Gout, lane);
    } else {
        // f64 accumulator: lane l, register v holds C[4v + l/16][l%16]
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * v + lq;
            Gout[i * 32 + li] = d00[v];
            Gout[i * 32 + 16 + li] = d01[v];
            Gout[(16 + li) * 32 + i] = d01[v];
            Gout[(16 + i) * 32 + 16 + li] = d11[v];
        }
    }
}

template <typename T, int NT, bool VEC>
__global__ __launch_bounds__(64) void tsqr_mixed_q(const TreeDesc d, T* __restrict__ A, int lda,
                                                   double* __restrict__ side, const int* __restrict__ ctl, int pass) {
    __shared__ double sums[12 * 64];
    if (!ctl[0] || (pass == 2 && !ctl[1])) return;
    // one workgroup per side superleaf (side_rows rows)
    LevelDesc view = d.lv[0];
    view.tile_rows = d.side_rows;
    view.tiles = d.side_leaves;
    mixed_q_body<T, NT, VEC>(view, d.n, A, lda, side + (int64_t)d.side_leaves * 1024,
                             side + (int64_t)blockIdx.x * 1024, sums, blockIdx.x);
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
        if (lane == 0) ctl[1] = 1;
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
