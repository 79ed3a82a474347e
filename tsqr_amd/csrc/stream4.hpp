// stream4.hpp -- the two streaming passes of the flat Gram path for FP32
// matrices (R pass: G = A^T A; Q pass: Q = A R^{-1}), on the 4x4x4 FP64
// matrix-core instruction.
//
// v_mfma_f64_4x4x4_4b_f64 computes four independent 4x4 products with
// k = 4.  Lane l = 16k + 4b + z feeds A_b[z][k] and B_b[k][z], and the
// result D_b[i][j] is in lane 16i + 4b + j.  On MI355X it issues at the full
// FP64 matrix rate (73 TF/s measured with independent accumulators, against
// 46-62 TF/s for v_mfma_f64_16x16x4f64), and 4 x 4 output tiles cover the
// upper triangle of a symmetric 32 x 32 Gram matrix with 36 of 64 tiles,
// where 16 x 16 tiles need 3 of 4: a quarter fewer products.
//
// Products of FP32 values are exact in FP64 and every sum is FP64, as in the
// 16 x 16 kernels of gram.hpp; only the order of the FP64 sums differs.
//
// Chunks are 64 rows.  A lane loads 4 consecutive rows of a column with one
// 16-byte load, so every load instruction reads 256 contiguous bytes of
// each of four columns.
#pragma once
#include <type_traits>

#include "gram.hpp"

namespace tsqr {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double mfma4x4(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// Rows [r, r + 4) of column c, all inside the matrix (no checks); NTL: a
// nontemporal load (6.6 against 6.3 TB/s for a plain read stream on MI355X).
template <bool VEC, bool NTL>
__device__ __forceinline__ f4v load_rows4(const float* __restrict__ A, int64_t lda, int64_t r, int c) {
    const float* p = A + (int64_t)c * lda + r;
    if constexpr (VEC && NTL) {
        return __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    } else if constexpr (VEC) {
        return *reinterpret_cast<const f4v*>(p);
    } else {
        return f4v{p[0], p[1], p[2], p[3]};
    }
}

// FP64 data: rows [r, r + 2) of column c (one 16-byte load), as load_rows4.
typedef double d2v __attribute__((ext_vector_type(2)));
template <bool VEC, bool NTL>
__device__ __forceinline__ d2v load_rows2(const double* __restrict__ A, int64_t lda, int64_t r, int c) {
    const double* p = A + (int64_t)c * lda + r;
    if constexpr (VEC && NTL) {
        return __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    } else if constexpr (VEC) {
        return *reinterpret_cast<const d2v*>(p);
    } else {
        return d2v{p[0], p[1]};
    }
}

// The 16-byte run of rows [r, r + RPL) of column c for either type.
template <typename T>
struct Rows16;
template <>
struct Rows16<float> {
    static constexpr int RPL = 4;
    typedef f4v vec;
    template <bool VEC, bool NTL>
    static __device__ __forceinline__ vec load(const float* A, int64_t lda, int64_t r, int c) {
        return load_rows4<VEC, NTL>(A, lda, r, c);
    }
};
template <>
struct Rows16<double> {
    static constexpr int RPL = 2;
    typedef d2v vec;
    template <bool VEC, bool NTL>
    static __device__ __forceinline__ vec load(const double* A, int64_t lda, int64_t r, int c) {
        return load_rows2<VEC, NTL>(A, lda, r, c);
    }
};

// Rows [r, r + 4) of column c, zero past row m.
__device__ __forceinline__ f4v load_rows4_tail(const float* __restrict__ A, int64_t lda, int64_t r, int c, int m) {
    const float* p = A + (int64_t)c * lda + r;
    f4v o;
#pragma unroll
    for (int s = 0; s < 4; ++s) o[s] = (r + s < m) ? p[s] : 0.0f;
    return o;
}

// End of a launch that sums a Gram matrix on 4 x 4 tiles (256 threads; acc
// holds this lane's tile-pair sums, lane 16k + 4b + z for row k of block b):
// the four blocks of each tile are added (lanes l, l^4, l^8, l^12), the
// workgroup's waves in wave order in L.G, then stream_finish.
template <int NC>
__device__ __forceinline__ void gram4_finish(GramLds& L, double (&acc)[NC * (NC + 1) / 2],
                                             double* __restrict__ partial, int* __restrict__ tk, int* bcast) {
    constexpr int NP = NC * (NC + 1) / 2;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        acc[p] += __shfl_xor(acc[p], 4);
        acc[p] += __shfl_xor(acc[p], 8);
    }
    for (int e = t; e < 32 * 33; e += 256) (&L.G[0][0])[e] = 0.0;
    __syncthreads();
    const int i = lane >> 4, j = lane & 3;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        if (wv == s && (lane & 12) == 0) {
            int p = 0;
#pragma unroll
            for (int I = 0; I < NC; ++I)
#pragma unroll
                for (int J = I; J < NC; ++J, ++p) {
                    L.G[4 * I + i][4 * J + j] += acc[p];
                    if (I != J) L.G[4 * J + j][4 * I + i] += acc[p];
                }
        }
        __syncthreads();
    }
    stream_finish(L, partial, tk, bcast);
}

// ---------------------------------------------------------------------------
// R pass.  Lane l holds rows RPL(l/4) .. +RPL-1 of the columns 4c + l%4
// (c < NC, NC = padded width / 4) of a chunk of 16 RPL rows (RPL = 4 for
// FP32 data, 2 for FP64: one 16-byte load per column); step s (0..RPL-1)
// feeds row RPL(l/4) + s, so block b of an instruction sums the rows
// RPL(4k + b) + s, k = 0..3.  Every
// instruction computes the same tile pair (I, J), I <= J, in all four
// blocks, on different rows; the blocks are added once, at the end
// (lanes l, l^4, l^8, l^12).  Workgroup b of P (4 waves) takes the chunks
// w, w + 4P, ... (w = 4b + wave), as tsqr_gram_stream does, and the partial
// sums and group tickets are those of stream_finish.
//
// Loads: three chunk buffers in turn, so that each chunk's loads are issued
// two chunks ahead of its products (the scheduling barriers keep that
// order).  The chunks are read with nontemporal loads, except for the last
// `cached` chunks of the sweep (the bottom of A): plain loads leave those in
// the memory-side cache, where the Q pass, which starts at the bottom, finds
// them (measured on MI355X: the nontemporal R pass is 0.09 ms faster over
// 3.06 GB, the cached bottom saves the Q pass 0.05 ms per 256 MB).
// FULL: n is the padded width, so no column is padding.  FP64 data: the
// products are rounded FP64 products, as in tsqr_gram_stream<double>.
template <int NC, bool VEC, bool FULL, typename T = float>
__global__ __launch_bounds__(256) void tsqr_gram_stream4(int m, int n, const T* __restrict__ A, int lda,
                                                         double* __restrict__ partial, int* __restrict__ tk,
                                                         int cached, const int* __restrict__ skip = nullptr) {
    constexpr int NP = NC * (NC + 1) / 2;
    using RW = Rows16<T>;
    using V = typename RW::vec;
    constexpr int RPL = RW::RPL;
    constexpr int CR = 16 * RPL;  // rows per chunk
    __shared__ GramLds L;
    __shared__ int bcast;
    if (skip && skip[0]) return;  // an earlier pass already took the flat path
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int W = 4 * (int)gridDim.x;
    const int w = 4 * (int)blockIdx.x + wv;
    const int z = lane & 3;
    const int rl = RPL * (lane >> 2);
    const int nfull = m / CR;
    double acc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = 0.0;
    // padding columns (>= n) load column 0 and count as zero
    bool in[NC];
    int col[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        in[c] = FULL || 4 * c + z < n;
        col[c] = in[c] ? 4 * c + z : 0;
    }
    auto chunk = [&](const V(&x)[NC]) {
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            double d[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) d[c] = (FULL || in[c]) ? (double)x[c][s] : 0.0;
            int p = 0;
#pragma unroll
            for (int I = 0; I < NC; ++I)
#pragma unroll
                for (int J = I; J < NC; ++J, ++p) acc[p] = mfma4x4(d[I], d[J], acc[p]);
        }
    };
    // steps [k0, k1) of this wave (chunk w + k W); past k1 - 1 a load
    // re-reads chunk k1 - 1 into a buffer that is not used (no branch
    // around loads)
    auto sweep = [&](auto ntl, int k0, int k1) {
        constexpr bool NTL = decltype(ntl)::value;
        auto load_chunk = [&](int k, V(&x)[NC]) {
            const int64_t r = (int64_t)(w + (k < k1 ? k : k1 - 1) * W) * CR + rl;
#pragma unroll
            for (int c = 0; c < NC; ++c) x[c] = RW::template load<VEC, NTL>(A, lda, r, col[c]);
        };
        V a[NC], b[NC], c[NC];
        load_chunk(k0, a);
        load_chunk(k0 + 1, b);
#pragma unroll 1
        for (int k = k0; k < k1; k += 3) {
            load_chunk(k + 2, c);
            __builtin_amdgcn_sched_barrier(0);
            chunk(a);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 >= k1) break;
            load_chunk(k + 3, a);
            __builtin_amdgcn_sched_barrier(0);
            chunk(b);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 2 >= k1) break;
            load_chunk(k + 4, b);
            __builtin_amdgcn_sched_barrier(0);
            chunk(c);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    if (w < nfull) {
        const int nch = (nfull - w + W - 1) / W;
        // first step whose chunk lies in the cached bottom of the sweep
        int kc = (nfull - cached - w + W - 1) / W;
        kc = kc < 0 ? 0 : (kc > nch ? nch : kc);
        if (kc > 0) sweep(std::true_type{}, 0, kc);
        if (kc < nch) sweep(std::false_type{}, kc, nch);
    }
    if (m % CR && w == nfull % W) {  // the short last chunk
        V x[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const T* p = A + (int64_t)col[c] * lda + (int64_t)nfull * CR + rl;
#pragma unroll
            for (int s = 0; s < RPL; ++s) x[c][s] = ((int64_t)nfull * CR + rl + s < m) ? p[s] : T(0);
        }
        chunk(x);
    }
    gram4_finish<NC>(L, acc, partial, tk, &bcast);
}

// ---------------------------------------------------------------------------
// Q pass of the flat path: Q = A X, X = R^{-1} (FP64, row-major 32 x 32,
// upper triangular), written over A, one rounding per entry.  Does nothing
// unless ctl[0] is set (the whole matrix passed the gate).
// Lane l holds rows 4(l%16) .. +3 of the columns 4I + l/16 (I < NC); step s
// feeds row 4(l%16) + s as the A operand (block b = (l/4)%4, row index
// i = l%4, k = l/16).  The B operand of the tile pair (I, J), I <= J, is
// X[4I + l/16][4J + l%4], held in registers for the whole launch.  The
// result of column group J, step s, in lane l is Q[4(4((l/4)%4) + l/16) + s]
// [4J + l%4], so the four steps give each lane 4 consecutive rows of one
// column: one 16-byte store.  A persistent grid sweeps the chunks from the
// bottom of A up: the R pass read them last, so part of them is still in
// the memory-side cache.
// flat_q4_body is the sweep of wave w of W; tsqr_flat_q4 the launch.
template <int NC, bool VEC, bool NTL>
__device__ __forceinline__ void flat_q4_body(int m, int n, float* __restrict__ A, int lda,
                                             const double* __restrict__ X, int W, int w) {
    constexpr int NP = NC * (NC + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int kc = lane >> 4;
    const int rl = 4 * (lane & 15);
    const int jc = lane & 3;
    const int ro = 4 * (4 * ((lane >> 2) & 3) + (lane >> 4));  // first output row of this lane
    double mb[NP];
    {
        int p = 0;
#pragma unroll
        for (int I = 0; I < NC; ++I)
#pragma unroll
            for (int J = I; J < NC; ++J, ++p) {
                const int r = 4 * I + kc, c = 4 * J + jc;
                mb[p] = (r < n && c < n) ? X[r * 32 + c] : 0.0;
            }
    }
    bool in[NC], out[NC];
    int col[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        in[c] = 4 * c + kc < n;
        col[c] = in[c] ? 4 * c + kc : 0;
        out[c] = 4 * c + jc < n;
    }
    // acc[s][J]: row ro + s, column 4J + jc
    auto chunk = [&](const f4v(&x)[NC], double (&acc)[4][NC]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            double d[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) d[c] = in[c] ? (double)x[c][s] : 0.0;
#pragma unroll
            for (int J = 0; J < NC; ++J) acc[s][J] = 0.0;
            int p = 0;
#pragma unroll
            for (int I = 0; I < NC; ++I)
#pragma unroll
                for (int J = I; J < NC; ++J, ++p) acc[s][J] = mfma4x4(d[I], mb[p], acc[s][J]);
        }
    };
    auto store_chunk = [&](int64_t r0, const double (&acc)[4][NC]) {
#pragma unroll
        for (int J = 0; J < NC; ++J) {
            if (!out[J]) continue;
            float* p = A + (int64_t)(4 * J + jc) * lda + r0 + ro;
            const f4v q = {(float)acc[0][J], (float)acc[1][J], (float)acc[2][J], (float)acc[3][J]};
            if constexpr (VEC) {
                *reinterpret_cast<f4v*>(p) = q;
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) p[s] = q[s];
            }
        }
    };
    const int nfull = m / 64;
    if (m % 64 && w == 0) {  // the short last chunk, bounds-checked
        f4v x[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = load_rows4_tail(A, lda, (int64_t)nfull * 64 + rl, col[c], m);
        double acc[4][NC];
        chunk(x, acc);
        const int64_t r0 = (int64_t)nfull * 64;
#pragma unroll
        for (int J = 0; J < NC; ++J) {
            if (!out[J]) continue;
            float* p = A + (int64_t)(4 * J + jc) * lda + r0 + ro;
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (r0 + ro + s < m) p[s] = (float)acc[s][J];
        }
    }
    if (w >= nfull) return;
    // chunk index of this wave's k-th step, from the bottom up
    auto chunk_of = [&](int k) { return nfull - 1 - (w + k * W); };
    const int nch = (nfull - w + W - 1) / W;
    f4v a[NC], b[NC];
    auto load_chunk = [&](int ch, f4v(&x)[NC]) {
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = load_rows4<VEC, NTL>(A, lda, (int64_t)ch * 64 + rl, col[c]);
    };
    // ping-pong buffers as in the R pass; past the last chunk a step reloads
    // the last one into the buffer it will not use (no branch around loads)
    load_chunk(chunk_of(0), a);
#pragma unroll 1
    for (int k = 0; k < nch; k += 2) {
        double acc[4][NC];
        load_chunk(chunk_of(k + 1 < nch ? k + 1 : nch - 1), b);
        __builtin_amdgcn_sched_barrier(0);
        chunk(a, acc);
        store_chunk((int64_t)chunk_of(k) * 64, acc);
        __builtin_amdgcn_sched_barrier(0);
        if (k + 1 >= nch) break;
        load_chunk(chunk_of(k + 2 < nch ? k + 2 : nch - 1), a);
        __builtin_amdgcn_sched_barrier(0);
        chunk(b, acc);
        store_chunk((int64_t)chunk_of(k + 1) * 64, acc);
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int NC, bool VEC, bool NTL>
__global__ __launch_bounds__(256) void tsqr_flat_q4(int m, int n, float* __restrict__ A, int lda,
                                                    const double* __restrict__ X, const int* __restrict__ ctl) {
    if (!ctl[0]) return;
    flat_q4_body<NC, VEC, NTL>(m, n, A, lda, X, 4 * (int)gridDim.x, 4 * (int)blockIdx.x + (int)(threadIdx.x >> 6));
}

// ---------------------------------------------------------------------------
// Q pass of the CholeskyQR2 path on 4 x 4 tiles (FP64 data): Q = A X in
// place (X = R^{-1} of pass 1 or R2^{-1} of pass 2, FP64 row-major 32 x 32,
// upper triangular) with the layout of tsqr_flat_q4 for RPL rows per lane
// (lane l holds rows RPL(l%16) .. +RPL-1 of the columns 4I + l/16; the result
// of column group J, step s, is Q[RPL(4((l/4)%4) + l/16) + s][4J + l%4]),
// and in the same pass the Gram matrix of the stored Q on the same
// instruction: the result layout is the R pass's operand layout (lane
// 16k + 4b + z holds column 4J + z of one row per step), so step s adds
// q[I] q[J] to the tile pair (I, J), I <= J, and the partial sums end as
// tsqr_gram_stream4's (gram4_finish).  A persistent grid (wave w takes the
// chunks w, w + 4P, ...) with two chunks of loads in flight.  Runs when the
// flat path was taken (ctl[0]); pass 2 only if the check asked for it
// (ctl[1]).
template <int NC, bool VEC, typename T>
__global__ __launch_bounds__(256) void tsqr_q4_gram(int m, int n, T* __restrict__ A, int lda,
                                                    const double* __restrict__ X, double* __restrict__ partial,
                                                    int* __restrict__ tk, const int* __restrict__ ctl, int pass) {
    constexpr int NP = NC * (NC + 1) / 2;
    using RW = Rows16<T>;
    using V = typename RW::vec;
    constexpr int RPL = RW::RPL;
    constexpr int CR = 16 * RPL;
    __shared__ GramLds L;
    __shared__ int bcast;
    if (!ctl[0] || (pass == 2 && !ctl[1])) return;
    const int lane = threadIdx.x & 63;
    const int W = 4 * (int)gridDim.x;
    const int w = 4 * (int)blockIdx.x + (threadIdx.x >> 6);
    const int kc = lane >> 4;
    const int rl = RPL * (lane & 15);
    const int jc = lane & 3;
    const int ro = RPL * (4 * ((lane >> 2) & 3) + (lane >> 4));  // first output row of this lane
    double mb[NP];
    {
        int p = 0;
#pragma unroll
        for (int I = 0; I < NC; ++I)
#pragma unroll
            for (int J = I; J < NC; ++J, ++p) {
                const int r = 4 * I + kc, c = 4 * J + jc;
                mb[p] = (r < n && c < n) ? X[r * 32 + c] : 0.0;
            }
    }
    bool in[NC], out[NC];
    int col[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        in[c] = 4 * c + kc < n;
        col[c] = in[c] ? 4 * c + kc : 0;
        out[c] = 4 * c + jc < n;
    }
    double g[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) g[p] = 0.0;
    // q[s][J]: row ro + s, column 4J + jc, as stored (zero in padded columns
    // and rows past m)
    auto chunk = [&](const V(&x)[NC], T (&q)[RPL][NC]) {
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            double d[NC], a[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                d[c] = in[c] ? (double)x[c][s] : 0.0;
                a[c] = 0.0;
            }
            int p = 0;
#pragma unroll
            for (int I = 0; I < NC; ++I)
#pragma unroll
                for (int J = I; J < NC; ++J, ++p) a[J] = mfma4x4(d[I], mb[p], a[J]);
#pragma unroll
            for (int J = 0; J < NC; ++J) q[s][J] = (T)a[J];
            p = 0;
#pragma unroll
            for (int I = 0; I < NC; ++I)
#pragma unroll
                for (int J = I; J < NC; ++J, ++p) g[p] = mfma4x4((double)q[s][I], (double)q[s][J], g[p]);
        }
    };
    auto store_chunk = [&](int64_t r0, const T (&q)[RPL][NC]) {
#pragma unroll
        for (int J = 0; J < NC; ++J) {
            if (!out[J]) continue;
            T* p = A + (int64_t)(4 * J + jc) * lda + r0 + ro;
            if constexpr (VEC) {
                V v;
#pragma unroll
                for (int s = 0; s < RPL; ++s) v[s] = q[s][J];
                *reinterpret_cast<V*>(p) = v;
            } else {
#pragma unroll
                for (int s = 0; s < RPL; ++s) p[s] = q[s][J];
            }
        }
    };
    const int nfull = m / CR;
    if (m % CR && w == nfull % W) {  // the short last chunk, bounds-checked
        V x[NC];
        const int64_t r0 = (int64_t)nfull * CR;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const T* p = A + (int64_t)col[c] * lda + r0 + rl;
#pragma unroll
            for (int s = 0; s < RPL; ++s) x[c][s] = (r0 + rl + s < m) ? p[s] : T(0);
        }
        T q[RPL][NC];
        chunk(x, q);
#pragma unroll
        for (int J = 0; J < NC; ++J) {
            if (!out[J]) continue;
            T* p = A + (int64_t)(4 * J + jc) * lda + r0 + ro;
#pragma unroll
            for (int s = 0; s < RPL; ++s)
                if (r0 + ro + s < m) p[s] = q[s][J];
        }
    }
    if (w < nfull) {
        const int nch = (nfull - w + W - 1) / W;
        auto load_chunk = [&](int k, V(&x)[NC]) {
            const int64_t r = (int64_t)(w + (k < nch ? k : nch - 1) * W) * CR + rl;
#pragma unroll
            for (int c = 0; c < NC; ++c) x[c] = RW::template load<VEC, true>(A, lda, r, col[c]);
        };
        // ping-pong buffers; past the last chunk a step reloads the last one
        // into the buffer it will not use (no branch around loads)
        V a[NC], b[NC];
        load_chunk(0, a);
#pragma unroll 1
        for (int k = 0; k < nch; k += 2) {
            T q[RPL][NC];
            load_chunk(k + 1, b);
            __builtin_amdgcn_sched_barrier(0);
            chunk(a, q);
            store_chunk((int64_t)(w + k * W) * CR, q);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 >= nch) break;
            load_chunk(k + 2, a);
            __builtin_amdgcn_sched_barrier(0);
            chunk(b, q);
            store_chunk((int64_t)(w + (k + 1) * W) * CR, q);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    gram4_finish<NC>(L, g, partial, tk, &bcast);
}

}  // namespace tsqr
