// batch.hpp -- batched TSQR of many independent tall-skinny matrices of
// uneven m and n (n <= 64) in a fixed, short sequence of launches.
//
// A batch is planned on the host (batch.hip) into
//   panel problems  m x n' column blocks (n' <= 32) with their own tree
//                   (TreeDesc), workspace slice (WsLayout), A, R and ld's;
//                   a matrix with n <= 32 is one panel, one with
//                   32 < n <= 64 is two: columns [0, 32) and [32, n);
//   couples         for the latter, the coupling of the two panels:
//                   R12 = Q1^T A2 and A2 -= Q1 R12 (block Gram-Schmidt,
//                   block_QR.h:88-93), done once for the plain method and
//                   twice (BCGS2) for the mixed one;
//   item lists      (problem, index) pairs: every superleaf of every panel,
//                   every level-1 tree block, every tile of each tree level,
//                   every row block of every couple.
// Each phase below is one launch over one item list, so the GPU is filled
// by rows of all matrices at once; per-matrix steps that need the whole
// matrix (the sum of the superleaf Grams, the residual check, the sum of the
// coupling products) are done by the last item of the matrix to finish
// (an atomic ticket with agent-scope release / acquire), so no wave ever
// waits for another.  The whole sequence holds no host synchronisation and
// no memset, so it can be captured once into a hipGraph and replayed.
//
// Phases for one panel stage (T = float or double; "mixed" is FP32 only):
//   batch_init       tickets and control words of every problem = 0
//   batch_mixed_gram (mixed) BF16x3 superleaf Grams; the last superleaf of
//                    a matrix sums them and runs the flat Cholesky + gate
//   batch_super_r    Householder superleaves (every superleaf of a plain
//                    problem; of a mixed problem only if its gate failed)
//   batch_tree_up    the ticketed walk up each problem's tree
//   batch_down       one launch per tree level
//   batch_mixed_q    (mixed) Q = A R^{-1} with the fused residual Gram, then
//   batch_mixed_check  one wave per matrix sums it and runs the check; the
//                    pair runs again for the second pass where a check asked
//                    for it
//   batch_super_q    Householder Q for the Householder problems
#pragma once
#include <type_traits>

#include "mixed.hpp"
#include "super.hpp"

namespace tsqr {

struct BatchProb {
    TreeDesc d;
    WsLayout w;    // offsets inside this problem's workspace slice
    int64_t ws;    // byte offset of the slice in the batch workspace
    void* A;       // first element of the panel
    void* R;       // first element of the panel's diagonal R block
    int lda, ldr;
    int N, vec;    // padded columns; A and lda allow 16-byte loads
    int mixed;     // 1: mixed method (FP32), 0: Householder
    int pad;
};

// Coupling of the two panels of one matrix with 32 < n <= 64.
struct BatchCouple {
    void* Q1;        // columns [0, 32) after the first panel's TSQR
    void* A2;        // columns [32, n)
    void* R12;       // R rows [0, 32), columns [32, n)
    int64_t ws;      // byte offset: leaves + 2 FP64 32 x 32 blocks, then 4 ints
    int m, n2, lda, ldr;
    int leaves, leaf_rows;
    int vec, pad;
};

// Per-couple workspace: block b < leaves is superleaf b's product, block
// `leaves` the summed product P (the projection's right factor), block
// `leaves + 1` the accumulated R12 in FP64; then the tickets.
inline size_t couple_ws_bytes(int leaves) { return align256(((size_t)leaves + 2) * 1024 * sizeof(double) + 64); }

template <int V>
using IntC = std::integral_constant<int, V>;
template <bool V>
using BoolC = std::integral_constant<bool, V>;

// Calls fn(IntC<N>, BoolC<VEC>) for the problem's padded width and alignment.
template <typename Fn>
__device__ __forceinline__ void dispatch_nv(int N, int vec, Fn&& fn) {
    if (N == 8) {
        if (vec) fn(IntC<8>{}, BoolC<true>{});
        else fn(IntC<8>{}, BoolC<false>{});
    } else if (N == 16) {
        if (vec) fn(IntC<16>{}, BoolC<true>{});
        else fn(IntC<16>{}, BoolC<false>{});
    } else {
        if (vec) fn(IntC<32>{}, BoolC<true>{});
        else fn(IntC<32>{}, BoolC<false>{});
    }
}

__device__ __forceinline__ char* prob_ws(const BatchProb& p, char* work) { return work + p.ws; }

template <typename T, int N>
struct SuperBytes {
    static constexpr int RPL = rows_per_lane<T>(N);
    static constexpr size_t value = sizeof(SuperLds<T, N, RPL>) + (size_t)N * 64 * RPL * sizeof(T);
};
template <typename T>
constexpr size_t super_lds_bytes() {
    constexpr size_t a = SuperBytes<T, 8>::value, b = SuperBytes<T, 16>::value, c = SuperBytes<T, 32>::value;
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}
template <typename T>
constexpr int batch_super_threads() {
    constexpr int a = super_tiles<T>(8), b = super_tiles<T>(16), c = super_tiles<T>(32);
    return 64 * (a > b ? (a > c ? a : c) : (b > c ? b : c));
}

// One wave adds `count` 32 x 32 FP64 blocks at src, in order, into L.G.
__device__ __forceinline__ void sum_blocks_wave(GramLds& L, const double* __restrict__ src, int count, int lane) {
    double s[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) s[e] = 0.0;
#pragma unroll 1
    for (int b = 0; b < count; ++b) {
        const double* G = src + (int64_t)b * 1024;
#pragma unroll
        for (int e = 0; e < 16; ++e) s[e] += G[lane + 64 * e];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = lane + 64 * e;
        L.G[idx >> 5][idx & 31] = s[e];
    }
    wave_sync();
}

// True for the wave whose increment completes `total` arrivals on *ticket;
// its later reads see every earlier arriver's writes.
__device__ __forceinline__ bool last_arriver(int* ticket, int total, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int old = 0;
    if (lane == 0) old = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    old = __builtin_amdgcn_readfirstlane(old);
    if (old != total - 1) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// ctl words of a problem (int): 0 flat CholeskyQR taken, 1 second pass
// asked, 2 check status, 8 Gram ticket.
constexpr int kCtlGramTicket = 8;

// ---------------------------------------------------------------------------
// Problem set-up: tickets and control words = 0; couples' tickets = 0;
// R21 = 0 for the two-panel matrices.  One workgroup per problem / couple.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void batch_init(const BatchProb* __restrict__ probs, int nprobs,
                                                   const BatchCouple* __restrict__ couples, int ncouples,
                                                   char* __restrict__ work) {
    const int b = blockIdx.x;
    if (b < nprobs) {
        const BatchProb& p = probs[b];
        char* ws = prob_ws(p, work);
        int* tickets = reinterpret_cast<int*>(ws + p.w.tickets);
        int* ctl = reinterpret_cast<int*>(ws + p.w.ctl);
        for (int i = threadIdx.x; i < p.d.total_ctr; i += blockDim.x) tickets[i] = 0;
        if (threadIdx.x < 64) ctl[threadIdx.x] = 0;
        return;
    }
    const int c = b - nprobs;
    if (c >= ncouples) return;
    const BatchCouple& q = couples[c];
    int* tk = reinterpret_cast<int*>(work + q.ws + ((size_t)q.leaves + 2) * 1024 * sizeof(double));
    if (threadIdx.x < 16) tk[threadIdx.x] = 0;
    // R21: rows [32, 32 + n2) of columns [0, 32), i.e. relative to R12 at
    // (0, 32): rows 32.., columns -32..-1
    T* R21 = reinterpret_cast<T*>(q.R12) + 32 - (int64_t)32 * q.ldr;
    for (int i = threadIdx.x; i < 32 * q.n2; i += blockDim.x) R21[(i % q.n2) + (int64_t)(i / q.n2) * q.ldr] = T(0);
}

// ---------------------------------------------------------------------------
// Householder superleaves (plain problems, and mixed ones whose gate failed).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(batch_super_threads<T>()) void batch_super_r(const BatchProb* __restrict__ probs,
                                                                         const int2* __restrict__ items,
                                                                         char* __restrict__ work) {
    __shared__ __align__(16) char smem[super_lds_bytes<T>()];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    if (reinterpret_cast<const int*>(ws + p.w.ctl)[0]) return;
    T* blocks = reinterpret_cast<T*>(ws + p.w.blocks);
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int N = decltype(NC)::value;
        constexpr bool VEC = decltype(VC)::value;
        constexpr int RPL = rows_per_lane<T>(N);
        auto& S = *reinterpret_cast<SuperLds<T, N, RPL>*>(smem);
        super_r_leaf<T, N, RPL, VEC>(p.d, reinterpret_cast<const T*>(p.A), p.lda, reinterpret_cast<T*>(p.R), p.ldr,
                                     blocks, S, it.y);
    });
}

template <typename T>
__global__ __launch_bounds__(batch_super_threads<T>()) void batch_super_q(const BatchProb* __restrict__ probs,
                                                                         const int2* __restrict__ items,
                                                                         char* __restrict__ work) {
    __shared__ __align__(16) char smem[super_lds_bytes<T>()];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    if (reinterpret_cast<const int*>(ws + p.w.ctl)[0]) return;
    const T* blocks = reinterpret_cast<const T*>(ws + p.w.blocks);
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int N = decltype(NC)::value;
        constexpr bool VEC = decltype(VC)::value;
        constexpr int RPL = rows_per_lane<T>(N);
        auto& S = *reinterpret_cast<SuperLds<T, N, RPL>*>(smem);
        T* stash = reinterpret_cast<T*>(smem + sizeof(SuperLds<T, N, RPL>));
        super_q_leaf<T, N, RPL, VEC>(p.d, reinterpret_cast<T*>(p.A), p.lda, blocks, S, stash, it.y);
    });
}

// ---------------------------------------------------------------------------
// Tree: walk up from every level-1 block; down sweep of one level.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void batch_tree_up(const BatchProb* __restrict__ probs,
                                                     const int2* __restrict__ items, char* __restrict__ work) {
    __shared__ __align__(16) char smem[sizeof(TileLds<T, 32>)];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    if (reinterpret_cast<const int*>(ws + p.w.ctl)[0]) return;
    T* blocks = reinterpret_cast<T*>(ws + p.w.blocks);
    int* tickets = reinterpret_cast<int*>(ws + p.w.tickets);
    dispatch_nv(p.N, 1, [&](auto NC, auto) {
        constexpr int N = decltype(NC)::value;
        auto& L = *reinterpret_cast<TileLds<T, N>*>(smem);
        tree_up_body<T, N, rows_per_lane<T>(N)>(p.d, reinterpret_cast<T*>(p.R), p.ldr, blocks, tickets, L, it.y);
    });
}

template <typename T>
__global__ __launch_bounds__(64) void batch_down(const BatchProb* __restrict__ probs, const int2* __restrict__ items,
                                                  int level, char* __restrict__ work) {
    __shared__ __align__(16) char smem[sizeof(TileLds<T, 32>)];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    if (reinterpret_cast<const int*>(ws + p.w.ctl)[0]) return;
    T* blocks = reinterpret_cast<T*>(ws + p.w.blocks);
    dispatch_nv(p.N, 1, [&](auto NC, auto) {
        constexpr int N = decltype(NC)::value;
        auto& L = *reinterpret_cast<TileLds<T, N>*>(smem);
        down_body<T, N, rows_per_lane<T>(N)>(p.d, level, blocks, L, it.y);
    });
}

// ---------------------------------------------------------------------------
// Mixed method (FP32).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double* prob_side(const BatchProb& p, char* work) {
    return reinterpret_cast<double*>(prob_ws(p, work) + p.w.side);
}

template <int UNUSED = 0>
__global__ __launch_bounds__(64) void batch_mixed_gram(const BatchProb* __restrict__ probs,
                                                        const int2* __restrict__ items, char* __restrict__ work,
                                                        double gate) {
    __shared__ double sums[12 * 64];
    __shared__ GramLds L;
    const int lane = threadIdx.x;
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    double* side = prob_side(p, work);
    const LevelDesc& L0 = p.d.lv[0];
    const int n = p.d.n;
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int NT = decltype(NC)::value > 16 ? 2 : 1;
        mixed_gram_body<NT, decltype(VC)::value>(L0, n, reinterpret_cast<const float*>(p.A), p.lda,
                                                 side + (int64_t)it.y * 1024, sums, it.y);
    });
    int* ctl = reinterpret_cast<int*>(ws + p.w.ctl);
    if (!last_arriver(ctl + kCtlGramTicket, L0.tiles, lane)) return;
    sum_blocks_wave(L, side, L0.tiles, lane);
    const int64_t T0 = L0.tiles;
    flat_factor<float>(L, n, reinterpret_cast<float*>(p.R), p.ldr, side + T0 * 1024, ctl, gate,
                       side + (T0 + 1) * 1024, lane);
}

// (at least two waves per SIMD: the register budget of the single-matrix
// tsqr_mixed_q)
template <int UNUSED = 0>
__global__ __launch_bounds__(64, 2) void batch_mixed_q(const BatchProb* __restrict__ probs,
                                                     const int2* __restrict__ items, char* __restrict__ work,
                                                     int pass) {
    __shared__ double sums[12 * 64];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    const int* ctl = reinterpret_cast<const int*>(ws + p.w.ctl);
    if (!p.mixed || !ctl[0] || (pass == 2 && !ctl[1])) return;
    double* side = prob_side(p, work);
    const LevelDesc& L0 = p.d.lv[0];
    const int64_t T0 = L0.tiles;
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int NT = decltype(NC)::value > 16 ? 2 : 1;
        mixed_q_body<NT, decltype(VC)::value>(L0, p.d.n, reinterpret_cast<float*>(p.A), p.lda, side + T0 * 1024,
                                              side + (int64_t)it.y * 1024, sums, it.y);
    });
}

// Residual check of one problem after a Q pass (one wave per problem; a
// launch of its own so that the Q pass keeps its registers for streaming).
template <int UNUSED = 0>
__global__ __launch_bounds__(64) void batch_mixed_check(const BatchProb* __restrict__ probs,
                                                         const int2* __restrict__ list, char* __restrict__ work,
                                                         int pass, double tol) {
    __shared__ GramLds L;
    const int lane = threadIdx.x;
    const BatchProb& p = probs[list[blockIdx.x].x];
    char* ws = prob_ws(p, work);
    int* ctl = reinterpret_cast<int*>(ws + p.w.ctl);
    if (!p.mixed || !ctl[0] || (pass == 2 && !ctl[1])) return;
    double* side = prob_side(p, work);
    const int64_t T0 = p.d.lv[0].tiles;
    sum_blocks_wave(L, side, (int)T0, lane);
    mixed_check_body(L, p.d.n, reinterpret_cast<float*>(p.R), p.ldr, side + T0 * 1024, side + (T0 + 1) * 1024,
                     side + (T0 + 2) * 1024, ctl, pass, tol, lane);
}

// ---------------------------------------------------------------------------
// Coupling of two panels: P = Q1^T A2 over a couple's rows (FP64 MFMA, FP32
// or FP64 inputs exact in FP64), summed by the last row block into R12;
// then A2 -= Q1 P.  `pass` 2 (mixed method) adds the second product to R12.
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__device__ __forceinline__ void cross_block(const BatchCouple& q, int blk, double* __restrict__ out, int lane) {
    const int col = lane & 15;
    const int rq = 4 * (lane >> 4);
    const int64_t r0 = (int64_t)blk * q.leaf_rows;
    const int64_t left = q.m - r0;
    const int rows = (int)(left < q.leaf_rows ? left : q.leaf_rows);
    const T* Qb = reinterpret_cast<const T*>(q.Q1) + r0;
    const T* Ab = reinterpret_cast<const T*>(q.A2) + r0;
    v4d c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
#pragma unroll 1
    for (int s0 = 0; s0 < rows; s0 += 16) {
        double q0[4], q1[4], a0[4], a1[4];
        load4<T, VEC>(q0, Qb, q.lda, s0 + rq, col, rows, 32);
        load4<T, VEC>(q1, Qb, q.lda, s0 + rq, col + 16, rows, 32);
        load4<T, VEC>(a0, Ab, q.lda, s0 + rq, col, rows, q.n2);
        load4<T, VEC>(a1, Ab, q.lda, s0 + rq, col + 16, rows, q.n2);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(q0[t], a0[t], c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(q0[t], a1[t], c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(q1[t], a0[t], c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(q1[t], a1[t], c11, 0, 0, 0);
        }
    }
    // f64 accumulator: lane l, register v holds C[4v + l/16][l%16]
    const int i0 = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = 4 * v + i0;
        out[i * 32 + col] = c00[v];
        out[i * 32 + 16 + col] = c01[v];
        out[(16 + i) * 32 + col] = c10[v];
        out[(16 + i) * 32 + 16 + col] = c11[v];
    }
}

template <typename T>
__global__ __launch_bounds__(64) void batch_cross(const BatchCouple* __restrict__ couples,
                                                   const int2* __restrict__ items, char* __restrict__ work,
                                                   int pass) {
    __shared__ GramLds L;
    const int lane = threadIdx.x;
    const int2 it = items[blockIdx.x];
    const BatchCouple& q = couples[it.x];
    double* blocks = reinterpret_cast<double*>(work + q.ws);
    int* tk = reinterpret_cast<int*>(work + q.ws + ((size_t)q.leaves + 2) * 1024 * sizeof(double));
    if (q.vec)
        cross_block<T, true>(q, it.y, blocks + (int64_t)it.y * 1024, lane);
    else
        cross_block<T, false>(q, it.y, blocks + (int64_t)it.y * 1024, lane);
    if (!last_arriver(tk + pass - 1, q.leaves, lane)) return;
    sum_blocks_wave(L, blocks, q.leaves, lane);
    double* P = blocks + (int64_t)q.leaves * 1024;
    double* Racc = blocks + (int64_t)(q.leaves + 1) * 1024;
    T* R12 = reinterpret_cast<T*>(q.R12);
#pragma unroll 1
    for (int e = 0; e < 16; ++e) {
        const int idx = lane + 64 * e;
        const int i = idx >> 5, j = idx & 31;
        const double v = L.G[i][j];
        P[idx] = v;
        const double r = (pass == 1 ? 0.0 : Racc[idx]) + v;
        Racc[idx] = r;
        if (j < q.n2) R12[i + (int64_t)j * q.ldr] = (T)r;
    }
}

// A2 rows of one block -= Q1 rows * P (32 x n2, FP64): FP64 MFMA with the
// accumulators initialised from A2 and B = -P, one rounding.
template <typename T, bool VEC>
__device__ __forceinline__ void project_block(const BatchCouple& q, int blk, const double* __restrict__ P, int lane) {
    const int li = lane & 15, lq = lane >> 4;
    const int64_t r0 = (int64_t)blk * q.leaf_rows;
    const int64_t left = q.m - r0;
    const int rows = (int)(left < q.leaf_rows ? left : q.leaf_rows);
    const T* Qb = reinterpret_cast<const T*>(q.Q1) + r0;
    T* Ab = reinterpret_cast<T*>(q.A2) + r0;
    double mb[8][2];  // -P in the B-operand layout of k-group kg (k = 4 kg + l/16)
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) mb[kg][jt] = -P[(4 * kg + lq) * 32 + 16 * jt + li];
#pragma unroll 1
    for (int s0 = 0; s0 < rows; s0 += 64) {
        // acc[t][jt] register v: row s0 + 16v + 4 lq + t, column 16 jt + li
        v4d acc[4][2];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double a[4];
                load4<T, VEC>(a, Ab, q.lda, s0 + 16 * v + 4 * lq, 16 * jt + li, rows, q.n2);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t][jt][v] = a[t];
            }
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) {
            double x[4];
            load4<T, VEC>(x, Qb, q.lda, s0 + 4 * li, 4 * kg + lq, rows, 32);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
                    acc[t][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[t], mb[kg][jt], acc[t][jt], 0, 0, 0);
        }
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int cc = 16 * jt + li;
            if (cc < q.n2) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    T o[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) o[t] = (T)acc[t][jt][v];
                    store_column<T, 4, VEC>(o, Ab + (int64_t)cc * q.lda + s0 + 16 * v + 4 * lq, rows,
                                            (int)(s0 + 16 * v + 4 * lq));
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64) void batch_project(const BatchCouple* __restrict__ couples,
                                                     const int2* __restrict__ items, char* __restrict__ work) {
    const int2 it = items[blockIdx.x];
    const BatchCouple& q = couples[it.x];
    const double* P = reinterpret_cast<const double*>(work + q.ws) + (int64_t)q.leaves * 1024;
    if (q.vec)
        project_block<T, true>(q, it.y, P, threadIdx.x);
    else
        project_block<T, false>(q, it.y, P, threadIdx.x);
}

}  // namespace tsqr
