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
//                   block_QR.h:88-93), done once for the plain method; the
//                   mixed method repeats it (BCGS2) for the matrices where
//                   the first projection removed more than half of some
//                   column's squared norm;
//   item lists      (problem, index) pairs: every superleaf of every panel,
//                   every level-1 tree block, every tile of each tree level,
//                   every row block of every couple.
// Each phase below is one launch over one item list, so the GPU is filled
// by rows of all matrices at once; the per-matrix steps that need the whole
// matrix (the sum of the superleaf Grams and its Cholesky, the residual
// check, the sum of the coupling products) are one wave per matrix in a
// launch of their own, so that the streaming launches keep their registers
// for loads in flight and no wave ever waits for another.  The whole
// sequence holds no host synchronisation and no memset, so it can be
// captured once into a hipGraph and replayed.
//
// Phases for one panel stage (T = float or double; "mixed" is FP32 only):
//   batch_init       tickets and control words of every problem = 0
//   batch_mixed_gram (mixed) BF16x3 superleaf Grams, then
//   batch_mixed_flat one wave per matrix sums them and runs the flat
//                    Cholesky + gate
//   batch_leaf_r     (plain) Householder leaves of one tile, one wave each
//                    (leaf1.hpp), reflectors stored over A
//   batch_super_r    (mixed) Householder superleaves of the mixed problems
//                    whose gate failed
//   batch_tree_up    (mixed) the ticketed walk up each problem's tree
//   batch_tree_level (plain) one tree level per launch
//   batch_down       one launch per tree level
//   batch_mixed_q    (mixed) Q = A R^{-1} with the fused residual Gram, then
//   batch_mixed_check  one wave per matrix sums it and runs the check; the
//                    pair runs again for the second pass where a check asked
//                    for it
//   batch_leaf_q     (plain) Q of every tile from its stored reflectors
//   batch_super_q    (mixed) Householder Q for the problems whose gate failed
#pragma once
#include <type_traits>

#include "gemm_mfma.hpp"
#include "leaf1.hpp"
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
    int pre_gram;  // 1: batch_project already wrote the superleaf Grams
};

// Coupling of the two panels of one matrix with 32 < n <= 64.
struct BatchCouple {
    void* Q1;        // columns [0, 32) after the first panel's TSQR
    void* A2;        // columns [32, n)
    void* R12;       // R rows [0, 32), columns [32, n)
    int64_t ws;      // byte offset: leaves + 2 FP64 32 x 32 blocks, then 4 ints
    int m, n2, lda, ldr;
    int leaves, leaf_rows;
    int vec;
    int mixed;      // 1: mixed method (FP32): BF16x3 coupling products
    int gram;       // 1: batch_project also writes the BF16x3 Gram matrix of
                    // each new A2 row block, the second panel's superleaf b,
                    // to the second panel's side block b (at byte side2)
    int64_t side2;
};

// Per-couple workspace: block b < leaves is row block b's product, block
// `leaves` the summed product P (the projection's right factor), block
// `leaves + 1` the accumulated R12 in FP64; then the squared column norms of
// A2 per row block (32 doubles each); then kCoupleInts ints (the
// reorthogonalisation flag).
__host__ __device__ inline size_t couple_norms_off(int leaves) { return ((size_t)leaves + 2) * 1024 * sizeof(double); }
__host__ __device__ inline size_t couple_ints_off(int leaves) {
    return couple_norms_off(leaves) + (size_t)leaves * 32 * sizeof(double);
}
constexpr int kCoupleReorth = 4;  // int index of the flag
constexpr int kCoupleInts = 16;
inline size_t couple_ws_bytes(int leaves) { return align256(couple_ints_off(leaves) + kCoupleInts * sizeof(int)); }

// Row blocks of the streaming batch passes (Gram, Q, coupling products and
// projection): a panel with S superleaves (a couple with S row blocks) is
// worked on by batch_groups(S) workgroups of four waves; wave w of
// workgroup g takes the rows of block 4g + w of 4 batch_groups(S) (the
// strided chunks of mixed_gram_body), and the workgroup adds its four
// waves' Gram or product blocks in wave order into block g, so the
// per-matrix sums read batch_groups(S) blocks.
__host__ __device__ inline int batch_groups(int S) { return (S + 1) / 2; }

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

// ctl words of a problem (int): 0 flat CholeskyQR taken, 1 second pass
// asked, 2 check status, 3 / 4 BF16 Q passes (mixed.hpp).

// ---------------------------------------------------------------------------
// Problem set-up: tickets and control words = 0; couples' tickets = 0;
// R21 = 0 for the two-panel matrices; the two Householder fallback lists
// (count, pad, then int2 items, at byte offsets list0 / list1) empty.  One
// workgroup per problem / couple.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void batch_init(const BatchProb* __restrict__ probs, int nprobs,
                                                   const BatchCouple* __restrict__ couples, int ncouples,
                                                   char* __restrict__ work, int64_t list0, int64_t list1) {
    const int b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0) {
        *reinterpret_cast<int*>(work + list0) = 0;
        *reinterpret_cast<int*>(work + list1) = 0;
    }
    if (b < nprobs) {
        const BatchProb& p = probs[b];
        char* ws = prob_ws(p, work);
        int* tickets = reinterpret_cast<int*>(ws + p.w.tickets);
        int* ctl = reinterpret_cast<int*>(ws + p.w.ctl);
        for (int i = threadIdx.x; i < p.d.total_ctr; i += blockDim.x) tickets[i] = 0;
        for (int i = threadIdx.x; i < kCtlWords; i += blockDim.x) ctl[i] = 0;
        return;
    }
    const int c = b - nprobs;
    if (c >= ncouples) return;
    const BatchCouple& q = couples[c];
    int* tk = reinterpret_cast<int*>(work + q.ws + couple_ints_off(q.leaves));
    if (threadIdx.x < kCoupleInts) tk[threadIdx.x] = 0;
    // R21: rows [32, 32 + n2) of columns [0, 32), i.e. relative to R12 at
    // (0, 32): rows 32.., columns -32..-1
    T* R21 = reinterpret_cast<T*>(q.R12) + 32 - (int64_t)32 * q.ldr;
    for (int i = threadIdx.x; i < 32 * q.n2; i += blockDim.x) R21[(i % q.n2) + (int64_t)(i / q.n2) * q.ldr] = T(0);
}

// ---------------------------------------------------------------------------
// Householder superleaves (plain problems, and mixed ones whose gate failed).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void super_r_item(const BatchProb* __restrict__ probs, int2 it, char* __restrict__ work,
                                             char* smem) {
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
__device__ __forceinline__ void super_q_item(const BatchProb* __restrict__ probs, int2 it, char* __restrict__ work,
                                             char* smem) {
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

// Both kernels take either a fixed item list (LIST = false: one item per
// workgroup) or a list built on the device (LIST = true: *count items,
// grid-stride), the superleaves of the mixed problems whose gate failed.
template <typename T, bool LIST>
__global__ __launch_bounds__(batch_super_threads<T>()) void batch_super_r(const BatchProb* __restrict__ probs,
                                                                         const int2* __restrict__ items,
                                                                         const int* __restrict__ count,
                                                                         char* __restrict__ work) {
    __shared__ __align__(16) char smem[super_lds_bytes<T>()];
    if constexpr (LIST) {
#pragma unroll 1
        for (int i = blockIdx.x; i < count[0]; i += gridDim.x) {
            super_r_item<T>(probs, items[i], work, smem);
            __syncthreads();
        }
    } else {
        super_r_item<T>(probs, items[blockIdx.x], work, smem);
    }
}

template <typename T, bool LIST>
__global__ __launch_bounds__(batch_super_threads<T>()) void batch_super_q(const BatchProb* __restrict__ probs,
                                                                         const int2* __restrict__ items,
                                                                         const int* __restrict__ count,
                                                                         char* __restrict__ work) {
    __shared__ __align__(16) char smem[super_lds_bytes<T>()];
    if constexpr (LIST) {
#pragma unroll 1
        for (int i = blockIdx.x; i < count[0]; i += gridDim.x) {
            super_q_item<T>(probs, items[i], work, smem);
            __syncthreads();
        }
    } else {
        super_q_item<T>(probs, items[blockIdx.x], work, smem);
    }
}

// Plain problems: leaves of one tile (one item per tile), one wave each.
template <typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void batch_leaf_r(
    const BatchProb* __restrict__ probs, const int2* __restrict__ items, char* __restrict__ work) {
    __shared__ __align__(16) char smem[sizeof(TileLds<T, 32>)];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    T* blocks = reinterpret_cast<T*>(prob_ws(p, work) + p.w.blocks);
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int N = decltype(NC)::value;
        constexpr bool VEC = decltype(VC)::value;
        auto& L = *reinterpret_cast<TileLds<T, N>*>(smem);
        if (p.d.lv[0].tiles == 1) {  // one tile: R to the caller, Q over A
            tsqr_single_body<T, N, rows_per_lane<T>(N), VEC>(p.d, reinterpret_cast<T*>(p.A), p.lda,
                                                              reinterpret_cast<T*>(p.R), p.ldr, L);
            return;
        }
        leaf_r_tile<T, N, rows_per_lane<T>(N), VEC, true>(p.d, reinterpret_cast<T*>(p.A), p.lda, blocks, L, it.y);
    });
}

template <typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void batch_leaf_q(
    const BatchProb* __restrict__ probs, const int2* __restrict__ items, char* __restrict__ work) {
    __shared__ __align__(16) char smem[sizeof(TileLds<T, 32>)];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    if (p.d.lv[0].tiles == 1) return;  // done by batch_leaf_r
    const T* blocks = reinterpret_cast<const T*>(prob_ws(p, work) + p.w.blocks);
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int N = decltype(NC)::value;
        constexpr bool VEC = decltype(VC)::value;
        auto& L = *reinterpret_cast<TileLds<T, N>*>(smem);
        leaf_q_tile<T, N, rows_per_lane<T>(N), VEC>(p.d, reinterpret_cast<T*>(p.A), p.lda,
                                                    leaf_s_block<T>(p.d, blocks, it.y), p.d.H, L, it.y);
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

// Plain problems: one tree level per launch over every problem deep enough
// (no tickets: an agent-scope release per block is an L2 write-back).
template <typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void batch_tree_level(
    const BatchProb* __restrict__ probs, const int2* __restrict__ items, int level, char* __restrict__ work) {
    __shared__ __align__(16) char smem[sizeof(TileLds<T, 32>)];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    T* blocks = reinterpret_cast<T*>(prob_ws(p, work) + p.w.blocks);
    dispatch_nv(p.N, 1, [&](auto NC, auto) {
        constexpr int N = decltype(NC)::value;
        auto& L = *reinterpret_cast<TileLds<T, N>*>(smem);
        tree_level_body<T, N, rows_per_lane<T>(N)>(p.d, level, reinterpret_cast<T*>(p.R), p.ldr, blocks, L, it.y);
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

// Item (problem, g), g < batch_groups(S): four waves, wave w on row block
// 4g + w of 4 batch_groups(S) (mixed_gram_body's strided chunks); the
// workgroup adds its four Gram slabs in wave order into side block g.
template <int UNUSED = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void batch_mixed_gram(
    const BatchProb* __restrict__ probs, const int2* __restrict__ items, char* __restrict__ work) {
    __shared__ double sums[4][12 * 64];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    if (p.pre_gram) return;  // written by batch_project
    double* side = prob_side(p, work);
    const int G = batch_groups(p.d.lv[0].tiles);
    const int wv = threadIdx.x >> 6;
    const int n = p.d.n;
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int NT = decltype(NC)::value > 16 ? 2 : 1;
        mixed_gram_body<NT, decltype(VC)::value>(p.d.m, n, reinterpret_cast<const float*>(p.A), p.lda, sums[wv],
                                                 4 * it.y + wv, 4 * G);
    });
    __syncthreads();
    gram32_combine(sums, side + (int64_t)it.y * 1024);
}

// One workgroup per problem, after batch_mixed_gram: the sum of its
// superleaf Grams (four waves, fixed order), then by one wave the flat
// Cholesky and the gate (a launch of its own so that the streaming Gram
// keeps its registers).  If the gate fails, the
// problem's superleaves go to the Householder list.
template <int UNUSED = 0>
__global__ __launch_bounds__(256) void batch_mixed_flat(const BatchProb* __restrict__ probs,
                                                         const int2* __restrict__ list, char* __restrict__ work,
                                                         double gate, int* __restrict__ fallback) {
    __shared__ GramLds L;
    const int lane = threadIdx.x;
    const int pi = list[blockIdx.x].x;
    const BatchProb& p = probs[pi];
    char* ws = prob_ws(p, work);
    double* side = prob_side(p, work);
    const LevelDesc& L0 = p.d.lv[0];
    int* ctl = reinterpret_cast<int*>(ws + p.w.ctl);
    sum_blocks_wg4(L, side, batch_groups(L0.tiles));
    if (threadIdx.x >= 64) return;
    const int64_t T0 = L0.tiles;
    const bool ok = flat_factor<float>(L, p.d.n, reinterpret_cast<float*>(p.R), p.ldr, side + T0 * 1024, ctl, gate,
                                       side + (T0 + 1) * 1024, lane, true);
    if (!ok) {
        int at = 0;
        if (lane == 0) at = __hip_atomic_fetch_add(fallback, L0.tiles, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        at = __builtin_amdgcn_readfirstlane(at);
        int2* dst = reinterpret_cast<int2*>(fallback + 2) + at;
        for (int t = lane; t < L0.tiles; t += 64) dst[t] = make_int2(pi, t);
    }
}

// Item (problem, g) by four waves, row blocks as in batch_mixed_gram; the
// Gram slabs of the stored rows go to side block g.
// (Two waves per SIMD: the register budget of the single-matrix Q pass.)
template <int UNUSED = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void batch_mixed_q(
    const BatchProb* __restrict__ probs, const int2* __restrict__ items, char* __restrict__ work, int pass) {
    __shared__ double sums[4][12 * 64];
    __shared__ float tr[4][32 * kTrPitch];
    const int2 it = items[blockIdx.x];
    const BatchProb& p = probs[it.x];
    char* ws = prob_ws(p, work);
    const int* ctl = reinterpret_cast<const int*>(ws + p.w.ctl);
    if (!p.mixed || !ctl[0] || (pass == 2 && !ctl[1])) return;
    double* side = prob_side(p, work);
    const int64_t T0 = p.d.lv[0].tiles;
    const int G = batch_groups((int)T0);
    const int wv = threadIdx.x >> 6;
    const bool bf16q = ctl[pass == 1 ? kCtlBf16Q1 : kCtlBf16Q2] != 0;
    dispatch_nv(p.N, p.vec, [&](auto NC, auto VC) {
        constexpr int NT = decltype(NC)::value > 16 ? 2 : 1;
        mixed_q_leaf<NT, decltype(VC)::value>(p.d.m, p.d.n, reinterpret_cast<float*>(p.A), p.lda, side + T0 * 1024,
                                              sums[wv], tr[wv], 4 * it.y + wv, 4 * G, bf16q);
    });
    __syncthreads();
    gram32_combine(sums, side + (int64_t)it.y * 1024);
}

// Residual check of one problem after a Q pass (one workgroup per problem:
// four waves sum the superleaf Grams, one runs the check; a launch of its
// own so that the Q pass keeps its registers for streaming).
template <int UNUSED = 0>
__global__ __launch_bounds__(256) void batch_mixed_check(const BatchProb* __restrict__ probs,
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
    sum_blocks_wg4(L, side, batch_groups((int)T0));
    if (threadIdx.x >= 64) return;
    mixed_check_body<float>(L, p.d.n, reinterpret_cast<float*>(p.R), p.ldr, side + T0 * 1024,
                            side + (T0 + 1) * 1024, side + (T0 + 2) * 1024, ctl, pass, tol, lane);
}

// ---------------------------------------------------------------------------
// Coupling of two panels: P = Q1^T A2 over a couple's rows (plain method:
// FP64 MFMA, FP32 or FP64 inputs exact in FP64; mixed method: BF16x3 with
// FP32 sums per 64 rows), summed per couple into R12 (batch_cross_sum); then
// A2 -= Q1 P.  `pass` 2 (mixed method) adds the second product to R12.
// ---------------------------------------------------------------------------

// P_b = Q1_b^T A2_b for row block `blk` of S, written to out (FP64 32 x 32,
// row-major), and the block's squared column norms of A2 to norms.  The row
// blocks partition the 64-row chunks by stride (block j: chunks j, j + S,
// ...), so a couple's waves sweep it together; only the sums over all
// blocks are used.  Lane l loads rows 4(l/16)..+3 of
// columns l%16 and 16 + l%16 of both matrices in each of a chunk's four
// 16-row slabs; the next chunk is in flight while the current one is
// multiplied.
template <typename T, bool VEC>
__device__ __forceinline__ void cross_block(const BatchCouple& q, int blk, int S, double* out, double* norms,
                                            int lane) {
    const int col = lane & 15;
    const int rq = 4 * (lane >> 4);
    const int chunks = (q.m + 63) / 64;
    const T* Q1 = reinterpret_cast<const T*>(q.Q1);
    const T* A2 = reinterpret_cast<const T*>(q.A2);
    v4d c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    double s0sq = 0.0, s1sq = 0.0;  // this lane's part of the squared column norms of A2
    auto mma4 = [&](const Raw4<T>& q0, const Raw4<T>& q1, const Raw4<T>& a0, const Raw4<T>& a1) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double x0 = (double)q0.v[t], x1 = (double)q1.v[t];
            const double y0 = (double)a0.v[t], y1 = (double)a1.v[t];
            s0sq = fmat(y0, y0, s0sq);
            s1sq = fmat(y1, y1, s1sq);
            c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y0, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y1, c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y0, c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y1, c11, 0, 0, 0);
        }
    };
    constexpr int D = 4;  // 16-row slabs per chunk
    auto load = [&](int ch, Raw4<T>(&x0)[D], Raw4<T>(&x1)[D], Raw4<T>(&y0)[D], Raw4<T>(&y1)[D]) {
        const int64_t r0 = (int64_t)ch * 64;
        const int lim = (int)(q.m - r0 < 64 ? q.m - r0 : 64);
        const T* Qb = Q1 + r0;
        const T* Ab = A2 + r0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            x0[k] = load_raw4<T, VEC>(Qb, q.lda, 16 * k + rq, col, lim, 32);
            x1[k] = load_raw4<T, VEC>(Qb, q.lda, 16 * k + rq, col + 16, lim, 32);
            y0[k] = load_raw4<T, VEC>(Ab, q.lda, 16 * k + rq, col, lim, q.n2);
            y1[k] = load_raw4<T, VEC>(Ab, q.lda, 16 * k + rq, col + 16, lim, q.n2);
        }
    };
    Raw4<T> cq0[D], cq1[D], ca0[D], ca1[D];
    int ch = blk;
    if (ch < chunks) load(ch, cq0, cq1, ca0, ca1);
#pragma unroll 1
    for (; ch < chunks; ch += S) {
        const int nx = ch + S;
        Raw4<T> pq0[D], pq1[D], pa0[D], pa1[D];
        if (nx < chunks) load(nx, pq0, pq1, pa0, pa1);
#pragma unroll
        for (int k = 0; k < D; ++k) mma4(cq0[k], cq1[k], ca0[k], ca1[k]);
        if (nx >= chunks) break;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            cq0[k] = pq0[k];
            cq1[k] = pq1[k];
            ca0[k] = pa0[k];
            ca1[k] = pa1[k];
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
    // column sums over the four row groups (lane bits 4 and 5)
    s0sq = fold_all<0>(fold_all<1>(s0sq));
    s1sq = fold_all<0>(fold_all<1>(s1sq));
    if (lane < 16) {
        norms[col] = s0sq;
        norms[16 + col] = s1sq;
    }
}

// The mixed method's coupling product (FP32 panels): as cross_block, with
// BF16x3 products (bf16x3.hpp) summed in FP32 over each 64-row chunk and the
// chunk sums added in FP64 (the accuracy of the mixed method's Gram
// matrices).  Lane l loads rows 8(l/16) .. +7 of columns l%16 and 16 + l%16
// of Q1 and A2 in each 32-row step, the same rows feeding the same BF16
// k-slots of both operands; the next chunk is in flight meanwhile.  Output
// and norms as cross_block's.
template <bool VEC>
__device__ __forceinline__ void cross_block_bf16(const BatchCouple& q, int blk, int S, double* out, double* norms,
                                                 int lane) {
    const int col = lane & 15;
    const int rq = 8 * (lane >> 4);
    const int chunks = (q.m + 63) / 64;
    const float* Q1 = reinterpret_cast<const float*>(q.Q1);
    const float* A2 = reinterpret_cast<const float*>(q.A2);
    double p64[4][4];  // [block (Q1 half, A2 half)][register]
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) p64[b][v] = 0.0;
    double s0sq = 0.0, s1sq = 0.0;
    // x[h][0..3]: Q1 columns col, col + 16, A2 columns col, col + 16 of step h
    auto load = [&](int ch, float(&x)[2][4][8]) {
        const int64_t r0 = (int64_t)ch * 64;
        const int lim = (int)(q.m - r0 < 64 ? q.m - r0 : 64);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            load8<VEC>(x[h][0], Q1 + r0, q.lda, 32 * h + rq, col, lim, 32);
            load8<VEC>(x[h][1], Q1 + r0, q.lda, 32 * h + rq, col + 16, lim, 32);
            load8<VEC>(x[h][2], A2 + r0, q.lda, 32 * h + rq, col, lim, q.n2);
            load8<VEC>(x[h][3], A2 + r0, q.lda, 32 * h + rq, col + 16, lim, q.n2);
        }
    };
    float cur[2][4][8];
    int ch = blk;
    if (ch < chunks) load(ch, cur);
#pragma unroll 1
    for (; ch < chunks; ch += S) {
        const int nx = ch + S;
        float nxt[2][4][8];
        if (nx < chunks) load(nx, nxt);
        f4v c[4] = {f4v{0, 0, 0, 0}, f4v{0, 0, 0, 0}, f4v{0, 0, 0, 0}, f4v{0, 0, 0, 0}};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                s0sq = fmat((double)cur[h][2][e], (double)cur[h][2][e], s0sq);
                s1sq = fmat((double)cur[h][3][e], (double)cur[h][3][e], s1sq);
            }
            const Split3 q0 = split3(cur[h][0]), q1 = split3(cur[h][1]);
            const Split3 a0 = split3(cur[h][2]), a1 = split3(cur[h][3]);
            c[0] = mma_bf16x3(q0, a0, c[0]);
            c[1] = mma_bf16x3(q0, a1, c[1]);
            c[2] = mma_bf16x3(q1, a0, c[2]);
            c[3] = mma_bf16x3(q1, a1, c[3]);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) p64[b][v] += (double)c[b][v];
        if (nx >= chunks) break;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int e = 0; e < 8; ++e) cur[h][a][e] = nxt[h][a][e];
    }
    // f32 accumulator: lane l, register v holds C[4(l/16) + v][l%16]
    const int i0 = 4 * (lane >> 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = i0 + v;
        out[i * 32 + col] = p64[0][v];
        out[i * 32 + 16 + col] = p64[1][v];
        out[(16 + i) * 32 + col] = p64[2][v];
        out[(16 + i) * 32 + 16 + col] = p64[3][v];
    }
    s0sq = fold_all<0>(fold_all<1>(s0sq));
    s1sq = fold_all<0>(fold_all<1>(s1sq));
    if (lane < 16) {
        norms[col] = s0sq;
        norms[16 + col] = s1sq;
    }
}

// Pass 1 also decides whether the second pass is needed ("twice is enough",
// Kahan-Parlett): with ||a_j||^2 the squared norm of column j of A2 and
// P_j the product's column j, ||a_j - Q1 P_j||^2 = ||a_j||^2 - ||P_j||^2
// (Q1 orthonormal); if every column keeps at least half its squared norm,
// one projection leaves A2 orthogonal to Q1 to the working precision and the
// second pass is skipped.  (The plain method never runs a second pass.)
// Item (couple, g), g < batch_groups(q.leaves), by four waves: wave w takes
// row block 4g + w of 4 batch_groups(q.leaves); the workgroup adds the four
// products and norm vectors in wave order into block g.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void batch_cross(
    const BatchCouple* __restrict__ couples, const int2* __restrict__ items, char* __restrict__ work, int pass) {
    __shared__ double red[4][1024 + 32];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int2 it = items[blockIdx.x];
    const BatchCouple& q = couples[it.x];
    double* blocks = reinterpret_cast<double*>(work + q.ws);
    double* norms = reinterpret_cast<double*>(work + q.ws + couple_norms_off(q.leaves));
    const int* tk = reinterpret_cast<const int*>(work + q.ws + couple_ints_off(q.leaves));
    if (pass == 2 && !tk[kCoupleReorth]) return;
    const int S = 4 * batch_groups(q.leaves);
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (q.mixed) {
            if (q.vec)
                cross_block_bf16<true>(q, 4 * it.y + wv, S, red[wv], red[wv] + 1024, lane);
            else
                cross_block_bf16<false>(q, 4 * it.y + wv, S, red[wv], red[wv] + 1024, lane);
            done = true;
        }
    }
    if (!done) {
        if (q.vec)
            cross_block<T, true>(q, 4 * it.y + wv, S, red[wv], red[wv] + 1024, lane);
        else
            cross_block<T, false>(q, 4 * it.y + wv, S, red[wv], red[wv] + 1024, lane);
    }
    __syncthreads();
    double* out = blocks + (int64_t)it.y * 1024;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int idx = t + 256 * e;
        out[idx] = ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx];
    }
    if (t < 32)
        norms[(int64_t)it.y * 32 + t] = ((red[0][1024 + t] + red[1][1024 + t]) + red[2][1024 + t]) + red[3][1024 + t];
}

// One workgroup per couple, after batch_cross: P = the sum of the row
// blocks' products (four waves, fixed order), R12 (+)= P by one wave.
template <typename T>
__global__ __launch_bounds__(256) void batch_cross_sum(const BatchCouple* __restrict__ couples,
                                                        const int2* __restrict__ list, char* __restrict__ work,
                                                        int pass) {
    __shared__ GramLds L;
    const int lane = threadIdx.x;
    const BatchCouple& q = couples[list[blockIdx.x].x];
    double* blocks = reinterpret_cast<double*>(work + q.ws);
    const double* norms = reinterpret_cast<const double*>(work + q.ws + couple_norms_off(q.leaves));
    int* tk = reinterpret_cast<int*>(work + q.ws + couple_ints_off(q.leaves));
    if (pass == 2 && !tk[kCoupleReorth]) return;
    const int nb = batch_groups(q.leaves);
    sum_blocks_wg4(L, blocks, nb);
    if (threadIdx.x >= 64) return;
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
    if (pass == 1) {
        // lane j < 32: ||a_j||^2 (row blocks in order) against ||P_j||^2
        const int j = lane & 31;
        double an = 0.0, pn = 0.0;
#pragma unroll 1
        for (int b = 0; b < nb; ++b) an += norms[(int64_t)b * 32 + j];
#pragma unroll 4
        for (int i = 0; i < 32; ++i) pn = fmat(L.G[i][j], L.G[i][j], pn);
        const bool lost = j < q.n2 && !(2.0 * pn <= an);  // NaN counts as lost
        const uint64_t any = __ballot(lost);
        if (lane == 0) tk[kCoupleReorth] = any ? 1 : 0;
    }
}

// FP32 form of project_block: the same products (v_mfma_f32_16x16x4f32,
// output register v of (t, jt) is row 16(l/16) + 4v + t, column 16 jt + l%16),
// with A2 moved through the wave's LDS tile `tw` (column c, row r at
// c * kTrPitch + r): it is loaded and stored in 16-byte column runs (lane l:
// rows 4(l%16) .. +3 of column 4p + l/16, 256 contiguous bytes per column and
// instruction) and exchanged with the accumulator layout in LDS, where each
// lane reads and writes the same positions.
template <bool VEC, bool GRAM>
__device__ __forceinline__ void project_block_f32(const BatchCouple& q, int blk, int S, const double* __restrict__ P,
                                                  int lane, double* sums, float* tw) {
    const int li = lane & 15, lq = lane >> 4;
    const int chunks = (q.m + 63) / 64;
    const float* Q1 = reinterpret_cast<const float*>(q.Q1);
    float* A2 = reinterpret_cast<float*>(q.A2);
    auto lim_of = [&](int ch) { return (int)(q.m - (int64_t)ch * 64 < 64 ? q.m - (int64_t)ch * 64 : 64); };
    float mb[8][2];  // -P in the B-operand layout of k-group kg (k = 4 kg + l/16)
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) mb[kg][jt] = (float)(-P[(4 * kg + lq) * 32 + 16 * jt + li]);
    float qa[8][4], ca[8][4];  // Q1 (A operand layout) and A2 (column runs) of a chunk
    auto load_block = [&](int ch) {
        const float* Qb = Q1 + (int64_t)ch * 64;
        const float* Ab = A2 + (int64_t)ch * 64;
        const int lim = lim_of(ch);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            load_run<float, 4, VEC>(qa[p], Qb, q.lda, 4 * li, 4 * p + lq, lim, 32);
            load_run<float, 4, VEC>(ca[p], Ab, q.lda, 4 * li, 4 * p + lq, lim, q.n2);
        }
    };
    GramAcc32 g;
    if constexpr (GRAM) {
        g.sums = sums;
        g.zero(lane);
    }
    if (blk < chunks) load_block(blk);
#pragma unroll 1
    for (int ch = blk; ch < chunks; ch += S) {
        wave_sync();  // the previous chunk's reads of the tile are done
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int t = 0; t < 4; ++t) tw[(4 * p + lq) * kTrPitch + 4 * li + t] = ca[p][t];
        wave_sync();
        typedef float v4 __attribute__((ext_vector_type(4)));
        v4 acc[4][2];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t][jt][v] = tw[(16 * jt + li) * kTrPitch + 16 * lq + 4 * v + t];
        float x[8][4];
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t) x[kg][t] = qa[kg][t];
        if (ch + S < chunks) load_block(ch + S);
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int jt = 0; jt < 2; ++jt)
                    acc[t][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[kg][t], mb[kg][jt], acc[t][jt], 0, 0, 0);
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int t = 0; t < 4; ++t) tw[(16 * jt + li) * kTrPitch + 16 * lq + 4 * v + t] = acc[t][jt][v];
        wave_sync();
        const int lim = lim_of(ch);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int c = 4 * p + lq;
            if (c < q.n2) {
                float o[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) o[t] = tw[c * kTrPitch + 4 * li + t];
                store_column<float, 4, VEC>(o, A2 + (int64_t)ch * 64 + (int64_t)c * q.lda + 4 * li, lim, 4 * li);
            }
        }
        if constexpr (GRAM) {
            // rows past the end and columns past n2 are zero in acc
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float x0[8], x1[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    x0[e] = acc[e & 3][0][2 * h + (e >> 2)];
                    x1[e] = acc[e & 3][1][2 * h + (e >> 2)];
                }
                g.template add<2>(x0, x1);
            }
            g.template flush<2>(lane);
        }
    }
}

// A2 rows of row block `blk` of S (the chunks blk, blk + S, ... of cross_block)
// -= Q1 rows * P on the matrix cores in the working precision
// (v_mfma_f32_16x16x4f32 for float, v_mfma_f64_16x16x4f64 for double),
// accumulators initialised from A2 and B = -P, one rounding.
// Rows in chunks of 64; lane l feeds row 4(l%16) + t of column 4 kg + l/16
// to MFMA t, so output register v of (t, jt) is row Mfma16<T>::row(v, l/16)
// * 4 + t (16(l/16) + 4v + t for float, 16v + 4(l/16) + t for double),
// column 16 jt + l%16, and each lane stores 4 consecutive rows per (v, jt).
// The next chunk's loads are issued before the current chunk's MFMAs.
//
// With GRAM (float, q.gram set) the new rows also go through the BF16x3
// Gram products of the mixed method (GramAcc32, FP32 per 64 rows, then FP64
// sums in `sums`) straight from the accumulators: lane (li, lq) holds rows
// 16 lq .. 16 lq + 15 of columns li and 16 + li of each 64-row chunk and
// feeds rows 16 lq + 8 h + s to k-slot s of the h-th 32-row product, the
// same assignment in every lane.  The Gram slab is left in `sums`.
template <typename T, bool VEC, bool GRAM = false>
__device__ __forceinline__ void project_block(const BatchCouple& q, int blk, int S, const double* __restrict__ P,
                                              int lane, double* sums = nullptr, float* tw = nullptr) {
    static_assert(!GRAM || sizeof(T) == 4, "the fused Gram is the mixed (FP32) method's");
    if constexpr (sizeof(T) == 4) {
        project_block_f32<VEC, GRAM>(q, blk, S, P, lane, sums, tw);
        return;
    }
    using M = Mfma16<T>;
    const int li = lane & 15, lq = lane >> 4;
    const int chunks = (q.m + 63) / 64;
    const T* Q1 = reinterpret_cast<const T*>(q.Q1);
    T* A2 = reinterpret_cast<T*>(q.A2);
    auto lim_of = [&](int ch) { return (int)(q.m - (int64_t)ch * 64 < 64 ? q.m - (int64_t)ch * 64 : 64); };
    T mb[8][2];  // -P in the B-operand layout of k-group kg (k = 4 kg + l/16)
#pragma unroll
    for (int kg = 0; kg < 8; ++kg)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) mb[kg][jt] = (T)(-P[(4 * kg + lq) * 32 + 16 * jt + li]);
    auto row_of = [&](int v) { return 4 * M::row(v, lq); };  // first of the 4 rows of register v
    Four<T> qa[8], cv[2][4];
    auto load_block = [&](int ch) {
        const T* Qb = Q1 + (int64_t)ch * 64;
        const T* Ab = A2 + (int64_t)ch * 64;
        const int lim = lim_of(ch);
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) qa[kg] = load_four<T, VEC>(Qb, q.lda, 4 * li, 4 * kg + lq, lim, 32);
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                cv[jt][v] = load_four<T, VEC>(Ab, q.lda, row_of(v), 16 * jt + li, lim, q.n2);
    };
    GramAcc32 g;
    if constexpr (GRAM) {
        g.sums = sums;
        g.zero(lane);
    }
    if (blk < chunks) load_block(blk);
#pragma unroll 1
    for (int ch = blk; ch < chunks; ch += S) {
        typename M::v4 acc[4][2];
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t][jt][v] = cv[jt][v].v[t];
        Four<T> x[8];
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) x[kg] = qa[kg];
        if (ch + S < chunks) load_block(ch + S);
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int jt = 0; jt < 2; ++jt) acc[t][jt] = M::mma(x[kg].v[t], mb[kg][jt], acc[t][jt]);
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Four<T> o;
#pragma unroll
                for (int t = 0; t < 4; ++t) o.v[t] = acc[t][jt][v];
                store_four<T, VEC>(o, A2 + (int64_t)ch * 64, q.lda, row_of(v), 16 * jt + li, lim_of(ch), q.n2);
            }
        }
        if constexpr (GRAM) {
            // rows past the end and columns past n2 are zero in acc
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float x0[8], x1[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    x0[e] = acc[e & 3][0][2 * h + (e >> 2)];
                    x1[e] = acc[e & 3][1][2 * h + (e >> 2)];
                }
                g.template add<2>(x0, x1);
            }
            g.template flush<2>(lane);
        }
    }
}

// Item (couple, g) by four waves, row blocks as in batch_cross.  With q.gram
// the four Gram slabs go, added in wave order, to the second panel's side
// block g (the flat path only sums them).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void batch_project(
    const BatchCouple* __restrict__ couples, const int2* __restrict__ items, char* __restrict__ work, int pass) {
    const int2 it = items[blockIdx.x];
    const BatchCouple& q = couples[it.x];
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    if (pass == 2 && !reinterpret_cast<const int*>(work + q.ws + couple_ints_off(q.leaves))[kCoupleReorth]) return;
    const double* P = reinterpret_cast<const double*>(work + q.ws) + (int64_t)q.leaves * 1024;
    const int blk = 4 * it.y + wv, S = 4 * batch_groups(q.leaves);
    __shared__ float tr[sizeof(T) == 4 ? 4 : 1][32 * kTrPitch];
    float* tw = tr[sizeof(T) == 4 ? wv : 0];
    if constexpr (sizeof(T) == 4) {
        if (q.gram) {
            __shared__ double sums[4][12 * 64];
            if (q.vec)
                project_block<T, true, true>(q, blk, S, P, lane, sums[wv], tw);
            else
                project_block<T, false, true>(q, blk, S, P, lane, sums[wv], tw);
            __syncthreads();
            gram32_combine(sums, reinterpret_cast<double*>(work + q.side2) + (int64_t)it.y * 1024);
            return;
        }
    }
    if (q.vec)
        project_block<T, true>(q, blk, S, P, lane, nullptr, tw);
    else
        project_block<T, false>(q, blk, S, P, lane, nullptr, tw);
}

}  // namespace tsqr
