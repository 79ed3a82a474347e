// leaf1.hpp -- Householder leaves of one tile each, one wavefront per tile.
//
// The superleaf kernels of super.hpp give each superleaf a workgroup of up
// to eight waves: the waves factor their tiles, then wave 0 alone factors the
// stack of their R factors while the others wait, and two tile states per
// wave do not fit the 256 VGPRs that an eight-wave workgroup leaves each
// wave, so the compiler spills to scratch memory.  With leaves of one tile
// the stack is simply the next tree level (arity H / n tiles per node, the
// same blocks), so a leaf kernel needs one wave and one tile state:
//   tsqr_leaf_r1  factor the tile (tile.hpp) and write its R to the
//                 parent's block.  WALK (1: the tree keeps reflectors, 2:
//                 explicit Q, tree_up_body's FORMQ): then the parent's
//                 ticket, and the
//                 last child to arrive walks up the tree (tree_up_body), so
//                 no wave waits and there is no separate tree launch; that
//                 costs an agent-scope release per tile (an L2 write-back on
//                 MI355X: 1.7 -> 4.3 ms over the 93312 tiles of 23.9M x 32),
//                 so tall matrices leave the tree to tsqr_tree_up instead,
//                 whose tickets start one level up.  STORE_U: the tile's
//                 reflectors (lower trapezoid, zero above) are written over
//                 the tile of A, as LAPACK's geqrf keeps them below R (A is
//                 overwritten by Q in the end anyway); without it (R only) A
//                 is read, never written.
//   tsqr_leaf_q1  loads the stored reflectors, forms the explicit tile Q by
//                 applying them in reverse (form_q_tile, kernelQR.h:227-290)
//                 and writes Q * S over the tile, with S the tile's n x n
//                 block of its parent's Q: read from the parent block after
//                 a down sweep, or (PATHS) formed as the product of its
//                 ancestors' explicit-Q blocks (path_s).  Reading the
//                 reflectors instead of factoring the tile again costs one
//                 more write and read of A and saves the factorisation, which
//                 is latency-bound (about 60k cycles per 256 x 32 FP32 tile).
// Both take the tiles listed in list[1 .. list[0]] (grid-stride) and do
// nothing if ctl[0] is set (the flat Gram path took the whole matrix).
#pragma once
#include "super.hpp"

namespace tsqr {

// factor_tile column sink that writes reflector J (this lane's rows) over
// column J of the tile.
template <typename T, bool VEC>
struct StoreColumnSink {
    T* a;
    int64_t lda;
    int rows, n, row0;
    template <int J, typename U, int R>
    __device__ __forceinline__ void column(const U (&u)[R]) const {
        if (J < n) store_column<U, R, false>(u, a + (int64_t)J * lda + row0, rows, row0);
    }
};

//
// Both kernels ask for two waves per SIMD (256 VGPRs): the factorisation is
// a chain of dependent reductions, and a second wave on the SIMD hides most
// of its latency; left alone the compiler took a few more registers and
// one wave per SIMD.
// Factors level-0 tile `leaf` and writes its R to the parent's block;
// STORE_U: the reflectors over the tile of A.  Returns the parent's index.
template <typename T, int N, int RPL, bool VEC, bool STORE_U>
__device__ __forceinline__ int leaf_r_tile(const TreeDesc& d, T* __restrict__ A, int lda, T* __restrict__ blocks,
                                           TileLds<T, N>& L, int leaf) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int rows = tile_rows_of(L0, leaf);
    T x[RPL][N];
    T* a = A + (int64_t)leaf * L0.tile_rows;
    // with the reflectors stored, 16-byte loads made the compiler spill
    // (register tuples for the loads on top of a full tile)
    load_rows<T, N, RPL, VEC && !STORE_U>(x, a, lda, rows, n, lane);
    if constexpr (STORE_U) {
        // each reflector column goes over A as soon as it is final
        const StoreColumnSink<T, VEC> sink{a, lda, rows, n, lane * RPL};
        factor_tile<T, N, RPL, 0>(x, L, lane, sink);
    } else {
        factor_tile<T, N, RPL>(x, L, lane);
    }
    const int parent = leaf / d.arity;
    const int slot = leaf - parent * d.arity;
    store_r<T, N>(L, blocks + d.lv[1].blk_off + (int64_t)parent * d.H * n + slot * n, d.H, rows < n ? rows : n, n,
                  lane);
    return parent;
}

// Q of level-0 tile `leaf` from its stored reflectors, times S (ld ldS,
// n x n), written over the tile.
template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void leaf_q_tile(const TreeDesc& d, T* __restrict__ A, int lda, const T* S, int ldS,
                                            TileLds<T, N>& L, int leaf) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const int rows = tile_rows_of(d.lv[0], leaf);
    T* a = A + (int64_t)leaf * d.lv[0].tile_rows;
    T x[RPL][N];
    load_rows<T, N, RPL, VEC>(x, a, lda, rows, n, lane);  // the reflectors
    form_q_tile<T, N, RPL>(x, L, lane);
    apply_s_store<T, N, RPL, VEC>(x, S, ldS, rows < n ? rows : n, n, a, lda, rows, L.R, lane);
}

// S of level-0 tile `leaf` after a down sweep: its n x n block of its
// parent's block (ld H).
template <typename T>
__device__ __forceinline__ const T* leaf_s_block(const TreeDesc& d, const T* blocks, int leaf) {
    const int parent = leaf / d.arity;
    const int slot = leaf - parent * d.arity;
    return blocks + d.lv[1].blk_off + (int64_t)parent * d.H * d.n + slot * d.n;
}

// Both kernels ask for two waves per SIMD (256 VGPRs): the factorisation is
// a chain of dependent reductions, and a second wave on the SIMD hides most
// of its latency; left alone the compiler took a few more registers and
// one wave per SIMD.
template <typename T, int N, int RPL, bool VEC, int WALK, bool STORE_U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void tsqr_leaf_r1(
    const TreeDesc d, T* __restrict__ A, int lda, T* __restrict__ R, int ldr, T* __restrict__ blocks,
    int* __restrict__ tickets, const int* __restrict__ list, const int* __restrict__ ctl) {
    __shared__ TileLds<T, N> L;
    if (ctl[0]) return;
    const int lane = threadIdx.x;
    const LevelDesc& L0 = d.lv[0];
    const int count = list[0];
#pragma unroll 1
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        const int parent = leaf_r_tile<T, N, RPL, VEC, STORE_U>(d, A, lda, blocks, L, list[1 + it]);
        if constexpr (WALK != 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            int old = 0;
            if (lane == 0)
                old = __hip_atomic_fetch_add(&tickets[d.lv[1].ctr_off + parent], 1, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            old = __builtin_amdgcn_readfirstlane(old);
            const int left = L0.tiles - parent * d.arity;
            const int nchild = left < d.arity ? left : d.arity;
            if (old == nchild - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                tree_up_body<T, N, RPL, WALK == 2>(d, R, ldr, blocks, tickets, L, parent);
            }
        }
    }
}

template <typename T, int N, int RPL, bool VEC, bool PATHS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void tsqr_leaf_q1(
    const TreeDesc d, T* __restrict__ A, int lda, const T* __restrict__ blocks, const int* __restrict__ list,
    const int* __restrict__ ctl) {
    __shared__ TileLds<T, N> L;
    __shared__ T buf[PATHS ? 3 : 1][N * N];  // path_s products
    if (ctl[0]) return;
    const int lane = threadIdx.x;
    const int count = list[0];
#pragma unroll 1
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        const int leaf = list[1 + it];
        if constexpr (PATHS) {
            const T* Sp = path_s<T, N>(d, blocks, leaf, buf[0], buf[1], buf[2], lane);
            leaf_q_tile<T, N, RPL, VEC>(d, A, lda, Sp, N, L, leaf);
        } else {
            leaf_q_tile<T, N, RPL, VEC>(d, A, lda, leaf_s_block<T>(d, blocks, leaf), d.H, L, leaf);
        }
    }
}

}  // namespace tsqr
