// super.hpp -- Householder path of a superleaf (ST = super_tiles tiles of H
// rows).
//
// Used for every superleaf when the caller asks for Householder leaves, and
// for the superleaves whose Gram factorisation was rejected (gram.hpp).
// One workgroup of ST wavefronts per superleaf:
//   tsqr_super_r  each wave factors its H-row tile (tile.hpp), the tile Rs
//                 are stacked in LDS, wave 0 factors the stack; that R goes
//                 to the superleaf's slot in its parent block (or to the
//                 caller at the root).  This is the first two levels of the
//                 reference's tree (kernelQR.h:341-388) inside one CU.
//   tsqr_super_q  the same factorisations again (identical code and input,
//                 so identical reflectors), then wave 0 forms the stack's Q
//                 and multiplies it by the superleaf's S factor from the tree,
//                 leaving one n x n factor per tile in LDS, and every wave
//                 writes Q_tile * factor over its rows of A.
// Q is always formed by applying the stored reflectors in reverse
// (form_q_tile), as the reference does (kernelQR.h:227-290).
#pragma once
#include "gram.hpp"
#include "tsqr_kernels.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

template <typename T, int N, int RPL>
struct SuperLds {
    static constexpr int H = 64 * RPL;
    static constexpr int ST = super_tiles<T>(N);
    static_assert(ST * N <= H, "the stacked tile Rs must fit one tile");
    TileLds<T, N> wave[ST];
    T stack[N * H];  // column-major, leading dimension H
};

// Rows of tile t of a superleaf with `rows` valid rows.
__device__ __forceinline__ int super_tile_rows(int rows, int t, int H) {
    const int left = rows - t * H;
    return left <= 0 ? 0 : (left < H ? left : H);
}

// Stacked height: every non-empty tile contributes min(tile rows, n).
template <int ST>
__device__ __forceinline__ int super_stack_rows(int rows, int n, int H) {
    int s = 0;
#pragma unroll
    for (int t = 0; t < ST; ++t) {
        const int r = super_tile_rows(rows, t, H);
        s += r < n ? r : n;
    }
    return s;
}

// Both kernels run a grid-stride loop over the Householder-path superleaves
// listed in list[1 .. list[0]], so they cost one short launch when the Gram
// path took every superleaf.
template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void super_r_leaf(const TreeDesc& d, const T* __restrict__ A, int lda, T* __restrict__ R,
                                             int ldr, T* __restrict__ blocks, SuperLds<T, N, RPL>& S, int leaf) {
    constexpr int H = 64 * RPL;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int rows = tile_rows_of(L0, leaf);
    const int trows = super_tile_rows(rows, w, H);
    T x[RPL][N];
    if (trows > 0) {
        load_rows<T, N, RPL, VEC>(x, A + (int64_t)leaf * L0.tile_rows + (int64_t)w * H, lda, trows, n, lane);
        factor_tile<T, N, RPL>(x, S.wave[w], lane);
        store_r<T, N>(S.wave[w], S.stack + w * n, H, trows < n ? trows : n, n, lane);
    }
    __syncthreads();
    if (w == 0) {
        const int srows = super_stack_rows<super_tiles<T>(N)>(rows, n, H);
        load_rows<T, N, RPL, true>(x, S.stack, H, srows, n, lane);
        factor_tile<T, N, RPL>(x, S.wave[0], lane);
        if (L0.tiles == 1) {
            store_r<T, N>(S.wave[0], R, ldr, n, n, lane);
        } else {
            const int parent = leaf / d.arity;
            const int slot = leaf - parent * d.arity;
            T* dst = blocks + d.lv[1].blk_off + (int64_t)parent * d.H * n + slot * n;
            store_r<T, N>(S.wave[0], dst, d.H, srows < n ? srows : n, n, lane);
        }
    }
}

template <typename T, int N, int RPL, bool VEC>
__global__ __launch_bounds__(64 * super_tiles<T>(N)) void tsqr_super_r(const TreeDesc d, const T* __restrict__ A,
                                                                         int lda, T* __restrict__ R, int ldr,
                                                                         T* __restrict__ blocks,
                                                                         const int* __restrict__ list,
                                                                         const int* __restrict__ ctl) {
    __shared__ SuperLds<T, N, RPL> S;
    if (ctl[0]) return;  // the flat Gram path took the whole matrix
    const int count = list[0];
#pragma unroll 1
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        super_r_leaf<T, N, RPL, VEC>(d, A, lda, R, ldr, blocks, S, list[1 + it]);
        __syncthreads();
    }
}

// tsqr_super_r and tsqr_tree_up<FORMQ> in one launch, for the paths where
// every superleaf is listed Householder (the flat paths' fallback): after a
// superleaf's R is in its level-1 parent's block, wave 0 takes the parent's
// ticket, and the last child to arrive walks up the tree from that parent
// (tree_up_body, in the LDS of wave 0).  Nothing waits.  Needs two or more
// tree levels.
template <typename T, int N, int RPL, bool VEC>
__global__ __launch_bounds__(64 * super_tiles<T>(N)) void tsqr_super_r_up(const TreeDesc d, const T* __restrict__ A,
                                                                            int lda, T* __restrict__ R, int ldr,
                                                                            T* __restrict__ blocks,
                                                                            int* __restrict__ tickets,
                                                                            const int* __restrict__ list,
                                                                            const int* __restrict__ ctl) {
    __shared__ SuperLds<T, N, RPL> S;
    if (ctl[0]) return;  // the flat Gram path took the whole matrix
    const int count = list[0];
    const LevelDesc& L0 = d.lv[0];
    const LevelDesc& L1 = d.lv[1];
#pragma unroll 1
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        const int leaf = list[1 + it];
        super_r_leaf<T, N, RPL, VEC>(d, A, lda, R, ldr, blocks, S, leaf);
        if (threadIdx.x < 64) {
            const int parent = leaf / d.arity;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            int old = 0;
            if (threadIdx.x == 0)
                old = __hip_atomic_fetch_add(&tickets[L1.ctr_off + parent], 1, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            old = __builtin_amdgcn_readfirstlane(old);
            const int left = L0.tiles - parent * d.arity;
            const int nchild = left < d.arity ? left : d.arity;
            if (old == nchild - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                tree_up_body<T, N, RPL, true>(d, R, ldr, blocks, tickets, S.wave[0], parent);
            }
        }
        __syncthreads();
    }
}

// PATHS: the tree above holds each node's explicit Q (tsqr_tree_up with
// FORMQ, no down sweep), and wave 0 forms this superleaf's S with path_s in
// the LDS of waves 1 to 3, which are idle between the two barriers.
template <typename T, int N, int RPL, bool VEC, bool PATHS = false>
__device__ __forceinline__ void super_q_leaf(const TreeDesc& d, T* __restrict__ A, int lda,
                                             const T* __restrict__ blocks, SuperLds<T, N, RPL>& S, T* stash,
                                             int leaf) {
    static_assert(SuperLds<T, N, RPL>::ST >= 4, "path_s borrows the LDS of waves 1 to 3");
    constexpr int H = 64 * RPL;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int rows = tile_rows_of(L0, leaf);
    const int trows = super_tile_rows(rows, w, H);
    T* a = A + (int64_t)leaf * L0.tile_rows + (int64_t)w * H;
    T x[RPL][N];
    if (trows > 0) {
        load_rows<T, N, RPL, VEC>(x, a, lda, trows, n, lane);
        factor_tile<T, N, RPL>(x, S.wave[w], lane);
        store_r<T, N>(S.wave[w], S.stack + w * n, H, trows < n ? trows : n, n, lane);
        form_q_tile<T, N, RPL>(x, S.wave[w], lane);
    }
    __syncthreads();
    if (w == 0) {
        store_rows<T, N, RPL, true>(x, stash, H, H, N, lane);
        const int srows = super_stack_rows<super_tiles<T>(N)>(rows, n, H);
        T y[RPL][N];
        load_rows<T, N, RPL, true>(y, S.stack, H, srows, n, lane);
        factor_tile<T, N, RPL>(y, S.wave[0], lane);
        form_q_tile<T, N, RPL>(y, S.wave[0], lane);
        if (L0.tiles == 1) {
            store_rows<T, N, RPL, true>(y, S.stack, H, srows, n, lane);
        } else if constexpr (PATHS) {
            const T* Sp = path_s<T, N>(d, blocks, leaf, S.wave[1].R, S.wave[2].R, S.wave[3].R, lane);
            apply_s_store<T, N, RPL, true>(y, Sp, N, srows < n ? srows : n, n, S.stack, H, srows, S.wave[0].R,
                                           lane);
        } else {
            const int parent = leaf / d.arity;
            const int slot = leaf - parent * d.arity;
            const T* Sb = blocks + d.lv[1].blk_off + (int64_t)parent * d.H * n + slot * n;
            apply_s_store<T, N, RPL, true>(y, Sb, d.H, srows < n ? srows : n, n, S.stack, H, srows,
                                           S.wave[0].R, lane);
        }
        load_rows<T, N, RPL, true>(x, stash, H, H, N, lane);
    }
    __syncthreads();
    if (trows > 0)
        apply_s_store<T, N, RPL, VEC>(x, S.stack + w * n, H, trows < n ? trows : n, n, a, lda, trows,
                                      S.wave[w].R, lane);
}

template <typename T, int N, int RPL, bool VEC, bool PATHS = false>
__global__ __launch_bounds__(64 * super_tiles<T>(N)) void tsqr_super_q(const TreeDesc d, T* __restrict__ A, int lda,
                                                                         const T* __restrict__ blocks,
                                                                         const int* __restrict__ list,
                                                                         const int* __restrict__ ctl) {
    __shared__ SuperLds<T, N, RPL> S;
    __shared__ T stash[N * 64 * RPL];  // wave 0's Q_tile while it works on the stack
    if (ctl[0]) return;  // the flat Gram path took the whole matrix
    const int count = list[0];
#pragma unroll 1
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        super_q_leaf<T, N, RPL, VEC, PATHS>(d, A, lda, blocks, S, stash, list[1 + it]);
        __syncthreads();
    }
}

}  // namespace tsqr
