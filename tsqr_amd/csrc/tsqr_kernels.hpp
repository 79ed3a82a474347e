// tsqr_kernels.hpp -- the TSQR up sweep, down sweep and leaf-Q kernels.
//
// Every workgroup is one wavefront that owns one tile of the tree (plan.hpp).
//
// tsqr_tree_up levels >= 1 in one launch: each wave factors a level-1
//              block, and the last child to finish a parent (atomic ticket)
//              factors the parent, so the rest of the reduction tree is
//              walked inside this one launch with no grid barrier and no
//              waiting (the reference spins on a global counter between
//              levels, kernelQR.h:367-375).  Blocks keep their reflectors in
//              place.  The root writes R to the caller and its own explicit
//              Q, which holds the n x n S factors of its children.
// tsqr_single  m <= one tile: factor, R, and Q in one wave.
// tsqr_down   levels root-1 .. 1: Q_tile (from the stored reflectors) times
//             the S factor handed down by the parent gives the S factors of
//             the tile's children, written over the tile's block.
//
// Level 0 (superleaves of A) is handled by gram.hpp and super.hpp.  Traffic
// per call: A is read twice and written once, plus the level >= 1 blocks
// (about m*n/(8*(arity-1)) elements) written and read a few times.  The
// reference instead round-trips every level's Q through a work array and,
// for m above one chunk, makes a third full pass over A (tsgemm_tsqr,
// TallShinnyQR.h:11-74, :126-127).
#pragma once
#include "plan.hpp"
#include "tile.hpp"

namespace tsqr {

template <typename T, int RPL>
struct Chunk16 {
    static constexpr int kElems = 16 / (int)sizeof(T);
    static constexpr int kChunks = RPL / kElems;
    static_assert(RPL % kElems == 0, "a lane's rows must be a whole number of 16-byte chunks");
    typedef T vec __attribute__((ext_vector_type(16 / sizeof(T))));
};

// x[s][c] = base[(row0 + s) + c*ld] for rows < `rows` and columns < ncols,
// zero elsewhere.  VEC: base and ld*sizeof(T) are 16-byte aligned.
template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void load_rows(T (&x)[RPL][N], const T* __restrict__ base, int64_t ld,
                                          int rows, int ncols, int lane) {
    using C = Chunk16<T, RPL>;
    const int row0 = lane * RPL;
    const bool full = row0 + RPL <= rows;
#pragma unroll
    for (int c = 0; c < N; ++c) {
        if (c < ncols) {
            const T* p = base + (int64_t)c * ld + row0;
            if (VEC && full) {
#pragma unroll
                for (int q = 0; q < C::kChunks; ++q) {
                    const typename C::vec v = *reinterpret_cast<const typename C::vec*>(p + q * C::kElems);
#pragma unroll
                    for (int e = 0; e < C::kElems; ++e) x[q * C::kElems + e][c] = v[e];
                }
            } else {
#pragma unroll
                for (int s = 0; s < RPL; ++s) x[s][c] = (row0 + s < rows) ? p[s] : T(0);
            }
        } else {
#pragma unroll
            for (int s = 0; s < RPL; ++s) x[s][c] = T(0);
        }
    }
}

template <typename T, int RPL, bool VEC, bool NT = false>
__device__ __forceinline__ void store_column(const T (&v)[RPL], T* __restrict__ p, int rows, int row0) {
    using C = Chunk16<T, RPL>;
    if (VEC && row0 + RPL <= rows) {
#pragma unroll
        for (int q = 0; q < C::kChunks; ++q) {
            typename C::vec w;
#pragma unroll
            for (int e = 0; e < C::kElems; ++e) w[e] = v[q * C::kElems + e];
            if constexpr (NT)
                __builtin_nontemporal_store(w, reinterpret_cast<typename C::vec*>(p + q * C::kElems));
            else
                *reinterpret_cast<typename C::vec*>(p + q * C::kElems) = w;
        }
    } else {
#pragma unroll
        for (int s = 0; s < RPL; ++s)
            if (row0 + s < rows) p[s] = v[s];
    }
}

template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void store_rows(const T (&x)[RPL][N], T* __restrict__ base, int64_t ld, int rows,
                                           int ncols, int lane) {
    const int row0 = lane * RPL;
#pragma unroll
    for (int c = 0; c < N; ++c) {
        if (c < ncols) {
            T v[RPL];
#pragma unroll
            for (int s = 0; s < RPL; ++s) v[s] = x[s][c];
            store_column<T, RPL, VEC>(v, base + (int64_t)c * ld + row0, rows, row0);
        }
    }
}

// dst[:, c] = Q_tile[:, 0:kmax] * S[0:kmax, c] for c < n.  S (global memory
// or LDS, leading dimension ldS) is first copied into the wave's N x N LDS
// staging area `Ls` with one batch of independent loads (zero beyond kmax
// rows and n columns), so the column loop reads broadcast LDS instead of
// waiting on one memory latency per column.
template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void apply_s_store(const T (&x)[RPL][N], const T* __restrict__ S, int ldS, int kmax,
                                              int n, T* __restrict__ dst, int64_t ldd, int rows, T* Ls,
                                              int lane) {
    constexpr int PER = (N * N + 63) / 64;
    T sv[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = lane + 64 * q;
        const int k = idx % N, c = idx / N;
        sv[q] = (idx < N * N && k < kmax && c < n) ? S[k + (int64_t)c * ldS] : T(0);
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = lane + 64 * q;
        if (idx < N * N) Ls[idx] = sv[q];
    }
    wave_sync();
    const int row0 = lane * RPL;
#pragma unroll 1
    for (int c = 0; c < n; ++c) {
        T acc[RPL];
#pragma unroll
        for (int s = 0; s < RPL; ++s) acc[s] = T(0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const T v = Ls[k + c * N];
#pragma unroll
            for (int s = 0; s < RPL; ++s) acc[s] = fmat(x[s][k], v, acc[s]);
        }
        store_column<T, RPL, VEC>(acc, dst + (int64_t)c * ldd + row0, rows, row0);
    }
    wave_sync();
}

// R (n x n, zero below the diagonal) from the wave's LDS to dst (ld),
// rows [0, nrows).
template <typename T, int N>
__device__ __forceinline__ void store_r(const TileLds<T, N>& L, T* __restrict__ dst, int64_t ld, int nrows, int n,
                                        int lane) {
    if (lane < nrows) {
#pragma unroll 1
        for (int c = 0; c < n; ++c) dst[lane + (int64_t)c * ld] = (lane <= c) ? L.R[lane + c * N] : T(0);
    }
}

__device__ __forceinline__ int tile_rows_of(const LevelDesc& L, int tile) {
    const int64_t left = L.rows - (int64_t)tile * L.tile_rows;
    return (int)(left < L.tile_rows ? left : L.tile_rows);
}

// Levels >= 1, one launch: a wave factors one level-1 block, keeps its
// reflectors in place and hands its R to the parent block; the last child
// of a parent to finish (ticket) goes on to factor the parent, and so on up
// to the root.  Nothing waits: a wave that is not last simply exits.  The
// root writes R to the caller and its explicit Q (the S factors of its
// children) over its block.  With FORMQ every node writes its explicit Q
// over its block instead of its reflectors, so that no down sweep is needed:
// a superleaf's S is then the product of the blocks of its ancestors' Q that
// lie on its path (path_s).  tree_up_body does this for level-1 block
// `tile`; the kernel below maps one block per workgroup.
template <typename T, int N, int RPL, bool FORMQ = false>
__device__ __forceinline__ void tree_up_body(const TreeDesc& d, T* __restrict__ R, int ldr, T* __restrict__ blocks,
                                             int* __restrict__ tickets, TileLds<T, N>& L, int tile) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    int level = 1;
    int rows = tile_rows_of(d.lv[1], tile);
    T* blk = blocks + d.lv[1].blk_off + (int64_t)tile * d.H * n;
    T x[RPL][N];
    load_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);
#pragma unroll 1
    for (;;) {
        factor_tile<T, N, RPL>(x, L, lane);
        const LevelDesc& Lv = d.lv[level];
        if (Lv.tiles == 1) break;
        const LevelDesc& P = d.lv[level + 1];
        const int parent = tile / d.arity;
        const int slot = tile - parent * d.arity;
        T* pblk = blocks + P.blk_off + (int64_t)parent * d.H * n;
        store_r<T, N>(L, pblk + slot * n, d.H, rows < n ? rows : n, n, lane);
        if constexpr (FORMQ) form_q_tile<T, N, RPL>(x, L, lane);
        store_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);

        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        int old = 0;
        if (lane == 0)
            old = __hip_atomic_fetch_add(&tickets[P.ctr_off + parent], 1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        old = __builtin_amdgcn_readfirstlane(old);
        const int left = Lv.tiles - parent * d.arity;
        const int nchild = left < d.arity ? left : d.arity;
        if (old != nchild - 1) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

        ++level;
        tile = parent;
        blk = pblk;
        rows = tile_rows_of(P, parent);
        load_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);
    }
    store_r<T, N>(L, R, ldr, n, n, lane);
    form_q_tile<T, N, RPL>(x, L, lane);
    store_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);
}

// A grid-stride loop over the level-1 blocks, so that a small grid serves
// any tree (and costs one short launch when skip[0] is set).
template <typename T, int N, int RPL, bool FORMQ = false>
__global__ __launch_bounds__(64) void tsqr_tree_up(const TreeDesc d, T* __restrict__ R, int ldr,
                                                    T* __restrict__ blocks, int* __restrict__ tickets,
                                                    const int* __restrict__ skip) {
    __shared__ TileLds<T, N> L;
    if (skip[0]) return;
#pragma unroll 1
    for (int tile = blockIdx.x; tile < d.lv[1].tiles; tile += gridDim.x)
        tree_up_body<T, N, RPL, FORMQ>(d, R, ldr, blocks, tickets, L, tile);
}

// One tree level per launch, no tickets (the pure Householder path of tall
// matrices): block `tile` of `level` is factored; its R goes to the parent
// block, its reflectors (or, with FORMQ, its explicit Q) stay over its block.
// The root writes R to the caller and its explicit Q over its block, as
// tree_up_body does.  Kernel boundaries order the levels, so no
// agent-scope release per block is needed (tree_up_body needs one per
// block: an L2 write-back on MI355X).
template <typename T, int N, int RPL, bool FORMQ = false>
__device__ __forceinline__ void tree_level_body(const TreeDesc& d, int level, T* __restrict__ R, int ldr,
                                                T* __restrict__ blocks, TileLds<T, N>& L, int tile) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const LevelDesc& Lv = d.lv[level];
    const int rows = tile_rows_of(Lv, tile);
    T* blk = blocks + Lv.blk_off + (int64_t)tile * d.H * n;
    T x[RPL][N];
    load_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);
    factor_tile<T, N, RPL>(x, L, lane);
    if (Lv.tiles == 1) {
        store_r<T, N>(L, R, ldr, n, n, lane);
        form_q_tile<T, N, RPL>(x, L, lane);
    } else {
        const LevelDesc& P = d.lv[level + 1];
        const int parent = tile / d.arity;
        const int slot = tile - parent * d.arity;
        store_r<T, N>(L, blocks + P.blk_off + (int64_t)parent * d.H * n + slot * n, d.H, rows < n ? rows : n, n,
                      lane);
        if constexpr (FORMQ) form_q_tile<T, N, RPL>(x, L, lane);
    }
    store_rows<T, N, RPL, true>(x, blk, d.H, rows, n, lane);
}

template <typename T, int N, int RPL, bool FORMQ = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void tsqr_tree_level(
    const TreeDesc d, int level, T* __restrict__ R, int ldr, T* __restrict__ blocks, const int* __restrict__ skip) {
    __shared__ TileLds<T, N> L;
    if (skip[0]) return;
    tree_level_body<T, N, RPL, FORMQ>(d, level, R, ldr, blocks, L, blockIdx.x);
}

// S of superleaf `leaf` after a FORMQ up sweep: B_1 B_2 ... B_top, B_l the
// n x n block (ld H) of the level-l ancestor's explicit Q at the rows of the
// level-(l-1) ancestor, accumulated from the root down (X <- B_l X).  One
// wave; X, Y and B are N x N LDS buffers (column-major, ld N), entries
// outside n x n zero.  Returns the buffer (X or Y) that holds the result;
// as with the down sweep, its rows beyond min(rows, n) of a short superleaf
// are not meaningful and the caller uses only the first min(rows, n).
template <typename T, int N>
__device__ __forceinline__ T* path_s(const TreeDesc& d, const T* __restrict__ blocks, int leaf, T* X, T* Y, T* B,
                                     int lane) {
    constexpr int PER = (N * N + 63) / 64;
    const int n = d.n;
    const int top = d.levels - 1;
    int64_t pw = 1;  // arity^(l - 1) for the level l being applied
    for (int l = 1; l < top; ++l) pw *= d.arity;
#pragma unroll 1
    for (int l = top; l >= 1; --l) {
        const int64_t child = leaf / pw;  // ancestor at level l - 1
        const int64_t anc = child / d.arity;
        const int slot = (int)(child - anc * d.arity);
        const T* blk = blocks + d.lv[l].blk_off + anc * d.H * n + (int64_t)slot * n;
        T* dst = l == top ? X : B;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int idx = lane + 64 * q;
            const int k = idx % N, c = idx / N;
            if (idx < N * N) dst[idx] = (k < n && c < n) ? blk[k + (int64_t)c * d.H] : T(0);
        }
        wave_sync();
        if (l < top) {
            // only the first min(rows, n) columns of the ancestor's Q (and
            // rows of X) exist; the block rows beyond a short child's R are
            // never written, so they are left out of the sum
            const int arows = tile_rows_of(d.lv[l], (int)anc);
            const int kmax = arows < n ? arows : n;
#pragma unroll 1
            for (int q = 0; q < PER; ++q) {
                const int idx = lane + 64 * q;
                const int i = idx % N, c = idx / N;
                T acc = T(0);
                if (idx < N * N) {
#pragma unroll 8
                    for (int k = 0; k < N; ++k)
                        if (k < kmax) acc = fmat(B[i + k * N], X[k + c * N], acc);
                    Y[idx] = acc;
                }
            }
            wave_sync();
            T* t = X;
            X = Y;
            Y = t;
        }
        pw /= d.arity;
    }
    return X;
}

// A single tile (m <= H): factor, R to the caller, Q over A (unless form_q
// is 0).
template <typename T, int N, int RPL, bool VEC>
__device__ __forceinline__ void tsqr_single_body(const TreeDesc& d, T* __restrict__ A, int lda, T* __restrict__ R,
                                                 int ldr, TileLds<T, N>& L, bool form_q = true) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const int rows = (int)d.lv[0].rows;
    T x[RPL][N];
    load_rows<T, N, RPL, VEC>(x, A, lda, rows, n, lane);
    factor_tile<T, N, RPL>(x, L, lane);
    store_r<T, N>(L, R, ldr, n, n, lane);
    if (!form_q) return;  // R only: A stays as it was
    form_q_tile<T, N, RPL>(x, L, lane);
    store_rows<T, N, RPL, VEC>(x, A, lda, rows, n, lane);
}

template <typename T, int N, int RPL, bool VEC>
__global__ __launch_bounds__(64) void tsqr_single(const TreeDesc d, T* __restrict__ A, int lda, T* __restrict__ R,
                                                   int ldr, int form_q) {
    __shared__ TileLds<T, N> L;
    tsqr_single_body<T, N, RPL, VEC>(d, A, lda, R, ldr, L, form_q != 0);
}

// Down sweep of one tile of `level` (1 <= level <= root - 1): its explicit
// Q times the S factor its parent holds for it, written over its block (the
// S factors of its own children).
template <typename T, int N, int RPL>
__device__ __forceinline__ void down_body(const TreeDesc& d, int level, T* __restrict__ blocks, TileLds<T, N>& L,
                                          int tile) {
    const int lane = threadIdx.x & 63;
    const int n = d.n;
    const LevelDesc& Lv = d.lv[level];
    const LevelDesc& P = d.lv[level + 1];
    const int rows = tile_rows_of(Lv, tile);
    T* myblk = blocks + Lv.blk_off + (int64_t)tile * d.H * n;
    T x[RPL][N];
    load_rows<T, N, RPL, true>(x, myblk, d.H, rows, n, lane);
    form_q_tile<T, N, RPL>(x, L, lane);
    const int parent = tile / d.arity;
    const int slot = tile - parent * d.arity;
    const T* S = blocks + P.blk_off + (int64_t)parent * d.H * n + slot * n;
    apply_s_store<T, N, RPL, true>(x, S, d.H, rows < n ? rows : n, n, myblk, d.H, rows, L.R, lane);
}

template <typename T, int N, int RPL>
__global__ __launch_bounds__(64) void tsqr_down(const TreeDesc d, int level, T* __restrict__ blocks,
                                                 const int* __restrict__ skip) {
    __shared__ TileLds<T, N> L;
    if (skip[0]) return;
    down_body<T, N, RPL>(d, level, blocks, L, blockIdx.x);
}

// One copy per translation unit (each dtype's TU launches its own).
namespace {
__global__ void zero_ints(int* __restrict__ p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}
}  // namespace

}  // namespace tsqr
