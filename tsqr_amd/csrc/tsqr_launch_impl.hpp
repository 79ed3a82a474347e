// tsqr_launch_impl.hpp -- body of tsqr_launch<T>: the dispatcher is
// instantiated in tsqr_f32.hip / tsqr_f64.hip, the launch sequence per
// (dtype, padded width) in tsqr_w{8,16,32}{s,d}.hip.
//
// Launch sequence for one call (m > H):
//   init          tickets = 0, superleaf flags = path, control words = 0
//   gram_stream   (auto path) FP64 Gram matrix of the whole matrix by a
//                 persistent grid (FP32: gram_stream4, on 4 x 4 tiles), its Cholesky and the gate in the same
//                 launch; if the whole matrix passes the gate this is the
//                 flat Gram path (Q = A R^{-1}) and the per-superleaf kernels
//                 and the tree below return at once
//   gram_g        (auto path when the flat gate rejects, gram-tree path) FP64
//                 Gram matrix of every superleaf
//   gram_chol     Cholesky R, R^{-1} and the gate per superleaf; flags the
//                 superleaves it rejects
//   super_r       Householder R of the flagged superleaves
//   tree_up       levels >= 1, one launch, ticketed walk to the root
//   down          levels root-1 .. 1, one launch each
//   gram_q        M = R^{-1} S, then Q = A M, for the Gram superleaves
//                 (FP32 flat path: flat_q4, Q = A R^{-1} by a persistent grid)
//   super_q       Householder Q for the flagged superleaves
// The CholeskyQR2 path replaces the Gram kernels by mixed_stream (FP32,
// BF16x3) or gram_stream (FP64) with the flat reduction and Cholesky, then
// the streamed Q pass mixed_qs + mixed_check (pass 1) and mixed_qs +
// mixed_check (pass 2, a no-op unless the check asks for it); every leaf is
// flagged Householder so that the tree path takes over if the flat gate
// rejects.
// A single tile (m <= H) is one tsqr_single launch.
// R only (LaunchOpts::r_only, the tsqr_r_* entry points): the sequence stops
// after the tree walk up, so A is only read.  The FP64 CholeskyQR2 path then
// forms the second pass's Gram matrix of A R1^{-1} by streaming A
// (tsqr_resid_stream without a residual) instead of writing Q1.
#pragma once
#include "../../include/tsqr_amd.h"
#include "check.hpp"
#include "leaf1.hpp"
#include "mixed.hpp"
#include "stream4.hpp"
#include "super.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

// Workgroups of the grid-stride Gram-tree kernels (gram_g, gram_chol):
// enough to fill the GPU.
constexpr int kFallbackGrid = 4096;
// Most superleaves for which a flat-path call walks up the tree inside the
// superleaf R kernel (tsqr_super_r_up).
constexpr int kFusedWalkLeaves = 256;
// Grid of the single-tile leaf kernels (leaf1.hpp) where they only back up
// a flat path: small, so that the launch costs little when they return at
// once; the pure Householder path launches one workgroup per tile.
constexpr int kFallbackTileGrid = 2048;

namespace {
// The workspace set-up of a call (WsInit), one thread per word.
__global__ void init_flags(const WsInit w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.count()) w.at(i);
}

}  // namespace

template <typename T, int N, bool VEC>
int launch_tsqr_n(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o,
                  hipStream_t st) {
    constexpr int RPL = rows_per_lane<T>(N);
    const bool timed = timing_enabled() && !o.untimed;
    if (d.lv[0].rows <= d.H) {
        if (timed) timing_mark(kPhaseLeafR, true, st);
        hipLaunchKernelGGL((tsqr_single<T, N, RPL, VEC>), dim3(1), dim3(64), 0, st, d, A, lda, R, ldr,
                           o.r_only ? 0 : 1);
        if (timed) timing_mark(kPhaseLeafR, false, st);
        return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
    }
    const WsLayout w = ws_layout(d, sizeof(T));
    char* base = reinterpret_cast<char*>(work);
    T* blocks = reinterpret_cast<T*>(base + w.blocks);
    int* tickets = reinterpret_cast<int*>(base + w.tickets);
    int* flags = reinterpret_cast<int*>(base + w.flags);
    int* list = reinterpret_cast<int*>(base + w.list);
    int* ctl = reinterpret_cast<int*>(base + w.ctl);
    double* partial = reinterpret_cast<double*>(base + w.partial);
    double* side = reinterpret_cast<double*>(base + w.side);
    const int leaves = d.lv[0].tiles;
    // R only, FP32: the flat FP64 Gram (exact products) needs no second
    // pass, so the CholeskyQR2 setting runs as AUTO
    const int path = (o.r_only && sizeof(T) == 4 && o.path == kLeafCholQR2) ? kLeafAuto : o.path;
    const bool mixed = path == kLeafCholQR2;
    // AUTO over two or more superleaves: the flat Gram path, with every
    // superleaf listed Householder in case its gate fails.  GRAM_TREE (and
    // AUTO over one superleaf): a Gram Cholesky per superleaf, Householder
    // where that is rejected.
    const bool flat_auto = path == kLeafAuto && leaves >= 2;
    const bool gram_tree = path == kLeafGramTree || (path == kLeafAuto && leaves < 2);
    const bool gram = flat_auto || gram_tree;
    const int force = gram_tree ? 0 : 1;
    constexpr int NT = N > 16 ? 2 : 1;

    // Short matrices (short_waves): the R pass without tickets on ssmall
    // workgroups of swaves waves, summed by the flat factor's 8 waves.
    // Where it is the call's first kernel and reads no control word (the
    // AUTO flat path, the FP64 CholeskyQR2 path), it also sets up the
    // workspace, one launch fewer.
    constexpr int kChunkRows = 16 * stream_slabs<T>();
    const int swaves = short_waves(d.m, kChunkRows);
    const bool small = swaves != 0;
    const int ssmall = short_groups(d.m, kChunkRows);
    WsInit ini{tickets, d.total_ctr, flags, list, leaves, force, ctl, false};
    const bool fold_init = small && (flat_auto || (mixed && sizeof(T) == 8));
    if (!fold_init) hipLaunchKernelGGL(init_flags, dim3((ini.count() + 255) / 256), dim3(256), 0, st, ini);
    ini.on = fold_init;
    // Householder superleaf kernels: a grid-stride loop over the list
    const int super_grid = leaves < 1024 ? leaves : 1024;
    double* rkeep = side + (size_t)(d.side_leaves + 1) * 1024;
    if (timed) timing_mark(kPhaseLeafR, true, st);
    // streaming Gram: sgroups workgroups, then sgroup_parts group partials
    int* stream_tk = ctl + kCtlStreamTickets;
    const int sgroups = stream_groups(d.m);
    const int sgroup_parts = stream_group_count(sgroups);
    double* stream_sum = partial + (size_t)sgroups * 1024;
    const double* X0 = side + (size_t)d.side_leaves * 1024;  // R^{-1} of the flat path
    // the R pass on 16 x 16 tiles (FP64 products) and the flat factor
    auto gram_and_factor = [&](int* tk, double gate, double* keep, const int* skip) {
        if (small) {
            if (swaves == 4)
                hipLaunchKernelGGL((tsqr_gram_short<T, NT, VEC, 4>), dim3(ssmall), dim3(256), 0, st, d, (const T*)A,
                                   lda, partial, skip, ini);
            else if (swaves == 8)
                hipLaunchKernelGGL((tsqr_gram_short<T, NT, VEC, 8>), dim3(ssmall), dim3(512), 0, st, d, (const T*)A,
                                   lda, partial, skip, ini);
            else
                hipLaunchKernelGGL((tsqr_gram_short<T, NT, VEC, 16>), dim3(ssmall), dim3(1024), 0, st, d,
                                   (const T*)A, lda, partial, skip, ini);
            ini.on = false;
            hipLaunchKernelGGL((tsqr_flat_short<T>), dim3(1), dim3(kFlatShortThreads), 0, st, d, ssmall,
                               (const double*)partial,
                               R, ldr, side, ctl, gate, keep, skip);
        } else {
            hipLaunchKernelGGL((tsqr_gram_stream<T, NT, VEC>), dim3(sgroups), dim3(256), 0, st, d, (const T*)A, lda,
                               partial, tk, skip);
            hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(kSumThreads), 0, st, d, sgroup_parts,
                               (const double*)stream_sum, R, ldr, side, ctl, gate, keep, 0, skip);
        }
    };
    if (mixed) {
        if constexpr (sizeof(T) == 4) {
            hipLaunchKernelGGL((tsqr_mixed_stream<NT, VEC>), dim3(sgroups), dim3(256), 0, st, d, (const float*)A, lda,
                               partial, stream_tk);
            hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(kSumThreads), 0, st, d, sgroup_parts,
                               (const double*)stream_sum, R, ldr, side, ctl, o.gate, rkeep, 1);
            // Escalation where the BF16x3 gate rejected the matrix: the FP64
            // Gram of the AUTO path (exact products) with the CholeskyQR2
            // gate mixed_gram64_gate; both kernels return at once when the
            // first pass succeeded (ctl[0]).  Pass 1 then forms Q1 with FP64
            // products and pass 2 restores the orthogonality.
            int* tk64 = ctl + kCtlStreamTickets2;
            if (use_stream4(d.m)) {
                const int g64 = stream_groups4(d.m);
                if (d.n == N)
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, true>), dim3(g64), dim3(256), 0, st, d.m, d.n,
                                       (const float*)A, lda, partial, tk64, kCachedChunks, (const int*)ctl);
                else
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, false>), dim3(g64), dim3(256), 0, st, d.m, d.n,
                                       (const float*)A, lda, partial, tk64, kCachedChunks, (const int*)ctl);
                hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(kSumThreads), 0, st, d,
                                   stream_group_count(g64), (const double*)(partial + (size_t)g64 * 1024), R, ldr,
                                   side, ctl, mixed_gram64_gate(o.gate), rkeep, 0, (const int*)ctl);
            } else {
                gram_and_factor(tk64, mixed_gram64_gate(o.gate), rkeep, (const int*)ctl);
            }
        } else {
            // FP64: the R pass on 4 x 4 tiles from kStream4MinRows rows (two
            // rows per lane and load), as the FP32 flat path's
            if (use_stream4(d.m)) {
                const int rg = stream_groups4(d.m);
                if (d.n == N)
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, true, T>), dim3(rg), dim3(256), 0, st, d.m, d.n,
                                       (const T*)A, lda, partial, stream_tk, 0, (const int*)nullptr);
                else
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, false, T>), dim3(rg), dim3(256), 0, st, d.m,
                                       d.n, (const T*)A, lda, partial, stream_tk, 0, (const int*)nullptr);
                hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(kSumThreads), 0, st, d, stream_group_count(rg),
                                   (const double*)(partial + (size_t)rg * 1024), R, ldr, side, ctl, o.gate, rkeep);
            } else {
                gram_and_factor(stream_tk, o.gate, rkeep, (const int*)nullptr);
            }
            if (o.r_only) {
                // second CholeskyQR pass without writing Q1 = A R1^{-1}: the
                // Gram matrix of the product streamed from A, then the check
                // (R = R1, or R2 R1 with R2 = chol(Q1^T Q1))
                hipLaunchKernelGGL((tsqr_resid_stream<T, double, NT, VEC, false>), dim3(sgroups), dim3(256), 0, st,
                                   d.m, d.n, (const T*)nullptr, 0, (const T*)A, lda, (const double*)X0, 32, 1,
                                   partial, ctl + kCtlStreamTickets2, (const int*)ctl);
                hipLaunchKernelGGL(tsqr_mixed_check<T>, dim3(1), dim3(kSumThreads), 0, st, d, sgroup_parts,
                                   (const double*)stream_sum, R, ldr, side, ctl, 1, o.tol);
            }
        }
    }
    if (flat_auto) {
        // FP32 from kStream4MinRows rows: the R pass on 4 x 4 tiles, with its
        // own grid
        if (sizeof(T) == 4 && use_stream4(d.m)) {
            if constexpr (sizeof(T) == 4) {
                const int rgroups = stream_groups4(d.m);
                if (d.n == N)
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, true>), dim3(rgroups), dim3(256), 0, st, d.m,
                                       d.n, (const float*)A, lda, partial, stream_tk, kCachedChunks);
                else
                    hipLaunchKernelGGL((tsqr_gram_stream4<N / 4, VEC, false>), dim3(rgroups), dim3(256), 0, st, d.m,
                                       d.n, (const float*)A, lda, partial, stream_tk, kCachedChunks);
                hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(kSumThreads), 0, st, d,
                                   stream_group_count(rgroups), (const double*)(partial + (size_t)rgroups * 1024), R,
                                   ldr, side, ctl, o.gate, (double*)nullptr);
            }
        } else {
            gram_and_factor(stream_tk, o.gate, (double*)nullptr, (const int*)nullptr);
        }
    }
    if (gram_tree) {
        const int gram_grid = leaves < kFallbackGrid ? leaves : kFallbackGrid;
        const int chol_grid = (leaves + 1) / 2 < kFallbackGrid ? (leaves + 1) / 2 : kFallbackGrid;
        hipLaunchKernelGGL((tsqr_gram_g<T, NT, VEC>), dim3(gram_grid), dim3(64), 0, st, d, (const T*)A, lda, side,
                           (const int*)ctl);
        hipLaunchKernelGGL((tsqr_gram_chol<T>), dim3(chol_grid), dim3(64), 0, st, d, R, ldr, blocks, side, flags, list,
                           (const int*)ctl, o.gate);
    }
    // Where the Householder kernels only back up a flat path, the tree keeps
    // every node's explicit Q and each superleaf forms its own S (path_s), so
    // no down-sweep launches (one per level) are queued for a path that
    // nearly never runs, and the superleaf R kernel walks up the tree itself
    // (tsqr_super_r_up).
    const bool flat_first = flat_auto || mixed;
    // single-tile leaves: one wave per tile (leaf1.hpp)
    const bool tile_leaves = d.lv[0].tile_rows == d.H;
    const int tile_grid = flat_first ? (leaves < kFallbackTileGrid ? leaves : kFallbackTileGrid) : leaves;
    // The fused walk saves a launch, which matters on short calls; on tall
    // matrices it measured slower (one wave walks while the rest of its
    // workgroup idles), and they keep the separate tree launch.
    // the walk up needs every leaf to take its parent's ticket, so not on the
    // Gram-tree path (its Gram leaves write their R in gram_chol)
    const bool fused_up = (flat_first || (tile_leaves && !gram_tree)) && d.levels >= 2 && leaves <= kFusedWalkLeaves;
    if (tile_leaves) {
        // the reflectors go over A unless A must stay as it is (R only)
        auto leaf_r = [&](auto walk) {
            constexpr int WK = decltype(walk)::value;
            if (o.r_only)
                hipLaunchKernelGGL((tsqr_leaf_r1<T, N, RPL, VEC, WK, false>), dim3(tile_grid), dim3(64), 0, st, d, A,
                                   lda, R, ldr, blocks, tickets, (const int*)list, (const int*)ctl);
            else
                hipLaunchKernelGGL((tsqr_leaf_r1<T, N, RPL, VEC, WK, true>), dim3(tile_grid), dim3(64), 0, st, d, A,
                                   lda, R, ldr, blocks, tickets, (const int*)list, (const int*)ctl);
        };
        if (fused_up && flat_first)
            leaf_r(std::integral_constant<int, 2>{});
        else if (fused_up)
            leaf_r(std::integral_constant<int, 1>{});
        else
            leaf_r(std::integral_constant<int, 0>{});
    } else if (fused_up)
        hipLaunchKernelGGL((tsqr_super_r_up<T, N, RPL, VEC>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0, st,
                           d, (const T*)A, lda, R, ldr, blocks, tickets, (const int*)list, (const int*)ctl);
    else
        hipLaunchKernelGGL((tsqr_super_r<T, N, RPL, VEC>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0, st, d,
                           (const T*)A, lda, R, ldr, blocks, (const int*)list, (const int*)ctl);
    if (timed) timing_mark(kPhaseLeafR, false, st);
    if (d.levels >= 2 && !fused_up) {
        if (timed) timing_mark(kPhaseTree, true, st);
        // where the tree only backs up a flat path, a small grid-stride grid
        const int l1 = d.lv[1].tiles;
        const int tree_grid = flat_first && l1 > kFallbackTileGrid ? kFallbackTileGrid : l1;
        if (flat_first)
            hipLaunchKernelGGL((tsqr_tree_up<T, N, RPL, true>), dim3(tree_grid), dim3(64), 0, st, d, R, ldr,
                               blocks, tickets, (const int*)ctl);
        else if (tile_leaves)
            // pure Householder over single-tile leaves: one launch per level,
            // no per-block tickets (the down sweep is queued anyway)
            for (int l = 1; l < d.levels; ++l)
                hipLaunchKernelGGL((tsqr_tree_level<T, N, RPL, false>), dim3(d.lv[l].tiles), dim3(64), 0, st, d,
                                   l, R, ldr, blocks, (const int*)ctl);
        else
            hipLaunchKernelGGL((tsqr_tree_up<T, N, RPL, false>), dim3(tree_grid), dim3(64), 0, st, d, R, ldr,
                               blocks, tickets, (const int*)ctl);
        if (timed) timing_mark(kPhaseTree, false, st);
    }
    if (o.r_only) return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
    if (timed) timing_mark(kPhaseDown, true, st);
    if (!flat_first)
        for (int l = d.levels - 2; l >= 1; --l)
            hipLaunchKernelGGL((tsqr_down<T, N, RPL>), dim3(d.lv[l].tiles), dim3(64), 0, st, d, l, blocks,
                               (const int*)ctl);
    if (timed) timing_mark(kPhaseDown, false, st);
    if (timed) timing_mark(kPhaseLeafQ, true, st);
    if (mixed) {
        // streamed Q passes (the grid of the 4 x 4 R pass: its partial sums
        // fit the workspace)
        const int qgroups = stream_groups4(d.m);
        for (int pass = 1; pass <= 2; ++pass) {
            int* qtk = ctl + (pass == 1 ? kCtlStreamTickets3 : kCtlStreamTickets4);
            bool q4 = false;
            if constexpr (sizeof(T) == 8) {
                if (use_stream4(d.m)) {
                    // FP64 data: products and the Gram of Q on 4 x 4 tiles
                    hipLaunchKernelGGL((tsqr_q4_gram<N / 4, VEC, T>), dim3(qgroups), dim3(256), 0, st, d.m, d.n, A,
                                       lda, X0, partial, qtk, (const int*)ctl, pass);
                    q4 = true;
                }
            }
            if (!q4)
                hipLaunchKernelGGL((tsqr_mixed_qs<T, NT, VEC>), dim3(qgroups), dim3(256), 0, st, d.m, d.n, A, lda,
                                   X0, partial, qtk, (const int*)ctl, pass);
            hipLaunchKernelGGL(tsqr_mixed_check<T>, dim3(1), dim3(kSumThreads), 0, st, d, stream_group_count(qgroups),
                               (const double*)(partial + (size_t)qgroups * 1024), R, ldr, side, ctl, pass, o.tol);
        }
    }
    if (flat_auto && sizeof(T) == 4) {
        // FP32 flat path: Q = A R^{-1} by a persistent grid (nothing if the
        // gate rejected the matrix)
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((tsqr_flat_q4<N / 4, VEC, true>), dim3(flat_q_groups(d.m)), dim3(256), 0, st, d.m,
                               d.n, (float*)A, lda, X0, (const int*)ctl);
    } else if (gram) {
        // few superleaves: split each into parts so that about
        // tuned_gramq_waves() (default 2048) waves share the rows
        const int tile_blocks = (d.lv[0].tile_rows + 63) / 64;
        const int qwaves = tuned_gramq_waves();
        int parts = (qwaves + leaves - 1) / leaves;
        parts = parts < 1 ? 1 : (parts > tile_blocks ? tile_blocks : parts);
        hipLaunchKernelGGL((tsqr_gram_q<T, VEC>), dim3(leaves * parts), dim3(64), 0, st, d, A, lda,
                           (const double*)side, (const T*)blocks, (const int*)flags, (const int*)ctl, parts);
    }
    if (tile_leaves && flat_first)
        hipLaunchKernelGGL((tsqr_leaf_q1<T, N, RPL, VEC, true>), dim3(tile_grid), dim3(64), 0, st, d, A, lda,
                           (const T*)blocks, (const int*)list, (const int*)ctl);
    else if (tile_leaves)
        hipLaunchKernelGGL((tsqr_leaf_q1<T, N, RPL, VEC, false>), dim3(tile_grid), dim3(64), 0, st, d, A, lda,
                           (const T*)blocks, (const int*)list, (const int*)ctl);
    else if (flat_first)
        hipLaunchKernelGGL((tsqr_super_q<T, N, RPL, VEC, true>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0,
                           st, d, A, lda, (const T*)blocks, (const int*)list, (const int*)ctl);
    else
        hipLaunchKernelGGL((tsqr_super_q<T, N, RPL, VEC, false>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0,
                           st, d, A, lda, (const T*)blocks, (const int*)list, (const int*)ctl);
    if (timed) timing_mark(kPhaseLeafQ, false, st);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

template <typename T, int N>
int launch_tsqr_v(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o,
                  hipStream_t st) {
    if (aligned16(A, lda, sizeof(T))) return launch_tsqr_n<T, N, true>(d, A, lda, R, ldr, work, o, st);
    return launch_tsqr_n<T, N, false>(d, A, lda, R, ldr, work, o, st);
}

// Each (dtype, padded width) pair is instantiated in a translation unit of
// its own (tsqr_w*.hip), so that the six compile in parallel.
#define TSQR_WIDTH_UNIT(T, N)                                                                                   \
    template int launch_tsqr_v<T, N>(const TreeDesc&, T*, int, T*, int, void*, const LaunchOpts&, hipStream_t)
extern TSQR_WIDTH_UNIT(float, 8);
extern TSQR_WIDTH_UNIT(float, 16);
extern TSQR_WIDTH_UNIT(float, 32);
extern TSQR_WIDTH_UNIT(double, 8);
extern TSQR_WIDTH_UNIT(double, 16);
extern TSQR_WIDTH_UNIT(double, 32);

template <typename T>
int tsqr_launch(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o, hipStream_t st) {
    switch (padded_cols(d.n)) {
        case 8:
            return launch_tsqr_v<T, 8>(d, A, lda, R, ldr, work, o, st);
        case 16:
            return launch_tsqr_v<T, 16>(d, A, lda, R, ldr, work, o, st);
        default:
            return launch_tsqr_v<T, 32>(d, A, lda, R, ldr, work, o, st);
    }
}

}  // namespace tsqr
