// tsqr_launch_impl.hpp -- body of tsqr_launch<T>, included once per dtype.
//
// Launch sequence for one call (m > H):
//   init          tickets = 0, superleaf flags = path
//   gram_g        (auto path) FP64 Gram matrix of every superleaf
//   gram_partial, gram_flat
//                 (auto path) sum of all superleaf Grams and its Cholesky;
//                 if the whole matrix passes the gate this is the flat Gram
//                 path (Q = A R^{-1}) and the per-superleaf kernels and the
//                 tree below return at once
//   gram_chol     (auto path) Cholesky R, R^{-1} and the gate per
//                 superleaf; flags the superleaves it rejects
//   super_r       Householder R of the flagged superleaves
//   tree_up       levels >= 1, one launch, ticketed walk to the root
//   down          levels root-1 .. 1, one launch each
//   gram_q        M = R^{-1} S, then Q = A M, for the Gram superleaves
//   super_q       Householder Q for the flagged superleaves
// A single tile (m <= H) is one tsqr_single launch.
#pragma once
#include "../../include/tsqr_amd.h"
#include "super.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

inline bool aligned16(const void* p, int ld, size_t elem) {
    return ((uintptr_t)p % 16 == 0) && (((size_t)ld * elem) % 16 == 0);
}

namespace {
// tickets = 0; flags = value; the Householder list is empty (value 0) or
// holds every superleaf (value 1); the flat-path flag is cleared.
__global__ void init_flags(int* __restrict__ tickets, int nt, int* __restrict__ flags, int* __restrict__ list,
                           int nf, int value, int* __restrict__ ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nt) tickets[i] = 0;
    if (i < nf) {
        flags[i] = value;
        if (value) list[1 + i] = i;
    }
    if (i == 0) {
        list[0] = value ? nf : 0;
        ctl[0] = 0;
    }
}
}  // namespace

template <typename T, int N, bool VEC>
int launch_tsqr_n(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o,
                  hipStream_t st) {
    constexpr int RPL = rows_per_lane<T>(N);
    const bool timed = timing_enabled();
    if (d.lv[0].rows <= d.H) {
        if (timed) timing_mark(kPhaseLeafR, true, st);
        hipLaunchKernelGGL((tsqr_single<T, N, RPL, VEC>), dim3(1), dim3(64), 0, st, d, A, lda, R, ldr);
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
    const bool gram = o.path != kLeafHouseholder;
    const int force = gram ? 0 : 1;
    constexpr int NT = N > 16 ? 2 : 1;

    const int ninit = d.total_ctr > leaves ? d.total_ctr : leaves;
    hipLaunchKernelGGL(init_flags, dim3((ninit + 255) / 256), dim3(256), 0, st, tickets, d.total_ctr, flags, list,
                       leaves, force, ctl);
    // Householder superleaf kernels: a grid-stride loop over the list
    const int super_grid = leaves < 1024 ? leaves : 1024;
    if (timed) timing_mark(kPhaseLeafR, true, st);
    if (gram) {
        hipLaunchKernelGGL((tsqr_gram_g<T, NT, VEC>), dim3(leaves), dim3(64), 0, st, d, (const T*)A, lda, side);
        if (leaves >= 2 && o.path == kLeafAuto) {
            const int parts = leaves < kGramPartials ? leaves : kGramPartials;
            hipLaunchKernelGGL(tsqr_gram_partial<>, dim3(parts), dim3(256), 0, st, leaves, (const double*)side,
                               partial);
            hipLaunchKernelGGL((tsqr_gram_flat<T>), dim3(1), dim3(256), 0, st, d, parts, (const double*)partial, R,
                               ldr, side, ctl, o.gate);
        }
        hipLaunchKernelGGL((tsqr_gram_chol<T>), dim3((leaves + 1) / 2), dim3(64), 0, st, d, R, ldr, blocks, side,
                           flags, list, (const int*)ctl, o.gate);
    }
    hipLaunchKernelGGL((tsqr_super_r<T, N, RPL, VEC>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0, st, d,
                       (const T*)A, lda, R, ldr, blocks, (const int*)list);
    if (timed) timing_mark(kPhaseLeafR, false, st);
    if (d.levels >= 2) {
        if (timed) timing_mark(kPhaseTree, true, st);
        hipLaunchKernelGGL((tsqr_tree_up<T, N, RPL>), dim3(d.lv[1].tiles), dim3(64), 0, st, d, R, ldr, blocks,
                           tickets, (const int*)ctl);
        if (timed) timing_mark(kPhaseTree, false, st);
    }
    if (timed) timing_mark(kPhaseDown, true, st);
    for (int l = d.levels - 2; l >= 1; --l)
        hipLaunchKernelGGL((tsqr_down<T, N, RPL>), dim3(d.lv[l].tiles), dim3(64), 0, st, d, l, blocks,
                           (const int*)ctl);
    if (timed) timing_mark(kPhaseDown, false, st);
    if (timed) timing_mark(kPhaseLeafQ, true, st);
    if (gram)
        hipLaunchKernelGGL((tsqr_gram_q<T, VEC>), dim3(leaves), dim3(64), 0, st, d, A, lda, (const double*)side,
                           (const T*)blocks, (const int*)flags, (const int*)ctl);
    hipLaunchKernelGGL((tsqr_super_q<T, N, RPL, VEC>), dim3(super_grid), dim3(64 * super_tiles<T>(N)), 0, st, d, A,
                       lda, (const T*)blocks, (const int*)list);
    if (timed) timing_mark(kPhaseLeafQ, false, st);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

template <typename T, int N>
int launch_tsqr_v(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o,
                  hipStream_t st) {
    if (aligned16(A, lda, sizeof(T))) return launch_tsqr_n<T, N, true>(d, A, lda, R, ldr, work, o, st);
    return launch_tsqr_n<T, N, false>(d, A, lda, R, ldr, work, o, st);
}

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
