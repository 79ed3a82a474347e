// tsqr_launch_impl.hpp -- body of tsqr_launch<T>, included once per dtype.
#pragma once
#include "../../include/tsqr_amd.h"
#include "tsqr_kernels.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

inline bool aligned16(const void* p, int ld, size_t elem) {
    return ((uintptr_t)p % 16 == 0) && (((size_t)ld * elem) % 16 == 0);
}

template <typename T, int N, bool VEC>
int launch_tsqr_n(const TreeDesc& d, T* A, int lda, T* R, int ldr, T* blocks, int* tickets, hipStream_t st) {
    constexpr int RPL = rows_per_lane<T>(N);
    const bool timed = timing_enabled();
    if (d.levels == 1) {
        if (timed) timing_mark(kPhaseUp, true, st);
        hipLaunchKernelGGL((tsqr_single<T, N, RPL, VEC>), dim3(1), dim3(64), 0, st, d, A, lda, R, ldr);
        if (timed) timing_mark(kPhaseUp, false, st);
        return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
    }
    if (d.total_ctr > 0)
        hipLaunchKernelGGL(zero_ints, dim3((d.total_ctr + 255) / 256), dim3(256), 0, st, tickets, d.total_ctr);
    if (timed) timing_mark(kPhaseUp, true, st);
    hipLaunchKernelGGL((tsqr_leaf_r<T, N, RPL, VEC>), dim3(d.lv[0].tiles), dim3(64), 0, st, d, (const T*)A, lda,
                       blocks);
    if (timed) timing_mark(kPhaseUp, false, st);
    if (timed) timing_mark(kPhaseTree, true, st);
    hipLaunchKernelGGL((tsqr_tree_up<T, N, RPL>), dim3(d.lv[1].tiles), dim3(64), 0, st, d, R, ldr, blocks, tickets);
    if (timed) timing_mark(kPhaseTree, false, st);
    if (d.levels >= 3) {
        if (timed) timing_mark(kPhaseDown, true, st);
        for (int l = d.levels - 2; l >= 1; --l)
            hipLaunchKernelGGL((tsqr_down<T, N, RPL>), dim3(d.lv[l].tiles), dim3(64), 0, st, d, l, blocks);
        if (timed) timing_mark(kPhaseDown, false, st);
    }
    if (timed) timing_mark(kPhaseLeafQ, true, st);
    hipLaunchKernelGGL((tsqr_leafq<T, N, RPL, VEC>), dim3(d.lv[0].tiles), dim3(64), 0, st, d, A, lda,
                       (const T*)blocks);
    if (timed) timing_mark(kPhaseLeafQ, false, st);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

template <typename T, int N>
int launch_tsqr_v(const TreeDesc& d, T* A, int lda, T* R, int ldr, T* blocks, int* tickets, hipStream_t st) {
    if (aligned16(A, lda, sizeof(T))) return launch_tsqr_n<T, N, true>(d, A, lda, R, ldr, blocks, tickets, st);
    return launch_tsqr_n<T, N, false>(d, A, lda, R, ldr, blocks, tickets, st);
}

template <typename T>
int tsqr_launch(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, hipStream_t st) {
    T* blocks = reinterpret_cast<T*>(work);
    int* tickets =
        reinterpret_cast<int*>(reinterpret_cast<char*>(work) + align256((size_t)d.block_elems * sizeof(T)));
    switch (padded_cols(d.n)) {
        case 8:
            return launch_tsqr_v<T, 8>(d, A, lda, R, ldr, blocks, tickets, st);
        case 16:
            return launch_tsqr_v<T, 16>(d, A, lda, R, ldr, blocks, tickets, st);
        default:
            return launch_tsqr_v<T, 32>(d, A, lda, R, ldr, blocks, tickets, st);
    }
}

}  // namespace tsqr
