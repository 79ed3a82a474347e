// batch_launch_impl.hpp -- body of batch_launch<T>, included once per dtype.
#pragma once
#include "../../include/tsqr_amd.h"
#include "batch.hpp"
#include "batch_launch.hpp"

namespace tsqr {

namespace {
template <typename T>
void launch_panel_stage(const BatchDev& b, const BatchStage& s, char* work, hipStream_t st) {
    const bool mixed = sizeof(T) == 4 && b.method == 1;
    if (s.leaves.count == 0) return;
    int* fallback = reinterpret_cast<int*>(work + s.fallback);
    // Householder superleaves: every one (plain), or the device-built list of
    // the mixed problems whose gate failed (grid-stride, usually empty)
    const int2* hh_items = mixed ? reinterpret_cast<const int2*>(fallback + 2) : s.leaves.items;
    const int* hh_count = mixed ? fallback : nullptr;
    const int hh_grid = mixed ? (s.leaves.count < 1024 ? s.leaves.count : 1024) : s.leaves.count;
    if constexpr (sizeof(T) == 4) {
        if (mixed) {
            hipLaunchKernelGGL(batch_mixed_gram<>, dim3(s.groups.count), dim3(256), 0, st, b.probs, s.groups.items,
                               work);
            hipLaunchKernelGGL(batch_mixed_flat<>, dim3(s.probs.count), dim3(256), 0, st, b.probs, s.probs.items,
                               work, b.gate, fallback);
        }
    }
    if (mixed)
        hipLaunchKernelGGL((batch_super_r<T, true>), dim3(hh_grid), dim3(batch_super_threads<T>()), 0, st, b.probs,
                           hh_items, hh_count, work);
    else
        hipLaunchKernelGGL(batch_leaf_r<T>, dim3(s.leaves.count), dim3(64), 0, st, b.probs, s.leaves.items, work);
    if (mixed) {
        if (s.level1.count)
            hipLaunchKernelGGL(batch_tree_up<T>, dim3(s.level1.count), dim3(64), 0, st, b.probs, s.level1.items, work);
    } else {
        for (int l = 1; l < (int)s.up.size(); ++l)
            if (s.up[l].count)
                hipLaunchKernelGGL(batch_tree_level<T>, dim3(s.up[l].count), dim3(64), 0, st, b.probs, s.up[l].items,
                                   l, work);
    }
    for (int l = (int)s.down.size() - 1; l >= 1; --l)
        if (s.down[l].count)
            hipLaunchKernelGGL(batch_down<T>, dim3(s.down[l].count), dim3(64), 0, st, b.probs, s.down[l].items, l,
                               work);
    if constexpr (sizeof(T) == 4) {
        if (mixed)
            for (int pass = 1; pass <= 2; ++pass) {
                hipLaunchKernelGGL(batch_mixed_q<>, dim3(s.groups.count), dim3(256), 0, st, b.probs, s.groups.items,
                                   work, pass);
                hipLaunchKernelGGL(batch_mixed_check<>, dim3(s.probs.count), dim3(256), 0, st, b.probs,
                                   s.probs.items, work, pass, b.tol);
            }
    }
    if (mixed)
        hipLaunchKernelGGL((batch_super_q<T, true>), dim3(hh_grid), dim3(batch_super_threads<T>()), 0, st, b.probs,
                           hh_items, hh_count, work);
    else
        hipLaunchKernelGGL(batch_leaf_q<T>, dim3(s.leaves.count), dim3(64), 0, st, b.probs, s.leaves.items, work);
}
}  // namespace

template <typename T>
int batch_launch(const BatchDev& b, void* work_v, hipStream_t st) {
    char* work = reinterpret_cast<char*>(work_v);
    if (sizeof(T) != 4 && b.method == 1) return TSQR_ERR_ARG;
    const int ninit = b.nprobs + b.ncouples;
    if (ninit)
        hipLaunchKernelGGL(batch_init<T>, dim3(ninit), dim3(256), 0, st, b.probs, b.nprobs, b.couples, b.ncouples,
                           work, b.stage[0].fallback, b.stage[1].fallback);
    launch_panel_stage<T>(b, b.stage[0], work, st);
    if (b.ncouples) {
        const int passes = b.method == 1 ? 2 : 1;
        for (int pass = 1; pass <= passes; ++pass) {
            hipLaunchKernelGGL(batch_cross<T>, dim3(b.couple_rows.count), dim3(256), 0, st, b.couples,
                               b.couple_rows.items, work, pass);
            hipLaunchKernelGGL(batch_cross_sum<T>, dim3(b.couple_list.count), dim3(256), 0, st, b.couples,
                               b.couple_list.items, work, pass);
            hipLaunchKernelGGL(batch_project<T>, dim3(b.couple_rows.count), dim3(256), 0, st, b.couples,
                               b.couple_rows.items, work, pass);
        }
    }
    launch_panel_stage<T>(b, b.stage[1], work, st);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

}  // namespace tsqr
