// batch_launch.hpp -- host-side description of a planned batch (device
// pointers into the plan's memory) and the launch sequence, instantiated per
// dtype in tsqr_f32.hip / tsqr_f64.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "plan.hpp"

namespace tsqr {

struct BatchProb;
struct BatchCouple;

struct ItemList {
    const int2* items = nullptr;  // device
    int count = 0;
};

// One stage of panel problems (all problems reference the batch-wide
// problem array by index).
struct BatchStage {
    ItemList probs;              // (problem, 0) for every problem of the stage
    ItemList leaves;             // every superleaf
    ItemList groups;             // (problem, g < batch_groups(superleaves)): the streaming passes
    ItemList level1;             // every level-1 tree block
    std::vector<ItemList> down;  // down[l]: every tile of level l (l >= 1)
    std::vector<ItemList> up;    // up[l]: every tile of level l >= 1, root included
    int64_t fallback = 0;        // workspace byte offset of the mixed method's Householder list
};

struct BatchDev {
    int method = 0;  // 0 Householder, 1 mixed (FP32)
    double gate = 1024.0, tol = 8e-7;
    const BatchProb* probs = nullptr;  // device
    int nprobs = 0;
    const BatchCouple* couples = nullptr;  // device
    int ncouples = 0;
    ItemList couple_rows;  // (couple, g < batch_groups(row blocks)) for every couple
    ItemList couple_list;  // (couple, 0) for every couple
    BatchStage stage[2];  // first panels (and whole matrices with n <= 32), second panels
};

template <typename T>
int batch_launch(const BatchDev& b, void* work, hipStream_t st);
extern template int batch_launch<float>(const BatchDev&, void*, hipStream_t);
extern template int batch_launch<double>(const BatchDev&, void*, hipStream_t);

}  // namespace tsqr
