// tsqr_launch.hpp -- entry from the host driver into the per-dtype kernel
// translation units (tsqr_f32.hip, tsqr_f64.hip), which are compiled in
// parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace tsqr {

// Columns handled by one kernel instantiation; n is padded up to the next
// of these with zero columns, which produce identity reflectors.
inline int padded_cols(int n) { return n <= 8 ? 8 : (n <= 16 ? 16 : 32); }

// Rows per lane: 512 bytes of tile per lane (128 VGPRs) for every shape.
template <typename T>
constexpr int rows_per_lane(int N) {
    return 512 / (N * (int)sizeof(T));
}

template <typename T>
inline bool plan_tsqr(int m, int n, TreeDesc* d) {
    return make_tree(m, n, 64 * rows_per_lane<T>(padded_cols(n)), d);
}

// Optional per-phase timing (tsqr_timing_* in the C ABI).  When enabled,
// the launchers bracket each phase with HIP events on the launch stream.
enum Phase {
    kPhaseUp = 0,
    kPhaseTree = 1,
    kPhaseDown = 2,
    kPhaseLeafQ = 3,
    kPhaseGemmTN = 4,
    kPhaseGemmNN = 5,
    kNumPhases = 6
};
bool timing_enabled();
void timing_mark(int phase, bool begin, hipStream_t st);

// Enqueues one TSQR (plan d) on stream st; work holds
// tree_workspace_bytes(d, sizeof(T)) bytes.  Returns a TSQR_* status.
template <typename T>
int tsqr_launch(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, hipStream_t st);

extern template int tsqr_launch<float>(const TreeDesc&, float*, int, float*, int, void*, hipStream_t);
extern template int tsqr_launch<double>(const TreeDesc&, double*, int, double*, int, void*, hipStream_t);

}  // namespace tsqr
