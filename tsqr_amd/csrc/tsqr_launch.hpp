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

// Householder tiles per level-0 superleaf: up to 8, and no more than fit
// their stacked N x N R factors into one H-row tile (H/N of them).
template <typename T>
constexpr int super_tiles(int N) {
    return (64 * rows_per_lane<T>(N)) / N < 8 ? (64 * rows_per_lane<T>(N)) / N : 8;
}

// Waves the level-0 passes aim to keep busy: two per SIMD of a 256-CU part.
constexpr int kTargetLeaves = 2048;

// Tiles per superleaf chosen by the autotuner (tsqr_set_leaf_tiles) for
// (dtype index 0 float / 1 double, padded width N, floor(log2 m)); 0 when
// the bucket has not been tuned.  Defined in tsqr_amd.hip.
int tuned_leaf_tiles(int dtype_index, int N, int m);
// Tree arity chosen by the autotuner (tsqr_set_tree_arity) for (dtype index,
// padded width N); 0 when untuned (full blocks, arity H / n).
int tuned_tree_arity(int dtype_index, int N);
// Waves the Gram Q pass aims for when superleaves are few (tsqr_set_stream_shape).
int tuned_gramq_waves();

// Superleaf height: the tuned number of tiles for this shape if there is
// one, else about m / kTargetLeaves rows, a whole number of tiles, between
// one tile and super_tiles tiles.  Large m gets full superleaves (fewer,
// longer streams and a smaller tree); small m gets one-tile superleaves so
// that the level-0 passes still fill the GPU.
template <typename T>
inline int leaf_rows_for(int m, int n) {
    const int N = padded_cols(n);
    const int H = 64 * rows_per_lane<T>(N);
    const int tuned = tuned_leaf_tiles(sizeof(T) == 4 ? 0 : 1, N, m);
    if (tuned > 0) return (tuned < super_tiles<T>(N) ? tuned : super_tiles<T>(N)) * H;
    const int64_t want = ((int64_t)m + kTargetLeaves - 1) / kTargetLeaves;
    int64_t tiles = (want + H - 1) / H;
    if (tiles < 1) tiles = 1;
    if (tiles > super_tiles<T>(N)) tiles = super_tiles<T>(N);
    return (int)tiles * H;
}

// Plans the tree of a call on leaf path `path` (LeafPath):
//   GRAM_TREE    superleaves of leaf_rows_for rows, each with a side block
//                (its Gram matrix, then R^{-1});
//   AUTO, HOUSEHOLDER, CHOLQR2  leaves of one tile (the Householder
//                kernels of leaf1.hpp), no per-leaf side blocks (the flat
//                passes stream the whole matrix).
template <typename T>
inline bool plan_tsqr(int m, int n, TreeDesc* d, int path) {
    const int N = padded_cols(n);
    const int H = 64 * rows_per_lane<T>(N);
    const int sl = leaf_rows_for<T>(m, n);
    const bool tiles = path != 2;  // kLeafGramTree (LeafPath, below)
    if (!make_tree(m, n, H, tiles ? H : sl, d, tuned_tree_arity(sizeof(T) == 4 ? 0 : 1, N))) return false;
    if (path != 2) d->side_leaves = 0;  // all but kLeafGramTree
    return true;
}

// How the level-0 superleaves are factored.
enum LeafPath {
    kLeafAuto = 0,         // flat Gram path if the whole matrix passes the gate, else Householder
    kLeafHouseholder = 1,  // Householder for every superleaf
    kLeafGramTree = 2,     // Gram path per superleaf (Householder where rejected) + Householder tree
    kLeafCholQR2 = 3,      // flat CholeskyQR with a residual check and an optional second pass
                           // (mixed.hpp): BF16x3 Gram for FP32 (the mixed-precision method),
                           // FP64 MFMA Gram for FP64; Householder tree if the gate rejects
};

static_assert(kLeafAuto == 0 && kLeafHouseholder == 1 && kLeafGramTree == 2 && kLeafCholQR2 == 3,
              "plan_tsqr uses these values");

struct LaunchOpts {
    int path;        // LeafPath
    double gate;     // largest accepted ||R||_F ||R^{-1}||_F on the Gram path
    double tol;      // mixed path: largest accepted ||Q^T Q - I||_F
    int r_only = 0;  // R only: A is read, never written (tsqr_r_*)
    int untimed = 0;  // record no phase events even while timing is enabled
};

// Optional per-phase timing (tsqr_timing_* in the C ABI).  When enabled,
// the launchers bracket each phase with HIP events on the launch stream.
enum Phase {
    kPhaseLeafR = 0,  // Gram R pass + Householder superleaf R pass
    kPhaseTree = 1,   // tree walk up
    kPhaseDown = 2,   // tree down sweep + Gram M products
    kPhaseLeafQ = 3,  // Gram Q pass + Householder superleaf Q pass
    kPhaseGemmTN = 4,
    kPhaseGemmNN = 5,
    kNumPhases = 6
};
bool timing_enabled();
void timing_mark(int phase, bool begin, hipStream_t st);

// 16-byte vector loads are possible at every column start.
inline bool aligned16(const void* p, int ld, size_t elem) {
    return ((uintptr_t)p % 16 == 0) && (((size_t)ld * elem) % 16 == 0);
}

// Enqueues the device-side checker's three Gram reductions (check.hpp) on
// st; work holds check_workspace_bytes(m) bytes.  On completion the group
// sums of A^T A, Q^T Q and E^T E (E = A - QR) start at byte offsets[0..2]
// of work: ceil(stream_groups(m) / kStreamGroup) blocks of 32 x 32 FP64
// each, to be added in order.  n <= 32.
template <typename T>
int check_launch(int m, int n, const T* A, int lda, const T* Q, int ldq, const T* R, int ldr, void* work,
                 size_t* offsets, hipStream_t st);
extern template int check_launch<float>(int, int, const float*, int, const float*, int, const float*, int, void*,
                                        size_t*, hipStream_t);
extern template int check_launch<double>(int, int, const double*, int, const double*, int, const double*, int,
                                         void*, size_t*, hipStream_t);

// Enqueues one TSQR (plan d) on stream st; work holds
// tree_workspace_bytes(d, sizeof(T)) bytes.  Returns a TSQR_* status.
template <typename T>
int tsqr_launch(const TreeDesc& d, T* A, int lda, T* R, int ldr, void* work, const LaunchOpts& o, hipStream_t st);

extern template int tsqr_launch<float>(const TreeDesc&, float*, int, float*, int, void*, const LaunchOpts&,
                                       hipStream_t);
extern template int tsqr_launch<double>(const TreeDesc&, double*, int, double*, int, void*, const LaunchOpts&,
                                        hipStream_t);

}  // namespace tsqr
