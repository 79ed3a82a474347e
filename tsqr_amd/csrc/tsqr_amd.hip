// tsqr_amd.hip -- host drivers and the C ABI of libtsqr_amd.so.
//
// tsqr_* : one TSQR of an m x n (n <= 32) column-major matrix; the kernels
//          (tsqr_kernels.hpp) are compiled per dtype in tsqr_f32.hip and
//          tsqr_f64.hip.
// qr_*   : the reference's recursive column split (block_QR.h:72-95) with
//          the MFMA coupling products of gemm_mfma.hpp.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/tsqr_amd.h"
#include "gemm_mfma.hpp"
#include "tsqr_launch.hpp"

namespace tsqr {

// ------------------------------------------------------------------------
// Phase timing: event pairs recorded on the launch stream, summed on read.
// ------------------------------------------------------------------------
namespace {
std::atomic<bool> g_timing{false};
std::mutex g_timing_mu;
struct EventPair {
    int phase;
    hipEvent_t begin, end;
};
std::vector<EventPair> g_events;
}  // namespace

bool timing_enabled() { return g_timing.load(std::memory_order_relaxed); }

void timing_mark(int phase, bool begin, hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    if (begin) {
        EventPair e{phase, nullptr, nullptr};
        if (hipEventCreate(&e.begin) != hipSuccess || hipEventCreate(&e.end) != hipSuccess) return;
        (void)hipEventRecord(e.begin, st);
        g_events.push_back(e);
    } else {
        for (auto it = g_events.rbegin(); it != g_events.rend(); ++it) {
            if (it->phase == phase) {
                (void)hipEventRecord(it->end, st);
                break;
            }
        }
    }
}

// ------------------------------------------------------------------------
// Tuning table (tsqr_set_leaf_tiles, tsqr_set_gemm_waves).
// ------------------------------------------------------------------------
namespace {
std::atomic<int> g_leaf_tiles[2][3][32];  // zero-initialised: nothing tuned
std::atomic<int> g_tn_waves{512};

inline int width_class(int N) { return N <= 8 ? 0 : (N <= 16 ? 1 : 2); }
inline int log2_floor(int64_t m) {
    int k = 0;
    while (k < 31 && (m >> (k + 1)) != 0) ++k;
    return k;
}
}  // namespace

int tuned_leaf_tiles(int dtype_index, int N, int m) {
    if (dtype_index < 0 || dtype_index > 1 || m < 1) return 0;
    return g_leaf_tiles[dtype_index][width_class(N)][log2_floor(m)].load(std::memory_order_relaxed);
}

namespace {
std::atomic<int> g_tree_arity[2][3];  // 0: full blocks (H / n)
// one workgroup of 4 waves per CU measured fastest for the FP64-MFMA-bound
// R pass on MI355X (23.9M x 32 FP32: 1.853 ms per call at 256, 1.873 at
// 512, 1.905 at 1024)
constexpr int kStreamDefaultGroups = 256;
std::atomic<int> g_stream_max_groups{kStreamDefaultGroups};
std::atomic<int> g_gramq_waves{2048};
}  // namespace

int tuned_tree_arity(int dtype_index, int N) {
    if (dtype_index < 0 || dtype_index > 1) return 0;
    return g_tree_arity[dtype_index][width_class(N)].load(std::memory_order_relaxed);
}
int tuned_stream_max_groups() { return g_stream_max_groups.load(std::memory_order_relaxed); }
int tuned_gramq_waves() { return g_gramq_waves.load(std::memory_order_relaxed); }

namespace {

struct Pool {
    void* ptr = nullptr;
    size_t bytes = 0;
};
// Per-device internal workspaces: one for the reference-shaped entry points
// (whose ldwork is short), one for the device-side checker.  A caller holds
// the pool's mutex from the moment it takes the buffer until its launches
// are enqueued (the checker: until its result is back), so concurrent calls
// from several host threads take turns instead of sharing the tickets and
// partial sums, and a buffer is never freed while a call still holds it.
constexpr int kMaxDevices = 64;
std::mutex g_pool_mu[kMaxDevices];
std::mutex g_check_mu[kMaxDevices];
Pool g_pool[kMaxDevices];
Pool g_check_pool[kMaxDevices];

int current_device(int* dev) {
    if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= kMaxDevices) return TSQR_ERR_HIP;
    return TSQR_OK;
}

// Grows p to at least `bytes` (never shrinks); the caller holds p's mutex.
// Work enqueued earlier on the old buffer is waited for before it is freed.
int pool_reserve(Pool& p, size_t bytes, void** out) {
    if (p.bytes < bytes) {
        if (p.ptr) {
            if (hipDeviceSynchronize() != hipSuccess) return TSQR_ERR_HIP;
            (void)hipFree(p.ptr);
            p.ptr = nullptr;
            p.bytes = 0;
        }
        if (hipMalloc(&p.ptr, bytes) != hipSuccess) {
            p.ptr = nullptr;
            return TSQR_ERR_ALLOC;
        }
        p.bytes = bytes;
    }
    *out = p.ptr;
    return TSQR_OK;
}

template <typename T>
int check_args(int m, int n, const T* A, int lda, const T* R, int ldr) {
    if (!A || !R || n < 1 || m < n || lda < m || ldr < n) return TSQR_ERR_ARG;
    return TSQR_OK;
}

template <typename T>
size_t tsqr_ws_bytes(int m, int n) {
    if (n < 1 || m < n || n > 32) return 0;
    // enough for the plan of every leaf path, so that the size does not
    // depend on the path a call takes
    size_t most = 0;
    for (int path = kLeafAuto; path <= kLeafCholQR2; ++path) {
        TreeDesc d;
        if (!plan_tsqr<T>(m, n, &d, path)) return 0;
        const size_t b = tree_workspace_bytes(d, sizeof(T));
        most = b > most ? b : most;
    }
    return most;
}

// Leaf path per dtype (tsqr_set_leaf_path).  FP32 defaults to the Gram path
// with the Householder fallback; FP64 has no wider type to accumulate the
// Gram matrix in, so it defaults to Householder leaves.
// The mixed entry points have their own gate and residual tolerance
// (tsqr_set_mixed).  FP64 defaults to CholeskyQR2 with an FP64 Gram: the
// gate keeps u64 kappa^2 small enough for the second pass to restore
// orthogonality, and the residual check asks for 4 n u64 with n = 32.
constexpr double kTolF64 = 4.0 * 32.0 * 1.1102230246251565e-16;
constexpr double kGateF64 = 1e5;
std::mutex g_opts_mu;
LaunchOpts g_opts_f32{kLeafAuto, 1024.0, 0.0};
LaunchOpts g_opts_f64{kLeafCholQR2, kGateF64, kTolF64};
LaunchOpts g_opts_mixed{kLeafCholQR2, kMixedGate, 8e-7};

template <typename T>
LaunchOpts current_opts() {
    std::lock_guard<std::mutex> lock(g_opts_mu);
    return sizeof(T) == 4 ? g_opts_f32 : g_opts_f64;
}

LaunchOpts mixed_opts() {
    std::lock_guard<std::mutex> lock(g_opts_mu);
    return g_opts_mixed;
}

// Runs one TSQR; work must hold tsqr_ws_bytes<T>(m, n) bytes.
template <typename T>
int tsqr_core(int m, int n, T* A, int lda, T* R, int ldr, void* work, hipStream_t st, bool mixed = false,
              bool r_only = false, bool untimed = false) {
    LaunchOpts o = mixed ? mixed_opts() : current_opts<T>();
    TreeDesc d;
    if (!plan_tsqr<T>(m, n, &d, o.path)) return TSQR_ERR_SHAPE;
    o.r_only = r_only ? 1 : 0;
    o.untimed = untimed ? 1 : 0;
    return tsqr_launch<T>(d, A, lda, R, ldr, work, o, st);
}

// ------------------------------------------------------------------------
// The entry point that allocates nothing (tsqr_s_stream_noalloc).  Its
// workspace is `scratch`, a fixed array of this library in the memory of
// every GPU it runs on, held until the process ends.  In host code the
// array's name only stands in for the symbol: the device address comes from
// hipGetSymbolAddress, looked up once per device on the first call with
// that device current, and only that address is handed to kernels.  The
// launches clear every ticket, flag and control word they use first
// (init_flags), so the contents left by the previous call do not matter;
// calls that overlap on one GPU share the array and race.
// ------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

int scratch_address(void** out) {
    int dev = 0;
    int e = current_device(&dev);
    if (e) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TSQR_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TSQR_OK;
}

template <typename T>
int tsqr_stream(int m, int n, T* A, int lda, T* R, int ldr, void* work, size_t work_bytes, hipStream_t st,
                bool mixed = false, bool r_only = false) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    if (n > 32) return TSQR_ERR_SHAPE;
    const size_t need = tsqr_ws_bytes<T>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    if (work_bytes < need || (need > 0 && !work)) return TSQR_ERR_WORKSPACE;
    return tsqr_core<T>(m, n, A, lda, R, ldr, work, st, mixed, r_only);
}

// R only (tsqr_r_*): the reference-shaped entry with the library's pool when
// the caller's ldwork is short.  A is read and never written.
template <typename T>
int tsqr_r_ref(int m, int n, const T* A, int lda, T* R, int ldr, T* work, int ldwork);

// ------------------------------------------------------------------------
// Recursive blocked QR.
// ------------------------------------------------------------------------

// Left block width of the split: n/2 for n = 32*2^k (block_QR.h:81-83),
// a multiple of 32 in general.
inline int split_cols(int n) { return 32 * ((n + 63) / 64); }

// Rows per split-K chunk of R12 = Q1^T A2: enough chunks that the launch has
// about W waves, each chunk a multiple of 16 rows and at least 64 rows.  W
// is g_tn_waves (512 unless tuned) for short panels and grows with m, about
// one wave per 512 rows up to 4 g_tn_waves (measured best at m = 27,648 and
// at m = 1,048,576; a launch has at most one wave per SIMD either way).
inline int tn_chunk_rows(int m, int n1, int n2) {
    const int tiles = ((n1 + 63) / 64) * ((n2 + 63) / 64);
    const int64_t base = g_tn_waves.load(std::memory_order_relaxed);
    int64_t waves = (int64_t)m / 512;
    waves = waves < base ? base : (waves > 4 * base ? 4 * base : waves);
    const int chunks = (int)((waves + tiles - 1) / tiles);
    int kc = (m + chunks - 1) / chunks;
    kc = (kc + 15) / 16 * 16;
    return kc < 64 ? 64 : kc;
}

// Largest chunks * n1 * n2 over the recursion (the partial sums of R12).
inline size_t tn_partial_elems(int m, int n) {
    if (n <= 32) return 0;
    const int n1 = split_cols(n), n2 = n - n1;
    const int kc = tn_chunk_rows(m, n1, n2);
    const size_t here = (size_t)((m + kc - 1) / kc) * n1 * n2;
    const size_t a = tn_partial_elems(m, n1), b = tn_partial_elems(m, n2);
    size_t mx = here > a ? here : a;
    return mx > b ? mx : b;
}

// Largest TSQR workspace over the panels the recursion factors (widths
// between 1 and 32).  Narrow panels can need more than a 32-wide one: the
// tuned superleaf height depends on the padded width.
template <typename T>
size_t panel_ws_bytes(int m, int n) {
    if (n <= 32) return tsqr_ws_bytes<T>(m, n);
    const int n1 = split_cols(n);
    const size_t a = panel_ws_bytes<T>(m, n1), b = panel_ws_bytes<T>(m, n - n1);
    return a > b ? a : b;
}

// Workspace of the recursive QR (byte offsets, 256-byte aligned): the panel
// TSQR workspace, the split-K partials (of the coupling products, and of the
// mixed method's final Gram matrix Q^T Q, FP64, for 32 < n <= kReorthMax),
// n1 x n2 elements for the mixed method's second coupling product, then the
// final reorthogonalisation's G (n x n FP64), D (n x n FP32) and its status
// words and residual.
struct QrLayout {
    size_t panel, parts, tmp, gram, dmat, stat, total;
};

// Rows per split-K chunk of the final Gram matrix Q^T Q (n x n).
inline int reorth_chunk_rows(int m, int n) { return tn_chunk_rows(m, n, n); }

template <typename T>
QrLayout qr_layout(int m, int n) {
    QrLayout L{};
    L.panel = 0;
    const size_t panel = panel_ws_bytes<T>(m, n);
    if (panel == 0) return L;
    // FP64 partials for the mixed method's coupling products (FP32 data)
    size_t parts = tn_partial_elems(m, n) * sizeof(double);
    const bool reorth = n > 32 && n <= kReorthMax;
    if (reorth) {
        const int kc = reorth_chunk_rows(m, n);
        const size_t g = (size_t)((m + kc - 1) / kc) * n * n * sizeof(double);
        parts = parts > g ? parts : g;
    }
    const int n1 = split_cols(n);
    L.parts = align256(panel);
    L.tmp = L.parts + align256(parts);
    L.gram = L.tmp + align256((size_t)n1 * (n - n1) * sizeof(T));
    L.dmat = L.gram + (reorth ? align256((size_t)n * n * sizeof(double)) : 0);
    L.stat = L.dmat + (reorth ? align256((size_t)n * n * sizeof(float)) : 0);
    L.total = L.stat + (reorth ? 256 : 0);
    return L;
}

template <typename T>
size_t qr_ws_bytes(int m, int n) {
    if (n < 1 || m < n) return 0;
    if (n <= 32) return tsqr_ws_bytes<T>(m, n);
    return qr_layout<T>(m, n).total;
}

// Final CholeskyQR pass of the mixed-precision recursive QR over the whole
// Q (32 < n <= kReorthMax): G = Q^T Q from the BF16x3 split-K products with
// each 32-row step added to FP64 sums (tiles on and above the diagonal),
// e = ||G - I||_F; if
// e > tol, R_c = chol(G), Q <- Q R_c^{-1} in place and R <- R_c R
// (gemm_mfma.hpp).  R_c overwrites G.
int reorth(int m, int n, float* Q, int ldq, float* R, int ldr, char* ws, const QrLayout& L, double tol,
           hipStream_t st) {
    double* part = reinterpret_cast<double*>(ws + L.parts);
    double* G = reinterpret_cast<double*>(ws + L.gram);
    float* D = reinterpret_cast<float*>(ws + L.dmat);
    int* stat = reinterpret_cast<int*>(ws + L.stat);
    double* resid = reinterpret_cast<double*>(ws + L.stat + 16);
    const int kc = reorth_chunk_rows(m, n);
    const int chunks = (m + kc - 1) / kc;
    const bool vec = aligned16(Q, ldq, sizeof(float));
    const dim3 gtn(chunks, (n + 63) / 64, (n + 63) / 64);
    if (vec)
        hipLaunchKernelGGL((gemm_tn_bf16x3<true, double>), gtn, dim3(64), 0, st, m, n, n, (const float*)Q,
                           (int64_t)ldq, (const float*)Q, (int64_t)ldq, kc, part, 1);
    else
        hipLaunchKernelGGL((gemm_tn_bf16x3<false, double>), gtn, dim3(64), 0, st, m, n, n, (const float*)Q,
                           (int64_t)ldq, (const float*)Q, (int64_t)ldq, kc, part, 1);
    hipLaunchKernelGGL((gemm_tn_reduce<double, double>), dim3((n * n + 15) / 16), dim3(256), 0, st, n, n, chunks,
                       (const double*)part, G, (int64_t)n, 0, 1);
    hipLaunchKernelGGL(reorth_chol<>, dim3(1), dim3(kReorthThreads), 0, st, n, (const double*)G, G, stat, resid, tol);
    hipLaunchKernelGGL(reorth_inv<>, dim3(n), dim3(kReorthMax), 0, st, n, (const double*)G, D, (const int*)stat);
    hipLaunchKernelGGL(reorth_rupdate<float>, dim3(n), dim3(kReorthMax), 0, st, n, (const double*)G, R, (int64_t)ldr,
                       (const int*)stat);
    // Q <- Q - Q Dneg, the last column group first (reorth_apply)
    const dim3 gap((m + 63) / 64);
    for (int c0 = ((n - 1) / 64) * 64; c0 >= 0; c0 -= 64) {
        const int N = n - c0 < 64 ? n - c0 : 64;
        if (vec)
            hipLaunchKernelGGL(reorth_apply<true>, gap, dim3(64), 0, st, m, c0 + N, N, Q, (int64_t)ldq, c0,
                               (const float*)D, (int64_t)n, (const int*)stat);
        else
            hipLaunchKernelGGL(reorth_apply<false>, gap, dim3(64), 0, st, m, c0 + N, N, Q, (int64_t)ldq, c0,
                               (const float*)D, (int64_t)n, (const int*)stat);
    }
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

// One coupling step: part = chunk products of Q1^T A2, C = their sum (or
// C += it), A2 -= Q1 (the new sum), as the reference's two gemm_blockqr
// calls (block_QR.h:88-93).
// bf16: FP32 data on the BF16 matrix cores with BF16x3 split operands
// (gemm_tn_bf16x3 / gemm_nn_bf16x3), the mixed-precision entry points.
template <typename T>
int couple(int m, int n1, int n2, const T* A1, T* A2, int lda, T* C, int ldc, T* part, hipStream_t st,
           bool bf16 = false) {
    const int kc = tn_chunk_rows(m, n1, n2);
    const int chunks = (m + kc - 1) / kc;
    const bool timed = timing_enabled();
    const bool vec = ((uintptr_t)A1 % 16 == 0) && (((size_t)lda * sizeof(T)) % 16 == 0);
    if constexpr (sizeof(T) == 4) {
        if (bf16) {
            const dim3 gtn(chunks, (n1 + 63) / 64, (n2 + 63) / 64), gnn((m + 63) / 64, (n2 + 63) / 64);
            if (timed) timing_mark(kPhaseGemmTN, true, st);
            // each 32-row step summed into FP64 (part holds FP64 partials):
            // the error of R12 sets how far the coupled blocks stay from
            // orthogonal, amplified by cond(R22)
            double* part64 = reinterpret_cast<double*>(part);
            if (vec)
                hipLaunchKernelGGL((gemm_tn_bf16x3<true, double>), gtn, dim3(64), 0, st, m, n1, n2, A1, (int64_t)lda,
                                   (const T*)A2, (int64_t)lda, kc, part64);
            else
                hipLaunchKernelGGL((gemm_tn_bf16x3<false, double>), gtn, dim3(64), 0, st, m, n1, n2, A1,
                                   (int64_t)lda, (const T*)A2, (int64_t)lda, kc, part64);
            hipLaunchKernelGGL((gemm_tn_reduce<double, T>), dim3((n1 * n2 + 15) / 16), dim3(256), 0, st, n1, n2,
                               chunks, (const double*)part64, C, (int64_t)ldc, 0);
            if (timed) timing_mark(kPhaseGemmTN, false, st);
            if (timed) timing_mark(kPhaseGemmNN, true, st);
            if (vec)
                hipLaunchKernelGGL(gemm_nn_bf16x3<true>, gnn, dim3(64), 0, st, m, n1, n2, A1, (int64_t)lda,
                                   (const T*)C, (int64_t)ldc, A2, (int64_t)lda);
            else
                hipLaunchKernelGGL(gemm_nn_bf16x3<false>, gnn, dim3(64), 0, st, m, n1, n2, A1, (int64_t)lda,
                                   (const T*)C, (int64_t)ldc, A2, (int64_t)lda);
            if (timed) timing_mark(kPhaseGemmNN, false, st);
            return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
        }
    }
    if (timed) timing_mark(kPhaseGemmTN, true, st);
    if (vec)
        hipLaunchKernelGGL((gemm_tn_mfma<T, true>), dim3(chunks, (n1 + 63) / 64, (n2 + 63) / 64), dim3(64), 0, st, m,
                           n1, n2, A1, (int64_t)lda, (const T*)A2, (int64_t)lda, kc, part);
    else
        hipLaunchKernelGGL((gemm_tn_mfma<T, false>), dim3(chunks, (n1 + 63) / 64, (n2 + 63) / 64), dim3(64), 0, st, m,
                           n1, n2, A1, (int64_t)lda, (const T*)A2, (int64_t)lda, kc, part);
    hipLaunchKernelGGL(gemm_tn_reduce<T>, dim3((n1 * n2 + 15) / 16), dim3(256), 0, st, n1, n2, chunks,
                       (const T*)part, C, (int64_t)ldc, 0);
    if (timed) timing_mark(kPhaseGemmTN, false, st);
    if (timed) timing_mark(kPhaseGemmNN, true, st);
    if (vec)
        hipLaunchKernelGGL((gemm_nn_mfma<T, true>), dim3((m + 63) / 64, (n2 + 63) / 64), dim3(64), 0, st, m, n1, n2,
                           A1, (int64_t)lda, (const T*)C, (int64_t)ldc, A2, (int64_t)lda);
    else
        hipLaunchKernelGGL((gemm_nn_mfma<T, false>), dim3((m + 63) / 64, (n2 + 63) / 64), dim3(64), 0, st, m, n1, n2,
                           A1, (int64_t)lda, (const T*)C, (int64_t)ldc, A2, (int64_t)lda);
    if (timed) timing_mark(kPhaseGemmNN, false, st);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

// Recursive blocked QR (block_QR.h:72-95).  mixed (FP32): the panels are the
// mixed-precision TSQR, the coupling products run on the BF16 matrix cores
// with BF16x3 split operands and FP32 accumulation, and the coupling is
// applied twice (block classical Gram-Schmidt with reorthogonalisation,
// BCGS2): the second product C2 = Q1^T A2 is subtracted again and added to
// R12, which restores the orthogonality of the column blocks that one pass
// loses in proportion to cond(A).  tmp holds n1 x n2 elements.
template <typename T>
int qr_rec(int m, int n, T* A, int lda, T* R, int ldr, char* ws, size_t panel_bytes, T* part, T* tmp,
           hipStream_t st, bool mixed) {
    if (n <= 32) return tsqr_core<T>(m, n, A, lda, R, ldr, ws, st, mixed);
    const int n1 = split_cols(n), n2 = n - n1;
    T* A1 = A;
    T* A2 = A + (int64_t)n1 * lda;
    T* R11 = R;
    T* R12 = R + (int64_t)n1 * ldr;
    T* R22 = R + (int64_t)n1 * ldr + n1;
    int e = qr_rec<T>(m, n1, A1, lda, R11, ldr, ws, panel_bytes, part, tmp, st, mixed);
    if (e) return e;
    e = couple<T>(m, n1, n2, A1, A2, lda, R12, ldr, part, st, mixed);
    if (e) return e;
    if (mixed) {
        e = couple<T>(m, n1, n2, A1, A2, lda, tmp, n1, part, st, true);
        if (e) return e;
        hipLaunchKernelGGL(add_block<T>, dim3((n1 * n2 + 255) / 256), dim3(256), 0, st, (const T*)tmp, n1, R12,
                           (int64_t)ldr, n1, n2);
    }
    if (hipGetLastError() != hipSuccess) return TSQR_ERR_HIP;
    return qr_rec<T>(m, n2, A2, lda, R22, ldr, ws, panel_bytes, part, tmp, st, mixed);
}

template <typename T>
int qr_stream(int m, int n, T* A, int lda, T* R, int ldr, void* work, size_t work_bytes, hipStream_t st,
              bool mixed = false) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = qr_ws_bytes<T>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    if (work_bytes < need || !work) return TSQR_ERR_WORKSPACE;
    if (n <= 32) return tsqr_core<T>(m, n, A, lda, R, ldr, work, st, mixed);
    const QrLayout L = qr_layout<T>(m, n);
    char* ws = reinterpret_cast<char*>(work);
    // every R21 block of the recursion at once (the panels write their own
    // diagonal blocks, zeros included)
    hipLaunchKernelGGL(zero_lower<T>, dim3((unsigned)(((int64_t)n * n + 255) / 256)), dim3(256), 0, st, R,
                       (int64_t)ldr, n);
    e = qr_rec<T>(m, n, A, lda, R, ldr, ws, L.parts, reinterpret_cast<T*>(ws + L.parts),
                  reinterpret_cast<T*>(ws + L.tmp), st, mixed);
    if (e) return e;
    if constexpr (sizeof(T) == 4) {
        if (mixed && n <= kReorthMax) return reorth(m, n, A, lda, R, ldr, ws, L, mixed_opts().tol, st);
    }
    return TSQR_OK;
}

// Eigenvalues of the symmetric n x n matrix S (row-major, ld 32, n <= 32;
// destroyed) by cyclic Jacobi rotations, into ev[0..n).
void sym_eigenvalues(double (*S)[32], int n, double* ev) {
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += S[i][i] * S[i][i];
            for (int j = i + 1; j < n; ++j) off += S[i][j] * S[i][j];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (S[p][q] == 0.0) continue;
                // rotation that zeroes S[p][q] (Golub and Van Loan, Algorithm 8.5.1)
                const double tau = (S[q][q] - S[p][p]) / (2.0 * S[p][q]);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double skp = S[k][p], skq = S[k][q];
                    S[k][p] = c * skp - sn * skq;
                    S[k][q] = sn * skp + c * skq;
                }
                for (int k = 0; k < n; ++k) {
                    const double spk = S[p][k], sqk = S[q][k];
                    S[p][k] = c * spk - sn * sqk;
                    S[q][k] = sn * spk + c * sqk;
                }
            }
    }
    for (int i = 0; i < n; ++i) ev[i] = S[i][i];
}

// norms[0..3] of tsqr_check_* from the device Gram reductions.
template <typename T>
int check_impl(int m, int n, const T* A, int lda, const T* Q, int ldq, const T* R, int ldr, double* norms,
               hipStream_t st) {
    if (!A || !Q || !R || !norms || n < 1 || m < n || lda < m || ldq < m || ldr < n) return TSQR_ERR_ARG;
    if (n > 32) return TSQR_ERR_SHAPE;
    int dev = 0;
    int e = current_device(&dev);
    if (e) return e;
    // the pool is held until the sums have been copied back
    std::lock_guard<std::mutex> lock(g_check_mu[dev]);
    void* work = nullptr;
    e = pool_reserve(g_check_pool[dev], check_workspace_bytes(m), &work);
    if (e) return e;
    size_t off[3];
    e = check_launch<T>(m, n, A, lda, Q, ldq, R, ldr, work, off, st);
    if (e) return e;
    const int groups = stream_group_count(stream_groups(m));
    std::vector<double> host((size_t)3 * groups * 1024);
    for (int k = 0; k < 3; ++k)
        if (hipMemcpyAsync(host.data() + (size_t)k * groups * 1024, reinterpret_cast<const char*>(work) + off[k],
                           (size_t)groups * 1024 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
            return TSQR_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TSQR_ERR_HIP;
    double G[3][32][32];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 32; ++i)
            for (int j = 0; j < 32; ++j) {
                double sum = 0.0;
                for (int g = 0; g < groups; ++g) sum += host[((size_t)k * groups + g) * 1024 + i * 32 + j];
                G[k][i][j] = sum;
            }
    double ev[32];
    double trA = 0.0, trE = 0.0, fQ = 0.0;
    for (int i = 0; i < n; ++i) {
        trA += G[0][i][i];
        trE += G[2][i][i];
        for (int j = 0; j < n; ++j) {
            const double d = G[1][i][j] - (i == j ? 1.0 : 0.0);
            fQ += d * d;
        }
    }
    auto lmax = [&](int k, bool shift) {
        for (int i = 0; i < n; ++i) G[k][i][i] -= shift ? 1.0 : 0.0;
        sym_eigenvalues(G[k], n, ev);
        double mx = 0.0;
        for (int i = 0; i < n; ++i) mx = std::fmax(mx, shift ? std::fabs(ev[i]) : ev[i]);
        return mx;
    };
    const double a2 = lmax(0, false), q2 = lmax(1, true), e2 = lmax(2, false);
    norms[0] = trA > 0.0 ? std::sqrt(std::fmax(trE, 0.0) / trA) : 0.0;
    norms[1] = std::sqrt(fQ);
    norms[2] = a2 > 0.0 ? std::sqrt(e2 / a2) : 0.0;
    norms[3] = q2;
    return TSQR_OK;
}

// ------------------------------------------------------------------------
// Least squares and Q^T B on the implicit Q = A R^{-1} (R-only mode).
// ------------------------------------------------------------------------

// C (n1 x n2, ldc) = X^T Y over m rows: split-K MFMA products and their
// ordered sum (the coupling kernels of qr_rec).
template <typename T>
int tn_product(int m, int n1, int n2, const T* X, int ldx, const T* Y, int ldy, T* C, int ldc, T* part,
               hipStream_t st) {
    const int kc = tn_chunk_rows(m, n1, n2);
    const int chunks = (m + kc - 1) / kc;
    const dim3 grid(chunks, (n1 + 63) / 64, (n2 + 63) / 64);
    if (aligned16(X, ldx, sizeof(T)) && aligned16(Y, ldy, sizeof(T)))
        hipLaunchKernelGGL((gemm_tn_mfma<T, true>), grid, dim3(64), 0, st, m, n1, n2, X, (int64_t)ldx, Y,
                           (int64_t)ldy, kc, part);
    else
        hipLaunchKernelGGL((gemm_tn_mfma<T, false>), grid, dim3(64), 0, st, m, n1, n2, X, (int64_t)ldx, Y,
                           (int64_t)ldy, kc, part);
    hipLaunchKernelGGL(gemm_tn_reduce<T>, dim3((n1 * n2 + 15) / 16), dim3(256), 0, st, n1, n2, chunks,
                       (const T*)part, C, (int64_t)ldc, 0);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

// C (m x N, ldc) -= X (m x K, ldx) B (K x N, ldb).
template <typename T>
int nn_update(int m, int K, int N, const T* X, int ldx, const T* B, int ldb, T* C, int ldc, hipStream_t st) {
    const dim3 grid((m + 63) / 64, (N + 63) / 64);
    if (aligned16(X, ldx, sizeof(T)) && aligned16(C, ldc, sizeof(T)))
        hipLaunchKernelGGL((gemm_nn_mfma<T, true>), grid, dim3(64), 0, st, m, K, N, X, (int64_t)ldx, B,
                           (int64_t)ldb, C, (int64_t)ldc);
    else
        hipLaunchKernelGGL((gemm_nn_mfma<T, false>), grid, dim3(64), 0, st, m, K, N, X, (int64_t)ldx, B,
                           (int64_t)ldb, C, (int64_t)ldc);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

// Workspace of tsqr_lstsq_* / tsqr_apply_qt_*: the R-only TSQR's, then R
// (n x n), the split-K partials of an n x nrhs product, and two n x nrhs
// blocks.
struct SolveLayout {
    size_t tsqr, R, part, C, D, total;
};
template <typename T>
SolveLayout solve_layout(int m, int n, int nrhs) {
    SolveLayout L;
    L.tsqr = 0;
    L.R = align256(tsqr_ws_bytes<T>(m, n));
    const int kc = tn_chunk_rows(m, n, nrhs);
    const size_t chunks = (size_t)((m + kc - 1) / kc);
    L.part = L.R + align256((size_t)n * n * sizeof(T));
    L.C = L.part + align256(chunks * n * nrhs * sizeof(T));
    L.D = L.C + align256((size_t)n * nrhs * sizeof(T));
    L.total = L.D + align256((size_t)n * nrhs * sizeof(T));
    return L;
}

template <typename T>
int solve_args(int m, int n, int nrhs, const T* A, int lda, const void* B, int ldb, const void* X, int ldx) {
    if (!A || !B || !X || n < 1 || m < n || nrhs < 1 || lda < m || ldb < m || ldx < n) return TSQR_ERR_ARG;
    if (n > 32) return TSQR_ERR_SHAPE;
    return TSQR_OK;
}

// min ||A X - B||_F by the corrected seminormal equations (Bjorck): R from
// the R-only TSQR, X = R^{-1} R^{-T} A^T B, then one refinement step with
// the residual E = B - A X: X += R^{-1} R^{-T} A^T E.  B is overwritten by
// the final residual.
template <typename T>
int lstsq_stream(int m, int n, int nrhs, const T* A, int lda, T* B, int ldb, T* X, int ldx, void* work,
                 size_t work_bytes, hipStream_t st) {
    int e = solve_args<T>(m, n, nrhs, A, lda, B, ldb, X, ldx);
    if (e) return e;
    const SolveLayout L = solve_layout<T>(m, n, nrhs);
    if (!work || work_bytes < L.total) return TSQR_ERR_WORKSPACE;
    char* w = reinterpret_cast<char*>(work);
    T* R = reinterpret_cast<T*>(w + L.R);
    T* part = reinterpret_cast<T*>(w + L.part);
    T* C = reinterpret_cast<T*>(w + L.C);
    T* D = reinterpret_cast<T*>(w + L.D);
    const dim3 tgrid((nrhs + 255) / 256);
    e = tsqr_core<T>(m, n, const_cast<T*>(A), lda, R, n, w + L.tsqr, st, false, true);
    if (e) return e;
    e = tn_product<T>(m, n, nrhs, A, lda, B, ldb, C, n, part, st);  // C = A^T B
    if (e) return e;
    hipLaunchKernelGGL(tri_normal_solve<T>, tgrid, dim3(256), 0, st, n, nrhs, (const T*)R, (int64_t)n, (const T*)C,
                       (int64_t)n, X, (int64_t)ldx, 0, 0);
    e = nn_update<T>(m, n, nrhs, A, lda, X, ldx, B, ldb, st);  // B = E = B - A X
    if (e) return e;
    e = tn_product<T>(m, n, nrhs, A, lda, B, ldb, C, n, part, st);  // C = A^T E
    if (e) return e;
    hipLaunchKernelGGL(tri_normal_solve<T>, tgrid, dim3(256), 0, st, n, nrhs, (const T*)R, (int64_t)n, (const T*)C,
                       (int64_t)n, D, (int64_t)n, 0, 0);  // D = correction
    hipLaunchKernelGGL(add_block<T>, dim3((n * nrhs + 255) / 256), dim3(256), 0, st, (const T*)D, n, X,
                       (int64_t)ldx, n, nrhs);
    return nn_update<T>(m, n, nrhs, A, lda, D, n, B, ldb, st);  // B = B - A D
}

// C (n x k) = Q^T B for the implicit Q = A R^{-1} of tsqr_r_* (R given).
template <typename T>
int apply_qt_stream(int m, int n, int k, const T* A, int lda, const T* R, int ldr, const T* B, int ldb, T* C, int ldc,
                    void* work, size_t work_bytes, hipStream_t st) {
    int e = solve_args<T>(m, n, k, A, lda, B, ldb, C, ldc);
    if (e) return e;
    if (!R || ldr < n) return TSQR_ERR_ARG;
    const SolveLayout L = solve_layout<T>(m, n, k);
    if (!work || work_bytes < L.total) return TSQR_ERR_WORKSPACE;
    char* w = reinterpret_cast<char*>(work);
    T* part = reinterpret_cast<T*>(w + L.part);
    T* Cw = reinterpret_cast<T*>(w + L.C);
    e = tn_product<T>(m, n, k, A, lda, B, ldb, Cw, n, part, st);  // A^T B
    if (e) return e;
    hipLaunchKernelGGL(tri_normal_solve<T>, dim3((k + 255) / 256), dim3(256), 0, st, n, k, R, (int64_t)ldr,
                       (const T*)Cw, (int64_t)n, C, (int64_t)ldc, 0, 1);
    return hipGetLastError() == hipSuccess ? TSQR_OK : TSQR_ERR_HIP;
}

// Runs fn(ws, bytes) on the caller's work array of the reference-shaped
// entry points (ldwork * n elements, test_tsqr.cu:42) if it holds `need`
// bytes at a 256-byte aligned address, else on the per-device pool, whose
// mutex is held until fn has enqueued its launches.
template <typename T, typename F>
int with_ref_workspace(int n, T* work, int ldwork, size_t need, F&& fn) {
    const size_t have = (work && ldwork > 0) ? (size_t)ldwork * (size_t)n * sizeof(T) : 0;
    if (have >= need && ((uintptr_t)work % 256) == 0) return fn((void*)work, have);
    int dev = 0;
    int e = current_device(&dev);
    if (e) return e;
    std::lock_guard<std::mutex> lock(g_pool_mu[dev]);
    void* ws = nullptr;
    e = pool_reserve(g_pool[dev], need, &ws);
    if (e) return e;
    return fn(ws, need);
}

template <typename T>
int tsqr_r_ref(int m, int n, const T* A, int lda, T* R, int ldr, T* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = tsqr_ws_bytes<T>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return tsqr_stream<T>(m, n, const_cast<T*>(A), lda, R, ldr, ws, bytes, nullptr, false, true);
    });
}

}  // namespace
}  // namespace tsqr

using namespace tsqr;

extern "C" {

const char* tsqr_status_string(int status) {
    switch (status) {
        case TSQR_OK:
            return "ok";
        case TSQR_ERR_ARG:
            return "invalid argument (null pointer, m < n, n < 1, lda < m or ldr < n)";
        case TSQR_ERR_SHAPE:
            return "shape not supported by this entry point";
        case TSQR_ERR_WORKSPACE:
            return "workspace smaller than required";
        case TSQR_ERR_HIP:
            return "HIP runtime error";
        case TSQR_ERR_ALLOC:
            return "internal workspace allocation failed";
        case TSQR_ERR_SCRATCH:
            return "problem needs more workspace than the 16 MiB fixed scratch buffer of tsqr_s_stream_noalloc";
        default:
            return "unknown status";
    }
}

int tsqr_check_s(int m, int n, const float* A, int lda, const float* Q, int ldq, const float* R, int ldr,
                 double* norms, void* stream) {
    return check_impl<float>(m, n, A, lda, Q, ldq, R, ldr, norms, (hipStream_t)stream);
}

int tsqr_check_d(int m, int n, const double* A, int lda, const double* Q, int ldq, const double* R, int ldr,
                 double* norms, void* stream) {
    return check_impl<double>(m, n, A, lda, Q, ldq, R, ldr, norms, (hipStream_t)stream);
}

size_t tsqr_solve_workspace_bytes(int m, int n, int nrhs, int dtype) {
    if (n < 1 || n > 32 || m < n || nrhs < 1) return 0;
    if (dtype == TSQR_FLOAT) return solve_layout<float>(m, n, nrhs).total;
    if (dtype == TSQR_DOUBLE) return solve_layout<double>(m, n, nrhs).total;
    return 0;
}

int tsqr_lstsq_s_stream(int m, int n, int nrhs, const float* A, int lda, float* B, int ldb, float* X, int ldx,
                        void* work, size_t work_bytes, void* stream) {
    return lstsq_stream<float>(m, n, nrhs, A, lda, B, ldb, X, ldx, work, work_bytes, (hipStream_t)stream);
}
int tsqr_lstsq_d_stream(int m, int n, int nrhs, const double* A, int lda, double* B, int ldb, double* X, int ldx,
                        void* work, size_t work_bytes, void* stream) {
    return lstsq_stream<double>(m, n, nrhs, A, lda, B, ldb, X, ldx, work, work_bytes, (hipStream_t)stream);
}
int tsqr_apply_qt_s_stream(int m, int n, int k, const float* A, int lda, const float* R, int ldr, const float* B,
                           int ldb, float* C, int ldc, void* work, size_t work_bytes, void* stream) {
    return apply_qt_stream<float>(m, n, k, A, lda, R, ldr, B, ldb, C, ldc, work, work_bytes, (hipStream_t)stream);
}
int tsqr_apply_qt_d_stream(int m, int n, int k, const double* A, int lda, const double* R, int ldr, const double* B,
                           int ldb, double* C, int ldc, void* work, size_t work_bytes, void* stream) {
    return apply_qt_stream<double>(m, n, k, A, lda, R, ldr, B, ldb, C, ldc, work, work_bytes, (hipStream_t)stream);
}

int tsqr_timing_enable(int enable) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    for (auto& e : g_events) {
        (void)hipEventDestroy(e.begin);
        (void)hipEventDestroy(e.end);
    }
    g_events.clear();
    g_timing.store(enable != 0);
    return TSQR_OK;
}

int tsqr_timing_read(double* ms, int* count, int nphases) {
    if (!ms || !count || nphases < 1) return TSQR_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_timing_mu);
    for (int p = 0; p < nphases; ++p) {
        ms[p] = 0.0;
        count[p] = 0;
    }
    int status = TSQR_OK;
    for (auto& e : g_events) {
        float t = 0.f;
        if (hipEventSynchronize(e.end) != hipSuccess || hipEventElapsedTime(&t, e.begin, e.end) != hipSuccess)
            status = TSQR_ERR_HIP;
        else if (e.phase < nphases) {
            ms[e.phase] += t;
            count[e.phase] += 1;
        }
        (void)hipEventDestroy(e.begin);
        (void)hipEventDestroy(e.end);
    }
    g_events.clear();
    return status;
}

int tsqr_set_leaf_path(int dtype, int path, double gate) {
    if ((path != TSQR_LEAF_AUTO && path != TSQR_LEAF_HOUSEHOLDER && path != TSQR_LEAF_GRAM_TREE &&
         path != TSQR_LEAF_CHOLQR2) ||
        !(gate >= 1.0))
        return TSQR_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_opts_mu);
    if (dtype == TSQR_FLOAT)
        g_opts_f32 = LaunchOpts{path, gate, g_opts_mixed.tol};
    else if (dtype == TSQR_DOUBLE)
        g_opts_f64 = LaunchOpts{path, gate, kTolF64};
    else
        return TSQR_ERR_ARG;
    return TSQR_OK;
}

int tsqr_leaf_stats(const void* work, int m, int n, int dtype, int* gram_leaves, int* householder_leaves,
                    int* flat) {
    if (!work || !gram_leaves || !householder_leaves || !flat || (dtype != TSQR_FLOAT && dtype != TSQR_DOUBLE))
        return TSQR_ERR_ARG;
    if (n < 1 || m < n || n > 32) return TSQR_ERR_SHAPE;
    TreeDesc d;
    const size_t elem = dtype == TSQR_FLOAT ? sizeof(float) : sizeof(double);
    const int path = dtype == TSQR_FLOAT ? current_opts<float>().path : current_opts<double>().path;
    const bool ok = dtype == TSQR_FLOAT ? plan_tsqr<float>(m, n, &d, path) : plan_tsqr<double>(m, n, &d, path);
    if (!ok) return TSQR_ERR_SHAPE;
    *gram_leaves = 0;
    *householder_leaves = 0;
    *flat = 0;
    if (d.lv[0].rows <= d.H) {  // single tile: always Householder
        *householder_leaves = 1;
        return TSQR_OK;
    }
    const WsLayout w = ws_layout(d, elem);
    std::vector<int> flags(d.lv[0].tiles);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(flags.data(), reinterpret_cast<const char*>(work) + w.flags, flags.size() * sizeof(int),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return TSQR_ERR_HIP;
    for (int f : flags) (f ? *householder_leaves : *gram_leaves) += 1;
    if (hipMemcpy(flat, reinterpret_cast<const char*>(work) + w.ctl, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return TSQR_ERR_HIP;
    if (*flat) {  // the whole matrix took a flat Gram path (AUTO or CHOLQR2)
        *gram_leaves = d.lv[0].tiles;
        *householder_leaves = 0;
    }
    return TSQR_OK;
}

int tsqr_set_mixed(double gate, double tol) {
    if (!(gate >= 1.0) || !(tol > 0.0)) return TSQR_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_opts_mu);
    g_opts_mixed.gate = gate;
    g_opts_mixed.tol = tol;
    return TSQR_OK;
}

int tsqr_mixed_stats(const void* work, int m, int n, int* flat, int* passes, int* check, double* residual) {
    if (!work || !flat || !passes || !check || !residual) return TSQR_ERR_ARG;
    if (n < 1 || m < n || n > 32) return TSQR_ERR_SHAPE;
    TreeDesc d;
    if (!plan_tsqr<float>(m, n, &d, mixed_opts().path)) return TSQR_ERR_SHAPE;
    *flat = 0;
    *passes = 0;
    *check = 0;
    *residual = 0.0;
    if (d.lv[0].rows <= d.H) return TSQR_OK;  // single tile: Householder
    const WsLayout w = ws_layout(d, sizeof(float));
    int ctl[3];
    double e[2];
    const char* base = reinterpret_cast<const char*>(work);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(ctl, base + w.ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(e, base + w.side + (size_t)(d.side_leaves + 2) * 1024 * sizeof(double), sizeof(e),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return TSQR_ERR_HIP;
    *flat = ctl[0];
    if (ctl[0]) {
        *passes = ctl[1] ? 2 : 1;
        *check = ctl[2];
        *residual = ctl[1] ? e[1] : e[0];
    }
    return TSQR_OK;
}

int tsqr_mixed_q_forms(const void* work, int m, int n, int* bf16_pass1, int* bf16_pass2) {
    if (!work || !bf16_pass1 || !bf16_pass2) return TSQR_ERR_ARG;
    if (n < 1 || m < n || n > 32) return TSQR_ERR_SHAPE;
    TreeDesc d;
    if (!plan_tsqr<float>(m, n, &d, mixed_opts().path)) return TSQR_ERR_SHAPE;
    *bf16_pass1 = *bf16_pass2 = 0;
    if (d.lv[0].rows <= d.H) return TSQR_OK;  // single tile: Householder
    const WsLayout w = ws_layout(d, sizeof(float));
    int ctl[5];
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(ctl, reinterpret_cast<const char*>(work) + w.ctl, sizeof(ctl), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return TSQR_ERR_HIP;
    if (ctl[0]) {
        *bf16_pass1 = ctl[kCtlBf16Q1] ? 1 : 0;
        *bf16_pass2 = (ctl[1] && ctl[kCtlBf16Q2]) ? 1 : 0;
    }
    return TSQR_OK;
}

int tsqr_set_leaf_tiles(int dtype, int n, int log2m, int tiles) {
    if ((dtype != TSQR_FLOAT && dtype != TSQR_DOUBLE) || n < 1 || n > 32 || log2m < 0 || log2m > 31 || tiles < 0 ||
        tiles > 8)
        return TSQR_ERR_ARG;
    g_leaf_tiles[dtype == TSQR_FLOAT ? 0 : 1][width_class(padded_cols(n))][log2m].store(tiles);
    return TSQR_OK;
}

int tsqr_leaf_tiles(int dtype, int m, int n) {
    if (n < 1 || n > 32 || m < n) return TSQR_ERR_ARG;
    const int N = padded_cols(n);
    if (dtype == TSQR_FLOAT) return leaf_rows_for<float>(m, n) / (64 * rows_per_lane<float>(N));
    if (dtype == TSQR_DOUBLE) return leaf_rows_for<double>(m, n) / (64 * rows_per_lane<double>(N));
    return TSQR_ERR_ARG;
}

int tsqr_set_tree_arity(int dtype, int n, int arity) {
    if ((dtype != TSQR_FLOAT && dtype != TSQR_DOUBLE) || n < 1 || n > 32 || arity < 0 || arity == 1 || arity > 64)
        return TSQR_ERR_ARG;
    g_tree_arity[dtype == TSQR_FLOAT ? 0 : 1][width_class(padded_cols(n))].store(arity);
    return TSQR_OK;
}

int tsqr_tree_arity(int dtype, int m, int n) {
    if (n < 1 || n > 32 || m < n) return TSQR_ERR_ARG;
    TreeDesc d;
    if (dtype == TSQR_FLOAT)
        return plan_tsqr<float>(m, n, &d, current_opts<float>().path) ? d.arity : TSQR_ERR_SHAPE;
    if (dtype == TSQR_DOUBLE)
        return plan_tsqr<double>(m, n, &d, current_opts<double>().path) ? d.arity : TSQR_ERR_SHAPE;
    return TSQR_ERR_ARG;
}

int tsqr_set_stream_shape(int max_groups, int q_waves) {
    if (max_groups < 0 || max_groups > kStreamMaxGroups || q_waves < 0 || q_waves > (1 << 16)) return TSQR_ERR_ARG;
    g_stream_max_groups.store(max_groups ? max_groups : kStreamDefaultGroups);
    g_gramq_waves.store(q_waves ? q_waves : 2048);
    return TSQR_OK;
}

int tsqr_set_gemm_waves(int waves) {
    if (waves < 1 || waves > 1 << 20) return TSQR_ERR_ARG;
    g_tn_waves.store(waves);
    return TSQR_OK;
}

size_t tsqr_workspace_bytes(int m, int n, int dtype) {
    if (dtype == TSQR_FLOAT) return tsqr_ws_bytes<float>(m, n);
    if (dtype == TSQR_DOUBLE) return tsqr_ws_bytes<double>(m, n);
    return 0;
}

size_t qr_workspace_bytes(int m, int n, int dtype) {
    if (dtype == TSQR_FLOAT) return qr_ws_bytes<float>(m, n);
    if (dtype == TSQR_DOUBLE) return qr_ws_bytes<double>(m, n);
    return 0;
}

int tsqr_s_stream(int m, int n, float* A, int lda, float* R, int ldr, void* work, size_t work_bytes, void* stream) {
    return tsqr_stream<float>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream);
}
int tsqr_d_stream(int m, int n, double* A, int lda, double* R, int ldr, void* work, size_t work_bytes,
                  void* stream) {
    return tsqr_stream<double>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream);
}
int tsqr_s_stream_noalloc(int m, int n, float* A, int lda, float* R, int ldr, void* stream) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    if (n > 32) return TSQR_ERR_SHAPE;
    const size_t need = tsqr_ws_bytes<float>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    if (need > kScratchBytes) return TSQR_ERR_SCRATCH;
    void* ws = nullptr;
    e = scratch_address(&ws);
    if (e) return e;
    return tsqr_core<float>(m, n, A, lda, R, ldr, ws, (hipStream_t)stream, false, false, true);
}
int tsqr_mixed_s_stream(int m, int n, float* A, int lda, float* R, int ldr, void* work, size_t work_bytes,
                        void* stream) {
    return tsqr_stream<float>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream, true);
}
int qr_s_stream(int m, int n, float* A, int lda, float* R, int ldr, void* work, size_t work_bytes, void* stream) {
    return qr_stream<float>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream);
}
int qr_d_stream(int m, int n, double* A, int lda, double* R, int ldr, void* work, size_t work_bytes, void* stream) {
    return qr_stream<double>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream);
}
int qr_mixed_s_stream(int m, int n, float* A, int lda, float* R, int ldr, void* work, size_t work_bytes,
                      void* stream) {
    return qr_stream<float>(m, n, A, lda, R, ldr, work, work_bytes, (hipStream_t)stream, true);
}

int tsqr_r_s(int m, int n, const float* A, int lda, float* R, int ldr, float* work, int ldwork) {
    return tsqr_r_ref<float>(m, n, A, lda, R, ldr, work, ldwork);
}
int tsqr_r_d(int m, int n, const double* A, int lda, double* R, int ldr, double* work, int ldwork) {
    return tsqr_r_ref<double>(m, n, A, lda, R, ldr, work, ldwork);
}
int tsqr_r_s_stream(int m, int n, const float* A, int lda, float* R, int ldr, void* work, size_t work_bytes,
                    void* stream) {
    return tsqr_stream<float>(m, n, const_cast<float*>(A), lda, R, ldr, work, work_bytes, (hipStream_t)stream,
                              false, true);
}
int tsqr_r_d_stream(int m, int n, const double* A, int lda, double* R, int ldr, void* work, size_t work_bytes,
                    void* stream) {
    return tsqr_stream<double>(m, n, const_cast<double*>(A), lda, R, ldr, work, work_bytes, (hipStream_t)stream,
                               false, true);
}
int tsqr_s(int m, int n, float* A, int lda, float* R, int ldr, float* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = tsqr_ws_bytes<float>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return tsqr_stream<float>(m, n, A, lda, R, ldr, ws, bytes, nullptr);
    });
}
int tsqr_mixed_s(int m, int n, float* A, int lda, float* R, int ldr, float* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = tsqr_ws_bytes<float>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return tsqr_stream<float>(m, n, A, lda, R, ldr, ws, bytes, nullptr, true);
    });
}
int tsqr_d(int m, int n, double* A, int lda, double* R, int ldr, double* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = tsqr_ws_bytes<double>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return tsqr_stream<double>(m, n, A, lda, R, ldr, ws, bytes, nullptr);
    });
}
int qr_s(int m, int n, float* A, int lda, float* R, int ldr, float* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = qr_ws_bytes<float>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return qr_stream<float>(m, n, A, lda, R, ldr, ws, bytes, nullptr);
    });
}
int qr_mixed_s(int m, int n, float* A, int lda, float* R, int ldr, float* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = qr_ws_bytes<float>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return qr_stream<float>(m, n, A, lda, R, ldr, ws, bytes, nullptr, true);
    });
}
int qr_d(int m, int n, double* A, int lda, double* R, int ldr, double* work, int ldwork) {
    int e = check_args(m, n, A, lda, R, ldr);
    if (e) return e;
    const size_t need = qr_ws_bytes<double>(m, n);
    if (need == 0) return TSQR_ERR_SHAPE;
    return with_ref_workspace(n, work, ldwork, need, [&](void* ws, size_t bytes) {
        return qr_stream<double>(m, n, A, lda, R, ldr, ws, bytes, nullptr);
    });
}

}  // extern "C"
