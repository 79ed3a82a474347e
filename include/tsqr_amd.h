/*
 * tsqr_amd.h -- C ABI of the MI355X (gfx950) TSQR / recursive blocked QR
 * library (libtsqr_amd.so).
 *
 * Every entry point replaces one template of the reference library
 * dajunHuang/tsqr, which is header-only C++ with no extern "C" symbols:
 *
 *   template <typename T>
 *   void tsqr(int m, int n, T *A, int lda, T *R, int ldr, T *d_work,
 *             int ldwork);                         // TallShinnyQR.h:86-87
 *   template <typename T>
 *   void qr(int m, int n, T *A, int lda, T *R, int ldr, T *d_work,
 *           int ldwork);                           // block_QR.h:72-73
 *
 * instantiated for float and double (TallShinnyQR.h:150-153,
 * block_QR.h:96-99).  Semantics kept from the reference:
 *   - A, R and the work array are device pointers owned by the caller;
 *   - A is m x n, column-major, leading dimension lda, and is overwritten by
 *     the explicit thin Q (same lda);
 *   - R is n x n, column-major, leading dimension ldr, written in full with
 *     zeros below the diagonal (the reference's qr leaves the lower-left
 *     block unwritten; here it is zeroed);
 *   - the call is asynchronous with respect to the host; the plain entry
 *     points run on the null stream, like the reference's launches.
 * Differences (documented, deliberate):
 *   - every call returns a status code instead of assert()/printf()/throw
 *     (TallShinnyQR.h:92-103, utils.h:47-55);
 *   - the accepted domain is m >= n >= 1 for any m (the reference needs
 *     n | m and caps m at (NUM_SM*BS/n)*NUM_SM*BS, TallShinnyQR.h:100);
 *     tsqr_* takes n <= 32 (the reference's effective limit, kernelQR.h:7);
 *     qr_* takes any n <= m;
 *   - the work array of the reference-shaped entry points holds ldwork*n
 *     elements; if that is smaller than tsqr_workspace_bytes(), the library
 *     uses a per-device internal buffer instead; a call holds that buffer's
 *     lock until its launches are enqueued on the null stream, so calls
 *     from several host threads take turns on it (a large enough work
 *     array, or the *_stream entry points, avoid the lock);
 *   - a zero column yields a zero R diagonal entry and an identity reflector
 *     (LAPACK dlarfg convention) where the reference divides by zero.
 */
#ifndef TSQR_AMD_H
#define TSQR_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TSQR_OK 0
#define TSQR_ERR_ARG (-1)       /* null pointer, m < n, n < 1, lda < m, ldr < n */
#define TSQR_ERR_SHAPE (-2)     /* shape outside what the entry point supports */
#define TSQR_ERR_WORKSPACE (-3) /* explicit workspace smaller than required */
#define TSQR_ERR_HIP (-4)       /* a HIP runtime call or kernel launch failed */
#define TSQR_ERR_ALLOC (-5)     /* internal workspace allocation failed */
#define TSQR_ERR_SCRATCH (-6)   /* tsqr_s_stream_noalloc: workspace above its fixed 16 MiB buffer */

/* dtype codes follow the reference harness CLI (test_tsqr.cu:118-128). */
#define TSQR_FLOAT 1
#define TSQR_DOUBLE 2

/* Level-0 leaf paths (tsqr_set_leaf_path). */
#define TSQR_LEAF_AUTO 0        /* flat Gram path if the whole matrix passes the gate, else Householder */
#define TSQR_LEAF_HOUSEHOLDER 1 /* Householder leaves and tree only */
#define TSQR_LEAF_GRAM_TREE 2   /* Gram leaves (Householder where rejected), Householder tree */
#define TSQR_LEAF_CHOLQR2 3     /* flat CholeskyQR + residual check + optional second pass, else tree */

/* Human-readable text for a status code. */
const char *tsqr_status_string(int status);

/* Bytes of device workspace tsqr_* needs for an m x n problem of the given
 * dtype (0 for an invalid shape).  Replaces the caller-side sizing
 * ldwork = 2*NUM_SM*BLOCK_SIZE (test_tsqr.cu:42). */
size_t tsqr_workspace_bytes(int m, int n, int dtype);

/* Same for the recursive qr_* entry points. */
size_t qr_workspace_bytes(int m, int n, int dtype);

/* Tall-skinny Householder QR, replaces tsqr<float>/tsqr<double>
 * (TallShinnyQR.h:86-149).  work holds ldwork*n elements. */
int tsqr_s(int m, int n, float *A, int lda, float *R, int ldr, float *work, int ldwork);
int tsqr_d(int m, int n, double *A, int lda, double *R, int ldr, double *work, int ldwork);

/* Stream-ordered variants with an explicit byte-sized workspace; stream is a
 * hipStream_t (NULL = null stream).  Return TSQR_ERR_WORKSPACE instead of
 * falling back to the internal buffer. */
int tsqr_s_stream(int m, int n, float *A, int lda, float *R, int ldr, void *work, size_t work_bytes,
                  void *stream);
int tsqr_d_stream(int m, int n, double *A, int lda, double *R, int ldr, void *work, size_t work_bytes,
                  void *stream);

/* tsqr_s_stream without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  The
 * workspace is a fixed 16 MiB array of the library in the memory of each GPU
 * it runs on (held until the process ends), so the call allocates nothing
 * and calls no allocating routine; problems whose tsqr_workspace_bytes(m, n,
 * TSQR_FLOAT) exceeds 16 MiB are rejected with TSQR_ERR_SCRATCH (n = 32:
 * up to about 200,000 rows).  Same kernels and results as tsqr_s_stream.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several streams or host threads at once) race and must be ordered by
 * the caller.  The first call on each device looks up the array's device
 * address (hipGetSymbolAddress), so make it outside graph capture. */
int tsqr_s_stream_noalloc(int m, int n, float *A, int lda, float *R, int ldr, void *stream);

/* Recursive blocked QR, replaces qr<float>/qr<double> (block_QR.h:72-95):
 * n <= 32 is one tsqr; otherwise the columns are split in two, the halves
 * are coupled by R12 = Q1^T A2 and A2 -= Q1 R12, and each half recurses. */
int qr_s(int m, int n, float *A, int lda, float *R, int ldr, float *work, int ldwork);
int qr_d(int m, int n, double *A, int lda, double *R, int ldr, double *work, int ldwork);
int qr_s_stream(int m, int n, float *A, int lda, float *R, int ldr, void *work, size_t work_bytes, void *stream);
int qr_d_stream(int m, int n, double *A, int lda, double *R, int ldr, void *work, size_t work_bytes,
                void *stream);

/* Mixed-precision TSQR (FP32 in and out, n <= 32), the tsqr<float> contract
 * (TallShinnyQR.h:86-149) with a different method; no counterpart in the
 * reference.  The Gram matrices are formed on the BF16 matrix cores with
 * exact three-way BF16 splits of the FP32 data (six products per pair, FP32
 * accumulation over 128 rows, FP64 beyond), R = chol(A^T A) and
 * Q = A R^{-1} (FP64 products, one rounding), followed by an FP32 residual
 * check ||Q^T Q - I||_F <= tol computed in the same pass; if the check fails
 * a second CholeskyQR pass (CholeskyQR2) re-orthogonalises Q and updates R.
 * If the Cholesky breaks down or ||R||_F ||R^{-1}||_F exceeds the gate, the
 * call runs the Householder tree instead.  Same workspace as tsqr_s. */
int tsqr_mixed_s(int m, int n, float *A, int lda, float *R, int ldr, float *work, int ldwork);
int tsqr_mixed_s_stream(int m, int n, float *A, int lda, float *R, int ldr, void *work, size_t work_bytes,
                        void *stream);

/* Mixed-precision recursive blocked QR (FP32 in and out), the qr<float>
 * contract (block_QR.h:72-95) with a different method; no counterpart in the
 * reference.  The same column split as qr_s; every panel of <= 32 columns is
 * tsqr_mixed_s, and the coupling of the column blocks is applied twice
 * (block classical Gram-Schmidt with reorthogonalisation): R12 = Q1^T A2,
 * A2 -= Q1 R12, then C = Q1^T A2, A2 -= Q1 C, R12 += C (BF16x3 matrix-core
 * products, FP64 sums for Q1^T A2).  For 32 < n <= 128 a final CholeskyQR
 * pass over the whole Q follows: G = Q^T Q (FP64 sums), and if
 * ||G - I||_F exceeds the mixed tolerance, Q <- Q chol(G)^{-1} and
 * R <- chol(G) R; this removes the loss of orthogonality between column
 * blocks that grows with cond(R22) times the coupling products' error, so
 * ||I - Q^T Q||_F stays below the tolerance independently of cond(A).
 * Workspace: qr_workspace_bytes(m, n, TSQR_FLOAT). */
int qr_mixed_s(int m, int n, float *A, int lda, float *R, int ldr, float *work, int ldwork);
int qr_mixed_s_stream(int m, int n, float *A, int lda, float *R, int ldr, void *work, size_t work_bytes,
                      void *stream);

/* Gate (>= 1, default 1024) and residual tolerance (> 0, default 8e-7) of
 * the mixed entry points, process-wide.  The gate is the largest
 * ||R||_F ||R^{-1}||_F the BF16x3 Gram pass accepts; above it the FP64 Gram
 * pass accepts up to max(gate, 1e7), or up to the gate itself when the gate
 * is below its default (a gate of 1 forces the Householder tree). */
int tsqr_set_mixed(double gate, double tol);

/* After a tsqr_mixed_s* call on the same stream, with its workspace: *flat =
 * 1 if the CholeskyQR path ran (0: Householder tree), *passes = 1 or 2,
 * *check = 0 if the final residual met tol (1: it did not after the second
 * pass, 2: the second Cholesky broke down), *residual = the final
 * ||Q^T Q - I||_F estimate.  Synchronises the device.  Diagnostic. */
int tsqr_mixed_stats(const void *work, int m, int n, int *flat, int *passes, int *check, double *residual);

/* After a tsqr_mixed_s* call on the same stream, with its workspace: which
 * form the Q passes took (no counterpart in the reference).  *bf16_pass1 = 1
 * if pass 1 formed Q1 = A R1^{-1} with BF16x3 products on the BF16 matrix
 * cores (R1 well conditioned: kappa_est <= 2n), 0 if with FP64 products;
 * *bf16_pass2 likewise for the second pass (0 when it did not run).
 * Synchronises the device.  Diagnostic. */
int tsqr_mixed_q_forms(const void *work, int m, int n, int *bf16_pass1, int *bf16_pass2);

/* Batched TSQR of many independent matrices of uneven shape (no counterpart
 * in the reference, which factors one matrix per call).  Matrix i is m[i] x
 * n[i] (1 <= n[i] <= 64, m[i] >= n[i]) column-major at the device pointer
 * A[i] (leading dimension lda[i]) and is overwritten by its explicit Q; its
 * n[i] x n[i] R goes to R[i] (ldr[i]).  n[i] <= 32 is one TSQR; for
 * 32 < n[i] <= 64 the columns are split at 32 and coupled by one block
 * Gram-Schmidt step (R12 = Q1^T A2, A2 -= Q1 R12, FP64 MFMA products), as
 * the recursive qr does (block_QR.h:72-95).
 * method TSQR_BATCH_PLAIN: every panel is the Householder TSQR (the
 *        reference's method), in the working precision (float or double);
 * method TSQR_BATCH_MIXED (float only): every panel is the mixed-precision
 *        TSQR of tsqr_mixed_s (BF16x3 CholeskyQR, residual check, optional
 *        second pass, Householder if the gate rejects) and the coupling step
 *        is applied twice (BCGS2).
 * tsqr_batch_create validates the shapes, plans the trees and the device
 * work lists of (matrix, row block) items once and copies them to device
 * memory it owns; the pointers are fixed for the life of the plan.
 * tsqr_batch_run enqueues the whole batch on `stream` (a few launches, one
 * per phase, each over every matrix; no host synchronisation, no memset),
 * tsqr_batch_graph_launch captures the same launches into a hipGraph on
 * its first call for a given workspace and replays the graph on `stream`.
 * Calls on one plan must not overlap.  work: tsqr_batch_workspace_bytes. */
#define TSQR_BATCH_PLAIN 0
#define TSQR_BATCH_MIXED 1
typedef struct tsqr_batch tsqr_batch;
int tsqr_batch_create(int dtype, int method, int count, const int *m, const int *n, void *const *A, const int *lda,
                      void *const *R, const int *ldr, tsqr_batch **plan);
size_t tsqr_batch_workspace_bytes(const tsqr_batch *plan);
int tsqr_batch_run(tsqr_batch *plan, void *work, size_t work_bytes, void *stream);
int tsqr_batch_graph_launch(tsqr_batch *plan, void *work, size_t work_bytes, void *stream);
/* Gate and residual tolerance of a mixed plan (defaults 1024 and 8e-7). */
int tsqr_batch_set_mixed(tsqr_batch *plan, double gate, double tol);
/* After a run (synchronises the device): panels that took the CholeskyQR
 * path, panels that ran Householder, panels that needed the second pass and
 * panels whose final residual check failed. */
int tsqr_batch_stats(const tsqr_batch *plan, const void *work, int *cholesky, int *householder, int *second_pass,
                     int *check_failed);
/* After a mixed run (synchronises the device): panels whose first Q pass ran
 * on the BF16 matrix cores (BF16x3 products), and panels whose second pass
 * did. */
int tsqr_batch_q_forms(const tsqr_batch *plan, const void *work, int *bf16_pass1, int *bf16_pass2);
int tsqr_batch_destroy(tsqr_batch *plan);

/* Chooses how level-0 superleaves (up to 8 Householder tiles of rows each)
 * are factored for one dtype, process-wide.  No counterpart in the
 * reference, whose tiles are always Householder (kernelQR.h:69-291).
 *   TSQR_LEAF_HOUSEHOLDER  every tile is a Householder QR and Q is formed
 *                          from the stored reflectors (the reference method);
 *   TSQR_LEAF_GRAM_TREE    a superleaf's R comes from the Cholesky factor of
 *                          its FP64-accumulated Gram matrix and its Q rows
 *                          from A (R^{-1} S) in FP64, unless the Cholesky
 *                          breaks down or ||R||_F ||R^{-1}||_F exceeds gate,
 *                          in which case that superleaf is Householder; the
 *                          tree above the superleaves is Householder;
 *   TSQR_LEAF_AUTO         the Gram matrix of the whole matrix (FP64
 *                          accumulation); if it passes the gate, R is its
 *                          Cholesky factor and Q = A R^{-1} (FP64), with no
 *                          tree (the same error bounds), else every
 *                          superleaf is Householder (over a single
 *                          superleaf: as GRAM_TREE);
 *   TSQR_LEAF_CHOLQR2      the whole matrix: R1 = chol(A^T A), Q1 = A R1^{-1}
 *                          with ||Q1^T Q1 - I||_F measured in the same pass;
 *                          above the tolerance a second CholeskyQR pass runs
 *                          (CholeskyQR2); if the Cholesky breaks down or
 *                          ||R||_F ||R^{-1}||_F exceeds the gate, the
 *                          Householder tree.  FP32: the BF16x3 Gram of the
 *                          mixed entry points, tolerance 8e-7; FP64: FP64
 *                          Gram, tolerance 4*32*u64.
 * Defaults: FP32 TSQR_LEAF_AUTO with gate 1024, FP64 TSQR_LEAF_CHOLQR2 with
 * gate 1e5 (u64*kappa^2 small enough for the second pass to restore
 * orthogonality).  Below the gate the Gram paths' backward error and loss of
 * orthogonality stay at the working roundoff.  Calls already enqueued are not
 * affected. */
int tsqr_set_leaf_path(int dtype, int path, double gate);

/* Tuning table, replacing the reference's compile-time NUM_SM / BLOCK_SIZE
 * (kernelQR.h:3-11); written by the autotuner (tsqr_amd/autotune.py), and
 * callable directly.  Process-wide; a change affects calls planned after it,
 * including the workspace size, so query tsqr_workspace_bytes afterwards.
 *   tsqr_set_leaf_tiles  tiles per level-0 superleaf (1..8, clipped to what
 *                        the tile shape allows; 0 = the built-in rule) for
 *                        matrices of this dtype and padded width of n
 *                        (8, 16 or 32) with floor(log2 m) == log2m;
 *   tsqr_leaf_tiles      the tiles per superleaf a call would use now;
 *   tsqr_set_gemm_waves  waves the split-K R12 = Q1^T A2 launch of the
 *                        recursive QR aims for (default 512). */
int tsqr_set_leaf_tiles(int dtype, int n, int log2m, int tiles);
int tsqr_leaf_tiles(int dtype, int m, int n);
int tsqr_set_gemm_waves(int waves);

/* More knobs of the same table (no counterpart in the reference, whose tree
 * fan-in is fixed by BLOCK_SIZE / n, kernelQR.h:3-11, 321-388):
 *   tsqr_set_tree_arity    children per tree node for this dtype and padded
 *                          width (2 .. H/n; 0 = H/n, full H-row blocks,
 *                          the default); fewer children give a deeper tree
 *                          of smaller nodes;
 *   tsqr_tree_arity        the arity a call would use now (after clipping);
 *   tsqr_set_stream_shape  largest grid of the streaming Gram passes (1 ..
 *                          1024 workgroups of 4 waves; 0 = 256, one per
 *                          CU, the fastest measured on MI355X) and the
 *                          waves the Gram Q pass aims for when superleaves
 *                          are few (0 = 2048).
 * Process-wide, for calls planned afterwards (workspace sizes included). */
int tsqr_set_tree_arity(int dtype, int n, int arity);
int tsqr_tree_arity(int dtype, int m, int n);
int tsqr_set_stream_shape(int max_groups, int q_waves);

/* After a tsqr_*_stream call that used `work` has been enqueued: waits for
 * the device, counts the level-0 superleaves that took the Gram path and the
 * Householder path, and sets *flat to 1 if the call took the flat Gram path.
 * Diagnostic; no counterpart in the reference. */
int tsqr_leaf_stats(const void *work, int m, int n, int dtype, int *gram_leaves, int *householder_leaves,
                    int *flat);

/* R only (no counterpart in the reference, whose tsqr always overwrites A
 * with Q; TallShinnyQR.h:86-149 is the call it otherwise mirrors).  Writes
 * R (n x n, ldr, zeros below the diagonal, the same R as tsqr_s / tsqr_d up
 * to rounding) and never writes A: the Q pass and the down sweep are
 * skipped, so A is read once on the flat Gram path (twice for FP64, whose
 * second CholeskyQR pass streams the Gram matrix of A R1^-1 without
 * storing it).  Workspace as tsqr_s / tsqr_d (tsqr_workspace_bytes). */
int tsqr_r_s(int m, int n, const float *A, int lda, float *R, int ldr, float *work, int ldwork);
int tsqr_r_d(int m, int n, const double *A, int lda, double *R, int ldr, double *work, int ldwork);
int tsqr_r_s_stream(int m, int n, const float *A, int lda, float *R, int ldr, void *work, size_t work_bytes,
                    void *stream);
int tsqr_r_d_stream(int m, int n, const double *A, int lda, double *R, int ldr, void *work, size_t work_bytes,
                    void *stream);

/* Least squares and Q^T B on the implicit Q of the R-only mode (no
 * counterpart in the reference; the most common use of a tall-skinny R).
 * n <= 32, m >= n, A of full column rank; every pointer on the device.
 *
 * tsqr_lstsq_*: X (n x nrhs, ldx) = argmin ||A X - B||_F by the corrected
 * seminormal equations (Bjorck): R from tsqr_r_*, X = R^-1 R^-T A^T B, then
 * one refinement step with the residual.  A is read only; B (m x nrhs) is
 * overwritten by the final residual B - A X.
 * tsqr_apply_qt_*: C (n x k, ldc) = Q^T B with Q = A R^-1 (R from tsqr_r_*
 * on the same A).
 * Both take work of tsqr_solve_workspace_bytes(m, n, nrhs or k, dtype)
 * bytes and run on `stream`. */
size_t tsqr_solve_workspace_bytes(int m, int n, int nrhs, int dtype);
int tsqr_lstsq_s_stream(int m, int n, int nrhs, const float *A, int lda, float *B, int ldb, float *X, int ldx,
                        void *work, size_t work_bytes, void *stream);
int tsqr_lstsq_d_stream(int m, int n, int nrhs, const double *A, int lda, double *B, int ldb, double *X, int ldx,
                        void *work, size_t work_bytes, void *stream);
int tsqr_apply_qt_s_stream(int m, int n, int k, const float *A, int lda, const float *R, int ldr, const float *B,
                           int ldb, float *C, int ldc, void *work, size_t work_bytes, void *stream);
int tsqr_apply_qt_d_stream(int m, int n, int k, const double *A, int lda, const double *R, int ldr,
                           const double *B, int ldb, double *C, int ldc, void *work, size_t work_bytes, void *stream);

/* Device-side FP64 checker of a factorisation, replaces check_QR_accuracy
 * (utils.h:317-455), whose m x m gesvd workspace limits it to m <= ~46k.
 * A is the input (m x n, lda), Q and R the computed factors (ldq, ldr),
 * n <= 32.  Three streaming FP64 reductions on the GPU form the n x n Gram
 * matrices A^T A, Q^T Q and E^T E with E = A - QR (formed in FP64 on the
 * matrix cores); the host then takes their traces and symmetric
 * eigenvalues.  Writes
 *   norms[0] = ||A - QR||_F / ||A||_F   norms[1] = ||I - Q^T Q||_F
 *   norms[2] = ||A - QR||_2 / ||A||_2   norms[3] = ||I - Q^T Q||_2
 * (the reference prints norms[2] and norms[3] / n, utils.h:380-381).
 * Runs on `stream` (0 = default), uses an internal per-device workspace
 * and waits for its own work before returning. */
int tsqr_check_s(int m, int n, const float *A, int lda, const float *Q, int ldq, const float *R, int ldr,
                 double *norms, void *stream);
int tsqr_check_d(int m, int n, const double *A, int lda, const double *Q, int ldq, const double *R, int ldr,
                 double *norms, void *stream);

/* Per-phase timing (no counterpart in the reference, which only times whole
 * calls with CUDA events, test_tsqr.cu:62-90).  While enabled, every call
 * brackets each launch phase with HIP events on its stream:
 *   phase 0 tsqr leaf R pass (or the whole single-tile call),
 *   1 tsqr tree walk up, 2 tsqr down sweep and leaf M products,
 *   3 tsqr leaf Q pass, 4 qr R12 = Q1^T A2, 5 qr A2 -= Q1 R12.
 * tsqr_timing_read waits for the recorded events, writes the summed
 * milliseconds and the number of intervals per phase (nphases <= 6 entries)
 * and forgets them.  Enabling or disabling also forgets them.  Meant for
 * benchmarks; it serialises nothing while recording. */
int tsqr_timing_enable(int enable);
int tsqr_timing_read(double *ms, int *count, int nphases);

#ifdef __cplusplus
}
#endif

#endif /* TSQR_AMD_H */
