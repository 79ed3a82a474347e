// gram.hpp -- the Gram (CholeskyQR) path for well-conditioned superleaves.
//
// A superleaf is a block of between one and super_tiles tiles of H
// consecutive rows of A (H = tile height of the Householder kernels,
// super_tiles = min(8, H/N); the height is chosen from m by leaf_rows_for).
// For such a block A_b the R factor is unique up to
// the signs of its rows, so it can be obtained either by Householder
// reflections (super.hpp) or from the Gram matrix:
//   G_b = A_b^T A_b   accumulated in FP64 on the matrix cores
//                     (v_mfma_f64_16x16x4f64; products of FP32 inputs are
//                     exact in FP64),
//   R_b = chol(G_b)   FP64, one wavefront, positive diagonal.
// With FP64 accumulation the relative error of R_b is O(u64 * kappa(A_b)^2),
// below FP32 roundoff while kappa(A_b) <= ~1e4.  The tree receives R_b
// rounded to the working precision (R~_b); the Q side is
//   Q_b = A_b * (R_b^{-1} S_b)
// with S_b the n x n block of the tree's Q for this leaf, accumulated in FP64
// and rounded once.  A_b R_b^{-1} is orthonormal to O(u64 kappa^2), and the
// residual A_b - (A_b R_b^{-1}) R~_b = (A_b R_b^{-1})(R_b - R~_b) is at the
// working roundoff, so neither error grows with kappa in that range.  Each
// superleaf computes
//   kappa_est = ||R_b||_F * ||R_b^{-1}||_F  (an upper bound of kappa_2(A_b))
// and takes the Householder path instead (super.hpp) when the Cholesky
// breaks down or kappa_est exceeds the gate.
//
// MFMA operand layout (16x16x4 f64): lane l feeds A[i = l%16][k = l/16] and
// B[k = l/16][j = l%16]; the accumulator register v of lane l holds
// C[4v + l/16][l%16].  A lane loads 4 consecutive rows of one column with one
// 16-byte load; MFMA number t (t = 0..3) takes row 4*(l/16)+t as its k index,
// and the four MFMAs together cover 16 rows.
#pragma once
#include "tsqr_kernels.hpp"

namespace tsqr {

typedef double v4d __attribute__((ext_vector_type(4)));

// LDS of one Gram wave: the 32 x 32 Gram / R matrix (row-major, padded) and
// one broadcast row.
struct GramLds {
    double G[32][33];
    double row[32];
    double rdiag[32];  // 1 / R_jj
};

// Four consecutive rows of one column in the storage type.
template <typename T>
struct Raw4 {
    T v[4];
};

// Rows [r0, r0+4) of column c of the (rows x ncols) block at base, zero
// outside; the caller converts to double when the value is consumed.
template <typename T, bool VEC>
__device__ __forceinline__ Raw4<T> load_raw4(const T* __restrict__ base, int64_t ld, int r0, int c, int rows,
                                             int ncols) {
    Raw4<T> o;
    if (c < ncols) {
        const T* p = base + (int64_t)c * ld + r0;
        if (VEC && r0 + 4 <= rows) {
            if constexpr (sizeof(T) == 4) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                const f4 q = *reinterpret_cast<const f4*>(p);
#pragma unroll
                for (int t = 0; t < 4; ++t) o.v[t] = q[t];
            } else {
                typedef double d2 __attribute__((ext_vector_type(2)));
                const d2 q0 = *reinterpret_cast<const d2*>(p);
                const d2 q1 = *reinterpret_cast<const d2*>(p + 2);
                o.v[0] = q0[0];
                o.v[1] = q0[1];
                o.v[2] = q1[0];
                o.v[3] = q1[1];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) o.v[t] = (r0 + t < rows) ? p[t] : T(0);
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) o.v[t] = T(0);
    }
    return o;
}

// Loads rows [r0, r0+4) of column c of the (rows x ncols) block at base into
// v (as double), zero outside.
template <typename T, bool VEC>
__device__ __forceinline__ void load4(double (&v)[4], const T* __restrict__ base, int64_t ld, int r0, int c,
                                      int rows, int ncols) {
    if (c < ncols) {
        const T* p = base + (int64_t)c * ld + r0;
        if (VEC && r0 + 4 <= rows) {
            if constexpr (sizeof(T) == 4) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                const f4 q = *reinterpret_cast<const f4*>(p);
#pragma unroll
                for (int t = 0; t < 4; ++t) v[t] = (double)q[t];
            } else {
                typedef double d2 __attribute__((ext_vector_type(2)));
                const d2 q0 = *reinterpret_cast<const d2*>(p);
                const d2 q1 = *reinterpret_cast<const d2*>(p + 2);
                v[0] = q0[0];
                v[1] = q0[1];
                v[2] = q1[0];
                v[3] = q1[1];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = (r0 + t < rows) ? (double)p[t] : 0.0;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = 0.0;
    }
}

// Cholesky of L.G (n x n, n <= 32): on return lane c of each half-wave
// holds column c of R (R^T R = G, positive diagonal) in g[].  The two
// half-waves (lanes 0-31, 32-63) factor two independent matrices, each in
// its own GramLds.  Columns >= n are treated as an identity block.  Returns
// false if a pivot is not positive and finite.
__device__ __forceinline__ bool chol_wave(GramLds& L, int n, int lane, double (&g)[32]) {
    const int c = lane & 31;
    const bool hi = lane >= 32;
#pragma unroll
    for (int r = 0; r < 32; ++r) g[r] = (r < n && c < n) ? L.G[r][c] : (r == c ? 1.0 : 0.0);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const double p_lo = readlane_t<double>(g[j], j);
        const double p_hi = readlane_t<double>(g[j], 32 + j);
        double piv = hi ? p_hi : p_lo;
        if (!(piv > 0.0) || !(piv < __builtin_inf())) {
            ok = false;
            piv = 1.0;
        }
        const double rd = 1.0 / __builtin_sqrt(piv);
        // row j of R: R[j][c] = G[j][c] / sqrt(G[j][j]) for c >= j
        g[j] = (c >= j) ? g[j] * rd : 0.0;
        L.row[c] = g[j];
        if (c == j) L.rdiag[j] = rd;
        wave_sync();
        // G[r][c] -= R[j][r] R[j][c] for j < r <= c
#pragma unroll
        for (int r = j + 1; r < 32; ++r) g[r] = (r <= c) ? fmat(-L.row[r], g[j], g[r]) : g[r];
        wave_sync();
    }
    return ok;
}

// Copies R (lane c of a half-wave holds column c in g[]) to L.G, row-major.
__device__ __forceinline__ void r_to_lds(const double (&g)[32], GramLds& L, int lane) {
    const int c = lane & 31;
#pragma unroll
    for (int r = 0; r < 32; ++r) L.G[r][c] = g[r];
    wave_sync();
}

// Inverse of the upper triangular R in L.G (row-major), with 1/R_ii in
// L.rdiag: lane c back-substitutes R x = e_c, so x[] is column c of R^{-1}.
// The dot products run as four interleaved partial sums.
__device__ __forceinline__ void tri_inv_wave(const GramLds& L, int lane, double (&x)[32]) {
    const int c = lane & 31;
#pragma unroll
    for (int i = 31; i >= 0; --i) {
        double p[4] = {(i == c) ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = i + 1; k < 32; ++k) p[(k - i) & 3] = fmat(-L.G[i][k], x[k], p[(k - i) & 3]);
        x[i] = (i <= c) ? ((p[0] + p[1]) + (p[2] + p[3])) * L.rdiag[i] : 0.0;
    }
}

__device__ __forceinline__ int leaf_rows_of(const LevelDesc& L, int leaf) {
    const int64_t left = L.rows - (int64_t)leaf * L.tile_rows;
    return (int)(left < L.tile_rows ? left : L.tile_rows);
}

// Sum over the 32 lanes of each half-wave (lane bits 4..0).
__device__ __forceinline__ double sum_half(double v) {
    v = fold_all<1>(v);
    v = fold_all<2>(v);
    v = fold_all<3>(v);
    v = fold_all<4>(v);
    return fold_all<5>(v);
}

// Pass 1a of the Gram path, one wavefront per superleaf: G = A_b^T A_b in
// FP64 (MFMA), written to side[leaf] (32 x 32, row-major).  A streaming
// kernel: the only state is the 12 accumulator registers and the slabs in
// flight.  A grid-stride loop over the superleaves; does nothing if skip[0]
// is set (the flat path succeeded).
template <typename T, int NT, bool VEC>
__device__ __forceinline__ void gram_g_leaf(const TreeDesc& d, const T* __restrict__ A, int lda,
                                            double* __restrict__ side, int leaf) {
    const int lane = threadIdx.x;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int rows = leaf_rows_of(L0, leaf);
    const T* base = A + (int64_t)leaf * L0.tile_rows;
    v4d acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
    const int col = lane & 15;
    const int rq = 4 * (lane >> 4);
    constexpr int D = 4;  // slabs of 16 rows per batch of loads
    int s0 = 0;
    if (VEC && rows >= D * 16) {
        // Batches of D*16 rows, double buffered: the loads of the next batch
        // are in flight while the MFMAs consume the current one.
        Raw4<T> c0[D], c1[D];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            c0[q] = load_raw4<T, true>(base, lda, 16 * q + rq, col, rows, n);
            if constexpr (NT == 2) c1[q] = load_raw4<T, true>(base, lda, 16 * q + rq, col + 16, rows, n);
        }
#pragma unroll 1
        for (;;) {
            const int next = s0 + D * 16;
            const bool more = next + D * 16 <= rows;
            Raw4<T> p0[D], p1[D];
            if (more) {
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    p0[q] = load_raw4<T, true>(base, lda, next + 16 * q + rq, col, rows, n);
                    if constexpr (NT == 2)
                        p1[q] = load_raw4<T, true>(base, lda, next + 16 * q + rq, col + 16, rows, n);
                }
            }
#pragma unroll
            for (int q = 0; q < D; ++q) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const double x0 = (double)c0[q].v[t];
                    acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, x0, acc00, 0, 0, 0);
                    if constexpr (NT == 2) {
                        const double x1 = (double)c1[q].v[t];
                        acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, x1, acc01, 0, 0, 0);
                        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, x1, acc11, 0, 0, 0);
                    }
                }
            }
            s0 = next;
            if (!more) break;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                c0[q] = p0[q];
                if constexpr (NT == 2) c1[q] = p1[q];
            }
        }
    }
#pragma unroll 1
    for (; s0 < rows; s0 += 16) {
        double a0[4], a1[4];
        load4<T, VEC>(a0, base, lda, s0 + rq, col, rows, n);
        if constexpr (NT == 2) load4<T, VEC>(a1, base, lda, s0 + rq, col + 16, rows, n);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], a0[t], acc00, 0, 0, 0);
            if constexpr (NT == 2) {
                acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], a1[t], acc01, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], a1[t], acc11, 0, 0, 0);
            }
        }
    }
    // G row-major 32 x 32; lane l, register v holds row 4v + l/16, column l%16
    double* G = side + (int64_t)leaf * 1024;
    const int i0 = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = 4 * v + i0;
        G[i * 32 + col] = acc00[v];
        G[i * 32 + 16 + col] = NT == 2 ? acc01[v] : 0.0;
        G[(16 + col) * 32 + i] = NT == 2 ? acc01[v] : 0.0;
        G[(16 + i) * 32 + 16 + col] = NT == 2 ? acc11[v] : 0.0;
    }
}

template <typename T, int NT, bool VEC>
__global__ __launch_bounds__(64) void tsqr_gram_g(const TreeDesc d, const T* __restrict__ A, int lda,
                                                   double* __restrict__ side, const int* __restrict__ skip) {
    if (skip[0]) return;  // the flat path already has R and R^{-1}
#pragma unroll 1
    for (int leaf = blockIdx.x; leaf < d.lv[0].tiles; leaf += gridDim.x) gram_g_leaf<T, NT, VEC>(d, A, lda, side, leaf);
}

// Pass 1b, two superleaves per wavefront (one per half-wave), a grid-stride
// loop over the superleaf pairs:
// R = chol(G), X = R^{-1}, kappa_est = ||R||_F ||R^{-1}||_F.
// Success: R rounded to T goes to the superleaf's slot in its parent block
// (or to the caller's R if the superleaf is the root) and X (FP64, 32 x 32,
// row-major) replaces G in side[leaf]; flags[leaf] stays 0.
// Failure (pivot not positive, or kappa_est > gate): flags[leaf] = 1, the
// superleaf is appended to list (list[0] counts them, in no particular
// order), and nothing else is written; the Householder superleaf kernels
// take over.
template <typename T>
__device__ __forceinline__ void gram_chol_pair(const TreeDesc& d, T* __restrict__ R, int ldr, T* __restrict__ blocks,
                                               double* __restrict__ side, int* __restrict__ flags,
                                               int* __restrict__ list, double gate, GramLds (&Lh)[2], int pair) {
    const int lane = threadIdx.x;
    const int half = lane >> 5;
    const int c = lane & 31;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int leaf = 2 * pair + half;
    const bool real = leaf < L0.tiles;
    GramLds& L = Lh[half];
    double* G = side + (int64_t)(real ? leaf : 0) * 1024;
#pragma unroll 8
    for (int r = 0; r < 32; ++r) L.G[r][c] = real ? G[r * 32 + c] : (r == c ? 1.0 : 0.0);
    wave_sync();
    double nr;
    bool ok;
    {
        double g[32];
        ok = chol_wave(L, n, lane, g);
        nr = 0.0;
#pragma unroll
        for (int r = 0; r < 32; ++r) nr = fmat(g[r], g[r], nr);
        r_to_lds(g, L, lane);
    }
    double x[32];
    tri_inv_wave(L, lane, x);
    double nx = 0.0;
#pragma unroll
    for (int r = 0; r < 32; ++r) nx = fmat(x[r], x[r], nx);
    if (c >= n) {
        nr = 0.0;
        nx = 0.0;
    }
    const double kest = __builtin_sqrt(sum_half(nr) * sum_half(nx));
    ok = ok && (kest <= gate);  // false for NaN too
    if (!real) return;
    if (!ok) {
        if (c == 0) {
            flags[leaf] = 1;
            const int at = __hip_atomic_fetch_add(&list[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            list[1 + at] = leaf;
        }
        return;
    }
    T* dst;
    int64_t ld;
    if (L0.tiles == 1) {
        dst = R;
        ld = ldr;
    } else {
        const int parent = leaf / d.arity;
        const int slot = leaf - parent * d.arity;
        dst = blocks + d.lv[1].blk_off + (int64_t)parent * d.H * n + slot * n;
        ld = d.H;
    }
    if (c < n) {
#pragma unroll 4
        for (int r = 0; r < n; ++r) dst[r + (int64_t)c * ld] = (r <= c) ? (T)L.G[r][c] : T(0);
    }
    // X row-major: lane c holds column c
#pragma unroll
    for (int r = 0; r < 32; ++r) G[r * 32 + c] = x[r];
}

template <typename T>
__global__ __launch_bounds__(64) void tsqr_gram_chol(const TreeDesc d, T* __restrict__ R, int ldr,
                                                      T* __restrict__ blocks, double* __restrict__ side,
                                                      int* __restrict__ flags, int* __restrict__ list,
                                                      const int* __restrict__ ctl, double gate) {
    __shared__ GramLds Lh[2];
    if (ctl[0]) return;  // the flat path already has R and R^{-1}
    const int pairs = (d.lv[0].tiles + 1) / 2;
#pragma unroll 1
    for (int pair = blockIdx.x; pair < pairs; pair += gridDim.x) {
        gram_chol_pair<T>(d, R, ldr, blocks, side, flags, list, gate, Lh, pair);
        wave_sync();  // Lh is reused by the next pair
    }
}

// Pass 2 of the Gram path, one wavefront per superleaf.
// Prologue: M = X S on the matrix cores (X = R^{-1} from side[leaf], S the
// superleaf's block of the tree's Q; at the root, and on the flat path, S
// is the identity and M is X itself, loaded in the same layout).  The MFMA
// accumulator layout of M (lane l, register v of tile (i0, j0) holds
// M[i0 + 4v + l/16][j0 + l%16]) is exactly the B-operand layout of the
// k-group k0 = i0 + 4v, so M stays in registers as mb[k-group][j-tile].
// Rows: blocks of 64 rows; lane l loads rows 4(l%16)..+3 of column
// k0 + l/16 with one 16-byte load, MFMA number t takes row 4(l%16)+t, and
// the accumulator register v of (t, j-tile) is row 16v + 4(l/16) + t,
// column j0 + l%16 of the output, so each lane stores 4 consecutive rows of
// one column per (v, j-tile).  Products and sums in FP64, one rounding.
// The grid is superleaves x parts: workgroup b takes part b % parts (a run
// of whole 64-row blocks) of superleaf b / parts, so that few superleaves
// still give many waves.  On the flat path every superleaf runs; otherwise
// those flagged Householder are left to tsqr_super_q.
template <typename T, bool VEC>
__global__ __launch_bounds__(64) void tsqr_gram_q(const TreeDesc d, T* __restrict__ A, int lda,
                                                   const double* __restrict__ side, const T* __restrict__ blocks,
                                                   const int* __restrict__ flags, const int* __restrict__ ctl,
                                                   int parts) {
    // Highest superleaves first: the streaming R pass reads them last, so
    // part of them is still in the memory-side cache.
    const int item = (int)gridDim.x - 1 - (int)blockIdx.x;
    const int leaf = item / parts;
    const int part = item - leaf * parts;
    const bool flat = ctl[0] != 0;
    if (!flat && flags[leaf]) return;
    const int lane = threadIdx.x;
    const int n = d.n;
    const LevelDesc& L0 = d.lv[0];
    const int li = lane & 15, lq = lane >> 4;
    const int leaf_rows = leaf_rows_of(L0, leaf);
    const int per = (((leaf_rows + 63) >> 6) + parts - 1) / parts * 64;  // rows per part
    const int first = part * per;
    if (first >= leaf_rows) return;
    // ---- M = X S ----
    // flat path: X = R^{-1} of the whole matrix (side block T0) and S = I
    const double* X = side + (int64_t)(flat ? d.side_leaves : leaf) * 1024;  // row-major
    const bool root = flat || L0.tiles == 1;
    const T* S = nullptr;
    if (!root) {
        const int parent = leaf / d.arity;
        const int slot = leaf - parent * d.arity;
        S = blocks + d.lv[1].blk_off + (int64_t)parent * d.H * n + slot * n;  // ld H
    }
    double mb[8][2];
    if (root) {
        // S = I_n: M is X with the columns past n cleared, read straight
        // into the B-operand layout (k-group kg holds row 4 kg + l/16)
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) {
                const int jc = 16 * jt + li;
                mb[kg][jt] = jc < n ? X[(4 * kg + lq) * 32 + jc] : 0.0;
            }
    } else {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
#pragma unroll
            for (int jt = 0; jt < 2; ++jt) {
                v4d acc = {0, 0, 0, 0};
#pragma unroll
                for (int kg = 0; kg < 8; ++kg) {
                    const int k = 4 * kg + lq;
                    const double xa = X[(16 * it + li) * 32 + k];
                    const int jc = 16 * jt + li;
                    const double sb = (k < n && jc < n) ? (double)S[k + (int64_t)jc * d.H] : 0.0;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, sb, acc, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) mb[4 * it + v][jt] = acc[v];
            }
        }
    }
    // ---- Q rows = A rows * M ----
    const int rows = leaf_rows - first < per ? leaf_rows - first : per;
    T* base = A + (int64_t)leaf * L0.tile_rows + first;
    const int nk = (n + 3) / 4;  // k-groups holding real columns
    // raw[kg]: this lane's 4 input rows of column 4 kg + l/16 for the block
    // being computed; the next block's loads are issued before the MFMAs.
    Raw4<T> raw[8];
#pragma unroll
    for (int kg = 0; kg < 8; ++kg) {
        const int k = 4 * kg + lq;
        if (kg < nk) raw[kg] = load_raw4<T, VEC>(base, lda, 4 * li, k < n ? k : n - 1, rows, n);
    }
#pragma unroll 1
    for (int r0 = 0; r0 < rows; r0 += 64) {
        double a[8][4];
#pragma unroll
        for (int kg = 0; kg < 8; ++kg)
#pragma unroll
            for (int t = 0; t < 4; ++t) a[kg][t] = (double)raw[kg].v[t];
        const int r1 = r0 + 64;
        if (r1 < rows) {
#pragma unroll
            for (int kg = 0; kg < 8; ++kg) {
                const int k = 4 * kg + lq;
                if (kg < nk) raw[kg] = load_raw4<T, VEC>(base, lda, r1 + 4 * li, k < n ? k : n - 1, rows, n);
            }
        }
        v4d acc[4][2];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = v4d{0, 0, 0, 0};
        // On the flat path M = R^{-1} is upper triangular: its rows 16..31
        // are zero in columns 0..15, so those MFMAs are skipped.
#pragma unroll
        for (int kg = 0; kg < 8; ++kg) {
            if (kg < nk) {
                const bool lower_zero = flat && kg >= 4;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (!lower_zero)
                        acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][0], acc[t][0], 0, 0, 0);
                    acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][t], mb[kg][1], acc[t][1], 0, 0, 0);
                }
            }
        }
        // acc[t][jt][v] = Q[r0 + 16v + 4 lq + t][16 jt + li]
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) {
            const int cc = 16 * jt + li;
            if (cc < n) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    T q[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) q[t] = (T)acc[t][jt][v];
                    store_column<T, 4, VEC>(q, base + (int64_t)cc * lda + r0 + 16 * v + 4 * lq, rows,
                                            r0 + 16 * v + 4 * lq);
                }
            }
        }
    }
}

// Sums `count` 32 x 32 FP64 blocks at src in index order (256 threads):
// out[e] is entry t + 256 e of the sum.  Eight blocks are in flight per
// step (the blocks come from other CUs, so every load goes to memory),
// the last step's too: its loads past the end reread the last block and
// add zero, so a short sum is one round trip to memory, not one per block.
__device__ __forceinline__ void sum_blocks(const double* __restrict__ src, int count, int t, double (&out)[4]) {
    double s[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int b = 0; b < count; b += 8) {
        double v[8][4];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int bq = b + q < count ? b + q : count - 1;
#pragma unroll
            for (int e = 0; e < 4; ++e) v[q][e] = src[(int64_t)bq * 1024 + t + 256 * e];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) s[e] += b + q < count ? v[q][e] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) out[e] = s[e];
}

// L.G = sum of the nparts 32 x 32 FP64 partials, added in index order
// (kSumThreads = 256 threads, thread t owns entries t + 256 e; sum_blocks).
// The caller synchronises before reading L.G.
constexpr int kSumThreads = 256;
__device__ __forceinline__ void sum_partials(GramLds& L, int nparts, const double* __restrict__ partial, int t) {
    double s[4];
    sum_blocks(partial, nparts, t, s);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int idx = t + 256 * e;
        L.G[idx >> 5][idx & 31] = s[e];
    }
}

// L.G = the sum of `count` 32 x 32 FP64 blocks at src (block b at b * stride
// blocks), by a workgroup of four waves (256 threads): wave w adds the
// blocks [w c / 4, (w + 1) c / 4) in order, four blocks in flight (they were
// written by other CUs, so every load goes to memory), then the four wave
// sums are added in wave order in LDS, so the result is fixed for a given
// count.  Every thread of the workgroup must call it.
__device__ __forceinline__ void sum_blocks_wg4(GramLds& L, const double* __restrict__ src, int count,
                                               int stride = 1) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int b0 = (int)((int64_t)count * w / 4), b1 = (int)((int64_t)count * (w + 1) / 4);
    double s[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) s[e] = 0.0;
    // four blocks in flight per step, the last step's too (its loads past
    // b1 reread block b1 - 1 and add zero)
#pragma unroll 1
    for (int b = b0; b < b1; b += 4) {
        double v[4][16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int bq = b + q < b1 ? b + q : b1 - 1;
#pragma unroll
            for (int e = 0; e < 16; ++e) v[q][e] = src[(int64_t)bq * stride * 1024 + lane + 64 * e];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 16; ++e) s[e] += b + q < b1 ? v[q][e] : 0.0;
    }
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        if (w == k) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int idx = lane + 64 * e;
                L.G[idx >> 5][idx & 31] = (k == 0 ? 0.0 : L.G[idx >> 5][idx & 31]) + s[e];
            }
        }
        __syncthreads();
    }
}

// Cholesky factor and its inverse in one sweep, by one wave.  Lane c < 32
// starts from column c of the Gram matrix in L.G (identity outside the
// leading n x n block) and ends with column c of R (zeros below the
// diagonal); lane 32 + c starts from column c of the identity and, by the
// same row operations (forward substitution with R^T), ends with column c
// of R^{-T}, i.e. row c of R^{-1}.  Step j scales row j by 1/R_jj and
// subtracts R_jr times it from every row r > j.  The step's critical path
// is short: row j + 1 is updated first, with R_j,j+1 read from lane j + 1,
// and the next pivot's 1/sqrt is formed right away; the other R_jr reach
// every lane through L.row (LDS instructions of a wave execute in order, so
// the next step's write cannot pass this step's reads).  On MI355X: 12.7k
// cycles against 15.7k with the whole row broadcast before each pivot, bit
// for bit the same results.  The whole symmetric Schur complement is
// updated (no per-lane select); its part below the diagonal is zeroed as
// each step reaches it.  Returns false (in every lane) if a pivot is not
// positive and finite; the factor is then meaningless.
__device__ __forceinline__ bool chol_inv_wave(GramLds& L, int n, int lane, double (&g)[32]) {
    const int c = lane & 31;
    const bool hi = lane >= 32;
#pragma unroll
    for (int r = 0; r < 32; ++r) g[r] = (!hi && r < n && c < n) ? L.G[r][c] : (r == c ? 1.0 : 0.0);
    bool ok = true;
    auto pivot = [&](int j) {
        double piv = readlane_t<double>(g[j], j);
        if (!(piv > 0.0) || !(piv < __builtin_inf())) {
            ok = false;
            piv = 1.0;
        }
        return rsqrt_fp64(piv);
    };
    double rd = pivot(0);
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        g[j] = (hi || c >= j) ? g[j] * rd : 0.0;
        if (!hi) L.row[c] = g[j];
        if (j + 1 < 32) {
            g[j + 1] = fmat(-readlane_t<double>(g[j], j + 1), g[j], g[j + 1]);
            rd = pivot(j + 1);
        }
        wave_sync();
#pragma unroll
        for (int r = j + 2; r < 32; ++r) g[r] = fmat(-L.row[r], g[j], g[r]);
    }
    return ok;
}

// Cholesky of the whole-matrix Gram matrix in L.G by one wave (lanes
// 0..63).  If it succeeds and ||R||_F ||R^{-1}||_F <= gate: R (rounded to
// T) goes to the caller, R^{-1} (FP64, row-major) to X, R itself in FP64
// (row-major) to rkeep (if not null) and ctl[0] = 1; with bf16q, also
// ctl[kCtlBf16Q1] = (kappa_est <= 2n), the mixed path's test for a Q pass on
// the BF16 matrix cores.  Otherwise nothing is written.  Returns whether the
// gate passed (the same in every lane).
template <typename T>
__device__ __forceinline__ bool flat_factor(GramLds& L, int n, T* __restrict__ R, int ldr, double* __restrict__ X,
                                            int* __restrict__ ctl, double gate, double* __restrict__ rkeep,
                                            int lane, bool bf16q = false) {
    const int c = lane & 31;
    const bool hi = lane >= 32;
    double g[32];
    bool ok = chol_inv_wave(L, n, lane, g);
    // ||R||_F^2 over lanes 0..n-1, ||R^{-1}||_F^2 over lanes 32..32+n-1
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 32; ++r) s = fmat(g[r], g[r], s);
    if (c >= n) s = 0.0;
    s = sum_half(s);
    const double kest = __builtin_sqrt(readlane_t<double>(s, 0) * readlane_t<double>(s, 32));
    ok = ok && (kest <= gate);
    if (!ok) return ok;
    if (!hi) {
        if (c < n) {
#pragma unroll
            for (int r = 0; r < 32; ++r)
                if (r < n) R[r + (int64_t)c * ldr] = (r <= c) ? (T)g[r] : T(0);
        }
        if (rkeep) {
#pragma unroll
            for (int r = 0; r < 32; ++r) rkeep[r * 32 + c] = g[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 32; ++r) X[c * 32 + r] = g[r];
        if (c == 0) {
            ctl[0] = 1;
            if (bf16q) ctl[kCtlBf16Q1] = (kest <= 2.0 * n) ? 1 : 0;
        }
    }
    return ok;
}

// Flat Gram path, stage 2 (one workgroup of kSumThreads threads): G = sum of the
// partials in order, then flat_factor with X = side block T0.  If the gate
// passes, the per-superleaf kernels and the tree do nothing and the Q pass
// forms Q = A R^{-1}; otherwise ctl[0] stays 0 and the tree path runs.
template <typename T>
__global__ __launch_bounds__(kSumThreads) void tsqr_gram_flat(const TreeDesc d, int nparts,
                                                        const double* __restrict__ partial, T* __restrict__ R,
                                                        int ldr, double* __restrict__ side, int* __restrict__ ctl,
                                                        double gate, double* __restrict__ rkeep,
                                                        int bf16q = 0, const int* __restrict__ skip = nullptr) {
    __shared__ GramLds L;
    if (skip && skip[0]) return;  // an earlier pass already took the flat path
    const int t = threadIdx.x;
    sum_partials(L, nparts, partial, t);
    __syncthreads();
    if (t >= 64) return;
    flat_factor<T>(L, d.n, R, ldr, side + (int64_t)d.side_leaves * 1024, ctl, gate, rkeep, t, bf16q != 0);
}

// ---------------------------------------------------------------------------
// Flat Gram path, streaming pass (tsqr_gram_stream).
//
// The whole-matrix Gram G = A^T A does not need per-superleaf sums, so the
// rows are streamed by a persistent grid instead: workgroup b (4 waves) of P
// owns wave indices w = 4b + 0..3, and wave w accumulates the chunks (64
// rows of FP32, 32 of FP64) w, w + 4P, w + 8P, ... in FP64 MFMA registers (the interleaving keeps
// every wave on nearby addresses and all of them busy to the end).  The
// workgroup then sums its 4 waves in LDS (in wave order) into partial[b].
// Workgroups form groups (stream_group_size); the last of a group to finish
// (ticket; no wave ever waits on another) sums the group's partials in index
// order into partial[P + g].  tsqr_gram_flat then sums the
// stream_group_count(P) group partials in order and factors the result.
// Every sum runs in a fixed order, so the result does not depend on which
// workgroup arrives last.  P depends only on m (stream_groups, plan.hpp).

// Slabs of 16 rows per streamed chunk: 64 rows of FP32 or 32 rows of FP64,
// so that a chunk and its prefetched successor take 64 VGPRs either way.
template <typename T>
constexpr int stream_slabs() {
    return sizeof(T) == 4 ? 4 : 2;
}

// Rows [r0, r0 + 4) of column c, all inside the matrix and c a real column
// (no checks).  The streaming kernel loads the padding columns c >= n as
// column 0 and drops them when it converts (gram_chunk's in0 / in1): a
// select right after the load, or a branch around it, made the compiler wait
// for every load in flight, prefetches included, before the next
// instruction.
template <typename T, bool VEC>
__device__ __forceinline__ Raw4<T> load_col4(const T* __restrict__ base, int64_t ld, int r0, int c) {
    Raw4<T> o;
    const T* p = base + (int64_t)c * ld + r0;
    if constexpr (VEC && sizeof(T) == 4) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 q = *reinterpret_cast<const f4*>(p);
#pragma unroll
        for (int t = 0; t < 4; ++t) o.v[t] = q[t];
    } else if constexpr (VEC) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 q0 = *reinterpret_cast<const d2*>(p);
        const d2 q1 = *reinterpret_cast<const d2*>(p + 2);
        o.v[0] = q0[0];
        o.v[1] = q0[1];
        o.v[2] = q1[0];
        o.v[3] = q1[1];
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) o.v[t] = p[t];
    }
    return o;
}

// FP64 MFMA Gram update of one chunk held as S slabs of 16 rows (lane l:
// rows 16q + 4(l/16) .. +3 of columns l%16 and 16 + l%16); values of a
// lane whose column is padding (in0 / in1 false) count as zero.
template <typename T, int NT, int S>
__device__ __forceinline__ void gram_chunk(const Raw4<T> (&c0)[S], const Raw4<T> (&c1)[S], v4d& a00, v4d& a01,
                                           v4d& a11, bool in0 = true, bool in1 = true) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double x0 = in0 ? (double)c0[q].v[t] : 0.0;
            a00 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, x0, a00, 0, 0, 0);
            if constexpr (NT == 2) {
                const double x1 = in1 ? (double)c1[q].v[t] : 0.0;
                a01 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, x1, a01, 0, 0, 0);
                a11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, x1, a11, 0, 0, 0);
            }
        }
    }
}

// One ticket per workgroup: returns (in every thread) whether this
// workgroup is the last of `count` to arrive at *tk.  The caller's global
// writes are released before the ticket and, for the last arriver, the
// other workgroups' writes are acquired after it.
__device__ __forceinline__ bool last_group(int* tk, int count, int* bcast) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const int old = __hip_atomic_fetch_add(tk, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *bcast = old;
    }
    __syncthreads();
    const bool last = *bcast == count - 1;
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}

// Tail of a streaming Gram launch (256 threads, P = gridDim.x workgroups),
// after the workgroup's sum is in L.G (synchronised): L.G goes to
// partial[blockIdx.x]; the last workgroup of each group of G =
// stream_group_size(P) (ticket tk[g]) sums the group's partials in index
// order into partial[P + g].  Returns (in every thread) whether this
// workgroup wrote a group sum.
__device__ __forceinline__ bool stream_finish(GramLds& L, double* __restrict__ partial, int* __restrict__ tk,
                                              int* bcast) {
    const int t = threadIdx.x;
    const int P = gridDim.x;
    double* mine = partial + (int64_t)blockIdx.x * 1024;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int idx = t + 256 * e;
        mine[idx] = L.G[idx >> 5][idx & 31];
    }
    const int G = stream_group_size(P);
    const int g = blockIdx.x / G;
    const int g0 = g * G;
    const int members = P - g0 < G ? P - g0 : G;
    if (!last_group(tk + g, members, bcast)) return false;
    double sum[4];
    sum_blocks(partial + (int64_t)g0 * 1024, members, t, sum);
    double* gsum = partial + (int64_t)(P + g) * 1024;
#pragma unroll
    for (int e = 0; e < 4; ++e) gsum[t + 256 * e] = sum[e];
    return true;
}

// End of a streaming Gram launch (256 threads, 4 waves, P = gridDim.x
// workgroups): each wave holds its FP64 MFMA accumulators of the blocks
// (0,0), (0,1), (1,1) of a 32 x 32 Gram matrix (lane l, register v: row
// 4v + l/16, column l%16 of the block).  The workgroup sums its waves in
// wave order in L.G (stream_lds_sum), then stream_reduce writes it to
// partial[blockIdx.x] and runs stream_finish, whose result it returns.
template <int NT>
__device__ __forceinline__ void stream_lds_sum(GramLds& L, const v4d& a00, const v4d& a01, const v4d& a11) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int col = lane & 15;
    const int i0 = lane >> 4;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        if (wv == s) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 4 * v + i0;
                const double g01 = NT == 2 ? a01[v] : 0.0;
                const double g11 = NT == 2 ? a11[v] : 0.0;
                if (s == 0) {
                    L.G[i][col] = a00[v];
                    L.G[i][16 + col] = g01;
                    L.G[16 + col][i] = g01;
                    L.G[16 + i][16 + col] = g11;
                } else {
                    L.G[i][col] += a00[v];
                    L.G[i][16 + col] += g01;
                    L.G[16 + col][i] += g01;
                    L.G[16 + i][16 + col] += g11;
                }
            }
        }
        __syncthreads();
    }
}

template <int NT>
__device__ __forceinline__ bool stream_reduce(GramLds& L, const v4d& a00, const v4d& a01, const v4d& a11,
                                              double* __restrict__ partial, int* __restrict__ tk, int* bcast) {
    stream_lds_sum<NT>(L, a00, a01, a11);
    return stream_finish(L, partial, tk, bcast);
}

// The sweep of a streaming Gram launch: wave w of W sums the products of
// its chunks (w, w + W, ...) into its FP64 MFMA accumulators a00, a01, a11.
template <typename T, int NT, bool VEC>
__device__ __forceinline__ void gram_stream_acc(const TreeDesc& d, const T* __restrict__ A, int lda, int W, int w,
                                                v4d& a00, v4d& a01, v4d& a11) {
    const int lane = threadIdx.x & 63;
    const int m = d.m, n = d.n;
    const int col = lane & 15;
    const int rq = 4 * (lane >> 4);
    constexpr int S = stream_slabs<T>();
    constexpr int CR = 16 * S;  // rows per chunk
    const int nfull = m / CR;   // complete chunks
    // padding columns (n < 32) load column 0 and are dropped in gram_chunk
    const bool in0 = col < n, in1 = col + 16 < n;
    const int c0col = in0 ? col : 0, c1col = in1 ? col + 16 : 0;
    // Two chunk buffers in turn (ping-pong): the loads of the next chunk are
    // issued before the MFMAs of the current one and consumed after them.
    // The scheduling barriers keep that order; without them the compiler
    // folded the two buffers into one and each wave waited for its loads
    // right before using them.  The last step reloads a chunk it already
    // has instead of branching around the loads.
    auto load_chunk = [&](int ch, Raw4<T>(&x0)[S], Raw4<T>(&x1)[S]) {
#pragma unroll
        for (int q = 0; q < S; ++q) {
            x0[q] = load_col4<T, VEC>(A, lda, CR * ch + 16 * q + rq, c0col);
            if constexpr (NT == 2) x1[q] = load_col4<T, VEC>(A, lda, CR * ch + 16 * q + rq, c1col);
        }
    };
    if (w < nfull) {
        // this wave's chunks are w + k W, k = 0 .. nch - 1, taken in pairs
        // with no exit inside the pair, so the compiler keeps every load
        // where it is written
        const int nch = (nfull - w + W - 1) / W;
        Raw4<T> a0[S], a1[S], b0[S], b1[S];
        load_chunk(w, a0, a1);
#pragma unroll 1
        for (int k = 0; k + 1 < nch; k += 2) {
            load_chunk(w + (k + 1) * W, b0, b1);
            __builtin_amdgcn_sched_barrier(0);
            gram_chunk<T, NT, S>(a0, a1, a00, a01, a11, in0, in1);
            __builtin_amdgcn_sched_barrier(0);
            const int ka = k + 2 < nch ? k + 2 : nch - 1;  // the last pair reloads a chunk it has
            load_chunk(w + ka * W, a0, a1);
            __builtin_amdgcn_sched_barrier(0);
            gram_chunk<T, NT, S>(b0, b1, a00, a01, a11, in0, in1);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (nch & 1) gram_chunk<T, NT, S>(a0, a1, a00, a01, a11, in0, in1);
    }
    if (m % CR && w == nfull % W) {  // the short last chunk, bounds-checked loads
        Raw4<T> c0[S], c1[S];
#pragma unroll
        for (int q = 0; q < S; ++q) {
            c0[q] = load_raw4<T, false>(A, lda, CR * nfull + 16 * q + rq, col, m, n);
            if constexpr (NT == 2) c1[q] = load_raw4<T, false>(A, lda, CR * nfull + 16 * q + rq, col + 16, m, n);
        }
        gram_chunk<T, NT, S>(c0, c1, a00, a01, a11);
    }
}

// partial: P + stream_group_count(P) blocks of 32 x 32 FP64; tk: the
// kCtlStreamTickets ticket words, zero on entry.
template <typename T, int NT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void tsqr_gram_stream(const TreeDesc d, const T* __restrict__ A, int lda,
                                                         double* __restrict__ partial, int* __restrict__ tk,
                                                         const int* __restrict__ skip = nullptr) {
    __shared__ GramLds L;
    __shared__ int bcast;
    if (skip && skip[0]) return;  // an earlier pass already took the flat path
    v4d a00 = {0, 0, 0, 0}, a01 = {0, 0, 0, 0}, a11 = {0, 0, 0, 0};
    gram_stream_acc<T, NT, VEC>(d, A, lda, 4 * (int)gridDim.x, 4 * (int)blockIdx.x + (int)(threadIdx.x >> 6), a00,
                                a01, a11);
    stream_reduce<NT>(L, a00, a01, a11, partial, tk, &bcast);
}

// ---------------------------------------------------------------------------
// Short matrices (plan.hpp, short_waves): the R pass without tickets on
// workgroups of W = 4, 8 or 16 waves with one chunk per wave
// (tsqr_gram_short), the fewest waves that keep the grid within
// kShortGroups; each workgroup sums its waves in LDS and writes
// partial[blockIdx.x] (its three distinct 16 x 16 blocks), and
// tsqr_flat_short sums the P partials with 8 waves (all in flight at once)
// and factors.  Measured on MI355X for
// 27648 x 32: the ticketed 4-wave launch spends about 8 of its 12-15 us
// after its products (the release fence before each ticket, about 3.5 us,
// the acquire and the group sums); on 16-wave workgroups (27 CUs) the FP64
// products alone take 7.5 us, on 8-wave ones (54 CUs) about half.

// A short partial holds the three distinct 16 x 16 blocks of a symmetric
// 32 x 32 Gram matrix, (0,0), (0,1) and (1,1), row-major one after the
// other: three quarters of the bytes the flat sum reads from other CUs.
constexpr int kShortPartial = 768;

// Sums the W waves' FP64 MFMA accumulators (stream_lds_sum's layout) in
// four LDS blocks (wave w adds into Q[w % 4], in the order w / 4), then
// writes the total, ((Q0 + Q1) + Q2) + Q3, to out as a short partial.
template <int NT, int W>
__device__ __forceinline__ void stream_lds_write(GramLds (&Q)[4], const v4d& a00, const v4d& a01, const v4d& a11,
                                                 double* __restrict__ out) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int col = lane & 15;
    const int i0 = lane >> 4;
    GramLds& L = Q[wv & 3];
#pragma unroll 1
    for (int s = 0; s < W / 4; ++s) {
        if ((wv >> 2) == s) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 4 * v + i0;
                const double g01 = NT == 2 ? a01[v] : 0.0;
                const double g11 = NT == 2 ? a11[v] : 0.0;
                if (s == 0) {
                    L.G[i][col] = a00[v];
                    L.G[i][16 + col] = g01;
                    L.G[16 + col][i] = g01;
                    L.G[16 + i][16 + col] = g11;
                } else {
                    L.G[i][col] += a00[v];
                    L.G[i][16 + col] += g01;
                    L.G[16 + col][i] += g01;
                    L.G[16 + i][16 + col] += g11;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < (kShortPartial + 64 * W - 1) / (64 * W); ++h) {
        const int idx = t + 64 * W * h;
        if (idx < kShortPartial) {
            const int blk = idx >> 8, i = (idx >> 4) & 15, j = idx & 15;
            const int r = (blk == 2 ? 16 : 0) + i, c = (blk == 0 ? 0 : 16) + j;
            out[idx] = ((Q[0].G[r][c] + Q[1].G[r][c]) + Q[2].G[r][c]) + Q[3].G[r][c];
        }
    }
}

// partial: gridDim.x blocks of 32 x 32 FP64.  ini.on: the grid also sets
// up the call's workspace (the launch is the call's first, and nothing in it
// reads the words it writes).
template <typename T, int NT, bool VEC, int W>
__global__ __launch_bounds__(64 * W) void tsqr_gram_short(const TreeDesc d, const T* __restrict__ A, int lda,
                                                          double* __restrict__ partial,
                                                          const int* __restrict__ skip, const WsInit ini) {
    __shared__ GramLds Q[4];
    if (ini.on)
        for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < ini.count(); i += (int)(gridDim.x * blockDim.x))
            ini.at(i);
    if (skip && skip[0]) return;  // an earlier pass already took the flat path
    v4d a00 = {0, 0, 0, 0}, a01 = {0, 0, 0, 0}, a11 = {0, 0, 0, 0};
    gram_stream_acc<T, NT, VEC>(d, A, lda, W * (int)gridDim.x, W * (int)blockIdx.x + (int)(threadIdx.x >> 6), a00,
                                a01, a11);
    stream_lds_write<NT, W>(Q, a00, a01, a11, partial + (int64_t)blockIdx.x * 1024);
}

// Q[0].G = the sum of `count` short partials at src (kShortPartial entries
// each, one per 1024-double slot) by 8 waves (512 threads): wave w adds the
// partials [w c / 8, (w + 1) c / 8) in order, all of them in flight at once
// for c <= 64 (eight per wave), then the wave sums are added in four LDS
// blocks (wave w into Q[w % 4], in the order w / 4) and those as
// ((Q0 + Q1) + Q2) + Q3, and the symmetric 32 x 32 matrix is filled in.
// Every thread of the workgroup must call it; Q[0] is synchronised on
// return.
constexpr int kFlatShortThreads = 512;
__device__ __forceinline__ void sum_short_wg8(GramLds (&Q)[4], const double* __restrict__ src, int count) {
    constexpr int E = kShortPartial / 64;  // entries per lane
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int b0 = count * w / 8, b1 = count * (w + 1) / 8;
    double s[E];
#pragma unroll
    for (int e = 0; e < E; ++e) s[e] = 0.0;
    // eight partials in flight per step, the last step's too (its loads past
    // b1 reread partial b1 - 1 and add zero)
#pragma unroll 1
    for (int b = b0; b < b1; b += 8) {
        double v[8][E];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int bq = b + q < b1 ? b + q : b1 - 1;
#pragma unroll
            for (int e = 0; e < E; ++e) v[q][e] = src[(int64_t)bq * 1024 + lane + 64 * e];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int e = 0; e < E; ++e) s[e] += b + q < b1 ? v[q][e] : 0.0;
    }
    // the four LDS blocks hold the partial layout flat (entry idx at
    // G[idx / 32][idx % 32])
    GramLds& L = Q[w & 3];
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        if ((w >> 2) == k) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int idx = lane + 64 * e;
                L.G[idx >> 5][idx & 31] = (k == 0 ? 0.0 : L.G[idx >> 5][idx & 31]) + s[e];
            }
        }
        __syncthreads();
    }
    double sum[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int idx = t + kFlatShortThreads * h;
        const int r = idx >> 5, c = idx & 31;
        sum[h] = idx < kShortPartial ? ((Q[0].G[r][c] + Q[1].G[r][c]) + Q[2].G[r][c]) + Q[3].G[r][c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int idx = t + kFlatShortThreads * h;
        if (idx < kShortPartial) {
            const int blk = idx >> 8, i = (idx >> 4) & 15, j = idx & 15;
            if (blk == 0) {
                Q[0].G[i][j] = sum[h];
            } else if (blk == 1) {
                Q[0].G[i][16 + j] = sum[h];
                Q[0].G[16 + j][i] = sum[h];
            } else {
                Q[0].G[16 + i][16 + j] = sum[h];
            }
        }
    }
    __syncthreads();
}

// Flat factor of a short matrix (one workgroup of 8 waves): G = the sum of
// the nparts partials of tsqr_gram_short, then flat_factor as
// tsqr_gram_flat.
template <typename T>
__global__ __launch_bounds__(kFlatShortThreads) void tsqr_flat_short(const TreeDesc d, int nparts,
                                                         const double* __restrict__ partial, T* __restrict__ R,
                                                         int ldr, double* __restrict__ side, int* __restrict__ ctl,
                                                         double gate, double* __restrict__ rkeep,
                                                         const int* __restrict__ skip = nullptr) {
    __shared__ GramLds Q[4];
    if (skip && skip[0]) return;  // an earlier pass already took the flat path
    sum_short_wg8(Q, partial, nparts);
    if (threadIdx.x >= 64) return;
    flat_factor<T>(Q[0], d.n, R, ldr, side + (int64_t)d.side_leaves * 1024, ctl, gate, rkeep, threadIdx.x);
}

}  // namespace tsqr
