/*
 * Oracle body, included twice by tsqr_oracle.c: once with REAL=float/SUF=s and
 * once with REAL=double/SUF=d.  TEST INFRASTRUCTURE ONLY: this is the CPU
 * restatement of the reference algorithm that the HIP library is checked
 * against.  It is never linked into, called by, or used as a fallback of the
 * product library.
 *
 * Every function below follows the reference file:line named in its comment.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* Householder QR of one h x n tile, explicit thin Q.
 * Follows kernelQR.h:69-291 (qr_kernel):
 *   - reflector: u = x/nu; u_j += sign(u_j); u *= 1/sqrt(|u_j|)   (:105-156)
 *     so ||u||^2 = 2 and H = I - u u^T; R_jj = -sign(x_jj)*nu      (:140-142)
 *   - trailing update x_c -= (u^T x_c) u for c > j                  (:163-196)
 *   - R extraction: strict upper from the tile, diagonal from RR,
 *     zeros below; the tile's strict upper part is then zeroed     (:203-220)
 *   - explicit Q: column c is H_0 ... H_{n-1} e_c, reflectors applied in
 *     reverse order, skipping j > c                                 (:227-290)
 * Q (ld ldq) receives the h x n explicit Q; R (ld ldr) receives n x n.
 * The reference computes 1.0/norm in double for float (:126); kept here.
 * Divergence: a zero column gives NaN in the reference (division by a
 * zero norm, :123-126); here it gives tau = 0 (H = I) and R_jj = 0 like
 * LAPACK dlarfg.  Only the summation order differs otherwise (the
 * reference reduces with per-lane partials and a warp butterfly).
 */
void FN(oracle_qr_tile)(int h, int n, const REAL *A, long lda, REAL *Q, long ldq,
                        REAL *R, long ldr) {
    REAL RR[1024];
    REAL *U = (REAL *)malloc(sizeof(REAL) * (size_t)h * n);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < h; ++i) U[i + (size_t)c * h] = A[i + c * lda];

    for (int j = 0; j < n; ++j) {
        REAL *u = U + (size_t)j * h;
        REAL nu = 0;
        for (int i = j; i < h; ++i) nu += u[i] * u[i];
        REAL norm = (REAL)sqrt((double)nu);
        if (norm == 0) {
            /* LAPACK convention: H_j = I, R_jj = 0; u_j = 0 marks it */
            RR[j] = 0;
            for (int i = j; i < h; ++i) u[i] = 0;
            continue;
        }
        REAL scale = (REAL)(1.0 / (double)norm);
        for (int i = j; i < h; ++i) u[i] *= scale;
        REAL u1 = u[j];
        u[j] += (u1 >= 0) ? (REAL)1 : (REAL)-1;
        RR[j] = (u1 >= 0) ? -norm : norm;
        scale = (REAL)1 / (REAL)sqrt((double)fabs((double)u[j]));
        for (int i = j; i < h; ++i) u[i] *= scale;
        for (int c = j + 1; c < n; ++c) {
            REAL *x = U + (size_t)c * h;
            REAL utx = 0;
            for (int i = j; i < h; ++i) utx += u[i] * x[i];
            for (int i = j; i < h; ++i) x[i] -= utx * u[i];
        }
    }
    /* R extraction and zeroing of the strict upper part of U */
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            if (i < c) {
                R[i + c * ldr] = (i < h) ? U[i + (size_t)c * h] : 0;
                if (i < h) U[i + (size_t)c * h] = 0;
            } else if (i == c) {
                R[i + c * ldr] = RR[c];
            } else {
                R[i + c * ldr] = 0;
            }
        }
    }
    /* explicit Q: reverse application, skipping reflectors j > c */
    REAL *q = (REAL *)malloc(sizeof(REAL) * (size_t)h);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < h; ++i) q[i] = (i == c) ? (REAL)1 : (REAL)0;
        for (int j = n - 1; j >= 0; --j) {
            if (c < j) continue;
            const REAL *u = U + (size_t)j * h;
            REAL utq = 0;
            for (int i = 0; i < h; ++i) utq += u[i] * q[i];
            for (int i = 0; i < h; ++i) q[i] -= utq * u[i];
        }
        for (int i = 0; i < h; ++i) Q[i + c * ldq] = q[i];
    }
    free(q);
    free(U);
}

/* One tsqr_kernel chunk (kernelQR.h:321-496) over an m x n matrix A (in/out),
 * block size bs.  Up-sweep: level heights H_{l+1} = ceil(H_l/bs)*n, tile b of
 * a level covers rows [b*bs, b*bs + min(bs, H_l - b*bs)); its R is stored in
 * place in A rows [b*n, b*n+n), or in R at the top level (:341-388).  Tiles
 * are processed in ascending b, which is the order that makes the in-place
 * R stores well defined.  Down-sweep (:390-492): at each level below the top,
 * A rows [b*bs, ...) := Q_{l,b} * A[b*n : b*n+n, :]; the top tile copies its
 * Q into A (:457-480). */
static int FN(oracle_tsqr_chunk)(int m, int n, REAL *A, long lda, REAL *R, long ldr, int bs) {
    int heights[64];
    int nlev = 0;
    int H = m;
    /* Q tiles of every level: level l has ceil(H_l/bs) tiles of bs x n */
    REAL **qs = (REAL **)calloc(64, sizeof(REAL *));
    while (H > n || nlev == 0) {
        if (nlev >= 63) return -2;
        heights[nlev] = H;
        int nblk = (H + bs - 1) / bs;
        qs[nlev] = (REAL *)calloc((size_t)nblk * bs * n, sizeof(REAL));
        if (H <= bs) {
            FN(oracle_qr_tile)(H, n, A, lda, qs[nlev], bs, R, ldr);
        } else {
            /* Tile b's R goes to rows [b*n, b*n+n) of A.  In the sequential
             * order those rows belong to tiles already factored (b*n < b*bs),
             * so staging the Rs and copying them after the loop gives the
             * same values and lets the tiles run in parallel. */
            REAL *Rst = (REAL *)malloc(sizeof(REAL) * (size_t)nblk * n * n);
#pragma omp parallel for schedule(dynamic, 4)
            for (int b = 0; b < nblk; ++b) {
                int h = H - b * bs < bs ? H - b * bs : bs;
                FN(oracle_qr_tile)(h, n, A + (long)b * bs, lda, qs[nlev] + (size_t)b * bs * n, bs,
                                   Rst + (size_t)b * n, (long)nblk * n);
            }
            for (int c = 0; c < n; ++c)
                for (long i = 0; i < (long)nblk * n; ++i) A[i + c * lda] = Rst[i + (size_t)c * nblk * n];
            free(Rst);
        }
        H = nblk * n;
        ++nlev;
    }
    for (int l = nlev - 1; l >= 0; --l) {
        int Hl = heights[l];
        int nblk = (Hl + bs - 1) / bs;
        if (Hl > bs) {
            REAL *work = (REAL *)malloc(sizeof(REAL) * (size_t)Hl * n);
#pragma omp parallel for schedule(static)
            for (int b = 0; b < nblk; ++b) {
                int h = Hl - b * bs < bs ? Hl - b * bs : bs;
                const REAL *qt = qs[l] + (size_t)b * bs * n; /* ld bs */
                const REAL *qn = A + (long)b * n;            /* ld lda */
                /* block_gemm (kernelQR.h:38-59): k loop ascending */
                for (int c = 0; c < n; ++c)
                    for (int i = 0; i < h; ++i) {
                        REAL sum = 0;
                        for (int k = 0; k < n; ++k) sum += qt[i + (long)k * bs] * qn[k + c * lda];
                        work[(long)b * bs + i + (size_t)c * Hl] = sum;
                    }
            }
            for (int c = 0; c < n; ++c)
                for (int i = 0; i < Hl; ++i) A[i + c * lda] = work[i + (size_t)c * Hl];
            free(work);
        } else {
            for (int c = 0; c < n; ++c)
                for (int i = 0; i < Hl; ++i) A[i + c * lda] = qs[l][i + (long)c * bs];
        }
    }
    for (int l = 0; l < nlev; ++l) free(qs[l]);
    free(qs);
    return 0;
}

/* tsqr host driver (TallShinnyQR.h:86-149).  m <= num_sm*bs: one chunk.
 * Otherwise: chunk pass with per-chunk R stacked at rows g*n of a work
 * matrix W1 (:119-120), TSQR of W1 giving R and W1 := Q2 (:122-124), then
 * A_g := A_g * W1[g*n:(g+1)*n, :] (tsgemm_tsqr, :11-74, :126-127).
 * Returns 0, or -1 for arguments outside the reference's domain
 * (m >= n, bs % n == 0, n | rows of each chunk, m <= (M/n)*M). */
int FN(oracle_tsqr)(int m, int n, REAL *A, long lda, REAL *R, long ldr, int bs, int num_sm) {
    if (n < 1 || m < n || bs % n != 0 || lda < m || ldr < n) return -1;
    long M = (long)num_sm * bs;
    if ((long)m > (M / n) * M) return -1;
    int grid_num = (int)((m + M - 1) / M);
    if (grid_num == 1) {
        if ((m % bs) % n != 0) return -1;
        return FN(oracle_tsqr_chunk)(m, n, A, lda, R, ldr, bs);
    }
    if ((m % M) % n != 0) return -1;
    int hw = grid_num * n;
    REAL *W1 = (REAL *)calloc((size_t)hw * n, sizeof(REAL));
    for (int g = 0; g < grid_num; ++g) {
        long r0 = (long)g * M;
        int mg = (int)((m - r0) < M ? (m - r0) : M);
        int st = FN(oracle_tsqr_chunk)(mg, n, A + r0, lda, W1 + (long)g * n, hw, bs);
        if (st) {
            free(W1);
            return st;
        }
    }
    int st = FN(oracle_tsqr_chunk)(hw, n, W1, hw, R, ldr, bs);
    if (st) {
        free(W1);
        return st;
    }
    for (int g = 0; g < grid_num; ++g) {
        long r0 = (long)g * M;
        int mg = (int)((m - r0) < M ? (m - r0) : M);
        const REAL *B = W1 + (long)g * n; /* ld hw */
#pragma omp parallel
        {
            REAL *row = (REAL *)malloc(sizeof(REAL) * n);
#pragma omp for schedule(static)
            for (int i = 0; i < mg; ++i) {
                for (int c = 0; c < n; ++c) {
                    REAL sum = 0;
                    for (int k = 0; k < n; ++k) sum += A[r0 + i + (long)k * lda] * B[k + (long)c * hw];
                    row[c] = sum;
                }
                for (int c = 0; c < n; ++c) A[r0 + i + (long)c * lda] = row[c];
            }
            free(row);
        }
    }
    free(W1);
    return 0;
}

/* Recursive blocked QR (block_QR.h:72-95): n <= nb -> tsqr; else n1 = n/2,
 * qr(A1) -> (Q1, R11); R12 = Q1^T A2 (gemm_blockqr trans_A=1, k = m,
 * :89-90); A2 -= Q1 R12 (:91-92); qr(A2) -> (Q2, R22).  The reference leaves
 * R21 unwritten; here it is set to zero.  Requires n <= nb or n = nb*2^k. */
int FN(oracle_qr)(int m, int n, REAL *A, long lda, REAL *R, long ldr, int bs, int num_sm, int nb) {
    if (n <= nb) return FN(oracle_tsqr)(m, n, A, lda, R, ldr, bs, num_sm);
    if (n % 2 != 0 || (n / 2) % nb != 0) return -1;
    int n1 = n / 2;
    REAL *A1 = A, *A2 = A + (long)n1 * lda;
    REAL *R11 = R, *R12 = R + (long)n1 * ldr, *R22 = R + (long)n1 * ldr + n1;
    int st = FN(oracle_qr)(m, n1, A1, lda, R11, ldr, bs, num_sm, nb);
    if (st) return st;
#pragma omp parallel for collapse(2) schedule(static)
    for (int r = 0; r < n1; ++r)
        for (int c = 0; c < n1; ++c) {
            REAL sum = 0;
            for (int i = 0; i < m; ++i) sum += A1[(long)r * lda + i] * A2[i + (long)c * lda];
            R12[r + (long)c * ldr] = sum;
        }
#pragma omp parallel
    {
        REAL *row = (REAL *)malloc(sizeof(REAL) * n1);
#pragma omp for schedule(static)
        for (int i = 0; i < m; ++i) {
            for (int c = 0; c < n1; ++c) {
                REAL sum = 0;
                for (int k = 0; k < n1; ++k) sum += A1[i + (long)k * lda] * R12[k + (long)c * ldr];
                row[c] = -sum + A2[i + (long)c * lda];
            }
            for (int c = 0; c < n1; ++c) A2[i + (long)c * lda] = row[c];
        }
        free(row);
    }
    for (int c = 0; c < n1; ++c)
        for (int r = n1; r < n; ++r) R[r + (long)c * ldr] = 0;
    return FN(oracle_qr)(m, n1, A2, lda, R22, ldr, bs, num_sm, nb);
}

#undef FN
#undef CAT
#undef CAT2
