// tile.hpp -- one wavefront factors one row tile with Householder reflectors.
//
// Register layout.  A tile is H = 64*RPL rows by N columns.  Lane l holds the
// RPL consecutive rows l*RPL .. l*RPL+RPL-1 of every column in x[s][c], so a
// column of the tile is read from column-major HBM with one 16-byte vector
// load per lane (RPL*sizeof(T) == 16 for the default shapes).  Rows past the
// valid height are zero; a zero row never changes under a reflector, so a
// ragged tile needs no other special case.
//
// Reflector convention (kernelQR.h:105-156): for column j with x = rows >= j,
//   nu = ||x||, s = (x_j >= 0) ? +1 : -1, R_jj = -s*nu,
//   u  = (x + s*nu*e_j) / sqrt(nu*(nu + |x_j|))      so ||u||^2 = 2,
//   H_j = I - u u^T.
// The reference forms u by scaling x by 1/nu and then by 1/sqrt(|u_j|); the
// closed form above is the same vector.  The scalars nu, s, 1/sqrt(...) and
// the row-j entries of R are computed in double for both float and double
// tiles.  One fused wave reduction per reflector yields d_c = x^T x_c for all
// trailing columns c >= j at once (d_j = nu^2), and then
//   u^T x_c = (d_c + s*nu*x_jc) / sqrt(nu*(nu + |x_j|)).
// A zero column (nu == 0) gives u = 0, i.e. H_j = I and R_jj = 0, as LAPACK's
// dlarfg does; the reference divides by zero there (kernelQR.h:123-126).
//
// After factor_tile the registers of column j hold u_j on rows >= j and zero
// above, which is exactly what the reference keeps in shared memory before
// forming Q (kernelQR.h:203-220), and R (upper triangle, N x N, column-major)
// sits in the per-wave LDS area.
//
// form_q_tile then overwrites the reflectors with the explicit thin Q =
// H_0 H_1 ... H_{N-1} [I; 0], applying the reflectors in reverse and touching
// column c only with reflectors j <= c (kernelQR.h:227-290); column j of Q
// takes the registers of u_j once u_j is no longer needed.
#pragma once
#include "wave.hpp"

namespace tsqr {

template <typename T, int N>
struct TileLds {
    T row[N];    // row j of the tile, broadcast to every lane
    T d[N];      // reduced dot products d_c
    T w[N];      // u^T x_c for the trailing columns
    T R[N * N];  // upper triangle of R, column-major, leading dimension N
};

// ---------------------------------------------------------------------------
// Factorization: reflector J applied to the trailing columns J+1..N-1.
// ---------------------------------------------------------------------------
template <typename T, int N, int RPL, int J>
__device__ __forceinline__ void factor_step(T (&x)[RPL][N], TileLds<T, N>& L, int lane) {
    constexpr int V = N - J;  // columns J..N-1 take part in the fused reduction
    constexpr int VP = pow2_ceil(V);
    const int row0 = lane * RPL;
    __builtin_amdgcn_sched_barrier(0);

    T um[RPL];  // column J restricted to rows >= J
#pragma unroll
    for (int s = 0; s < RPL; ++s) um[s] = (row0 + s >= J) ? x[s][J] : T(0);

    T p[VP];
#pragma unroll
    for (int c = 0; c < V; ++c) {
        T acc = um[0] * x[0][J + c];
#pragma unroll
        for (int s = 1; s < RPL; ++s) acc = fmat(um[s], x[s][J + c], acc);
        p[c] = acc;
    }
#pragma unroll
    for (int c = V; c < VP; ++c) p[c] = T(0);
    wave_sum_transposed<VP>(p, lane);
    const int cl = column_of_lane<VP>(lane);
    const bool mine = primary_lane<VP>(lane) && cl < V;

    if (mine) L.d[cl] = p[0];
    constexpr int owner = J / RPL;
    constexpr int slot = J % RPL;
    if (lane == owner) {
#pragma unroll
        for (int c = J; c < N; ++c) L.row[c] = x[slot][c];
    }
    wave_sync();

    // nu = sqrt(djj) and gamma = 1/sqrt(den) from the reciprocal square
    // root (two Newton steps): a shorter dependent chain than a rounded
    // square root and a division, and this chain is on every reflector's
    // critical path.  A zero column gives nu = 0 and gamma = 0; a NaN
    // propagates.
    const double djj = (double)L.d[0];
    const double xjj = (double)L.row[J];
    const double nu = djj * (djj > 0.0 ? rsqrt_fp64(djj) : 0.0);
    const double snu = (xjj >= 0.0) ? nu : -nu;
    const double den = nu * (nu + __builtin_fabs(xjj));
    const double gamma = den > 0.0 ? rsqrt_fp64(den) : 0.0;
    if (mine) {
        const double xjc = (double)L.row[J + cl];
        if (cl == 0) {
            L.R[J + J * N] = (T)(-snu);
        } else {
            const double utx = ((double)p[0] + snu * xjc) * gamma;  // u^T x_c
            L.w[cl] = (T)utx;
            L.R[J + (J + cl) * N] = (T)(xjc - utx * gamma * (xjj + snu));
        }
    }
    wave_sync();

    // u on this lane's rows; rows above J are zero.
    const T g = (T)gamma;
    const T ujj = (T)(gamma * (xjj + snu));
#pragma unroll
    for (int s = 0; s < RPL; ++s) um[s] = (row0 + s > J) ? g * um[s] : (row0 + s == J ? ujj : T(0));

#pragma unroll
    for (int c = 1; c < V; ++c) {
        const T w = L.w[c];
#pragma unroll
        for (int s = 0; s < RPL; ++s) x[s][J + c] = fmat(-w, um[s], x[s][J + c]);
    }
#pragma unroll
    for (int s = 0; s < RPL; ++s) x[s][J] = um[s];
    wave_sync();
}

// No-op column sink of factor_tile.
struct NoColumnSink {
    template <int J, typename T, int RPL>
    __device__ __forceinline__ void column(const T (&)[RPL]) const {}
};

// sink.column<J>(u) receives this lane's rows of reflector J as soon as it
// is final, so a caller that keeps the reflectors (in memory) stores each
// column while the rest of the tile is factored and its registers are free
// from then on.
template <typename T, int N, int RPL, int J = 0, typename Sink = NoColumnSink>
__device__ __forceinline__ void factor_tile(T (&x)[RPL][N], TileLds<T, N>& L, int lane, const Sink& sink = Sink{}) {
    if constexpr (J < N) {
        factor_step<T, N, RPL, J>(x, L, lane);
        T u[RPL];
#pragma unroll
        for (int s = 0; s < RPL; ++s) u[s] = x[s][J];
        sink.template column<J, T, RPL>(u);
        factor_tile<T, N, RPL, J + 1, Sink>(x, L, lane, sink);
    }
}

// ---------------------------------------------------------------------------
// Explicit Q from the stored reflectors, in place, reverse order.
// ---------------------------------------------------------------------------
template <typename T, int N, int RPL, int J>
__device__ __forceinline__ void qform_step(T (&x)[RPL][N], TileLds<T, N>& L, int lane) {
    constexpr int V = N - 1 - J;  // Q columns J+1..N-1 already formed
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (V > 0) {
        constexpr int VP = pow2_ceil(V);
        T p[VP];
#pragma unroll
        for (int c = 0; c < V; ++c) {
            T acc = x[0][J] * x[0][J + 1 + c];
#pragma unroll
            for (int s = 1; s < RPL; ++s) acc = fmat(x[s][J], x[s][J + 1 + c], acc);
            p[c] = acc;
        }
#pragma unroll
        for (int c = V; c < VP; ++c) p[c] = T(0);
        wave_sum_transposed<VP>(p, lane);
        const int cl = column_of_lane<VP>(lane);
        if (primary_lane<VP>(lane) && cl < V) L.w[cl] = p[0];
        wave_sync();
#pragma unroll
        for (int c = 0; c < V; ++c) {
            const T w = L.w[c];
#pragma unroll
            for (int s = 0; s < RPL; ++s) x[s][J + 1 + c] = fmat(-w, x[s][J], x[s][J + 1 + c]);
        }
        wave_sync();
    }
    // Q e_J = e_J - u_J * u_JJ; rows above J become zero.
    const T ujj = readlane_t<T>(x[J % RPL][J], J / RPL);
    const int row0 = lane * RPL;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
        const int r = row0 + s;
        x[s][J] = (r == J) ? (T(1) - ujj * ujj) : -ujj * x[s][J];
    }
}

template <typename T, int N, int RPL, int J = N - 1>
__device__ __forceinline__ void form_q_tile(T (&x)[RPL][N], TileLds<T, N>& L, int lane) {
    if constexpr (J >= 0) {
        qform_step<T, N, RPL, J>(x, L, lane);
        form_q_tile<T, N, RPL, J - 1>(x, L, lane);
    }
}

}  // namespace tsqr
