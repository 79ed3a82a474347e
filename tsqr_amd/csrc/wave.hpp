// wave.hpp -- wave64 cross-lane primitives for gfx950 (CDNA4).
//
// The per-tile Householder kernel keeps one tile row set per lane, so every
// column norm and reflector dot product is a sum over the 64 lanes of a
// wavefront.  The reference does the same sums with a 32-lane __shfl_xor
// butterfly, one reduction at a time (kernelQR.h:30-36, :110-122, :171-181).
// Here V independent sums are reduced together with a *transposed* butterfly:
// at each lane bit the pair (p[2i], p[2i+1]) is folded so that lanes with the
// bit clear keep the even value and lanes with the bit set keep the odd one.
// One exchange plus one add retires two partial sums, so V sums cost about
// 2V instructions instead of 6V, and the result of sum c lands in the lanes
// whose bits spell c (see column_of_lane).
//
// Lane bits are visited in the order 5, 4, 3, 2, 1, 0:
//   bit 5  v_permlane32_swap   (exchange lane halves)
//   bit 4  v_permlane16_swap   (exchange 16-lane rows 0<->1, 2<->3)
//   bit 3  DPP row_mirror      (lane i <-> 15-i inside a row)
//   bit 2  DPP row_half_mirror (lane i <-> 7-i inside a half row)
//   bit 1  DPP quad_perm [2,3,0,1]
//   bit 0  DPP quad_perm [1,0,3,2]
// Each pairing joins lanes that differ in that bit, so the composition of the
// six steps covers all 64 lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsqr {

constexpr int kWave = 64;

// Fused multiply-add in the operand type (__builtin_fma is the double one).
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

constexpr int pow2_ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
constexpr int log2_exact(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// DPP controls (GFX9 encoding).
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum of the two values returned by a permlane swap of (a, b): lanes in the
// "low" group get a_low + a_high, lanes in the "high" group b_low + b_high.
__device__ __forceinline__ float swap32_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a),
                                              __builtin_bit_cast(uint32_t, b), false, false);
    return __builtin_bit_cast(float, (uint32_t)r[0]) + __builtin_bit_cast(float, (uint32_t)r[1]);
}
__device__ __forceinline__ float swap16_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, a),
                                              __builtin_bit_cast(uint32_t, b), false, false);
    return __builtin_bit_cast(float, (uint32_t)r[0]) + __builtin_bit_cast(float, (uint32_t)r[1]);
}
__device__ __forceinline__ double swap32_sum(double a, double b) {
    uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    double x = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0]);
    double y = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
    return x + y;
}
__device__ __forceinline__ double swap16_sum(double a, double b) {
    uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    double x = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0]);
    double y = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
    return x + y;
}

// Lane bit visited at step k (k = 0..5).
constexpr int step_bit(int k) { return 5 - k; }

// One transposed step: returns the kept half of (a, b) plus the partner's.
template <int STEP, typename T>
__device__ __forceinline__ T fold_pair(T a, T b, int lane) {
    if constexpr (STEP == 0) {
        return swap32_sum(a, b);
    } else if constexpr (STEP == 1) {
        return swap16_sum(a, b);
    } else {
        constexpr int bit = step_bit(STEP);
        const bool hi = (lane >> bit) & 1;
        T keep = hi ? b : a;
        T send = hi ? a : b;
        constexpr int ctrl = bit == 3 ? kDppRowMirror
                           : bit == 2 ? kDppRowHalfMirror
                           : bit == 1 ? kDppQuadXor2
                                      : kDppQuadXor1;
        return keep + dpp_move<ctrl>(send);
    }
}

// One plain all-reduce step over lane bit step_bit(STEP).
template <int STEP, typename T>
__device__ __forceinline__ T fold_all(T a) {
    if constexpr (STEP == 0) {
        return swap32_sum(a, a);
    } else if constexpr (STEP == 1) {
        return swap16_sum(a, a);
    } else {
        constexpr int bit = step_bit(STEP);
        constexpr int ctrl = bit == 3 ? kDppRowMirror
                           : bit == 2 ? kDppRowHalfMirror
                           : bit == 1 ? kDppQuadXor2
                                      : kDppQuadXor1;
        return a + dpp_move<ctrl>(a);
    }
}

template <int STEP, int CNT, typename T>
__device__ __forceinline__ void fold_steps(T* p, int lane) {
    if constexpr (CNT > 1) {
#pragma unroll
        for (int i = 0; i < CNT / 2; ++i) p[i] = fold_pair<STEP>(p[2 * i], p[2 * i + 1], lane);
        fold_steps<STEP + 1, CNT / 2>(p, lane);
    } else if constexpr (STEP < 6) {
        p[0] = fold_all<STEP>(p[0]);
        fold_steps<STEP + 1, 1>(p, lane);
    }
}

// Reduces VP (a power of two, <= 64) per-lane partial sums over the wave.
// On return p[0] of every lane holds the full sum number column_of_lane<VP>().
template <int VP, typename T>
__device__ __forceinline__ T wave_sum_transposed(T* p, int lane) {
    static_assert((VP & (VP - 1)) == 0 && VP <= 64, "VP must be a power of two <= 64");
    fold_steps<0, VP>(p, lane);
    return p[0];
}

// Which of the VP sums lane `lane` holds after wave_sum_transposed<VP>.
template <int VP>
__device__ __forceinline__ int column_of_lane(int lane) {
    constexpr int levels = log2_exact(VP);
    int c = 0;
#pragma unroll
    for (int k = 0; k < levels; ++k) c |= ((lane >> step_bit(k)) & 1) << k;
    return c;
}

// True for exactly one lane per column after wave_sum_transposed<VP>: the one
// whose bits not used by the column index are all zero.
template <int VP>
__device__ __forceinline__ bool primary_lane(int lane) {
    constexpr int levels = log2_exact(VP);
    int mask = 0;
#pragma unroll
    for (int k = levels; k < 6; ++k) mask |= 1 << step_bit(k);
    return (lane & mask) == 0;
}

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane);
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int lane) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// 1/sqrt(x) for x > 0 finite: the hardware estimate and two Newton steps
// (y <- y + y (1/2 - (x/2) y^2)), within a few ulp of the rounded value.
__device__ __forceinline__ double rsqrt_fp64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
#pragma unroll
    for (int i = 0; i < 2; ++i) y = __builtin_fma(y, __builtin_fma(-h * y, y, 0.5), y);
    return y;
}

// Orders LDS traffic between the lanes of one wavefront.  LDS instructions of
// a wave execute in order, so no hardware barrier is needed; the fences keep
// the compiler from moving memory operations across this point, and it must
// not be an s_barrier: the superleaf kernels run several waves that
// synchronise independently.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace tsqr
