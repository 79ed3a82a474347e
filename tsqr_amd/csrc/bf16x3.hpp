// bf16x3.hpp -- FP32-accurate products on the BF16 matrix cores.
//
// An FP32 value x is split exactly into three BF16 pieces
//   x = h + m + l,  h = bf16(x), m = bf16(x - h), l = x - h - m
// (round to nearest each time; |m| <= 2^-9 |x|, |l| <= 2^-18 |x|, and l is
// exact in BF16 because x - h - m has at most 8 significant bits).  A
// product x*y is then the six leading terms
//   hh + hm + mh + hl + lh + mm          (dropped: ml, lm, ll < 2^-26 |xy|)
// each a BF16 x BF16 product, exact in FP32, summed by
// v_mfma_f32_16x16x32_bf16 in FP32.
//
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds A[i = l%16][k =
// 8(l/16) + s] and B[k = 8(l/16) + s][j = l%16] for s = 0..7 (one bf16x8
// each); the FP32 accumulator register v of lane l holds C[4(l/16) + v][l%16].
#pragma once
#include <hip/hip_runtime.h>

namespace tsqr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

struct Split3 {
    bf16x8 h, m, l;
};

__device__ __forceinline__ Split3 split3(const float (&x)[8]) {
    Split3 s;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const __bf16 h = (__bf16)x[t];
        const float r = x[t] - (float)h;
        const __bf16 m = (__bf16)r;
        s.h[t] = h;
        s.m[t] = m;
        s.l[t] = (__bf16)(r - (float)m);
    }
    return s;
}

// The same split of FP64 values (the pieces carry the leading 24 bits; the
// residuals are formed in FP64).
__device__ __forceinline__ Split3 split3(const double (&x)[8]) {
    Split3 s;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const __bf16 h = (__bf16)(float)x[t];
        const double r = x[t] - (double)(float)h;
        const __bf16 m = (__bf16)(float)r;
        s.h[t] = h;
        s.m[t] = m;
        s.l[t] = (__bf16)(float)(r - (double)(float)m);
    }
    return s;
}

__device__ __forceinline__ f4v mma_bf16(const bf16x8& a, const bf16x8& b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// acc += A^T B over 32 rows to ~FP32 accuracy, smallest terms first.
__device__ __forceinline__ f4v mma_bf16x3(const Split3& a, const Split3& b, f4v acc) {
    acc = mma_bf16(a.m, b.m, acc);
    acc = mma_bf16(a.h, b.l, acc);
    acc = mma_bf16(a.l, b.h, acc);
    acc = mma_bf16(a.h, b.m, acc);
    acc = mma_bf16(a.m, b.h, acc);
    return mma_bf16(a.h, b.h, acc);
}

}  // namespace tsqr
