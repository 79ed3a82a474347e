// tsqr_compat.hpp -- the reference's C++ entry points over the C ABI.
//
// The reference declares header-only templates, explicitly instantiated for
// float and double:
//   template <typename T> void tsqr(int m, int n, T *A, int lda, T *R, int ldr,
//                                   T *d_work, int ldwork);   // TallShinnyQR.h:86-87
//   template <typename T> void qr(int m, int n, T *A, int lda, T *R, int ldr,
//                                 T *d_work, int ldwork);     // block_QR.h:72-73
// and its callers size the workspace with its own macros,
// ldwork = 2 * NUM_SM * BLOCK_SIZE (test_tsqr.cu:42, kernelQR.h:3-11).
// Including this header instead gives the same templates, forwarding to
// libtsqr_amd.so.  They return nothing, as the reference's do; a failing
// status is reported through tsqr_compat_error (default: print and abort,
// like the reference's assert / CUDA_CHECK), or can be read with
// tsqr_compat_last_status().  When the caller's ldwork is too small for the
// library, the library uses its own per-device buffer, so the reference's
// sizing keeps working for every m.
#pragma once
#include <cstdio>
#include <cstdlib>

#include "tsqr_amd.h"

#ifndef NUM_SM
#define NUM_SM 256  // MI355X compute units
#endif
#ifndef BLOCK_SIZE
#define BLOCK_SIZE 256  // rows of an FP32 Householder tile at n = 32
#endif

inline int& tsqr_compat_last_status() {
    static thread_local int s = TSQR_OK;
    return s;
}

inline void tsqr_compat_error(const char* what, int status) {
    std::fprintf(stderr, "%s: %s (status %d)\n", what, tsqr_status_string(status), status);
    std::abort();
}

inline void tsqr_compat_check(const char* what, int status) {
    tsqr_compat_last_status() = status;
    if (status != TSQR_OK) tsqr_compat_error(what, status);
}

template <typename T>
void tsqr(int m, int n, T* A, int lda, T* R, int ldr, T* d_work, int ldwork);
template <typename T>
void qr(int m, int n, T* A, int lda, T* R, int ldr, T* d_work, int ldwork);

template <>
inline void tsqr<float>(int m, int n, float* A, int lda, float* R, int ldr, float* d_work, int ldwork) {
    tsqr_compat_check("tsqr<float>", tsqr_s(m, n, A, lda, R, ldr, d_work, ldwork));
}
template <>
inline void tsqr<double>(int m, int n, double* A, int lda, double* R, int ldr, double* d_work, int ldwork) {
    tsqr_compat_check("tsqr<double>", tsqr_d(m, n, A, lda, R, ldr, d_work, ldwork));
}
template <>
inline void qr<float>(int m, int n, float* A, int lda, float* R, int ldr, float* d_work, int ldwork) {
    tsqr_compat_check("qr<float>", qr_s(m, n, A, lda, R, ldr, d_work, ldwork));
}
template <>
inline void qr<double>(int m, int n, double* A, int lda, double* R, int ldr, double* d_work, int ldwork) {
    tsqr_compat_check("qr<double>", qr_d(m, n, A, lda, R, ldr, d_work, ldwork));
}
