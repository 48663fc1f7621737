// FP64 device linear algebra for B formation (see linalg.hip).
#pragma once
#include "common.hpp"

namespace nopt {
// C = alpha * A * op(B) + beta * C; op(B) = B (K x N, ldb) or, with transB,
// the transpose of B (N x K, ldb). Row-major.
hipError_t gemm_f64(bool transB, int M, int N, int K, double alpha, const double* A, long lda,
                    const double* B, long ldb, double beta, double* C, long ldc, hipStream_t s);
// In-place lower Cholesky A = L L^T (strict upper part zeroed). *d_info is
// set to j+1 for the first non-positive pivot j, else 0.
hipError_t cholesky_f64(double* A, long lda, int n, int* d_info, hipStream_t s);
// B (n x ncol) <- L^-1 B for lower-triangular L (n x n).
hipError_t trsm_lower_f64(const double* L, long ldl, int n, double* B, long ldb, long ncol,
                          hipStream_t s);
}  // namespace nopt
