// FP64 dense linear algebra for B formation on the device: a general MFMA
// GEMM, a blocked Cholesky factorisation of the Coulomb metric and the
// blocked forward substitution B = L^-1 (P|mn).
//
// These replace NOPT's metric route, S05_calc + dsyevd (src/matr.cpp:205-250,
// 372-401) and the per-build V^-1/2 products (src/RI.cpp:128-138, 152-156):
// the metric is factored once per geometry and folded into B.
#include "linalg.hpp"

namespace nopt {

// C = alpha * A * op(B) + beta * C, FP64, 64x64 tiles, 4 waves of 32x32.
template <bool TB>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, double alpha,
                                                  const double* __restrict__ A, long lda,
                                                  const double* __restrict__ B, long ldb,
                                                  double beta, double* __restrict__ C, long ldc) {
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64 + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  const long m0 = (long)blockIdx.y * 64, n0 = (long)blockIdx.x * 64;
  d4_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += 16) {
    {
      const int r = t >> 2, kk = (t & 3) * 4;
      const long gm = m0 + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int gk = k0 + kk + j;
        As[kk + j][r] = (gm < M && gk < K) ? A[gm * lda + gk] : 0.0;
      }
    }
    if (!TB) {
      const int kk = t >> 4, c = (t & 15) * 4;
      const int gk = k0 + kk;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long gn = n0 + c + j;
        Bs[kk][c + j] = (gk < K && gn < N) ? B[(long)gk * ldb + gn] : 0.0;
      }
    } else {
      const int c = t >> 2, kk = (t & 3) * 4;
      const long gn = n0 + c;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int gk = k0 + kk + j;
        Bs[kk + j][c] = (gn < N && gk < K) ? B[gn * ldb + gk] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[ks + lane / 16][wr * 32 + i * 16 + lane % 16];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[ks + lane / 16][wc * 32 + j * 16 + lane % 16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long row = m0 + wr * 32 + i * 16 + acc_row<double>(lane, v);
        const long col = n0 + wc * 32 + j * 16 + lane % 16;
        if (row < M && col < N) {
          double r = alpha * acc[i][j][v];
          if (beta != 0.0) r += beta * C[row * ldc + col];
          C[row * ldc + col] = r;
        }
      }
}

hipError_t gemm_f64(bool transB, int M, int N, int K, double alpha, const double* A, long lda,
                    const double* B, long ldb, double beta, double* C, long ldc,
                    hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  if (grid.y > 65535) return hipErrorInvalidValue;
  if (transB)
    k_gemm_f64<true><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else
    k_gemm_f64<false><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  return hipGetLastError();
}

// ---------------------------------------------------------------- Cholesky
constexpr int NB = 64;

// Factor the kb x kb diagonal block at (k0,k0) in place (lower), zeroing its
// strict upper part. A non-positive pivot sets *info = k0 + j + 1 (first one
// wins) and the pivot is replaced by 1 so that the kernels stay finite.
__global__ __launch_bounds__(256) void k_potrf_diag(double* A, long lda, int k0, int kb,
                                                    int* info) {
  __shared__ double L[NB][NB + 1];
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
    int i = e / kb, j = e % kb;
    L[i][j] = (j <= i) ? A[(long)(k0 + i) * lda + k0 + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    if (threadIdx.x == 0) {
      double d = L[j][j];
      if (!(d > 0.0)) {
        atomicCAS(info, 0, k0 + j + 1);
        d = 1.0;
      }
      L[j][j] = sqrt(d);
    }
    __syncthreads();
    const double piv = L[j][j];
    for (int i = j + 1 + threadIdx.x; i < kb; i += blockDim.x) L[i][j] /= piv;
    __syncthreads();
    // rank-1 update of the trailing lower triangle
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
      int i = e / kb, c = e % kb;
      if (c > j && i >= c) L[i][c] -= L[i][j] * L[c][j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
    int i = e / kb, j = e % kb;
    A[(long)(k0 + i) * lda + k0 + j] = (j <= i) ? L[i][j] : 0.0;
  }
}

// Rows i >= k0+kb of the panel: x L_kk^T = a, x overwrites a.
__global__ __launch_bounds__(256) void k_trsm_panel(double* A, long lda, int n, int k0, int kb) {
  __shared__ double L[NB][NB + 1];
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
    int i = e / kb, j = e % kb;
    L[i][j] = A[(long)(k0 + i) * lda + k0 + j];
  }
  __syncthreads();
  const long row = (long)k0 + kb + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double* a = A + row * lda + k0;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < kb) {
      double v = a[j];
#pragma unroll
      for (int l = 0; l < NB; ++l)
        if (l < j) v -= x[l] * L[j][l];
      x[j] = v / L[j][j];
      a[j] = x[j];
    }
  }
}

// Columns of B (rows k0..k0+kb): L_kk x = b, x overwrites b.
__global__ __launch_bounds__(256) void k_trsm_cols(const double* Lm, long ldl, double* B, long ldb,
                                                   long ncol, int k0, int kb) {
  __shared__ double L[NB][NB + 1];
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
    int i = e / kb, j = e % kb;
    L[i][j] = (j <= i) ? Lm[(long)(k0 + i) * ldl + k0 + j] : 0.0;
  }
  __syncthreads();
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < kb) {
      double v = B[(long)(k0 + j) * ldb + c];
#pragma unroll
      for (int l = 0; l < NB; ++l)
        if (l < j) v -= L[j][l] * x[l];
      x[j] = v / L[j][j];
      B[(long)(k0 + j) * ldb + c] = x[j];
    }
  }
}

__global__ void k_zero_upper(double* A, long lda, int n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const long i = e / n, j = e % n;
  if (j > i) A[i * lda + j] = 0.0;
}

__global__ void k_set_int(int* p, int v) { *p = v; }

hipError_t cholesky_f64(double* A, long lda, int n, int* d_info, hipStream_t s) {
  hipError_t e;
  k_set_int<<<1, 1, 0, s>>>(d_info, 0);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = n - k0 < NB ? n - k0 : NB;
    k_potrf_diag<<<1, 256, 0, s>>>(A, lda, k0, kb, d_info);
    const int rest = n - k0 - kb;
    if (rest > 0) {
      k_trsm_panel<<<(rest + 255) / 256, 256, 0, s>>>(A, lda, n, k0, kb);
      const double* P = A + (long)(k0 + kb) * lda + k0;
      double* T = A + (long)(k0 + kb) * lda + k0 + kb;
      e = gemm_f64(true, rest, rest, kb, -1.0, P, lda, P, lda, 1.0, T, lda, s);
      if (e != hipSuccess) return e;
    }
  }
  k_zero_upper<<<(int)(((long)n * n + 255) / 256), 256, 0, s>>>(A, lda, n);
  return hipGetLastError();
}

hipError_t trsm_lower_f64(const double* L, long ldl, int n, double* B, long ldb, long ncol,
                          hipStream_t s) {
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = n - k0 < NB ? n - k0 : NB;
    k_trsm_cols<<<(int)((ncol + 255) / 256), 256, 0, s>>>(L, ldl, B, ldb, ncol, k0, kb);
    const int rest = n - k0 - kb;
    if (rest > 0) {
      // B[k0+kb:, :] -= L[k0+kb:, k0:k0+kb] * B[k0:k0+kb, :], in column chunks
      // so that the grid stays within its limits
      const long chunk = 65535L * 64;
      for (long c0 = 0; c0 < ncol; c0 += chunk) {
        const int nc = (int)((ncol - c0) < chunk ? (ncol - c0) : chunk);
        hipError_t e = gemm_f64(false, rest, nc, kb, -1.0, L + (long)(k0 + kb) * ldl + k0, ldl,
                                B + (long)k0 * ldb + c0, ldb, 1.0, B + (long)(k0 + kb) * ldb + c0,
                                ldb, s);
        if (e != hipSuccess) return e;
      }
    }
  }
  return hipGetLastError();
}

}  // namespace nopt
