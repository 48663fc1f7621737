/* ORACLE -- test infrastructure only (tests/, smoke() and the CPU-baseline
 * leg of bench.py); the product never links or calls it.
 *
 * FP64 CPU restatement of NOPT's RI-J/K Fock build, kept in NOPT's own data
 * layout and operation order so that it can also serve as the timed CPU
 * baseline of "NOPT's CPU path":
 *   oracle_s05          S05_calc (src/matr.cpp:372-401) with lapack_diag =
 *                       dsyevd (src/matr.cpp:205-250): V^-1/2 and V^+1/2 by
 *                       eigendecomposition, eigenvalues <= 0 dropped
 *   oracle_par_dgemm    nopt_par_dgemm (include/blas_link.h:54-116): the M
 *                       rows split evenly over the OpenMP threads, a
 *                       single-threaded BLAS call per thread
 *   oracle_jk_nopt      DM_to_F_transform_RI (src/RI.cpp:115-176): gamma =
 *                       DM.RI_AA, two V^-1/2 products, J = RI_AA.c, the
 *                       half-transform MO.RI_AA, the metric product with
 *                       V^-1/2, the [i][m][Q] -> [m][i][Q] transpose and the
 *                       full (not symmetric-rank-k) K product
 *   oracle_jk_nopt_prepared  the same build with the metric folded into the
 *                       stored tensor once, RI_B = RI_AA.V^-1/2 (the CPU
 *                       "B-prepared" variant): gamma = DM.RI_B, J = RI_B.gamma,
 *                       the half-transform MO.RI_B, the transpose and the full
 *                       K product, i.e. src/RI.cpp:120-171 without the V^-1/2
 *                       products of :128-138 and :152-156
 *   oracle_dm           gen_HF_DM (src/SCF.cpp:63-73)
 * BLAS/LAPACK: Intel MKL through libmkl_rt (NOPT supports MKL, Makefile:22-26),
 * with prototypes declared here because no CBLAS/LAPACKE headers are present.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { RowMajor = 101, NoTrans = 111, Trans = 112 };
void cblas_dgemm(int layout, int ta, int tb, int M, int N, int K, double alpha, const double* A,
                 int lda, const double* B, int ldb, double beta, double* C, int ldc);
void dsyevd(const char* jobz, const char* uplo, const int* n, double* a, const int* lda, double* w,
            double* work, const int* lwork, int* iwork, const int* liwork, int* info);
int MKL_Set_Threading_Layer(int code);
void MKL_Set_Num_Threads(int n);
int MKL_Get_Max_Threads(void);

static int g_threads = 1;

/* threads used by the row split; MKL runs on the GNU OpenMP layer */
void oracle_set_threads(int n) {
  static int layer_set = 0;
  if (!layer_set) {
    MKL_Set_Threading_Layer(3); /* MKL_THREADING_GNU */
    layer_set = 1;
  }
  g_threads = n > 0 ? n : 1;
  omp_set_num_threads(g_threads);
  MKL_Set_Num_Threads(g_threads);
}

void oracle_par_dgemm(int ta, int tb, int M, int N, int K, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, double* C, int ldc) {
  if (M == 0 || N == 0) return;
  if (K == 0) {
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) C[(long)i * ldc + j] *= beta;
    return;
  }
  int nt = g_threads;
  int* idx = (int*)malloc(sizeof(int) * (nt + 1));
  for (int i = 0; i <= nt; ++i) idx[i] = (int)(((long)i * M) / nt);
  int saved = MKL_Get_Max_Threads();
  MKL_Set_Num_Threads(1);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    long inc_a = (ta == NoTrans) ? lda : 1;
    int rows = idx[t + 1] - idx[t];
    if (rows > 0)
      cblas_dgemm(RowMajor, ta, tb, rows, N, K, alpha, A + (long)idx[t] * inc_a, lda, B, ldb, beta,
                  C + (long)idx[t] * ldc, ldc);
  }
  MKL_Set_Num_Threads(saved);
  free(idx);
}

/* V^-1/2 (M) and V^+1/2 (P) of the symmetric n x n matrix S (S is
 * overwritten by its eigenvectors). Returns the number of dropped
 * eigenvalues. */
int oracle_s05(double* S, double* M, double* P, int n) {
  double* w = (double*)malloc(sizeof(double) * n);
  int lwork = 2 * n * n + 6 * n + 1, liwork = 5 * n + 3, info = 0;
  double* work = (double*)malloc(sizeof(double) * (size_t)lwork);
  int* iwork = (int*)malloc(sizeof(int) * liwork);
  dsyevd("V", "L", &n, S, &n, w, work, &lwork, iwork, &liwork, &info);
  free(work);
  free(iwork);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n * n);
  memset(M, 0, sizeof(double) * (size_t)n * n);
  memset(P, 0, sizeof(double) * (size_t)n * n);
  int dropped = 0;
  for (int i = 0; i < n; ++i) {
    if (w[i] > 0.0) {
      P[(long)i * (n + 1)] = sqrt(w[i]);
      M[(long)i * (n + 1)] = 1.0 / P[(long)i * (n + 1)];
    } else {
      ++dropped;
    }
  }
  /* row i of S is eigenvector i (column-major LAPACK on a row-major array) */
  cblas_dgemm(RowMajor, NoTrans, NoTrans, n, n, n, 1.0, P, n, S, n, 0.0, tmp, n);
  cblas_dgemm(RowMajor, Trans, NoTrans, n, n, n, 1.0, S, n, tmp, n, 0.0, P, n);
  cblas_dgemm(RowMajor, NoTrans, NoTrans, n, n, n, 1.0, M, n, S, n, 0.0, tmp, n);
  cblas_dgemm(RowMajor, Trans, NoTrans, n, n, n, 1.0, S, n, tmp, n, 0.0, M, n);
  free(tmp);
  free(w);
  return dropped;
}

/* DM = sum_{k<n_el} MO_k MO_k^T, rows of MO are orbitals. */
void oracle_dm(double* DM, const double* MO, int n, int n_el) {
  memset(DM, 0, sizeof(double) * (size_t)n * n);
  for (int k = 0; k < n_el; ++k)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) DM[(long)i * n + j] += MO[(long)k * n + i] * MO[(long)k * n + j];
}

/* J, K (n_ao x n_ao) from DM (n_ao x n_ao) and the first n_c rows of MO.
 * RI_AA: [(m*n_ao+n)*naux+P]; Vm05: naux x naux; RI_CA, RI_CA2: scratch of
 * n_c*n_ao*naux doubles each; RI_J, RI_J2: naux doubles each. */
int oracle_jk_nopt(double* J, double* K, const double* DM, const double* MO, long n_c, long n_ao,
                   long naux, const double* RI_AA, const double* Vm05, double* RI_CA,
                   double* RI_CA2, double* RI_J, double* RI_J2) {
  memset(RI_J, 0, sizeof(double) * naux);
  /* gamma: plain (BLAS-threaded) GEMM with M = 1 */
  cblas_dgemm(RowMajor, NoTrans, NoTrans, 1, (int)naux, (int)(n_ao * n_ao), 1.0, DM,
              (int)(n_ao * n_ao), RI_AA, (int)naux, 0.0, RI_J, (int)naux);
  oracle_par_dgemm(NoTrans, NoTrans, (int)naux, 1, (int)naux, 1.0, Vm05, (int)naux, RI_J, 1, 0.0,
                   RI_J2, 1);
  oracle_par_dgemm(NoTrans, NoTrans, (int)naux, 1, (int)naux, 1.0, Vm05, (int)naux, RI_J2, 1, 0.0,
                   RI_J, 1);
  oracle_par_dgemm(NoTrans, NoTrans, (int)(n_ao * n_ao), 1, (int)naux, 1.0, RI_AA, (int)naux, RI_J,
                   1, 0.0, J, 1);
  oracle_par_dgemm(NoTrans, NoTrans, (int)n_c, (int)(naux * n_ao), (int)n_ao, 1.0, MO, (int)n_ao,
                   RI_AA, (int)(naux * n_ao), 0.0, RI_CA, (int)(naux * n_ao));
  oracle_par_dgemm(NoTrans, NoTrans, (int)(n_c * n_ao), (int)naux, (int)naux, 1.0, RI_CA, (int)naux,
                   Vm05, (int)naux, 0.0, RI_CA2, (int)naux);
#pragma omp parallel for num_threads(g_threads)
  for (long b = 0; b < n_ao; ++b)
    for (long a = 0; a < n_c; ++a)
      memcpy(RI_CA + (b * n_c + a) * naux, RI_CA2 + (a * n_ao + b) * naux, sizeof(double) * naux);
  oracle_par_dgemm(NoTrans, Trans, (int)n_ao, (int)n_ao, (int)(naux * n_c), 1.0, RI_CA,
                   (int)(naux * n_c), RI_CA, (int)(naux * n_c), 0.0, K, (int)n_ao);
  return 1;
}

/* RI_B = RI_AA.V^-1/2 ([(m*n_ao+n)*naux+P], one row-split GEMM, once per
 * geometry) and the build over it: J = RI_B (RI_B^T DM), K = Y Y^T with
 * Y[m][i][P] = sum_n MO[i][n] RI_B[(n*n_ao+m)*naux+P] (RI_B is symmetric in
 * m, n). RI_CA: scratch of 2*n_c*n_ao*naux doubles; RI_J: naux doubles. */
void oracle_prepare_b(double* RI_B, const double* RI_AA, const double* Vm05, long n_ao, long naux) {
  oracle_par_dgemm(NoTrans, NoTrans, (int)(n_ao * n_ao), (int)naux, (int)naux, 1.0, RI_AA, (int)naux, Vm05,
                   (int)naux, 0.0, RI_B, (int)naux);
}

int oracle_jk_nopt_prepared(double* J, double* K, const double* DM, const double* MO, long n_c, long n_ao,
                            long naux, const double* RI_B, double* RI_CA, double* RI_J) {
  double* RI_CT = RI_CA + n_c * n_ao * naux;
  memset(RI_J, 0, sizeof(double) * naux);
  cblas_dgemm(RowMajor, NoTrans, NoTrans, 1, (int)naux, (int)(n_ao * n_ao), 1.0, DM, (int)(n_ao * n_ao), RI_B,
              (int)naux, 0.0, RI_J, (int)naux);
  oracle_par_dgemm(NoTrans, NoTrans, (int)(n_ao * n_ao), 1, (int)naux, 1.0, RI_B, (int)naux, RI_J, 1, 0.0, J, 1);
  oracle_par_dgemm(NoTrans, NoTrans, (int)n_c, (int)(naux * n_ao), (int)n_ao, 1.0, MO, (int)n_ao, RI_B,
                   (int)(naux * n_ao), 0.0, RI_CA, (int)(naux * n_ao));
#pragma omp parallel for num_threads(g_threads)
  for (long b = 0; b < n_ao; ++b)
    for (long a = 0; a < n_c; ++a)
      memcpy(RI_CT + (b * n_c + a) * naux, RI_CA + (a * n_ao + b) * naux, sizeof(double) * naux);
  oracle_par_dgemm(NoTrans, Trans, (int)n_ao, (int)n_ao, (int)(naux * n_c), 1.0, RI_CT, (int)(naux * n_c), RI_CT,
                   (int)(naux * n_c), 0.0, K, (int)n_ao);
  return 1;
}
