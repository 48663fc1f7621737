// RI-J/K Fock build for gfx950: kernels, device context and the C entry
// points of include/nopt_jk.h.
//
// Replaces DM_to_F_transform_RI (src/RI.cpp:115-176). NOPT's sequence is
//   gamma = D.RI_AA; c = V^-1/2 V^-1/2 gamma; J = RI_AA.c;
//   X = MO.RI_AA; X~ = X.V^-1/2; transpose; K = Y.Y^T (full GEMM).
// Here the metric is folded into B once (B = L^-1 (P|mn), linalg.hip), so a
// build is
//   half-transform   X[Q][m][i] = sum_n B[Q][m][n] C[i][n]   (MFMA GEMM)
//                    with gamma_Q = sum_mn B[Q][m][n] D[m][n] fused into the
//                    same pass over B (per-tile partial sums)
//   Coulomb          J[m][n] = sum_Q gamma_Q B[Q][m][n]      (HBM-bound)
//   exchange         K[m][n] = sum_Q sum_i X[Q][m][i] X[Q][n][i]  (MFMA
//                    SYRK over lower-triangle tiles, split over Q)
//   finalize         deterministic reduction of the Q-split partials in FP64
//                    and symmetric write of J and K.
//
// HBM layout (row-major): B64[Q][m][ld64] FP64, full square with zero padding
// columns (ld64 a multiple of 16); the mixed-precision kernels round or split
// it on the fly; X[Q][m][ldx] in the build's
// precision; J/K partials FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../../include/nopt_jk.h"
#include "../host/errors.hpp"
#include "../host/order.hpp"
#include "common.hpp"
#include "bf16.hpp"
#include "linalg.hpp"
#include "ozaki.hpp"

using nopt::set_error;

namespace {

constexpr int BM = 128;     // half-transform rows per tile
constexpr int KT = 128;     // exchange tile edge
constexpr int NSPLIT_J_MIN = 16; // minimum Q-split of the Coulomb pass
constexpr int JCHUNK_MAX = 4096; // gamma values cached in LDS per split

// ------------------------------------------------------------------ copies
__global__ void k_pack_B(const double* __restrict__ B64, long ld64, int nbf, long npair, long naux,
                         double* __restrict__ Bp) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= naux * npair) return;
  const long Q = e / npair, p = e % npair;
  long a = (long)((std::sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= p) ++a;
  while (a * (a + 1) / 2 > p) --a;
  const long b = p - a * (a + 1) / 2;
  Bp[e] = B64[(Q * nbf + a) * ld64 + b];
}

// ----------------------------------------------------- active-space integrals
// R[Q][a2*na + a1] = sum_m C1[a1][m] sum_n B_Q[m][n] C2[a2][n] (C1, C2: na x
// nbf, orbitals as rows). One block per Q; each wave takes rows m = w, w+4, ...
// with its lanes over n, sums the row's na products across the wave (butterfly,
// then lane 0's value for all lanes, so every lane holds the same sum), and
// each lane adds its share of the na x na outer product with column m of C1.
// The four waves' partial R are added in a fixed order.
template <int NA>
__global__ __launch_bounds__(256) void k_active_half(const double* __restrict__ B64, long ld64, int nbf,
                                                     const double* __restrict__ C1, const double* __restrict__ C2,
                                                     int na, double* __restrict__ R) {
  constexpr int PER = (NA * NA + 63) / 64;
  __shared__ double part[4][PER * 64];
  const int Q = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double* Bq = B64 + (long)Q * nbf * ld64;
  double r[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) r[k] = 0.0;
  for (int m = w; m < nbf; m += 4) {
    double u[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) u[a] = 0.0;
    for (int n = lane; n < nbf; n += 64) {
      const double b = Bq[(long)m * ld64 + n];
#pragma unroll
      for (int a = 0; a < NA; ++a)
        if (a < na) u[a] += b * C2[(long)a * nbf + n];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      double v = u[a];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      u[a] = __shfl(v, 0, 64);
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = lane + 64 * k, a2 = e / NA, a1 = e % NA;
      if (a2 < na && a1 < na) {
        double ua = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a) ua = (a == a2) ? u[a] : ua;  // register select, no dynamic index
        r[k] += C1[(long)a1 * nbf + m] * ua;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) part[w][lane + 64 * k] = r[k];
  __syncthreads();
  for (int e = threadIdx.x; e < na * na; e += 256) {
    const int i = (e / na) * NA + e % na;
    R[(long)Q * na * na + e] = ((part[0][i] + part[1][i]) + part[2][i]) + part[3][i];
  }
}

// ints[e1][e2] = sum_Q R[Q][e1] R[Q][e2], e = a2*na + a1 (nn = na*na).
__global__ void k_active_ints(const double* __restrict__ R, long naux, int nn, double* __restrict__ ints) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nn * nn) return;
  const int e1 = (int)(idx / nn), e2 = (int)(idx % nn);
  double s = 0.0;
  for (long Q = 0; Q < naux; ++Q) s += R[Q * nn + e1] * R[Q * nn + e2];
  ints[idx] = s;
}

// dst[r][c] (ldd) = src[r'][c'] (lds) for r < rows, c < cols; zero elsewhere
// in [0, rows_pad) x [0, ldd); r' = rperm[r], c' = cperm[c] (each optional)
template <typename T>
__device__ __forceinline__ void k_pad_copy_body(const VB vb, const double* __restrict__ src, long lds, int rows,
                                                int cols, T* __restrict__ dst, long ldd, int rows_pad,
                                                const int* __restrict__ rperm, const int* __restrict__ cperm) {
  const long e = (long)vb.x * blockDim.x + threadIdx.x;
  if (e >= (long)rows_pad * ldd) return;
  const long r = e / ldd, c = e % ldd;
  double v = (r < rows && c < cols) ? src[(rperm ? rperm[r] : r) * lds + (cperm ? cperm[c] : c)] : 0.0;
  dst[e] = (T)v;
}
template <typename T>
__global__ void k_pad_copy(const double* __restrict__ src, long lds, int rows, int cols,
                           T* __restrict__ dst, long ldd, int rows_pad, const int* __restrict__ rperm,
                           const int* __restrict__ cperm) {
  k_pad_copy_body<T>(grid_vb(), src, lds, rows, cols, dst, ldd, rows_pad, rperm, cperm);
}
// work-list form (FP64 copies)
struct ItemPad {
  const double* src;
  long lds;
  int rows;
  int cols;
  double* dst;
  long ldd;
  int rows_pad;
  const int* rperm;
  const int* cperm;
  int gx, gy, gz;
};
__global__ void k_pad_copy_list(const ItemPad* __restrict__ items, const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemPad& a = items[i];
  k_pad_copy_body<double>(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz), a.src, a.lds, a.rows, a.cols, a.dst,
                          a.ldd, a.rows_pad, a.rperm, a.cperm);
}

// --------------------------------------------------- half-transform f64/f32
// X[Q][m][col] = sum_n B[Q][m][n] Cp[col][n] for one (m-tile, Q, column group)
// and gpart[Q][mtile] = sum over the tile of B[Q][m][n] D[m][n].
template <typename T>
struct Vec8 {
  T v[8];
};

template <typename T>
__device__ __forceinline__ void load8(const T* p, T* out) {
  if constexpr (sizeof(T) == 8) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double2 x = q[i];
      out[2 * i] = x.x;
      out[2 * i + 1] = x.y;
    }
  } else {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      float4 x = q[i];
      out[4 * i] = x.x;
      out[4 * i + 1] = x.y;
      out[4 * i + 2] = x.z;
      out[4 * i + 3] = x.w;
    }
  }
}

// B is always read in FP64; the FP32 variant rounds it while staging into LDS,
// so the fused gamma partials are FP64 in both variants.
template <typename T, int NT>
__global__ __launch_bounds__(256) void k_half(const double* __restrict__ B, long ld, int nbf,
                                              const T* __restrict__ Cp, const double* __restrict__ Dp,
                                              T* __restrict__ X, long ldx, double* __restrict__ gpart,
                                              int mtiles) {
  constexpr int BN = 16 * NT, BK = 16;
  using acc_t = typename std::conditional<sizeof(T) == 8, d4_t, f4_t>::type;
  __shared__ T As[BK][BM + 2];
  __shared__ T Cs[BK][BN + 2];
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int mtile = blockIdx.x;
  const long Q = blockIdx.y;
  const int cg = blockIdx.z;
  const int m0 = mtile * BM;
  const int ar = t >> 1, ak = (t & 1) * 8;
  const bool arow_ok = (m0 + ar) < nbf;
  const double* Arow = B + ((long)Q * nbf + m0 + ar) * ld + ak;
  const double* Drow = Dp + (long)(m0 + ar) * ld + ak;
  const bool c_ok = t < 2 * BN;
  const T* Crow = Cp + ((long)cg * BN + (t >> 1)) * ld + ak;
  acc_t acc[2][NT];
  d4_t acc64[2][NT];  // used by the FP32 variant only
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      acc[i][j] = acc_t{0, 0, 0, 0};
      acc64[i][j] = d4_t{0, 0, 0, 0};
    }
  double rd = 0.0;
  double a_reg[8];
  T c_reg[8];
  double d_reg[8];
  auto gload = [&](int k0) {
    if (arow_ok) {
      load8<double>(Arow + k0, a_reg);
      const double2* dq = reinterpret_cast<const double2*>(Drow + k0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double2 x = dq[i];
        d_reg[2 * i] = x.x;
        d_reg[2 * i + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        a_reg[i] = 0;
        d_reg[i] = 0;
      }
    }
    if (c_ok) load8<T>(Crow + k0, c_reg);
  };
  gload(0);
  for (int k0 = 0; k0 < ld; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      As[ak + i][ar] = (T)a_reg[i];
      rd += a_reg[i] * d_reg[i];
    }
    if (c_ok) {
#pragma unroll
      for (int i = 0; i < 8; ++i) Cs[ak + i][t >> 1] = c_reg[i];
    }
    __syncthreads();
    if (k0 + BK < ld) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T a[2], b[NT];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[ks + lane / 16][w * 32 + i * 16 + lane % 16];
#pragma unroll
      for (int j = 0; j < NT; ++j) b[j] = Cs[ks + lane / 16][j * 16 + lane % 16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          if constexpr (sizeof(T) == 8)
            acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
          else
            acc[i][j] = mfma_f32(a[i], b[j], acc[i][j]);
        }
    }
    if constexpr (sizeof(T) == 4) {
      // FP32 chains of BK terms (rounded to nearest by the MFMA), accumulated
      // across chains in FP64
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
#pragma unroll
          for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)acc[i][j][v];
          acc[i][j] = acc_t{0, 0, 0, 0};
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + w * 32 + i * 16 + acc_row<T>(lane, v);
      if (m < nbf) {
        T* xr = X + ((long)Q * nbf + m) * ldx + (long)cg * BN + lane % 16;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          if constexpr (sizeof(T) == 4)
            xr[j * 16] = (T)acc64[i][j][v];
          else
            xr[j * 16] = acc[i][j][v];
        }
      }
    }
  if (cg == 0) {
    double s = block_sum256(rd, red);
    if (t == 0) gpart[Q * mtiles + mtile] = s;
  }
}

// ------------------------------------------------------------------ gamma
// gamma[d][Q] = sum_mn B[Q][m][n] D_d[m][n] for ND densities (D_d at
// Dp + d*dstride) over the lower triangle of B (off-diagonal weight 2): one
// block per auxiliary function, row pairs (p, nbf-1-p) in 8-column chunks, so
// B is read once for all ND densities.
template <int ND>
__global__ __launch_bounds__(256) void k_gamma_f64(const double* __restrict__ B, long ld, int nbf,
                                                   const double* __restrict__ Dp, long ldd, long dstride,
                                                   double* __restrict__ gamma, long gstride) {
  __shared__ double red[4];
  const long Q = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = 0.0;
  const int npairs = (nbf + 1) / 2;
  for (int p = w; p < npairs; p += 4) {
    const int m1 = p, m2 = nbf - 1 - p;
    const int c1 = m1 / 8 + 1, c2 = (m2 != m1) ? m2 / 8 + 1 : 0;
    for (int ch = lane; ch < c1 + c2; ch += 64) {
      const int m = ch < c1 ? m1 : m2;
      const int n0 = 8 * (ch < c1 ? ch : ch - c1);
      double bv[8];
      load8<double>(B + ((long)Q * nbf + m) * ld + n0, bv);
#pragma unroll
      for (int j = 0; j < 8; ++j) bv[j] *= (n0 + j < m) ? 2.0 : (n0 + j == m ? 1.0 : 0.0);
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double dr[8];
        load8<double>(Dp + d * dstride + (long)m * ldd + n0, dr);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[d] += bv[j] * dr[j];
      }
    }
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double tot = block_sum256(acc[d], red);
    if (threadIdx.x == 0) gamma[d * gstride + Q] = tot;
  }
}

__global__ void k_gamma(const double* __restrict__ gpart, int mtiles, int naux,
                        double* __restrict__ gamma) {
  const int Q = blockIdx.x * blockDim.x + threadIdx.x;
  if (Q >= naux) return;
  double s = 0.0;
  for (int i = 0; i < mtiles; ++i) s += gpart[(long)Q * mtiles + i];
  gamma[Q] = s;
}

// ---------------------------------------------------------------- Coulomb
// Jpart[d][s][m][n] = sum_{Q in split s} gamma[d][Q] B[Q][m][n] over the lower
// triangle, for ND densities at once (B is read once). A block owns the row
// pair (p, nbf-1-p): nbf+1 elements, read as aligned pairs of columns (a
// row's last pair may include its first upper element, which lies inside the
// row because ld > nbf and is not written). The Q loop is unrolled so that UNR
// loads per thread are in flight.
template <int ND>
__global__ __launch_bounds__(256) void k_jpart(const double* __restrict__ B, long ld, int nbf, int naux,
                                               const double* __restrict__ gamma, long gstride,
                                               double* __restrict__ Jpart, long jstride) {
  constexpr int UNR = ND == 1 ? 8 : 4;
  __shared__ double gs[JCHUNK_MAX];  // the host keeps ND * (q1 - q0) <= JCHUNK_MAX
  const int p = blockIdx.x, s = blockIdx.y, ns = gridDim.y;
  const int q0 = (int)((long)naux * s / ns), q1 = (int)((long)naux * (s + 1) / ns);
  const int len = q1 - q0;
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x)
#pragma unroll
    for (int d = 0; d < ND; ++d) gs[d * len + q - q0] = gamma[d * gstride + q];
  __syncthreads();
  auto g = [&](int d, int i) { return gs[d * len + i]; };
  const int m1 = p, m2 = nbf - 1 - p;
  const int s1 = (m1 + 2) / 2, s2 = (m2 != m1) ? (m2 + 2) / 2 : 0;
  const long qstride = (long)nbf * ld;
  for (int slot = threadIdx.x; slot < s1 + s2; slot += blockDim.x) {
    const int row = slot < s1 ? m1 : m2;
    const int col = 2 * (slot < s1 ? slot : slot - s1);
    const double* bp = B + (long)q0 * qstride + (long)row * ld + col;
    double a0[ND], a1[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) a0[d] = a1[d] = 0.0;
    int q = q0;
    for (; q + UNR <= q1; q += UNR) {
      double2 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) v[u] = *reinterpret_cast<const double2*>(bp + (long)u * qstride);
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          a0[d] += g(d, q - q0 + u) * v[u].x;
          a1[d] += g(d, q - q0 + u) * v[u].y;
        }
      bp += (long)UNR * qstride;
    }
    for (; q < q1; ++q) {
      const double2 v = *reinterpret_cast<const double2*>(bp);
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        a0[d] += g(d, q - q0) * v.x;
        a1[d] += g(d, q - q0) * v.y;
      }
      bp += qstride;
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double* Js = Jpart + d * jstride + (long)s * nbf * nbf;
      Js[(long)row * nbf + col] = a0[d];
      if (col + 1 <= row) Js[(long)row * nbf + col + 1] = a1[d];
    }
  }
}

// --------------------------------------------------------------- exchange
__device__ __forceinline__ void tri_tile(int t, int& bi, int& bj) {
  bi = (int)((std::sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  while (bi * (bi + 1) / 2 > t) --bi;
  bj = t - bi * (bi + 1) / 2;
}

// Kpart[s][tile][r][c] = sum_{Q in split s} sum_i X[Q][bi*KT+r][i] X[Q][bj*KT+c][i]
template <typename T>
__global__ __launch_bounds__(256) void k_syrk(const T* __restrict__ X, long ldx, int nbf, int naux,
                                              int kpad, double* __restrict__ Kpart) {
  constexpr int BK = 16;
  using acc_t = typename std::conditional<sizeof(T) == 8, d4_t, f4_t>::type;
  __shared__ T As[BK][KT + 2];
  __shared__ T Bs[BK][KT + 2];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  int bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  const int s = blockIdx.y, ns = gridDim.y;
  const int q0 = (int)((long)naux * s / ns), q1 = (int)((long)naux * (s + 1) / ns);
  const int r = t >> 1, kk = (t & 1) * 8;
  const bool ra = bi * KT + r < nbf, rb = bj * KT + r < nbf;
  acc_t acc[4][4];
  d4_t acc64[4][4];  // used by the FP32 variant only
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = acc_t{0, 0, 0, 0};
      acc64[i][j] = d4_t{0, 0, 0, 0};
    }
  T a_reg[8], b_reg[8];
  auto gload = [&](int q, int i0) {
    const T* xa = X + ((long)q * nbf + bi * KT + r) * ldx + i0 + kk;
    const T* xb = X + ((long)q * nbf + bj * KT + r) * ldx + i0 + kk;
    if (ra)
      load8<T>(xa, a_reg);
    else
#pragma unroll
      for (int i = 0; i < 8; ++i) a_reg[i] = 0;
    if (rb)
      load8<T>(xb, b_reg);
    else
#pragma unroll
      for (int i = 0; i < 8; ++i) b_reg[i] = 0;
  };
  if (q0 < q1) gload(q0, 0);
  for (int q = q0; q < q1; ++q) {
    for (int i0 = 0; i0 < kpad; i0 += BK) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        As[kk + i][r] = a_reg[i];
        Bs[kk + i][r] = b_reg[i];
      }
      __syncthreads();
      int nq = q, ni = i0 + BK;
      if (ni >= kpad) {
        ni = 0;
        ++nq;
      }
      if (nq < q1) gload(nq, ni);
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[ks + lane / 16][wr * 64 + i * 16 + lane % 16];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[ks + lane / 16][wc * 64 + j * 16 + lane % 16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if constexpr (sizeof(T) == 8)
              acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
            else
              acc[i][j] = mfma_f32(a[i], b[j], acc[i][j]);
          }
      }
      if constexpr (sizeof(T) == 4) {
        // FP32 chains of BK terms, accumulated in FP64
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)acc[i][j][v];
            acc[i][j] = acc_t{0, 0, 0, 0};
          }
      }
    }
  }
  double* Kt = Kpart + ((long)s * gridDim.x + blockIdx.x) * KT * KT;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = wr * 64 + i * 16 + acc_row<T>(lane, v);
        const int cc = wc * 64 + j * 16 + lane % 16;
        if constexpr (sizeof(T) == 4)
          Kt[rr * KT + cc] = acc64[i][j][v];
        else
          Kt[rr * KT + cc] = acc[i][j][v];
      }
}

// --------------------------------------------------------------- finalize
// J = sum_s Jpart[s], K = sum_s Kpart[s] over the lower triangle, written to
// both triangles of J, K (ld nbf). Either output may be null. With jlive (the
// integer Coulomb pair), J is zero in the 8-column slots (m, n/8) whose jlive
// byte is 0: B is zero there for every Q, and the J kernels skip them.
struct ItemFin {
  const double* Jpart;
  int nsj;
  const uint8_t* jlive;
  int nsr;
  const double* Kpart;
  int nsk;
  int ntiles;
  int kt;
  int nbf;
  const int* perm;
  double* J;
  double* K;
  int gx, gy, gz;
};
__device__ __forceinline__ void k_finalize_body(const VB vb, const double* __restrict__ Jpart, int nsj,
                                                const uint8_t* __restrict__ jlive, int nsr,
                                                const double* __restrict__ Kpart, int nsk, int ntiles, int kt,
                                                int nbf, const int* __restrict__ perm, double* __restrict__ J,
                                                double* __restrict__ K) {
  const int KT = kt;  // exchange tile edge of the partials
  const long e = (long)vb.x * blockDim.x + threadIdx.x;
  if (e >= (long)nbf * nbf) return;
  const int m = (int)(e / nbf), n = (int)(e % nbf);
  if (n > m) return;
  // partials in the order of the integer forms (perm[new] = old)
  const long pm = perm ? perm[m] : m, pn = perm ? perm[n] : n;
  if (J) {
    double s = 0.0;
    if (!jlive || jlive[(long)m * nsr + n / 8])
      for (int k = 0; k < nsj; ++k) s += Jpart[((long)k * nbf + m) * nbf + n];
    J[pm * nbf + pn] = s;
    J[pn * nbf + pm] = s;
  }
  if (K) {
    const int bi = m / KT, bj = n / KT;
    const int tile = bi * (bi + 1) / 2 + bj;
    const int r = m % KT, c = n % KT;
    double s = 0.0;
    for (int k = 0; k < nsk; ++k) s += Kpart[(((long)k * ntiles + tile) * KT + r) * KT + c];
    K[pm * nbf + pn] = s;
    K[pn * nbf + pm] = s;
  }
}
__global__ void k_finalize(const double* __restrict__ Jpart, int nsj, const uint8_t* __restrict__ jlive, int nsr,
                           const double* __restrict__ Kpart, int nsk, int ntiles, int kt, int nbf,
                           const int* __restrict__ perm, double* __restrict__ J, double* __restrict__ K) {
  k_finalize_body(grid_vb(), Jpart, nsj, jlive, nsr, Kpart, nsk, ntiles, kt, nbf, perm, J, K);
}
__global__ void k_finalize_list(const ItemFin* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemFin& a = items[i];
  k_finalize_body(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.Jpart, a.nsj, a.jlive, a.nsr, a.Kpart, a.nsk, a.ntiles, a.kt, a.nbf, a.perm, a.J, a.K);
}


// ----------------------------------------------------------- Fock, energy
__global__ void k_fock_energy(const double* __restrict__ H, const double* __restrict__ D,
                              const double* __restrict__ J, const double* __restrict__ K, double cj,
                              double ck, long nn, double* __restrict__ F,
                              double* __restrict__ epart) {
  __shared__ double red[4];
  double acc = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (long)gridDim.x * blockDim.x) {
    const double h = H[e];
    double f = h;
    if (J) f += cj * J[e];
    if (K) f -= ck * K[e];
    if (F) F[e] = f;
    acc += D[e] * (f + h);
  }
  double s = block_sum256(acc, red);
  if (threadIdx.x == 0) epart[blockIdx.x] = s;
}

__global__ void k_sum(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += v[i];
  double s = block_sum256(acc, red);
  if (threadIdx.x == 0) *out = s;
}

// D[m][n] = sum_i C[i][m] C[i][n]
__global__ void k_density(const double* __restrict__ C, int nocc, int nbf, double* __restrict__ D) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nbf * nbf) return;
  const int m = (int)(e / nbf), n = (int)(e % nbf);
  double s = 0.0;
  for (int i = 0; i < nocc; ++i) s += C[(long)i * nbf + m] * C[(long)i * nbf + n];
  D[e] = s;
}

}  // namespace

// ================================================================ context
namespace {
constexpr int kMaxBatch = NOPT_MAX_BATCH;  // densities per batched build
// largest full-row digit planes of B a context keeps (larger: the packed
// triangle alone feeds the half-transform)
constexpr size_t kFullRowsMax = (size_t)24 << 30;
}  // namespace

struct nopt_jk_ctx {
  int device = 0;
  int ncu = 256;  // compute units of the device
  long nbf = 0, naux = 0, npair = 0;
  long ld64 = 0;
  int mtiles = 0, ktn = 0, ntiles = 0;
  hipStream_t stream = nullptr;
  // second stream for the bandwidth-bound Coulomb pair, which runs beside the
  // MFMA-bound exchange; fork/join events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_half = nullptr, ev_join = nullptr;
  double* B64 = nullptr;
  int flags = 0;        // NOPT_CTX_*: which forms of B the context keeps
  bool b_ready = false; // B formed
  // per-build work buffers (grown on demand)
  long cap_ldx = 0;
  void* X = nullptr;  // naux*nbf*ldx elements of the largest type
  size_t X_bytes = 0;
  double *gpart = nullptr, *gamma = nullptr, *Kpart = nullptr;
  void* Jpart = nullptr;  // J partials: (densities x splits) x nbf x nbf
  size_t Jpart_bytes = 0;
  int nsk = 0;
  int nsj = 0;  // Q-splits of the FP64 J partials (0: the smallest the LDS gamma cache allows, at least 16)
  size_t Kpart_bytes = 0;
  double* Dp = nullptr;  // padded densities, kMaxBatch slots of dstride
  long dstride = 0;
  void* Cp = nullptr;    // padded orbital factor
  size_t Cp_bytes = 0;
  // host-staging slots of the host-memory entry points (slot 0 also serves
  // the Fock/energy helpers)
  double *dJ = nullptr, *dK = nullptr, *dD = nullptr, *dC = nullptr, *dH = nullptr,
         *dF = nullptr, *dG = nullptr, *epart = nullptr, *dE = nullptr;
  long dC_cap = 0;
  double *xD[kMaxBatch - 1] = {}, *xC[kMaxBatch - 1] = {}, *xJ[kMaxBatch - 1] = {}, *xK[kMaxBatch - 1] = {};
  long xD_cap[kMaxBatch - 1] = {}, xC_cap[kMaxBatch - 1] = {}, xJ_cap[kMaxBatch - 1] = {},
       xK_cap[kMaxBatch - 1] = {};
  int* info = nullptr;
  // integer forms of B: the row exponents (Be, and as bytes Be8[Q*ld8 + n]),
  // the packed lower triangle Bs below, and, when bfull, the full rows as
  // four digit planes (4 x naux*nbf x ld8) for the half-transform; the
  // per-build digit planes of the orbital factor (4 x ldx x ld8)
  long ld8 = 0;
  int mtiles64 = 0;
  bool bfull = false;
  int8_t* Bd = nullptr;
  int* Be = nullptr;
  int8_t* Be8 = nullptr;
  void* Cd = nullptr;
  size_t Cd_bytes = 0;
  void* Ce = nullptr;
  size_t Ce_bytes = 0;
  void* Xd = nullptr;  // 4 digit planes of X (naux*nbf rows x ldk)
  size_t Xd_bytes = 0;
  void* Xe = nullptr;
  size_t Xe_bytes = 0;
  int ntiles8 = 0, nsk8 = 0;  // 64x64 exchange tiles of the integer path
  // the lower triangle of B as packed rows in 16-column segments (N28 as
  // int32, then d4 as bytes: 5*splane bytes), read by the Coulomb pair and,
  // without bfull, by the half-transform; the segment and slot tables, gamma
  // partials per slot block, Q-splits of the gamma and J kernels
  int8_t* Bs = nullptr;
  long splane = 0;
  int nseg = 0;
  int* segs = nullptr;   // segment table (m << 16 | n0), used to pack Bs
  int* slots = nullptr;  // every 8-column slot (m << 16 | n0)
  int nslots = 0, nsg = 0, nsj8 = 0;
  // what the kernels walk: the segments and slots in which some B_Q holds a
  // nonzero value (all of them until B is formed), jlive[m*nsr + n/8] = 1
  // for a live slot; and the half-transform's slab masks (naux x nmb words)
  int* lsegs = nullptr;
  int* lslots = nullptr;
  int nlseg = 0, nlslots = 0, nsr = 0;
  uint8_t* jlive = nullptr;
  uint32_t* Bm = nullptr;
  int nmb = 0;
  // the order of the functions in B's integer forms (band_order of B's
  // weights, so that exact-zero blocks gather): perm[new] = old, pinv its
  // inverse; the integer builds read D and C and write J and K through it
  int* perm = nullptr;
  int* pinv = nullptr;
  // when the integer Coulomb pair starts on the side stream: 0 after the
  // half-transform; 1 gamma before it (beside it), J after it; 2 both before
  int cplace = 0;
  double* gslot = nullptr;
  // optional per-phase timing: events for up to kTimedBuilds builds
  bool timing = false;
  int timed = 0;
  std::vector<hipEvent_t> tev;
  // completion of the last work enqueued on the context's buffers, on
  // whatever stream it ran: the next call waits for it on its own stream
  hipEvent_t ev_done = nullptr;
  bool done_set = false;
  // HIP graphs that hold the addresses of the work buffers: while any exists
  // the buffers keep their addresses (no growth, no tuning change) and
  // destroy is deferred to the last graph's destruction
  int pins = 0;
  bool zombie = false;
};

namespace {
constexpr int kPhases = 6;  // staging, half-transform, gamma, J partials, exchange, J output
constexpr int kTimedBuilds = 512;
}  // namespace

namespace {

const char* hip_msg(hipError_t e) { return hipGetErrorString(e); }

#define HIPCHK(call)                                                                       \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess)                                                                  \
      return set_error(_e == hipErrorOutOfMemory ? NOPT_ENOMEM : NOPT_EHIP,                \
                       std::string(#call) + ": " + hip_msg(_e));                          \
  } while (0)

// returns a HIP error from a helper that itself returns hipError_t
#define CHK_(call)                    \
  do {                                \
    hipError_t e_ = (call);           \
    if (e_ != hipSuccess) return e_;  \
  } while (0)

int dev_alloc(void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(NOPT_ENOMEM, std::string("device allocation of ") + std::to_string(bytes) +
                                      " bytes for " + what + " failed: " + hip_msg(e));
  }
  return NOPT_OK;
}

template <typename P>
void dev_free(P*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Grows a device buffer to at least `bytes`.
// set between nopt_graph_capture_begin and _end on this thread: device buffers
// may not grow while a stream is being captured
thread_local bool t_capturing = false;
// contexts whose builds were recorded by the capture in progress on this thread
thread_local std::vector<nopt_jk_ctx*> t_captured;

int ensure(const nopt_jk_ctx* c, void** p, size_t* cap, size_t bytes, const char* what) {
  if (*cap >= bytes && *p) return NOPT_OK;
  if (t_capturing)
    return set_error(NOPT_ESTATE, std::string("the ") + what +
                                      " buffer would grow during graph capture; run one build of the same shape "
                                      "before capturing");
  if (c->pins > 0)
    return set_error(NOPT_ESTATE, std::string("the ") + what +
                                      " buffer would grow while a captured HIP graph refers to it; destroy the "
                                      "graph first");
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  int rc = dev_alloc(p, bytes, what);
  if (rc == NOPT_OK) *cap = bytes;
  return rc;
}

// Column tiling of the integer half-transform: 16*nt columns per group,
// nt in {2, 4, 6}, with the least padding.
int nt_for_i8(long ncol, int* ngroups) {
  const long t = std::max(1L, (ncol + 15) / 16);
  if (t <= 6) {
    *ngroups = 1;
    return t <= 2 ? 2 : t <= 4 ? 4 : 6;
  }
  const long g6 = (t + 5) / 6, g4 = (t + 3) / 4;
  if (4 * g4 < 6 * g6) {
    *ngroups = (int)g4;
    return 4;
  }
  *ngroups = (int)g6;
  return 6;
}

int nt_for(long nocc, int* ngroups) {
  const long t = (nocc + 15) / 16;
  int nt = t <= 1 ? 1 : t <= 2 ? 2 : t <= 4 ? 4 : t <= 6 ? 6 : 8;
  *ngroups = (int)((t + nt - 1) / nt);
  return nt;
}

template <typename T>
hipError_t launch_half(int nt, dim3 grid, hipStream_t s, const double* B, long ld, int nbf, const T* Cp,
                       const double* Dp, T* X, long ldx, double* gpart, int mtiles) {
#define NOPT_HALF(NTV) \
  case NTV:            \
    k_half<T, NTV><<<grid, 256, 0, s>>>(B, ld, nbf, Cp, Dp, X, ldx, gpart, mtiles); \
    break;
  switch (nt) {
    NOPT_HALF(1)
    NOPT_HALF(2)
    NOPT_HALF(4)
    NOPT_HALF(6)
    NOPT_HALF(8)
    default:
      return hipErrorInvalidValue;
  }
#undef NOPT_HALF
  return hipGetLastError();
}

int check_ctx(nopt_jk_ctx* c, bool need_B) {
  if (!c) return set_error(NOPT_EINVAL, "null context");
  if (need_B && !c->b_ready) return set_error(NOPT_ESTATE, "B tensor not set (call nopt_jk_set_B first)");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return set_error(NOPT_EHIP, std::string("hipSetDevice: ") + hip_msg(e));
  return NOPT_OK;
}

// Every call that enqueues work on a context's buffers waits, on its own
// stream, for the last work enqueued on them (order_in) and records its own
// completion (order_out), so that builds on different streams (the default
// stream, torch's stream, the context's own) run one after the other instead
// of sharing the work buffers. Nothing is recorded or awaited while a capture
// is in progress on this thread; the capture notes the context instead, and
// the graph's launch does the ordering.
int order_in(nopt_jk_ctx* c, hipStream_t s) {
  if (t_capturing || !c->done_set) return NOPT_OK;
  HIPCHK(hipStreamWaitEvent(s, c->ev_done, 0));
  return NOPT_OK;
}

int order_out(nopt_jk_ctx* c, hipStream_t s) {
  if (t_capturing) {
    if (std::find(t_captured.begin(), t_captured.end(), c) == t_captured.end()) t_captured.push_back(c);
    return NOPT_OK;
  }
  HIPCHK(hipEventRecord(c->ev_done, s));
  c->done_set = true;
  return NOPT_OK;
}

// Q-splits of the integer Coulomb pair: about four blocks per CU over the
// slots the kernels walk (at most 32 J partials, which the output sums).
void coulomb_split_defaults(nopt_jk_ctx* c) {
  const int nsb = nopt::gamma_slot_blocks(c->nlslots);
  const int nbj = (c->nlslots + 256 * nopt::kJSPL - 1) / (256 * nopt::kJSPL);
  c->nsg = (int)std::max(1L, std::min(c->naux, (4L * c->ncu + nsb - 1) / nsb));
  c->nsj8 = (int)std::max(1L, std::min({c->naux, 32L, (4L * c->ncu + nbj - 1) / nbj}));
}

// Once B's integer forms are written: the Coulomb segments and slots that
// hold a nonzero value for some Q (the tables the kernels walk, and the J
// map of the finalize), the half-transform's slab masks, and the Coulomb
// split counts for the live tables.
int find_live(nopt_jk_ctx* c) {
  const long na = c->naux;
  const int nbf = (int)c->nbf;
  int* live = nullptr;
  int rc = dev_alloc((void**)&live, sizeof(int) * c->nseg, "segment liveness");
  if (rc) return rc;
  auto run = [&]() -> int {
    hipStream_t s = c->stream;
    HIPCHK(hipMemsetAsync(live, 0, sizeof(int) * c->nseg, s));
    HIPCHK(nopt::seg_live(c->Bs, c->splane, c->nseg, na, live, s));
    if (c->Bm && c->Bd) HIPCHK(nopt::slab_mask_rows(c->Bd, na * c->nbf * c->ld8, c->ld8, nbf, na, c->Bm, s));
    if (c->Bm && !c->Bd)
      HIPCHK(nopt::slab_mask_packed(reinterpret_cast<const int*>(c->Bs), 16L * c->nseg, nbf, na, c->Bm, s));
    std::vector<int> h((size_t)c->nseg);
    HIPCHK(hipMemcpyAsync(h.data(), live, sizeof(int) * c->nseg, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> segs, slots;
    std::vector<uint8_t> jl((size_t)nbf * c->nsr, 0);
    long k = 0;  // segments in the order of coulomb_seg_table
    for (int m = 0; m < nbf; ++m)
      for (int n0 = 0; n0 <= m; n0 += 16, ++k) {
        if (h[k]) segs.push_back((m << 16) | n0);
        for (int half = 0; half < 2; ++half) {
          const int n8 = n0 + 8 * half;
          if (n8 <= m && (h[k] >> half & 1)) {
            slots.push_back((m << 16) | n8);
            jl[(size_t)m * c->nsr + n8 / 8] = 1;
          }
        }
      }
    if (segs.empty()) segs.push_back(0);  // B all zeros: one segment of zeros keeps the grids valid
    if (slots.empty()) slots.push_back(0);
    HIPCHK(hipMemcpyAsync(c->lsegs, segs.data(), sizeof(int) * segs.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->lslots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->jlive, jl.data(), jl.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->nlseg = (int)segs.size();
    c->nlslots = (int)slots.size();
    coulomb_split_defaults(c);
    return NOPT_OK;
  };
  rc = run();
  dev_free(live);
  return rc;
}

// Forms B from packed lower-triangle rows in host memory, host[P][m(m+1)/2+n]:
// raw (P|mn) solved against the Cholesky factor dL of the metric (B = L^-1
// (P|mn)), or an already fitted tensor when dL is null. The rows m stream
// through the device in chunks (whole rows, at most kChunkBytes), so the
// packed tensor is never resident: each chunk is copied in, solved, and
// written into the kept forms of B: FP64 rows, and the digit planes (split
// from the FP64 rows, or, without them, from the chunks in a second pass once
// every row maximum is known).
constexpr size_t kChunkBytes = (size_t)4 << 30;

int form_B(nopt_jk_ctx* c, const double* host, const double* dL) {
  const bool keep64 = !(c->flags & NOPT_CTX_INT8_ONLY), keep8 = !(c->flags & NOPT_CTX_FP64_ONLY);
  const long na = c->naux, nbf = c->nbf, rows = na * nbf;
  const long cp = std::max(nbf, std::min(c->npair, (long)(kChunkBytes / (sizeof(double) * na))));
  int rc = NOPT_OK;
  c->b_ready = false;
  if (keep64 && !c->B64) rc = dev_alloc((void**)&c->B64, sizeof(double) * rows * c->ld64, "B (FP64)");
  if (!rc && keep8 && !c->Be) {
    if (c->bfull) rc = dev_alloc((void**)&c->Bd, (size_t)4 * rows * c->ld8, "B digit planes");
    if (!rc) rc = dev_alloc((void**)&c->Be, sizeof(int) * rows, "B row exponents");
    if (!rc) rc = dev_alloc((void**)&c->Be8, (size_t)na * c->ld8, "B row exponents (bytes)");
    c->splane = na * 16L * c->nseg;
    if (!rc) rc = dev_alloc((void**)&c->Bs, (size_t)5 * c->splane, "packed lower triangle of B");
    c->nmb = (int)((nbf + 31) / 32);
    if (!rc && c->ld8 <= 2048) rc = dev_alloc((void**)&c->Bm, sizeof(uint32_t) * na * c->nmb, "slab masks of B");
    if (!rc) rc = dev_alloc((void**)&c->perm, sizeof(int) * nbf, "order of the functions");
    if (!rc) rc = dev_alloc((void**)&c->pinv, sizeof(int) * nbf, "inverse order of the functions");
  }
  if (rc) return rc;
  if (keep8) {
    std::vector<int> hp((size_t)nbf), hi((size_t)nbf);
    nopt::band_order(host, na, nbf, hp.data());
    for (long i = 0; i < nbf; ++i) hi[(size_t)hp[(size_t)i]] = (int)i;
    if (hipMemcpy(c->perm, hp.data(), sizeof(int) * nbf, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->pinv, hi.data(), sizeof(int) * nbf, hipMemcpyHostToDevice) != hipSuccess)
      return set_error(NOPT_EHIP, "form_B: copy of the function order failed");
  }
  double* dA = nullptr;
  unsigned long long* rmax = nullptr;
  rc = dev_alloc((void**)&dA, sizeof(double) * na * cp, "chunk of the packed tensor");
  if (!rc && keep8 && !keep64) rc = dev_alloc((void**)&rmax, sizeof(unsigned long long) * rows, "row maxima");
  auto run = [&]() -> int {
    hipStream_t s = c->stream;
    if (keep64) HIPCHK(hipMemsetAsync(c->B64, 0, sizeof(double) * rows * c->ld64, s));
    if (rmax) {
      if (c->Bd) HIPCHK(hipMemsetAsync(c->Bd, 0, (size_t)4 * rows * c->ld8, s));
      HIPCHK(hipMemsetAsync(c->Bs, 0, (size_t)5 * c->splane, s));  // the scatter leaves the padding
      HIPCHK(hipMemsetAsync(rmax, 0, sizeof(unsigned long long) * rows, s));
    }
    // pass 1 forms the FP64 rows or the row maxima, pass 2 (integer forms
    // without FP64 rows) the packed triangle and the digit planes
    for (int pass = 0; pass < (rmax ? 2 : 1); ++pass) {
      for (int m0 = 0; m0 < (int)nbf;) {
        int m1 = m0 + 1;
        while (m1 < (int)nbf && (long)(m1 + 1) * (m1 + 2) / 2 - (long)m0 * (m0 + 1) / 2 <= cp) ++m1;
        const long p0 = (long)m0 * (m0 + 1) / 2, w = (long)m1 * (m1 + 1) / 2 - p0;
        HIPCHK(hipMemcpy2DAsync(dA, sizeof(double) * w, host + p0, sizeof(double) * c->npair, sizeof(double) * w,
                                na, hipMemcpyHostToDevice, s));
        if (dL) HIPCHK(nopt::trsm_lower_f64(dL, na, (int)na, dA, w, w, s));
        if (keep64) {
          HIPCHK(nopt::unpack_packed_f64(dA, w, m0, m1, (int)nbf, na, c->B64, c->ld64, s));
        } else if (pass == 0) {
          HIPCHK(nopt::rowmax_packed(dA, w, m0, m1, (int)nbf, na, rmax, s, c->pinv));
        } else {
          HIPCHK(nopt::scatter_segs_chunk(dA, w, m0, m1, (int)nbf, na, c->Be, c->pinv, c->nseg, c->Bs, c->splane, s));
          if (c->Bd)
            HIPCHK(nopt::split_packed_i8(dA, w, m0, m1, (int)nbf, na, c->Be, 4, c->Bd, c->ld8, rows * c->ld8, s,
                                         c->pinv));
        }
        m0 = m1;
      }
      if (rmax && pass == 0) HIPCHK(nopt::rowmax_exponents(rmax, rows, c->Be, s));
    }
    if (keep64 && keep8) {
      // the exponents (and digit planes) from the FP64 rows, then the packed triangle
      HIPCHK(nopt::split_rows_i8(c->B64, c->ld64, (int)nbf, rows, c->Bd ? 4 : 0, c->Bd, c->ld8, rows,
                                 rows * c->ld8, c->Be, s, c->perm, nbf, c->perm));
      HIPCHK(nopt::pack_segs_rows(c->B64, c->ld64, (int)nbf, na, c->Be, c->segs, c->nseg, c->Bs, c->splane, s,
                                  c->perm));
    }
    if (keep8) HIPCHK(nopt::expo_bytes(c->Be, (int)nbf, c->ld8, na, c->Be8, s));
    HIPCHK(hipStreamSynchronize(s));
    return NOPT_OK;
  };
  if (!rc) rc = run();
  dev_free(dA);
  dev_free(rmax);
  if (!rc && keep8) rc = find_live(c);
  if (!rc) c->b_ready = true;
  return rc;
}

// One density of a batched build (device pointers). D: nbf x nbf; C: nocc x
// nbf (orbitals as rows); J, K: nbf x nbf outputs, either may be null.
struct BuildItem {
  const double* D;
  const double* C;
  long nocc;
  double* J;
  double* K;
};

template <typename T>
void stage_rows(const double* src, long lds, int rows, int cols, T* dst, long ldd, int rows_pad, hipStream_t s,
                const int* rperm = nullptr, const int* cperm = nullptr) {
  const long n = (long)rows_pad * ldd;
  if (n > 0)
    k_pad_copy<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, lds, rows, cols, dst, ldd, rows_pad, rperm, cperm);
}

// Work buffers of one build: the context's own (grown on demand) or carved
// out of a caller-independent scratch area (the allocation-free entry point).
struct WorkBytes {
  size_t X = 0, Cp = 0, Cd = 0, Ce = 0, Xd = 0, Xe = 0, Jpart = 0, Kpart = 0, Dp = 0, gamma = 0, gpart = 0,
         gslot = 0;
};
struct Work {
  void *X = nullptr, *Cp = nullptr, *Cd = nullptr, *Ce = nullptr, *Xd = nullptr, *Xe = nullptr, *Jpart = nullptr;
  double *Kpart = nullptr, *Dp = nullptr, *gamma = nullptr, *gpart = nullptr, *gslot = nullptr;
};

// The context's buffers, grown to the sizes of this build (the fixed-size
// ones were allocated by create and cover every build).
int context_work(nopt_jk_ctx* c, const WorkBytes& b, Work* w) {
  int rc = NOPT_OK;
  if (b.X) {
    size_t cap = c->X_bytes;
    rc = ensure(c, &c->X, &cap, b.X, "half-transformed tensor X");
    if (rc) return rc;
    c->X_bytes = cap;
  }
  if (b.Cp) {
    size_t cap = c->Cp_bytes;
    rc = ensure(c, &c->Cp, &cap, b.Cp, "padded orbital factor");
    if (rc) return rc;
    c->Cp_bytes = cap;
  }
  if (b.Cd) {
    size_t b1 = c->Cd_bytes, b2 = c->Ce_bytes;
    rc = ensure(c, &c->Cd, &b1, b.Cd, "orbital digit planes");
    if (!rc) rc = ensure(c, &c->Ce, &b2, b.Ce, "orbital exponents");
    if (rc) return rc;
    c->Cd_bytes = b1;
    c->Ce_bytes = b2;
  }
  if (b.Xd) {
    size_t b3 = c->Xd_bytes, b4 = c->Xe_bytes;
    rc = ensure(c, &c->Xd, &b3, b.Xd, "exchange digit planes");
    if (!rc) rc = ensure(c, &c->Xe, &b4, b.Xe, "exchange row exponents");
    if (rc) return rc;
    c->Xd_bytes = b3;
    c->Xe_bytes = b4;
  }
  {
    size_t kb = c->Kpart_bytes;
    rc = ensure(c, (void**)&c->Kpart, &kb, b.Kpart, "K partials");
    if (rc) return rc;
    c->Kpart_bytes = kb;
  }
  if (b.Jpart) {
    size_t jb = c->Jpart_bytes;
    rc = ensure(c, &c->Jpart, &jb, b.Jpart, "J partials");
    if (rc) return rc;
    c->Jpart_bytes = jb;
  }
  w->X = c->X;
  w->Cp = c->Cp;
  w->Cd = c->Cd;
  w->Ce = c->Ce;
  w->Xd = c->Xd;
  w->Xe = c->Xe;
  w->Jpart = c->Jpart;
  w->Kpart = c->Kpart;
  w->Dp = c->Dp;
  w->gamma = c->gamma;
  w->gpart = c->gpart;
  w->gslot = c->gslot;
  return NOPT_OK;
}

// The same buffers as consecutive 256-byte aligned pieces of `base` (bytes
// available: `cap`); nothing is allocated.
int carve_work(void* base, size_t cap, const WorkBytes& b, Work* w) {
  char* p = (char*)base;
  size_t used = 0;
  auto take = [&](size_t bytes) -> void* {
    if (!bytes) return nullptr;
    void* r = p + used;
    used += (bytes + 255) & ~(size_t)255;
    return r;
  };
  w->X = take(b.X);
  w->Cp = take(b.Cp);
  w->Cd = take(b.Cd);
  w->Ce = take(b.Ce);
  w->Xd = take(b.Xd);
  w->Xe = take(b.Xe);
  w->Jpart = take(b.Jpart);
  w->Kpart = (double*)take(b.Kpart);
  w->Dp = (double*)take(b.Dp);
  w->gamma = (double*)take(b.gamma);
  w->gpart = (double*)take(b.gpart);
  w->gslot = (double*)take(b.gslot);
  if (used > cap)
    return set_error(NOPT_EINVAL, "this build needs " + std::to_string(used) + " bytes of work space; the static "
                                  "scratch area holds " + std::to_string(cap) + " (use nopt_jk_build_device)");
  return NOPT_OK;
}

// Enqueues the J/K builds of nd densities that share B. The orbital factors
// are stacked into one padded factor, so the half-transform reads B (or its
// digit planes) once for all of them; gamma and J are formed for every
// density in one pass over B each; K is one triangular product per density.
// Phases (for the optional timing): staging, half-transform, gamma, J
// partials, exchange (K products and K output), J output.
// A recorded build (the ragged plan): the integer build's launches as
// work-list items tagged with their phase (launches of one phase do not
// depend on each other across systems; phases run in order).
constexpr int kListPad = nopt::kListOzakiKinds, kListFin = nopt::kListOzakiKinds + 1;
struct RecItem {
  int phase;
  nopt::ListRec r;
};
struct Rec {
  std::vector<RecItem> items;
  template <typename T>
  void add(int phase, int kind, int blocks, int threads, const T& a) {
    RecItem x;
    x.phase = phase;
    x.r.kind = kind;
    x.r.cfg = 0;
    x.r.blocks = blocks;
    x.r.threads = threads;
    x.r.args.resize(sizeof(T));
    std::memcpy(x.r.args.data(), &a, sizeof(T));
    items.push_back(std::move(x));
  }
  void add(int phase, const nopt::ListRec& r) { items.push_back(RecItem{phase, r}); }
};

// With rec, the integer build is recorded instead of enqueued (no timing, no
// side stream; the fused exchange split only).
int enqueue_batch(nopt_jk_ctx* c, int nd, const BuildItem* it, int prec, hipStream_t s, void* work_area = nullptr,
                  size_t work_area_bytes = 0, Rec* rec = nullptr) {
  const int nbf = (int)c->nbf, naux = (int)c->naux;
  if (prec < 0 || prec > 4) return set_error(NOPT_EINVAL, "unknown precision mode");
  if (nd < 1 || nd > kMaxBatch) return set_error(NOPT_EINVAL, "batch size must be between 1 and NOPT_MAX_BATCH");
  const bool f64 = prec == NOPT_PREC_FP64, i8 = prec == NOPT_PREC_INT8;
  const bool bf = prec == NOPT_PREC_BF16X3 || prec == NOPT_PREC_BF16X6;
  if (i8 && !c->Bs)
    return set_error(NOPT_ESTATE, "the context keeps no integer form of B (NOPT_CTX_FP64_ONLY): integer mode unavailable");
  if (!i8 && !c->B64)
    return set_error(NOPT_ESTATE, "the context keeps no FP64 B (NOPT_CTX_INT8_ONLY): only the integer mode is available");
  const long unit = bf ? 32 : 16;  // column padding of each orbital block
  long off[kMaxBatch + 1];
  int jl[kMaxBatch];
  int nj = 0;
  off[0] = 0;
  for (int d = 0; d < nd; ++d) {
    const BuildItem& x = it[d];
    if (x.nocc < 0 || x.nocc > c->nbf) return set_error(NOPT_EINVAL, "nocc out of range");
    const bool wk = x.K && x.nocc > 0;
    if (wk && !x.C) return set_error(NOPT_EINVAL, "K requested without an orbital factor");
    if (x.J && !x.D) return set_error(NOPT_EINVAL, "J requested without a density");
    off[d + 1] = off[d] + (wk ? round_up(x.nocc, unit) : 0);
    if (x.J) jl[nj++] = d;
  }
  const long ncol = off[nd];
  const bool anyK = ncol > 0;
  // one density with J and K in a non-integer mode: gamma rides on the
  // half-transform (row dots of the B panels already in registers)
  const bool fused = !i8 && nd == 1 && nj == 1 && anyK;
  int ngroups = 1;
  const int nt = i8 ? nt_for_i8(anyK ? ncol : 1, &ngroups) : nt_for(anyK ? ncol : 1, &ngroups);
  const long ldx = (long)nt * 16 * ngroups;  // stacked, padded orbital count
  const long ldb = i8 ? c->ld8 : c->ld64;    // leading dimension of the staged D and C
  const int mt = i8 ? c->mtiles64 : c->mtiles;
  const size_t esz = (f64 || i8) ? sizeof(double) : sizeof(float);
  const long rows = c->naux * c->nbf;
  const long nn = (long)nbf * nbf;
  int rc;
  // integer mode: the half-transform writes the digit planes of X directly,
  // X itself is never stored. Each density's orbital columns are cut at the
  // column groups into segments (at most nopt::kMaxXSeg in all), and each
  // segment's rows get their own exponent and digit planes; pairs of
  // auxiliary functions share the exponent of a row m and lie side by side
  // (npq rows of pairs), so the exchange streams 2*len orbitals per exponent,
  // one launch per segment into its own range of K partials. (Triples of
  // functions pad less for 81 orbitals, 256 columns per 3 against 192 per 2,
  // but their half-transform, one block of 12 waves per CU, was slower than
  // the exchange gained: 535 against 557 builds/s at C20H42.)
  const long bn = 16L * nt;
  int nxs = 0, sfirst[kMaxBatch + 1];  // segments of density d: sfirst[d] .. sfirst[d+1]-1
  struct Seg { int off, len; };
  Seg segl[nopt::kMaxXSeg];
  bool xfits = true;
  for (int d = 0; d < nd; ++d) {
    sfirst[d] = nxs;
    if (off[d + 1] == off[d]) continue;
    const long a = off[d], b = off[d] + it[d].nocc;
    for (long g = a / bn; g * bn < b; ++g) {
      const long s0 = std::max(a, g * bn), s1 = std::min(b, (g + 1) * bn);
      if (nxs == nopt::kMaxXSeg) {
        xfits = false;
        break;
      }
      segl[nxs++] = Seg{(int)s0, (int)(s1 - s0)};
    }
  }
  sfirst[nd] = nxs;
  const bool xfused = i8 && anyK && xfits;
  const int qpb = xfused ? 2 : 1;
  const long npq = (naux + qpb - 1) / qpb, rowsq = npq * nbf;
  auto ldk_of = [&](long len) { return xfused ? round_up(qpb * len, 64) : round_up(len, 64); };
  long xoff[nopt::kMaxXSeg + 1];  // digit-plane offsets (bytes) of each segment
  xoff[0] = 0;
  for (int k = 0; k < (xfused ? nxs : 0); ++k) xoff[k + 1] = xoff[k] + 4 * rowsq * ldk_of(segl[k].len);
  if (rec && (!i8 || (anyK && !xfused)))
    return set_error(NOPT_EINVAL, "a recorded (ragged plan) build needs the integer mode and at most 8 column-group "
                                  "segments");
  int segs_max = 1;  // most segments of one density (K partial ranges)
  for (int d = 0; d < nd && xfused; ++d) segs_max = std::max(segs_max, sfirst[d + 1] - sfirst[d]);
  // Coulomb splits: nj gamma chunks of one split must fit the LDS cache
  const int nsj_min = !nj ? 0 : (naux + JCHUNK_MAX / nj - 1) / (JCHUNK_MAX / nj);
  const int nsj = !nj ? 0 : i8 ? c->nsj8 : std::max(nsj_min, c->nsj > 0 ? c->nsj : NSPLIT_J_MIN);
  // bytes of each work buffer this build uses
  WorkBytes wb;
  if (anyK) {
    if (!xfused) wb.X = esz * rows * ldx;
    wb.Cp = esz * ldx * ldb;
    if (i8) {
      wb.Cd = (size_t)4 * ldx * c->ld8;
      wb.Ce = sizeof(int) * ldx;
      long ldk_max = 0;
      for (int d = 0; d < nd; ++d)
        if (off[d + 1] > off[d]) ldk_max = std::max(ldk_max, ldk_of(it[d].nocc));
      // fused: planes of every segment side by side; otherwise one density
      // at a time
      wb.Xd = xfused ? (size_t)xoff[nxs] : (size_t)4 * rows * ldk_max;
      wb.Xe = sizeof(int) * (xfused ? rowsq * nxs : rows);
    }
  }
  // K partials: a range per segment of one density, reused density after
  // density; a recorded build runs every exchange before any K output, so
  // there each segment of the batch has its own range
  const int kranges = rec ? std::max(nxs, 1) : segs_max;
  wb.Kpart = sizeof(double) * (i8 ? (size_t)kranges * c->nsk8 * c->ntiles8 * nopt::kSyrkI8Tile * nopt::kSyrkI8Tile
                                  : (size_t)c->nsk * c->ntiles * KT * KT);
  if (nj) wb.Jpart = sizeof(double) * nj * nsj * nn;
  wb.Dp = sizeof(double) * std::max(nj, 1) * c->dstride;
  wb.gamma = sizeof(double) * std::max(nj, 1) * naux;
  wb.gpart = sizeof(double) * naux * c->mtiles64;
  wb.gslot = sizeof(double) * nopt::gamma_partial_blocks(c->nlslots, c->nlseg) * std::max(nj, 1) * naux;
  Work w;
  rc = work_area ? carve_work(work_area, work_area_bytes, wb, &w) : context_work(c, wb, &w);
  if (rc) return rc;
  double* Jpart = (double*)w.Jpart;
  const bool tm = !rec && c->timing && c->timed < kTimedBuilds;
  auto mark = [&](int ph) {
    if (tm) (void)hipEventRecord(c->tev[(size_t)c->timed * (kPhases + 1) + ph], s);
  };
  // Coulomb pair (gamma, J partials) and the J output. Without timing they
  // run on the side stream, forked after the staging (after the
  // half-transform in the integer mode, or when gamma rides on it) and joined
  // before the build ends, so that they overlap the exchange.
  const bool overlap = !rec && !tm && nj > 0 && anyK;
  const unsigned nelj = (unsigned)((nn + 255) / 256);  // finalize blocks
  // the integer mode works on the functions in the order of B's integer
  // forms (perm[new] = old): densities are read through it, orbitals'
  // columns too, and J, K written back through it
  const int* perm = i8 ? c->perm : nullptr;
  auto pad_rows = [&](int phase, const double* src, long lds, int r, int cols, double* dst, long ldd, int rpad,
                      const int* rperm = nullptr, const int* cperm = nullptr) {
    const long n = (long)rpad * ldd;
    if (n <= 0) return;
    if (rec)
      rec->add(phase, kListPad, (int)((n + 255) / 256), 256,
               ItemPad{src, lds, r, cols, dst, ldd, rpad, rperm, cperm, (int)((n + 255) / 256), 1, 1});
    else
      stage_rows<double>(src, lds, r, cols, dst, ldd, rpad, s, rperm, cperm);
  };
  auto split_rows = [&](const double* src, long ld, int cols, long rows_src, int8_t* planes, long ldo, long rows_dst,
                        long pstride, int* expo, const int* cperm = nullptr) -> hipError_t {
    if (!rec)
      return nopt::split_rows_i8(src, ld, cols, rows_src, 4, planes, ldo, rows_dst, pstride, expo, s, nullptr, 0,
                                 cperm);
    if (rows_dst <= 0) return hipSuccess;
    nopt::ListRec r;
    CHK_(nopt::split_rows_item(src, ld, cols, rows_src, 4, planes, ldo, rows_dst, pstride, expo, &r, cperm));
    rec->add(1, r);
    return hipSuccess;
  };
  auto finalize = [&](int phase, const double* Jp, int nsj_, const double* Kp, int nsk_, int ntl, int kt,
                      double* J, double* K, hipStream_t st) {
    const uint8_t* jl = Jp && i8 ? c->jlive : nullptr;  // J partials of the integer pair
    if (rec)
      rec->add(phase, kListFin, (int)nelj, 256,
               ItemFin{Jp, nsj_, jl, c->nsr, Kp, nsk_, ntl, kt, nbf, perm, J, K, (int)nelj, 1, 1});
    else
      k_finalize<<<nelj, 256, 0, st>>>(Jp, nsj_, jl, c->nsr, Kp, nsk_, ntl, kt, nbf, perm, J, K);
  };
  hipStream_t sj = overlap ? c->side : s;
  auto cmark = [&](int ph) {
    if (!overlap) mark(ph);
  };
  auto coulomb_gamma = [&]() -> hipError_t {
    if (fused) {
      k_gamma<<<(naux + 255) / 256, 256, 0, sj>>>(w.gpart, c->mtiles, naux, w.gamma);
    } else if (i8 && rec) {
      nopt::ListRec r[2];
      CHK_(nopt::gamma_slots_items(nj, c->nsg, c->Bs, c->splane, 16L * c->nseg, c->Be, nbf, naux, c->lslots,
                                   c->nlslots, c->lsegs, c->nlseg, w.Dp, ldb, c->dstride, w.gslot, w.gamma, naux, r));
      rec->add(3, r[0]);
      rec->add(4, r[1]);
      return hipSuccess;
    } else if (i8) {
      CHK_(nopt::launch_gamma_slots(nj, c->nsg, sj, c->Bs, c->splane, 16L * c->nseg, c->Be, nbf, naux, c->lslots,
                                    c->nlslots, c->lsegs, c->nlseg, w.Dp, ldb, c->dstride, w.gslot, w.gamma, naux));
    } else {
#define NOPT_GF(ND)                                                                                         \
  case ND:                                                                                                  \
    k_gamma_f64<ND><<<naux, 256, 0, sj>>>(c->B64, c->ld64, nbf, w.Dp, ldb, c->dstride, w.gamma, naux); \
    break;
      switch (nj) { NOPT_GF(1) NOPT_GF(2) NOPT_GF(3) NOPT_GF(4) }
#undef NOPT_GF
    }
    CHK_(hipGetLastError());
    cmark(3);
    return hipSuccess;
  };
  auto coulomb_j = [&]() -> hipError_t {
    if (i8 && rec) {
      nopt::ListRec r;
      CHK_(nopt::jpart_slots_item(nj, nsj, c->Bs, c->splane, 16L * c->nseg, c->Be, nbf, naux, c->lslots,
                                  c->nlslots, c->lsegs, c->nlseg, w.gamma, naux, Jpart, (long)nsj * nn, &r));
      rec->add(5, r);
      return hipSuccess;
    }
    if (i8)
      return nopt::launch_jpart_slots(nj, nsj, sj, c->Bs, c->splane, 16L * c->nseg, c->Be, nbf, naux, c->lslots,
                                      c->nlslots, c->lsegs, c->nlseg, w.gamma, naux, Jpart, (long)nsj * nn);
    dim3 gj((nbf + 1) / 2, nsj);
#define NOPT_JP(ND)                                                                                     \
  case ND:                                                                                              \
    k_jpart<ND><<<gj, 256, 0, sj>>>(c->B64, c->ld64, nbf, naux, w.gamma, naux, Jpart, (long)nsj * nn); \
    break;
    switch (nj) { NOPT_JP(1) NOPT_JP(2) NOPT_JP(3) NOPT_JP(4) }
#undef NOPT_JP
    return hipGetLastError();
  };
  auto coulomb = [&]() -> hipError_t {
    CHK_(coulomb_gamma());
    return coulomb_j();
  };
  auto coulomb_out = [&]() -> hipError_t {
    for (int k = 0; k < nj; ++k)
      finalize(6, Jpart + (long)k * nsj * nn, nsj, nullptr, 0, 1, KT, it[jl[k]].J, nullptr, sj);
    return hipGetLastError();
  };
  // 0: no J; 1: forked after staging; 2: in line (after the half-transform)
  int coulomb_at = nj ? 2 : 0;
  mark(0);
  // densities of the J list into slots 0..nj-1 (ld = ldb); orbital blocks
  // into the stacked factor at column offsets off[d] (zero rows between)
  for (int k = 0; k < nj; ++k) pad_rows(0, it[jl[k]].D, nbf, nbf, nbf, w.Dp + k * c->dstride, ldb, nbf, perm, perm);
  if (anyK) {
    for (int d = 0; d < nd; ++d) {
      const int pad = (int)(off[d + 1] - off[d]);
      if (!pad) continue;
      const int r = (int)it[d].nocc;
      if (i8)
        HIPCHK(split_rows(it[d].C, nbf, nbf, r, (int8_t*)w.Cd + off[d] * c->ld8, c->ld8, pad, ldx * c->ld8,
                          (int*)w.Ce + off[d], perm));
      else if (f64)
        stage_rows<double>(it[d].C, nbf, r, nbf, (double*)w.Cp + off[d] * ldb, ldb, pad, s);
      else
        stage_rows<float>(it[d].C, nbf, r, nbf, (float*)w.Cp + off[d] * ldb, ldb, pad, s);
    }
    const int tail = (int)(ldx - ncol);
    if (i8)
      HIPCHK(split_rows(w.Dp, ldb, 0, 0, (int8_t*)w.Cd + ncol * c->ld8, c->ld8, tail, ldx * c->ld8,
                        (int*)w.Ce + ncol));
    else if (f64)
      stage_rows<double>(w.Dp, ldb, 0, 0, (double*)w.Cp + ncol * ldb, ldb, tail, s);
    else
      stage_rows<float>(w.Dp, ldb, 0, 0, (float*)w.Cp + ncol * ldb, ldb, tail, s);
  }
  HIPCHK(hipGetLastError());
  // the integer Coulomb pair starts where cplace says (tuned: beside the
  // half-transform it takes CUs from it, after it it shares the device with
  // the exchange)
  if (overlap && !fused && !i8) {
    HIPCHK(hipEventRecord(c->ev_fork, s));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    HIPCHK(coulomb());
    HIPCHK(coulomb_out());
    coulomb_at = 1;
  }
  const bool i8early = overlap && i8 && c->cplace > 0;
  if (i8early) {
    HIPCHK(hipEventRecord(c->ev_fork, s));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    HIPCHK(coulomb_gamma());
    if (c->cplace == 2) {
      HIPCHK(coulomb_j());
      HIPCHK(coulomb_out());
    }
    coulomb_at = 1;
  }
  mark(1);
  if (anyK) {
    dim3 gh(mt, naux, ngroups);
    hipError_t e;
    if (i8) {
      nopt::XSplit xs{};
      if (xfused)
        for (int k = 0; k < nxs; ++k) {
          const long ldk = ldk_of(segl[k].len);
          xs.off[k] = segl[k].off;
          xs.len[k] = segl[k].len;
          xs.ldk[k] = ldk;
          xs.xplane[k] = rowsq * ldk;
          xs.xd[k] = (int8_t*)w.Xd + xoff[k];
          xs.xe[k] = (int*)w.Xe + (long)k * rowsq;
          xs.nseg = k + 1;
        }
      if (xfused) gh.y = (unsigned)npq;
      const nopt::HalfB hb{c->Bd, rows * c->ld8, reinterpret_cast<const int*>(c->Bs), 16L * c->nseg, c->Be8, c->Be,
                           c->Bm, c->nmb};
      if (rec) {
        nopt::ListRec r;
        e = nopt::half_i8_item(nt, qpb, gh, hb, c->ld8, nbf, naux, (const int8_t*)w.Cd, ldx * c->ld8,
                               (const int*)w.Ce, nullptr, ldx, xs, &r);
        if (e == hipSuccess) rec->add(2, r);
      } else {
        e = nopt::launch_half_i8(nt, qpb, gh, s, hb, c->ld8, nbf, naux, (const int8_t*)w.Cd, ldx * c->ld8,
                                 (const int*)w.Ce, xfused ? nullptr : (double*)w.X, ldx, xs);
      }
    }
    else if (f64)
      e = launch_half<double>(nt, gh, s, c->B64, ldb, nbf, (const double*)w.Cp, w.Dp, (double*)w.X, ldx,
                              w.gpart, c->mtiles);
    else if (prec == NOPT_PREC_FP32)
      e = launch_half<float>(nt, gh, s, c->B64, ldb, nbf, (const float*)w.Cp, w.Dp, (float*)w.X, ldx,
                             w.gpart, c->mtiles);
    else
      e = nopt::launch_half_bf16(prec == NOPT_PREC_BF16X3 ? 2 : 3, nt, gh, s, c->B64, ldb, nbf,
                                 (const float*)w.Cp, w.Dp, (float*)w.X, ldx, w.gpart, c->mtiles);
    HIPCHK(e);
  }
  mark(2);
  if (overlap && (fused || i8) && !(i8early && c->cplace == 2)) {
    HIPCHK(hipEventRecord(c->ev_half, s));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_half, 0));
    if (i8early) {
      HIPCHK(coulomb_j());
    } else {
      HIPCHK(coulomb());
    }
    HIPCHK(coulomb_out());
    coulomb_at = 1;
  }
  if (coulomb_at == 2) HIPCHK(coulomb());
  mark(4);
  for (int d = 0; d < nd; ++d) {
    if (!it[d].K) continue;
    const int kpad = (int)(off[d + 1] - off[d]);
    if (!kpad) {  // no occupied orbitals: K = 0
      pad_rows(8, w.Dp, nbf, 0, 0, it[d].K, nbf, nbf);
      continue;
    }
    const unsigned nel = (unsigned)((nn + 255) / 256);
    if (i8) {
      // the exchange runs over the rows of the digit planes: groups of qpb
      // functions (npq of them) per segment when fused, single functions of
      // the whole density otherwise; each segment fills its own nsk8 ranges
      // of K partials and the finalize sums them all
      const size_t kseg = (size_t)c->nsk8 * c->ntiles8 * nopt::kSyrkI8Tile * nopt::kSyrkI8Tile;
      const int kfirst = rec ? sfirst[d] : 0;  // the density's first K partial range
      int nsum = 1;
      if (xfused) {
        nsum = sfirst[d + 1] - sfirst[d];
        for (int k = sfirst[d]; k < sfirst[d + 1]; ++k) {
          const long ldk = ldk_of(segl[k].len);
          const int8_t* xd = (const int8_t*)w.Xd + xoff[k];
          const int* xe = (const int*)w.Xe + (long)k * rowsq;
          double* kp = w.Kpart + (k - sfirst[d] + kfirst) * kseg;
          if (rec) {
            nopt::ListRec r;
            HIPCHK(nopt::syrk_i8_item(c->ntiles8, c->nsk8, xd, rowsq * ldk, xe, ldk, nbf, (int)npq, kp, &r));
            rec->add(7, r);
          } else {
            HIPCHK(nopt::launch_syrk_i8(c->ntiles8, c->nsk8, s, xd, rowsq * ldk, xe, ldk, nbf, (int)npq, kp));
          }
        }
      } else {
        const long ldk = ldk_of(it[d].nocc);
        HIPCHK(nopt::split_rows_i8((const double*)w.X + off[d], ldx, (int)it[d].nocc, rows, 4, (int8_t*)w.Xd, ldk,
                                   rows, rows * ldk, (int*)w.Xe, s));
        HIPCHK(nopt::launch_syrk_i8(c->ntiles8, c->nsk8, s, (const int8_t*)w.Xd, rows * ldk, (const int*)w.Xe, ldk,
                                    nbf, (int)naux, w.Kpart));
      }
      finalize(8, nullptr, 0, w.Kpart + kfirst * kseg, nsum * c->nsk8, c->ntiles8, nopt::kSyrkI8Tile, nullptr,
               it[d].K, s);
    } else {
      dim3 gk(c->ntiles, c->nsk);
      if (f64)
        k_syrk<double><<<gk, 256, 0, s>>>((const double*)w.X + off[d], ldx, nbf, naux, kpad, w.Kpart);
      else if (prec == NOPT_PREC_FP32)
        k_syrk<float><<<gk, 256, 0, s>>>((const float*)w.X + off[d], ldx, nbf, naux, kpad, w.Kpart);
      else
        HIPCHK(nopt::launch_syrk_bf16(prec == NOPT_PREC_BF16X3 ? 2 : 3, gk, s, (const float*)w.X + off[d], ldx,
                                      nbf, naux, kpad, w.Kpart));
      k_finalize<<<nel, 256, 0, s>>>(nullptr, 0, nullptr, 0, w.Kpart, c->nsk, c->ntiles, KT, nbf, nullptr, nullptr,
                                     it[d].K);
    }
    HIPCHK(hipGetLastError());
  }
  mark(5);
  if (coulomb_at == 2) HIPCHK(coulomb_out());
  if (overlap) {
    HIPCHK(hipEventRecord(c->ev_join, c->side));
    HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
  }
  mark(6);
  if (tm) ++c->timed;
  return NOPT_OK;
}

int enqueue_build(nopt_jk_ctx* c, const double* dD, const double* dC, long nocc, double* dJ, double* dK, int prec,
                  hipStream_t s) {
  const BuildItem x{dD, dC, nocc, dJ, dK};
  return enqueue_batch(c, 1, &x, prec, s);
}

int upload(nopt_jk_ctx* c, double** dst, long* cap, const double* src, long count) {
  if (*cap < count || !*dst) {
    dev_free(*dst);
    int rc = dev_alloc((void**)dst, sizeof(double) * std::max(count, 1L), "host staging copy");
    if (rc) return rc;
    *cap = count;
  }
  HIPCHK(hipMemcpyAsync(*dst, src, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
  return NOPT_OK;
}

}  // namespace

namespace {
// Work space of the allocation-free build (nopt_jk_build_device_static): one
// array per GPU, part of the library's code object, held until the process
// ends. In host code `scratch` only names the symbol; its device address on
// each GPU comes from hipGetSymbolAddress (scratch_address) and is the only
// form in which it reaches kernels or copies.
constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array (the device must be
// current): looked up on the first call on each device and kept.
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(NOPT_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(NOPT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hip_msg(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return NOPT_OK;
}
}  // namespace

extern "C" {

int nopt_jk_create(int device, long nbf, long naux, nopt_jk_ctx** out) {
  return nopt_jk_create_ex(device, nbf, naux, NOPT_CTX_DEFAULT, out);
}

int nopt_jk_create_ex(int device, long nbf, long naux, int flags, nopt_jk_ctx** out) {
  nopt::clear_error();
  if (!out) return set_error(NOPT_EINVAL, "nopt_jk_create: null output pointer");
  *out = nullptr;
  const int form = flags & ~NOPT_CTX_COMPACT;
  if ((form != NOPT_CTX_DEFAULT && form != NOPT_CTX_FP64_ONLY && form != NOPT_CTX_INT8_ONLY) ||
      (form == NOPT_CTX_FP64_ONLY && (flags & NOPT_CTX_COMPACT)))
    return set_error(NOPT_EINVAL,
                     "nopt_jk_create_ex: flags must be one of NOPT_CTX_DEFAULT, _FP64_ONLY, _INT8_ONLY, the first "
                     "or last with NOPT_CTX_COMPACT");
  if (nbf <= 0 || naux <= 0) return set_error(NOPT_EINVAL, "nopt_jk_create: nbf and naux must be positive");
  if (nbf > 2046) return set_error(NOPT_EINVAL, "nopt_jk_create: nbf above 2046 is not supported");
  if (naux > 65535) return set_error(NOPT_EINVAL, "nopt_jk_create: naux above 65535 is not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_error(NOPT_EHIP, "nopt_jk_create: no HIP device available");
  if (device < 0 || device >= ndev) return set_error(NOPT_EINVAL, "nopt_jk_create: bad device index");
  auto* c = new (std::nothrow) nopt_jk_ctx;
  if (!c) return set_error(NOPT_ENOMEM, "nopt_jk_create: host allocation failed");
  c->device = device;
  c->flags = flags;
  c->nbf = nbf;
  c->naux = naux;
  c->npair = nbf * (nbf + 1) / 2;
  c->ld64 = round_up(nbf + 1, 16);  // ld > nbf: see k_jpart
  c->mtiles = (int)((nbf + BM - 1) / BM);
  c->ld8 = round_up(nbf, 64);  // int8 MFMA k-steps of 64
  // the full-row digit planes of the half-transform, unless large or declined
  c->bfull = !(flags & NOPT_CTX_COMPACT) && 4.0 * naux * nbf * c->ld8 <= (double)kFullRowsMax;
  c->mtiles64 = (int)((nbf + 63) / 64);
  c->ktn = (int)((nbf + KT - 1) / KT);
  c->ntiles = c->ktn * (c->ktn + 1) / 2;
  c->nsk = std::max(1, std::min((int)naux, (1024 + c->ntiles - 1) / c->ntiles));
  {
    // integer exchange: by default Q-splits in multiples of 8 (as many
    // blocks on every XCD), sized so that the blocks of an XCD fill whole
    // rounds of its resident slots (two blocks per CU)
    const int kt8 = (int)((nbf + nopt::kSyrkI8Tile - 1) / nopt::kSyrkI8Tile);
    c->ntiles8 = kt8 * (kt8 + 1) / 2;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
      ncu = 256;
    const long slots = 2L * ncu / 8;  // per XCD
    int best = 8;
    double best_eff = 0.0;
    for (int nsp = 8; nsp <= 64 && nsp <= std::max(8L, naux); nsp += 8) {
      const long per = (long)nsp * c->ntiles8 / 8;
      const double eff = (double)per / (double)(slots * ((per + slots - 1) / slots));
      if (eff > best_eff + 0.02) {
        best_eff = eff;
        best = nsp;
      }
    }
    c->nsk8 = best;
  }
  int rc = NOPT_OK;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_half, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    return set_error(NOPT_EHIP, std::string("nopt_jk_create: ") + hip_msg(e));
  }
  const long nn = nbf * nbf;
  if (!rc) rc = dev_alloc((void**)&c->gpart, sizeof(double) * naux * c->mtiles64, "gamma partials");
  if (!rc) rc = dev_alloc((void**)&c->gamma, sizeof(double) * kMaxBatch * naux, "gamma");
  if (!rc) {
    c->Kpart_bytes = sizeof(double) * std::max((long)c->nsk * c->ntiles * KT * KT,
                                               (long)c->nsk8 * c->ntiles8 * nopt::kSyrkI8Tile * nopt::kSyrkI8Tile);
    rc = dev_alloc((void**)&c->Kpart, c->Kpart_bytes, "K partials");
  }
  c->dstride = nbf * std::max(c->ld64, c->ld8);
  if (!rc) rc = dev_alloc((void**)&c->Dp, sizeof(double) * kMaxBatch * c->dstride, "padded densities");
  if (!rc) rc = dev_alloc((void**)&c->dJ, sizeof(double) * nn, "J");
  if (!rc) rc = dev_alloc((void**)&c->dK, sizeof(double) * nn, "K");
  if (!rc) rc = dev_alloc((void**)&c->dD, sizeof(double) * nn, "D");
  if (!rc) rc = dev_alloc((void**)&c->dH, sizeof(double) * nn, "H");
  if (!rc) rc = dev_alloc((void**)&c->dF, sizeof(double) * nn, "F");
  if (!rc) rc = dev_alloc((void**)&c->epart, sizeof(double) * 1024, "energy partials");
  if (!rc) rc = dev_alloc((void**)&c->dE, sizeof(double), "energy");
  if (!rc) rc = dev_alloc((void**)&c->info, sizeof(int), "info");
  if (!rc) {
    c->nseg = nopt::coulomb_seg_count((int)nbf);
    std::vector<int> stab((size_t)c->nseg);
    nopt::coulomb_seg_table((int)nbf, stab.data());
    c->nslots = nopt::coulomb_slot_count((int)nbf);
    std::vector<int> tab((size_t)c->nslots);
    nopt::coulomb_slot_table((int)nbf, tab.data());
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      c->ncu = ncu;
    c->nlslots = c->nslots;
    coulomb_split_defaults(c);
    c->nsr = (int)((nbf + 7) / 8);
    const std::vector<uint8_t> ones((size_t)nbf * c->nsr, 1);
    rc = dev_alloc((void**)&c->slots, sizeof(int) * c->nslots, "Coulomb slot table");
    if (!rc) rc = dev_alloc((void**)&c->segs, sizeof(int) * c->nseg, "Coulomb segment table");
    if (!rc) rc = dev_alloc((void**)&c->lslots, sizeof(int) * c->nslots, "Coulomb live slot table");
    if (!rc) rc = dev_alloc((void**)&c->lsegs, sizeof(int) * c->nseg, "Coulomb live segment table");
    if (!rc) rc = dev_alloc((void**)&c->jlive, ones.size(), "Coulomb live slot map");
    if (!rc)
      rc = dev_alloc((void**)&c->gslot,
                     sizeof(double) * nopt::gamma_partial_blocks(c->nslots, c->nseg) * kMaxBatch * naux,
                     "gamma partials");
    c->nlseg = c->nseg;
    if (!rc && (hipMemcpy(c->slots, tab.data(), sizeof(int) * c->nslots, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->segs, stab.data(), sizeof(int) * c->nseg, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->lslots, tab.data(), sizeof(int) * c->nslots, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->lsegs, stab.data(), sizeof(int) * c->nseg, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->jlive, ones.data(), ones.size(), hipMemcpyHostToDevice) != hipSuccess))
      rc = set_error(NOPT_EHIP, "nopt_jk_create: copy of the slot tables failed");
  }
  if (rc) {
    nopt_jk_destroy(c);
    return rc;
  }
  *out = c;
  return NOPT_OK;
}

void nopt_jk_destroy(nopt_jk_ctx* c) {
  if (!c) return;
  if (c->pins > 0) {  // a graph still refers to the buffers: freed with the last such graph
    c->zombie = true;
    return;
  }
  (void)hipSetDevice(c->device);
  if (c->done_set) (void)hipEventSynchronize(c->ev_done);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->B64);
  dev_free(c->X);
  dev_free(c->gpart);
  dev_free(c->gamma);
  dev_free(c->Jpart);
  dev_free(c->Kpart);
  dev_free(c->Dp);
  dev_free(c->Cp);
  dev_free(c->dJ);
  dev_free(c->dK);
  dev_free(c->dD);
  dev_free(c->dC);
  dev_free(c->dH);
  dev_free(c->dF);
  dev_free(c->dG);
  dev_free(c->epart);
  dev_free(c->dE);
  dev_free(c->info);
  dev_free(c->Bd);
  dev_free(c->Be);
  dev_free(c->Be8);
  dev_free(c->Cd);
  dev_free(c->Ce);
  dev_free(c->Xd);
  dev_free(c->Xe);
  dev_free(c->slots);
  dev_free(c->segs);
  dev_free(c->lslots);
  dev_free(c->lsegs);
  dev_free(c->jlive);
  dev_free(c->Bm);
  dev_free(c->perm);
  dev_free(c->pinv);
  dev_free(c->gslot);
  dev_free(c->Bs);
  for (int i = 0; i < kMaxBatch - 1; ++i) {
    dev_free(c->xD[i]);
    dev_free(c->xC[i]);
    dev_free(c->xJ[i]);
    dev_free(c->xK[i]);
  }
  for (auto& ev : c->tev) (void)hipEventDestroy(ev);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_half) (void)hipEventDestroy(c->ev_half);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_done) (void)hipEventDestroy(c->ev_done);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int nopt_jk_get_tuning(const nopt_jk_ctx* c, int* params) {
  nopt::clear_error();
  if (!c || !params) return set_error(NOPT_EINVAL, "nopt_jk_get_tuning: null argument");
  params[0] = c->nsk;
  params[1] = c->nsj;
  params[2] = c->nsk8;
  params[3] = c->nsg;
  params[4] = c->nsj8;
  params[5] = c->cplace;
  return NOPT_OK;
}

int nopt_jk_set_tuning(nopt_jk_ctx* c, const int* params) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!params) return set_error(NOPT_EINVAL, "nopt_jk_set_tuning: null argument");
  if (c->pins > 0)
    return set_error(NOPT_ESTATE, "nopt_jk_set_tuning: a captured HIP graph refers to this context's buffers");
  const int na = (int)c->naux;
  if (params[0] < 1 || params[0] > na || params[1] < 0 || params[1] > na || params[2] < 1 ||
      params[2] > std::max(8, na) || params[3] < 1 || params[3] > na || params[4] < 1 || params[4] > na ||
      params[5] < 0 || params[5] > 2)
    return set_error(NOPT_EINVAL, "nopt_jk_set_tuning: split counts out of range (see nopt_jk.h)");
  c->nsk = params[0];
  c->nsj = params[1];
  c->nsk8 = params[2];
  c->nsg = params[3];
  c->nsj8 = params[4];
  c->cplace = params[5];
  return NOPT_OK;
}

int nopt_jk_autotune(nopt_jk_ctx* c, const double* dD, const double* dC, long nocc, int prec, int reps,
                     double* best_ms, void* stream) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!dD || !dC || reps < 1) return set_error(NOPT_EINVAL, "nopt_jk_autotune: null input or reps < 1");
  if (t_capturing) return set_error(NOPT_ESTATE, "nopt_jk_autotune: not during graph capture");
  if (c->pins > 0)
    return set_error(NOPT_ESTATE, "nopt_jk_autotune: a captured HIP graph refers to this context's buffers");
  hipStream_t s = (hipStream_t)stream;
  rc = order_in(c, s);
  if (rc) return rc;
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // mean time of `reps` builds of the current settings, after one warm-up
  auto measure = [&](float* ms) -> int {
    int r = enqueue_build(c, dD, dC, nocc, c->dJ, c->dK, prec, s);
    if (r) return r;
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps && !r; ++i) r = enqueue_build(c, dD, dC, nocc, c->dJ, c->dK, prec, s);
    if (r) return r;
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    return NOPT_OK;
  };
  const int na = (int)c->naux;
  const bool i8 = prec == NOPT_PREC_INT8;
  // coordinate descent over the split counts the mode uses
  int* knobs[4];
  std::vector<int> cand[4];
  int nk = 0;
  auto pow2 = [&](int lo, int hi) {
    std::vector<int> v;
    for (int x = lo; x <= hi; x *= 2) v.push_back(x);
    return v;
  };
  if (i8) {
    knobs[nk] = &c->nsk8;
    for (int x = 8; x <= std::min(128, std::max(8, na)); x += 8) cand[nk].push_back(x);
    // and the largest counts that give each CU at most 1, 2 or 3 exchange
    // blocks: with multiples of 8 alone some CUs can hold one block more
    // than the rest, and the kernel waits for them
    for (int b = 1; b <= 3; ++b) {
      const int x = b * c->ncu / std::max(1, c->ntiles8);
      if (x >= 1 && x <= std::max(8, na) && x % 8) cand[nk].push_back(x);
    }
    ++nk;
    knobs[nk] = &c->nsj8;
    cand[nk++] = pow2(1, std::min(64, na));
    knobs[nk] = &c->nsg;
    cand[nk++] = pow2(1, std::min(64, na));
    knobs[nk] = &c->cplace;
    cand[nk++] = {0, 1, 2};
  } else {
    knobs[nk] = &c->nsk;
    cand[nk++] = pow2(1, std::min(128, na));
    knobs[nk] = &c->nsj;
    cand[nk++] = pow2(1, std::min(128, na));
  }
  // two rounds: the split counts interact (the Coulomb pair shares the
  // device with the exchange), so the first knob is tried again once the
  // others have moved
  float best = 0.f;
  rc = measure(&best);
  for (int round = 0; round < 2 && !rc; ++round)
    for (int k = 0; k < nk && !rc; ++k) {
      int keep = *knobs[k];
      for (int v : cand[k]) {
        if (v == keep) continue;
        *knobs[k] = v;
        float ms = 0.f;
        rc = measure(&ms);
        if (rc) break;
        if (ms < best) {
          best = ms;
          keep = v;
        }
      }
      *knobs[k] = keep;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = timing;
  if (!rc) rc = order_out(c, s);
  if (!rc && best_ms) *best_ms = best;
  return rc;
}

int nopt_jk_set_timing(nopt_jk_ctx* c, int enable) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (enable && c->tev.empty()) {
    c->tev.resize((size_t)kTimedBuilds * (kPhases + 1));
    for (auto& ev : c->tev) HIPCHK(hipEventCreate(&ev));
  }
  c->timing = enable != 0;
  c->timed = 0;
  return NOPT_OK;
}

int nopt_jk_read_timing(nopt_jk_ctx* c, double* ms, long* builds) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!ms) return set_error(NOPT_EINVAL, "nopt_jk_read_timing: null output");
  for (int p = 0; p < kPhases; ++p) ms[p] = 0.0;
  if (!c->tev.empty())
    for (int b = 0; b < c->timed; ++b) {
      const hipEvent_t* e = &c->tev[(size_t)b * (kPhases + 1)];
      HIPCHK(hipEventSynchronize(e[kPhases]));
      for (int p = 0; p < kPhases; ++p) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e[p], e[p + 1]));
        ms[p] += t;
      }
    }
  if (builds) *builds = c->timed;
  c->timed = 0;
  return NOPT_OK;
}

void* nopt_jk_stream(nopt_jk_ctx* c) { return c ? (void*)c->stream : nullptr; }

struct nopt_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<nopt_jk_ctx*> ctxs;  // contexts whose buffers the graph refers to (pinned)
};

int nopt_graph_capture_begin(int device, void* stream) {
  nopt::clear_error();
  if (!stream) return set_error(NOPT_EINVAL, "nopt_graph_capture_begin: a non-default stream is required");
  if (t_capturing) return set_error(NOPT_ESTATE, "nopt_graph_capture_begin: already capturing on this thread");
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  t_capturing = true;
  t_captured.clear();
  return NOPT_OK;
}

int nopt_graph_capture_end(int device, void* stream, nopt_graph** out) {
  nopt::clear_error();
  if (!stream || !out) return set_error(NOPT_EINVAL, "nopt_graph_capture_end: null stream or output");
  *out = nullptr;
  if (!t_capturing) return set_error(NOPT_ESTATE, "nopt_graph_capture_end: no capture in progress");
  t_capturing = false;
  std::vector<nopt_jk_ctx*> ctxs;
  ctxs.swap(t_captured);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return set_error(NOPT_EHIP, std::string("hipGraphInstantiate: ") + hip_msg(e));
  }
  auto* gr = new (std::nothrow) nopt_graph;
  if (!gr) {
    (void)hipGraphExecDestroy(ex);
    (void)hipGraphDestroy(g);
    return set_error(NOPT_ENOMEM, "nopt_graph_capture_end: host allocation failed");
  }
  gr->device = device;
  gr->graph = g;
  gr->exec = ex;
  gr->ctxs.swap(ctxs);
  for (nopt_jk_ctx* c : gr->ctxs) ++c->pins;
  *out = gr;
  return NOPT_OK;
}

int nopt_graph_launch(nopt_graph* g, void* stream) {
  nopt::clear_error();
  if (!g) return set_error(NOPT_EINVAL, "nopt_graph_launch: null graph");
  HIPCHK(hipSetDevice(g->device));
  hipStream_t s = (hipStream_t)stream;
  for (nopt_jk_ctx* c : g->ctxs) {
    int rc = order_in(c, s);
    if (rc) return rc;
  }
  HIPCHK(hipGraphLaunch(g->exec, s));
  for (nopt_jk_ctx* c : g->ctxs) {
    int rc = order_out(c, s);
    if (rc) return rc;
  }
  return NOPT_OK;
}

void nopt_graph_destroy(nopt_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  for (nopt_jk_ctx* c : g->ctxs) {  // the graph may still run: wait before unpinning
    if (c->done_set) (void)hipEventSynchronize(c->ev_done);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (nopt_jk_ctx* c : g->ctxs)
    if (--c->pins == 0 && c->zombie) nopt_jk_destroy(c);
  delete g;
}

void* nopt_dev_malloc(int device, size_t bytes) {
  void* p = nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error(NOPT_ENOMEM, "nopt_dev_malloc failed");
    return nullptr;
  }
  return p;
}

void nopt_dev_free(void* p) {
  if (p) (void)hipFree(p);
}

// kind: 0 host->device, 1 device->host, 2 device->device
int nopt_dev_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  nopt::clear_error();
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                   : hipMemcpyDeviceToDevice;
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return NOPT_OK;
}

int nopt_dev_synchronize(int device) {
  nopt::clear_error();
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipDeviceSynchronize());
  return NOPT_OK;
}

int nopt_jk_set_B(nopt_jk_ctx* c, const double* eri3, const double* metric) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!eri3 || !metric) return set_error(NOPT_EINVAL, "nopt_jk_set_B: null input");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  const long na = c->naux;
  double* dV = nullptr;
  rc = dev_alloc((void**)&dV, sizeof(double) * na * na, "metric");
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dV, metric, sizeof(double) * na * na, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = nopt::cholesky_f64(dV, na, (int)na, c->info, c->stream);
    int info = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&info, c->info, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      rc = set_error(NOPT_EHIP, std::string("metric factorisation: ") + hip_msg(e));
    else if (info != 0)
      rc = set_error(NOPT_ECHOL, "Coulomb metric is not positive definite (pivot " + std::to_string(info - 1) + ")");
  }
  if (!rc) rc = form_B(c, eri3, dV);
  dev_free(dV);
  return rc;
}

int nopt_jk_set_B_fitted(nopt_jk_ctx* c, const double* Bp) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!Bp) return set_error(NOPT_EINVAL, "nopt_jk_set_B_fitted: null input");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  return form_B(c, Bp, nullptr);
}

int nopt_jk_get_B(nopt_jk_ctx* c, double* Bp) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!Bp) return set_error(NOPT_EINVAL, "nopt_jk_get_B: null output");
  if (!c->B64) return set_error(NOPT_ESTATE, "nopt_jk_get_B: the context keeps no FP64 B (NOPT_CTX_INT8_ONLY)");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  double* dA = nullptr;
  const long total = c->naux * c->npair;
  rc = dev_alloc((void**)&dA, sizeof(double) * total, "packed B");
  if (rc) return rc;
  k_pack_B<<<(unsigned)((total + 255) / 256), 256, 0, c->stream>>>(c->B64, c->ld64, (int)c->nbf, c->npair,
                                                                   c->naux, dA);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(Bp, dA, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_free(dA);
  if (e != hipSuccess) return set_error(NOPT_EHIP, std::string("nopt_jk_get_B: ") + hip_msg(e));
  return NOPT_OK;
}

int nopt_jk_active_ints(nopt_jk_ctx* c, const double* C1, const double* C2, long na, double* ints) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!C1 || !C2 || !ints) return set_error(NOPT_EINVAL, "nopt_jk_active_ints: null pointer");
  if (na < 1 || na > 32) return set_error(NOPT_EINVAL, "nopt_jk_active_ints: need 1 <= na <= 32");
  if (!c->B64) return set_error(NOPT_ESTATE, "nopt_jk_active_ints: the context keeps no FP64 B (NOPT_CTX_INT8_ONLY)");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  const long nbf = c->nbf, naux = c->naux, nn = na * na;
  double *dC = nullptr, *dR = nullptr, *dI = nullptr;
  rc = dev_alloc((void**)&dC, sizeof(double) * 2 * na * nbf, "active orbitals");
  if (!rc) rc = dev_alloc((void**)&dR, sizeof(double) * naux * nn, "active half-transform");
  if (!rc) rc = dev_alloc((void**)&dI, sizeof(double) * nn * nn, "active integrals");
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipMemcpyAsync(dC, C1, sizeof(double) * na * nbf, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dC + na * nbf, C2, sizeof(double) * na * nbf, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      const double* d1 = dC;
      const double* d2 = dC + na * nbf;
      if (na <= 8)
        k_active_half<8><<<(unsigned)naux, 256, 0, c->stream>>>(c->B64, c->ld64, (int)nbf, d1, d2, (int)na, dR);
      else if (na <= 16)
        k_active_half<16><<<(unsigned)naux, 256, 0, c->stream>>>(c->B64, c->ld64, (int)nbf, d1, d2, (int)na, dR);
      else
        k_active_half<32><<<(unsigned)naux, 256, 0, c->stream>>>(c->B64, c->ld64, (int)nbf, d1, d2, (int)na, dR);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      k_active_ints<<<(unsigned)((nn * nn + 255) / 256), 256, 0, c->stream>>>(dR, naux, (int)nn, dI);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ints, dI, sizeof(double) * nn * nn, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  dev_free(dC);
  dev_free(dR);
  dev_free(dI);
  if (rc) return rc;
  if (e != hipSuccess) return set_error(NOPT_EHIP, std::string("nopt_jk_active_ints: ") + hip_msg(e));
  return NOPT_OK;
}

int nopt_jk_zero_blocks(const nopt_jk_ctx* c, double out[3]) {
  if (!c || !out) return set_error(NOPT_EINVAL, "nopt_jk_zero_blocks: null argument");
  out[0] = out[1] = out[2] = 1.0;
  if (!c->b_ready || !c->Bs) return NOPT_OK;
  out[0] = (double)c->nlseg / c->nseg;
  out[1] = (double)c->nlslots / c->nslots;
  if (c->Bm) {
    std::vector<uint32_t> h((size_t)c->naux * c->nmb);
    if (hipSetDevice(c->device) != hipSuccess ||
        hipMemcpy(h.data(), c->Bm, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return set_error(NOPT_EHIP, "nopt_jk_zero_blocks: copy of the slab masks failed");
    long bits = 0;
    for (uint32_t v : h) bits += __builtin_popcount(v);
    out[2] = (double)bits / ((double)h.size() * (double)(c->ld8 / 64));
  }
  return NOPT_OK;
}

long nopt_jk_device_bytes(const nopt_jk_ctx* c) {
  if (!c) return 0;
  long b = 0;
  if (c->B64) b += sizeof(double) * c->naux * c->nbf * c->ld64;
  if (c->Bd) b += 4L * c->naux * c->nbf * c->ld8;
  if (c->Bs) b += sizeof(int) * c->naux * c->nbf + c->naux * c->ld8 + 5L * c->splane;
  if (c->Bm) b += sizeof(uint32_t) * c->naux * c->nmb;
  b += (long)(c->X_bytes + c->Cp_bytes + c->Cd_bytes + c->Ce_bytes + c->Xd_bytes + c->Xe_bytes + c->Jpart_bytes);
  b += (long)c->Kpart_bytes;
  b += sizeof(double) * (c->naux * c->mtiles64 + kMaxBatch * c->naux + kMaxBatch * c->dstride + 5 * c->nbf * c->nbf);
  return b;
}

int nopt_jk_build_device(nopt_jk_ctx* c, const double* dD, const double* dC, long nocc, double* dJ,
                         double* dK, int prec, void* stream) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
  rc = order_in(c, s);
  if (!rc) rc = enqueue_build(c, dD, dC, nocc, dJ, dK, prec, s);
  if (!rc) rc = order_out(c, s);
  return rc;
}

int nopt_jk_build_device_static(nopt_jk_ctx* c, const double* dD, const double* dC, long nocc, double* dJ,
                                double* dK, int prec, void* stream) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  void* work = nullptr;
  rc = scratch_address(c->device, &work);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
  const BuildItem x{dD, dC, nocc, dJ, dK};
  rc = order_in(c, s);
  if (!rc) rc = enqueue_batch(c, 1, &x, prec, s, work, kScratchBytes);
  if (!rc) rc = order_out(c, s);
  return rc;
}

long nopt_jk_static_scratch_bytes(void) { return (long)kScratchBytes; }

int nopt_jk_build_batch_device(nopt_jk_ctx* c, int nd, const double* const* dD, const double* const* dC,
                               const long* nocc, double* const* dJ, double* const* dK, int prec, void* stream) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (nd < 1 || nd > kMaxBatch) return set_error(NOPT_EINVAL, "batch size must be between 1 and NOPT_MAX_BATCH");
  if (!nocc) return set_error(NOPT_EINVAL, "nopt_jk_build_batch_device: null nocc array");
  BuildItem it[kMaxBatch];
  for (int d = 0; d < nd; ++d)
    it[d] = BuildItem{dD ? dD[d] : nullptr, dC ? dC[d] : nullptr, nocc[d], dJ ? dJ[d] : nullptr,
                      dK ? dK[d] : nullptr};
  hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
  rc = order_in(c, s);
  if (!rc) rc = enqueue_batch(c, nd, it, prec, s);
  if (!rc) rc = order_out(c, s);
  return rc;
}

int nopt_jk_build_ragged_device(int nsys, nopt_jk_ctx* const* ctxs, const int* nd, const double* const* dD,
                                const double* const* dC, const long* nocc, double* const* dJ, double* const* dK,
                                int prec, void* stream) {
  nopt::clear_error();
  if (nsys < 1 || !ctxs || !nd || !nocc) return set_error(NOPT_EINVAL, "nopt_jk_build_ragged_device: empty batch");
  const int dev = ctxs[0] ? ctxs[0]->device : 0;
  for (int k = 0; k < nsys; ++k) {  // every argument checked before any work is enqueued
    int rc = check_ctx(ctxs[k], true);
    if (rc) return rc;
    if (ctxs[k]->device != dev)
      return set_error(NOPT_EINVAL, "nopt_jk_build_ragged_device: all contexts must be on one device");
    if (nd[k] < 1 || nd[k] > kMaxBatch)
      return set_error(NOPT_EINVAL, "nopt_jk_build_ragged_device: 1 <= nd[k] <= NOPT_MAX_BATCH");
  }
  HIPCHK(hipSetDevice(dev));
  hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
  long first = 0;
  for (int k = 0; k < nsys; ++k) {
    nopt_jk_ctx* c = ctxs[k];
    BuildItem it[kMaxBatch];
    for (int d = 0; d < nd[k]; ++d) {
      const long i = first + d;
      it[d] = BuildItem{dD ? dD[i] : nullptr, dC ? dC[i] : nullptr, nocc[i], dJ ? dJ[i] : nullptr,
                        dK ? dK[i] : nullptr};
    }
    first += nd[k];
    int rc = order_in(c, s);
    if (!rc) rc = enqueue_batch(c, nd[k], it, prec, s);
    if (!rc) rc = order_out(c, s);
    if (rc) return rc;
  }
  return NOPT_OK;
}

}  // extern "C"

struct nopt_ragged {
  int device = 0;
  std::vector<nopt_jk_ctx*> ctxs;
  void* table = nullptr;  // item arrays and block offsets of every launch
  struct Launch {
    int phase, kind, cfg, n, blocks;
    size_t items, first;  // byte offsets into table
  };
  std::vector<Launch> launches;
  // the Coulomb phases (gamma, J, J output) run on a second stream beside
  // the exchange, forked after the half-transform and joined at the end
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {
int launch_ragged_list(const nopt_ragged::Launch& l, const char* base, hipStream_t s) {
  const void* items = base + l.items;
  const int* first = reinterpret_cast<const int*>(base + l.first);
  if (l.kind == kListPad) {
    k_pad_copy_list<<<l.blocks, 256, 0, s>>>((const ItemPad*)items, first, l.n);
  } else if (l.kind == kListFin) {
    k_finalize_list<<<l.blocks, 256, 0, s>>>((const ItemFin*)items, first, l.n);
  } else {
    HIPCHK(nopt::launch_list(l.kind, l.cfg, items, first, l.n, l.blocks, s));
  }
  HIPCHK(hipGetLastError());
  return NOPT_OK;
}
}  // namespace

extern "C" {

int nopt_ragged_create(int nsys, nopt_jk_ctx* const* ctxs, const int* nd, const double* const* dD,
                       const double* const* dC, const long* nocc, double* const* dJ, double* const* dK,
                       nopt_ragged** out) {
  nopt::clear_error();
  if (!out) return set_error(NOPT_EINVAL, "nopt_ragged_create: null output");
  *out = nullptr;
  if (nsys < 1 || !ctxs || !nd || !nocc) return set_error(NOPT_EINVAL, "nopt_ragged_create: empty batch");
  if (t_capturing) return set_error(NOPT_ESTATE, "nopt_ragged_create: not during graph capture");
  const int dev = ctxs[0] ? ctxs[0]->device : 0;
  for (int k = 0; k < nsys; ++k) {
    int rc = check_ctx(ctxs[k], true);
    if (rc) return rc;
    if (ctxs[k]->device != dev) return set_error(NOPT_EINVAL, "nopt_ragged_create: all contexts must be on one device");
    if (nd[k] < 1 || nd[k] > kMaxBatch) return set_error(NOPT_EINVAL, "nopt_ragged_create: 1 <= nd[k] <= NOPT_MAX_BATCH");
    for (int j = 0; j < k; ++j)
      if (ctxs[j] == ctxs[k]) return set_error(NOPT_EINVAL, "nopt_ragged_create: a context appears twice");
  }
  HIPCHK(hipSetDevice(dev));
  // record every system's build (buffers grow here, outside any capture)
  Rec rec;
  long first = 0;
  for (int k = 0; k < nsys; ++k) {
    nopt_jk_ctx* c = ctxs[k];
    if (c->done_set) HIPCHK(hipEventSynchronize(c->ev_done));
    BuildItem it[kMaxBatch];
    for (int d = 0; d < nd[k]; ++d) {
      const long i = first + d;
      it[d] = BuildItem{dD ? dD[i] : nullptr, dC ? dC[i] : nullptr, nocc[i], dJ ? dJ[i] : nullptr,
                        dK ? dK[i] : nullptr};
    }
    first += nd[k];
    const size_t before = rec.items.size();
    int rc = enqueue_batch(c, nd[k], it, NOPT_PREC_INT8, c->stream, nullptr, 0, &rec);
    if (rc) return rc;
    for (size_t i = before; i < rec.items.size(); ++i)
      if (rec.items[i].r.blocks < 1) return set_error(NOPT_EINVAL, "nopt_ragged_create: empty work item");
  }
  // launches: the items of one (phase, kernel, configuration), in phase order
  std::stable_sort(rec.items.begin(), rec.items.end(), [](const RecItem& a, const RecItem& b) {
    if (a.phase != b.phase) return a.phase < b.phase;
    if (a.r.kind != b.r.kind) return a.r.kind < b.r.kind;
    return a.r.cfg < b.r.cfg;
  });
  auto* p = new (std::nothrow) nopt_ragged;
  if (!p) return set_error(NOPT_ENOMEM, "nopt_ragged_create: host allocation failed");
  p->device = dev;
  std::vector<char> host;
  auto align = [&](size_t a) { host.resize((host.size() + a - 1) / a * a); };
  for (size_t i = 0; i < rec.items.size();) {
    size_t j = i;
    while (j < rec.items.size() && rec.items[j].phase == rec.items[i].phase &&
           rec.items[j].r.kind == rec.items[i].r.kind && rec.items[j].r.cfg == rec.items[i].r.cfg)
      ++j;
    nopt_ragged::Launch l{rec.items[i].phase, rec.items[i].r.kind, rec.items[i].r.cfg, (int)(j - i), 0, 0, 0};
    align(256);
    l.items = host.size();
    for (size_t k = i; k < j; ++k) host.insert(host.end(), rec.items[k].r.args.begin(), rec.items[k].r.args.end());
    align(256);
    l.first = host.size();
    std::vector<int> fb(j - i + 1, 0);
    long tot = 0;
    for (size_t k = i; k < j; ++k) {
      fb[k - i] = (int)tot;
      tot += rec.items[k].r.blocks;
    }
    if (tot > 0x7fffffffL) {
      delete p;
      return set_error(NOPT_EINVAL, "nopt_ragged_create: too many blocks in one launch");
    }
    fb[j - i] = (int)tot;
    l.blocks = (int)tot;
    const char* fp = reinterpret_cast<const char*>(fb.data());
    host.insert(host.end(), fp, fp + sizeof(int) * fb.size());
    p->launches.push_back(l);
    i = j;
  }
  int rc = dev_alloc(&p->table, std::max<size_t>(host.size(), 256), "ragged plan tables");
  if (!rc && hipMemcpy(p->table, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(NOPT_EHIP, "nopt_ragged_create: copy of the work lists failed");
  if (!rc && (hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess))
    rc = set_error(NOPT_EHIP, "nopt_ragged_create: stream or event creation failed");
  if (rc) {
    dev_free(p->table);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->side) (void)hipStreamDestroy(p->side);
    delete p;
    return rc;
  }
  for (int k = 0; k < nsys; ++k) {
    p->ctxs.push_back(ctxs[k]);
    ++ctxs[k]->pins;
  }
  *out = p;
  return NOPT_OK;
}

int nopt_ragged_launch(nopt_ragged* p, void* stream) {
  nopt::clear_error();
  if (!p) return set_error(NOPT_EINVAL, "nopt_ragged_launch: null plan");
  HIPCHK(hipSetDevice(p->device));
  hipStream_t s = (hipStream_t)stream;
  for (nopt_jk_ctx* c : p->ctxs) {
    int rc = order_in(c, s);
    if (rc) return rc;
  }
  // phases 0-2 (staging, C digits, half-transform) and 7-8 (exchange, K
  // output) on the caller's stream, 3-6 (gamma, J, J output) on the side
  // stream from the end of the half-transform on
  bool forked = false;
  for (const auto& l : p->launches) {
    const bool coulomb = l.phase >= 3 && l.phase <= 6;
    if (coulomb && !forked) {
      HIPCHK(hipEventRecord(p->ev_fork, s));
      HIPCHK(hipStreamWaitEvent(p->side, p->ev_fork, 0));
      forked = true;
    }
    int rc = launch_ragged_list(l, (const char*)p->table, coulomb ? p->side : s);
    if (rc) return rc;
  }
  if (forked) {
    HIPCHK(hipEventRecord(p->ev_join, p->side));
    HIPCHK(hipStreamWaitEvent(s, p->ev_join, 0));
  }
  for (nopt_jk_ctx* c : p->ctxs) {
    int rc = order_out(c, s);
    if (rc) return rc;
  }
  return NOPT_OK;
}

int nopt_ragged_launches(const nopt_ragged* p) { return p ? (int)p->launches.size() : 0; }

void nopt_ragged_destroy(nopt_ragged* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  for (nopt_jk_ctx* c : p->ctxs)
    if (c->done_set) (void)hipEventSynchronize(c->ev_done);
  if (p->side) (void)hipStreamSynchronize(p->side);
  dev_free(p->table);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  for (nopt_jk_ctx* c : p->ctxs)
    if (--c->pins == 0 && c->zombie) nopt_jk_destroy(c);
  delete p;
}

int nopt_jk_build_batch(nopt_jk_ctx* c, int nd, const double* const* D, const double* const* C, const long* nocc,
                        double* const* J, double* const* K, int prec) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (nd < 1 || nd > kMaxBatch) return set_error(NOPT_EINVAL, "batch size must be between 1 and NOPT_MAX_BATCH");
  if (!nocc) return set_error(NOPT_EINVAL, "nopt_jk_build_batch: null nocc array");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  const long nn = c->nbf * c->nbf;
  BuildItem it[kMaxBatch];
  for (int d = 0; d < nd; ++d) {
    const double* Dh = D ? D[d] : nullptr;
    const double* Ch = C ? C[d] : nullptr;
    double* Jh = J ? J[d] : nullptr;
    double* Kh = K ? K[d] : nullptr;
    const long no = nocc[d];
    if (no < 0 || no > c->nbf) return set_error(NOPT_EINVAL, "nocc out of range");
    if (Jh && !Dh) return set_error(NOPT_EINVAL, "J requested without a density");
    if (Kh && no > 0 && !Ch) return set_error(NOPT_EINVAL, "K requested without an orbital factor");
    // slot d: device staging of D, C and the outputs
    double** sD = d ? &c->xD[d - 1] : &c->dD;
    double** sC = d ? &c->xC[d - 1] : &c->dC;
    double** sJ = d ? &c->xJ[d - 1] : &c->dJ;
    double** sK = d ? &c->xK[d - 1] : &c->dK;
    long dcap = nn, jcap = nn, kcap = nn;
    long* capD = d ? &c->xD_cap[d - 1] : &dcap;
    long* capC = d ? &c->xC_cap[d - 1] : &c->dC_cap;
    long* capJ = d ? &c->xJ_cap[d - 1] : &jcap;
    long* capK = d ? &c->xK_cap[d - 1] : &kcap;
    if (Dh) {
      rc = upload(c, sD, capD, Dh, nn);
      if (rc) return rc;
    }
    if (Ch && no > 0) {
      rc = upload(c, sC, capC, Ch, no * c->nbf);
      if (rc) return rc;
    }
    if (Jh && (*capJ < nn || !*sJ)) {
      dev_free(*sJ);
      rc = dev_alloc((void**)sJ, sizeof(double) * nn, "J");
      if (rc) return rc;
      *capJ = nn;
    }
    if (Kh && (*capK < nn || !*sK)) {
      dev_free(*sK);
      rc = dev_alloc((void**)sK, sizeof(double) * nn, "K");
      if (rc) return rc;
      *capK = nn;
    }
    it[d] = BuildItem{Dh ? *sD : nullptr, (Ch && no > 0) ? *sC : nullptr, no, Jh ? *sJ : nullptr,
                      Kh ? *sK : nullptr};
  }
  rc = enqueue_batch(c, nd, it, prec, c->stream);
  if (rc) return rc;
  for (int d = 0; d < nd; ++d) {
    if (J && J[d]) HIPCHK(hipMemcpyAsync(J[d], it[d].J, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    if (K && K[d]) HIPCHK(hipMemcpyAsync(K[d], it[d].K, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return NOPT_OK;
}

int nopt_jk_build(nopt_jk_ctx* c, const double* D, const double* C, long nocc, double* J, double* K,
                  int prec) {
  return nopt_jk_build_batch(c, 1, &D, &C, &nocc, &J, &K, prec);
}

int nopt_jk_fock_energy(nopt_jk_ctx* c, const double* H, const double* D, const double* J,
                        const double* K, double cj, double ck, double* F, double* E) {
  nopt::clear_error();
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!H || !D) return set_error(NOPT_EINVAL, "nopt_jk_fock_energy: H and D are required");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  const long nn = c->nbf * c->nbf;
  HIPCHK(hipMemcpyAsync(c->dH, H, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dD, D, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  if (J) HIPCHK(hipMemcpyAsync(c->dJ, J, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  if (K) HIPCHK(hipMemcpyAsync(c->dK, K, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  const int nblk = (int)std::min<long>(1024, (nn + 255) / 256);
  k_fock_energy<<<nblk, 256, 0, c->stream>>>(c->dH, c->dD, J ? c->dJ : nullptr, K ? c->dK : nullptr, cj, ck,
                                             nn, c->dF, c->epart);
  k_sum<<<1, 256, 0, c->stream>>>(c->epart, nblk, c->dE);
  HIPCHK(hipGetLastError());
  if (F) HIPCHK(hipMemcpyAsync(F, c->dF, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  double e = 0.0;
  HIPCHK(hipMemcpyAsync(&e, c->dE, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (E) *E = e;
  return NOPT_OK;
}

namespace {
// E, G = 4 C F and F on device buffers, enqueued on s (the context's D, J, K
// buffers serve as scratch). dG and dF may be null.
int enqueue_energy_gradient(nopt_jk_ctx* c, const double* dH, const double* dC, long nocc, int prec, double* dE,
                            double* dG, double* dF, hipStream_t s) {
  const long nbf = c->nbf, nn = nbf * nbf;
  double* F = dF ? dF : c->dF;
  k_density<<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(dC, (int)nocc, (int)nbf, c->dD);
  HIPCHK(hipGetLastError());
  int rc = enqueue_build(c, c->dD, dC, nocc, c->dJ, c->dK, prec, s);
  if (rc) return rc;
  const int nblk = (int)std::min<long>(1024, (nn + 255) / 256);
  k_fock_energy<<<nblk, 256, 0, s>>>(dH, c->dD, c->dJ, c->dK, 2.0, 1.0, nn, F, c->epart);
  k_sum<<<1, 256, 0, s>>>(c->epart, nblk, dE);
  HIPCHK(hipGetLastError());
  // G (nocc x nbf) = 4 C F
  if (dG) HIPCHK(nopt::gemm_f64(false, (int)nocc, (int)nbf, (int)nbf, 4.0, dC, nbf, F, nbf, 0.0, dG, nbf, s));
  return NOPT_OK;
}
}  // namespace

int nopt_jk_energy_gradient(nopt_jk_ctx* c, const double* H, const double* C, long nocc, int prec,
                            double* E, double* G, double* F) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!H || !C || nocc <= 0 || nocc > c->nbf)
    return set_error(NOPT_EINVAL, "nopt_jk_energy_gradient: H, C and 0 < nocc <= nbf are required");
  rc = order_in(c, c->stream);
  if (rc) return rc;
  const long nbf = c->nbf, nn = nbf * nbf;
  rc = upload(c, &c->dC, &c->dC_cap, C, nocc * nbf);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->dH, H, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  if (G && !c->dG) {
    rc = dev_alloc((void**)&c->dG, sizeof(double) * nn, "gradient");
    if (rc) return rc;
  }
  rc = enqueue_energy_gradient(c, c->dH, c->dC, nocc, prec, c->dE, G ? c->dG : nullptr, nullptr, c->stream);
  if (rc) return rc;
  if (G) HIPCHK(hipMemcpyAsync(G, c->dG, sizeof(double) * nocc * nbf, hipMemcpyDeviceToHost, c->stream));
  if (F) HIPCHK(hipMemcpyAsync(F, c->dF, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  double e = 0.0;
  HIPCHK(hipMemcpyAsync(&e, c->dE, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (E) *E = e;
  return NOPT_OK;
}

int nopt_jk_energy_gradient_device(nopt_jk_ctx* c, const double* dH, const double* dC, long nocc, int prec,
                                   double* dE, double* dG, double* dF, void* stream) {
  nopt::clear_error();
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!dH || !dC || !dE || nocc <= 0 || nocc > c->nbf)
    return set_error(NOPT_EINVAL, "nopt_jk_energy_gradient_device: H, C, E and 0 < nocc <= nbf are required");
  hipStream_t s = (hipStream_t)stream;
  rc = order_in(c, s);
  if (!rc) rc = enqueue_energy_gradient(c, dH, dC, nocc, prec, dE, dG, dF, s);
  if (!rc) rc = order_out(c, s);
  return rc;
}

}  // extern "C"
