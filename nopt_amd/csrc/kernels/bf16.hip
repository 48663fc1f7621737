// Split-bf16 MFMA kernels for the half-transform and the exchange SYRK.
//
// An FP32 value is written as a sum of NS bf16 pieces (x = p0 + p1 [+ p2]);
// with NS = 3 the split of an FP32 value is exact (8 + 8 + 8 significant
// bits). A product a*b is formed from the piece products whose combined weight
// is at least 2^-16 (NS = 2: three products, "bf16x3") or 2^-24 (NS = 3: six
// products, "bf16x6"), smallest first, on v_mfma_f32_16x16x32_bf16. Each
// 32-wide k-step is one FP32 chain, added into FP64 accumulators, so the
// accuracy is set by the representation, not by the accumulation length.
#include "bf16.hpp"

namespace nopt {
namespace {

constexpr int BM = 128;  // rows per half-transform tile
constexpr int KT = 128;  // exchange tile edge
constexpr int BK = 32;   // k per MFMA step
constexpr int PK = BK + 8;

template <int NS>
__device__ __forceinline__ void split_n(float x, uint16_t* p) {
  p[0] = f2bf(x);
  if constexpr (NS >= 2) {
    float r = x - bf2f(p[0]);
    p[1] = f2bf(r);
    if constexpr (NS >= 3) {
      r -= bf2f(p[1]);
      p[2] = f2bf(r);
    }
  }
}

// Piece products (ia, ib), smallest weight first.
template <int NS>
struct Prods;
template <>
struct Prods<2> {
  static constexpr int n = 3;
  static constexpr int a[3] = {1, 0, 0};
  static constexpr int b[3] = {0, 1, 0};
};
template <>
struct Prods<3> {
  static constexpr int n = 6;
  static constexpr int a[6] = {2, 1, 0, 1, 0, 0};
  static constexpr int b[6] = {0, 1, 2, 0, 1, 0};
};

// Stages 16 consecutive k values of one row into NS LDS piece arrays.
template <int NS, int ROWS>
__device__ __forceinline__ void stage16(uint16_t (*S)[ROWS][PK], int row, int k, const float* v) {
  uint16_t pc[NS][16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    uint16_t t[NS];
    split_n<NS>(v[i], t);
#pragma unroll
    for (int s = 0; s < NS; ++s) pc[s][i] = t[s];
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    uint4* d = reinterpret_cast<uint4*>(&S[s][row][k]);
    d[0] = __builtin_bit_cast(uint4, *reinterpret_cast<uint16_t(*)[8]>(&pc[s][0]));
    d[1] = __builtin_bit_cast(uint4, *reinterpret_cast<uint16_t(*)[8]>(&pc[s][8]));
  }
}

__device__ __forceinline__ void load8d(const double* p, double* o) {
  const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double2 x = q[i];
    o[2 * i] = x.x;
    o[2 * i + 1] = x.y;
  }
}
__device__ __forceinline__ void load8f(const float* p, float* o) {
  const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    float4 x = q[i];
    o[4 * i] = x.x;
    o[4 * i + 1] = x.y;
    o[4 * i + 2] = x.z;
    o[4 * i + 3] = x.w;
  }
}

// X[Q][m][col] = sum_n B[Q][m][n] Cp[col][n] (B FP64, Cp FP32), and the FP64
// gamma partials gpart[Q][mtile] = sum over the tile of B[Q][m][n] D[m][n].
template <int NS, int NT>
__global__ __launch_bounds__(256) void k_half_bf16(const double* __restrict__ B, long ld, int nbf,
                                                   const float* __restrict__ Cp,
                                                   const double* __restrict__ Dp,
                                                   float* __restrict__ X, long ldx,
                                                   double* __restrict__ gpart, int mtiles) {
  constexpr int BN = 16 * NT;
  __shared__ __attribute__((aligned(16))) uint16_t As[NS][BM][PK];
  __shared__ __attribute__((aligned(16))) uint16_t Cs[NS][BN][PK];
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int mtile = blockIdx.x;
  const long Q = blockIdx.y;
  const int cg = blockIdx.z;
  const int m0 = mtile * BM;
  const int ar = t >> 1, ak = (t & 1) * 16;
  const bool arow_ok = (m0 + ar) < nbf;
  const double* Arow = B + ((long)Q * nbf + m0 + ar) * ld + ak;
  const double* Drow = Dp + (long)(m0 + ar) * ld + ak;
  const bool c_ok = t < 2 * BN;
  const float* Crow = Cp + ((long)cg * BN + (t >> 1)) * ld + ak;
  f4_t acc[2][NT];
  d4_t acc64[2][NT];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      acc[i][j] = f4_t{0, 0, 0, 0};
      acc64[i][j] = d4_t{0, 0, 0, 0};
    }
  double rd = 0.0;
  double a_reg[16], d_reg[16];
  float c_reg[16];
  auto gload = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool kin = k0 + ak + 8 * h + 8 <= ld;
      if (arow_ok && kin) {
        load8d(Arow + k0 + 8 * h, a_reg + 8 * h);
        load8d(Drow + k0 + 8 * h, d_reg + 8 * h);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          a_reg[8 * h + i] = 0.0;
          d_reg[8 * h + i] = 0.0;
        }
      }
      if (c_ok) {
        if (kin)
          load8f(Crow + k0 + 8 * h, c_reg + 8 * h);
        else
#pragma unroll
          for (int i = 0; i < 8; ++i) c_reg[8 * h + i] = 0.f;
      }
    }
  };
  gload(0);
  for (int k0 = 0; k0 < ld; k0 += BK) {
    __syncthreads();
    {
      float af[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        af[i] = (float)a_reg[i];
        rd += a_reg[i] * d_reg[i];
      }
      stage16<NS, BM>(As, ar, ak, af);
    }
    if (c_ok) stage16<NS, BN>(Cs, t >> 1, ak, c_reg);
    __syncthreads();
    if (k0 + BK < ld) gload(k0 + BK);
    uint4 av[NS][2];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        av[s][i] = *reinterpret_cast<const uint4*>(&As[s][w * 32 + i * 16 + lane % 16][8 * (lane / 16)]);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      uint4 bv[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        bv[s] = *reinterpret_cast<const uint4*>(&Cs[s][j * 16 + lane % 16][8 * (lane / 16)]);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int p = 0; p < Prods<NS>::n; ++p)
          acc[i][j] = mfma_bf16(av[Prods<NS>::a[p]][i], bv[Prods<NS>::b[p]], acc[i][j]);
#pragma unroll
        for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)acc[i][j][v];
        acc[i][j] = f4_t{0, 0, 0, 0};
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + w * 32 + i * 16 + 4 * (lane / 16) + v;
      if (m < nbf) {
        float* xr = X + ((long)Q * nbf + m) * ldx + (long)cg * BN + lane % 16;
#pragma unroll
        for (int j = 0; j < NT; ++j) xr[j * 16] = (float)acc64[i][j][v];
      }
    }
  if (cg == 0) {
    double s = block_sum256(rd, red);
    if (t == 0) gpart[Q * mtiles + mtile] = s;
  }
}

__device__ __forceinline__ void tri_tile(int t, int& bi, int& bj) {
  bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  while (bi * (bi + 1) / 2 > t) --bi;
  bj = t - bi * (bi + 1) / 2;
}

// Kpart[s][tile][r][c] = sum_{Q in split s} sum_i X[Q][bi*KT+r][i] X[Q][bj*KT+c][i]
template <int NS>
__global__ __launch_bounds__(256) void k_syrk_bf16(const float* __restrict__ X, long ldx, int nbf,
                                                   int naux, int kpad, double* __restrict__ Kpart) {
  __shared__ __attribute__((aligned(16))) uint16_t As[NS][KT][PK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[NS][KT][PK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  int bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  const int s = blockIdx.y, ns = gridDim.y;
  const int q0 = (int)((long)naux * s / ns), q1 = (int)((long)naux * (s + 1) / ns);
  const int r = t >> 1, kk = (t & 1) * 16;
  const bool ra = bi * KT + r < nbf, rb = bj * KT + r < nbf;
  f4_t acc[4][4];
  d4_t acc64[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = f4_t{0, 0, 0, 0};
      acc64[i][j] = d4_t{0, 0, 0, 0};
    }
  float a_reg[16], b_reg[16];
  auto gload = [&](int q, int i0) {
    const float* xa = X + ((long)q * nbf + bi * KT + r) * ldx + i0 + kk;
    const float* xb = X + ((long)q * nbf + bj * KT + r) * ldx + i0 + kk;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool inr = i0 + kk + 8 * h + 8 <= ldx;  // ldx is a multiple of 16
      if (ra && inr)
        load8f(xa + 8 * h, a_reg + 8 * h);
      else
#pragma unroll
        for (int i = 0; i < 8; ++i) a_reg[8 * h + i] = 0.f;
      if (rb && inr)
        load8f(xb + 8 * h, b_reg + 8 * h);
      else
#pragma unroll
        for (int i = 0; i < 8; ++i) b_reg[8 * h + i] = 0.f;
    }
  };
  if (q0 < q1) gload(q0, 0);
  for (int q = q0; q < q1; ++q) {
    for (int i0 = 0; i0 < kpad; i0 += BK) {
      __syncthreads();
      stage16<NS, KT>(As, r, kk, a_reg);
      stage16<NS, KT>(Bs, r, kk, b_reg);
      __syncthreads();
      int nq = q, ni = i0 + BK;
      if (ni >= kpad) {
        ni = 0;
        ++nq;
      }
      if (nq < q1) gload(nq, ni);
      uint4 av[NS][4];
#pragma unroll
      for (int sp = 0; sp < NS; ++sp)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          av[sp][i] = *reinterpret_cast<const uint4*>(&As[sp][wr * 64 + i * 16 + lane % 16][8 * (lane / 16)]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uint4 bv[NS];
#pragma unroll
        for (int sp = 0; sp < NS; ++sp)
          bv[sp] = *reinterpret_cast<const uint4*>(&Bs[sp][wc * 64 + j * 16 + lane % 16][8 * (lane / 16)]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
          for (int p = 0; p < Prods<NS>::n; ++p)
            acc[i][j] = mfma_bf16(av[Prods<NS>::a[p]][i], bv[Prods<NS>::b[p]], acc[i][j]);
#pragma unroll
          for (int v = 0; v < 4; ++v) acc64[i][j][v] += (double)acc[i][j][v];
          acc[i][j] = f4_t{0, 0, 0, 0};
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
        const int rr = wr * 64 + i * 16 + 4 * (lane / 16) + v;
        const int cc = wc * 64 + j * 16 + lane % 16;
        Kt[rr * KT + cc] = acc64[i][j][v];
      }
}

}  // namespace

hipError_t launch_half_bf16(int ns, int nt, dim3 grid, hipStream_t s, const double* B, long ld, int nbf,
                            const float* Cp, const double* Dp, float* X, long ldx, double* gpart,
                            int mtiles) {
#define NOPT_HB(NSV, NTV)                                                                      \
  if (ns == NSV && nt == NTV) {                                                               \
    k_half_bf16<NSV, NTV><<<grid, 256, 0, s>>>(B, ld, nbf, Cp, Dp, X, ldx, gpart, mtiles);   \
    return hipGetLastError();                                                                 \
  }
  NOPT_HB(2, 1) NOPT_HB(2, 2) NOPT_HB(2, 4) NOPT_HB(2, 6) NOPT_HB(2, 8)
  NOPT_HB(3, 1) NOPT_HB(3, 2) NOPT_HB(3, 4) NOPT_HB(3, 6) NOPT_HB(3, 8)
#undef NOPT_HB
  return hipErrorInvalidValue;
}

hipError_t launch_syrk_bf16(int ns, dim3 grid, hipStream_t s, const float* X, long ldx, int nbf, int naux,
                            int kpad, double* Kpart) {
  if (ns == 2)
    k_syrk_bf16<2><<<grid, 256, 0, s>>>(X, ldx, nbf, naux, kpad, Kpart);
  else if (ns == 3)
    k_syrk_bf16<3><<<grid, 256, 0, s>>>(X, ldx, nbf, naux, kpad, Kpart);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace nopt
