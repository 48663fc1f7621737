// Integer-split (Ozaki-type) half-transform on the int8 matrix cores.
//
// Every row r of an operand is scaled by a power of two, 2^-e_r with
// e_r = exponent of the row's largest magnitude, and the scaled row
// v = x 2^-e_r (|v| < 1) is written as signed base-128 digits
//   v = 2^-6 d_0 + 2^-13 d_1 + 2^-20 d_2 + ... ,  d_k = rint(...) in [-64, 64],
// each digit rounded to nearest, so the truncation error is unbiased and
// below 2^-(6+7S) of the row maximum for S digits. A product of two split
// operands is the weighted sum of digit-plane products; each digit-plane
// product is an exact int32 dot product on v_mfma_i32_16x16x64_i8, so the
// accumulation itself adds no rounding error at all. Digit pairs (k, l) of
// equal weight k+l share one int32 accumulator.
//
// The half-transform keeps the pairs with k + l <= 3 of four digits each
// (relative truncation about 2^-28); the fifth digit plane of B is kept for
// the FP64 gamma partials (gamma_Q = sum_mn B D needs more than 28 bits).
#include "ozaki.hpp"

namespace nopt {
namespace {

typedef int v4i_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4i_t mfma_i8(const uint4& a, const uint4& b, v4i_t c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(v4i_t, a), __builtin_bit_cast(v4i_t, b), c,
                                               0, 0, 0);
}

// row maximum -> exponent e with |x| * 2^-e < 1 (0 for an all-zero row)
__device__ __forceinline__ int row_exponent(double mx) {
  if (!(mx > 0.0)) return 0;
  int e;
  (void)frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1)
  return e;
}

// Splits the rows of a row-major FP64 matrix (rows x cols, ld) into ns int8
// digit planes planes[k][row][ldo] (zero padded to ldo) and row exponents.
// One wavefront per row.
__global__ __launch_bounds__(64) void k_split_rows(const double* __restrict__ src, long ld, int cols,
                                                   long rows_src, int ns, int8_t* __restrict__ planes,
                                                   long ldo, long rows_dst, int* __restrict__ expo) {
  const long row = blockIdx.x;
  const int lane = threadIdx.x;
  const bool live = row < rows_src;
  const double* s = src + row * ld;
  double mx = 0.0;
  if (live)
    for (int c = lane; c < cols; c += 64) mx = fmax(mx, fabs(s[c]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const int e = row_exponent(mx);
  if (lane == 0) expo[row] = e;
  const long plane = rows_dst * ldo;
  for (int c = lane; c < ldo; c += 64) {
    double r = (live && c < cols) ? ldexp(s[c], 6 - e) : 0.0;  // v * 2^6, |r| <= 64
    for (int k = 0; k < ns; ++k) {
      const double d = rint(r);
      planes[k * plane + row * ldo + c] = (int8_t)d;
      r = (r - d) * 128.0;
    }
  }
}

// X[Q][m][col] = 2^(e_Qm + f_col) sum_{k+l<=3} 2^(-12-7(k+l)) <Bd_k[Q][m][:], Cd_l[col][:]>
// gpart[Q][mtile] = sum over the tile rows of sum_n B[Q][m][n] D[m][n] with B
// rebuilt exactly from its five digit planes.
// Tile: 64 rows (16 per wave) x 16*NT columns, k-steps of 64.
template <int NT>
__global__ __launch_bounds__(256) void k_half_i8(const int8_t* __restrict__ Bd, long bplane, const int* __restrict__ Be,
                                                 long ld8, int nbf, const int8_t* __restrict__ Cd, long cplane,
                                                 const int* __restrict__ Ce, const double* __restrict__ Dp, long ldd,
                                                 float* __restrict__ X, long ldx, double* __restrict__ gpart,
                                                 int mtiles) {
  constexpr int BMI = 64, BN = 16 * NT, KS = 64, PAD = 16;
  __shared__ __attribute__((aligned(16))) int8_t As[4][BMI][KS + PAD];
  __shared__ __attribute__((aligned(16))) int8_t Cs[4][BN][KS + PAD];
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int mtile = blockIdx.x;
  const long Q = blockIdx.y;
  const int cg = blockIdx.z;
  const int m0 = mtile * BMI;
  // loader role: row ar of the tile, 16-wide k segment aq
  const int ar = t >> 2, aq = (t & 3) * 16;
  const bool arow_ok = m0 + ar < nbf;
  const long arow = Q * nbf + m0 + ar;
  const int8_t* Ap = Bd + arow * ld8 + aq;
  const double* Drow = Dp + (long)(m0 + ar) * ldd + aq;
  v4i_t acc[4][NT];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[c][j] = v4i_t{0, 0, 0, 0};
  double rd_hi = 0.0, rd_lo = 0.0;
  uint4 a_reg[5];
  double d_reg[16];
  constexpr int CCH = BN * 4 * (KS / 16);  // 16-byte chunks of the C tile
  uint4 c_reg[(CCH + 255) / 256];
  auto gload = [&](int k0) {
    if (arow_ok) {
#pragma unroll
      for (int p = 0; p < 5; ++p) a_reg[p] = *reinterpret_cast<const uint4*>(Ap + p * bplane + k0);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const double2 x = *reinterpret_cast<const double2*>(Drow + k0 + 2 * i);
        d_reg[2 * i] = x.x;
        d_reg[2 * i + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int p = 0; p < 5; ++p) a_reg[p] = uint4{0, 0, 0, 0};
#pragma unroll
      for (int i = 0; i < 16; ++i) d_reg[i] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < (CCH + 255) / 256; ++u) {
      const int e = t + 256 * u;
      if (e < CCH) {
        const int col = e / 16, p = (e / 4) % 4, seg = (e % 4) * 16;
        c_reg[u] = *reinterpret_cast<const uint4*>(Cd + p * cplane + ((long)cg * BN + col) * ld8 + k0 + seg);
      }
    }
  };
  gload(0);
  for (int k0 = 0; k0 < ld8; k0 += KS) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) *reinterpret_cast<uint4*>(&As[p][ar][aq]) = a_reg[p];
#pragma unroll
    for (int u = 0; u < (CCH + 255) / 256; ++u) {
      const int e = t + 256 * u;
      if (e < CCH) {
        const int col = e / 16, p = (e / 4) % 4, seg = (e % 4) * 16;
        *reinterpret_cast<uint4*>(&Cs[p][col][seg]) = c_reg[u];
      }
    }
    // FP64 gamma partials from the exact five-digit value of B:
    // v 2^6 = d0 + 2^-7 d1 + ... = 2^-7 hi + 2^-28 lo with
    // hi = 128 d0 + d1 and lo = 2^14 d2 + 2^7 d3 + d4 (both exact integers)
    {
      const int8_t* b0 = reinterpret_cast<const int8_t*>(&a_reg[0]);
      const int8_t* b1 = reinterpret_cast<const int8_t*>(&a_reg[1]);
      const int8_t* b2 = reinterpret_cast<const int8_t*>(&a_reg[2]);
      const int8_t* b3 = reinterpret_cast<const int8_t*>(&a_reg[3]);
      const int8_t* b4 = reinterpret_cast<const int8_t*>(&a_reg[4]);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int hi = 128 * (int)b0[i] + (int)b1[i];
        const int lo = 16384 * (int)b2[i] + 128 * (int)b3[i] + (int)b4[i];
        rd_hi += (double)hi * d_reg[i];
        rd_lo += (double)lo * d_reg[i];
      }
    }
    __syncthreads();
    if (k0 + KS < ld8) gload(k0 + KS);
    uint4 av[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
      av[p] = *reinterpret_cast<const uint4*>(&As[p][w * 16 + lane % 16][16 * (lane / 16)]);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      uint4 bv[4];
#pragma unroll
      for (int p = 0; p < 4; ++p)
        bv[p] = *reinterpret_cast<const uint4*>(&Cs[p][j * 16 + lane % 16][16 * (lane / 16)]);
      acc[0][j] = mfma_i8(av[0], bv[0], acc[0][j]);
      acc[1][j] = mfma_i8(av[0], bv[1], acc[1][j]);
      acc[1][j] = mfma_i8(av[1], bv[0], acc[1][j]);
      acc[2][j] = mfma_i8(av[0], bv[2], acc[2][j]);
      acc[2][j] = mfma_i8(av[1], bv[1], acc[2][j]);
      acc[2][j] = mfma_i8(av[2], bv[0], acc[2][j]);
      acc[3][j] = mfma_i8(av[0], bv[3], acc[3][j]);
      acc[3][j] = mfma_i8(av[1], bv[2], acc[3][j]);
      acc[3][j] = mfma_i8(av[2], bv[1], acc[3][j]);
      acc[3][j] = mfma_i8(av[3], bv[0], acc[3][j]);
    }
  }
  // epilogue: X in FP64 from the class sums, stored as FP32
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int m = m0 + w * 16 + 4 * (lane / 16) + v;
    if (m < nbf) {
      const int em = Be[Q * nbf + m];
      float* xr = X + ((long)Q * nbf + m) * ldx + (long)cg * BN + lane % 16;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int col = cg * BN + j * 16 + lane % 16;
        const double s = (double)acc[0][j][v] + 0x1p-7 * (double)acc[1][j][v] + 0x1p-14 * (double)acc[2][j][v] +
                         0x1p-21 * (double)acc[3][j][v];
        xr[j * 16] = (float)ldexp(s, em + Ce[col] - 12);
      }
    }
  }
  if (cg == 0) {
    const double part = arow_ok ? ldexp(rd_hi, Be[arow] - 13) + ldexp(rd_lo, Be[arow] - 34) : 0.0;
    const double s = block_sum256(part, red);
    if (t == 0) gpart[Q * mtiles + mtile] = s;
  }
}

}  // namespace

hipError_t split_rows_i8(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                         long rows_dst, int* expo, hipStream_t s) {
  if (rows_dst <= 0) return hipSuccess;
  k_split_rows<<<(unsigned)rows_dst, 64, 0, s>>>(src, ld, cols, rows_src, ns, planes, ldo, rows_dst, expo);
  return hipGetLastError();
}

hipError_t launch_half_i8(int nt, dim3 grid, hipStream_t s, const int8_t* Bd, long bplane, const int* Be, long ld8,
                          int nbf, const int8_t* Cd, long cplane, const int* Ce, const double* Dp, long ldd,
                          float* X, long ldx, double* gpart, int mtiles) {
#define NOPT_HI8(NTV)                                                                                     \
  case NTV:                                                                                               \
    k_half_i8<NTV><<<grid, 256, 0, s>>>(Bd, bplane, Be, ld8, nbf, Cd, cplane, Ce, Dp, ldd, X, ldx, gpart, \
                                        mtiles);                                                          \
    break;
  switch (nt) {
    NOPT_HI8(1)
    NOPT_HI8(2)
    NOPT_HI8(4)
    NOPT_HI8(6)
    NOPT_HI8(8)
    default:
      return hipErrorInvalidValue;
  }
#undef NOPT_HI8
  return hipGetLastError();
}

}  // namespace nopt
