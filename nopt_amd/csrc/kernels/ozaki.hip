// Integer-split (Ozaki-type) kernels on the int8 matrix cores.
//
// Every row r of an operand is scaled by a power of two, 2^-e_r with
// e_r = exponent of the row's largest magnitude, and the scaled row
// v = x 2^-e_r (|v| < 1) is written as signed base-128 digits
//   x 2^(6-e_r) = d_0 + 2^-7 d_1 + 2^-14 d_2 + ... ,  d_k = rint(...) in [-64, 64],
// each digit rounded to nearest, so the truncation error is unbiased and at
// most 2^(e_r - 7S), i.e. 2^(1-7S) of the row maximum for S digits (2^-27
// for 4 digits, 2^-34 for 5). A product of two split
// operands is the weighted sum of digit-plane products; every digit-plane
// product is an exact int32 dot product on v_mfma_i32_16x16x64_i8, so the
// accumulation adds no rounding at all. Digit pairs of equal weight (k+l)
// share one int32 accumulator ("class").
//
//   half-transform  X = B C^T     4 digits each, pairs k+l <= 4 (13 pairs)
//   exchange        K = sum_Q X_Q X_Q^T   4 digits, pairs k+l <= 4 with
//                   k,l <= 3 (13 pairs); the k+l = 4 terms keep the
//                   same-number cross terms on the diagonal, which do not
//                   average out
//   Coulomb gamma   gamma_Q = sum_mn B_Qmn D_mn from 5 digits of B (35 bits)
// Class sums are combined in FP64 and scaled by the row exponents.
#include "ozaki.hpp"

#include <cstring>
#include <type_traits>
#include <utility>

namespace nopt {
namespace {

typedef int v4i_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4i_t mfma_i8(const uint4& a, const uint4& b, v4i_t c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(v4i_t, a), __builtin_bit_cast(v4i_t, b), c,
                                               0, 0, 0);
}

// f(std::integral_constant<int, i>) for i = 0..N-1, unrolled at compile time
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// row maximum -> exponent e with |x| 2^-e < 1 (0 for an all-zero row); at
// least kMinExp, so that every exponent fits a byte (a row below 2^-121 keeps
// its digits at the scale 2^-120: absolute resolution 2^-147)
constexpr int kMinExp = -120;
__device__ __forceinline__ int row_exponent(double mx) {
  if (!(mx > 0.0)) return 0;
  int e;
  (void)frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1)
  return max(e, kMinExp);
}

// The five balanced base-128 digits of x 2^(6-e) as k_split_rows forms them,
// returned as N28 = sum_{k<4} d_k 2^(7(3-k)) (the first four digits, |N28| <=
// 2^27) and d4: N35 = 128 N28 + d4 = rint(x 2^(34-e)).
__device__ __forceinline__ void digits5(double x, int e, int& n28, int& d4) {
  double r = ldexp(x, 6 - e);
  int n = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double d = rint(r);
    n = n * 128 + (int)d;
    r = (r - d) * 128.0;
  }
  n28 = n;
  d4 = (int)rint(r);
}

// first 16-column segment of row m in the packed lower triangle Bs (row r
// holds r/16 + 1 segments; the layout is described with k_pack_segs)
__host__ __device__ __forceinline__ long seg0(int m) {
  const long a = m / 16, b = m % 16;
  return m + 8 * a * (a - 1) + a * b;
}

// Splits the rows of a row-major FP64 matrix (rows_src x cols, ld) into ns
// int8 digit planes planes[k*pstride + row*ldo + col] (rows_dst rows, zero
// padded) and row exponents. One wavefront per row. Optional reorders:
// output row r reads source row (r/rper)*rper + rperm[r%rper], output column
// c source column cperm[c].
__device__ __forceinline__ void k_split_rows_body(const VB vb, const double* __restrict__ src, long ld, int cols,
                                                   long rows_src, int ns, int8_t* __restrict__ planes,
                                                   long ldo, long pstride, int* __restrict__ expo,
                                                   const int* __restrict__ rperm, long rper,
                                                   const int* __restrict__ cperm) {
  const long row = vb.x;
  const int lane = threadIdx.x;
  const bool live = row < rows_src;
  const double* s = src + (rperm ? row / rper * rper + rperm[row % rper] : row) * ld;
  double mx = 0.0;
  if (live)
    for (int c = lane; c < cols; c += 64) mx = fmax(mx, fabs(s[c]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const int e = row_exponent(mx);
  if (lane == 0) expo[row] = e;
  for (int c = lane; c < ldo; c += 64) {
    double r = (live && c < cols) ? ldexp(s[cperm ? cperm[c] : c], 6 - e) : 0.0;  // |r| < 64
    for (int k = 0; k < ns; ++k) {
      const double d = rint(r);
      planes[k * pstride + row * ldo + c] = (int8_t)d;
      r = (r - d) * 128.0;
    }
  }
}
__global__ __launch_bounds__(64) void k_split_rows(const double* __restrict__ src, long ld, int cols,
                                                   long rows_src, int ns, int8_t* __restrict__ planes,
                                                   long ldo, long pstride, int* __restrict__ expo,
                                                   const int* __restrict__ rperm, long rper,
                                                   const int* __restrict__ cperm) {
  k_split_rows_body(grid_vb(), src, ld, cols, rows_src, ns, planes, ldo, pstride, expo, rperm, rper, cperm);
}
__global__ __launch_bounds__(64) void k_split_rows_list(const ItemSplit* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemSplit& a = items[i];
  k_split_rows_body(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.src, a.ld, a.cols, a.rows_src, a.ns, a.planes, a.ldo, a.pstride, a.expo, nullptr, 0, a.cperm);
}


// ------------------------------------------- rows of packed symmetric chunks
// A chunk of a symmetric tensor in packed lower-triangle form,
// A[P*cp + m(m+1)/2 + n - m0(m0+1)/2] for the rows m0 <= m < m1 (n <= m), holds
// each element of the full rows (P, m) and (P, n). A block takes a 64 x 64
// tile (rows mb.., columns nb..) of one P through LDS, so that the element
// (m, n) and its mirror (n, m) are both written along contiguous columns.
// Modes: 0 writes FP64 rows (out64, ld), 1 folds |x| into the row maxima
// (bit patterns of non-negative doubles order like their values), 2 writes
// ns int8 digits with the rows' exponents (the digits k_split_rows forms),
// 3 writes the packed lower triangle Bs (planes = Bs, pstride = splane,
// ld = sq) of the reordered functions: the element (m, n) goes to row
// max(a, b), column min(a, b) with a, b = pinv[m], pinv[n], as digits5 of the
// value at that row's exponent. With pinv, modes 1 and 2 write the row and
// column pinv[m], pinv[n] as well (the function m of B is function pinv[m]
// of the kept forms).
template <int MODE>
__global__ __launch_bounds__(256) void k_packed_tiles(const double* __restrict__ A, long cp, int m0, int m1, int nbf,
                                                      double* __restrict__ out64, long ld,
                                                      unsigned long long* __restrict__ rowmax,
                                                      const int* __restrict__ expo, int ns,
                                                      int8_t* __restrict__ planes, long pstride,
                                                      const int* __restrict__ pinv) {
  auto fn = [&](int i) { return pinv ? pinv[i] : i; };
  __shared__ double tile[64][65];
  const int nb = blockIdx.x * 64, mb = m0 + blockIdx.y * 64;
  if (nb > mb + 63) return;  // above the diagonal: nothing stored (uniform per block)
  const long P = blockIdx.z;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long p0 = (long)m0 * (m0 + 1) / 2;
  const double* Ap = A + P * cp;
  double v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r = w + 4 * k, m = mb + r, n = nb + c;
    v[k] = (m < m1 && n <= m) ? Ap[(long)m * (m + 1) / 2 - p0 + n] : 0.0;
    tile[r][c] = v[k];
  }
  // the element itself: row m, column n
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int m = mb + w + 4 * k, n = nb + c;
    const bool ok = m < m1 && n <= m;
    if constexpr (MODE == 0) {
      if (ok) out64[(P * nbf + m) * ld + n] = v[k];
    } else if constexpr (MODE == 1) {
      double mx = fabs(v[k]);
      for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
      if (c == 0 && m < m1) atomicMax(rowmax + P * nbf + fn(m), (unsigned long long)__double_as_longlong(mx));
    } else if constexpr (MODE == 2) {
      if (ok) {
        const long row = P * nbf + fn(m);
        double r = ldexp(v[k], 6 - expo[row]);
        for (int d = 0; d < ns; ++d) {
          const double dg = rint(r);
          planes[d * pstride + row * ld + fn(n)] = (int8_t)dg;
          r = (r - dg) * 128.0;
        }
      }
    } else {
      if (ok) {
        const int a = fn(m), b = fn(n), r = max(a, b), cc = min(a, b);
        int n28, d4;
        digits5(v[k], expo[P * nbf + r], n28, d4);
        const long el = P * ld + 16 * seg0(r) + cc;
        reinterpret_cast<int*>(planes)[el] = n28;
        planes[4 * pstride + el] = (int8_t)d4;
      }
    }
  }
  if constexpr (MODE == 3) return;  // each stored element is one element of the triangle
  __syncthreads();
  // the mirror: row n = nb + rr, column m = mb + c (strictly below the diagonal)
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int rr = w + 4 * k, n = nb + rr, m = mb + c;
    const bool ok = m < m1 && n < m;
    const double x = ok ? tile[c][rr] : 0.0;
    if constexpr (MODE == 0) {
      if (ok) out64[(P * nbf + n) * ld + m] = x;
    } else if constexpr (MODE == 1) {
      double mx = fabs(x);
      for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
      if (c == 0 && n < nbf) atomicMax(rowmax + P * nbf + fn(n), (unsigned long long)__double_as_longlong(mx));
    } else if constexpr (MODE == 2) {
      if (ok) {
        const long row = P * nbf + fn(n);
        double r = ldexp(x, 6 - expo[row]);
        for (int d = 0; d < ns; ++d) {
          const double dg = rint(r);
          planes[d * pstride + row * ld + fn(m)] = (int8_t)dg;
          r = (r - dg) * 128.0;
        }
      }
    }
  }
}

__global__ void k_rowmax_exponent(const unsigned long long* __restrict__ rowmax, long rows, int* __restrict__ expo) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) expo[r] = row_exponent(__longlong_as_double((long long)rowmax[r]));
}

// ------------------------------------------------ packed lower triangle of B
// B's lower triangle as packed rows cut into segments of 16 columns: segment
// s = (row m, columns n0 .. n0+15), n0 a multiple of 16, with the digits of
// row m (exponent Be[Q*nbf + m]) and zeros beyond column m. Row m holds
// m/16 + 1 segments, starting at segment seg0(m) = sum_{r<m} (r/16 + 1), and
// the segments of all rows follow each other, so one B_Q is a contiguous run
// of sq = 16*nseg values and the element (Q, m, n <= m) sits at
// e = Q*sq + 16*seg0(m) + n. seg_code[s] = (m << 16) | n0. A value is kept as
// its 35-bit integer N35 = rint(B 2^(34-e)) = 128 N28 + d4 (digits5): the
// int32 N28, the value to 28 bits, at ((int*)Bs)[e] (the first 4*splane
// bytes), and the signed byte d4 at Bs[4*splane + e]. The Coulomb kernels
// rebuild N35 exactly as fma(N28, 128, d4); the half-transform reads N28
// alone.

// Bs for the segments [s0, s1) (the rows of one chunk) with the row exponents
// Be: SRC 0 from FP64 rows B[(Q*nbf + m)*ld + n] (with perm, the row and
// column m, n of Bs read the row and column perm[m], perm[n] of B), SRC 1
// from a packed chunk A[Q*cp + m(m+1)/2 - m0(m0+1)/2 + n] of the rows
// m0 <= m < m1. Grid: (s1 - s0)/256 x naux, one segment per thread.
template <int SRC>
__global__ __launch_bounds__(256) void k_pack_segs(const double* __restrict__ A, long ld, int m0, int nbf,
                                                   const int* __restrict__ Be, const int* __restrict__ seg_code,
                                                   int s0, int s1, long sq, int8_t* __restrict__ Bs, long splane,
                                                   const int* __restrict__ perm) {
  const int s = s0 + blockIdx.x * 256 + threadIdx.x;
  if (s >= s1) return;
  const long Q = blockIdx.y;
  const int code = seg_code[s], m = code >> 16, n0 = code & 0xffff;
  const int e = Be[Q * nbf + m];
  const double* src = SRC == 0 ? A + (Q * nbf + (perm ? perm[m] : m)) * ld
                               : A + Q * ld + (long)m * (m + 1) / 2 - (long)m0 * (m0 + 1) / 2;
  int part[16];
  uint32_t low[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int n28 = 0, d4 = 0;
    const int n = n0 + j;
    if (n <= m) digits5(src[SRC == 0 && perm ? perm[n] : n], e, n28, d4);
    part[j] = n28;
    low[j / 4] |= (uint32_t)(d4 & 255) << (8 * (j % 4));
  }
  const long el = Q * sq + 16L * s;  // element of the first column
  uint4* p32 = reinterpret_cast<uint4*>(reinterpret_cast<int*>(Bs) + el);
#pragma unroll
  for (int c = 0; c < 4; ++c)
    p32[c] = uint4{(uint32_t)part[4 * c], (uint32_t)part[4 * c + 1], (uint32_t)part[4 * c + 2], (uint32_t)part[4 * c + 3]};
  *reinterpret_cast<uint4*>(Bs + 4 * splane + el) = uint4{low[0], low[1], low[2], low[3]};
}

// the row exponents as bytes, Be8[Q*ld8 + n] (zero padding to ld8)
__global__ void k_expo8(const int* __restrict__ Be, int nbf, long ld8, long naux, int8_t* __restrict__ Be8) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= naux * ld8) return;
  const long Q = i / ld8, n = i % ld8;
  Be8[i] = n < nbf ? (int8_t)Be[Q * nbf + n] : (int8_t)0;
}

// ------------------------------------------------------ blocks of exact zeros
// Where B's integer forms hold nothing but zeros. An element whose value lies
// below the resolution of its row (2^-28 of the row maximum in the digit
// planes, 2^-35 in Bs) is stored as an exact zero, and for a molecule of some
// extent most pairs (m, n) of distant functions are such zeros in every B_Q.
// The kernels leave these blocks out: adding zero products changes no bit of
// an exact integer sum, and a J element that is zero in every B_Q is zero.

// live[s] |= bit 0 (bit 1) when the first (last) eight columns of Coulomb
// segment s hold a nonzero N28 or d4 for some function in the range of
// blockIdx.y. Grid: nseg/256 x function ranges; live zeroed before.
__global__ __launch_bounds__(256) void k_seg_live(const int8_t* __restrict__ Bs, long splane, long sq, int nseg,
                                                  int naux, int* __restrict__ live) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nseg) return;
  const int q0 = (int)((long)naux * blockIdx.y / gridDim.y), q1 = (int)((long)naux * (blockIdx.y + 1) / gridDim.y);
  uint32_t bits = 0;
  for (int q = q0; q < q1 && bits != 3u; ++q) {
    const long e = (long)q * sq + 16L * s;
    const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const int*>(Bs) + e);
    const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
    const uint4 l = *reinterpret_cast<const uint4*>(Bs + 4 * splane + e);
    if ((a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w | l.x | l.y) != 0u) bits |= 1u;
    if ((c.x | c.y | c.z | c.w | d.x | d.y | d.z | d.w | l.z | l.w) != 0u) bits |= 2u;
  }
  if (bits) atomicOr(live + s, (int)bits);
}

// Bm[Q*nmb + mb] bit k: the rows 32 mb .. 32 mb + 31 of function Q hold a
// nonzero digit in the columns 64 k .. 64 k + 63 of the full-row planes Bd
// (4 planes of stride bplane, rows of ld8 bytes). Grid: nmb x naux.
__global__ __launch_bounds__(256) void k_slab_mask_rows(const int8_t* __restrict__ Bd, long bplane, long ld8, int nbf,
                                                        int nmb, uint32_t* __restrict__ Bm) {
  __shared__ uint32_t red;
  const int mb = blockIdx.x;
  const long Q = blockIdx.y;
  if (threadIdx.x == 0) red = 0u;
  __syncthreads();
  const int rows = min(32, nbf - 32 * mb), cpr = (int)(ld8 / 16), per = rows * cpr;
  uint32_t bits = 0;
  for (int i = threadIdx.x; i < 4 * per; i += 256) {
    const int p = i / per, r = (i % per) / cpr, c = i % cpr;
    const uint4 v = *reinterpret_cast<const uint4*>(Bd + p * bplane + (Q * nbf + 32 * mb + r) * ld8 + 16 * c);
    if ((v.x | v.y | v.z | v.w) != 0u) bits |= 1u << (c / 4);
  }
  if (bits) atomicOr(&red, bits);
  __syncthreads();
  if (threadIdx.x == 0) Bm[Q * nmb + mb] = red;
}

// The same masks for the packed loader, from the N28 of Bs (int32 plane Bp,
// sq values per function): (m, n <= m) from row m, (m, n > m) from the stored
// (n, m), which the loader rescales (a zero stays zero, a nonzero value counts
// as live). The second loop walks the stored rows n so that consecutive
// threads read consecutive columns m. Grid: nmb x naux.
__global__ __launch_bounds__(256) void k_slab_mask_packed(const int* __restrict__ Bp, long sq, int nbf, int nmb,
                                                          uint32_t* __restrict__ Bm) {
  __shared__ uint32_t red;
  const int mb = blockIdx.x, mlo = 32 * mb;
  const long Q = blockIdx.y;
  if (threadIdx.x == 0) red = 0u;
  __syncthreads();
  const int rows = min(32, nbf - mlo);
  const int* B = Bp + Q * sq;
  uint32_t bits = 0;
  for (int i = threadIdx.x; i < rows * nbf; i += 256) {
    const int m = mlo + i / nbf, n = i % nbf;
    if (n <= m && B[16 * seg0(m) + n] != 0) bits |= 1u << (n / 64);
  }
  for (int i = threadIdx.x; i < nbf * 32; i += 256) {
    const int n = i / 32, m = mlo + i % 32;
    if (m < mlo + rows && n > m && B[16 * seg0(n) + m] != 0) bits |= 1u << (n / 64);
  }
  if (bits) atomicOr(&red, bits);
  __syncthreads();
  if (threadIdx.x == 0) Bm[Q * nmb + mb] = red;
}

// Maximum over each row of 16 lanes, in every lane of the row: rotations by
// 8, 4, 2 and 1 within the row (DPP row_ror, register moves).
template <int N>
__device__ __forceinline__ int ror16(int x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x120 + N, 0xf, 0xf, false);
}
__device__ __forceinline__ int row_max16(int x) {
  x = max(x, ror16<8>(x));
  x = max(x, ror16<4>(x));
  x = max(x, ror16<2>(x));
  return max(x, ror16<1>(x));
}

// ------------------------------------------------------------ half-transform
// LDS position of 16-byte chunk c (of a 64-byte k-slab) in row r: an XOR
// swizzle under which the operand reads (ds_read_b128, lane l: row l % 16 of
// a 16-row block, chunk l / 16) hit no bank conflict on gfx950. Measured with
// SQ_LDS_BANK_CONFLICT: c ^ ((r >> 1) & 3), and 16 rows read 4 lanes per
// row in order, are free of conflicts; c ^ (r & 3), c ^ ((r >> 2) & 3),
// c ^ (((r >> 1) + (r >> 3)) & 3) and no swizzle at all cost 4 cycles per
// read. The loaders write whole 64-byte rows per 4 (or 16) lanes, which no
// swizzle changes.
__device__ __forceinline__ int swz(int r, int c) { return c ^ ((r >> 1) & 3); }

// X[Q][m][col] = 2^(e_Qm + f_col - 12) sum_{k+l<=4} 2^(-7(k+l)) <Bd_k[Q][m][:], Cd_l[col][:]>
// for k, l < 4 (13 digit pairs; the products of equal k+l share one exact
// int32 accumulator), combined in FP64. Tile: 64 rows x 16*NT columns; the NW
// waves form a (NW/WC) x WC grid, each with IT = 4/(NW/WC) row tiles and
// JT = NT/WC column tiles of 16, so that a wave reads IT + JT operand tiles
// from LDS per 13*IT*JT products. k-slabs of 64 bytes, LDS double buffered
// (one barrier per slab) with the next B slab prefetched into registers while
// the current one is multiplied. NW = 8 halves the accumulators per wave so
// that twice as many waves hide the memory latency. QPB = 2 gives a block two
// auxiliary functions, Q = 2 blockIdx.y and Q + 1 (NW waves each, one shared C
// slab); with the fused split their rows share one exponent per m and are
// written side by side, so that the exchange reads 2 x len orbitals per
// exponent (Q + 1 = naux contributes zeros). BMI rows per tile (64, or 32 for
// two functions, so that two blocks still share a CU); WPE waves per SIMD.
// PK reads B from the packed lower triangle instead (Bp = the N28 of Bs, sq
// values per function, Be8 the row exponents as bytes, row stride ld8): an
// element (m, n) with n <= m is N28 of row m as stored; one with n > m is
// the stored (n, m), N28 at the scale of row n, brought to the scale of row
// m by a power of two (rounded to nearest on a right shift). Each thread
// turns four consecutive columns of a row into one dword per digit plane. In
// a slab left of the tile's diagonal block the lanes walk rows of the stored
// triangle (16-byte loads); right of it and on it the lanes walk the tile's
// rows, and each loads one column of four stored rows (the lanes of a
// function read contiguous dwords).
template <int NT, int WC, int NW, int QPB, int BMI, int WPE, bool PK>
__device__ __forceinline__ void k_half_i8_body(const VB vb, 
    const int8_t* __restrict__ Bd, long bplane, const int* __restrict__ Bp, long sq, const int8_t* __restrict__ Be8,
    const int* __restrict__ Be, const uint32_t* __restrict__ Bm, int nmb, long ld8, int nbf, int naux,
    const int8_t* __restrict__ Cd, long cplane, const int* __restrict__ Ce, double* __restrict__ X, long ldx,
    const XSplit& xs) {
  constexpr int BN = 16 * NT, KS = 64, NWT = NW * QPB, NTH = 64 * NWT;
  constexpr int WR = NW / WC, IT = BMI / (16 * WR), JT = NT / WC;
  static_assert(NT % WC == 0 && NW % WC == 0 && IT * WR * 16 == BMI, "tile split");
  constexpr int AB = 4 * BMI * KS, CB = 4 * BN * KS;  // bytes per B (one Q) and C buffer
  constexpr int SB = QPB * AB + CB;                   // bytes per stage
  constexpr int ACQ = 16 * BMI;                       // 16-byte B chunks per Q (4 planes x BMI rows x 4)
  constexpr int APT = ACQ * QPB / NTH;                // per thread
  constexpr int CCH = 4 * BN * 4;                     // 16-byte C chunks per slab
  constexpr int CPT = (CCH + NTH - 1) / NTH;
  static_assert(ACQ * QPB % NTH == 0 && NTH % (4 * BMI) == 0, "B slab must split evenly over the block");
  __shared__ __attribute__((aligned(16))) int8_t sm[2][SB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int qq = w / NW, w8 = w % NW;  // which Q of the block, wave within it
  const int wr = w8 / WC, wc = w8 % WC;
  const int m0 = vb.x * BMI;
  const long Qb = (long)vb.y * QPB;  // first Q of the block
  const long Q = Qb + qq;                  // this wave's Q
  const int cg = vb.z;
  // B loader: chunk e = t + NTH u is chunk e%4 of row (e/4)%BMI in plane
  // (e%ACQ)/(4 BMI) of the block's Q number e/ACQ (rows beyond nbf and a Q
  // beyond naux load zeros); the row and chunk do not depend on u
  const int ar = (t >> 2) % BMI, ac = t & 3;
  const int8_t* ap[APT];
  bool aok[APT];
  if constexpr (!PK) {
#pragma unroll
    for (int u = 0; u < APT; ++u) {
      const int e = t + NTH * u, qe = e / ACQ, pe = (e % ACQ) / (4 * BMI);
      aok[u] = m0 + ar < nbf && Qb + qe < naux;
      ap[u] = Bd + ((Qb + qe) * nbf + m0 + ar) * ld8 + 16 * ac + pe * bplane;
    }
  }
  // packed loader: GPT groups of four columns per thread and slab. Row-walk
  // (slabs left of the diagonal block): group e = t + NTH u is columns
  // 4 (e % 16) .. +3 of tile row (e / 16) % BMI of function e / (16 BMI).
  // Column-walk: lane l takes tile row l % BMI of function l / BMI, and the
  // wave's group index is w + NWT u (columns 4 (w + NWT u) .. +3).
  constexpr int GPT = QPB * BMI * 16 / NTH;
  static_assert(!PK || (GPT * NTH == QPB * BMI * 16 && QPB * BMI == 64 && NWT * GPT == 16), "packed loader shape");
  const int tq = lane / BMI, tr = lane % BMI;  // column-walk: function and tile row of the lane
  const int wu = __builtin_amdgcn_readfirstlane(w);
  // buffer resources from the block's first function (int32 byte offsets;
  // the ranges end with the last function, so loads beyond read zeros): the
  // N28 values, and the exponent bytes
  const long nfq = PK ? min((long)QPB, naux - Qb) : 0;  // functions of this block
  const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<int*>(Bp + (PK ? Qb * sq : 0)), (short)0, (int)(4 * sq * nfq), 0x00020000);
  const __amdgpu_buffer_rsrc_t rE = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<int8_t*>(Be8 + (PK ? Qb * ld8 : 0)), (short)0, (int)(ld8 * nfq), 0x00020000);
  const int tqs = tq * (int)sq, tqe = tq * (int)ld8;
  const bool tok = PK && m0 + tr < nbf && Qb + tq < naux;
  const int em = tok ? (int)(int8_t)__builtin_amdgcn_raw_buffer_load_b8(rE, tqe + m0 + tr, 0, 0) : 0;
  int doff[PK ? GPT : 1];
  bool dok[PK ? GPT : 1];
  if constexpr (PK) {
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
      const int e = t + NTH * u, q = e / (16 * BMI), r = (e / 16) % BMI, g = e % 16;
      dok[u] = m0 + r < nbf && Qb + q < naux;
      doff[u] = q * (int)sq + 16 * (int)seg0(dok[u] ? m0 + r : 0) + 4 * g;
    }
  }
  uint4 praw[PK ? GPT : 1];   // N28 of the group's four columns
  uint32_t pexp[PK ? GPT : 1];  // column-walk: exponent bytes of the four stored rows
  int pn0 = 0, ptype = 0;       // slab of the prefetched values and its walk (0 rows, 1 columns, 2 diagonal block)
  const int8_t* Cbase = Cd + (long)cg * BN * ld8;
  // C loader: chunk e = t + NTH u < CCH of the slab is chunk (e & 3) of column
  // (e >> 2) % BN in plane e / (4 BN); global and LDS offsets per u
  int c_g[CPT], c_s[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int e = min(t + NTH * u, CCH - 1);
    const int c = e & 3, col = (e >> 2) % BN, p = e / (4 * BN);
    c_g[u] = (int)(p * cplane + (long)col * ld8 + 16 * c);
    c_s[u] = QPB * AB + (p * BN + col) * KS + 16 * swz(col, c);
  }
  constexpr bool c_even = CCH % NTH == 0;
  const bool c_last = c_even || t + NTH * (CPT - 1) < CCH;  // this thread has a chunk in the last round
  // prefetch registers; the C chunks sit in two arrays of at most 16 dwords
  // indexed with compile-time constants, so that both stay in registers
  constexpr int CH = (CPT + 1) / 2;
  uint4 a_reg[APT], c_lo[CH], c_hi[CPT - CH > 0 ? CPT - CH : 1];

  // packed slab at columns n0..n0+63: raw N28 (and exponent bytes) into registers
  auto gload_packed = [&](int n0) __attribute__((always_inline)) {
    pn0 = n0;
    ptype = n0 + KS - 1 <= m0 ? 0 : n0 >= m0 + BMI - 1 ? 1 : 2;
    if (ptype == 0) {
#pragma unroll
      for (int u = 0; u < GPT; ++u)
        if (dok[u]) {
          const auto x = __builtin_amdgcn_raw_buffer_load_b128(rB, 4 * (doff[u] + n0), 0, 0);
          praw[u] = uint4{x[0], x[1], x[2], x[3]};
        } else {
          praw[u] = uint4{0, 0, 0, 0};
        }
      return;
    }
    const int m = m0 + tr;
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
      const int n4 = n0 + 4 * (wu + NWT * u);  // wave-uniform
      pexp[u] = tok && n4 < nbf ? __builtin_amdgcn_raw_buffer_load_b32(rE, tqe, n4, 0) : 0u;
      uint32_t v[4];
      if (ptype == 1) {
        // every stored row n >= m: the row offset is wave-uniform (scalar)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n4 + j;
          const int so = __builtin_amdgcn_readfirstlane(64 * (int)seg0(min(n, nbf - 1)));
          v[j] = tok && n < nbf ? __builtin_amdgcn_raw_buffer_load_b32(rB, 4 * (tqs + m), so, 0) : 0u;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n4 + j;
          const int r = max(m, n), c = min(m, n);
          v[j] = tok && n < nbf ? __builtin_amdgcn_raw_buffer_load_b32(rB, 4 * (tqs + 16 * (int)seg0(r) + c), 0, 0)
                                : 0u;
        }
      }
      praw[u] = uint4{v[0], v[1], v[2], v[3]};
    }
  };
  // the prefetched slab as four digit planes in LDS (one dword per plane and group)
  auto lstore_packed = [&](int8_t* As) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
      int q, r, g;
      int v[4] = {(int)praw[u].x, (int)praw[u].y, (int)praw[u].z, (int)praw[u].w};
      if (ptype == 0) {
        const int e = t + NTH * u;
        q = e / (16 * BMI), r = (e / 16) % BMI, g = e % 16;
      } else {
        q = tq, r = tr, g = wu + NWT * u;
        const int m = m0 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = pn0 + 4 * g + j;
          // stored at row n (scale e_n) when n > m: times 2^(e_n - e_m)
          // (right of the diagonal block n >= m, and e_n = e_m for n = m)
          const int d = (ptype == 1 || n > m) ? (int)(int8_t)(pexp[u] >> (8 * j)) - em : 0;
          const int k = min(max(-d, 0), 31);
          v[j] = d > 0 ? (int)((uint32_t)v[j] << d) : (v[j] + (int)((1u << k) >> 1)) >> k;
        }
      }
      // balanced digits of each value (bytes d0..d3 of one dword), then a
      // 4 x 4 byte transpose into one dword per plane
      uint32_t f[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int M = v[j] + 1056832;
        const uint32_t uu = (((uint32_t)M >> 14 & 127u) << 8) | (((uint32_t)M >> 7 & 127u) << 16) |
                            (((uint32_t)M & 127u) << 24);
        const uint32_t ff = uu ^ 0x40404000u;
        f[j] = ff | ((ff & 0x40404000u) << 1) | ((uint32_t)(M >> 21) & 255u);
      }
      const uint32_t lo01 = __builtin_amdgcn_perm(f[1], f[0], 0x05010400u);
      const uint32_t hi01 = __builtin_amdgcn_perm(f[1], f[0], 0x07030602u);
      const uint32_t lo23 = __builtin_amdgcn_perm(f[3], f[2], 0x05010400u);
      const uint32_t hi23 = __builtin_amdgcn_perm(f[3], f[2], 0x07030602u);
      const uint32_t pl[4] = {__builtin_amdgcn_perm(lo23, lo01, 0x05040100u),
                              __builtin_amdgcn_perm(lo23, lo01, 0x07060302u),
                              __builtin_amdgcn_perm(hi23, hi01, 0x05040100u),
                              __builtin_amdgcn_perm(hi23, hi01, 0x07060302u)};
      int8_t* base = As + q * AB + r * KS + 16 * swz(r, g >> 2) + 4 * (g & 3);
#pragma unroll
      for (int p = 0; p < 4; ++p) *reinterpret_cast<uint32_t*>(base + p * BMI * KS) = pl[p];
    }
  };
#define NOPT_GLOAD_A(k0)                                                                                 \
  do {                                                                                                   \
    if constexpr (PK) gload_packed(k0);                                                                  \
    else                                                                                                 \
      _Pragma("unroll") for (int u = 0; u < APT; ++u) a_reg[u] =                                         \
          aok[u] ? *reinterpret_cast<const uint4*>(ap[u] + (k0)) : uint4{0, 0, 0, 0};                     \
  } while (0)
#define NOPT_GLOAD_C(k0)                                                                                 \
  do {                                                                                                   \
    static_for<CPT>([&](auto U) {                                                                        \
      constexpr int u_ = decltype(U)::value;                                                             \
      const uint4 v_ = (u_ < CPT - 1 || c_last) ? *reinterpret_cast<const uint4*>(Cbase + c_g[u_] + (k0)) \
                                                : uint4{0, 0, 0, 0};                                     \
      if constexpr (u_ < CH) c_lo[u_ < CH ? u_ : 0] = v_;                                                \
      else c_hi[u_ < CH ? 0 : u_ - CH] = v_;                                                             \
    });                                                                                                  \
  } while (0)
#define NOPT_LSTORE(buf)                                                                                 \
  do {                                                                                                   \
    int8_t* As_ = sm[buf];                                                                               \
    if constexpr (!PK)                                                                                   \
      _Pragma("unroll") for (int u = 0; u < APT; ++u) {                                                  \
        const int e_ = t + NTH * u;                                                                      \
        *reinterpret_cast<uint4*>(As_ + (e_ / ACQ) * AB + (((e_ % ACQ) / (4 * BMI)) * BMI + ar) * KS +   \
                                  16 * swz(ar, ac)) = a_reg[u];                                          \
      }                                                                                                  \
    static_for<CPT>([&](auto U) {                                                                        \
      constexpr int u_ = decltype(U)::value;                                                             \
      uint4* d_ = reinterpret_cast<uint4*>(As_ + c_s[u_]);                                               \
      if (u_ < CPT - 1 || c_last) {                                                                      \
        if constexpr (u_ < CH) *d_ = c_lo[u_ < CH ? u_ : 0];                                             \
        else *d_ = c_hi[u_ < CH ? 0 : u_ - CH];                                                          \
      }                                                                                                  \
    });                                                                                                  \
  } while (0)
  v4i_t acc[IT][JT][5];
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int j = 0; j < JT; ++j)
#pragma unroll
      for (int c = 0; c < 5; ++c) acc[i][j][c] = v4i_t{0, 0, 0, 0};
  // the k-slabs to multiply: every slab (ld8 <= 2048: at most 32) less those
  // in which the block's rows of B hold no nonzero digit (Bm: a 32-bit mask
  // per function and 32 rows); a zero slab adds nothing to the class sums,
  // so leaving it out changes no bit of X
  const int nks = (int)(ld8 / KS);
  uint32_t todo = nks >= 32 ? 0xffffffffu : (1u << nks) - 1u;
  if (Bm) {
    uint32_t any = 0;
#pragma unroll
    for (int q = 0; q < QPB; ++q)
#pragma unroll
      for (int h = 0; h < BMI / 32; ++h) {
        const int mb = m0 / 32 + h;
        if (Qb + q < naux && mb < nmb) any |= Bm[(Qb + q) * nmb + mb];
      }
    todo &= any;
  }
  todo = __builtin_amdgcn_readfirstlane(todo);
  bool have = todo != 0;
  if (have) {
    const int k0 = __builtin_ctz(todo) * KS;
    todo &= todo - 1u;
    NOPT_GLOAD_A(k0);
    if constexpr (PK) lstore_packed(sm[0]);
    NOPT_GLOAD_C(k0);
    NOPT_LSTORE(0);
    __syncthreads();
  }
  const int lr = lane & 15, lc = lane >> 4;
  for (int ks = 0; have; ++ks) {
    const int cur = ks & 1;
    const bool more = todo != 0;
    const int kn = more ? __builtin_ctz(todo) * KS : 0;  // the next slab
    // the B slab (HBM) is fetched before the products, the small C slab
    // (L2-resident) after them (measured faster than fetching both first)
    if (more) NOPT_GLOAD_A(kn);
    const int8_t* As = sm[cur] + qq * AB;
    const int8_t* Cs = sm[cur] + QPB * AB;
    uint4 av[IT][4];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int r = wr * 16 * IT + i * 16 + lr;
#pragma unroll
      for (int p = 0; p < 4; ++p) av[i][p] = *reinterpret_cast<const uint4*>(As + (p * BMI + r) * KS + 16 * swz(r, lc));
    }
    // one digit plane l of the C operand at a time: pairs (k, l), k+l <= 4;
    // each plane's LDS read is issued before the products of the previous
    // plane, so that its latency hides behind them (two operand registers
    // in turn; the int32 sums are exact in any order)
    auto cplane_ptr = [&](int j, int l) {
      const int col = wc * 16 * JT + j * 16 + lr;
      return reinterpret_cast<const uint4*>(Cs + col * KS + 16 * swz(col, lc) + l * BN * KS);
    };
    uint4 bq[2];
    bq[0] = *cplane_ptr(0, 0);
    // scheduling groups (DS read 0x100, MFMA 0x008): the operand reads first,
    // then per plane one read of the next plane and the plane's products (not
    // for the packed loader, whose conversion leaves no registers to spare)
    if constexpr (!PK) __builtin_amdgcn_sched_group_barrier(0x100, 4 * IT + 1, 0);
    static_for<4 * JT>([&](auto S) {
      constexpr int s_ = decltype(S)::value, j = s_ / 4, l = s_ % 4;
      if constexpr (s_ + 1 < 4 * JT) bq[(s_ + 1) & 1] = *cplane_ptr((s_ + 1) / 4, (s_ + 1) % 4);
      const uint4 bv = bq[s_ & 1];
#pragma unroll
      for (int i = 0; i < IT; ++i) {
        v4i_t* A = acc[i][j];
#pragma unroll
        for (int k = 0; k + l <= 4 && k < 4; ++k) A[k + l] = mfma_i8(av[i][k], bv, A[k + l]);
      }
      if constexpr (!PK && s_ + 1 < 4 * JT) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if constexpr (!PK) __builtin_amdgcn_sched_group_barrier(0x008, IT * (l < 2 ? 4 : 5 - l), 0);
    });
    if (more) {
      // packed: the B slab is converted and stored before the C slab is
      // fetched (fewer registers live at once)
      if constexpr (PK) lstore_packed(sm[cur ^ 1]);
      NOPT_GLOAD_C(kn);
      NOPT_LSTORE(cur ^ 1);
    }
    __syncthreads();
    todo &= todo - 1u;
    have = more;
  }
#undef NOPT_GLOAD_A
#undef NOPT_GLOAD_C
#undef NOPT_LSTORE
  // X values in FP64 (rows m0 + rowl(i, v), columns colb(j) of the group)
  auto rowl = [&](int i, int v) { return wr * 16 * IT + i * 16 + 4 * lc + v; };
  auto colb = [&](int j) { return wc * 16 * JT + j * 16 + lr; };
  double xv[IT][JT][4];
  // 2^28 x the digit sum: hi = 128 A0 + A1 is exact in int32 (|A0| <= 4096 ld8,
  // ld8 <= 2048), so is mid = 128 A2 + A3 for ld8 <= 1024 (|A2| <= 3 * 4096 ld8),
  // and every partial sum below 2^52 is exact in FP64
  // the row and column exponents, all loads issued at once (indices clamped
  // into range; rows beyond nbf and a Q beyond naux give zeros below)
  int ebe[IT][4], ece[JT];
  {
    const long Qc = Q < naux ? Q : naux - 1;
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) ebe[i][v] = Be[Qc * nbf + min(m0 + rowl(i, v), nbf - 1)];
#pragma unroll
    for (int j = 0; j < JT; ++j) ece[j] = Ce[cg * BN + colb(j)];
  }
  auto combine = [&](auto MID) {
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + rowl(i, v);
        const bool live = m < nbf && Q < naux;
        const int em = ebe[i][v];
#pragma unroll
        for (int j = 0; j < JT; ++j) {
          const v4i_t* A = acc[i][j];
          const int hi = A[0][v] * 128 + A[1][v];
          double sv;
          if constexpr (decltype(MID)::value)
            sv = fma((double)hi, 0x1p21, fma((double)(A[2][v] * 128 + A[3][v]), 128.0, (double)A[4][v]));
          else
            sv = fma((double)hi, 0x1p21, fma((double)A[2][v], 0x1p14, fma((double)A[3][v], 128.0, (double)A[4][v])));
          xv[i][j][v] = live ? ldexp(sv, em + ece[j] - 40) : 0.0;
        }
      }
  };
  if (ld8 <= 1024)
    combine(std::true_type{});
  else
    combine(std::false_type{});
  if (xs.nseg == 0) {
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + rowl(i, v);
        if (m < nbf && Q < naux) {
          double* xr = X + ((long)Q * nbf + m) * ldx + cg * BN;
#pragma unroll
          for (int j = 0; j < JT; ++j) xr[colb(j)] = xv[i][j][v];
        }
      }
    return;
  }
  // fused digit split (the segments inside the block's column group): per
  // segment the row exponent (the
  // largest frexp exponent over the segment's columns: 16 lanes of a wave,
  // then the WC waves of a row and the QPB functions of the block via LDS),
  // then each value as the integer N = rint(x 2^(27-e)), |N| <= 2^27, in
  // balanced base-128 digits N = 2^21 d0 + 2^14 d1 + 2^7 d2 + d3 (|d| <= 64),
  // staged in LDS (row blockIdx.y*nbf + m: the block's functions side by
  // side, len columns each) as one dword per column holding the four digits
  // (byte k = digit k; rows of ldk + 4 dwords, so that the four row groups
  // of a wave write to different banks), and written out per 16 columns:
  // four 4 x 4 byte transposes give the 16 bytes of each digit plane.
  // The staging tile overlays the operand buffers: both stages when the row
  // exponents stay in registers, else the first, with the exponents of the
  // QPB x WC waves of a row in the second (no LDS beyond the operand
  // buffers).
  constexpr bool kRegExp = WC == 1 && QPB == 1;
  constexpr int kLdkMax = (QPB * BN + 63) / 64 * 64;
  static_assert(4 * BMI * (kLdkMax + 4) <= (kRegExp ? 2 : 1) * SB && QPB * WC * BMI * 4 <= SB, "epilogue LDS");
  constexpr int kNoExp = -4096;  // exponent of an all-zero row segment
  uint32_t* st = reinterpret_cast<uint32_t*>(&sm[0][0]);
  int(*rexp)[BMI] = reinterpret_cast<int(*)[BMI]>(&sm[1][0]);
  const long orow = (long)vb.y * nbf;  // first output row of the block
  for (int d = 0; d < xs.nseg; ++d) {
    if (xs.off[d] / BN != cg) continue;  // a segment of another column group (uniform per block)
    const int o = xs.off[d] - cg * BN, len = xs.len[d], pad = (len + 15) & ~15;  // o: within the group
    const int ldk = (int)xs.ldk[d], srow = ldk + 4;
    const int nch = BMI * srow / 4;  // 16-byte chunks of the staging tile
    for (int i = t; i < nch; i += NTH) reinterpret_cast<uint4*>(st)[i] = uint4{0, 0, 0, 0};
    int rx[IT][4];
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int ex = kNoExp;
#pragma unroll
        for (int j = 0; j < JT; ++j) {
          const int c0 = wc * 16 * JT + j * 16;
          const double x = xv[i][j][v];
          int k;
          (void)frexp(x, &k);
          if (c0 >= o && c0 < o + pad && x != 0.0) ex = max(ex, k);
        }
        ex = row_max16(ex);
        if constexpr (kRegExp)
          rx[i][v] = ex;
        else if (lr == 0)
          rexp[qq * WC + wc][rowl(i, v)] = ex;
      }
    // orders the zeroing of the staging tile (and the row exponents of the
    // other waves) before the digit writes below
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rowl(i, v);
        const int m = m0 + row;
        int e;
        if constexpr (kRegExp) {
          e = rx[i][v];
        } else {
          e = rexp[0][row];
#pragma unroll
          for (int q = 1; q < QPB * WC; ++q) e = max(e, rexp[q][row]);
        }
        if (e == kNoExp) e = 0;
        if (qq == 0 && wc == 0 && lr == 0 && m < nbf) xs.xe[d][orow + m] = e;
#pragma unroll
        for (int j = 0; j < JT; ++j) {
          const int c0 = wc * 16 * JT + j * 16;
          const int cl = c0 + lr - o;  // column within the segment
          if (c0 >= o && c0 < o + pad && cl < len) {
            // the balanced digits of N at once: M = N + 64 (1 + 2^7 + 2^14)
            // holds d3 + 64, d2 + 64, d1 + 64 as unsigned 7-bit fields at
            // bits 0, 7, 14 and d0 above bit 21; each field u becomes the
            // byte u - 64 (u ^ 64, bit 6 copied into bit 7)
            const int M = (int)rint(ldexp(xv[i][j][v], 27 - e)) + 1056832;
            const uint32_t u = (((uint32_t)M >> 14 & 127u) << 8) | (((uint32_t)M >> 7 & 127u) << 16) |
                               (((uint32_t)M & 127u) << 24);
            const uint32_t f = u ^ 0x40404000u;
            st[row * srow + qq * len + cl] = f | ((f & 0x40404000u) << 1) | ((uint32_t)(M >> 21) & 255u);
          }
        }
      }
    __syncthreads();
    const int cpr = ldk / 16;  // 16-column units per row
    for (int i = t; i < BMI * cpr; i += NTH) {
      const int row = i / cpr, ch = i % cpr;
      const int m = m0 + row;
      if (m >= nbf) continue;
      const uint4* src = reinterpret_cast<const uint4*>(st + row * srow + 16 * ch);
      uint32_t pl[4][4];  // pl[plane][a]: the plane's bytes of columns 16 ch + 4a .. 4a+3
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const uint4 q = src[a];
        // byte k of __builtin_amdgcn_perm(x, y, sel) is byte sel_k of y (0..3) or x (4..7)
        const uint32_t lo01 = __builtin_amdgcn_perm(q.y, q.x, 0x05010400u);  // x0 y0 x1 y1
        const uint32_t hi01 = __builtin_amdgcn_perm(q.y, q.x, 0x07030602u);  // x2 y2 x3 y3
        const uint32_t lo23 = __builtin_amdgcn_perm(q.w, q.z, 0x05010400u);  // z0 w0 z1 w1
        const uint32_t hi23 = __builtin_amdgcn_perm(q.w, q.z, 0x07030602u);  // z2 w2 z3 w3
        pl[0][a] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);          // x0 y0 z0 w0
        pl[1][a] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);          // x1 y1 z1 w1
        pl[2][a] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);          // x2 y2 z2 w2
        pl[3][a] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);          // x3 y3 z3 w3
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
        *reinterpret_cast<uint4*>(xs.xd[d] + p * xs.xplane[d] + (orow + m) * (long)ldk + 16 * ch) =
            uint4{pl[p][0], pl[p][1], pl[p][2], pl[p][3]};
    }
    __syncthreads();
  }
}
template <int NT, int WC, int NW, int QPB, int BMI, int WPE, bool PK>
__global__ __launch_bounds__(64 * NW * QPB) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_half_i8(
    const int8_t* __restrict__ Bd, long bplane, const int* __restrict__ Bp, long sq, const int8_t* __restrict__ Be8,
    const int* __restrict__ Be, const uint32_t* __restrict__ Bm, int nmb, long ld8, int nbf, int naux,
    const int8_t* __restrict__ Cd, long cplane, const int* __restrict__ Ce, double* __restrict__ X, long ldx,
    const XSplit xs) {
  k_half_i8_body<NT, WC, NW, QPB, BMI, WPE, PK>(grid_vb(),
      Bd, bplane, Bp, sq, Be8, Be, Bm, nmb, ld8, nbf, naux, Cd, cplane, Ce, X, ldx, xs);
}
template <int NT, int WC, int NW, int QPB, int BMI, int WPE, bool PK>
__global__ __launch_bounds__(64 * NW * QPB) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void k_half_i8_list(const ItemHalf* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemHalf& a = items[i];
  k_half_i8_body<NT, WC, NW, QPB, BMI, WPE, PK>(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.Bd, a.bplane, a.Bp, a.sq, a.Be8, a.Be, a.Bm, a.nmb, a.ld8, a.nbf, a.naux, a.Cd, a.cplane, a.Ce, a.X, a.ldx,
      a.xs);
}


// ------------------------------------------------- Coulomb over packed rows
// The lower triangle of every B_Q is cut into 8-column slots (m, n0), n0 <= m,
// listed in `slots` as (m << 16) | n0. Both Coulomb kernels read the int32
// parts and the bytes of a slot (two 16-byte and one 8-byte load) and
// rebuild the integers N35 = 128 N28 + d4 exactly in FP64.
struct Dig8 {
  uint4 p[2];  // N28 of the 8 columns
  uint2 b;     // their last digits d4
};
// element e = Q*sq + 16*seg0(m) + n0 of the packed layout
__device__ __forceinline__ void load_dig8(const int8_t* Bs, long splane, long e, Dig8& d) {
  const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const int*>(Bs) + e);
  d.p[0] = p[0];
  d.p[1] = p[1];
  d.b = *reinterpret_cast<const uint2*>(Bs + 4 * splane + e);
}
// the two 8-column halves of a 16-column segment
__device__ __forceinline__ void load_dig16(const int8_t* Bs, long splane, long e, Dig8& lo, Dig8& hi) {
  const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const int*>(Bs) + e);
  lo.p[0] = p[0];
  lo.p[1] = p[1];
  hi.p[0] = p[2];
  hi.p[1] = p[3];
  const uint4 b = *reinterpret_cast<const uint4*>(Bs + 4 * splane + e);
  lo.b = uint2{b.x, b.y};
  hi.b = uint2{b.z, b.w};
}
__device__ __forceinline__ void unpack8(const Dig8& d, double (&v)[8]) {
  const int* q = reinterpret_cast<const int*>(d.p);
  const int8_t* b = reinterpret_cast<const int8_t*>(&d.b);
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = fma((double)q[j], 128.0, (double)b[j]);
}

// Sum over the 64 lanes of a wave, valid in every lane: rotations within each
// row of 16 lanes (DPP row_ror 8, 4, 2, 1: register moves, no LDS round
// trip), then the four row sums read from lanes 0, 16, 32, 48 and added in
// that order (a fixed order: reproducible).
template <int N>
__device__ __forceinline__ double row_ror(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x120 + N, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x120 + N, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}
__device__ __forceinline__ double lane_read(double x, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
__device__ __forceinline__ double wave_sum_rows(double v) {
  v += row_ror<8>(v);
  v += row_ror<4>(v);
  v += row_ror<2>(v);
  v += row_ror<1>(v);
  return ((lane_read(v, 0) + lane_read(v, 16)) + lane_read(v, 32)) + lane_read(v, 48);
}

// gpart[sb][d][Q] = sum over the slots of slot block sb of B_Q,mn D_d,mn
// (off-diagonal weight 2), Q in split blockIdx.y. Each wave takes one Q at a
// time over all 64*SPL slots of the block (SPL per lane, the weighted density
// values held in registers) and reduces across its lanes; the digits of its
// next Q are loaded while the current one is reduced.
template <int ND, int SPL, bool SEG>
__device__ __forceinline__ void k_gamma_slots_body(const VB vb, const int8_t* __restrict__ Bs, long splane, long sq,
                                                     const int* __restrict__ Be, int nbf, int naux,
                                                     const int* __restrict__ slots, int nslots,
                                                     const double* __restrict__ Dp, long ldd, long dstride,
                                                     double* __restrict__ gpart) {
  const int sb = vb.x, sp = vb.y, ns = vb.gy;
  const int q0 = (int)((long)naux * sp / ns), q1 = (int)((long)naux * (sp + 1) / ns);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  static_assert(!SEG || SPL == 2, "a segment is two slots");
  int mrow[SPL], ncol[SPL];
  double dv[ND][SPL][8];
  // every lane loads (empty slots and columns beyond the row read a valid
  // element and weigh it zero), so that the loads issue back to back
  bool okk[SPL];
#pragma unroll
  for (int k = 0; k < SPL; ++k) {
    // SEG: the table lists 16-column segments, lane's slots are its two halves
    const int sl = SEG ? sb * 64 + lane : sb * 64 * SPL + k * 64 + lane;
    okk[k] = sl < nslots;
    const int code = slots[min(sl, nslots - 1)];
    mrow[k] = okk[k] ? code >> 16 : 0;  // row 0 stands in for an empty slot (zero weights)
    ncol[k] = okk[k] ? (code & 0xffff) + (SEG ? 8 * k : 0) : 0;
  }
#pragma unroll
  for (int k = 0; k < SPL; ++k)
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int n = ncol[k] + j, m = mrow[k];
        const double x = Dp[d * dstride + (long)m * ldd + min(n, m)];
        dv[d][k][j] = (okk[k] && n <= m) ? (n < m ? 2.0 : 1.0) * x : 0.0;
      }
  Dig8 cur[SPL], nxt[SPL];
  int ecur[SPL], enxt[SPL];
  int boff[SPL];  // element offset of the slot in a packed B_Q
#pragma unroll
  for (int k = 0; k < SPL; ++k) boff[k] = (int)(16 * seg0(mrow[k])) + ncol[k];
  auto fetch = [&](int q, Dig8 (&dst)[SPL], int (&e)[SPL]) {
    if constexpr (SEG) {
      load_dig16(Bs, splane, (long)q * sq + boff[0], dst[0], dst[1]);
      e[0] = e[1] = Be[(long)q * nbf + mrow[0]];
    } else {
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        load_dig8(Bs, splane, (long)q * sq + boff[k], dst[k]);
        e[k] = Be[(long)q * nbf + mrow[k]];
      }
    }
  };
  int q = q0 + w;
  if (q < q1) fetch(q, cur, ecur);
  for (; q < q1; q += 4) {
    if (q + 4 < q1) fetch(q + 4, nxt, enxt);
    double acc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[d] = 0.0;
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      double bv[8];
      unpack8(cur[k], bv);
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double sacc = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) sacc += bv[j] * dv[d][k][j];
        acc[d] += ldexp(sacc, ecur[k] - 34);
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double tot = wave_sum_rows(acc[d]);
      if (lane == 0) gpart[((long)sb * ND + d) * naux + q] = tot;
    }
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      cur[k] = nxt[k];
      ecur[k] = enxt[k];
    }
  }
}
template <int ND, int SPL, bool SEG>
__global__ __launch_bounds__(256) void k_gamma_slots(const int8_t* __restrict__ Bs, long splane, long sq,
                                                     const int* __restrict__ Be, int nbf, int naux,
                                                     const int* __restrict__ slots, int nslots,
                                                     const double* __restrict__ Dp, long ldd, long dstride,
                                                     double* __restrict__ gpart) {
  k_gamma_slots_body<ND, SPL, SEG>(grid_vb(), Bs, splane, sq, Be, nbf, naux, slots, nslots, Dp, ldd, dstride, gpart);
}
template <int ND, int SPL, bool SEG>
__global__ __launch_bounds__(256) void k_gamma_slots_list(const ItemGamma* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemGamma& a = items[i];
  k_gamma_slots_body<ND, SPL, SEG>(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.Bs, a.splane, a.sq, a.Be, a.nbf, a.naux, a.slots, a.nslots, a.Dp, a.ldd, a.dstride, a.gpart);
}


// gamma[d][Q] = sum_sb gpart[sb][d][Q]. A block takes 64 consecutive Q of one
// density with four waves: wave p adds the slot blocks sb = p, p+4, p+8, ...
// (eight loads in flight at a time), then the four sums are added in the
// order p = 0, 1, 2, 3, so that the result is reproducible.
__device__ __forceinline__ void k_gamma_sum_body(const VB vb, const double* __restrict__ gpart, int nsb, int nd,
    int naux,
                                                   double* __restrict__ gamma, long gstride) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63, p = threadIdx.x >> 6;
  const int d = vb.y, q = vb.x * 64 + lane;
  double s = 0.0;
  if (q < naux) {
    const double* g = gpart + (long)d * naux + q;
    const long st = (long)nd * naux;  // between consecutive slot blocks
    int b = p;
    for (; b + 28 < nsb; b += 32) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = g[(long)(b + 4 * u) * st];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; b < nsb; b += 4) s += g[(long)b * st];
  }
  part[p][lane] = s;
  __syncthreads();
  if (p == 0 && q < naux) gamma[d * gstride + q] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}
__global__ __launch_bounds__(256) void k_gamma_sum(const double* __restrict__ gpart, int nsb, int nd, int naux,
                                                   double* __restrict__ gamma, long gstride) {
  k_gamma_sum_body(grid_vb(), gpart, nsb, nd, naux, gamma, gstride);
}
__global__ __launch_bounds__(256) void k_gamma_sum_list(const ItemGammaSum* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemGammaSum& a = items[i];
  k_gamma_sum_body(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.gpart, a.nsb, a.nd, a.naux, a.gamma, a.gstride);
}


// Jpart[d][split][m][n] = sum_{Q in split} gamma_d[Q] B_Q,mn over the slots:
// each thread owns SPL slots (8 columns each) and accumulates over the Q of
// its split in registers, loading the digits of the next Q while the current
// one is accumulated.
template <int ND, int SPL, bool SEG>
__device__ __forceinline__ void k_jpart_slots_body(const VB vb, const int8_t* __restrict__ Bs, long splane, long sq,
                                                     const int* __restrict__ Be, int nbf, int naux,
                                                     const int* __restrict__ slots, int nslots,
                                                     const double* __restrict__ gamma, long gstride,
                                                     double* __restrict__ Jpart, long jstride) {
  const int sb = vb.x, sp = vb.y, ns = vb.gy;
  const int q0 = (int)((long)naux * sp / ns), q1 = (int)((long)naux * (sp + 1) / ns);
  int mrow[SPL], ncol[SPL];
  bool live[SPL];
  double acc[ND][SPL][8];
#pragma unroll
  for (int k = 0; k < SPL; ++k) {
    const int sl = SEG ? sb * 256 + (int)threadIdx.x : sb * 256 * SPL + k * 256 + threadIdx.x;
    const int code = sl < nslots ? slots[min(sl, nslots - 1)] : -1;
    live[k] = code >= 0;
    mrow[k] = live[k] ? code >> 16 : 0;  // row 0 stands in for an empty slot (result not written)
    ncol[k] = live[k] ? (code & 0xffff) + (SEG ? 8 * k : 0) : 0;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[d][k][j] = 0.0;
  }
  Dig8 cur[SPL], nxt[SPL];
  int ecur[SPL], enxt[SPL];
  int boff[SPL];  // element offset of the slot in a packed B_Q
#pragma unroll
  for (int k = 0; k < SPL; ++k) boff[k] = (int)(16 * seg0(mrow[k])) + ncol[k];
  auto fetch = [&](int q, Dig8 (&dst)[SPL], int (&e)[SPL]) {
    if constexpr (SEG) {
      load_dig16(Bs, splane, (long)q * sq + boff[0], dst[0], dst[1]);
      e[0] = e[1] = Be[(long)q * nbf + mrow[0]];
    } else {
#pragma unroll
      for (int k = 0; k < SPL; ++k) {
        load_dig8(Bs, splane, (long)q * sq + boff[k], dst[k]);
        e[k] = Be[(long)q * nbf + mrow[k]];
      }
    }
  };
  if (q0 < q1) fetch(q0, cur, ecur);
  for (int q = q0; q < q1; ++q) {
    if (q + 1 < q1) fetch(q + 1, nxt, enxt);
    double g[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) g[d] = gamma[d * gstride + q];
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      double bv[8];
      unpack8(cur[k], bv);
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double gs = ldexp(g[d], ecur[k] - 34);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[d][k][j] += gs * bv[j];
      }
    }
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
      cur[k] = nxt[k];
      ecur[k] = enxt[k];
    }
  }
#pragma unroll
  for (int k = 0; k < SPL; ++k) {
    if (!live[k]) continue;
    const int m = mrow[k];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double* Js = Jpart + d * jstride + ((long)sp * nbf + m) * nbf;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (ncol[k] + j <= m) Js[ncol[k] + j] = acc[d][k][j];
    }
  }
}
template <int ND, int SPL, bool SEG>
__global__ __launch_bounds__(256) void k_jpart_slots(const int8_t* __restrict__ Bs, long splane, long sq,
                                                     const int* __restrict__ Be, int nbf, int naux,
                                                     const int* __restrict__ slots, int nslots,
                                                     const double* __restrict__ gamma, long gstride,
                                                     double* __restrict__ Jpart, long jstride) {
  k_jpart_slots_body<ND, SPL, SEG>(grid_vb(),
      Bs, splane, sq, Be, nbf, naux, slots, nslots, gamma, gstride, Jpart, jstride);
}
template <int ND, int SPL, bool SEG>
__global__ __launch_bounds__(256) void k_jpart_slots_list(const ItemJ* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemJ& a = items[i];
  k_jpart_slots_body<ND, SPL, SEG>(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.Bs, a.splane, a.sq, a.Be, a.nbf, a.naux, a.slots, a.nslots, a.gamma, a.gstride, a.Jpart, a.jstride);
}


// ------------------------------------------------------------------ exchange
__device__ __forceinline__ void tri_tile(int t, int& bi, int& bj) {
  bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  while (bi * (bi + 1) / 2 > t) --bi;
  bj = t - bi * (bi + 1) / 2;
}

// Kpart[s][tile][r][c] = sum_{Q in split s} 2^(e_Qm + e_Qn - 12) sum_{k+l<=4}
//   2^(-7(k+l)) <Xd_k[Q][m][:], Xd_l[Q][n][:]>,  m = 64 bi + r, n = 64 bj + c,
// k, l < 4 (13 digit pairs in 5 exact int32 classes, combined in FP64 once
// per auxiliary function; with the pair rows of the fused half-transform a
// "function" here is a pair of them). Tile 64x64: NW waves in a (NW/2) x 2
// grid (NW = 8: 16x32 each, four waves per SIMD); the operands stream as
// 64-byte k-slabs over (Q, orbital chunk), double buffered in swizzled LDS
// with one barrier per slab; each thread loads the row exponents of its own
// accumulator rows and columns at the start of a function's last slab.
constexpr int KT8 = 64;
//
// Blocks are numbered so that each XCD (workgroups go to the 8 XCDs round
// robin) owns a contiguous range of the (tile, split) pairs, tile fastest:
// the tiles of one split read the same X rows at about the same time, from
// that XCD's L2. The grid is 1-D, ntiles * nsplit blocks; when that is not a
// multiple of 8, the first (ntiles * nsplit) % 8 XCDs take one pair more.
// BUF: the operand and exponent loads as buffer loads (32-bit offsets: a
// digit plane and the exponents below 2^31 bytes, which the launchers check),
// else as loads through 64-bit addresses.
template <int NW, bool BUF>
__device__ __forceinline__ void k_syrk_i8_body(const VB vb, 
    const int8_t* __restrict__ Xd, long xplane, const int* __restrict__ Xe, long ldk, int nbf, int naux,
    int ntiles, int nsplit, double* __restrict__ Kpart) {
  // NW waves in a (NW/2) x 2 grid: each wave 64/(NW/2) rows (IT row tiles) x 32 columns
  constexpr int KS = 64, TB = 4 * KT8 * KS;  // bytes of one operand tile (4 planes)
  constexpr int NTH = 64 * NW, IT = 64 / (NW / 2) / 16, OPT = 512 / NTH;  // operands loaded per thread
  static_assert(IT >= 1 && OPT >= 1 && IT * 16 * (NW / 2) == 64, "tile split");
  __shared__ __attribute__((aligned(16))) int8_t sm[2][2 * TB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  const int lr = lane & 15, lc = lane >> 4;
  const int total = ntiles * nsplit, per_xcd = total / 8, xcd = vb.x % 8;
  const int gidx = xcd * per_xcd + min(xcd, total % 8) + vb.x / 8;
  const int tile = gidx % ntiles, sp = gidx / ntiles, ns = nsplit;
  int bi, bj;
  tri_tile(tile, bi, bj);
  const int q0 = (int)((long)naux * sp / ns), q1 = (int)((long)naux * (sp + 1) / ns);
  const int nkq = (int)(ldk / KS);  // k-slabs per auxiliary function
  const int nsl = (q1 - q0) * nkq;
  // loader: row (t/4)%64 of operand tile (t/256 + o) (o < OPT; tile 0 rows
  // bi, tile 1 rows bj), chunk t%4, all 4 planes
  const int ar = (t >> 2) & 63, ac = t & 3, op0 = (OPT == 2) ? 0 : (t >> 8);
  // rows beyond nbf load row nbf - 1 and store zeros: every lane issues the
  // same loads, so that the waits count them exactly. Buffer form: a
  // resource per digit plane, the lane's row and chunk as the vector offset,
  // the slab's offset (the same for the whole block) as the scalar one.
  __amdgpu_buffer_rsrc_t rX[4];
  if constexpr (BUF)
#pragma unroll
    for (int p = 0; p < 4; ++p)
      rX[p] = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(Xd + p * xplane), (short)0, (int)xplane,
                                                0x00020000);
  long voff[OPT];
  bool rok[OPT];
#pragma unroll
  for (int o = 0; o < OPT; ++o) {
    const int m = ((op0 + o) == 0 ? bi : bj) * KT8 + ar;
    rok[o] = m < nbf;
    voff[o] = (long)min(m, nbf - 1) * ldk + 16 * ac;
  }
  const long qstride = (long)nbf * ldk;
  uint4 o_reg[OPT][4];
  auto gload = [&](int sl) {
    const int q = q0 + sl / nkq, k0 = (sl % nkq) * KS;
    if constexpr (BUF) {
      const int so = __builtin_amdgcn_readfirstlane((int)(q * qstride + k0));
#pragma unroll
      for (int o = 0; o < OPT; ++o)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const auto x = __builtin_amdgcn_raw_buffer_load_b128(rX[p], (int)voff[o], so, 0);
          o_reg[o][p] = uint4{x[0], x[1], x[2], x[3]};
        }
    } else {
      const int8_t* base = Xd + q * qstride + k0;
#pragma unroll
      for (int o = 0; o < OPT; ++o)
#pragma unroll
        for (int p = 0; p < 4; ++p) o_reg[o][p] = *reinterpret_cast<const uint4*>(base + p * xplane + voff[o]);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int o = 0; o < OPT; ++o) {
      int8_t* Ts = sm[buf] + (op0 + o) * TB;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        *reinterpret_cast<uint4*>(Ts + (p * KT8 + ar) * KS + 16 * swz(ar, ac)) =
            rok[o] ? o_reg[o][p] : uint4{0, 0, 0, 0};
    }
  };
  // exponents of this thread's accumulator rows (IT x 4) and columns (2)
  int ea[IT][4], eb[2];
  __amdgpu_buffer_rsrc_t rXe;
  if constexpr (BUF)
    rXe = __builtin_amdgcn_make_buffer_rsrc(const_cast<int*>(Xe), (short)0, (int)(4L * naux * nbf), 0x00020000);
  auto eload = [&](int q, int (&xa)[IT][4], int (&xb)[2]) {
    // rows beyond nbf read row nbf - 1 (their products are zero)
    auto ld = [&](int m) {
      if constexpr (BUF)
        return (int)__builtin_amdgcn_raw_buffer_load_b32(rXe, 4 * min(m, nbf - 1),
                                                         __builtin_amdgcn_readfirstlane(4 * q * nbf), 0);
      else
        return Xe[(long)q * nbf + min(m, nbf - 1)];
    };
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) xa[i][v] = ld(bi * KT8 + wr * 16 * IT + i * 16 + 4 * lc + v);
#pragma unroll
    for (int j = 0; j < 2; ++j) xb[j] = ld(bj * KT8 + wc * 32 + j * 16 + lr);
  };
  // a wave whose block lies wholly above the diagonal of a diagonal tile
  // computes nothing there (k_finalize reads only the lower triangle)
  const bool upper = bi == bj && (wr + 1) * 16 * IT <= wc * 32;
  v4i_t acc[IT][2][5];
  d4_t k64[IT][2];
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      k64[i][j] = d4_t{0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < 5; ++c) acc[i][j][c] = v4i_t{0, 0, 0, 0};
    }
  if (nsl > 0) {
    gload(0);
    lstore(0);
  }
  __syncthreads();
  for (int sl = 0; sl < nsl; ++sl) {
    const int cur = sl & 1;
    const bool more = sl + 1 < nsl;
    const bool qend = (sl % nkq) == nkq - 1;
    // the exponents of this function's rows, needed after its last slab,
    // requested before the next slab, so that waiting for them leaves the
    // slab's loads in flight
    if (qend) eload(q0 + sl / nkq, ea, eb);
    // (the last slab loads itself again, unused: every path then issues
    // the same loads, and the wait for the exponents leaves the slab's in
    // flight)
    gload(more ? sl + 1 : sl);
    const int8_t* As = sm[cur];
    const int8_t* Bs = sm[cur] + TB;
    if (!upper) {
      uint4 av[IT][4];
#pragma unroll
      for (int i = 0; i < IT; ++i) {
        const int r = wr * 16 * IT + i * 16 + lr;
#pragma unroll
        for (int p = 0; p < 4; ++p) av[i][p] = *reinterpret_cast<const uint4*>(As + (p * KT8 + r) * KS + 16 * swz(r, lc));
      }
      // the column operand one digit plane at a time, the next plane's LDS
      // read issued before the current plane's products (two registers in
      // turn, scheduling groups: one LDS read, then the products)
      auto bplane = [&](int j, int l) {
        const int cidx = wc * 32 + j * 16 + lr;
        return *reinterpret_cast<const uint4*>(Bs + cidx * KS + 16 * swz(cidx, lc) + l * KT8 * KS);
      };
      uint4 bq[2];
      bq[0] = bplane(0, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 4 * IT + 1, 0);
      static_for<8>([&](auto S) {
        constexpr int s_ = decltype(S)::value, j = s_ / 4, l = s_ % 4;
        if constexpr (s_ + 1 < 8) bq[(s_ + 1) & 1] = bplane((s_ + 1) / 4, (s_ + 1) % 4);
        const uint4 bv = bq[s_ & 1];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
          v4i_t* A = acc[i][j];
#pragma unroll
          for (int k = 0; k + l <= 4 && k < 4; ++k) A[k + l] = mfma_i8(av[i][k], bv, A[k + l]);
        }
        if constexpr (s_ + 1 < 8) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, IT * (l < 2 ? 4 : 5 - l), 0);
      });
      if (qend) {
        // combine the class sums of this auxiliary function: with |digit| <= 64
        // and ldk <= 1024, hi = 128 A0 + A1 and mid = 128 A2 + A3 are exact in
        // int32 and 2^21 hi + 2^7 mid + A4 (= 2^28 x the digit sum) in FP64
#pragma unroll
        for (int i = 0; i < IT; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            v4i_t* A = acc[i][j];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int hi = A[0][v] * 128 + A[1][v], mid = A[2][v] * 128 + A[3][v];
              const double sv = fma((double)hi, 0x1p21, fma((double)mid, 128.0, (double)A[4][v]));
              k64[i][j][v] += ldexp(sv, ea[i][v] + eb[j] - 40);
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) A[c] = v4i_t{0, 0, 0, 0};
          }
      }
    }
    if (more) lstore(cur ^ 1);
    __syncthreads();
  }
  double* Kt = Kpart + ((long)sp * ntiles + tile) * KT8 * KT8;
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = wr * 16 * IT + i * 16 + 4 * lc + v;
        const int cc = wc * 32 + j * 16 + lr;
        Kt[rr * KT8 + cc] = k64[i][j][v];
      }
}
template <int NW, bool BUF>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 2, NW / 2))) void k_syrk_i8(
    const int8_t* __restrict__ Xd, long xplane, const int* __restrict__ Xe, long ldk, int nbf, int naux,
    int ntiles, int nsplit, double* __restrict__ Kpart) {
  k_syrk_i8_body<NW, BUF>(grid_vb(), Xd, xplane, Xe, ldk, nbf, naux, ntiles, nsplit, Kpart);
}
template <int NW, bool BUF>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 2, NW / 2)))
void k_syrk_i8_list(const ItemSyrk* __restrict__ items,
    const int* __restrict__ first, int n) {
  const int i = list_item(first, n, (int)blockIdx.x);
  const ItemSyrk& a = items[i];
  k_syrk_i8_body<NW, BUF>(list_vb((int)blockIdx.x - first[i], a.gx, a.gy, a.gz),
      a.Xd, a.xplane, a.Xe, a.ldk, a.nbf, a.naux, a.ntiles, a.nsplit, a.Kpart);
}


}  // namespace

hipError_t split_rows_i8(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                         long rows_dst, long pstride, int* expo, hipStream_t s, const int* rperm, long rper,
                         const int* cperm) {
  if (rows_dst <= 0) return hipSuccess;
  if (rows_src > rows_dst || cols > ldo || pstride < rows_dst * ldo || (rperm && (rper <= 0 || rows_src % rper)))
    return hipErrorInvalidValue;
  k_split_rows<<<(unsigned)rows_dst, 64, 0, s>>>(src, ld, cols, rows_src, ns, planes, ldo, pstride, expo, rperm,
                                                 rper, cperm);
  return hipGetLastError();
}

namespace {
bool packed_chunk_ok(long cp, int m0, int m1, int nbf, long naux) {
  return m0 >= 0 && m1 <= nbf && m0 < m1 && naux > 0 && naux <= 65535 &&
         (long)m1 * (m1 + 1) / 2 - (long)m0 * (m0 + 1) / 2 <= cp;
}
dim3 packed_grid(int m0, int m1, long naux) { return dim3((m1 + 63) / 64, (m1 - m0 + 63) / 64, (unsigned)naux); }
}  // namespace

hipError_t unpack_packed_f64(const double* A, long cp, int m0, int m1, int nbf, long naux, double* B, long ld,
                             hipStream_t s) {
  if (!packed_chunk_ok(cp, m0, m1, nbf, naux) || ld < nbf) return hipErrorInvalidValue;
  k_packed_tiles<0><<<packed_grid(m0, m1, naux), 256, 0, s>>>(A, cp, m0, m1, nbf, B, ld, nullptr, nullptr, 0,
                                                               nullptr, 0, nullptr);
  return hipGetLastError();
}

hipError_t rowmax_packed(const double* A, long cp, int m0, int m1, int nbf, long naux, unsigned long long* rowmax,
                         hipStream_t s, const int* pinv) {
  if (!packed_chunk_ok(cp, m0, m1, nbf, naux)) return hipErrorInvalidValue;
  k_packed_tiles<1><<<packed_grid(m0, m1, naux), 256, 0, s>>>(A, cp, m0, m1, nbf, nullptr, 0, rowmax, nullptr, 0,
                                                               nullptr, 0, pinv);
  return hipGetLastError();
}

hipError_t rowmax_exponents(const unsigned long long* rowmax, long rows, int* expo, hipStream_t s) {
  if (rows > 0) k_rowmax_exponent<<<(unsigned)((rows + 255) / 256), 256, 0, s>>>(rowmax, rows, expo);
  return hipGetLastError();
}

hipError_t split_packed_i8(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* expo, int ns,
                           int8_t* planes, long ldo, long pstride, hipStream_t s, const int* pinv) {
  if (!packed_chunk_ok(cp, m0, m1, nbf, naux) || ldo < nbf || ns < 1 || ns > 8) return hipErrorInvalidValue;
  if (pstride < naux * nbf * ldo) return hipErrorInvalidValue;
  k_packed_tiles<2><<<packed_grid(m0, m1, naux), 256, 0, s>>>(A, cp, m0, m1, nbf, nullptr, ldo, nullptr, expo, ns,
                                                               planes, pstride, pinv);
  return hipGetLastError();
}

hipError_t scatter_segs_chunk(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* Be,
                              const int* pinv, int nseg, int8_t* Bs, long splane, hipStream_t s) {
  if (!packed_chunk_ok(cp, m0, m1, nbf, naux) || nseg != seg0(nbf) || splane < naux * 16L * nseg)
    return hipErrorInvalidValue;
  k_packed_tiles<3><<<packed_grid(m0, m1, naux), 256, 0, s>>>(A, cp, m0, m1, nbf, nullptr, 16L * nseg, nullptr, Be,
                                                               0, Bs, splane, pinv);
  return hipGetLastError();
}

namespace {
hipError_t half_i8_check(int nt, int qpb, dim3 grid, const HalfB& b, long ld8, int nbf, int naux, long ldx,
                         const XSplit& xs) {
  // ld8 <= 2048 keeps 128 A0 + A1 exact in int32 (the epilogue's class combine)
  if (ld8 % 64 != 0 || ld8 > 2048 || (long)grid.z * 16 * nt > ldx || qpb < 1 || qpb > 2)
    return hipErrorInvalidValue;
  if ((long)grid.y * qpb < naux || (long)(grid.y - 1) * qpb >= naux) return hipErrorInvalidValue;
  if (!b.Be || (!b.Bd && (!b.Bp || !b.Be8 || b.sq < 16 * seg0(nbf)))) return hipErrorInvalidValue;
  if (xs.nseg < 0 || xs.nseg > kMaxXSeg) return hipErrorInvalidValue;
  const long bn = 16L * nt, ldk_max = (qpb * bn + 63) / 64 * 64;
  for (int d = 0; d < xs.nseg; ++d) {
    const long o = xs.off[d], pad = (xs.len[d] + 15) & ~15;
    if (o % 16 || xs.len[d] <= 0 || xs.ldk[d] % 64 || xs.ldk[d] > ldk_max || xs.ldk[d] < (long)qpb * xs.len[d] ||
        o / bn != (o + pad - 1) / bn || o + pad > (long)grid.z * bn || !xs.xd[d] || !xs.xe[d])
      return hipErrorInvalidValue;
  }
  if (nt != 2 && nt != 4 && nt != 6) return hipErrorInvalidValue;
  return hipSuccess;
}
// The half-transform's configurations, one per (packed, nt, qpb):
// X(PK, NT, WC, NW, QPB, BMI, WPE). 64-row tiles of one function per block
// (two blocks per CU), or 32-row tiles of two functions (two blocks per CU);
// the packed loader wants one tile row per lane (QPB * BMI = 64).
#define NOPT_HALF_CONFIGS(X)   \
  X(false, 2, 1, 4, 1, 64, 2)  \
  X(false, 4, 2, 8, 1, 64, 4)  \
  X(false, 6, 2, 8, 1, 64, 4)  \
  X(false, 2, 1, 4, 2, 64, 4)  \
  X(false, 4, 2, 4, 2, 32, 4)  \
  X(false, 6, 2, 4, 2, 32, 4)  \
  X(true, 2, 1, 4, 1, 64, 2)   \
  X(true, 4, 2, 8, 1, 64, 4)   \
  X(true, 6, 2, 8, 1, 64, 4)   \
  X(true, 2, 1, 2, 2, 32, 4)   \
  X(true, 4, 2, 4, 2, 32, 4)   \
  X(true, 6, 2, 4, 2, 32, 4)
constexpr int half_cfg(bool pk, int nt, int qpb) { return (pk ? 64 : 0) + nt * 4 + qpb; }
// rows per block and threads of a configuration (0: none)
int half_i8_bmi(int nt, int qpb, bool pk) {
#define NOPT_HALF_BMI(PK_, NT_, WC_, NW_, QPB_, BMI_, WPE_) \
  if (half_cfg(pk, nt, qpb) == half_cfg(PK_, NT_, QPB_)) return BMI_;
  NOPT_HALF_CONFIGS(NOPT_HALF_BMI)
#undef NOPT_HALF_BMI
  return 0;
}
int half_i8_threads(int nt, int qpb, bool pk) {
#define NOPT_HALF_NTH(PK_, NT_, WC_, NW_, QPB_, BMI_, WPE_) \
  if (half_cfg(pk, nt, qpb) == half_cfg(PK_, NT_, QPB_)) return 64 * NW_ * QPB_;
  NOPT_HALF_CONFIGS(NOPT_HALF_NTH)
#undef NOPT_HALF_NTH
  return 0;
}

template <typename T>
void put_args(ListRec* r, const T& a) {
  r->args.resize(sizeof(T));
  std::memcpy(r->args.data(), &a, sizeof(T));
}
}  // namespace

hipError_t launch_half_i8(int nt, int qpb, dim3 grid, hipStream_t s, const HalfB& b, long ld8, int nbf, int naux,
                          const int8_t* Cd, long cplane, const int* Ce, double* X, long ldx, const XSplit& xs) {
  const hipError_t ce = half_i8_check(nt, qpb, grid, b, ld8, nbf, naux, ldx, xs);
  if (ce != hipSuccess) return ce;
#define NOPT_HALF_I8(PK_, NT_, WC_, NW_, QPB_, BMI_, WPE_)                                                  \
  case half_cfg(PK_, NT_, QPB_):                                                                          \
    k_half_i8<NT_, WC_, NW_, QPB_, BMI_, WPE_, PK_><<<dim3((nbf + BMI_ - 1) / BMI_, grid.y, grid.z),       \
                                                     64 * NW_ * QPB_, 0, s>>>(b.Bd, b.bplane, b.Bp, b.sq,  \
                                                                              b.Be8, b.Be, b.Bm, b.nmb, ld8, \
                                                                              nbf, naux, Cd, cplane, Ce, X, \
                                                                              ldx, xs);                     \
    break;
  switch (half_cfg(!b.Bd, nt, qpb)) {
    NOPT_HALF_CONFIGS(NOPT_HALF_I8)
    default: return hipErrorInvalidValue;
  }
#undef NOPT_HALF_I8
  return hipGetLastError();
}

int coulomb_seg_count(int nbf) { return (int)seg0(nbf); }

void coulomb_seg_table(int nbf, int* out) {
  long i = 0;
  for (int m = 0; m < nbf; ++m)
    for (int n0 = 0; n0 <= m; n0 += 16) out[i++] = (m << 16) | n0;
}

hipError_t pack_segs_rows(const double* B, long ld, int nbf, long naux, const int* Be, const int* seg_code, int nseg,
                          int8_t* Bs, long splane, hipStream_t s, const int* perm) {
  if (naux <= 0 || naux > 65535 || splane < naux * 16L * nseg || ld < nbf) return hipErrorInvalidValue;
  k_pack_segs<0><<<dim3((nseg + 255) / 256, (unsigned)naux), 256, 0, s>>>(B, ld, 0, nbf, Be, seg_code, 0, nseg,
                                                                         16L * nseg, Bs, splane, perm);
  return hipGetLastError();
}

hipError_t pack_segs_chunk(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* Be,
                           const int* seg_code, int nseg, int8_t* Bs, long splane, hipStream_t s) {
  if (!packed_chunk_ok(cp, m0, m1, nbf, naux) || splane < naux * 16L * nseg) return hipErrorInvalidValue;
  const int s0 = (int)seg0(m0), s1 = (int)seg0(m1);
  k_pack_segs<1><<<dim3((s1 - s0 + 255) / 256, (unsigned)naux), 256, 0, s>>>(A, cp, m0, nbf, Be, seg_code, s0, s1,
                                                                             16L * nseg, Bs, splane, nullptr);
  return hipGetLastError();
}

hipError_t expo_bytes(const int* Be, int nbf, long ld8, long naux, int8_t* Be8, hipStream_t s) {
  if (ld8 < nbf) return hipErrorInvalidValue;
  k_expo8<<<(unsigned)((naux * ld8 + 255) / 256), 256, 0, s>>>(Be, nbf, ld8, naux, Be8);
  return hipGetLastError();
}

hipError_t seg_live(const int8_t* Bs, long splane, int nseg, long naux, int* live, hipStream_t s) {
  if (nseg <= 0 || naux <= 0 || splane < naux * 16L * nseg) return hipErrorInvalidValue;
  const unsigned ranges = (unsigned)std::min(naux, 64L);
  k_seg_live<<<dim3((nseg + 255) / 256, ranges), 256, 0, s>>>(Bs, splane, 16L * nseg, nseg, (int)naux, live);
  return hipGetLastError();
}

hipError_t slab_mask_rows(const int8_t* Bd, long bplane, long ld8, int nbf, long naux, uint32_t* Bm,
                          hipStream_t s) {
  if (ld8 % 64 || ld8 > 2048 || ld8 < nbf || naux <= 0 || naux > 65535 || bplane < naux * nbf * ld8)
    return hipErrorInvalidValue;
  const int nmb = (nbf + 31) / 32;
  k_slab_mask_rows<<<dim3(nmb, (unsigned)naux), 256, 0, s>>>(Bd, bplane, ld8, nbf, nmb, Bm);
  return hipGetLastError();
}

hipError_t slab_mask_packed(const int* Bp, long sq, int nbf, long naux, uint32_t* Bm, hipStream_t s) {
  if (nbf > 2048 || naux <= 0 || naux > 65535 || sq < 16 * seg0(nbf)) return hipErrorInvalidValue;
  const int nmb = (nbf + 31) / 32;
  k_slab_mask_packed<<<dim3(nmb, (unsigned)naux), 256, 0, s>>>(Bp, sq, nbf, nmb, Bm);
  return hipGetLastError();
}

int coulomb_slot_count(int nbf) {
  long n = 0;
  for (int m = 0; m < nbf; ++m) n += m / 8 + 1;
  return (int)n;
}

void coulomb_slot_table(int nbf, int* out) {
  long i = 0;
  for (int m = 0; m < nbf; ++m)
    for (int n0 = 0; n0 <= m; n0 += 8) out[i++] = (m << 16) | n0;
}

int gamma_slot_blocks(int nslots) { return (nslots + 64 * kGammaSPL - 1) / (64 * kGammaSPL); }

hipError_t launch_gamma_slots(int nd, int nsplit, hipStream_t s, const int8_t* Bs, long splane, long sq, const int* Be,
                              int nbf, int naux, const int* slots, int nslots, const int* segs, int nseg,
                              const double* Dp, long ldd, long dstride, double* gpart, double* gamma, long gstride) {
  // one or two densities: a 16-column segment per lane (16-byte loads);
  // more: 8-column slots
  const int nsb = nd <= 2 ? (nseg + 63) / 64 : gamma_slot_blocks(nslots);
  const dim3 grid(nsb, nsplit);
  switch (nd) {
    case 1: k_gamma_slots<1, 2, true><<<grid, 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, segs, nseg, Dp, ldd, dstride, gpart); break;
    case 2: k_gamma_slots<2, 2, true><<<grid, 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, segs, nseg, Dp, ldd, dstride, gpart); break;
    case 3: k_gamma_slots<3, kGammaSPL, false><<<grid, 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, slots, nslots, Dp, ldd, dstride, gpart); break;
    case 4: k_gamma_slots<4, kGammaSPL, false><<<grid, 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, slots, nslots, Dp, ldd, dstride, gpart); break;
    default: return hipErrorInvalidValue;
  }
  k_gamma_sum<<<dim3((naux + 63) / 64, nd), 256, 0, s>>>(gpart, (int)grid.x, nd, naux, gamma, gstride);
  return hipGetLastError();
}

int gamma_partial_blocks(int nslots, int nseg) { return std::max(gamma_slot_blocks(nslots), (nseg + 63) / 64); }

hipError_t launch_jpart_slots(int nd, int nsplit, hipStream_t s, const int8_t* Bs, long splane, long sq, const int* Be,
                              int nbf, int naux, const int* slots, int nslots, const int* segs, int nseg,
                              const double* gamma, long gstride, double* Jpart, long jstride) {
  switch (nd) {
    case 1: k_jpart_slots<1, 2, true><<<dim3((nseg + 255) / 256, nsplit), 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, segs, nseg, gamma, gstride, Jpart, jstride); break;
    case 2: k_jpart_slots<2, 2, true><<<dim3((nseg + 255) / 256, nsplit), 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, segs, nseg, gamma, gstride, Jpart, jstride); break;
    case 3: k_jpart_slots<3, 1, false><<<dim3((nslots + 255) / 256, nsplit), 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, slots, nslots, gamma, gstride, Jpart, jstride); break;
    case 4: k_jpart_slots<4, 1, false><<<dim3((nslots + 255) / 256, nsplit), 256, 0, s>>>(Bs, splane, sq, Be, nbf, naux, slots, nslots, gamma, gstride, Jpart, jstride); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------- work lists
hipError_t split_rows_item(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                           long rows_dst, long pstride, int* expo, ListRec* r, const int* cperm) {
  if (rows_src > rows_dst || cols > ldo || pstride < rows_dst * ldo || rows_dst <= 0) return hipErrorInvalidValue;
  r->kind = kListSplit;
  r->cfg = 0;
  r->blocks = (int)rows_dst;
  r->threads = 64;
  put_args(r, ItemSplit{src, ld, cols, rows_src, ns, planes, ldo, pstride, expo, cperm, (int)rows_dst, 1, 1});
  return hipSuccess;
}

hipError_t half_i8_item(int nt, int qpb, dim3 grid, const HalfB& b, long ld8, int nbf, int naux, const int8_t* Cd,
                        long cplane, const int* Ce, double* X, long ldx, const XSplit& xs, ListRec* r) {
  const hipError_t ce = half_i8_check(nt, qpb, grid, b, ld8, nbf, naux, ldx, xs);
  if (ce != hipSuccess) return ce;
  const bool pk = !b.Bd;
  const int bmi = half_i8_bmi(nt, qpb, pk);
  const int gx = (nbf + bmi - 1) / bmi;
  r->kind = kListHalf;
  r->cfg = half_cfg(pk, nt, qpb);
  r->blocks = gx * (int)grid.y * (int)grid.z;
  r->threads = half_i8_threads(nt, qpb, pk);
  put_args(r, ItemHalf{b.Bd, b.bplane, b.Bp, b.sq, b.Be8, b.Be, b.Bm, b.nmb, ld8, nbf, naux, Cd, cplane, Ce, X, ldx,
                       xs, gx, (int)grid.y, (int)grid.z});
  return hipSuccess;
}

hipError_t gamma_slots_items(int nd, int nsplit, const int8_t* Bs, long splane, long sq, const int* Be, int nbf,
                             int naux, const int* slots, int nslots, const int* segs, int nseg, const double* Dp,
                             long ldd, long dstride, double* gpart, double* gamma, long gstride, ListRec* r) {
  if (nd < 1 || nd > 4 || nsplit < 1) return hipErrorInvalidValue;
  const bool seg = nd <= 2;
  const int nsb = seg ? (nseg + 63) / 64 : gamma_slot_blocks(nslots);
  r[0].kind = kListGamma;
  r[0].cfg = nd;
  r[0].blocks = nsb * nsplit;
  r[0].threads = 256;
  put_args(&r[0], ItemGamma{Bs, splane, sq, Be, nbf, naux, seg ? segs : slots, seg ? nseg : nslots, Dp, ldd, dstride,
                            gpart, nsb, nsplit, 1});
  r[1].kind = kListGammaSum;
  r[1].cfg = 0;
  r[1].blocks = (naux + 63) / 64 * nd;
  r[1].threads = 256;
  put_args(&r[1], ItemGammaSum{gpart, nsb, nd, naux, gamma, gstride, (naux + 63) / 64, nd, 1});
  return hipSuccess;
}

hipError_t jpart_slots_item(int nd, int nsplit, const int8_t* Bs, long splane, long sq, const int* Be, int nbf,
                            int naux, const int* slots, int nslots, const int* segs, int nseg, const double* gamma,
                            long gstride, double* Jpart, long jstride, ListRec* r) {
  if (nd < 1 || nd > 4 || nsplit < 1) return hipErrorInvalidValue;
  const bool seg = nd <= 2;
  const int gx = seg ? (nseg + 255) / 256 : (nslots + 255) / 256;
  r->kind = kListJ;
  r->cfg = nd;
  r->blocks = gx * nsplit;
  r->threads = 256;
  put_args(r, ItemJ{Bs, splane, sq, Be, nbf, naux, seg ? segs : slots, seg ? nseg : nslots, gamma, gstride, Jpart,
                    jstride, gx, nsplit, 1});
  return hipSuccess;
}

// the exchange's buffer-load form: 32-bit offsets cover a digit plane and the exponents
bool syrk_buf(long xplane, int nbf, int naux) { return xplane <= 0x7fffffffL && 4L * naux * nbf <= 0x7fffffffL; }

hipError_t syrk_i8_item(int ntiles, int nsplit, const int8_t* Xd, long xplane, const int* Xe, long ldk, int nbf,
                        int naux, double* Kpart, ListRec* r) {
  if (nsplit < 1 || ntiles <= 0 || ldk % 64 != 0 || ldk > 1024) return hipErrorInvalidValue;
  r->kind = kListSyrk;
  r->cfg = syrk_buf(xplane, nbf, naux) ? 5 : 4;
  r->blocks = ntiles * nsplit;
  r->threads = 256;
  put_args(r, ItemSyrk{Xd, xplane, Xe, ldk, nbf, naux, ntiles, nsplit, Kpart, ntiles * nsplit, 1, 1});
  return hipSuccess;
}

hipError_t launch_list(int kind, int cfg, const void* items, const int* first, int n, int blocks, hipStream_t s) {
  if (n < 1 || blocks < 1) return hipErrorInvalidValue;
  switch (kind) {
    case kListSplit:
      k_split_rows_list<<<blocks, 64, 0, s>>>((const ItemSplit*)items, first, n);
      break;
    case kListHalf: {
      const ItemHalf* it = (const ItemHalf*)items;
#define NOPT_HALF_L(PK_, NT_, WC_, NW_, QPB_, BMI_, WPE_)                                       \
  case half_cfg(PK_, NT_, QPB_):                                                               \
    k_half_i8_list<NT_, WC_, NW_, QPB_, BMI_, WPE_, PK_><<<blocks, 64 * NW_ * QPB_, 0, s>>>(it, first, n); \
    break;
      switch (cfg) {
        NOPT_HALF_CONFIGS(NOPT_HALF_L)
        default: return hipErrorInvalidValue;
      }
#undef NOPT_HALF_L
      break;
    }
    case kListGamma: {
      const ItemGamma* it = (const ItemGamma*)items;
      switch (cfg) {
        case 1: k_gamma_slots_list<1, 2, true><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 2: k_gamma_slots_list<2, 2, true><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 3: k_gamma_slots_list<3, kGammaSPL, false><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 4: k_gamma_slots_list<4, kGammaSPL, false><<<blocks, 256, 0, s>>>(it, first, n); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case kListGammaSum:
      k_gamma_sum_list<<<blocks, 256, 0, s>>>((const ItemGammaSum*)items, first, n);
      break;
    case kListJ: {
      const ItemJ* it = (const ItemJ*)items;
      switch (cfg) {
        case 1: k_jpart_slots_list<1, 2, true><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 2: k_jpart_slots_list<2, 2, true><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 3: k_jpart_slots_list<3, 1, false><<<blocks, 256, 0, s>>>(it, first, n); break;
        case 4: k_jpart_slots_list<4, 1, false><<<blocks, 256, 0, s>>>(it, first, n); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case kListSyrk:
      if (cfg == 5)
        k_syrk_i8_list<4, true><<<blocks, 256, 0, s>>>((const ItemSyrk*)items, first, n);
      else
        k_syrk_i8_list<4, false><<<blocks, 256, 0, s>>>((const ItemSyrk*)items, first, n);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_syrk_i8(int ntiles, int nsplit, hipStream_t s, const int8_t* Xd, long xplane, const int* Xe,
                          long ldk, int nbf, int naux, double* Kpart) {
  if (nsplit < 1 || ntiles <= 0 || ldk % 64 != 0 || ldk > 1024) return hipErrorInvalidValue;
  // four waves of 32 x 32 (two waves per SIMD, two blocks per CU)
  if (syrk_buf(xplane, nbf, naux))
    k_syrk_i8<4, true><<<(unsigned)(ntiles * nsplit), 256, 0, s>>>(Xd, xplane, Xe, ldk, nbf, naux, ntiles, nsplit,
                                                                    Kpart);
  else
    k_syrk_i8<4, false><<<(unsigned)(ntiles * nsplit), 256, 0, s>>>(Xd, xplane, Xe, ldk, nbf, naux, ntiles, nsplit,
                                                                     Kpart);
  return hipGetLastError();
}

}  // namespace nopt
