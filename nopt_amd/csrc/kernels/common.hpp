// Shared device helpers for the gfx950 RI-J/K kernels.
//
// MFMA operand layouts (CDNA3/CDNA4, 64-lane wavefront), checked against a
// host reference by the GPU tests:
//   16x16x4 f64 / f32 : A lane l = A[l%16][l/16], B lane l = B[l/16][l%16]
//   16x16x32 bf16     : A lane l = A[l%16][8*(l/16)+j], j<8 (8 bf16 per lane)
//                       B lane l = B[8*(l/16)+j][l%16]
//   16x16 f32 accumulator : lane l item v = D[4*(l/16)+v][l%16]
//   16x16 f64 accumulator : lane l item v = D[l/16+4*v][l%16]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef uint16_t bf16_t;  // bf16 storage

#define NOPT_WAVE 64

// Row of accumulator item v held by `lane` (element size selects the layout).
template <typename T>
__device__ __forceinline__ int acc_row(int lane, int v) {
  return sizeof(T) == 8 ? (lane / 16 + 4 * v) : (4 * (lane / 16) + v);
}

__device__ __forceinline__ d4_t mfma_f64(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma_f32(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma_bf16(const uint4& a, const uint4& b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, a),
                                                 __builtin_bit_cast(bf8_t, b), c, 0, 0, 0);
}

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));
}

__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum over 256 threads; result valid in thread 0
__device__ __forceinline__ double block_sum256(double v, double* sh /* >= 4 */) {
  v = warp_sum(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  return r;
}

static inline long round_up(long x, long m) { return (x + m - 1) / m * m; }

// Work lists: one launch runs a kernel over the grids of several items (the
// systems of a ragged batch). Item i owns the blocks [first[i], first[i+1])
// of the 1-D launch; a kernel body sees its block of the item's own grid as
// a VB, x fastest.
struct VB {
  int x, y, z, gx, gy, gz;
};
// the item of block b: the largest i with first[i] <= b (first[0] = 0)
__device__ __forceinline__ int list_item(const int* __restrict__ first, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// the launch's own block as a VB
__device__ __forceinline__ VB grid_vb() {
  return VB{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}
__device__ __forceinline__ VB list_vb(int local, int gx, int gy, int gz) {
  return VB{local % gx, (local / gx) % gy, local / (gx * gy), gx, gy, gz};
}
