// Integer-split (Ozaki-type) int8 MFMA kernels (see ozaki.hip).
#pragma once
#include "common.hpp"

namespace nopt {
// Rows of src (rows_src x cols, ld) -> ns digit planes planes[k][row][ldo]
// (rows_dst >= rows_src rows, extra rows and columns zero) and exponents.
hipError_t split_rows_i8(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                         long rows_dst, int* expo, hipStream_t s);
// Half-transform X = B C^T from digit planes (B: 5 planes, C: 4 planes),
// FP32 output, FP64 gamma partials (grid: ceil(nbf/64) x naux x groups).
hipError_t launch_half_i8(int nt, dim3 grid, hipStream_t s, const int8_t* Bd, long bplane, const int* Be, long ld8,
                          int nbf, const int8_t* Cd, long cplane, const int* Ce, const double* Dp, long ldd,
                          float* X, long ldx, double* gpart, int mtiles);
}  // namespace nopt
