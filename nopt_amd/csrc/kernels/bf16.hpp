// Split-bf16 MFMA kernels (see bf16.hip).
#pragma once
#include "common.hpp"

namespace nopt {
// ns = 2 (bf16x3) or 3 (bf16x6); nt in {1, 2, 4, 6, 8}.
hipError_t launch_half_bf16(int ns, int nt, dim3 grid, hipStream_t s, const double* B, long ld, int nbf,
                            const float* Cp, const double* Dp, float* X, long ldx, double* gpart,
                            int mtiles);
hipError_t launch_syrk_bf16(int ns, dim3 grid, hipStream_t s, const float* X, long ldx, int nbf, int naux,
                            int kpad, double* Kpart);
}  // namespace nopt
