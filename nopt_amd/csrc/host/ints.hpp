// Host integral engine (McMurchie-Davidson), FP64. See ints.cpp.
#pragma once
#include <vector>

#include "basis.hpp"

namespace nopt {

// Boys function F_0..F_nmax at T (nmax <= 23).
void boys(int nmax, double T, double* F);

// Overlap, kinetic and nuclear-attraction matrices (n x n, row-major, any of
// the outputs may be null).
void one_electron(const std::vector<Shell>& sh, int natom, const int* Z, const double* xyz,
                  double* S, double* T, double* V);

// Coulomb metric (P|Q), naux x naux.
void eri2c(const std::vector<Shell>& aux, double* V);

// 3-centre (P|mn): packed out[P*npair + m(m+1)/2 + n] (m >= n), or NOPT's
// RI_AA layout out[(m*n+n)*naux + P] when full_nopt is set. The caller zeroes
// the output first.
void eri3c(const std::vector<Shell>& sh, const std::vector<Shell>& aux, double* out,
           bool full_nopt);

}  // namespace nopt
