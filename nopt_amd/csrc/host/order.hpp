// An order of the basis functions for B's integer forms.
#pragma once

namespace nopt {

// perm[new] = old: an order of the nbf functions in which functions whose
// products (m, n) carry weight in the packed three-index tensor T (row-major
// [nrow][nbf(nbf+1)/2], lower triangle m >= n) sit close together. (m, n) is
// an edge when max_P |T[P][mn]| exceeds 2^-30 times the geometric mean of
// max_P |T[P][mm]| and max_P |T[P][nn]|; the order is reverse Cuthill-McKee
// on that graph (breadth-first from a low-degree node at the end of a
// longest level structure, neighbours by increasing degree, then reversed),
// component after component. Deterministic (ties by index).
void band_order(const double* T, long nrow, long nbf, int* perm);

}  // namespace nopt
