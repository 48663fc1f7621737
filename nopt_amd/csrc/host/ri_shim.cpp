// NOPT-compatible RI entry points (include/nopt_ri.h) over the C-ABI of
// include/nopt_jk.h. One process-wide context stands in for NOPT's RI module
// globals (src/RI.cpp:16-32).
#include "../../../include/nopt_ri.h"

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../../include/nopt_ints.h"
#include "../../../include/nopt_jk.h"
#include "errors.hpp"

namespace {
nopt_jk_ctx* g_ctx = nullptr;
long g_n_ao = 0, g_naux = 0;
int g_prec = NOPT_PREC_FP64;

int attach_packed(const double* V, const double* packed, long n_ao, long naux, int device) {
  regen_RI_AA_release();
  nopt_jk_ctx* c = nullptr;
  int rc = nopt_jk_create(device, n_ao, naux, &c);
  if (rc) return -rc;
  rc = nopt_jk_set_B(c, packed, V);
  if (rc) {
    nopt_jk_destroy(c);
    return -rc;
  }
  g_ctx = c;
  g_n_ao = n_ao;
  g_naux = naux;
  return 0;
}
}  // namespace

int RI_gpu_attach(const double* AUX_V, const double* RI_AA, long n_ao, long aux_n_ao, int device) {
  nopt::clear_error();
  if (!AUX_V || !RI_AA || n_ao <= 0 || aux_n_ao <= 0)
    return -nopt::set_error(NOPT_EINVAL, "RI_gpu_attach: null input or bad size");
  const long npair = n_ao * (n_ao + 1) / 2;
  std::vector<double> packed;
  try {
    packed.resize((size_t)aux_n_ao * npair);
  } catch (const std::bad_alloc&) {
    return -nopt::set_error(NOPT_ENOMEM, "RI_gpu_attach: host allocation failed");
  }
  // [(m*n_ao+n)*naux + P] -> [P][m(m+1)/2 + n], m >= n
  for (long m = 0; m < n_ao; ++m)
    for (long n = 0; n <= m; ++n) {
      const double* src = RI_AA + (m * n_ao + n) * aux_n_ao;
      const long p = m * (m + 1) / 2 + n;
      for (long P = 0; P < aux_n_ao; ++P) packed[(size_t)P * npair + p] = src[P];
    }
  return attach_packed(AUX_V, packed.data(), n_ao, aux_n_ao, device);
}

int RI_gpu_attach_molecule(int natom, const int* Z, const double* xyz_bohr, const char* basis_gbs,
                           const char* aux_gbs, int pure, int device) {
  nopt::clear_error();
  nopt_mol* mol = nullptr;
  int rc = nopt_mol_create(natom, Z, xyz_bohr, basis_gbs, aux_gbs, pure, &mol);
  if (rc) return -rc;
  long nbf = 0, naux = 0;
  nopt_mol_dims(mol, &nbf, &naux, nullptr, nullptr);
  std::vector<double> V, A;
  try {
    V.resize((size_t)naux * naux);
    A.resize((size_t)naux * (nbf * (nbf + 1) / 2));
  } catch (const std::bad_alloc&) {
    nopt_mol_destroy(mol);
    return -nopt::set_error(NOPT_ENOMEM, "RI_gpu_attach_molecule: host allocation failed");
  }
  rc = nopt_mol_metric(mol, V.data());
  if (!rc) rc = nopt_mol_eri3(mol, A.data(), 0);
  nopt_mol_destroy(mol);
  if (rc) return -rc;
  return attach_packed(V.data(), A.data(), nbf, naux, device);
}

int RI_gpu_set_precision(int prec) {
  if (prec < NOPT_PREC_FP64 || prec > NOPT_PREC_INT8)
    return -nopt::set_error(NOPT_EINVAL, "RI_gpu_set_precision: unknown mode");
  g_prec = prec;
  return 0;
}

int RI_core_realloc(long n_c, long n_ao) {
  if (!g_ctx) return -nopt::set_error(NOPT_ESTATE, "RI_core_realloc: no RI context (RI_gpu_attach first)");
  if (n_c < 0 || n_c > n_ao || n_ao != g_n_ao)
    return -nopt::set_error(NOPT_EINVAL, "RI_core_realloc: n_c or n_ao does not match the context");
  return 0;
}

int regen_RI_AA_release() {
  if (g_ctx) nopt_jk_destroy(g_ctx);
  g_ctx = nullptr;
  g_n_ao = g_naux = 0;
  return 0;
}

int DM_to_F_transform_RI(double* J, double* K, double* DM, double* MO, long n_c, long n_ao) {
  nopt::clear_error();
  if (!g_ctx) return -nopt::set_error(NOPT_ESTATE, "DM_to_F_transform_RI: no RI context (RI_gpu_attach first)");
  if (n_ao != g_n_ao) return -nopt::set_error(NOPT_EINVAL, "DM_to_F_transform_RI: n_ao does not match");
  const int rc = nopt_jk_build(g_ctx, DM, MO, n_c, J, K, g_prec);
  return rc ? -rc : 1;
}

int calc_2el_MO_INTS_RI(long n_ao, double* DM_D, double* J, double* K, double* MO_INTS, double* MO_VEC,
                        double* MO_VEC1, double* MO_VEC2, long n_c, long n_a, long t) {
  nopt::clear_error();
  if (MO_INTS != nullptr) {  // src/RI.cpp:197-226, before the J/K early return
    if (!g_ctx) return -nopt::set_error(NOPT_ESTATE, "calc_2el_MO_INTS_RI: no RI context");
    if (n_ao != g_n_ao) return -nopt::set_error(NOPT_EINVAL, "calc_2el_MO_INTS_RI: n_ao does not match");
    if (!MO_VEC1 || !MO_VEC2 || n_a < 1)
      return -nopt::set_error(NOPT_EINVAL, "calc_2el_MO_INTS_RI: MO_INTS needs MO_VEC1, MO_VEC2 and n_a >= 1");
    std::vector<double> c1, c2;
    const double* r1 = MO_VEC1;
    const double* r2 = MO_VEC2;
    if (t) {  // orbitals as columns, [m][a] with leading dimension n_a: rows for the device
      try {
        c1.resize((size_t)n_a * n_ao);
        c2.resize((size_t)n_a * n_ao);
      } catch (const std::bad_alloc&) {
        return -nopt::set_error(NOPT_ENOMEM, "calc_2el_MO_INTS_RI: host allocation failed");
      }
      for (long m = 0; m < n_ao; ++m)
        for (long a = 0; a < n_a; ++a) {
          c1[(size_t)a * n_ao + m] = MO_VEC1[m * n_a + a];
          c2[(size_t)a * n_ao + m] = MO_VEC2[m * n_a + a];
        }
      r1 = c1.data();
      r2 = c2.data();
    }
    const int rc = nopt_jk_active_ints(g_ctx, r1, r2, n_a, MO_INTS);
    if (rc) return -rc;
  }
  if (J == nullptr || K == nullptr) return 0;  // src/RI.cpp:230-233 returns before J and K as well
  if (!g_ctx) return -nopt::set_error(NOPT_ESTATE, "calc_2el_MO_INTS_RI: no RI context");
  if (n_ao != g_n_ao) return -nopt::set_error(NOPT_EINVAL, "calc_2el_MO_INTS_RI: n_ao does not match");
  const int rc = nopt_jk_build(g_ctx, DM_D, MO_VEC, n_c, J, K, g_prec);
  return rc ? -rc : 0;
}
