// C entry points of include/nopt_ints.h: molecule set-up and host integrals.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../../include/nopt_ints.h"
#include "../../../include/nopt_jk.h"
#include "basis.hpp"
#include "errors.hpp"
#include "ints.hpp"

struct nopt_mol {
  int natom = 0;
  std::vector<int> Z;
  std::vector<double> xyz;
  std::vector<nopt::Shell> shells, aux;
  long nbf = 0, naux = 0, nelec = 0;
  double enuc = 0.0;
};

using nopt::set_error;

extern "C" {

void nopt_set_host_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int nopt_mol_create(int natom, const int* Z, const double* xyz, const char* basis_gbs,
                    const char* aux_gbs, int orbital_pure, nopt_mol** out) {
  nopt::clear_error();
  if (!out || natom <= 0 || !Z || !xyz || !basis_gbs || !aux_gbs)
    return set_error(NOPT_EINVAL, "nopt_mol_create: null argument or natom <= 0");
  *out = nullptr;
  try {
    std::vector<std::pair<std::string, nopt::ElementShells>> lib, alib;
    std::string e = nopt::read_gbs(basis_gbs, &lib);
    if (e.empty()) e = nopt::read_gbs(aux_gbs, &alib);
    if (!e.empty()) return set_error(NOPT_EIO, e);
    auto* m = new nopt_mol;
    m->natom = natom;
    m->Z.assign(Z, Z + natom);
    m->xyz.assign(xyz, xyz + 3 * natom);
    e = nopt::build_shells(lib, natom, Z, xyz, orbital_pure != 0, &m->shells);
    if (e.empty()) e = nopt::build_shells(alib, natom, Z, xyz, true, &m->aux);
    if (!e.empty()) {
      delete m;
      return set_error(NOPT_EINVAL, e);
    }
    for (auto& s : m->shells) m->nbf += s.size();
    for (auto& s : m->aux) m->naux += s.size();
    for (int a = 0; a < natom; ++a) m->nelec += Z[a];
    // nuclear repulsion as E_nuc_calc (src/molecule.cpp:2236-2249)
    for (int i = 0; i < natom; ++i)
      for (int j = i + 1; j < natom; ++j) {
        double d2 = 0;
        for (int k = 0; k < 3; ++k) d2 += (xyz[3 * i + k] - xyz[3 * j + k]) * (xyz[3 * i + k] - xyz[3 * j + k]);
        m->enuc += Z[i] * Z[j] / std::sqrt(d2);
      }
    *out = m;
    return NOPT_OK;
  } catch (const std::bad_alloc&) {
    return set_error(NOPT_ENOMEM, "nopt_mol_create: host allocation failed");
  }
}

void nopt_mol_destroy(nopt_mol* mol) { delete mol; }

int nopt_mol_dims(const nopt_mol* m, long* nbf, long* naux, long* nelec, double* enuc) {
  if (!m) return set_error(NOPT_EINVAL, "nopt_mol_dims: null molecule");
  if (nbf) *nbf = m->nbf;
  if (naux) *naux = m->naux;
  if (nelec) *nelec = m->nelec;
  if (enuc) *enuc = m->enuc;
  return NOPT_OK;
}

int nopt_mol_function_l(const nopt_mol* m, int* l_out) {
  if (!m || !l_out) return set_error(NOPT_EINVAL, "nopt_mol_function_l: null argument");
  long k = 0;
  for (auto& s : m->shells)
    for (int i = 0; i < s.size(); ++i) l_out[k++] = s.l;
  return NOPT_OK;
}

int nopt_mol_one_electron(const nopt_mol* m, double* S, double* T, double* V) {
  if (!m) return set_error(NOPT_EINVAL, "nopt_mol_one_electron: null molecule");
  nopt::one_electron(m->shells, m->natom, m->Z.data(), m->xyz.data(), S, T, V);
  return NOPT_OK;
}

int nopt_mol_metric(const nopt_mol* m, double* V) {
  if (!m || !V) return set_error(NOPT_EINVAL, "nopt_mol_metric: null argument");
  nopt::eri2c(m->aux, V);
  return NOPT_OK;
}

int nopt_mol_eri3(const nopt_mol* m, double* out, int layout) {
  if (!m || !out) return set_error(NOPT_EINVAL, "nopt_mol_eri3: null argument");
  if (layout != 0 && layout != 1) return set_error(NOPT_EINVAL, "nopt_mol_eri3: layout must be 0 or 1");
  size_t count = layout == 0 ? (size_t)m->naux * (m->nbf * (m->nbf + 1) / 2)
                             : (size_t)m->naux * m->nbf * m->nbf;
  std::memset(out, 0, count * sizeof(double));
  nopt::eri3c(m->shells, m->aux, out, layout == 1);
  return NOPT_OK;
}

}  // extern "C"
