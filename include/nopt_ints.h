/* Molecule set-up and host integrals for B-tensor formation (FP64).
 *
 * These entry points stand in for the libint2-backed integral layer that
 * NOPT's gen_RI_AA (src/RI.cpp:34-79) drives:
 *   nopt_mol_create        <- molecule_read_by_inp_par + basis_lib_read_gbs
 *                             (src/inp_par_read.cpp:2472-2600,
 *                              src/basis_lib_read.cpp:57-247)
 *   nopt_mol_one_electron  <- molecule::gen_1el_data S, T, V
 *                             (src/molecule.cpp:3757-3803)
 *   nopt_mol_metric        <- RI_V_calc (src/libint_link.cpp:2049-2107)
 *   nopt_mol_eri3          <- calc_2el_AO_INTS_RI (src/libint_link.cpp:2261-2350)
 * All return 0 on success or a nonzero NOPT_E* code (nopt_jk.h); the message
 * is in nopt_last_error(). Coordinates are in bohr.
 */
#ifndef NOPT_INTS_H
#define NOPT_INTS_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct nopt_mol nopt_mol;

/* orbital_pure = 1: spherical orbital basis (nbf counts 2l+1 per shell);
 * orbital_pure = 0: Cartesian, as NOPT reads it. The auxiliary basis is
 * always spherical. */
int nopt_mol_create(int natom, const int* Z, const double* xyz_bohr, const char* basis_gbs,
                    const char* aux_gbs, int orbital_pure, nopt_mol** out);
void nopt_mol_destroy(nopt_mol* mol);
int nopt_mol_dims(const nopt_mol* mol, long* nbf, long* naux, long* nelec, double* e_nuc);
/* angular momentum of each orbital basis function (nbf entries) */
int nopt_mol_function_l(const nopt_mol* mol, int* l_out);
/* S, T, V: nbf x nbf row-major; any may be NULL. */
int nopt_mol_one_electron(const nopt_mol* mol, double* S, double* T, double* V);
/* (P|Q): naux x naux row-major. */
int nopt_mol_metric(const nopt_mol* mol, double* V);
/* layout 0: packed out[P*npair + m(m+1)/2 + n], m >= n, npair = nbf(nbf+1)/2;
 * layout 1: NOPT RI_AA layout out[(m*nbf+n)*naux + P] (src/RI.cpp:73). */
int nopt_mol_eri3(const nopt_mol* mol, double* out, int layout);
/* OpenMP threads used by the host integral engine (0 = runtime default). */
void nopt_set_host_threads(int n);
const char* nopt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
