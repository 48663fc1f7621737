// Drop-in replacement for the Fock-build part of NOPT's RI layer
// (include/RI.h:6-23), backed by the GPU context of nopt_jk.h.
//
// The functions keep NOPT's names, argument lists and semantics, with C linkage
// here; the build functions are exported with NOPT's C++ linkage as well
// (include/RI.h:10-16), so an SCF or CASSCF loop written against NOPT's API
// compiles against either header and links to libnoptjk.so:
//   - D is normalised to 1 and MO rows are orbitals; K uses the first n_c rows
//     (DM_to_F_transform_RI, src/RI.cpp:115-176); J and K are overwritten.
//   - DM_to_F_transform_RI returns 1 on success like NOPT's; the others return
//     0. Where NOPT would exit or throw, these return a negative NOPT_E* code
//     and leave the message in nopt_last_error().
//   - One global context, like NOPT's module globals (src/RI.cpp:16-32); calls
//     come from one host thread.
//
// NOPT's gen_RI_AA(molecule*) computes (P|Q) and (mn|P) with libint2 and then
// hands them over with RI_gpu_attach (the integral layer stays on the host);
// standalone callers can use RI_gpu_attach_molecule with the integral engine
// of nopt_ints.h instead.
#ifndef NOPT_RI_H
#define NOPT_RI_H

#ifdef __cplusplus
extern "C" {
#endif

// Replaces the tail of gen_RI_AA (src/RI.cpp:55-72): AUX_V is the Coulomb
// metric (aux_n_ao x aux_n_ao, RI_V_calc) and RI_AA the raw 3-centre tensor in
// NOPT's layout RI_AA[(m*n_ao+n)*aux_n_ao+P] (calc_2el_AO_INTS_RI). The metric
// is factored and folded into B on the device; neither array is kept.
int RI_gpu_attach(const double* AUX_V, const double* RI_AA, long n_ao, long aux_n_ao, int device);

// Standalone set-up: integrals from the host engine of nopt_ints.h (coordinates
// in bohr, Cartesian orbital basis as NOPT reads it when pure == 0).
int RI_gpu_attach_molecule(int natom, const int* Z, const double* xyz_bohr, const char* basis_gbs,
                           const char* aux_gbs, int pure, int device);

// Arithmetic of the half-transform and exchange for the following builds
// (a NOPT_PREC_* value of nopt_jk.h; FP64 by default).
int RI_gpu_set_precision(int prec);

// NOPT's regen_RI_AA (src/RI.cpp:81-101) minus the rebuild: releases the
// device context so that the next RI_gpu_attach forms B for a new geometry.
int regen_RI_AA_release();

// The three build functions below are also exported with C++ linkage under
// the signatures of NOPT's include/RI.h:10-16, which NOPT's own callers link
// to. A translation unit that also includes RI.h (NOPT's src/RI.cpp) defines
// NOPT_RI_SETUP_ONLY before including this header: it then sees only the
// set-up functions above, since a C and a C++ declaration of one function
// cannot coexist.
#ifndef NOPT_RI_SETUP_ONLY
// NOPT's RI_core_realloc (src/RI.cpp:104-113): the device buffers grow on
// demand, so this only checks that a context exists.
int RI_core_realloc(long n_c, long n_ao);

// NOPT's DM_to_F_transform_RI (src/RI.cpp:115-176).
int DM_to_F_transform_RI(double* J, double* K, double* DM, double* MO, long n_c, long n_ao);

// NOPT's calc_2el_MO_INTS_RI (src/RI.cpp:178-292): first the active-space
// integrals MO_INTS[(a2 n_a + a1) n_a^2 + b2 n_a + b1] = sum_Q R_Q[a2][a1]
// R_Q[b2][b1], R_Q[a2][a1] = sum_mn MO_VEC1[a1][m] B_Q,mn MO_VEC2[a2][n]
// (orbitals as rows of n_ao, or as columns of n_a when t != 0; n_a <= 32;
// skipped when MO_INTS is NULL), then, unless J or K is NULL (returns 0 there,
// like NOPT), J from DM_D and K from the first n_c rows of MO_VEC.
int calc_2el_MO_INTS_RI(long n_ao, double* DM_D, double* J, double* K, double* MO_INTS, double* MO_VEC,
                        double* MO_VEC1, double* MO_VEC2, long n_c, long n_a, long t);
#endif  // NOPT_RI_SETUP_ONLY

#ifdef __cplusplus
}
#endif

#endif
