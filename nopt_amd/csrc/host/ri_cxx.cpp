// C++-linkage forwarders with the exact signatures of NOPT's include/RI.h:10-16,
// so that NOPT's own callers (molecule::calc_F_AO, src/molecule.cpp:4423-4425;
// CASSCF, src/CAS.cpp:461, :702; the SAD guess, src/sad_guess.cpp:200), which
// include RI.h and call these functions by their C++ (mangled) names, link
// against libnoptjk.so without any edit to RI.h. Each one calls the C-linkage
// entry point of include/nopt_ri.h.
//
// This file must not include nopt_ri.h: a C++ function and a C-linkage function
// of the same name and parameters cannot be declared in one namespace, so the
// C entry points are declared here inside a namespace of their own (C linkage
// names one function whatever the namespace of its declaration).

namespace nopt_ri_c {
extern "C" {
int RI_core_realloc(long n_c, long n_ao);
int DM_to_F_transform_RI(double* J, double* K, double* DM, double* MO, long n_c, long n_ao);
int calc_2el_MO_INTS_RI(long n_ao, double* DM_D, double* J, double* K, double* MO_INTS, double* MO_VEC,
                        double* MO_VEC1, double* MO_VEC2, long n_c, long n_a, long t);
}
}  // namespace nopt_ri_c

// include/RI.h:10
int RI_core_realloc(long n_c, long n_ao) { return nopt_ri_c::RI_core_realloc(n_c, n_ao); }

// include/RI.h:12
int DM_to_F_transform_RI(double* J, double* K, double* DM, double* MO, long n_c, long n_ao) {
  return nopt_ri_c::DM_to_F_transform_RI(J, K, DM, MO, n_c, n_ao);
}

// include/RI.h:14-16
int calc_2el_MO_INTS_RI(long n_ao, double* DM_D, double* J, double* K, double* MO_INTS, double* MO_VEC,
                        double* MO_VEC1, double* MO_VEC2, long n_c, long n_a, long t) {
  return nopt_ri_c::calc_2el_MO_INTS_RI(n_ao, DM_D, J, K, MO_INTS, MO_VEC, MO_VEC1, MO_VEC2, n_c, n_a, t);
}
