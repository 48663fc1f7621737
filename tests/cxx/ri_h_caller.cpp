// A caller written against NOPT's RI interface: the declarations below carry
// the names, parameter lists and C++ linkage of include/RI.h:10-16, so this
// file links against libnoptjk.so the way NOPT's own objects would
// (molecule::calc_F_AO, src/molecule.cpp:4423-4425; CASSCF, src/CAS.cpp:702).
//
// Without an attached RI context every call must report that (a negative
// status) instead of building; exit code 0 when all three do.
int RI_core_realloc(long n_c, long n_ao);
int DM_to_F_transform_RI(double* J, double* K, double* DM, double* MO, long n_c, long n_ao);
int calc_2el_MO_INTS_RI(long n_ao, double* DM_D, double* J, double* K, double* MO_INTS, double* MO_VEC,
                        double* MO_VEC1, double* MO_VEC2, long n_c, long n_a, long t);

int main() {
  double J[4] = {0}, K[4] = {0}, D[4] = {0}, C[4] = {0};
  const int a = RI_core_realloc(1, 2);
  const int b = DM_to_F_transform_RI(J, K, D, C, 1, 2);
  const int c = calc_2el_MO_INTS_RI(2, D, J, K, nullptr, C, nullptr, nullptr, 1, 0, 0);
  return (a < 0 && b < 0 && c < 0) ? 0 : 1;
}
