// Gaussian basis sets for the RI-J/K build: Psi4 .gbs reader, shell
// normalisation, Cartesian component order and real solid harmonics.
//
// Conventions kept from NOPT (libint2 shells):
//  * the orbital basis may be Cartesian (NOPT reads it with pure=false,
//    src/inp_par_read.cpp:2550-2552) or spherical; the auxiliary basis is
//    always spherical (pure=true, src/inp_par_read.cpp:2583-2585);
//  * Cartesian components of a shell of angular momentum l are ordered
//    x^i y^j z^k with i running l..0 and, inside, j running l-i..0 (the
//    libint2/CCA order), and every Cartesian component carries the
//    normalisation of x^l;
//  * spherical components are ordered m = -l..l.
// The .gbs parsing follows the Psi4 layout that NOPT's reader accepts
// (src/basis_lib_read.cpp:57-247): element blocks delimited by "****",
// shell lines "L nprim scale", then "exponent coefficient" lines.
#pragma once
#include <string>
#include <vector>

namespace nopt {

struct Shell {
  int l = 0;
  bool pure = false;
  int atom = 0;
  double center[3] = {0, 0, 0};
  std::vector<double> alpha;  // primitive exponents
  std::vector<double> coef;   // contraction coefficients with primitive and
                              // contraction normalisation folded in (x^l)
  int ncart() const { return (l + 1) * (l + 2) / 2; }
  int nsph() const { return 2 * l + 1; }
  int size() const { return pure ? nsph() : ncart(); }
};

struct ElementShells {
  struct Raw {
    int l;
    std::vector<double> alpha, coef;
  };
  std::vector<Raw> shells;
};

// Parses every element block of a .gbs file. Returns an empty string on
// success, otherwise a message.
std::string read_gbs(const std::string& path,
                     std::vector<std::pair<std::string, ElementShells>>* out);

// Element symbol for nuclear charge Z (1..18).
const char* element_symbol(int Z);

// Builds the shells of a molecule (atoms in bohr) from a parsed basis.
std::string build_shells(const std::vector<std::pair<std::string, ElementShells>>& lib,
                         int natom, const int* Z, const double* xyz, bool pure,
                         std::vector<Shell>* out);

// Cartesian exponents of component c of shell l in the libint2 order.
void cart_exponents(int l, int c, int* ix, int* iy, int* iz);
int cart_index(int l, int ix, int iy, int iz);

// Real solid harmonic S_lm as a sparse polynomial over Cartesian monomials of
// degree l: S_lm = sum_k coef_k x^ix y^iy z^iz. Normalised so that
// S_lm exp(-a r^2) carries the same norm as x^l exp(-a r^2).
struct SolidTerm {
  int ix, iy, iz;
  double c;
};
const std::vector<SolidTerm>& solid_harmonic(int l, int m);

// Dense Cartesian -> spherical transform for shell l: row m (0..2l) holds
// the coefficients over the ncart components.
std::vector<double> cart2sph_matrix(int l);

double double_factorial(int n);  // n!! with (-1)!! = 1

}  // namespace nopt
