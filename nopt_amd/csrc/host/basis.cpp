// Basis-set reader and shell algebra; see basis.hpp for the conventions.
#include "basis.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace nopt {

double double_factorial(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

static const char* kElements[] = {"",   "H",  "He", "Li", "Be", "B", "C",
                                  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al",
                                  "Si", "P",  "S",  "Cl", "Ar"};

const char* element_symbol(int Z) {
  if (Z < 1 || Z > 18) return nullptr;
  return kElements[Z];
}

static int am_from_letter(char c) {
  switch (std::toupper(c)) {
    case 'S': return 0;
    case 'P': return 1;
    case 'D': return 2;
    case 'F': return 3;
    case 'G': return 4;
    case 'H': return 5;
    case 'I': return 6;
    default: return -1;
  }
}

static double parse_fortran_double(std::string s) {
  for (auto& ch : s)
    if (ch == 'D' || ch == 'd') ch = 'E';
  return std::strtod(s.c_str(), nullptr);
}

std::string read_gbs(const std::string& path,
                     std::vector<std::pair<std::string, ElementShells>>* out) {
  std::ifstream in(path);
  if (!in) return "cannot open basis file " + path;
  std::vector<std::string> lines;
  for (std::string ln; std::getline(in, ln);) {
    auto p = ln.find('!');
    if (p != std::string::npos) ln = ln.substr(0, p);
    // trim
    size_t a = ln.find_first_not_of(" \t\r");
    if (a == std::string::npos) continue;
    size_t b = ln.find_last_not_of(" \t\r");
    lines.push_back(ln.substr(a, b - a + 1));
  }
  out->clear();
  size_t i = 0;
  while (i < lines.size()) {
    if (lines[i] == "****" || lines[i] == "spherical" || lines[i] == "cartesian") {
      ++i;
      continue;
    }
    // element header: "<Sym> 0"
    std::istringstream hs(lines[i]);
    std::string sym;
    int zero = -1;
    hs >> sym >> zero;
    if (zero != 0) return "malformed element header in " + path + ": " + lines[i];
    ElementShells el;
    ++i;
    while (i < lines.size() && lines[i] != "****") {
      std::istringstream ss(lines[i]);
      std::string am;
      int nprim = 0;
      double scale = 1.0;
      ss >> am >> nprim >> scale;
      if (nprim <= 0) return "malformed shell line in " + path + ": " + lines[i];
      bool sp = (am.size() == 2 && std::toupper(am[0]) == 'S' && std::toupper(am[1]) == 'P');
      ElementShells::Raw s1, s2;
      s1.l = sp ? 0 : am_from_letter(am[0]);
      s2.l = 1;
      if (s1.l < 0 || (!sp && am.size() != 1)) return "unknown shell type " + am;
      ++i;
      for (int k = 0; k < nprim; ++k, ++i) {
        if (i >= lines.size()) return "truncated shell in " + path;
        std::istringstream ps(lines[i]);
        std::string ea, ca, cb;
        ps >> ea >> ca >> cb;
        double e = parse_fortran_double(ea) * scale * scale;
        s1.alpha.push_back(e);
        s1.coef.push_back(parse_fortran_double(ca));
        if (sp) {
          s2.alpha.push_back(e);
          s2.coef.push_back(parse_fortran_double(cb));
        }
      }
      el.shells.push_back(s1);
      if (sp) el.shells.push_back(s2);
    }
    out->push_back({sym, el});
    ++i;
  }
  return "";
}

static void normalise(Shell& s) {
  const int l = s.l;
  const double dfl = double_factorial(2 * l - 1);
  for (size_t k = 0; k < s.alpha.size(); ++k) {
    double a = s.alpha[k];
    s.coef[k] *= std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * l) / std::sqrt(dfl);
  }
  double ov = 0.0;
  for (size_t i = 0; i < s.alpha.size(); ++i)
    for (size_t j = 0; j < s.alpha.size(); ++j) {
      double p = s.alpha[i] + s.alpha[j];
      ov += s.coef[i] * s.coef[j] * std::pow(M_PI / p, 1.5) * dfl / std::pow(2.0 * p, l);
    }
  double f = 1.0 / std::sqrt(ov);
  for (auto& c : s.coef) c *= f;
}

std::string build_shells(const std::vector<std::pair<std::string, ElementShells>>& lib,
                         int natom, const int* Z, const double* xyz, bool pure,
                         std::vector<Shell>* out) {
  out->clear();
  for (int a = 0; a < natom; ++a) {
    const char* sym = element_symbol(Z[a]);
    if (!sym) return "unsupported nuclear charge";
    const ElementShells* el = nullptr;
    for (auto& e : lib)
      if (strcasecmp(e.first.c_str(), sym) == 0) el = &e.second;
    if (!el) return std::string("element ") + sym + " missing from basis";
    for (auto& r : el->shells) {
      Shell s;
      s.l = r.l;
      s.pure = pure;
      s.atom = a;
      for (int d = 0; d < 3; ++d) s.center[d] = xyz[3 * a + d];
      s.alpha = r.alpha;
      s.coef = r.coef;
      normalise(s);
      out->push_back(s);
    }
  }
  return "";
}

void cart_exponents(int l, int c, int* ix, int* iy, int* iz) {
  int n = 0;
  for (int i = l; i >= 0; --i)
    for (int j = l - i; j >= 0; --j, ++n)
      if (n == c) {
        *ix = i;
        *iy = j;
        *iz = l - i - j;
        return;
      }
}

int cart_index(int l, int ix, int iy, int iz) {
  (void)iz;
  // position of (ix,iy) in the l..0 / (l-i)..0 order
  int i = l - ix;  // number of complete x-blocks before
  return i * (i + 1) / 2 + (l - ix - iy);
}

static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// Real solid harmonics from the closed form of Helgaker, Jorgensen & Olsen,
// "Molecular Electronic-Structure Theory", eqs. 6.4.47-6.4.50.
static std::vector<SolidTerm> make_solid(int l, int m) {
  const int am = std::abs(m);
  const double N = 1.0 / (std::pow(2.0, am) * fact(l)) *
                   std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
  std::map<int, double> acc;  // key ix*64+iy
  const int vm2 = (m < 0) ? 1 : 0;  // 2*v_m
  for (int t = 0; t <= (l - am) / 2; ++t)
    for (int u = 0; u <= t; ++u)
      for (int k = vm2; k <= am; k += 2) {  // k = 2v
        // v - v_m = (k - vm2)/2
        int sgn_exp = t + (k - vm2) / 2;
        double c = ((sgn_exp & 1) ? -1.0 : 1.0) * std::pow(0.25, t) * binom(l, t) *
                   binom(l - t, am + t) * binom(t, u) * binom(am, k);
        int ix = 2 * t + am - 2 * u - k;
        int iy = 2 * u + k;
        int iz = l - 2 * t - am;
        if (ix < 0 || iz < 0) continue;
        acc[ix * 64 + iy] += N * c;
      }
  std::vector<SolidTerm> out;
  for (auto& kv : acc) {
    if (std::fabs(kv.second) < 1e-15) continue;
    int ix = kv.first / 64, iy = kv.first % 64;
    out.push_back({ix, iy, l - ix - iy, kv.second});
  }
  // Rescale so the norm equals that of x^l: Gaussian moments of equal total
  // degree differ only by the double factorials of the exponent sums.
  double nrm = 0.0;
  for (auto& a : out)
    for (auto& b : out) {
      int sx = a.ix + b.ix, sy = a.iy + b.iy, sz = a.iz + b.iz;
      if ((sx | sy | sz) & 1) continue;
      nrm += a.c * b.c * double_factorial(sx - 1) * double_factorial(sy - 1) *
             double_factorial(sz - 1);
    }
  double f = std::sqrt(double_factorial(2 * l - 1) / nrm);
  for (auto& a : out) a.c *= f;
  return out;
}

const std::vector<SolidTerm>& solid_harmonic(int l, int m) {
  static std::mutex mu;
  static std::map<int, std::vector<SolidTerm>> cache;
  std::lock_guard<std::mutex> g(mu);
  int key = l * 64 + (m + 32);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  return cache[key] = make_solid(l, m);
}

std::vector<double> cart2sph_matrix(int l) {
  const int nc = (l + 1) * (l + 2) / 2;
  std::vector<double> T((2 * l + 1) * nc, 0.0);
  for (int m = -l; m <= l; ++m)
    for (auto& t : solid_harmonic(l, m)) T[(m + l) * nc + cart_index(l, t.ix, t.iy, t.iz)] = t.c;
  return T;
}

}  // namespace nopt
