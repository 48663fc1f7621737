// Host integral engine for B-tensor formation: overlap, kinetic, nuclear
// attraction, the 2-centre Coulomb metric (P|Q) and the 3-centre Coulomb
// integrals (P|mn), all in FP64.
//
// These replace the libint2 calls of NOPT's B build: RI_V_calc
// (src/libint_link.cpp:2049-2107) for (P|Q) and calc_2el_AO_INTS_RI
// (src/libint_link.cpp:2261-2350) for (mn|P), plus the one-electron matrices
// of molecule::gen_1el_data (src/molecule.cpp:3757-3803).
//
// Method: McMurchie-Davidson. A product of two Cartesian Gaussians is expanded
// in Hermite Gaussians (coefficients E^{ij}_t); a spherical auxiliary function
// is written directly as a Hermite derivative through Hobson's theorem,
//   S_lm(r_C) exp(-g r_C^2) = (2g)^-l S_lm(d/dC) exp(-g r_C^2),
// so only the Hermite terms of total degree l appear on the auxiliary side.
// The Coulomb kernels are the Hermite integrals R_{tuv} built from the Boys
// function F_n(T) by the usual recursion.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "basis.hpp"
#include "ints.hpp"

namespace nopt {

// ---------------------------------------------------------------- Boys -----
// F_n(T) on a grid T_k = k*h, h = 1/20, up to T = 36; off-grid values by an
// 8-term Taylor expansion (F_n' = -F_{n+1}), error below 1e-17 relative.
// For T >= 36 the asymptotic F_0 = sqrt(pi/T)/2 (erfc(6) ~ 2e-17) and upward
// recursion, which is stable for T > n.
namespace {
constexpr int kBoysNMax = 24;        // highest order served
constexpr int kBoysTaylor = 8;       // Taylor terms
constexpr double kBoysH = 1.0 / 20;  // grid step
constexpr double kBoysTMax = 36.0;
constexpr int kBoysGrid = 721;  // 36 / h + 1

struct BoysTable {
  std::vector<double> f;  // [k][n], n < kBoysNMax + kBoysTaylor
  int ncol = kBoysNMax + kBoysTaylor;
  BoysTable() {
    f.resize((size_t)kBoysGrid * ncol);
    for (int k = 0; k < kBoysGrid; ++k) {
      long double T = (long double)k * kBoysH;
      int N = ncol - 1;
      // series F_N(T) = e^-T sum_j (2T)^j / ((2N+1)(2N+3)...(2N+2j+1))
      long double term = 1.0L / (2 * N + 1), sum = term;
      for (int j = 1; j < 2000; ++j) {
        term *= 2.0L * T / (2 * N + 2 * j + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
      }
      long double eT = expl(-T);
      long double Fn = eT * sum;
      f[(size_t)k * ncol + N] = (double)Fn;
      for (int n = N - 1; n >= 0; --n) {
        Fn = (2.0L * T * Fn + eT) / (2 * n + 1);
        f[(size_t)k * ncol + n] = (double)Fn;
      }
    }
  }
};
const BoysTable& boys_table() {
  static BoysTable t;
  return t;
}
}  // namespace

void boys(int nmax, double T, double* F) {
  if (T >= kBoysTMax) {
    double eT = std::exp(-T);
    F[0] = 0.5 * std::sqrt(M_PI / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
    return;
  }
  const BoysTable& tb = boys_table();
  int k = (int)(T / kBoysH + 0.5);
  double d = k * kBoysH - T;  // F(T) = sum_j F^{(j)}(T_k) (T-T_k)^j/j!, F^{(j)}_n = (-1)^j F_{n+j}
  const double* row = &tb.f[(size_t)k * tb.ncol];
  // top order by Taylor, lower orders by downward recursion
  double v = 0.0, fac = 1.0;
  for (int j = 0; j < kBoysTaylor; ++j) {
    v += row[nmax + j] * fac;
    fac *= d / (j + 1);
  }
  F[nmax] = v;
  double eT = std::exp(-T);
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + eT) / (2 * n + 1);
}

// ------------------------------------------------------- Hermite tables -----
// E[(i*(lb+1)+j)*(la+lb+1)+t] for one Cartesian direction.
static void hermite_E(int la, int lb, double p, double XPA, double XPB, double* E) {
  const int L = la + lb + 1;
  auto at = [&](int i, int j, int t) -> double& { return E[(i * (lb + 1) + j) * L + t]; };
  std::fill(E, E + (la + 1) * (lb + 1) * L, 0.0);
  const double o2p = 0.5 / p;
  at(0, 0, 0) = 1.0;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = XPA * (t <= i ? at(i, 0, t) : 0.0);
      if (t > 0) v += o2p * at(i, 0, t - 1);
      if (t + 1 <= i) v += (t + 1) * at(i, 0, t + 1);
      at(i + 1, 0, t) = v;
    }
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j < lb; ++j)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = XPB * (t <= i + j ? at(i, j, t) : 0.0);
        if (t > 0) v += o2p * at(i, j, t - 1);
        if (t + 1 <= i + j) v += (t + 1) * at(i, j, t + 1);
        at(i, j + 1, t) = v;
      }
}

// R^0_{tuv}(a, X, Y, Z) for t+u+v <= L in a box of side L+1.
struct RWork {
  std::vector<double> a, b;
  double F[64];
};

static const double* hermite_R(int L, double alpha, double X, double Y, double Z, RWork& w) {
  const int S = L + 1;
  const size_t box = (size_t)S * S * S;
  if (w.a.size() < box) {
    w.a.resize(box);
    w.b.resize(box);
  }
  double* prev = w.a.data();
  double* cur = w.b.data();
  boys(L, alpha * (X * X + Y * Y + Z * Z), w.F);
  const double m2a = -2.0 * alpha;
  double pw = std::pow(m2a, L);
  for (int n = L; n >= 0; --n) {
    const int top = L - n;
    cur[0] = pw * w.F[n];
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          if (t + u + v == 0) continue;
          double r;
          if (t > 0) {
            r = X * prev[((t - 1) * S + u) * S + v];
            if (t > 1) r += (t - 1) * prev[((t - 2) * S + u) * S + v];
          } else if (u > 0) {
            r = Y * prev[(t * S + u - 1) * S + v];
            if (u > 1) r += (u - 1) * prev[(t * S + u - 2) * S + v];
          } else {
            r = Z * prev[(t * S + u) * S + v - 1];
            if (v > 1) r += (v - 1) * prev[(t * S + u) * S + v - 2];
          }
          cur[(t * S + u) * S + v] = r;
        }
    std::swap(prev, cur);
    pw /= m2a;
  }
  return prev;
}

// ------------------------------------------------------ shell utilities -----
static void offsets(const std::vector<Shell>& sh, std::vector<long>* off, long* n) {
  off->resize(sh.size());
  long k = 0;
  for (size_t i = 0; i < sh.size(); ++i) {
    (*off)[i] = k;
    k += sh[i].size();
  }
  *n = k;
}

// Cached Cartesian -> spherical matrices, l <= 7.
static const std::vector<double>& c2s(int l) {
  static std::vector<std::vector<double>> tab = [] {
    std::vector<std::vector<double>> t;
    for (int k = 0; k <= 7; ++k) t.push_back(cart2sph_matrix(k));
    return t;
  }();
  return tab[l];
}

// Cartesian block (ncart_a x ncart_b x nm) -> output block (size_a x size_b x
// nm), transforming the a and b indices of pure shells.
static void transform_block(const Shell& a, const Shell& b, const double* cart, double* out,
                            int nm, std::vector<double>& tmp) {
  const int na = a.ncart(), nb = b.ncart();
  const double* src = cart;
  int ra = na;
  if (a.pure) {
    const auto& Ta = c2s(a.l);
    tmp.assign((size_t)a.nsph() * nb * nm, 0.0);
    for (int m = 0; m < a.nsph(); ++m)
      for (int c = 0; c < na; ++c) {
        double t = Ta[m * na + c];
        if (t == 0.0) continue;
        for (int j = 0; j < nb * nm; ++j) tmp[(size_t)m * nb * nm + j] += t * cart[(size_t)c * nb * nm + j];
      }
    src = tmp.data();
    ra = a.nsph();
  }
  if (b.pure) {
    const auto& Tb = c2s(b.l);
    const int sb = b.nsph();
    for (int i = 0; i < ra; ++i)
      for (int m = 0; m < sb; ++m) {
        double* o = out + ((size_t)i * sb + m) * nm;
        for (int k = 0; k < nm; ++k) o[k] = 0.0;
        for (int c = 0; c < nb; ++c) {
          double t = Tb[m * nb + c];
          if (t == 0.0) continue;
          const double* s = src + ((size_t)i * nb + c) * nm;
          for (int k = 0; k < nm; ++k) o[k] += t * s[k];
        }
      }
  } else {
    std::memcpy(out, src, sizeof(double) * ra * nb * nm);
  }
}

// ------------------------------------------------------- 1-e integrals -----
void one_electron(const std::vector<Shell>& sh, int natom, const int* Z, const double* xyz,
                  double* S, double* T, double* V) {
  std::vector<long> off;
  long n;
  offsets(sh, &off, &n);
  const int ns = (int)sh.size();
#pragma omp parallel
  {
    RWork rw;
    std::vector<double> Ex, Ey, Ez, bs, bt, bv, o1, o2, o3;
#pragma omp for schedule(dynamic)
    for (int ij = 0; ij < ns * ns; ++ij) {
      int ia = ij / ns, ib = ij % ns;
      if (ib > ia) continue;
      const Shell& A = sh[ia];
      const Shell& B = sh[ib];
      const int la = A.l, lb = B.l, na = A.ncart(), nb = B.ncart();
      const int lb2 = lb + 2, L2 = la + lb2 + 1;
      bs.assign(na * nb, 0.0);
      bt.assign(na * nb, 0.0);
      bv.assign(na * nb, 0.0);
      Ex.resize((la + 1) * (lb2 + 1) * L2);
      Ey.resize(Ex.size());
      Ez.resize(Ex.size());
      double AB[3];
      for (int d = 0; d < 3; ++d) AB[d] = A.center[d] - B.center[d];
      const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
      for (size_t pa = 0; pa < A.alpha.size(); ++pa)
        for (size_t pb = 0; pb < B.alpha.size(); ++pb) {
          const double a = A.alpha[pa], b = B.alpha[pb], p = a + b, mu = a * b / p;
          const double K = A.coef[pa] * B.coef[pb] * std::exp(-mu * rab2);
          double P[3];
          for (int d = 0; d < 3; ++d) P[d] = (a * A.center[d] + b * B.center[d]) / p;
          hermite_E(la, lb2, p, P[0] - A.center[0], P[0] - B.center[0], Ex.data());
          hermite_E(la, lb2, p, P[1] - A.center[1], P[1] - B.center[1], Ey.data());
          hermite_E(la, lb2, p, P[2] - A.center[2], P[2] - B.center[2], Ez.data());
          auto e = [&](const std::vector<double>& E, int i, int j, int t) {
            return E[(i * (lb2 + 1) + j) * L2 + t];
          };
          const double sp = std::pow(M_PI / p, 1.5);
          // kinetic 1-D factor: -1/2 d^2/dx^2 acting on the ket
          auto kin = [&](const std::vector<double>& E, int i, int j) {
            double v = -2.0 * b * b * e(E, i, j + 2, 0) + b * (2 * j + 1) * e(E, i, j, 0);
            if (j >= 2) v -= 0.5 * j * (j - 1) * e(E, i, j - 2, 0);
            return v;
          };
          // nuclear attraction: R for each nucleus
          const int Lab = la + lb;
          std::vector<double> Rn;
          for (int c = 0; c < natom; ++c) {
            const double* R = hermite_R(Lab, p, P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1],
                                        P[2] - xyz[3 * c + 2], rw);
            const int S3 = Lab + 1;
            const double f = -Z[c] * 2.0 * M_PI / p * K;
            for (int ca = 0; ca < na; ++ca) {
              int ax, ay, az;
              cart_exponents(la, ca, &ax, &ay, &az);
              for (int cb = 0; cb < nb; ++cb) {
                int bx, by, bz;
                cart_exponents(lb, cb, &bx, &by, &bz);
                double s = 0.0;
                for (int t = 0; t <= ax + bx; ++t)
                  for (int u = 0; u <= ay + by; ++u)
                    for (int v = 0; v <= az + bz; ++v)
                      s += e(Ex, ax, bx, t) * e(Ey, ay, by, u) * e(Ez, az, bz, v) *
                           R[(t * S3 + u) * S3 + v];
                bv[ca * nb + cb] += f * s;
              }
            }
          }
          for (int ca = 0; ca < na; ++ca) {
            int ax, ay, az;
            cart_exponents(la, ca, &ax, &ay, &az);
            for (int cb = 0; cb < nb; ++cb) {
              int bx, by, bz;
              cart_exponents(lb, cb, &bx, &by, &bz);
              double sx = e(Ex, ax, bx, 0), sy = e(Ey, ay, by, 0), sz = e(Ez, az, bz, 0);
              bs[ca * nb + cb] += K * sp * sx * sy * sz;
              bt[ca * nb + cb] += K * sp *
                                  (kin(Ex, ax, bx) * sy * sz + sx * kin(Ey, ay, by) * sz +
                                   sx * sy * kin(Ez, az, bz));
            }
          }
        }
      const int sa = A.size(), sb = B.size();
      o1.resize(sa * sb);
      o2.resize(sa * sb);
      o3.resize(sa * sb);
      std::vector<double> tb;
      transform_block(A, B, bs.data(), o1.data(), 1, tb);
      transform_block(A, B, bt.data(), o2.data(), 1, tb);
      transform_block(A, B, bv.data(), o3.data(), 1, tb);
      for (int i = 0; i < sa; ++i)
        for (int j = 0; j < sb; ++j) {
          long r = off[ia] + i, c = off[ib] + j;
          if (S) S[r * n + c] = S[c * n + r] = o1[i * sb + j];
          if (T) T[r * n + c] = T[c * n + r] = o2[i * sb + j];
          if (V) V[r * n + c] = V[c * n + r] = o3[i * sb + j];
        }
    }
  }
}

// ---------------------------------------------------- auxiliary helpers -----
// Spherical auxiliary shell as Hermite terms: for each m, the (tuv, s) list.
struct AuxHerm {
  int l;
  std::vector<std::vector<SolidTerm>> m;  // size 2l+1
};

static AuxHerm aux_herm(int l) {
  AuxHerm h;
  h.l = l;
  for (int m = -l; m <= l; ++m) h.m.push_back(solid_harmonic(l, m));
  return h;
}

// 2-centre metric (P|Q), naux x naux, full square.
void eri2c(const std::vector<Shell>& aux, double* V) {
  std::vector<long> off;
  long n;
  offsets(aux, &off, &n);
  const int ns = (int)aux.size();
  const double pi52 = 2.0 * std::pow(M_PI, 2.5);
#pragma omp parallel
  {
    RWork rw;
    std::vector<double> buf;
#pragma omp for schedule(dynamic)
    for (int ij = 0; ij < ns * ns; ++ij) {
      int ia = ij / ns, ic = ij % ns;
      if (ic > ia) continue;
      const Shell& A = aux[ia];
      const Shell& C = aux[ic];
      if (!A.pure || !C.pure) continue;  // auxiliary shells are always spherical
      AuxHerm ha = aux_herm(A.l), hc = aux_herm(C.l);
      const int L = A.l + C.l, S = L + 1;
      const int sa = A.nsph(), sc = C.nsph();
      buf.assign(sa * sc, 0.0);
      for (size_t pa = 0; pa < A.alpha.size(); ++pa)
        for (size_t pc = 0; pc < C.alpha.size(); ++pc) {
          const double a = A.alpha[pa], g = C.alpha[pc], ar = a * g / (a + g);
          const double* R = hermite_R(L, ar, A.center[0] - C.center[0],
                                      A.center[1] - C.center[1], A.center[2] - C.center[2], rw);
          double pref = pi52 / (a * g * std::sqrt(a + g)) * A.coef[pa] * C.coef[pc] *
                        std::pow(2.0 * a, -A.l) * std::pow(2.0 * g, -C.l) *
                        ((C.l & 1) ? -1.0 : 1.0);
          for (int ma = 0; ma < sa; ++ma)
            for (int mc = 0; mc < sc; ++mc) {
              double s = 0.0;
              for (auto& ta : ha.m[ma])
                for (auto& tc : hc.m[mc])
                  s += ta.c * tc.c * R[((ta.ix + tc.ix) * S + ta.iy + tc.iy) * S + ta.iz + tc.iz];
              buf[ma * sc + mc] += pref * s;
            }
        }
      for (int i = 0; i < sa; ++i)
        for (int j = 0; j < sc; ++j) {
          long r = off[ia] + i, c = off[ic] + j;
          V[r * n + c] = V[c * n + r] = buf[i * sc + j];
        }
    }
  }
}

// 3-centre integrals (P|mn) for m >= n, written as out[P*npair + m(m+1)/2 + n]
// (packed layout) or, with full_nopt set, NOPT's RI_AA layout
// out[(m*nbf+n)*naux + P] with both triangles filled (src/RI.cpp:73,
// src/libint_link.cpp:2261-2350).
void eri3c(const std::vector<Shell>& sh, const std::vector<Shell>& aux, double* out,
           bool full_nopt) {
  std::vector<long> off, aoff;
  long n, naux;
  offsets(sh, &off, &n);
  offsets(aux, &aoff, &naux);
  const long npair = n * (n + 1) / 2;
  const int ns = (int)sh.size(), nx = (int)aux.size();
  const double pi52 = 2.0 * std::pow(M_PI, 2.5);
  int lmax_aux = 0;
  for (auto& c : aux) lmax_aux = std::max(lmax_aux, c.l);
  std::vector<AuxHerm> herm;
  for (int l = 0; l <= lmax_aux; ++l) herm.push_back(aux_herm(l));
  // shell pairs with at least one significant primitive pair
  std::vector<std::pair<int, int>> pairs;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) pairs.push_back({a, b});

#pragma omp parallel
  {
    RWork rw;
    std::vector<double> Ex, Ey, Ez, W, cart, tmp, res;
    struct PP {
      double p, P[3], K;
      int off;  // into E arrays
    };
    std::vector<PP> pp;
#pragma omp for schedule(dynamic, 1)
    for (long ip = 0; ip < (long)pairs.size(); ++ip) {
      const Shell& A = sh[pairs[ip].first];
      const Shell& B = sh[pairs[ip].second];
      const int la = A.l, lb = B.l, na = A.ncart(), nb = B.ncart();
      const int Lab = la + lb, LE = Lab + 1;
      const int esz = (la + 1) * (lb + 1) * LE;
      double AB[3];
      for (int d = 0; d < 3; ++d) AB[d] = A.center[d] - B.center[d];
      const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
      pp.clear();
      Ex.clear();
      Ey.clear();
      Ez.clear();
      for (size_t pa = 0; pa < A.alpha.size(); ++pa)
        for (size_t pb = 0; pb < B.alpha.size(); ++pb) {
          const double a = A.alpha[pa], b = B.alpha[pb], p = a + b, mu = a * b / p;
          if (mu * rab2 > 46.0) continue;  // exp(-mu R^2) < 1e-20
          PP q;
          q.p = p;
          q.K = A.coef[pa] * B.coef[pb] * std::exp(-mu * rab2);
          for (int d = 0; d < 3; ++d) q.P[d] = (a * A.center[d] + b * B.center[d]) / p;
          q.off = (int)Ex.size();
          Ex.resize(Ex.size() + esz);
          Ey.resize(Ey.size() + esz);
          Ez.resize(Ez.size() + esz);
          hermite_E(la, lb, p, q.P[0] - A.center[0], q.P[0] - B.center[0], Ex.data() + q.off);
          hermite_E(la, lb, p, q.P[1] - A.center[1], q.P[1] - B.center[1], Ey.data() + q.off);
          hermite_E(la, lb, p, q.P[2] - A.center[2], q.P[2] - B.center[2], Ez.data() + q.off);
          pp.push_back(q);
        }
      if (pp.empty()) continue;  // all zero; output was zero-initialised
      // Cartesian exponent tables of the pair
      int ax[28], ay[28], az[28], bx[28], by[28], bz[28];
      for (int c = 0; c < na; ++c) cart_exponents(la, c, &ax[c], &ay[c], &az[c]);
      for (int c = 0; c < nb; ++c) cart_exponents(lb, c, &bx[c], &by[c], &bz[c]);
      const int sa = A.size(), sb = B.size();
      for (int ic = 0; ic < nx; ++ic) {
        const Shell& C = aux[ic];
        const int lc = C.l, nm = C.nsph();
        const AuxHerm& hc = herm[lc];
        const int L = Lab + lc, S = L + 1;
        cart.assign((size_t)na * nb * nm, 0.0);
        W.resize((size_t)LE * LE * LE * nm);
        for (size_t pc = 0; pc < C.alpha.size(); ++pc) {
          const double g = C.alpha[pc];
          const double gf = std::pow(2.0 * g, -lc) * ((lc & 1) ? -1.0 : 1.0) * C.coef[pc];
          for (const PP& q : pp) {
            const double p = q.p, ar = p * g / (p + g);
            const double* R = hermite_R(L, ar, q.P[0] - C.center[0], q.P[1] - C.center[1],
                                        q.P[2] - C.center[2], rw);
            const double pref = pi52 / (p * g * std::sqrt(p + g)) * q.K * gf;
            // W[tuv][m] = sum over S_lm terms of s * R_{t+tau,u+nu,v+phi}
            for (int t = 0; t <= Lab; ++t)
              for (int u = 0; u <= Lab - t; ++u)
                for (int v = 0; v <= Lab - t - u; ++v) {
                  double* w = &W[(((size_t)t * LE + u) * LE + v) * nm];
                  for (int m = 0; m < nm; ++m) {
                    double s = 0.0;
                    for (auto& tm : hc.m[m])
                      s += tm.c * R[((t + tm.ix) * S + u + tm.iy) * S + v + tm.iz];
                    w[m] = s * pref;
                  }
                }
            const double* ex = Ex.data() + q.off;
            const double* ey = Ey.data() + q.off;
            const double* ez = Ez.data() + q.off;
            double acc[16];
            for (int ca = 0; ca < na; ++ca)
              for (int cb = 0; cb < nb; ++cb) {
                for (int m = 0; m < nm; ++m) acc[m] = 0.0;
                const double* exr = ex + (ax[ca] * (lb + 1) + bx[cb]) * LE;
                const double* eyr = ey + (ay[ca] * (lb + 1) + by[cb]) * LE;
                const double* ezr = ez + (az[ca] * (lb + 1) + bz[cb]) * LE;
                const int tx = ax[ca] + bx[cb], ty = ay[ca] + by[cb], tz = az[ca] + bz[cb];
                for (int t = 0; t <= tx; ++t) {
                  const double e1 = exr[t];
                  for (int u = 0; u <= ty; ++u) {
                    const double e2 = e1 * eyr[u];
                    for (int v = 0; v <= tz; ++v) {
                      const double e3 = e2 * ezr[v];
                      const double* w = &W[(((size_t)t * LE + u) * LE + v) * nm];
                      for (int m = 0; m < nm; ++m) acc[m] += e3 * w[m];
                    }
                  }
                }
                double* dst = &cart[((size_t)ca * nb + cb) * nm];
                for (int m = 0; m < nm; ++m) dst[m] += acc[m];
              }
          }
        }
        // transform the (a,b) indices to the output representation
        res.resize((size_t)sa * sb * nm);
        transform_block(A, B, cart.data(), res.data(), nm, tmp);
        for (int i = 0; i < sa; ++i)
          for (int j = 0; j < sb; ++j) {
            long r = off[pairs[ip].first] + i, c = off[pairs[ip].second] + j;
            if (c > r) continue;  // diagonal shell pair: keep m >= n only
            for (int m = 0; m < nm; ++m) {
              const long P = aoff[ic] + m;
              const double v = res[((size_t)i * sb + j) * nm + m];
              if (full_nopt) {
                out[(r * n + c) * naux + P] = v;
                out[(c * n + r) * naux + P] = v;
              } else {
                out[P * npair + r * (r + 1) / 2 + c] = v;
              }
            }
          }
      }
    }
  }
}

}  // namespace nopt
