/* ORACLE -- test infrastructure only. Nothing in the product links, imports
 * or calls this file; it exists to check the product against.
 *
 * FP64 Gaussian integrals by the Obara-Saika recurrences, restating the
 * algorithm of libint2 v2.13.1 (commit 5fb07b4862f2, the version NOPT pins in
 * dockerbuild/Dockerfile:100,213-214) at the call sites NOPT uses:
 *   one-electron S, T, V     libint_link.cpp:31-150 (AO_1el_from_2shells)
 *   2-centre metric (P|Q)    libint_link.cpp:2049-2107 (RI_V_calc, xs_xs)
 *   3-centre (mn|P)          libint_link.cpp:2261-2350 (calc_2el_AO_INTS_RI,
 *                            RI_AA layout [(m*n_ao+n)*naux+P], both triangles)
 * The orbital basis is Cartesian (NOPT reads it with pure=false,
 * src/inp_par_read.cpp:2550-2552); every Cartesian component is normalised
 * like x^l, as libint2 does. Auxiliary shells are spherical: the caller
 * passes, per l, a matrix whose rows span the harmonic (traceless) subspace
 * of the degree-l Cartesian monomials.
 *
 * Recurrences: Obara & Saika, J. Chem. Phys. 84, 3963 (1986); horizontal
 * recurrence of Head-Gordon & Pople, J. Chem. Phys. 89, 5777 (1988).
 * Boys function by its power series (small T) and erf-based asymptotics with
 * upward recursion (large T); no tables.
 *
 * Parity pins: the three RI-RHF energies NOPT prints in its example logs (see
 * tests/test_oracle.py), closed forms for s-type integrals, and the
 * independent McMurchie-Davidson engine of the product.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 8          /* per-shell angular momentum limit */
#define MAXD (2 * MAXL) /* highest summed degree */

/* ---------------------------------------------------------------- Boys ---- */
static void boys_series(int nmax, double T, double* F) {
  if (T > 35.0) {
    double eT = exp(-T);
    F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
    return;
  }
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 5000; ++k) {
    term *= 2.0 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  double eT = exp(-T);
  F[nmax] = eT * sum;
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + eT) / (2 * n + 1);
}

double oracle_boys(int n, double T) {
  double F[64];
  boys_series(n, T, F);
  return F[n];
}

/* ------------------------------------------------------ monomial tables --- */
/* Monomials of degree d in the libint2 order (x power l..0, then y). The
 * cumulative index of a monomial of degree d is d(d+1)(d+2)/6 + position. */
static int mono_pos(int ix, int iy, int iz) {
  int d = ix + iy + iz, i = d - ix;
  return i * (i + 1) / 2 + (d - ix - iy);
}
static int mono_cum(int ix, int iy, int iz) {
  int d = ix + iy + iz;
  return d * (d + 1) * (d + 2) / 6 + mono_pos(ix, iy, iz);
}
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int ncum(int d) { return (d + 1) * (d + 2) * (d + 3) / 6; } /* degrees 0..d */
static void mono_of(int d, int pos, int* e) {
  e[0] = e[1] = e[2] = 0;
  int n = 0;
  for (int i = d; i >= 0; --i)
    for (int j = d - i; j >= 0; --j, ++n)
      if (n == pos) {
        e[0] = i;
        e[1] = j;
        e[2] = d - i - j;
        return;
      }
}

static double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

/* ------------------------------------------------------------- shells ---- */
typedef struct {
  int l;
  double c[3];
  int np;
  const double* a;
  double* k; /* normalised coefficients */
} shl;

/* normalise like libint2: primitives as x^l, then the contraction */
static void shl_init(shl* s, int l, const double* c, int np, const double* a, const double* k) {
  s->l = l;
  memcpy(s->c, c, sizeof(double) * 3);
  s->np = np;
  s->a = a;
  s->k = (double*)malloc(sizeof(double) * np);
  double df = dfact(2 * l - 1);
  for (int i = 0; i < np; ++i)
    s->k[i] = k[i] * pow(2.0 * a[i] / M_PI, 0.75) * pow(4.0 * a[i], 0.5 * l) / sqrt(df);
  double ov = 0.0;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      double p = a[i] + a[j];
      ov += s->k[i] * s->k[j] * pow(M_PI / p, 1.5) * df / pow(2.0 * p, l);
    }
  for (int i = 0; i < np; ++i) s->k[i] /= sqrt(ov);
}

static shl* load_shells(int nsh, const int* l, const double* center, const int* np,
                        const int* poff, const double* alpha, const double* coef) {
  shl* s = (shl*)malloc(sizeof(shl) * nsh);
  for (int i = 0; i < nsh; ++i)
    shl_init(&s[i], l[i], center + 3 * i, np[i], alpha + poff[i], coef + poff[i]);
  return s;
}
static void free_shells(shl* s, int n) {
  for (int i = 0; i < n; ++i) free(s[i].k);
  free(s);
}

/* ------------------------------------------------- one-electron, 1-D OS --- */
/* S[i][j] and T[i][j] (kinetic, -1/2 d^2/dx^2) for one direction. */
static void os_1d(int la, int lb, double a, double b, double XA, double XB,
                  double S[MAXL + 3][MAXL + 3], double T[MAXL + 3][MAXL + 3]) {
  double p = a + b, mu = a * b / p;
  double Px = (a * XA + b * XB) / p, PA = Px - XA, PB = Px - XB, AB = XA - XB;
  double o2p = 0.5 / p;
  memset(S, 0, sizeof(double) * (MAXL + 3) * (MAXL + 3));
  memset(T, 0, sizeof(double) * (MAXL + 3) * (MAXL + 3));
  S[0][0] = sqrt(M_PI / p) * exp(-mu * AB * AB);
  for (int i = 0; i <= la + 1; ++i)
    for (int j = 0; j <= lb + 1; ++j) {
      if (i == 0 && j == 0) continue;
      if (i > 0) {
        double v = PA * S[i - 1][j];
        if (i > 1) v += o2p * (i - 1) * S[i - 2][j];
        if (j > 0) v += o2p * j * S[i - 1][j - 1];
        S[i][j] = v;
      } else {
        double v = PB * S[i][j - 1];
        if (j > 1) v += o2p * (j - 1) * S[i][j - 2];
        S[i][j] = v;
      }
    }
  /* kinetic recurrences (Obara-Saika 1986, eqs. for T_ij) */
  T[0][0] = (a - 2.0 * a * a * (PA * PA + o2p)) * S[0][0];
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      if (i > 0) {
        int ii = i - 1;
        double v = PA * T[ii][j] + 2.0 * mu * S[i][j];
        if (ii > 0) v += o2p * ii * T[ii - 1][j] - mu / a * ii * S[ii - 1][j];
        if (j > 0) v += o2p * j * T[ii][j - 1];
        T[i][j] = v;
      } else {
        int jj = j - 1;
        double v = PB * T[i][jj] + 2.0 * mu * S[i][j];
        if (jj > 0) v += o2p * jj * T[i][jj - 1] - mu / b * jj * S[i][jj - 1];
        T[i][j] = v;
      }
    }
}

/* ------------------------------------------------ horizontal recurrence --- */
/* X[cum(e)][f] for |e| in [la, la+lb] -> out[a][b][f] (a, b Cartesian of
 * degrees la, lb), (a, b+1_i| = (a+1_i, b| + AB_i (a, b|. */
static void hrr(int la, int lb, const double* AB, int nf, const double* X, double* out) {
  int L = la + lb;
  if (lb == 0) {
    for (int a = 0; a < ncart(la); ++a) {
      int e[3];
      mono_of(la, a, e);
      memcpy(out + (size_t)a * nf, X + (size_t)mono_cum(e[0], e[1], e[2]) * nf, sizeof(double) * nf);
    }
    return;
  }
  /* level k holds (e, b| for |b| = k, |e| in [la, L-k]: index [cum(e)][pos(b)][f] */
  int ne = ncum(L);
  double* prev = (double*)calloc((size_t)ne * 1 * nf, sizeof(double));
  memcpy(prev, X, sizeof(double) * (size_t)ne * nf);
  for (int k = 1; k <= lb; ++k) {
    int nb = ncart(k), nbp = ncart(k - 1);
    double* cur = (double*)calloc((size_t)ne * nb * nf, sizeof(double));
    for (int d = la; d <= L - k; ++d)
      for (int pe = 0; pe < ncart(d); ++pe) {
        int e[3];
        mono_of(d, pe, e);
        int ce = mono_cum(e[0], e[1], e[2]);
        for (int pb = 0; pb < nb; ++pb) {
          int b[3];
          mono_of(k, pb, b);
          int i = b[0] ? 0 : (b[1] ? 1 : 2);
          int bp[3] = {b[0], b[1], b[2]};
          bp[i]--;
          int e1[3] = {e[0], e[1], e[2]};
          e1[i]++;
          int c1 = mono_cum(e1[0], e1[1], e1[2]);
          int pbp = mono_pos(bp[0], bp[1], bp[2]);
          for (int f = 0; f < nf; ++f)
            cur[((size_t)ce * nb + pb) * nf + f] = prev[((size_t)c1 * nbp + pbp) * nf + f] +
                                                   AB[i] * prev[((size_t)ce * nbp + pbp) * nf + f];
        }
      }
    free(prev);
    prev = cur;
  }
  int nb = ncart(lb);
  for (int a = 0; a < ncart(la); ++a) {
    int e[3];
    mono_of(la, a, e);
    int ce = mono_cum(e[0], e[1], e[2]);
    for (int b = 0; b < nb; ++b)
      memcpy(out + ((size_t)a * nb + b) * nf, prev + ((size_t)ce * nb + b) * nf, sizeof(double) * nf);
  }
  free(prev);
}

/* ------------------------------------------------- nuclear attraction ---- */
/* [e|V_C|0]^(m) by the vertical recurrence, contracted over primitives. */
static void nuclear_shellpair(const shl* A, const shl* B, int natom, const double* Zc,
                              const double* xyz, double* out /* [a][b] */) {
  int L = A->l + B->l, ne = ncum(L);
  double* X = (double*)calloc(ne, sizeof(double));
  double* I = (double*)malloc(sizeof(double) * ne * (L + 1));
  double AB[3];
  for (int d = 0; d < 3; ++d) AB[d] = A->c[d] - B->c[d];
  double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  for (int i = 0; i < A->np; ++i)
    for (int j = 0; j < B->np; ++j) {
      double a = A->a[i], b = B->a[j], p = a + b, mu = a * b / p;
      double P[3], PA[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A->c[d] + b * B->c[d]) / p;
        PA[d] = P[d] - A->c[d];
      }
      double kab = A->k[i] * B->k[j] * exp(-mu * rab2) * 2.0 * M_PI / p;
      for (int c = 0; c < natom; ++c) {
        double PC[3], F[64];
        for (int d = 0; d < 3; ++d) PC[d] = P[d] - xyz[3 * c + d];
        boys_series(L, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F);
        for (int m = 0; m <= L; ++m) I[m] = kab * F[m];
        for (int dgr = 1; dgr <= L; ++dgr)
          for (int pe = 0; pe < ncart(dgr); ++pe) {
            int e[3];
            mono_of(dgr, pe, e);
            int ce = mono_cum(e[0], e[1], e[2]);
            int k = e[0] ? 0 : (e[1] ? 1 : 2);
            int e1[3] = {e[0], e[1], e[2]};
            e1[k]--;
            int c1 = mono_cum(e1[0], e1[1], e1[2]);
            int c2 = -1;
            if (e1[k] > 0) {
              int e2[3] = {e1[0], e1[1], e1[2]};
              e2[k]--;
              c2 = mono_cum(e2[0], e2[1], e2[2]);
            }
            for (int m = 0; m <= L - dgr; ++m) {
              double v = PA[k] * I[c1 * (L + 1) + m] - PC[k] * I[c1 * (L + 1) + m + 1];
              if (c2 >= 0)
                v += e1[k] * 0.5 / p * (I[c2 * (L + 1) + m] - I[c2 * (L + 1) + m + 1]);
              I[ce * (L + 1) + m] = v;
            }
          }
        for (int e = 0; e < ne; ++e) X[e] -= Zc[c] * I[e * (L + 1)];
      }
    }
  hrr(A->l, B->l, AB, 1, X, out);
  free(X);
  free(I);
}

/* ------------------------------------------------ Coulomb, 3-centre VRR --- */
/* Contracted [e 0 | f] over primitives of (A,B) and of C, |e| in [0, la+lb],
 * |f| = lc (Cartesian); result X[cum(e)][pos(f)]. */
static void eri_vrr(const shl* A, const shl* B, const shl* C, double* X) {
  const int Le = A->l + B->l, Lf = C->l, Lm = Le + Lf;
  const int ne = ncum(Le), nfc = ncum(Lf), nf = ncart(Lf);
  const size_t stride_f = (size_t)(Lm + 1), stride_e = (size_t)nfc * stride_f;
  double* I = (double*)malloc(sizeof(double) * ne * stride_e);
  memset(X, 0, sizeof(double) * ne * nf);
  double AB[3];
  for (int d = 0; d < 3; ++d) AB[d] = A->c[d] - B->c[d];
  double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  for (int i = 0; i < A->np; ++i)
    for (int j = 0; j < B->np; ++j) {
      double a = A->a[i], b = B->a[j], p = a + b, mu = a * b / p;
      double P[3], PA[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A->c[d] + b * B->c[d]) / p;
        PA[d] = P[d] - A->c[d];
      }
      double kab = A->k[i] * B->k[j] * exp(-mu * rab2);
      for (int kq = 0; kq < C->np; ++kq) {
        double q = C->a[kq], rho = p * q / (p + q);
        double W[3], WP[3], WQ[3], PQ[3], F[64];
        for (int d = 0; d < 3; ++d) {
          W[d] = (p * P[d] + q * C->c[d]) / (p + q);
          WP[d] = W[d] - P[d];
          WQ[d] = W[d] - C->c[d];
          PQ[d] = P[d] - C->c[d];
        }
        boys_series(Lm, rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), F);
        double base = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * kab * C->k[kq];
#define IX(ce, cf, m) I[(size_t)(ce) * stride_e + (size_t)(cf) * stride_f + (m)]
        for (int m = 0; m <= Lm; ++m) IX(0, 0, m) = base * F[m];
        /* bra build-up with f = 0 */
        for (int dg = 1; dg <= Le; ++dg)
          for (int pe = 0; pe < ncart(dg); ++pe) {
            int e[3];
            mono_of(dg, pe, e);
            int ce = mono_cum(e[0], e[1], e[2]);
            int k = e[0] ? 0 : (e[1] ? 1 : 2);
            int e1[3] = {e[0], e[1], e[2]};
            e1[k]--;
            int c1 = mono_cum(e1[0], e1[1], e1[2]), c2 = -1;
            if (e1[k] > 0) {
              int e2[3] = {e1[0], e1[1], e1[2]};
              e2[k]--;
              c2 = mono_cum(e2[0], e2[1], e2[2]);
            }
            for (int m = 0; m <= Lm - dg; ++m) {
              double v = PA[k] * IX(c1, 0, m) + WP[k] * IX(c1, 0, m + 1);
              if (c2 >= 0) v += e1[k] * 0.5 / p * (IX(c2, 0, m) - rho / p * IX(c2, 0, m + 1));
              IX(ce, 0, m) = v;
            }
          }
        /* ket build-up (Q = C, so QC = 0) */
        for (int fg = 1; fg <= Lf; ++fg)
          for (int pf = 0; pf < ncart(fg); ++pf) {
            int f[3];
            mono_of(fg, pf, f);
            int cf = mono_cum(f[0], f[1], f[2]);
            int k = f[0] ? 0 : (f[1] ? 1 : 2);
            int f1[3] = {f[0], f[1], f[2]};
            f1[k]--;
            int cf1 = mono_cum(f1[0], f1[1], f1[2]), cf2 = -1;
            if (f1[k] > 0) {
              int f2[3] = {f1[0], f1[1], f1[2]};
              f2[k]--;
              cf2 = mono_cum(f2[0], f2[1], f2[2]);
            }
            for (int dg = 0; dg <= Le; ++dg)
              for (int pe = 0; pe < ncart(dg); ++pe) {
                int e[3];
                mono_of(dg, pe, e);
                int ce = mono_cum(e[0], e[1], e[2]), ce1 = -1;
                if (e[k] > 0) {
                  int em[3] = {e[0], e[1], e[2]};
                  em[k]--;
                  ce1 = mono_cum(em[0], em[1], em[2]);
                }
                for (int m = 0; m <= Lm - dg - fg; ++m) {
                  double v = WQ[k] * IX(ce, cf1, m + 1);
                  if (cf2 >= 0) v += f1[k] * 0.5 / q * (IX(ce, cf2, m) - rho / q * IX(ce, cf2, m + 1));
                  if (ce1 >= 0) v += e[k] * 0.5 / (p + q) * IX(ce1, cf1, m + 1);
                  IX(ce, cf, m) = v;
                }
              }
          }
        for (int ce = 0; ce < ne; ++ce)
          for (int pf = 0; pf < nf; ++pf) {
            int f[3];
            mono_of(Lf, pf, f);
            X[(size_t)ce * nf + pf] += IX(ce, mono_cum(f[0], f[1], f[2]), 0);
          }
#undef IX
      }
    }
  free(I);
}

static long basis_size(const shl* s, int n, int pure) {
  long k = 0;
  for (int i = 0; i < n; ++i) k += pure ? 2 * s[i].l + 1 : ncart(s[i].l);
  return k;
}

/* spherical transform of the f index of a [..][ncart(l)] block */
static const double* tf_of(const double* tf, int l) {
  size_t off = 0;
  for (int k = 0; k < l; ++k) off += (size_t)(2 * k + 1) * ncart(k);
  return tf + off;
}

/* ======================================================== entry points === */

/* Cartesian S, T, V (n_ao x n_ao). Zc, xyz: nuclear charges and positions. */
int oracle_one_electron(int nsh, const int* l, const double* center, const int* np,
                        const int* poff, const double* alpha, const double* coef, int natom,
                        const double* Zc, const double* xyz, double* S, double* T, double* V) {
  shl* sh = load_shells(nsh, l, center, np, poff, alpha, coef);
  long n = basis_size(sh, nsh, 0);
  long* off = (long*)malloc(sizeof(long) * nsh);
  long k = 0;
  for (int i = 0; i < nsh; ++i) {
    off[i] = k;
    k += ncart(sh[i].l);
  }
#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < nsh; ++ia)
    for (int ib = 0; ib <= ia; ++ib) {
      const shl *A = &sh[ia], *B = &sh[ib];
      int na = ncart(A->l), nb = ncart(B->l);
      double* bs = (double*)calloc(na * nb, sizeof(double));
      double* bt = (double*)calloc(na * nb, sizeof(double));
      double* bv = (double*)calloc(na * nb, sizeof(double));
      for (int i = 0; i < A->np; ++i)
        for (int j = 0; j < B->np; ++j) {
          double Sx[MAXL + 3][MAXL + 3], Tx[MAXL + 3][MAXL + 3], Sy[MAXL + 3][MAXL + 3],
              Ty[MAXL + 3][MAXL + 3], Sz[MAXL + 3][MAXL + 3], Tz[MAXL + 3][MAXL + 3];
          os_1d(A->l, B->l, A->a[i], B->a[j], A->c[0], B->c[0], Sx, Tx);
          os_1d(A->l, B->l, A->a[i], B->a[j], A->c[1], B->c[1], Sy, Ty);
          os_1d(A->l, B->l, A->a[i], B->a[j], A->c[2], B->c[2], Sz, Tz);
          double cc = A->k[i] * B->k[j];
          for (int pa = 0; pa < na; ++pa) {
            int ea[3];
            mono_of(A->l, pa, ea);
            for (int pb = 0; pb < nb; ++pb) {
              int eb[3];
              mono_of(B->l, pb, eb);
              double sx = Sx[ea[0]][eb[0]], sy = Sy[ea[1]][eb[1]], sz = Sz[ea[2]][eb[2]];
              bs[pa * nb + pb] += cc * sx * sy * sz;
              bt[pa * nb + pb] += cc * (Tx[ea[0]][eb[0]] * sy * sz + sx * Ty[ea[1]][eb[1]] * sz +
                                        sx * sy * Tz[ea[2]][eb[2]]);
            }
          }
        }
      nuclear_shellpair(A, B, natom, Zc, xyz, bv);
      for (int pa = 0; pa < na; ++pa)
        for (int pb = 0; pb < nb; ++pb) {
          long r = off[ia] + pa, c = off[ib] + pb;
          if (S) S[r * n + c] = S[c * n + r] = bs[pa * nb + pb];
          if (T) T[r * n + c] = T[c * n + r] = bt[pa * nb + pb];
          if (V) V[r * n + c] = V[c * n + r] = bv[pa * nb + pb];
        }
      free(bs);
      free(bt);
      free(bv);
    }
  free(off);
  free_shells(sh, nsh);
  return 0;
}

/* (P|Q) over spherical auxiliary functions; tf: concatenated per-l
 * (2l+1) x ncart(l) matrices for l = 0.. */
int oracle_metric(int nsh, const int* l, const double* center, const int* np, const int* poff,
                  const double* alpha, const double* coef, const double* tf, double* V) {
  shl* sh = load_shells(nsh, l, center, np, poff, alpha, coef);
  long n = basis_size(sh, nsh, 1);
  long* off = (long*)malloc(sizeof(long) * nsh);
  long k = 0;
  for (int i = 0; i < nsh; ++i) {
    off[i] = k;
    k += 2 * sh[i].l + 1;
  }
#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < nsh; ++ia)
    for (int ic = 0; ic <= ia; ++ic) {
      const shl *A = &sh[ia], *C = &sh[ic];
      int la = A->l, lc = C->l, nca = ncart(la), ncc = ncart(lc);
      /* (a 0 | c): B = A with a single unit-coefficient s primitive of
       * exponent 0 would be singular; use the bra as a one-centre pair by
       * treating B as an s function of zero exponent contribution: handled by
       * a dedicated shell with a = A's exponent split. Simplest exact route:
       * (a|c) = (a s0|c) with s0 = 1, i.e. B->np = 1, exponent 0, coef 1. */
      double zero = 0.0, one = 1.0;
      shl B;
      B.l = 0;
      memcpy(B.c, A->c, sizeof(double) * 3);
      B.np = 1;
      B.a = &zero;
      B.k = &one;
      int ne = ncum(la);
      double* X = (double*)malloc(sizeof(double) * ne * ncc);
      eri_vrr(A, &B, C, X);
      double* cart = (double*)malloc(sizeof(double) * nca * ncc);
      hrr(la, 0, (double[3]){0, 0, 0}, ncc, X, cart);
      const double* Ta = tf_of(tf, la);
      const double* Tc = tf_of(tf, lc);
      for (int ma = 0; ma < 2 * la + 1; ++ma)
        for (int mc = 0; mc < 2 * lc + 1; ++mc) {
          double s = 0.0;
          for (int a = 0; a < nca; ++a)
            for (int c = 0; c < ncc; ++c) s += Ta[ma * nca + a] * Tc[mc * ncc + c] * cart[a * ncc + c];
          long r = off[ia] + ma, cc = off[ic] + mc;
          V[r * n + cc] = V[cc * n + r] = s;
        }
      free(X);
      free(cart);
    }
  free(off);
  free_shells(sh, nsh);
  return 0;
}

/* (mn|P) in NOPT's RI_AA layout out[(m*n_ao+n)*naux+P], Cartesian orbital
 * basis, spherical auxiliary basis; both (m,n) and (n,m) are written. */
int oracle_eri3(int nsh, const int* l, const double* center, const int* np, const int* poff,
                const double* alpha, const double* coef, int nash, const int* al,
                const double* acenter, const int* anp, const int* apoff, const double* aalpha,
                const double* acoef, const double* tf, double* out) {
  shl* sh = load_shells(nsh, l, center, np, poff, alpha, coef);
  shl* ax = load_shells(nash, al, acenter, anp, apoff, aalpha, acoef);
  long n = basis_size(sh, nsh, 0), naux = basis_size(ax, nash, 1);
  long *off = (long*)malloc(sizeof(long) * nsh), *aoff = (long*)malloc(sizeof(long) * nash);
  long k = 0;
  for (int i = 0; i < nsh; ++i) {
    off[i] = k;
    k += ncart(sh[i].l);
  }
  k = 0;
  for (int i = 0; i < nash; ++i) {
    aoff[i] = k;
    k += 2 * ax[i].l + 1;
  }
#pragma omp parallel for schedule(dynamic)
  for (long ij = 0; ij < (long)nsh * nsh; ++ij) {
    int ia = (int)(ij / nsh), ib = (int)(ij % nsh);
    if (ib > ia) continue;
    const shl *A = &sh[ia], *B = &sh[ib];
    int la = A->l, lb = B->l, na = ncart(la), nb = ncart(lb);
    double AB[3];
    for (int d = 0; d < 3; ++d) AB[d] = A->c[d] - B->c[d];
    for (int ic = 0; ic < nash; ++ic) {
      const shl* C = &ax[ic];
      int lc = C->l, ncc = ncart(lc), nsc = 2 * lc + 1;
      int ne = ncum(la + lb);
      double* X = (double*)malloc(sizeof(double) * ne * ncc);
      eri_vrr(A, B, C, X);
      double* cart = (double*)malloc(sizeof(double) * na * nb * ncc);
      hrr(la, lb, AB, ncc, X, cart);
      const double* Tc = tf_of(tf, lc);
      for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
          for (int mc = 0; mc < nsc; ++mc) {
            double s = 0.0;
            for (int c = 0; c < ncc; ++c) s += Tc[mc * ncc + c] * cart[((size_t)a * nb + b) * ncc + c];
            long r = off[ia] + a, cc = off[ib] + b, P = aoff[ic] + mc;
            out[(r * n + cc) * naux + P] = s;
            out[(cc * n + r) * naux + P] = s;
          }
      free(X);
      free(cart);
    }
  }
  free(off);
  free(aoff);
  free_shells(sh, nsh);
  free_shells(ax, nash);
  return 0;
}
