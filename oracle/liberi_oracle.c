/*
 * CPU oracle for the Schwarz-screened RHF Fock build -- TEST INFRASTRUCTURE.
 *
 * Restates gms_libERI's `compute_integrals` (src/liberi_driver.F90:52-167):
 * class order, per-class pre-screen and quartet test, value cutoff and
 * digestion are the reference's; the ERI values come from an independent
 * McMurchie-Davidson engine (Helgaker/Jorgensen/Olsen, "Molecular
 * Electronic-Structure Theory", ch. 9.5 and 9.9).  Imported only by tests,
 * the smoke check and the benchmark's CPU-baseline leg.
 */
#include "liberi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/liberi_parameters.F90:38 (also the literal 5.0D-11 of e.g.
 * src/rhf/rot_axis/int0002.F90:403: the same double) */
#define REF_CUTOFF 5.00000000000000018e-11
/* 2 pi^(5/2), src/liberi_parameters.F90:20 */
static const double TWO_PI_52 = 34.98683665524971863;

/* ---------------------------------------------------------------- classes */
/* pair classes, as the members of shell_pair_container_t
 * (src/liberi_types.f90:94-109) */
enum { P_SS, P_SP, P_PP, P_SD, P_PD, P_DD, P_SF, P_PF, P_DF, P_FF, NPAIRCLASS };
static const int pair_l[NPAIRCLASS][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2},
                                          {2, 2}, {0, 3}, {1, 3}, {2, 3}, {3, 3}};

/* quartet test of the kernel loop */
enum { CMP_GT,   /* keep if x_bra*x_ket > cut  (e.g. int0001.F90:121) */
       CMP_GE,   /* skip if x_bra*x_ket < cut  (int0000.F90:107, int0101, int1111) */
       CMP_NONE  /* no quartet test            (int2222_rysgen.F90:428-431, int3333) */
};
/* pair pre-screen threshold: cutoff / maxval(xints) of ... */
enum { PS_CROSS, /* ... the other pair class (rot-axis kernels, int0001.F90:67,76) */
       PS_SELF   /* ... the same pair class (Rys/ERIC kernels, int3300_rysgen.F90) */
};

typedef struct {
  const char *name;
  int bra, ket; /* pair classes; diag when bra == ket */
  int cmp, ps;
} qclass;

/* dispatch order of src/liberi_driver.F90:69-158; screening flags read from
 * each kernel file (scut* lines and the `test` comparison) */
static const qclass CLASSES[ORACLE_NCLASS] = {
    {"int0000", P_SS, P_SS, CMP_GE, PS_SELF},   {"int0001", P_SS, P_SP, CMP_GT, PS_CROSS},
    {"int0011", P_SS, P_PP, CMP_GT, PS_CROSS},  {"int0111", P_SP, P_PP, CMP_GT, PS_CROSS},
    {"int0101", P_SP, P_SP, CMP_GE, PS_SELF},   {"int1111", P_PP, P_PP, CMP_GE, PS_SELF},
    {"int0002", P_SS, P_SD, CMP_GT, PS_CROSS},  {"int0012", P_SS, P_PD, CMP_GT, PS_CROSS},
    {"int0102", P_SP, P_SD, CMP_GT, PS_CROSS},  {"int0112", P_SP, P_PD, CMP_GT, PS_CROSS},
    {"int0022", P_SS, P_DD, CMP_GT, PS_CROSS},  {"int0122", P_SP, P_DD, CMP_GT, PS_CROSS},
    {"int1102", P_PP, P_SD, CMP_GT, PS_CROSS},  {"int0202", P_SD, P_SD, CMP_GT, PS_SELF},
    {"int2211", P_DD, P_PP, CMP_GT, PS_SELF},   {"int2121", P_PD, P_PD, CMP_GT, PS_SELF},
    {"int2120", P_PD, P_SD, CMP_GT, PS_SELF},   {"int2111", P_PD, P_PP, CMP_GT, PS_SELF},
    {"int2220", P_DD, P_SD, CMP_GT, PS_SELF},   {"int2221", P_DD, P_PD, CMP_GT, PS_SELF},
    {"int2222gen", P_DD, P_DD, CMP_NONE, PS_SELF},
    {"int3333", P_FF, P_FF, CMP_NONE, PS_SELF}, {"int0030", P_SS, P_SF, CMP_GT, PS_SELF},
    {"int0031", P_SS, P_PF, CMP_GT, PS_SELF},   {"int0032", P_SS, P_DF, CMP_GT, PS_SELF},
    {"int1030", P_SP, P_SF, CMP_GT, PS_SELF},   {"int3330", P_FF, P_SF, CMP_GT, PS_SELF},
    {"int3011", P_SF, P_PP, CMP_GT, PS_SELF},   {"int3020", P_SF, P_SD, CMP_GT, PS_SELF},
    {"int2130", P_PD, P_SF, CMP_GT, PS_SELF},   {"int2230", P_DD, P_SF, CMP_GT, PS_SELF},
    {"int3030", P_SF, P_SF, CMP_GT, PS_SELF},   {"int3110", P_PF, P_SP, CMP_GT, PS_SELF},
    {"int3111", P_PF, P_PP, CMP_GT, PS_SELF},   {"int3120", P_PF, P_SD, CMP_GT, PS_SELF},
    {"int3121", P_PF, P_PD, CMP_GT, PS_SELF},   {"int2231", P_DD, P_PF, CMP_GT, PS_SELF},
    {"int3130", P_PF, P_SF, CMP_GT, PS_SELF},   {"int3131", P_PF, P_PF, CMP_GT, PS_SELF},
    {"int3210", P_DF, P_SP, CMP_GT, PS_SELF},   {"int3331", P_FF, P_PF, CMP_GT, PS_SELF},
    {"int3211", P_DF, P_PP, CMP_GT, PS_SELF},   {"int3220", P_DF, P_SD, CMP_GT, PS_SELF},
    {"int3221", P_DF, P_PD, CMP_GT, PS_SELF},   {"int3222", P_DF, P_DD, CMP_GT, PS_SELF},
    {"int3230", P_DF, P_SF, CMP_GT, PS_SELF},   {"int3231", P_DF, P_PF, CMP_GT, PS_SELF},
    {"int3232", P_DF, P_DF, CMP_GT, PS_SELF},   {"int3300", P_FF, P_SS, CMP_GT, PS_SELF},
    {"int3310", P_FF, P_SP, CMP_GT, PS_SELF},   {"int3311", P_FF, P_PP, CMP_GT, PS_SELF},
    {"int3320", P_FF, P_SD, CMP_GT, PS_SELF},   {"int3321", P_FF, P_PD, CMP_GT, PS_SELF},
    {"int3322", P_FF, P_DD, CMP_GT, PS_SELF},   {"int3332", P_FF, P_DF, CMP_GT, PS_SELF},
};

int oracle_nclass(void) { return ORACLE_NCLASS; }
const char *oracle_class_name(int c) {
  return (c >= 0 && c < ORACLE_NCLASS) ? CLASSES[c].name : "";
}
int oracle_class_def(int c, int def[4]) {
  if (c < 0 || c >= ORACLE_NCLASS) return -1;
  def[0] = CLASSES[c].bra;
  def[1] = CLASSES[c].ket;
  def[2] = CLASSES[c].cmp;
  def[3] = CLASSES[c].ps;
  return 0;
}

/* ------------------------------------------------------- cartesian tables */
/* GAMESS component order: p x,y,z; d xx,yy,zz,xy,xz,yz; f xxx,yyy,zzz,xxy,
 * xxz,xyy,yyz,xzz,yzz,xyz (src/rhf/eric/int0030_ericgen.F90:49-57 iord,
 * src/rhf/rys/int3300_rysgen.F90 ix/iy/iz tables).  Scale: sqrt3 on the
 * mixed d, sqrt5 on the two-index f, sqrt15 on xyz (angl at :54-57). */
static const int NCART[4] = {1, 3, 6, 10};
static const int CART[4][10][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
    {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {1, 1, 1}}};
static double cart_scale(int l, int c) {
  if (l == 2) return c >= 3 ? sqrt(3.0) : 1.0;
  if (l == 3) return c == 9 ? sqrt(15.0) : (c >= 3 ? sqrt(5.0) : 1.0);
  return 1.0;
}

/* ------------------------------------------------------------------ Boys */
double oracle_boys(int m, double T) {
  if (T < 1e-300) return 1.0 / (2 * m + 1);
  if (T > 50.0) {
    /* F_0 = sqrt(pi/4T) erf(sqrt T); erf = 1 to 1e-23 here; upward recursion
     * F_{k+1} = ((2k+1) F_k - e^-T) / 2T is stable for large T */
    long double f = sqrtl(3.14159265358979323846264338327950288L / (4.0L * T));
    long double e = expl(-(long double)T);
    for (int k = 0; k < m; ++k) f = ((2 * k + 1) * f - e) / (2.0L * T);
    return (double)f;
  }
  /* F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)); all terms > 0 */
  long double term = 1.0L / (2 * m + 1), sum = term, tt = 2.0L * T;
  for (int k = 1; k < 1000; ++k) {
    term *= tt / (2 * m + 2 * k + 1);
    sum += term;
    if (term < sum * 1e-21L) break;
  }
  return (double)(sum * expl(-(long double)T));
}

/* The reference's large-argument form: past its switch point each kernel
 * takes F_m = (2m-1)!!/(2T)^m sqrt(pi/4T) instead of its tabulated Boys
 * values or Rys roots (src/rhf/rot_axis/int0122.F90:493-501 and the same
 * branch in every rotated-axis kernel, src/rhf/eric/int0030_ericgen.F90:300-311;
 * its Rys kernels switch to asymptotic roots, whose moments are these values,
 * e.g. src/rhf/rot_axis/int2222_rysgen.F90:642).  Switch points: T > 25
 * (t_max, src/liberi_boys.F90:22) for the rotated-axis and ERIC classes,
 * X >= 37, 43, 49, 55, 60, 65 for the Rys classes with 2..7 roots. */
typedef struct {
  double t;   /* switch point (INFINITY: exact everywhere) */
  int ge;     /* 1: asymptotic when T >= t; 0: when T > t */
} boys_switch;

static double boys_asymptotic(int m, double T) {
  long double f = sqrtl(3.14159265358979323846264338327950288L / (4.0L * T));
  for (int k = 0; k < m; ++k) f *= (2 * k + 1) / (2.0L * T);
  return (double)f;
}

static double boys_sw(int m, double T, const boys_switch *sw) {
  if (sw && (sw->ge ? T >= sw->t : T > sw->t)) return boys_asymptotic(m, T);
  return oracle_boys(m, T);
}

/* --------------------------------------------------------------- shells */
typedef struct {
  int l, nprim, bf0;
  double ctr[3];
  const double *expo, *coef;
} shell_t;

static void get_shell(const oracle_basis *b, int s, shell_t *sh) {
  sh->l = b->ang_mom[s] - 1;
  sh->nprim = b->contr_num[s];
  sh->bf0 = b->atom_loc[s] - 1;
  int a = b->atom_num[s] - 1;
  for (int x = 0; x < 3; ++x) sh->ctr[x] = b->coords[3 * a + x];
  int off = b->sh_loc[s] - 1;
  sh->expo = b->exponents + off;
  const double *c = sh->l == 0 ? b->coef_s : sh->l == 1 ? b->coef_p : sh->l == 2 ? b->coef_d : b->coef_f;
  sh->coef = c + off;
}

/* ----------------------------------------------------- McMurchie-Davidson */
#define LMAX 3
#define LTOT (4 * LMAX)
#define NT (2 * LMAX + 1)

/* E^{ij}_t for one dimension: E[i][j][t], i<=la, j<=lb, t<=i+j; E^{00}_0 = 1 */
static void hermite_e(int la, int lb, double p, double xpa, double xpb, double E[LMAX + 1][LMAX + 1][NT]) {
  memset(E, 0, sizeof(double) * (LMAX + 1) * (LMAX + 1) * NT);
  const double h = 0.5 / p;
  E[0][0][0] = 1.0;
  for (int i = 0; i <= la; ++i) {
    if (i > 0)
      for (int t = 0; t <= i; ++t) {
        double v = xpa * E[i - 1][0][t];
        if (t > 0) v += h * E[i - 1][0][t - 1];
        if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
        E[i][0][t] = v;
      }
    for (int j = 1; j <= lb; ++j)
      for (int t = 0; t <= i + j; ++t) {
        double v = 0.0;
        if (t <= i + j - 1) v += xpb * E[i][j - 1][t];
        if (t > 0) v += h * E[i][j - 1][t - 1];
        if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        E[i][j][t] = v;
      }
  }
}

/* R^0_{tuv}, t+u+v <= L, stored R[(t*(L+1)+u)*(L+1)+v] */
static void hermite_r(int L, double alpha, const double pq[3], double *R, double *work, const boys_switch *sw) {
  const int n1 = L + 1;
  double T = alpha * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
  /* work[n][t][u][v] */
#define W(n, t, u, v) work[(((n) * n1 + (t)) * n1 + (u)) * n1 + (v)]
  double m2a = -2.0 * alpha, fac = 1.0;
  for (int n = 0; n <= L; ++n) {
    W(n, 0, 0, 0) = fac * boys_sw(n, T, sw);
    fac *= m2a;
  }
  for (int n = L - 1; n >= 0; --n) {
    int top = L - n;
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          if (t + u + v == 0) continue;
          double val;
          if (t > 0) {
            val = pq[0] * W(n + 1, t - 1, u, v);
            if (t > 1) val += (t - 1) * W(n + 1, t - 2, u, v);
          } else if (u > 0) {
            val = pq[1] * W(n + 1, t, u - 1, v);
            if (u > 1) val += (u - 1) * W(n + 1, t, u - 2, v);
          } else {
            val = pq[2] * W(n + 1, t, u, v - 1);
            if (v > 1) val += (v - 1) * W(n + 1, t, u, v - 2);
          }
          W(n, t, u, v) = val;
        }
  }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) R[(t * n1 + u) * n1 + v] = W(0, t, u, v);
#undef W
}

/* normalized cartesian block of shells A,B,C,D (GAMESS order) */
static void md_quartet(const shell_t *A, const shell_t *B, const shell_t *C, const shell_t *D, double *out,
                       const boys_switch *sw) {
  const int na = NCART[A->l], nb = NCART[B->l], nc = NCART[C->l], nd = NCART[D->l];
  const int nab = na * nb, ncd = nc * nd, n = nab * ncd;
  const int L = A->l + B->l + C->l + D->l, n1 = L + 1;
  const int lab = A->l + B->l;
  memset(out, 0, sizeof(double) * n);
  double ab2 = 0, cd2 = 0;
  for (int x = 0; x < 3; ++x) {
    ab2 += (A->ctr[x] - B->ctr[x]) * (A->ctr[x] - B->ctr[x]);
    cd2 += (C->ctr[x] - D->ctr[x]) * (C->ctr[x] - D->ctr[x]);
  }
  double *R = malloc(sizeof(double) * n1 * n1 * n1);
  double *work = malloc(sizeof(double) * n1 * n1 * n1 * n1);
  /* per (ket component, tuv) intermediate */
  const int ntuv = n1 * n1 * n1;
  double *Wk = malloc(sizeof(double) * ncd * ntuv);
  double *Eab = malloc(sizeof(double) * nab * ntuv); /* E^{ab}_{tuv} (only t+u+v<=lab used) */
  double Ex[LMAX + 1][LMAX + 1][NT], Ey[LMAX + 1][LMAX + 1][NT], Ez[LMAX + 1][LMAX + 1][NT];
  double Fx[LMAX + 1][LMAX + 1][NT], Fy[LMAX + 1][LMAX + 1][NT], Fz[LMAX + 1][LMAX + 1][NT];

  for (int ia = 0; ia < A->nprim; ++ia)
    for (int ib = 0; ib < B->nprim; ++ib) {
      const double a = A->expo[ia], bb = B->expo[ib], p = a + bb;
      const double mu = a * bb / p;
      const double kab = A->coef[ia] * B->coef[ib] * exp(-mu * ab2);
      double P[3];
      for (int x = 0; x < 3; ++x) P[x] = (a * A->ctr[x] + bb * B->ctr[x]) / p;
      hermite_e(A->l, B->l, p, P[0] - A->ctr[0], P[0] - B->ctr[0], Ex);
      hermite_e(A->l, B->l, p, P[1] - A->ctr[1], P[1] - B->ctr[1], Ey);
      hermite_e(A->l, B->l, p, P[2] - A->ctr[2], P[2] - B->ctr[2], Ez);
      for (int ca = 0; ca < na; ++ca)
        for (int cb = 0; cb < nb; ++cb) {
          const int *la = CART[A->l][ca], *lb = CART[B->l][cb];
          double *e = Eab + (ca * nb + cb) * ntuv;
          for (int t = 0; t <= la[0] + lb[0]; ++t)
            for (int u = 0; u <= la[1] + lb[1]; ++u)
              for (int v = 0; v <= la[2] + lb[2]; ++v)
                e[(t * n1 + u) * n1 + v] = Ex[la[0]][lb[0]][t] * Ey[la[1]][lb[1]][u] * Ez[la[2]][lb[2]][v];
        }
      for (int ic = 0; ic < C->nprim; ++ic)
        for (int id = 0; id < D->nprim; ++id) {
          const double c = C->expo[ic], d = D->expo[id], q = c + d;
          const double nu = c * d / q;
          const double kcd = C->coef[ic] * D->coef[id] * exp(-nu * cd2);
          double Q[3], pq[3];
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C->ctr[x] + d * D->ctr[x]) / q;
            pq[x] = P[x] - Q[x];
          }
          const double alpha = p * q / (p + q);
          hermite_r(L, alpha, pq, R, work, sw);
          hermite_e(C->l, D->l, q, Q[0] - C->ctr[0], Q[0] - D->ctr[0], Fx);
          hermite_e(C->l, D->l, q, Q[1] - C->ctr[1], Q[1] - D->ctr[1], Fy);
          hermite_e(C->l, D->l, q, Q[2] - C->ctr[2], Q[2] - D->ctr[2], Fz);
          const double pref = TWO_PI_52 / (p * q * sqrt(p + q)) * kab * kcd;
          /* Wk[cd][tuv] = sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd} R_{t+tau,u+nu,v+phi} */
          for (int cc = 0; cc < nc; ++cc)
            for (int cdd = 0; cdd < nd; ++cdd) {
              const int *lc = CART[C->l][cc], *ld = CART[D->l][cdd];
              double *w = Wk + (cc * nd + cdd) * ntuv;
              for (int t = 0; t <= lab; ++t)
                for (int u = 0; u <= lab - t; ++u)
                  for (int v = 0; v <= lab - t - u; ++v) {
                    double s = 0.0;
                    for (int tau = 0; tau <= lc[0] + ld[0]; ++tau) {
                      double ex = Fx[lc[0]][ld[0]][tau] * ((tau & 1) ? -1.0 : 1.0);
                      for (int nn = 0; nn <= lc[1] + ld[1]; ++nn) {
                        double ey = ex * Fy[lc[1]][ld[1]][nn] * ((nn & 1) ? -1.0 : 1.0);
                        for (int ph = 0; ph <= lc[2] + ld[2]; ++ph) {
                          double ez = ey * Fz[lc[2]][ld[2]][ph] * ((ph & 1) ? -1.0 : 1.0);
                          s += ez * R[((t + tau) * n1 + (u + nn)) * n1 + (v + ph)];
                        }
                      }
                    }
                    w[(t * n1 + u) * n1 + v] = s;
                  }
            }
          for (int ca = 0; ca < na; ++ca)
            for (int cb = 0; cb < nb; ++cb) {
              const int *la = CART[A->l][ca], *lb = CART[B->l][cb];
              const double *e = Eab + (ca * nb + cb) * ntuv;
              for (int k = 0; k < ncd; ++k) {
                const double *w = Wk + k * ntuv;
                double s = 0.0;
                for (int t = 0; t <= la[0] + lb[0]; ++t)
                  for (int u = 0; u <= la[1] + lb[1]; ++u)
                    for (int v = 0; v <= la[2] + lb[2]; ++v) {
                      int idx = (t * n1 + u) * n1 + v;
                      s += e[idx] * w[idx];
                    }
                out[(ca * nb + cb) * ncd + k] += pref * s;
              }
            }
        }
    }
  for (int ca = 0; ca < na; ++ca)
    for (int cb = 0; cb < nb; ++cb)
      for (int cc = 0; cc < nc; ++cc)
        for (int cdd = 0; cdd < nd; ++cdd)
          out[((ca * nb + cb) * nc + cc) * nd + cdd] *=
              cart_scale(A->l, ca) * cart_scale(B->l, cb) * cart_scale(C->l, cc) * cart_scale(D->l, cdd);
  free(R);
  free(work);
  free(Wk);
  free(Eab);
}

int oracle_shell_quartet(const oracle_basis *b, int i, int j, int k, int l, double *out) {
  if (i < 0 || j < 0 || k < 0 || l < 0 || i >= b->nsh || j >= b->nsh || k >= b->nsh || l >= b->nsh) return -1;
  shell_t A, B, C, D;
  get_shell(b, i, &A);
  get_shell(b, j, &B);
  get_shell(b, k, &C);
  get_shell(b, l, &D);
  if (A.l > LMAX || B.l > LMAX || C.l > LMAX || D.l > LMAX) return -2;
  md_quartet(&A, &B, &C, &D, out, NULL);
  return 0;
}

int oracle_schwarz(const oracle_basis *b, double *sw) {
  const int nsh = b->nsh;
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int i = 0; i < nsh; ++i) {
    double *blk = malloc(sizeof(double) * 10000);
    for (int j = 0; j <= i; ++j) {
      shell_t A, B;
      get_shell(b, i, &A);
      get_shell(b, j, &B);
      if (A.l > LMAX || B.l > LMAX) {
        bad = 1;
        continue;
      }
      md_quartet(&A, &B, &A, &B, blk, NULL);
      const int na = NCART[A.l], nb = NCART[B.l];
      double mx = 0.0;
      for (int a = 0; a < na; ++a)
        for (int c = 0; c < nb; ++c) {
          double v = fabs(blk[((a * nb + c) * na + a) * nb + c]);
          if (v > mx) mx = v;
        }
      sw[(int64_t)i * (i + 1) / 2 + j] = sqrt(mx);
    }
    free(blk);
  }
  return bad ? -2 : 0;
}

/* ------------------------------------------------------------ pair lists */
/* Shell lists per L in ascending shell index (src/liberi_shell_pair.F90:301-344
 * leaves them sorted ascending); pair lists as in compute_*_pairs: XX classes
 * triangular (ii >= jj), XY classes the full product. */
typedef struct {
  int n;
  int *sa, *sb; /* shells (0-based) */
  double *x;    /* Schwarz value, basis%schwrz_int(ia(max)+min) */
} pairlist;

typedef struct {
  int nl[4];
  int *list[4];
  pairlist pl[NPAIRCLASS];
} screening;

static double schwarz_of(const oracle_basis *b, int s1, int s2) {
  int hi = s1 > s2 ? s1 : s2, lo = s1 > s2 ? s2 : s1;
  return b->schwrz[(int64_t)hi * (hi + 1) / 2 + lo];
}

static void build_screening(const oracle_basis *b, screening *S) {
  memset(S, 0, sizeof(*S));
  for (int l = 0; l < 4; ++l) {
    S->list[l] = malloc(sizeof(int) * (b->nsh + 1));
    for (int s = 0; s < b->nsh; ++s)
      if (b->ang_mom[s] - 1 == l) S->list[l][S->nl[l]++] = s;
  }
  for (int pc = 0; pc < NPAIRCLASS; ++pc) {
    int l1 = pair_l[pc][0], l2 = pair_l[pc][1];
    int n1 = S->nl[l1], n2 = S->nl[l2];
    int64_t n = (l1 == l2) ? (int64_t)n1 * (n1 + 1) / 2 : (int64_t)n1 * n2;
    pairlist *p = &S->pl[pc];
    p->n = (int)n;
    p->sa = malloc(sizeof(int) * (n + 1));
    p->sb = malloc(sizeof(int) * (n + 1));
    p->x = malloc(sizeof(double) * (n + 1));
    int64_t k = 0;
    if (l1 == l2) {
      for (int ii = 0; ii < n1; ++ii)
        for (int jj = 0; jj <= ii; ++jj) {
          p->sa[k] = S->list[l1][ii];
          p->sb[k] = S->list[l1][jj];
          p->x[k] = schwarz_of(b, p->sa[k], p->sb[k]);
          ++k;
        }
    } else {
      for (int ii = 0; ii < n1; ++ii)
        for (int jj = 0; jj < n2; ++jj) {
          p->sa[k] = S->list[l1][ii];
          p->sb[k] = S->list[l2][jj];
          p->x[k] = schwarz_of(b, p->sa[k], p->sb[k]);
          ++k;
        }
    }
  }
}

static void free_screening(screening *S) {
  for (int l = 0; l < 4; ++l) free(S->list[l]);
  for (int pc = 0; pc < NPAIRCLASS; ++pc) {
    free(S->pl[pc].sa);
    free(S->pl[pc].sb);
    free(S->pl[pc].x);
  }
}

static double maxval(const pairlist *p) {
  /* Fortran maxval of a non-empty array */
  double m = -INFINITY;
  for (int i = 0; i < p->n; ++i)
    if (p->x[i] > m) m = p->x[i];
  return m;
}

/* kept list (positions into the pair list) after the pair pre-screen
 * `xints(ij) .ge. cutoff/maxval(...)` (e.g. int0001.F90:65-84) */
static int prescreen(const pairlist *p, double scut, int *keep) {
  int n = 0;
  for (int i = 0; i < p->n; ++i)
    if (p->x[i] >= scut) keep[n++] = i;
  return n;
}

/* which classes the driver calls (src/liberi_driver.F90:69-159) */
static int class_dispatched(const screening *S, const qclass *q) {
  int lmax = 0;
  for (int k = 0; k < 2; ++k) {
    int pc = k ? q->ket : q->bra;
    if (pair_l[pc][0] > lmax) lmax = pair_l[pc][0];
    if (pair_l[pc][1] > lmax) lmax = pair_l[pc][1];
  }
  if (S->nl[lmax] == 0) return 0;
  /* a class over an empty pair list has nothing to do (the reference would
   * divide by zero there, int0000.F90:80-81) */
  if (S->pl[q->bra].n == 0 || S->pl[q->ket].n == 0) return 0;
  return 1;
}

static inline int quartet_passes(int cmp, double t, double cut) {
  if (cmp == CMP_GT) return t > cut;
  if (cmp == CMP_GE) return !(t < cut);
  return 1;
}

typedef struct {
  int nb, nk;
  int *kb, *kk; /* kept positions */
} kept_t;

static void class_kept(const screening *S, const qclass *q, double cut, kept_t *K) {
  const pairlist *pb = &S->pl[q->bra], *pk = &S->pl[q->ket];
  double sb, sk;
  if (q->ps == PS_CROSS) {
    sb = cut / maxval(pk);
    sk = cut / maxval(pb);
  } else {
    sb = cut / maxval(pb);
    sk = cut / maxval(pk);
  }
  K->kb = malloc(sizeof(int) * (pb->n + 1));
  K->kk = malloc(sizeof(int) * (pk->n + 1));
  K->nb = prescreen(pb, sb, K->kb);
  K->nk = (q->bra == q->ket) ? 0 : prescreen(pk, sk, K->kk);
}

int oracle_screen_counts(const oracle_basis *b, double cut, int64_t counts[ORACLE_NCLASS]) {
  screening S;
  build_screening(b, &S);
  for (int c = 0; c < ORACLE_NCLASS; ++c) {
    const qclass *q = &CLASSES[c];
    counts[c] = 0;
    if (!class_dispatched(&S, q)) continue;
    kept_t K;
    class_kept(&S, q, cut, &K);
    const pairlist *pb = &S.pl[q->bra], *pk = &S.pl[q->ket];
    int64_t cnt = 0;
    if (q->bra == q->ket) {
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cnt)
      for (int i = 0; i < K.nb; ++i)
        for (int j = 0; j <= i; ++j)
          cnt += quartet_passes(q->cmp, pb->x[K.kb[i]] * pb->x[K.kb[j]], cut);
    } else {
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cnt)
      for (int i = 0; i < K.nb; ++i)
        for (int j = 0; j < K.nk; ++j) cnt += quartet_passes(q->cmp, pb->x[K.kb[i]] * pk->x[K.kk[j]], cut);
    }
    counts[c] = cnt;
    free(K.kb);
    free(K.kk);
  }
  free_screening(&S);
  return 0;
}

/* -------------------------------------------------------------- digestion */
static inline int64_t tri(int64_t i, int64_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* DIRFCK_RHF restated over a whole shell block (src/rhf/rot_axis/int0000.F90:241-311):
 * every component integral v with |v| >= cutoff contributes
 *   F[ab] += 4 s v D[cd], F[cd] += 4 s v D[ab],
 *   F[ac] -= s v D[bd], F[bd] -= s v D[ac], F[ad] -= s v D[bc], F[bc] -= s v D[ad]
 * into the packed lower triangle, where s = 1/2 per coinciding shell pair
 * position (A==B, C==D, AB==CD).  Summed over all component quadruples of the
 * block this is the reference's per-unique-integral halving rule. */
static void digest(const shell_t *A, const shell_t *B, const shell_t *C, const shell_t *D, int sameAB,
                   int sameCD, int samePair, const double *blk, const double *Dm, double *F, double cut) {
  const int na = NCART[A->l], nb = NCART[B->l], nc = NCART[C->l], nd = NCART[D->l];
  double s = 1.0;
  if (sameAB) s *= 0.5;
  if (sameCD) s *= 0.5;
  if (samePair) s *= 0.5;
  for (int a = 0; a < na; ++a)
    for (int bq = 0; bq < nb; ++bq)
      for (int c = 0; c < nc; ++c)
        for (int d = 0; d < nd; ++d) {
          double v = blk[((a * nb + bq) * nc + c) * nd + d];
          if (fabs(v) < cut) continue;
          const double w = s * v;
          const int64_t i = A->bf0 + a, j = B->bf0 + bq, k = C->bf0 + c, l = D->bf0 + d;
          const int64_t ij = tri(i, j), kl = tri(k, l), ik = tri(i, k), jl = tri(j, l), il = tri(i, l),
                        jk = tri(j, k);
          F[ij] += 4.0 * w * Dm[kl];
          F[kl] += 4.0 * w * Dm[ij];
          F[ik] -= w * Dm[jl];
          F[jl] -= w * Dm[ik];
          F[il] -= w * Dm[jk];
          F[jk] -= w * Dm[il];
        }
}

int64_t oracle_fock(const oracle_basis *b, const double *Dm, double *F, double cut, int64_t counts[ORACLE_NCLASS],
                    int64_t stride, int nthreads) {
  return oracle_fock_classes(b, Dm, F, cut, counts, stride, nthreads, NULL);
}

/* switch point of class q in the reference (see boys_switch) */
static boys_switch reference_switch(const qclass *q) {
  boys_switch sw;
  const int c = (int)(q - CLASSES);
  if (c <= 13 || (c >= 22 && c <= 25)) { /* rotated-axis and ERIC kernels */
    sw.t = 25.0;
    sw.ge = 0;
  } else {
    static const double thr[8] = {0, 0, 37.0, 43.0, 49.0, 55.0, 60.0, 65.0};
    const int L = pair_l[q->bra][0] + pair_l[q->bra][1] + pair_l[q->ket][0] + pair_l[q->ket][1];
    sw.t = thr[L / 2 + 1];
    sw.ge = 1;
  }
  return sw;
}

int64_t oracle_fock_classes(const oracle_basis *b, const double *Dm, double *F, double cut,
                            int64_t counts[ORACLE_NCLASS], int64_t stride, int nthreads, const int32_t *enable) {
  return oracle_fock_classes_sw(b, Dm, F, cut, counts, stride, nthreads, enable, 0);
}

int64_t oracle_fock_classes_sw(const oracle_basis *b, const double *Dm, double *F, double cut,
                               int64_t counts[ORACLE_NCLASS], int64_t stride, int nthreads, const int32_t *enable,
                               int ref_switch) {
  for (int s = 0; s < b->nsh; ++s)
    if (b->ang_mom[s] < 1 || b->ang_mom[s] > 4) return -1;
  if (stride < 1) stride = 1;
  const int64_t nbf = b->num_bas, ntri = nbf * (nbf + 1) / 2;
  memset(F, 0, sizeof(double) * ntri); /* fock = 0 (src/liberi_driver.F90:64) */
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  screening S;
  build_screening(b, &S);
  int64_t evaluated = 0;
  for (int cidx = 0; cidx < ORACLE_NCLASS; ++cidx) {
    const qclass *q = &CLASSES[cidx];
    if (counts) counts[cidx] = 0;
    if (!class_dispatched(&S, q)) continue;
    if (enable && !enable[cidx]) continue;
    kept_t K;
    class_kept(&S, q, cut, &K);
    const pairlist *pb = &S.pl[q->bra], *pk = &S.pl[q->ket];
    const int diag = q->bra == q->ket;
    /* surviving quartets as (kept bra position, kept ket position) */
    int64_t cap = 1024, n = 0;
    int32_t *qb = malloc(sizeof(int32_t) * cap), *qk = malloc(sizeof(int32_t) * cap);
    for (int i = 0; i < K.nb; ++i) {
      int jn = diag ? i + 1 : K.nk;
      for (int j = 0; j < jn; ++j) {
        double xb = pb->x[K.kb[i]], xk = diag ? pb->x[K.kb[j]] : pk->x[K.kk[j]];
        if (!quartet_passes(q->cmp, xb * xk, cut)) continue;
        if (n == cap) {
          cap *= 2;
          qb = realloc(qb, sizeof(int32_t) * cap);
          qk = realloc(qk, sizeof(int32_t) * cap);
        }
        qb[n] = K.kb[i];
        qk[n] = diag ? K.kb[j] : K.kk[j];
        ++n;
      }
    }
    if (counts) counts[cidx] = n;
    const boys_switch sw = reference_switch(q);
    int64_t nev = 0;
#pragma omp parallel reduction(+ : nev)
    {
      double *blk = malloc(sizeof(double) * 10000);
      double *Fp = calloc(ntri, sizeof(double));
#pragma omp for schedule(dynamic, 16)
      for (int64_t t = 0; t < n; t += stride) {
        shell_t A, B, C, D;
        int sa = pb->sa[qb[t]], sb = pb->sb[qb[t]], sc = pk->sa[qk[t]], sd = pk->sb[qk[t]];
        get_shell(b, sa, &A);
        get_shell(b, sb, &B);
        get_shell(b, sc, &C);
        get_shell(b, sd, &D);
        md_quartet(&A, &B, &C, &D, blk, ref_switch ? &sw : NULL);
        digest(&A, &B, &C, &D, sa == sb, sc == sd, diag && qb[t] == qk[t], blk, Dm, Fp, cut);
        ++nev;
      }
#pragma omp critical
      for (int64_t i = 0; i < ntri; ++i) F[i] += Fp[i];
      free(Fp);
      free(blk);
    }
    evaluated += nev;
    free(qb);
    free(qk);
    free(K.kb);
    free(K.kk);
  }
  free_screening(&S);
  return evaluated;
}
