/*
 * astr_oracle.c -- CPU restatement of the ASTR per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (astr_amd/, include/)
 * links, loads or calls this file; it is the checker the HIP path is
 * compared against.
 *
 * Semantics restated: the single-file TGV solver
 * miniapps/tgv_solver_3d/tgvsolver.F90 (the reference's runnable hot path,
 * "astr.min") on top of src/derivative.F90, src/filter.F90 and the Thomas
 * routines of src/commfunc.F90, all with the ntype=3 (periodic, single
 * block) closures that a one-rank run uses.
 *
 * How each routine is pinned (besides itself):
 *   or_thomas_*, or_deriv_*, or_df_compact     -> reference Fortran df_compact
 *        (oracle/_ref build) on random lines; dense numpy solve of the same
 *        (n+3)-row system; exactness on polynomials of degree <= 6; modified
 *        wavenumber of Lele (1992) on a sine.
 *   or_filter_*, or_compact_filter             -> reference Fortran
 *        compact_filter; dense solve; constant preservation; transfer
 *        function T(w)=sum c_m 2cos(mw)/(1+2a cos w) on interior rows.
 *   or_df_compact_nt, or_compact_filter_nt     -> the reference's own
 *        fd_scheme_initiate/df_compact and compact_filter_initiate/
 *        compact_filter with ntype 1,2,4 (oracle/_ref), bit for bit;
 *        exactness of the closed derivative system on cubics (its
 *        3rd-order boundary row fails on quartics); the boundary
 *        filter's preservation of polynomials of degree <= 5 with the
 *        main path's beta_bound 0.98.
 *   or_rhscal, or_filterq, or_rk3_steps,
 *   or_stacal, or_tgv_init                      -> reference Fortran
 *        rhscal/filterq/rk3/stacal run through oracle/_ref on the same
 *        input; the reference-held history state.ref_128 (128^3, 100 steps);
 *        closed form KE(0)=1/8 for the unperturbed field.
 *   or_gauss                                    -> not a reference routine
 *        (input generator); checked for mean/variance and against the
 *        published Philox4x32-10 known-answer vectors.
 */
#include "astr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define HM OR_HM

/* constdef.F90:11-27 */
static const double num1d3 = 1.0 / 3.0, num7d9 = 7.0 / 9.0,
                    num1d36 = 1.0 / 36.0, num1d60 = 1.0 / 60.0;

/* ------------------------------------------------------------------ */
/* Thomas: commfunc.F90:752-774 (preprocess) and :790-813 (solve).     */
/* ac is stored row-major ac[3*i + (0,1,2)] for row i=0..n-1.          */
/* ------------------------------------------------------------------ */
void or_thomas_prep(int n, const double *a, const double *c, double *ac) {
  ac[0] = c[0];
  ac[1] = 0.0; /* never read: row 1 of the reference's solve is left as is */
  ac[2] = 0.0;
  for (int i = 1; i < n; i++) {
    double den = 1.0 - a[i] * ac[3 * (i - 1)];
    ac[3 * i + 0] = c[i] / den;
    ac[3 * i + 1] = 1.0 / den;
    ac[3 * i + 2] = a[i] / den;
  }
}

void or_thomas_solve(int n, const double *ac, double *d, double *x) {
  for (int i = 1; i < n; i++) d[i] = d[i] * ac[3 * i + 1] - d[i - 1] * ac[3 * i + 2];
  x[n - 1] = d[n - 1];
  for (int i = n - 2; i >= 0; i--) x[i] = d[i] - ac[3 * i] * x[i + 1];
}

/* ------------------------------------------------------------------ */
/* Compact 6th-order first derivative, scheme 643c, ntype=3:           */
/* derivative.F90:63-158 (LHS), :171-198 (df_compact), :210-306 (RHS). */
/* Unknowns x(-1..n+1) stored at index u=r+1.                          */
/* ------------------------------------------------------------------ */
int or_deriv_nrows(int n) { return n + 3; }

void or_deriv_lhs(int n, double *a, double *c) {
  int m = n + 3;
  for (int u = 0; u < m; u++) { a[u] = num1d3; c[u] = num1d3; }
  a[0] = 0.0; c[0] = 0.0;         /* explicit central row at i_0=-1 */
  a[m - 1] = 0.0; c[m - 1] = 0.0; /* explicit central row at i_m=n+1 */
}

/* f[t+HM] = f(t), t=-HM..n+HM */
void or_deriv_rhs(int n, const double *f, double *d) {
  const double *F = f + HM;
  int m = n + 3;
  /* rows i_0=-1 and i_m=n+1: 6th-order explicit, derivative.F90:255-260,287-292 */
  int js[2] = {-1, n + 1};
  int us[2] = {0, m - 1};
  for (int e = 0; e < 2; e++) {
    int j = js[e];
    double var1 = F[j + 1] - F[j - 1];
    double var2 = F[j + 2] - F[j - 2];
    double var3 = F[j + 3] - F[j - 3];
    d[us[e]] = 0.75 * var1 - 0.15 * var2 + num1d60 * var3;
  }
  /* rows 0..n: 6th-order compact, derivative.F90:296-304 */
  for (int j = 0; j <= n; j++) {
    double var1 = F[j + 1] - F[j - 1];
    double var2 = F[j + 2] - F[j - 2];
    d[j + 1] = num7d9 * var1 + num1d36 * var2;
  }
}

void or_df_compact(int n, const double *f, double *df) {
  int m = n + 3;
  double *a = malloc(sizeof(double) * m), *c = malloc(sizeof(double) * m);
  double *ac = malloc(sizeof(double) * 3 * m), *d = malloc(sizeof(double) * m);
  double *x = malloc(sizeof(double) * m);
  or_deriv_lhs(n, a, c);
  or_thomas_prep(m, a, c, ac);
  or_deriv_rhs(n, f, d);
  or_thomas_solve(m, ac, d, x);
  for (int j = 0; j <= n; j++) df[j] = x[j + 1];
  free(a); free(c); free(ac); free(d); free(x);
}

/* ------------------------------------------------------------------ */
/* Compact 10th-order filter, ntype=3: filter.F90:31-100 (LHS),        */
/* :112-144 (compact_filter), :156-285 (RHS), :299-432 (coefficients). */
/* Unknowns x(-3..n+3) stored at u=r+3.                                */
/* coefh is row-major coefh[k*11+j], k=0..2, j=0..10.                  */
/* ------------------------------------------------------------------ */
int or_filter_nrows(int n) { return n + 7; }

void or_filter_coefs(double alfa, double bh, double bb, double *coef10i, double *coefh) {
  (void)bb; /* beta_bound enters only physical-boundary rows (ntype 1,2,4) */
  coef10i[0] = (193.0 + 126.0 * alfa) / 512.0;
  coef10i[1] = (105.0 + 302.0 * alfa) / 512.0;
  coef10i[2] = (-15.0 + 30.0 * alfa) / 128.0;
  coef10i[3] = (45.0 - 90.0 * alfa) / 1024.0;
  coef10i[4] = (-5.0 + 10.0 * alfa) / 512.0;
  coef10i[5] = (1.0 - 2.0 * alfa) / 1024.0;
  double *h0 = coefh, *h1 = coefh + 11, *h2 = coefh + 22;
  h0[0] = (-1.0 + 1.0 * bh) / 1024.0;
  h0[1] = (5.0 - 5.0 * bh) / 512.0;
  h0[2] = (979.0 + 45.0 * bh) / 1024.0;
  h0[3] = (15.0 + 113.0 * bh) / 128.0;
  h0[4] = (-105.0 + 105.0 * bh) / 512.0;
  h0[5] = (63.0 - 63.0 * bh) / 256.0;
  h0[6] = (-105.0 + 105.0 * bh) / 512.0;
  h0[7] = (15.0 - 15.0 * bh) / 128.0;
  h0[8] = (-45.0 + 45.0 * bh) / 1024.0;
  h0[9] = (5.0 - 5.0 * bh) / 512.0;
  h0[10] = (-1.0 + 1.0 * bh) / 1024.0;
  h1[0] = (1.0 - 2.0 * alfa) / 1024.0;
  h1[1] = (-5.0 + 10.0 * alfa) / 512.0;
  h1[2] = (45.0 + 934.0 * alfa) / 1024.0;
  h1[3] = (113.0 + 30.0 * alfa) / 128.0;
  h1[4] = (105.0 + 302.0 * alfa) / 512.0;
  h1[5] = (-63.0 + 126.0 * alfa) / 256.0;
  h1[6] = (105.0 - 210.0 * alfa) / 512.0;
  h1[7] = (-15.0 + 30.0 * alfa) / 128.0;
  h1[8] = (45.0 - 90.0 * alfa) / 1024.0;
  h1[9] = (-5.0 + 10.0 * alfa) / 512.0;
  h1[10] = (1.0 - 2.0 * alfa) / 1024.0;
  h2[0] = (-1.0 + 2.0 * alfa) / 1024.0;
  h2[1] = (5.0 - 10.0 * alfa) / 512.0;
  h2[2] = (-45.0 + 90.0 * alfa) / 1024.0;
  h2[3] = (15.0 + 98.0 * alfa) / 128.0;
  h2[4] = (407.0 + 210.0 * alfa) / 512.0;
  h2[5] = (63.0 + 130.0 * alfa) / 256.0;
  h2[6] = (-105.0 + 210.0 * alfa) / 512.0;
  h2[7] = (15.0 - 30.0 * alfa) / 128.0;
  h2[8] = (-45.0 + 90.0 * alfa) / 1024.0;
  h2[9] = (5.0 - 10.0 * alfa) / 512.0;
  h2[10] = (-1.0 + 2.0 * alfa) / 1024.0;
}

void or_filter_lhs(int n, double alfa, double bh, double *a, double *c) {
  int m = n + 7;
  for (int u = 0; u < m; u++) { a[u] = alfa; c[u] = alfa; }
  a[0] = bh; c[0] = bh;         /* filter.F90:90-91 */
  a[m - 1] = bh; c[m - 1] = bh; /* filter.F90:93-94 */
}

void or_filter_rhs(int n, const double *coef10i, const double *coefh, const double *f,
                   double *d) {
  const double *F = f + HM;
  int i0 = -3, im_ = n + 3;
  /* rows -3,-2,-1: filter.F90:212-218; stencil f(i_0-2+j), j=0..10 */
  for (int k = i0; k <= i0 + 2; k++) {
    double v = 0.0;
    for (int j = 0; j <= 10; j++) v = v + coefh[(k - i0) * 11 + j] * F[i0 - 2 + j];
    d[k - i0] = v;
  }
  /* rows n+1..n+3: filter.F90:255-261; stencil f(i_m+2-j) */
  for (int k = im_ - 2; k <= im_; k++) {
    double v = 0.0;
    for (int j = 0; j <= 10; j++) v = v + coefh[(im_ - k) * 11 + j] * F[im_ + 2 - j];
    d[k - i0] = v;
  }
  /* rows 0..n: filter.F90:271-283 */
  for (int j = 0; j <= n; j++) {
    double var0 = F[j] + F[j];
    double var1 = F[j + 1] + F[j - 1];
    double var2 = F[j + 2] + F[j - 2];
    double var3 = F[j + 3] + F[j - 3];
    double var4 = F[j + 4] + F[j - 4];
    double var5 = F[j + 5] + F[j - 5];
    d[j - i0] = coef10i[0] * var0 + coef10i[1] * var1 + coef10i[2] * var2 +
                coef10i[3] * var3 + coef10i[4] * var4 + coef10i[5] * var5;
  }
}

void or_compact_filter(int n, double alfa, double bh, double bb, const double *f,
                       double *ff) {
  int m = n + 7;
  double c10[6], ch[33];
  or_filter_coefs(alfa, bh, bb, c10, ch);
  double *a = malloc(sizeof(double) * m), *c = malloc(sizeof(double) * m);
  double *ac = malloc(sizeof(double) * 3 * m), *d = malloc(sizeof(double) * m);
  double *x = malloc(sizeof(double) * m);
  or_filter_lhs(n, alfa, bh, a, c);
  or_thomas_prep(m, a, c, ac);
  or_filter_rhs(n, c10, ch, f, d);
  or_thomas_solve(m, ac, d, x);
  for (int j = 0; j <= n; j++) ff[j] = x[j + 3];
  free(a); free(c); free(ac); free(d); free(x);
}

/* ------------------------------------------------------------------ */
/* Lines with physical boundaries: boundary type ntype (1: physical    */
/* boundary at 0, 2: at n, 3: interfaces at both ends, 4: physical at  */
/* both ends), as fd_scheme_initiate / compact_fd_rhs                  */
/* (derivative.F90:63-158, :210-306) and compact_filter_initiate /     */
/* compact_filter / compact_filter_rhs (filter.F90:31-285) set them up. */
/* Unknowns x(i_0..i_m) stored at u = r - i_0.                         */
/* ------------------------------------------------------------------ */
static int left_phys(int ntype) { return ntype == 1 || ntype == 4; }
static int right_phys(int ntype) { return ntype == 2 || ntype == 4; }

int or_df_compact_nt(int n, int ntype, const double *f, double *df) {
  if (ntype < 1 || ntype > 4) return -1;
  const double *F = f + HM;
  const int i0 = left_phys(ntype) ? 0 : -1;      /* derivative.F90:74-93 */
  const int im_ = right_phys(ntype) ? n : n + 1;
  const int m = im_ - i0 + 1;
  double *a = malloc(sizeof(double) * m), *c = malloc(sizeof(double) * m);
  double *ac = malloc(sizeof(double) * 3 * m), *d = malloc(sizeof(double) * m);
  double *x = malloc(sizeof(double) * m);
  /* LHS, derivative.F90:106-136 */
  for (int u = 0; u < m; u++) { a[u] = num1d3; c[u] = num1d3; }
  a[0] = c[0] = 0.0;
  a[m - 1] = c[m - 1] = 0.0;
  if (left_phys(ntype)) { a[0] = c[0] = 2.0; a[1] = c[1] = 0.25; }
  if (right_phys(ntype)) { a[m - 1] = c[m - 1] = 2.0; a[m - 2] = c[m - 2] = 0.25; }
  /* RHS, derivative.F90:229-304 */
  int is, ie, j;
  if (left_phys(ntype)) {
    is = i0 + 2;
    j = i0 + 1;
    d[j - i0] = 0.75 * (F[j + 1] - F[j - 1]);
    j = i0;
    d[j - i0] = -2.5 * F[j] + 2.0 * F[j + 1] + 0.5 * F[j + 2];
  } else {
    is = i0 + 1;
    j = i0;
    double var1 = F[j + 1] - F[j - 1], var2 = F[j + 2] - F[j - 2], var3 = F[j + 3] - F[j - 3];
    d[j - i0] = 0.75 * var1 - 0.15 * var2 + num1d60 * var3;
  }
  if (right_phys(ntype)) {
    ie = im_ - 2;
    j = im_ - 1;
    d[j - i0] = 0.75 * (F[j + 1] - F[j - 1]);
    j = im_;
    d[j - i0] = 2.5 * F[j] - 2.0 * F[j - 1] - 0.5 * F[j - 2];
  } else {
    ie = im_ - 1;
    j = im_;
    double var1 = F[j + 1] - F[j - 1], var2 = F[j + 2] - F[j - 2], var3 = F[j + 3] - F[j - 3];
    d[j - i0] = 0.75 * var1 - 0.15 * var2 + num1d60 * var3;
  }
  for (j = is; j <= ie; j++) {
    double var1 = F[j + 1] - F[j - 1], var2 = F[j + 2] - F[j - 2];
    d[j - i0] = num7d9 * var1 + num1d36 * var2;
  }
  or_thomas_prep(m, a, c, ac);
  or_thomas_solve(m, ac, d, x);
  for (j = 0; j <= n; j++) df[j] = x[j - i0];
  free(a); free(c); free(ac); free(d); free(x);
  return 0;
}

/* filter_coefficient_cal (filter.F90:299-432): the rows the physical
 * boundary uses, coefb(0:2,0:6) row-major as coefb[k*7+j], coef6i(0:3),
 * coef8i(0:4) */
void or_filter_coefs_b(double alfa, double bb, double *coefb, double *coef6i, double *coef8i) {
  coef6i[0] = (11.0 + 10.0 * alfa) / 32.0;
  coef6i[1] = (15.0 + 34.0 * alfa) / 64.0;
  coef6i[2] = (-3.0 + 6.0 * alfa) / 32.0;
  coef6i[3] = (1.0 - 2.0 * alfa) / 64.0;
  coef8i[0] = (93.0 + 70.0 * alfa) / 256.0;
  coef8i[1] = (7.0 + 18.0 * alfa) / 32.0;
  coef8i[2] = (-7.0 + 14.0 * alfa) / 64.0;
  coef8i[3] = (1.0 - 2.0 * alfa) / 32.0;
  coef8i[4] = (-1.0 + 2.0 * alfa) / 256.0;
  double *b0 = coefb, *b1 = coefb + 7, *b2 = coefb + 14;
  b2[0] = (-1.0 + 2.0 * alfa) / 64.0;
  b2[1] = (3.0 + 26.0 * alfa) / 32.0;
  b2[2] = (49.0 + 30.0 * alfa) / 64.0;
  b2[3] = (5.0 + 6.0 * alfa) / 16.0;
  b2[4] = (-15.0 + 30.0 * alfa) / 64.0;
  b2[5] = (3.0 - 6.0 * alfa) / 32.0;
  b2[6] = (-1.0 + 2.0 * alfa) / 64.0;
  b1[0] = (1.0 + 62.0 * alfa) / 64.0;
  b1[1] = (29.0 + 6.0 * alfa) / 32.0;
  b1[2] = (15.0 + 34.0 * alfa) / 64.0;
  b1[3] = (-5.0 + 10.0 * alfa) / 16.0;
  b1[4] = (15.0 - 30.0 * alfa) / 64.0;
  b1[5] = (-3.0 + 6.0 * alfa) / 32.0;
  b1[6] = (1.0 - 2.0 * alfa) / 64.0;
  b0[0] = (63.0 + 1.0 * bb) / 64.0;
  b0[1] = (3.0 + 29.0 * bb) / 32.0;
  b0[2] = (-15.0 + 15.0 * bb) / 64.0;
  b0[3] = (5.0 - 5.0 * bb) / 16.0;
  b0[4] = (-15.0 + 15.0 * bb) / 64.0;
  b0[5] = (3.0 - 3.0 * bb) / 32.0;
  b0[6] = (-1.0 + 1.0 * bb) / 64.0;
}

/* The LHS takes beta_halo at an interface end (filter.F90:90-94 with the
 * main path's beter_halo) and 0.98 at a physical end (filter.F90:47-66). */
int or_compact_filter_nt(int n, int ntype, double alfa, double bh, double bb, const double *f,
                         double *ff) {
  if (ntype < 1 || ntype > 4) return -1;
  const double *F = f + HM;
  const int i0 = left_phys(ntype) ? 0 : -3;
  const int im_ = right_phys(ntype) ? n : n + 3;
  const double b0 = left_phys(ntype) ? 0.98 : bh, bm = right_phys(ntype) ? 0.98 : bh;
  const int m = im_ - i0 + 1;
  double c10[6], ch[33], cb[21], c6[4], c8[5];
  or_filter_coefs(alfa, bh, bb, c10, ch);
  or_filter_coefs_b(alfa, bb, cb, c6, c8);
  double *a = malloc(sizeof(double) * m), *c = malloc(sizeof(double) * m);
  double *ac = malloc(sizeof(double) * 3 * m), *d = malloc(sizeof(double) * m);
  double *x = malloc(sizeof(double) * m);
  for (int u = 0; u < m; u++) { a[u] = alfa; c[u] = alfa; }
  a[0] = c[0] = b0;
  a[m - 1] = c[m - 1] = bm;
  int is, ie, j, k;
  if (left_phys(ntype)) {        /* filter.F90:175-205 */
    is = i0 + 5;
    for (k = i0; k <= i0 + 2; k++) {
      double v = 0.0;
      for (j = 0; j <= 6; j++) v = v + cb[(k - i0) * 7 + j] * F[i0 + j];
      d[k - i0] = v;
    }
    j = i0 + 3;
    d[j - i0] = c6[0] * (F[j] + F[j]) + c6[1] * (F[j + 1] + F[j - 1]) +
                c6[2] * (F[j + 2] + F[j - 2]) + c6[3] * (F[j + 3] + F[j - 3]);
    j = i0 + 4;
    d[j - i0] = c8[0] * (F[j] + F[j]) + c8[1] * (F[j + 1] + F[j - 1]) +
                c8[2] * (F[j + 2] + F[j - 2]) + c8[3] * (F[j + 3] + F[j - 3]) +
                c8[4] * (F[j + 4] + F[j - 4]);
  } else {                       /* filter.F90:207-219 */
    is = i0 + 3;
    for (k = i0; k <= i0 + 2; k++) {
      double v = 0.0;
      for (j = 0; j <= 10; j++) v = v + ch[(k - i0) * 11 + j] * F[i0 - 2 + j];
      d[k - i0] = v;
    }
  }
  if (right_phys(ntype)) {       /* filter.F90:221-250 */
    ie = im_ - 5;
    for (k = im_ - 2; k <= im_; k++) {
      double v = 0.0;
      for (j = 0; j <= 6; j++) v = v + cb[(im_ - k) * 7 + j] * F[im_ - j];
      d[k - i0] = v;
    }
    j = im_ - 3;
    d[j - i0] = c6[0] * (F[j] + F[j]) + c6[1] * (F[j + 1] + F[j - 1]) +
                c6[2] * (F[j + 2] + F[j - 2]) + c6[3] * (F[j + 3] + F[j - 3]);
    j = im_ - 4;
    d[j - i0] = c8[0] * (F[j] + F[j]) + c8[1] * (F[j + 1] + F[j - 1]) +
                c8[2] * (F[j + 2] + F[j - 2]) + c8[3] * (F[j + 3] + F[j - 3]) +
                c8[4] * (F[j + 4] + F[j - 4]);
  } else {                       /* filter.F90:252-262 */
    ie = im_ - 3;
    for (k = im_ - 2; k <= im_; k++) {
      double v = 0.0;
      for (j = 0; j <= 10; j++) v = v + ch[(im_ - k) * 11 + j] * F[im_ + 2 - j];
      d[k - i0] = v;
    }
  }
  for (j = is; j <= ie; j++) {   /* filter.F90:271-283 */
    double var0 = F[j] + F[j], var1 = F[j + 1] + F[j - 1], var2 = F[j + 2] + F[j - 2];
    double var3 = F[j + 3] + F[j - 3], var4 = F[j + 4] + F[j - 4], var5 = F[j + 5] + F[j - 5];
    d[j - i0] = c10[0] * var0 + c10[1] * var1 + c10[2] * var2 + c10[3] * var3 + c10[4] * var4 +
                c10[5] * var5;
  }
  or_thomas_prep(m, a, c, ac);
  or_thomas_solve(m, ac, d, x);
  for (j = 0; j <= n; j++) ff[j] = x[j - i0];
  if (left_phys(ntype)) ff[0] = F[0];   /* filter.F90:141-142 */
  if (right_phys(ntype)) ff[n] = F[n];
  free(a); free(c); free(ac); free(d); free(x);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Field state and the miniapp stage (tgvsolver.F90)                   */
/* ------------------------------------------------------------------ */
typedef struct line_sys {
  int n, m;
  double *ac;
} line_sys;

struct or_state {
  or_cfg c;
  long N;
  int n[3];     /* im, jm, km */
  long st[3];   /* strides */
  int ndir;     /* 3, or 2 when km==0 */
  double const1, const2, const5, const6;
  double *q, *prims, *qrhs, *dvel, *dtmp, *qsave, *sigma, *qflux;
  line_sys dsys[3], fsys[3];
  double c10[6], ch[33];
  int nstep;
};

static void sys_init(line_sys *s, int n, int filt, const or_cfg *c) {
  s->n = n;
  s->m = filt ? n + 7 : n + 3;
  double *a = malloc(sizeof(double) * s->m), *cc = malloc(sizeof(double) * s->m);
  if (filt) or_filter_lhs(n, c->alfa_filter, c->beta_halo, a, cc);
  else or_deriv_lhs(n, a, cc);
  s->ac = malloc(sizeof(double) * 3 * s->m);
  or_thomas_prep(s->m, a, cc, s->ac);
  free(a); free(cc);
}

or_state *or_create(const or_cfg *cfg) {
  or_state *s = calloc(1, sizeof(or_state));
  s->c = *cfg;
  s->n[0] = cfg->im; s->n[1] = cfg->jm; s->n[2] = cfg->km;
  s->ndir = cfg->km > 0 ? 3 : 2;
  s->st[0] = 1; s->st[1] = cfg->im + 1; s->st[2] = (long)(cfg->im + 1) * (cfg->jm + 1);
  s->N = (long)(cfg->im + 1) * (cfg->jm + 1) * (cfg->km + 1);
  double g = cfg->gamma, ma = cfg->mach;
  /* tgvsolver.F90:189-194 */
  s->const1 = 1.0 / (g * (g - 1.0) * ma * ma);
  s->const2 = g * ma * ma;
  s->const5 = (g - 1.0) * ma * ma;
  s->const6 = 1.0 / (g - 1.0);
  long N = s->N;
  s->q = calloc(5 * N, sizeof(double));
  s->prims = calloc(6 * N, sizeof(double));
  s->qrhs = calloc(5 * N, sizeof(double));
  s->dvel = calloc(9 * N, sizeof(double));
  s->dtmp = calloc(3 * N, sizeof(double));
  s->qsave = calloc(5 * N, sizeof(double));
  s->sigma = calloc(6 * N, sizeof(double));
  s->qflux = calloc(3 * N, sizeof(double));
  for (int d = 0; d < s->ndir; d++) {
    sys_init(&s->dsys[d], s->n[d], 0, cfg);
    sys_init(&s->fsys[d], s->n[d], 1, cfg);
  }
  or_filter_coefs(cfg->alfa_filter, cfg->beta_halo, cfg->beta_bound, s->c10, s->ch);
  return s;
}

void or_destroy(or_state *s) {
  if (!s) return;
  free(s->q); free(s->prims); free(s->qrhs); free(s->dvel); free(s->dtmp);
  free(s->qsave); free(s->sigma); free(s->qflux);
  for (int d = 0; d < s->ndir; d++) { free(s->dsys[d].ac); free(s->fsys[d].ac); }
  free(s);
}

long or_npoints(const or_state *s) { return s->N; }
double *or_q(or_state *s) { return s->q; }
double *or_prims(or_state *s) { return s->prims; }
double *or_qrhs(or_state *s) { return s->qrhs; }
double *or_dvel(or_state *s) { return s->dvel; }
double *or_dtmp(or_state *s) { return s->dtmp; }

/* periodic halo map, bchomovec: var(i)=var(im+i) for i<0, var(i-im) for i>im
 * (tgvsolver.F90:371-440) */
static inline int wrapi(int t, int n) { return t < 0 ? t + n : (t > n ? t - n : t); }

/* q2fvar (tgvsolver.F90:92-121) and thermal_scar (:73-90) */
static inline void q2fvar_pt(const or_state *s, const double *q5, double *r, double *u,
                             double *v, double *w, double *p, double *t) {
  double den = q5[0];
  *r = den;
  *u = q5[1] / den;
  *v = q5[2] / den;
  *w = q5[3] / den;
  *p = (q5[4] - 0.5 * den * ((*u) * (*u) + (*v) * (*v) + (*w) * (*w))) / s->const6;
  *t = (*p) / den * s->const2;
}

void or_q2fvar_all(or_state *s) {
  long N = s->N;
#pragma omp parallel for
  for (long e = 0; e < N; e++) {
    double q5[5];
    for (int m = 0; m < 5; m++) q5[m] = s->q[m * N + e];
    q2fvar_pt(s, q5, &s->prims[e], &s->prims[N + e], &s->prims[2 * N + e],
              &s->prims[3 * N + e], &s->prims[4 * N + e], &s->prims[5 * N + e]);
  }
}

void or_set_q(or_state *s, const double *q) {
  memcpy(s->q, q, sizeof(double) * 5 * s->N);
  or_q2fvar_all(s);
}

/* ---- counter-based Gaussian noise: Philox4x32-10 (Salmon et al., SC'11)
 *      + Box-Muller.  Input generator shared with the device path. ---- */
static void philox4x32_10(uint32_t ctr[4], uint32_t key[2]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * ctr[0], p1 = (uint64_t)M1 * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ ctr[1] ^ k0, n2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

void or_philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {in[0], in[1], in[2], in[3]}, k[2] = {key[0], key[1]};
  philox4x32_10(c, k);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

double or_gauss(uint64_t seed, uint32_t i, uint32_t j, uint32_t k, uint32_t c) {
  uint32_t ctr[4] = {i, j, k, c};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox4x32_10(ctr, key);
  const double two53 = 9007199254740992.0;
  double u1 = ((double)(((uint64_t)(ctr[0] >> 5) << 26) | (uint64_t)(ctr[1] >> 6)) + 0.5) / two53;
  double u2 = ((double)(((uint64_t)(ctr[2] >> 5) << 26) | (uint64_t)(ctr[3] >> 6)) + 0.5) / two53;
  const double twopi = 6.283185307179586476925286766559;
  return sqrt(-2.0 * log(u1)) * cos(twopi * u2);
}

/* TGV initial field, tgvsolver.F90:172-234 (flowfield_init): prims set from
 * the closed form, q from fvar2q (:48-71).  Optional velocity noise (not in
 * the reference) is added on i,j,k<n and copied to the duplicate planes. */
void or_tgv_init(or_state *s, double amp, uint64_t seed) {
  const double pi = 4.0 * atan(1.0);
  int im = s->n[0], jm = s->n[1], km = s->n[2];
  long N = s->N;
  double pinf = 1.0 / s->const2;
#pragma omp parallel for collapse(2)
  for (int k = 0; k <= km; k++)
    for (int j = 0; j <= jm; j++)
      for (int i = 0; i <= im; i++) {
        long e = i + s->st[1] * j + s->st[2] * k;
        double x1 = 2.0 * pi / (double)im * (double)i;
        double x2 = 2.0 * pi / (double)jm * (double)j;
        double x3 = km > 0 ? 2.0 * pi / (double)km * (double)k : 0.0;
        double rho = 1.0;
        double u = sin(x1) * cos(x2) * cos(x3);
        double v = -cos(x1) * sin(x2) * cos(x3);
        double w = 0.0;
        if (amp != 0.0) {
          int ii = i == im ? 0 : i, jj = j == jm ? 0 : j, kk = (km > 0 && k == km) ? 0 : k;
          u += amp * or_gauss(seed, ii, jj, kk, 0);
          v += amp * or_gauss(seed, ii, jj, kk, 1);
          if (km > 0) w += amp * or_gauss(seed, ii, jj, kk, 2);
        }
        double p = pinf + 1.0 / 16.0 * (cos(2.0 * x1) + cos(2.0 * x2)) * (cos(2.0 * x3) + 2.0);
        double t = p / rho * s->const2;
        s->prims[e] = rho; s->prims[N + e] = u; s->prims[2 * N + e] = v;
        s->prims[3 * N + e] = w; s->prims[4 * N + e] = p; s->prims[5 * N + e] = t;
        double var1 = 0.5 * (u * u + v * v + w * w);
        s->q[e] = rho; s->q[N + e] = rho * u; s->q[2 * N + e] = rho * v;
        s->q[3 * N + e] = rho * w; s->q[4 * N + e] = p * s->const6 + rho * var1;
      }
  s->nstep = 0;
}

/* Solve one line: f[t+HM] t=-HM..n+HM -> out[0..n] */
static void line_deriv(const line_sys *sy, const double *f, double *d, double *x,
                       double *out) {
  or_deriv_rhs(sy->n, f, d);
  or_thomas_solve(sy->m, sy->ac, d, x);
  for (int j = 0; j <= sy->n; j++) out[j] = x[j + 1];
}

static void line_filter(const or_state *s, const line_sys *sy, const double *f, double *d,
                        double *x, double *out) {
  or_filter_rhs(sy->n, s->c10, s->ch, f, d);
  or_thomas_solve(sy->m, sy->ac, d, x);
  for (int j = 0; j <= sy->n; j++) out[j] = x[j + 3];
}

/* line base index of line number L in direction dir */
static inline long line_base(const or_state *s, int dir, long L) {
  int n0 = s->n[0] + 1, n1 = s->n[1] + 1;
  if (dir == 0) { long j = L % n1, k = L / n1; return s->st[1] * j + s->st[2] * k; }
  if (dir == 1) { long i = L % n0, k = L / n0; return i + s->st[2] * k; }
  { long i = L % n0, j = L / n0; return i + s->st[1] * j; }
}
static inline long nlines(const or_state *s, int dir) {
  return s->N / (s->n[dir] + 1);
}

#define MAXL 4200

/* gradcal, tgvsolver.F90:450-537 */
static void gradcal(or_state *s) {
  long N = s->N;
  const double dd[3] = {s->c.dx, s->c.dy, s->c.dz};
  for (int dir = 0; dir < s->ndir; dir++) {
    const line_sys *sy = &s->dsys[dir];
    int n = s->n[dir];
    long st = s->st[dir], nl = nlines(s, dir);
#pragma omp parallel
    {
      double f[MAXL], d[MAXL], x[MAXL], df[MAXL];
#pragma omp for
      for (long L = 0; L < nl; L++) {
        long b = line_base(s, dir, L);
        for (int col = 0; col < 4; col++) {
          const double *src = col < 3 ? s->prims + (1 + col) * N : s->prims + 5 * N;
          for (int t = -HM; t <= n + HM; t++) f[t + HM] = src[b + st * wrapi(t, n)];
          line_deriv(sy, f, d, x, df);
          double *dst = col < 3 ? s->dvel + (col + 3 * dir) * N : s->dtmp + dir * N;
          for (int t = 0; t <= n; t++) dst[b + st * t] = df[t] / dd[dir];
        }
      }
    }
  }
  if (s->ndir == 2) {
    for (int m = 0; m < 3; m++) memset(s->dvel + (m + 6) * N, 0, sizeof(double) * N);
    memset(s->dtmp + 2 * N, 0, sizeof(double) * N);
  }
}

/* convection, tgvsolver.F90:539-633.  Halo q is fvar2q of the halo
 * primitives, as bchomo (tgvsolver.F90:271-369) leaves it. */
static void convection(or_state *s) {
  long N = s->N;
  const double dd[3] = {s->c.dx, s->c.dy, s->c.dz};
  for (int dir = 0; dir < s->ndir; dir++) {
    const line_sys *sy = &s->dsys[dir];
    int n = s->n[dir];
    long st = s->st[dir], nl = nlines(s, dir);
#pragma omp parallel
    {
      double fc[5][MAXL], d[MAXL], x[MAXL], df[MAXL];
#pragma omp for
      for (long L = 0; L < nl; L++) {
        long b = line_base(s, dir, L);
        for (int t = -HM; t <= n + HM; t++) {
          long e = b + st * wrapi(t, n);
          double qq[5];
          double rho = s->prims[e], u = s->prims[N + e], v = s->prims[2 * N + e],
                 w = s->prims[3 * N + e], p = s->prims[4 * N + e];
          if (t < 0 || t > n) {
            /* fvar2q with pressure, tgvsolver.F90:59-68 */
            double var1 = 0.5 * (u * u + v * v + w * w);
            qq[0] = rho; qq[1] = rho * u; qq[2] = rho * v; qq[3] = rho * w;
            qq[4] = p * s->const6 + rho * var1;
          } else {
            for (int m = 0; m < 5; m++) qq[m] = s->q[m * N + e];
          }
          double uu = dir == 0 ? u : (dir == 1 ? v : w);
          fc[0][t + HM] = qq[0] * uu;
          fc[1][t + HM] = qq[1] * uu;
          fc[2][t + HM] = qq[2] * uu;
          fc[3][t + HM] = qq[3] * uu;
          fc[1 + dir][t + HM] = qq[1 + dir] * uu + p;
          fc[4][t + HM] = (qq[4] + p) * uu;
        }
        for (int m = 0; m < 5; m++) {
          line_deriv(sy, fc[m], d, x, df);
          double *dst = s->qrhs + m * N;
          for (int t = 0; t <= n; t++) dst[b + st * t] = dst[b + st * t] + df[t] / dd[dir];
        }
      }
    }
  }
}

/* Sutherland, tgvsolver.F90:123-140 */
static inline double miucal(const or_state *s, double t) {
  double tc = s->c.suth, tc1 = 1.0 + tc;
  return t * sqrt(t) * tc1 / (t + tc);
}

/* diffusion, tgvsolver.F90:635-789 */
static void diffusion(or_state *s) {
  long N = s->N;
  const double dd[3] = {s->c.dx, s->c.dy, s->c.dz};
#pragma omp parallel for
  for (long e = 0; e < N; e++) {
    const double *dv = s->dvel;
#define DV(m, n_) dv[((m) + 3 * (n_)) * N + e]
    double s11 = DV(0, 0);
    double s12 = 0.5 * (DV(0, 1) + DV(1, 0));
    double s13 = 0.5 * (DV(0, 2) + DV(2, 0));
    double s22 = DV(1, 1);
    double s23 = 0.5 * (DV(1, 2) + DV(2, 1));
    double s33 = DV(2, 2);
#undef DV
    double skk = num1d3 * (s11 + s22 + s33);
    double miu = miucal(s, s->prims[5 * N + e]) / s->c.reynolds;
    double miu2 = 2.0 * miu;
    double hcc = (miu / s->c.prandtl) / s->const5;
    double *sg = s->sigma;
    sg[e] = miu2 * (s11 - skk);
    sg[N + e] = miu2 * s12;
    sg[2 * N + e] = miu2 * s13;
    sg[3 * N + e] = miu2 * (s22 - skk);
    sg[4 * N + e] = miu2 * s23;
    sg[5 * N + e] = miu2 * (s33 - skk);
    double u = s->prims[N + e], v = s->prims[2 * N + e], w = s->prims[3 * N + e];
    s->qflux[e] = hcc * s->dtmp[e] + sg[e] * u + sg[N + e] * v + sg[2 * N + e] * w;
    s->qflux[N + e] = hcc * s->dtmp[N + e] + sg[N + e] * u + sg[3 * N + e] * v + sg[4 * N + e] * w;
    s->qflux[2 * N + e] =
        hcc * s->dtmp[2 * N + e] + sg[2 * N + e] * u + sg[4 * N + e] * v + sg[5 * N + e] * w;
  }
  /* columns per direction: x (s11,s12,s13,q1), y (s12,s22,s23,q2), z (s13,s23,s33,q3) */
  static const int col[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  for (int dir = 0; dir < s->ndir; dir++) {
    const line_sys *sy = &s->dsys[dir];
    int n = s->n[dir];
    long st = s->st[dir], nl = nlines(s, dir);
#pragma omp parallel
    {
      double f[MAXL], d[MAXL], x[MAXL], df[MAXL];
#pragma omp for
      for (long L = 0; L < nl; L++) {
        long b = line_base(s, dir, L);
        for (int c = 0; c < 4; c++) {
          const double *src = c < 3 ? s->sigma + col[dir][c] * N : s->qflux + dir * N;
          for (int t = -HM; t <= n + HM; t++) f[t + HM] = src[b + st * wrapi(t, n)];
          line_deriv(sy, f, d, x, df);
          double *dst = s->qrhs + (1 + c) * N;
          for (int t = 0; t <= n; t++) dst[b + st * t] = dst[b + st * t] + df[t] / dd[dir];
        }
      }
    }
  }
}

/* rhscal, tgvsolver.F90:962-992 */
void or_rhscal(or_state *s) {
  long N = s->N;
  memset(s->qrhs, 0, sizeof(double) * 5 * N);
  gradcal(s);
  convection(s);
#pragma omp parallel for
  for (long e = 0; e < 5 * N; e++) s->qrhs[e] = -s->qrhs[e];
  diffusion(s);
}

/* filterq, tgvsolver.F90:791-883: x, then y, then z, in place */
void or_filterq(or_state *s) {
  long N = s->N;
  for (int dir = 0; dir < s->ndir; dir++) {
    const line_sys *sy = &s->fsys[dir];
    int n = s->n[dir];
    long st = s->st[dir], nl = nlines(s, dir);
#pragma omp parallel
    {
      double f[MAXL], d[MAXL], x[MAXL], ff[MAXL];
#pragma omp for
      for (long L = 0; L < nl; L++) {
        long b = line_base(s, dir, L);
        for (int m = 0; m < 5; m++) {
          double *qm = s->q + m * N;
          for (int t = -HM; t <= n + HM; t++) f[t + HM] = qm[b + st * wrapi(t, n)];
          line_filter(s, sy, f, d, x, ff);
          for (int t = 0; t <= n; t++) qm[b + st * t] = ff[t];
        }
      }
    }
  }
}

/* stacal, tgvsolver.F90:893-942: sums over i,j,k=1..n.  Summed in long
 * double per k-plane, then planes in order (deterministic). */
void or_stacal(or_state *s, double *ke, double *ens) {
  long N = s->N;
  int im = s->n[0], jm = s->n[1], km = s->n[2];
  int kmax = km > 0 ? km : 0, kmin = km > 0 ? 1 : 0;
  long double tke = 0.0L, enst = 0.0L;
  long double *pk = calloc(kmax + 1, sizeof(long double)), *pe = calloc(kmax + 1, sizeof(long double));
#pragma omp parallel for
  for (int k = kmin; k <= kmax; k++) {
    long double a = 0.0L, b = 0.0L;
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        long e = i + s->st[1] * j + s->st[2] * k;
        double rho = s->prims[e], u = s->prims[N + e], v = s->prims[2 * N + e],
               w = s->prims[3 * N + e];
        double var1 = u * u + v * v + w * w;
        a += rho * var1;
        const double *dv = s->dvel;
        double o1 = dv[(2 + 3 * 1) * N + e] - dv[(1 + 3 * 2) * N + e];
        double o2 = dv[(0 + 3 * 2) * N + e] - dv[(2 + 3 * 0) * N + e];
        double o3 = dv[(1 + 3 * 0) * N + e] - dv[(0 + 3 * 1) * N + e];
        b += rho * (o1 * o1 + o2 * o2 + o3 * o3);
      }
    pk[k] = a; pe[k] = b;
  }
  for (int k = kmin; k <= kmax; k++) { tke += pk[k]; enst += pe[k]; }
  free(pk); free(pe);
  double cnt = (double)im * (double)jm * (double)(km > 0 ? km : 1);
  *ke = (double)(0.5L * tke / cnt);
  *ens = (double)(0.5L * enst / cnt);
}

/* diss_rate_cal, statistic.F90:994-1043: sum over i,j,k=1..n of
 * 2 miu (s11^2+s22^2+s33^2 + 2(s12^2+s13^2+s23^2) - div^2/3), miu =
 * miucal(tmp)/reynolds (the nondimensional branch, :1015-1016), divided by
 * im*jm*km (:1037).  Summed like or_stacal. */
void or_diss_rate(or_state *s, double *diss) {
  long N = s->N;
  int im = s->n[0], jm = s->n[1], km = s->n[2];
  int kmax = km > 0 ? km : 0, kmin = km > 0 ? 1 : 0;
  long double tot = 0.0L;
  long double *pd = calloc(kmax + 1, sizeof(long double));
#pragma omp parallel for
  for (int k = kmin; k <= kmax; k++) {
    long double a = 0.0L;
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        long e = i + s->st[1] * j + s->st[2] * k;
        const double *dv = s->dvel;
#define DV(m, n_) dv[((m) + 3 * (n_)) * N + e]
        double miu = miucal(s, s->prims[5 * N + e]) / s->c.reynolds;
        double s11 = DV(0, 0);
        double s12 = 0.5 * (DV(0, 1) + DV(1, 0));
        double s13 = 0.5 * (DV(0, 2) + DV(2, 0));
        double s22 = DV(1, 1);
        double s23 = 0.5 * (DV(1, 2) + DV(2, 1));
        double s33 = DV(2, 2);
#undef DV
        double div = s11 + s22 + s33;
        a += 2.0 * miu * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23) -
                          num1d3 * (div * div));
      }
    pd[k] = a;
  }
  for (int k = kmin; k <= kmax; k++) tot += pd[k];
  free(pd);
  double cnt = (double)im * (double)jm * (double)(km > 0 ? km : 1);
  *diss = (double)(tot / cnt);
}

/* rk3, tgvsolver.F90:994-1082, with statcal's three statistics of each
 * step's first stage (KE and enstrophy as stacal, the dissipation rate as
 * diss_rate_cal) */
void or_rk3_steps3(or_state *s, int nsteps, double *ke, double *ens, double *diss) {
  static const double rk[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  long N = s->N;
  for (int it = 0; it < nsteps; it++) {
    for (int st = 0; st < 3; st++) {
      or_rhscal(s);
      if (st == 0) {
        double a, b;
        or_stacal(s, &a, &b);
        if (ke) ke[it] = a;
        if (ens) ens[it] = b;
        if (diss) or_diss_rate(s, &diss[it]);
        memcpy(s->qsave, s->q, sizeof(double) * 5 * N);
      }
      const double c0 = rk[st][0], c1 = rk[st][1], c2 = rk[st][2], dt = s->c.dt;
#pragma omp parallel for
      for (long e = 0; e < 5 * N; e++)
        s->q[e] = c0 * s->qsave[e] + c1 * s->q[e] + c2 * s->qrhs[e] * dt;
      or_filterq(s);
      or_q2fvar_all(s);
    }
    s->nstep++;
  }
}

/* rk3, tgvsolver.F90:994-1082 */
void or_rk3_steps(or_state *s, int nsteps, double *ke, double *ens) {
  or_rk3_steps3(s, nsteps, ke, ens, NULL);
}
