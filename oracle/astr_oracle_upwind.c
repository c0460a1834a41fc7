/*
 * astr_oracle_upwind.c -- CPU restatement of the pointwise pieces of the
 * reference's upwind-biased convective path (conschm with an odd first digit
 * and suffix 'e': convrsduwd, src/solver.F90:548-1201):
 *   or_sw_split      flux_steger_warming, src/riemann.F90:23-166 (nondimen,
 *                    no species): the Steger-Warming split with the
 *                    eigenvalue smoothing lmda+ = (lmda + sqrt(lmda^2 +
 *                    eps^2))/2, eps = 0.04
 *   or_recons        recons_exp, src/flux.F90:269-383, periodic branch: the
 *                    interface value from the 8-point window f(1:8) =
 *                    points i-3..i+4 (upwind side) by scheme -1 (f(4)),
 *                    0 SUW7 (:418), 1 WENO7 (:820), 2 WENO7-Z (:977),
 *                    3 MP7 (:498), 5 MP7-LD (:636), 6 ROUND (:717)
 *   or_chardecomp    chardecomp, src/solver.F90:1958-2162 (no species): the
 *                    left and right eigenvectors of the Roe-averaged state
 *                    between two points
 * and of the compact upwind-biased path (suffix 'c': convrsdcmp,
 * src/solver.F90:1271-1937), scheme 543 on periodic lines:
 *   or_cuw_lu        compact_flux_initiate, src/flux.F90:33-118: the
 *                    tridiagonal LHS a, c and its LU factors
 *   or_cuw_flux      flux_compact + compact_flux_rhs, src/flux.F90:125-254:
 *                    the linear 5th-order compact interface flux of a line
 *   or_mp5           MP5, src/flux.F90:434-495, as mplimiter calls it
 * The numerical constants are the reference's own decimal literals (its
 * 16-digit roundings of 13/12, 1/36, ...), so that the restatement and the
 * HIP path round as the reference does.  Expressions keep the reference's
 * left-to-right order; the file is built with -ffp-contract=off.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (astr_amd/, include/)
 * links, loads or calls this file; it is the checker the HIP path's upwind
 * pass is compared against.
 *
 * How it is pinned (tests/test_oracle_upwind.py): the split against the
 * textbook Steger-Warming form written in numpy from the eigen-decomposition
 * A = R diag(lmda) L of the Euler flux Jacobian, F+ + F- = F, and the
 * supersonic limits; SUW7 and MP7 on smooth data against exact polynomial
 * reproduction of the flux difference up to degree 7; WENO7 and WENO7-Z
 * against their design order on a smooth periodic function and their
 * non-oscillation on a step; MP7-LD with its bandwidth factor 0 against the
 * centred 8-point interpolant; ROUND against its normalised-variable limits;
 * chardecomp against LEV REV = I and the Roe-state eigen-relations.  The
 * compact flux against a dense solve of its own system written from the
 * scheme's definition, polynomial exactness of the interface value, its
 * convergence order on a smooth periodic function and the Fourier symbol
 * of the interior rows; MP5 against the linear value on smooth monotone
 * data and against the monotonicity bounds on a step.
 */
#include <math.h>
#include <stdlib.h>

#include "astr_oracle.h"

/* flux_steger_warming at one point (riemann.F90:45-162): q the conservative
 * variables, pv = rho, u, v, w, p, T, dxi the metric row of the direction,
 * jac the Jacobian */
void or_sw_split(double gamma, double mach, double jac, const double *dxi, const double *q,
                 const double *pv, double *fp, double *fm) {
  const double eps = 0.04;
  const double rho = pv[0], u = pv[1], v = pv[2], w = pv[3], prs = pv[4], tmp = pv[5];
  const double uu = dxi[0] * u + dxi[1] * v + dxi[2] * w;
  const double var0 = 1.0 / sqrt(dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2]);
  const double gpd[3] = {dxi[0] * var0, dxi[1] * var0, dxi[2] * var0};
  const double gm2 = 0.5 / gamma;
  const double css = sqrt(tmp) / mach; /* sos, fludyna.F90:850-851 */
  const double csa = css / var0;
  const double lmach = uu / csa;
  const double lmda[5] = {uu, uu, uu, uu + csa, uu - csa};
  double lp[5], lm[5];
  for (int k = 0; k < 5; k++) {
    lp[k] = 0.5 * (lmda[k] + sqrt(lmda[k] * lmda[k] + eps * eps));
    lm[k] = lmda[k] - lp[k];
  }
  double full[5];
  full[0] = jac * q[0] * uu;
  full[1] = jac * (q[1] * uu + dxi[0] * prs);
  full[2] = jac * (q[2] * uu + dxi[1] * prs);
  full[3] = jac * (q[3] * uu + dxi[2] * prs);
  full[4] = jac * (q[4] + prs) * uu;
  if (lmach >= 1.0) {
    for (int m = 0; m < 5; m++) { fp[m] = full[m]; fm[m] = 0.0; }
    return;
  }
  if (lmach <= -1.0) {
    for (int m = 0; m < 5; m++) { fp[m] = 0.0; fm[m] = full[m]; }
    return;
  }
  const double fhi = 0.5 * (gamma - 1.0) * (u * u + v * v + w * w);
  const double jro = jac * rho;
  for (int side = 0; side < 2; side++) {
    const double *l = side == 0 ? lp : lm;
    double *f = side == 0 ? fp : fm;
    const double var1 = l[0];
    const double var2 = l[3] - l[4];
    const double var3 = 2.0 * l[0] - l[3] - l[4];
    const double var4 = var1 - var3 * gm2;
    f[0] = jro * var4;
    f[1] = jro * (var4 * u + var2 * css * gpd[0] * gm2);
    f[2] = jro * (var4 * v + var2 * css * gpd[1] * gm2);
    f[3] = jro * (var4 * w + var2 * css * gpd[2] * gm2);
    f[4] = jac * (var1 * q[4] +
                  rho * (var2 * uu * var0 * css * gm2 - var3 * (fhi + css * css) * gm2 / (gamma - 1.0)));
  }
}

/* minmod2, minmod4: commfunc.F90:699-738 */
static double minmod2(double a, double b) {
  if (a > 0.0 && b > 0.0) return fmin(fabs(a), fabs(b));
  if (a < 0.0 && b < 0.0) return -1.0 * fmin(fabs(a), fabs(b));
  return 0.0;
}
static double minmod4(double a, double b, double c, double d) {
  if (a > 0.0 && b > 0.0 && c > 0.0 && d > 0.0) return fmin(fmin(fmin(fabs(a), fabs(b)), fabs(c)), fabs(d));
  if (a < 0.0 && b < 0.0 && c < 0.0 && d < 0.0)
    return -1.0 * fmin(fmin(fmin(fabs(a), fabs(b)), fabs(c)), fabs(d));
  return 0.0;
}
static double min3(double a, double b, double c) { return fmin(fmin(a, b), c); }
static double max3(double a, double b, double c) { return fmax(fmax(a, b), c); }

/* SUW7 (flux.F90:418-424), u = u(1:7) */
static double suw7(const double *u) {
  return (-3.0 * u[0] + 25.0 * u[1] - 101.0 * u[2] + 319.0 * u[3] + 214.0 * u[4] - 38.0 * u[5] +
          4.0 * u[6]) / 420.0;
}

/* the smoothness indicators and candidate values of WENO7 / WENO7-Z
 * (flux.F90:842-875, 999-1032), u = u(1:7) */
static void weno7_parts(const double *u, double uh[4], double bt[4]) {
  const double *v = u - 1; /* v[1..7] = u(1:7) */
  uh[0] = -0.25 * v[1] + 1.083333333333333 * v[2] - 1.916666666666667 * v[3] + 2.083333333333333 * v[4];
  uh[1] = 8.333333333333333e-2 * v[2] - 4.166666666666667e-1 * v[3] + 1.083333333333333 * v[4] + 0.25 * v[5];
  uh[2] = -8.333333333333333e-2 * v[3] + 5.833333333333333e-1 * v[4] + 5.833333333333333e-1 * v[5] -
          8.333333333333333e-2 * v[6];
  uh[3] = 0.25 * v[4] + 1.083333333333333 * v[5] - 4.166666666666667e-1 * v[6] + 8.333333333333333e-2 * v[7];
  double a, b, c;
  a = -2.0 * v[1] + 9.0 * v[2] - 18.0 * v[3] + 11.0 * v[4];
  b = -1.0 * v[1] + 4.0 * v[2] - 5.0 * v[3] + 2.0 * v[4];
  c = -1.0 * v[1] + 3.0 * v[2] - 3.0 * v[3] + 1.0 * v[4];
  bt[0] = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = v[2] - 6.0 * v[3] + 3.0 * v[4] + 2.0 * v[5];
  b = v[3] - 2.0 * v[4] + v[5];
  c = -1.0 * v[2] + 3.0 * v[3] - 3.0 * v[4] + 1.0 * v[5];
  bt[1] = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = -2.0 * v[3] - 3.0 * v[4] + 6.0 * v[5] - 1.0 * v[6];
  b = v[3] - 2.0 * v[4] + v[5];
  c = -1.0 * v[3] + 3.0 * v[4] - 3.0 * v[5] + 1.0 * v[6];
  bt[2] = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = -11.0 * v[4] + 18.0 * v[5] - 9.0 * v[6] + 2.0 * v[7];
  b = 2.0 * v[4] - 5.0 * v[5] + 4.0 * v[6] - v[7];
  c = -1.0 * v[4] + 3.0 * v[5] - 3.0 * v[6] + 1.0 * v[7];
  bt[3] = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
}

static const double kC7[4] = {2.857142857142857e-2, 3.428571428571429e-1, 5.142857142857143e-1,
                              1.142857142857143e-1};

static double weno7_mix(const double uh[4], const double al[4]) {
  const double s = al[0] + al[1] + al[2] + al[3];
  const double w0 = al[0] / s, w1 = al[1] / s, w2 = al[2] / s, w3 = al[3] / s;
  return w0 * uh[0] + w1 * uh[1] + w2 * uh[2] + w3 * uh[3];
}

/* WENO7 (flux.F90:820-892) */
static double weno7(const double *u) {
  const double eps = 1.0e-6;
  double uh[4], bt[4], al[4];
  weno7_parts(u, uh, bt);
  for (int k = 0; k < 4; k++) al[k] = kC7[k] / ((bt[k] + eps) * (bt[k] + eps));
  return weno7_mix(uh, al);
}

/* WENO7-Z (flux.F90:977-1055) */
static double weno7z(const double *u) {
  const double eps = 1.0e-6;
  double uh[4], bt[4], al[4];
  weno7_parts(u, uh, bt);
  const double tau7 = fabs(bt[3] - bt[0]);
  for (int k = 0; k < 4; k++) al[k] = kC7[k] + kC7[k] * (tau7 / ((bt[k] + eps) * (bt[k] + eps)));
  return weno7_mix(uh, al);
}

/* the MP limiter of MP7 / MP7-LD (flux.F90:515-546, 669-703): u = u(0:5) with
 * u(3) the upwind point, ul the linear interface value */
static double mp_limit(const double *u, double ul) {
  const double dm1 = u[1] - 2.0 * u[2] + u[3];
  const double d0 = u[2] - 2.0 * u[3] + u[4];
  const double d1 = u[3] - 2.0 * u[4] + u[5];
  const double dhm1 = minmod4(4.0 * dm1 - d0, 4.0 * d0 - dm1, dm1, d0);
  const double dh0 = minmod4(4.0 * d0 - d1, 4.0 * d1 - d0, d0, d1);
  const double uUL = u[3] + 4.0 * (u[3] - u[2]);
  const double uAV = 0.5 * (u[3] + u[4]);
  const double uMD = uAV - 0.5 * dh0;
  const double uLC = u[3] + 0.5 * (u[3] - u[2]) + 1.333333333333333 * dhm1;
  const double uMIN = fmax(min3(u[3], u[4], uMD), min3(u[3], uUL, uLC));
  const double uMAX = fmin(max3(u[3], u[4], uMD), max3(u[3], uUL, uLC));
  return ul + minmod2(uMIN - ul, uMAX - ul);
}

/* MP7 (flux.F90:498-551), u = u(0:6) */
static double mp7(const double *u) {
  const double ul = (-3.0 * u[0] + 25.0 * u[1] - 101.0 * u[2] + 319.0 * u[3] + 214.0 * u[4] - 38.0 * u[5] +
                     4.0 * u[6]) / 420.0;
  const double ump = u[3] + minmod2(u[4] - u[3], 4.0 * (u[3] - u[2]));
  if ((ul - u[3]) * (ul - ump) >= 1.0e-10) return mp_limit(u, ul);
  return ul;
}

/* MP7-LD (flux.F90:636-708), u = u(0:7), bandwidth factor bw, shock flag */
static double mp7ld(const double *u, double bw, int shock) {
  const double vadp = 3.571428571428571e-3 * bw;
  const double b0 = 1.0 * vadp - 7.142857142857143e-3;
  const double b1 = -7.0 * vadp + 5.952380952380952e-2;
  const double b2 = 21.0 * vadp - 0.240476190476190;
  const double b3 = -35.0 * vadp + 0.759523809523809;
  const double b4 = 35.0 * vadp + 0.509523809523809;
  const double b5 = -21.0 * vadp - 9.047619047619047e-2;
  const double b6 = 7.0 * vadp + 9.523809523809525e-3;
  const double b7 = -1.0 * vadp;
  const double ul = b0 * u[0] + b1 * u[1] + b2 * u[2] + b3 * u[3] + b4 * u[4] + b5 * u[5] + b6 * u[6] + b7 * u[7];
  return shock ? mp_limit(u, ul) : ul;
}

/* ROUND (flux.F90:717-752), u = u(1:3) */
static double round3(const double *u) {
  const double eps = 1.0e-16;
  const double z = (u[1] - u[0] + eps) / (u[2] - u[0] + eps);
  const double a1 = 1.0 + 12.0 * z * z;
  const double a2 = 1.0 + 5.0 * (z - 1.0) * (z - 1.0);
  const double pl = 1100.0 * (z - 0.05) * (z - 0.05) * (z - 0.05) * (0.47 - z) * (0.47 - z) * (0.47 - z);
  const double pr = 18000.0 * (z - 0.55) * (z - 0.55) * (z - 0.55) * (0.97 - z) * (0.97 - z) * (0.97 - z) *
                    (0.97 - z) * (0.97 - z);
  const double p1 = 0.833333333333333 * z + 0.333333333333333 + fmax(pl, 0.0) + fmax(pr, 0.0);
  const double p2 = 1.5 * z;
  const double p3 = 0.5 * z + 0.5;
  const double w1 = 1.0 / a1 / a1 / a1 / a1;
  const double w2 = 1.0 / a2 / a2 / a2 / a2 / a2 / a2 / a2 / a2;
  const double g = (p1 * (1.0 - w1) + p2 * w1) * (1.0 - w2) + p3 * w2;
  return g * (u[2] - u[0]) + u[0];
}

/* recons_exp's periodic branch (flux.F90:305-335): f = f(1:8); returns NAN
 * for a scheme the reference does not define */
double or_recons(int scheme, const double *f, double bfac, int shock) {
  switch (scheme) {
    case -1: return f[3];
    case 0: return suw7(f);
    case 1: return weno7(f);
    case 2: return weno7z(f);
    case 3: return mp7(f);
    case 5: return mp7ld(f, bfac, shock);
    case 6: return round3(f + 2);
    default: return NAN;
  }
}

/* MP5 (flux.F90:434-495) as mplimiter calls it for a periodic direction
 * (flux.F90:363-379): u = u(1:5) with u(3) the upwind point, ul the compact
 * scheme's linear interface value, discont the shock flag */
double or_mp5(const double *u, double ul, int discont) {
  const double ump = u[2] + minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
  if (discont && (ul - u[2]) * (ul - ump) >= 1.0e-10) {
    const double w[6] = {0.0, u[0], u[1], u[2], u[3], u[4]};  /* w(1:5) = u(1:5) */
    return mp_limit(w, ul);
  }
  return ul;
}

/* compact_flux_initiate (flux.F90:33-118) for scheme 543 on a periodic line
 * (ntype 3): rows i_0 = -2 .. i_m = n+1, a(:) = 0.5 - bfac/6 and
 * c(:) = 1/6 + bfac/6 for the upwind ('+') flux, the two swapped for '-',
 * a = c = 0 on both end rows; ac[3 r + 0..2] = ac(1..3, r+1), r = 0..n+3,
 * by tridiagonal_thomas_proprocess (commfunc.F90:752-774).  Row 0 holds
 * ac(1,1) = c(1) only (the solver never reads its ac(2:3, 1)). */
void or_cuw_lu(int n, double bfac, int plus, double *ac) {
  const double one6 = 1.0 / 6.0;
  const double av = plus ? 0.5 - one6 * bfac : one6 + one6 * bfac;
  const double cv = plus ? one6 + one6 * bfac : 0.5 - one6 * bfac;
  const int N = n + 4;
  for (int r = 0; r < N; r++) {
    const int end = r == 0 || r == N - 1;
    const double a = end ? 0.0 : av, c = end ? 0.0 : cv;
    if (r == 0) {
      ac[0] = c;
      ac[1] = 1.0;
      ac[2] = 0.0;
      continue;
    }
    ac[3 * r] = c / (1.0 - a * ac[3 * (r - 1)]);
    ac[3 * r + 1] = 1.0 / (1.0 - a * ac[3 * (r - 1)]);
    ac[3 * r + 2] = a / (1.0 - a * ac[3 * (r - 1)]);
  }
}

/* flux_compact (flux.F90:125-150) with compact_flux_rhs (flux.F90:163-254),
 * periodic line: f = f(-4:n+4) at f[t + 4], fh = fh(-1:n) at fh[i + 1].
 * End rows -2 and n+1: the explicit 6th-order interface value; rows -1..n:
 * the 5th-order compact upwind right-hand side; then the Thomas sweeps of
 * tridiagonal_thomas_solver (commfunc.F90:790-813). */
void or_cuw_flux(int n, double bfac, int plus, const double *ac, const double *f, double *fh) {
  const int N = n + 4;
  double *d = malloc(sizeof(double) * 2 * N), *x = d + N;
  const double c37 = 37.0 / 60.0, c2 = 2.0 / 15.0, c1 = 1.0 / 60.0;
  const double k118 = 1.0 / 18.0, k136 = 1.0 / 36.0, k1918 = 19.0 / 18.0, k59 = 5.0 / 9.0, k936 = 9.0 / 36.0;
#define F(j) f[(j) + 4]
  for (int r = 0; r < N; r++) {
    const int j = r - 2;
    if (r == 0 || r == N - 1) {
      const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
      d[r] = c37 * v1 - c2 * v2 + c1 * v3;
    } else if (plus) {
      d[r] = (k118 - k136 * bfac) * F(j - 1) + (k1918 - k936 * bfac) * F(j) + (k59 + k936 * bfac) * F(j + 1) +
             k136 * bfac * F(j + 2);
    } else {
      d[r] = (k118 - k136 * bfac) * F(j + 2) + (k1918 - k936 * bfac) * F(j + 1) + (k59 + k936 * bfac) * F(j) +
             k136 * bfac * F(j - 1);
    }
  }
#undef F
  for (int r = 1; r < N; r++) d[r] = d[r] * ac[3 * r + 1] - d[r - 1] * ac[3 * r + 2];
  x[N - 1] = d[N - 1];
  for (int r = N - 2; r >= 0; r--) x[r] = d[r] - ac[3 * r] * x[r + 1];
  for (int i = -1; i <= n; i++) fh[i + 1] = x[i + 2];
  free(d);
}

/* chardecomp (solver.F90:1958-2162, no species): l, r = rho, u, v, w, p,
 * E (= q5), dxi row (3) of the two points; rev, lev row-major 5 x 5
 * (rev[5 a + b] = REV(a+1, b+1)).  Returns -1 when the averaged metric row
 * is zero (the reference stops). */
int or_chardecomp(double gamma, const double *l, const double *r, double *rev, double *lev) {
  const double rero = 1.0e-12;
#define LEV(a, b) lev[5 * ((a)-1) + (b)-1]
#define REV(a, b) rev[5 * ((a)-1) + (b)-1]
  const double WRoe = sqrt(l[0]) / (sqrt(l[0]) + sqrt(r[0]));
  const double WRoe1 = 1.0 - WRoe;
  const double u1 = WRoe * l[1] + WRoe1 * r[1];
  const double u2 = WRoe * l[2] + WRoe1 * r[2];
  const double u3 = WRoe * l[3] + WRoe1 * r[3];
  const double KRoe = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
  const double HL = (l[5] + l[4]) / l[0];
  const double HR = (r[5] + r[4]) / r[0];
  const double HRoe = WRoe * HL + WRoe1 * HR;
  const double Css = sqrt((gamma - 1.0) * (HRoe - KRoe));
  const double rcs = 1.0 / Css;
  const double var1 = 0.5 * (l[6] + r[6]);
  const double var2 = 0.5 * (l[7] + r[7]);
  const double var3 = 0.5 * (l[8] + r[8]);
  const double var4 = 1.0 / sqrt(var1 * var1 + var2 * var2 + var3 * var3);
  const double g1 = var1 * var4, g2 = var2 * var4, g3 = var3 * var4;
  const double ugp = u1 * g1 + u2 * g2 + u3 * g3;
  const double b1 = (gamma - 1.0) / (Css * Css);
  const double b2 = 1.0 + 2.0 * b1 * KRoe - b1 * HRoe;
  LEV(1, 1) = 0.5 * (b2 + ugp * rcs);
  LEV(1, 2) = -0.5 * (b1 * u1 + g1 * rcs);
  LEV(1, 3) = -0.5 * (b1 * u2 + g2 * rcs);
  LEV(1, 4) = -0.5 * (b1 * u3 + g3 * rcs);
  LEV(1, 5) = 0.5 * b1;
  REV(1, 1) = 1.0;
  REV(2, 1) = u1 - Css * g1;
  REV(3, 1) = u2 - Css * g2;
  REV(4, 1) = u3 - Css * g3;
  REV(5, 1) = HRoe - ugp * Css;
  LEV(2, 1) = 1.0 - b2;
  LEV(2, 2) = b1 * u1;
  LEV(2, 3) = b1 * u2;
  LEV(2, 4) = b1 * u3;
  LEV(2, 5) = -b1;
  REV(1, 2) = 1.0;
  REV(2, 2) = u1;
  REV(3, 2) = u2;
  REV(4, 2) = u3;
  REV(5, 2) = HRoe - 1.0 / b1;
  if (fabs(var1) > rero) {
    const double rgp = 1.0 / g1;
    LEV(3, 1) = (ugp * g2 - u2) * rgp;
    LEV(3, 2) = -g2;
    LEV(3, 3) = (1.0 - g2 * g2) * rgp;
    LEV(3, 4) = -(g2 * g3 * rgp);
    LEV(3, 5) = 0.0;
    LEV(4, 1) = (ugp * g3 - u3) * rgp;
    LEV(4, 2) = -g3;
    LEV(4, 3) = -(g2 * g3 * rgp);
    LEV(4, 4) = (1.0 - g3 * g3) * rgp;
    LEV(4, 5) = 0.0;
    REV(1, 3) = 0.0; REV(2, 3) = -g2; REV(3, 3) = g1; REV(4, 3) = 0.0; REV(5, 3) = u2 * g1 - u1 * g2;
    REV(1, 4) = 0.0; REV(2, 4) = -g3; REV(3, 4) = 0.0; REV(4, 4) = g1; REV(5, 4) = u3 * g1 - u1 * g3;
  } else if (fabs(var2) > rero) {
    const double rgp = 1.0 / g2;
    LEV(3, 1) = (ugp * g1 - u1) * rgp;
    LEV(3, 2) = (1.0 - g1 * g1) * rgp;
    LEV(3, 3) = -g1;
    LEV(3, 4) = -(g1 * g3 * rgp);
    LEV(3, 5) = 0.0;
    LEV(4, 1) = (ugp * g3 - u3) * rgp;
    LEV(4, 2) = -(g1 * g3 * rgp);
    LEV(4, 3) = -g3;
    LEV(4, 4) = (1.0 - g3 * g3) * rgp;
    LEV(4, 5) = 0.0;
    REV(1, 3) = 0.0; REV(2, 3) = g2; REV(3, 3) = -g1; REV(4, 3) = 0.0; REV(5, 3) = u1 * g2 - u2 * g1;
    REV(1, 4) = 0.0; REV(2, 4) = 0.0; REV(3, 4) = -g3; REV(4, 4) = g2; REV(5, 4) = u3 * g2 - u2 * g3;
  } else if (fabs(var3) > rero) {
    const double rgp = 1.0 / g3;
    LEV(3, 1) = (ugp * g1 - u1) * rgp;
    LEV(3, 2) = (1.0 - g1 * g1) * rgp;
    LEV(3, 3) = -(g1 * g2 * rgp);
    LEV(3, 4) = -g1;
    LEV(3, 5) = 0.0;
    LEV(4, 1) = (ugp * g2 - u2) * rgp;
    LEV(4, 2) = -(g1 * g2 * rgp);
    LEV(4, 3) = (1.0 - g2 * g2) * rgp;
    LEV(4, 4) = -g2;
    LEV(4, 5) = 0.0;
    REV(1, 3) = 0.0; REV(2, 3) = g3; REV(3, 3) = 0.0; REV(4, 3) = -g1; REV(5, 3) = u1 * g3 - u3 * g1;
    REV(1, 4) = 0.0; REV(2, 4) = 0.0; REV(3, 4) = g3; REV(4, 4) = -g2; REV(5, 4) = u2 * g3 - u3 * g2;
  } else {
    return -1;
  }
  LEV(5, 1) = 0.5 * (b2 - ugp * rcs);
  LEV(5, 2) = -0.5 * (b1 * u1 - g1 * rcs);
  LEV(5, 3) = -0.5 * (b1 * u2 - g2 * rcs);
  LEV(5, 4) = -0.5 * (b1 * u3 - g3 * rcs);
  LEV(5, 5) = 0.5 * b1;
  REV(1, 5) = 1.0;
  REV(2, 5) = u1 + Css * g1;
  REV(3, 5) = u2 + Css * g2;
  REV(4, 5) = u3 + Css * g3;
  REV(5, 5) = HRoe + ugp * Css;
#undef LEV
#undef REV
  return 0;
}
