// upwind.hip -- the reference's explicit upwind-biased convective term
// (convrsduwd, src/solver.F90:548-1201) and its shock sensor (ducrossensor,
// src/commcal.F90:196-360) on gfx950, for the main path (stored primitives,
// qswap's halo sources) with periodic directions on one rank.
//
// Per direction two kernels:
//   k_uw_split   the Steger-Warming split (flux_steger_warming,
//                src/riemann.F90:23-166) of every point of every line,
//                rows -4..n+4 (the halo rows from qswap's halo q and the
//                primitives q2fvar gives it), into a split buffer laid out
//                so that the next kernel reads it coalesced: rows
//                contiguous for x-lines, lines contiguous for y- and
//                z-lines;
//   k_uw_face_*  per interface i = -1..n the 8-point windows of F+ and F-
//                (iwind8, solver.F90:1213-1234), optionally projected on
//                the Roe-state characteristics (chardecomp,
//                solver.F90:1958-2162), reconstructed (recons_exp,
//                src/flux.F90:269-383) and summed, then qrhs(i) =
//                qrhs(i) + Fh(i) - Fh(i-1).  x-lines: one lane per
//                interface, 63 points per wavefront, Fh(i-1) from the
//                neighbouring lane; y- and z-lines: one lane per line and
//                a run of 16 points, Fh(i-1) carried in registers.
// The sensor: k_ssf (the Ducros ratio from the stored velocity gradient and
// the pressure with its halo values) and k_lshock (its maximum over the
// cross of offsets -4..5 against shkcrt).
//
// Expressions keep the reference's left-to-right order and its decimal
// constants, without FMA contraction.
#include <hip/hip_runtime.h>

#include "upwind.hpp"

#pragma clang fp contract(off)

namespace astr {
namespace {

constexpr int kRun = 16;  // points per lane of a y- or z-line

// ---------------------------------------------------------------------------
// pointwise pieces
// ---------------------------------------------------------------------------

// flux_steger_warming at one point (riemann.F90:45-162, nondimen).  The
// branch results are chosen value by value so that fp and fm stay in
// registers.
__device__ __forceinline__ void sw_split(double gamma, double mach, double jac, const double dx[3],
                                         const double q[5], const double pv[6], double fp[5],
                                         double fm[5]) {
  const double eps = 0.04;
  const double rho = pv[0], u = pv[1], v = pv[2], w = pv[3], prs = pv[4], tmp = pv[5];
  const double uu = dx[0] * u + dx[1] * v + dx[2] * w;
  const double var0 = 1.0 / sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
  const double gpd0 = dx[0] * var0, gpd1 = dx[1] * var0, gpd2 = dx[2] * var0;
  const double gm2 = 0.5 / gamma;
  const double css = sqrt(tmp) / mach;
  const double csa = css / var0;
  const double lmach = uu / csa;
  const bool sup = lmach >= 1.0, sdown = lmach <= -1.0;
  if (sup || sdown) {
    double full[5];
    full[0] = jac * q[0] * uu;
    full[1] = jac * (q[1] * uu + dx[0] * prs);
    full[2] = jac * (q[2] * uu + dx[1] * prs);
    full[3] = jac * (q[3] * uu + dx[2] * prs);
    full[4] = jac * (q[4] + prs) * uu;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      fp[m] = sup ? full[m] : 0.0;
      fm[m] = sup ? 0.0 : full[m];
    }
    return;
  }
  const double l0 = uu, l3 = uu + csa, l4 = uu - csa;
  const double p0 = 0.5 * (l0 + sqrt(l0 * l0 + eps * eps));
  const double p3 = 0.5 * (l3 + sqrt(l3 * l3 + eps * eps));
  const double p4 = 0.5 * (l4 + sqrt(l4 * l4 + eps * eps));
  const double fhi = 0.5 * (gamma - 1.0) * (u * u + v * v + w * w);
  const double jro = jac * rho;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const double e0 = side == 0 ? p0 : l0 - p0;
    const double e3 = side == 0 ? p3 : l3 - p3;
    const double e4 = side == 0 ? p4 : l4 - p4;
    const double var1 = e0;
    const double var2 = e3 - e4;
    const double var3 = 2.0 * e0 - e3 - e4;
    const double var4 = var1 - var3 * gm2;
    const double f0 = jro * var4;
    const double f1 = jro * (var4 * u + var2 * css * gpd0 * gm2);
    const double f2 = jro * (var4 * v + var2 * css * gpd1 * gm2);
    const double f3 = jro * (var4 * w + var2 * css * gpd2 * gm2);
    const double f4 = jac * (var1 * q[4] + rho * (var2 * uu * var0 * css * gm2 -
                                                  var3 * (fhi + css * css) * gm2 / (gamma - 1.0)));
    if (side == 0) { fp[0] = f0; fp[1] = f1; fp[2] = f2; fp[3] = f3; fp[4] = f4; }
    else { fm[0] = f0; fm[1] = f1; fm[2] = f2; fm[3] = f3; fm[4] = f4; }
  }
}

// minmod2, minmod4 (commfunc.F90:699-738)
__device__ __forceinline__ double minmod2(double a, double b) {
  if (a > 0.0 && b > 0.0) return fmin(fabs(a), fabs(b));
  if (a < 0.0 && b < 0.0) return -1.0 * fmin(fabs(a), fabs(b));
  return 0.0;
}
__device__ __forceinline__ double minmod4(double a, double b, double c, double d) {
  if (a > 0.0 && b > 0.0 && c > 0.0 && d > 0.0) return fmin(fmin(fmin(fabs(a), fabs(b)), fabs(c)), fabs(d));
  if (a < 0.0 && b < 0.0 && c < 0.0 && d < 0.0)
    return -1.0 * fmin(fmin(fmin(fabs(a), fabs(b)), fabs(c)), fabs(d));
  return 0.0;
}

// SUW7 (flux.F90:418-424) on u(1:7) = f[0..6]
__device__ __forceinline__ double suw7(const double* f) {
  return (-3.0 * f[0] + 25.0 * f[1] - 101.0 * f[2] + 319.0 * f[3] + 214.0 * f[4] - 38.0 * f[5] +
          4.0 * f[6]) / 420.0;
}

// WENO7 (flux.F90:820-892) and WENO7-Z (:977-1055) on u(1:7) = f[0..6]
__device__ __forceinline__ double weno7(const double* f, bool z) {
  const double eps = 1.0e-6;
  const double u1 = f[0], u2 = f[1], u3 = f[2], u4 = f[3], u5 = f[4], u6 = f[5], u7 = f[6];
  const double uh1 = -0.25 * u1 + 1.083333333333333 * u2 - 1.916666666666667 * u3 + 2.083333333333333 * u4;
  const double uh2 = 8.333333333333333e-2 * u2 - 4.166666666666667e-1 * u3 + 1.083333333333333 * u4 + 0.25 * u5;
  const double uh3 = -8.333333333333333e-2 * u3 + 5.833333333333333e-1 * u4 + 5.833333333333333e-1 * u5 -
                     8.333333333333333e-2 * u6;
  const double uh4 = 0.25 * u4 + 1.083333333333333 * u5 - 4.166666666666667e-1 * u6 + 8.333333333333333e-2 * u7;
  double a, b, c;
  a = -2.0 * u1 + 9.0 * u2 - 18.0 * u3 + 11.0 * u4;
  b = -1.0 * u1 + 4.0 * u2 - 5.0 * u3 + 2.0 * u4;
  c = -1.0 * u1 + 3.0 * u2 - 3.0 * u3 + 1.0 * u4;
  const double bt1 = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = u2 - 6.0 * u3 + 3.0 * u4 + 2.0 * u5;
  b = u3 - 2.0 * u4 + u5;
  c = -1.0 * u2 + 3.0 * u3 - 3.0 * u4 + 1.0 * u5;
  const double bt2 = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = -2.0 * u3 - 3.0 * u4 + 6.0 * u5 - 1.0 * u6;
  b = u3 - 2.0 * u4 + u5;
  c = -1.0 * u3 + 3.0 * u4 - 3.0 * u5 + 1.0 * u6;
  const double bt3 = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  a = -11.0 * u4 + 18.0 * u5 - 9.0 * u6 + 2.0 * u7;
  b = 2.0 * u4 - 5.0 * u5 + 4.0 * u6 - u7;
  c = -1.0 * u4 + 3.0 * u5 - 3.0 * u6 + 1.0 * u7;
  const double bt4 = 2.777777777777778e-2 * (a * a) + 1.083333333333333 * (b * b) + 1.084722222222222 * (c * c);
  const double C1 = 2.857142857142857e-2, C2 = 3.428571428571429e-1, C3 = 5.142857142857143e-1,
               C4 = 1.142857142857143e-1;
  double a1, a2, a3, a4;
  if (z) {
    const double tau7 = fabs(bt4 - bt1);
    a1 = C1 + C1 * (tau7 / ((bt1 + eps) * (bt1 + eps)));
    a2 = C2 + C2 * (tau7 / ((bt2 + eps) * (bt2 + eps)));
    a3 = C3 + C3 * (tau7 / ((bt3 + eps) * (bt3 + eps)));
    a4 = C4 + C4 * (tau7 / ((bt4 + eps) * (bt4 + eps)));
  } else {
    a1 = C1 / ((bt1 + eps) * (bt1 + eps));
    a2 = C2 / ((bt2 + eps) * (bt2 + eps));
    a3 = C3 / ((bt3 + eps) * (bt3 + eps));
    a4 = C4 / ((bt4 + eps) * (bt4 + eps));
  }
  const double s = a1 + a2 + a3 + a4;
  const double w1 = a1 / s, w2 = a2 / s, w3 = a3 / s, w4 = a4 / s;
  return w1 * uh1 + w2 * uh2 + w3 * uh3 + w4 * uh4;
}

// the MP limiter of MP7 / MP7-LD (flux.F90:515-546, 669-703), u(0:5) = f[0..5]
__device__ __forceinline__ double mp_limit(const double* u, double ul) {
  const double dm1 = u[1] - 2.0 * u[2] + u[3];
  const double d0 = u[2] - 2.0 * u[3] + u[4];
  const double d1 = u[3] - 2.0 * u[4] + u[5];
  const double dhm1 = minmod4(4.0 * dm1 - d0, 4.0 * d0 - dm1, dm1, d0);
  const double dh0 = minmod4(4.0 * d0 - d1, 4.0 * d1 - d0, d0, d1);
  const double uUL = u[3] + 4.0 * (u[3] - u[2]);
  const double uAV = 0.5 * (u[3] + u[4]);
  const double uMD = uAV - 0.5 * dh0;
  const double uLC = u[3] + 0.5 * (u[3] - u[2]) + 1.333333333333333 * dhm1;
  const double uMIN = fmax(fmin(fmin(u[3], u[4]), uMD), fmin(fmin(u[3], uUL), uLC));
  const double uMAX = fmin(fmax(fmax(u[3], u[4]), uMD), fmax(fmax(u[3], uUL), uLC));
  return ul + minmod2(uMIN - ul, uMAX - ul);
}

// MP7 (flux.F90:498-551), u(0:6) = f[0..6]
__device__ __forceinline__ double mp7(const double* u) {
  const double ul = suw7(u);
  const double ump = u[3] + minmod2(u[4] - u[3], 4.0 * (u[3] - u[2]));
  if ((ul - u[3]) * (ul - ump) >= 1.0e-10) return mp_limit(u, ul);
  return ul;
}

// MP7-LD (flux.F90:636-708), u(0:7) = f[0..7]
__device__ __forceinline__ double mp7ld(const double* u, double bw, bool shock) {
  const double vadp = 3.571428571428571e-3 * bw;
  const double b0 = 1.0 * vadp - 7.142857142857143e-3;
  const double b1 = -7.0 * vadp + 5.952380952380952e-2;
  const double b2 = 21.0 * vadp - 0.240476190476190;
  const double b3 = -35.0 * vadp + 0.759523809523809;
  const double b4 = 35.0 * vadp + 0.509523809523809;
  const double b5 = -21.0 * vadp - 9.047619047619047e-2;
  const double b6 = 7.0 * vadp + 9.523809523809525e-3;
  const double b7 = -1.0 * vadp;
  const double ul =
      b0 * u[0] + b1 * u[1] + b2 * u[2] + b3 * u[3] + b4 * u[4] + b5 * u[5] + b6 * u[6] + b7 * u[7];
  return shock ? mp_limit(u, ul) : ul;
}

// ROUND (flux.F90:717-752), u(1:3) = f[2..4]
__device__ __forceinline__ double round3(const double* f) {
  const double eps = 1.0e-16;
  const double u1 = f[2], u2 = f[3], u3 = f[4];
  const double z = (u2 - u1 + eps) / (u3 - u1 + eps);
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
  return g * (u3 - u1) + u1;
}

// recons_exp's periodic branch (flux.F90:305-335) for scheme SCH (checked
// on the host), one instantiation per scheme so that each kernel holds one
// scheme's code
template <int SCH>
__device__ __forceinline__ double recons(const double* f, double bfac, bool shock) {
  if constexpr (SCH == -1) return f[3];
  else if constexpr (SCH == 0) return suw7(f);
  else if constexpr (SCH == 1) return weno7(f, false);
  else if constexpr (SCH == 2) return weno7(f, true);
  else if constexpr (SCH == 3) return mp7(f);
  else if constexpr (SCH == 5) return mp7ld(f, bfac, shock);
  else return round3(f);
}

// chardecomp (solver.F90:1958-2162, no species); l, r = rho, u, v, w, p, E,
// dxi row; rev, lev row-major 5 x 5.  A zero averaged metric row (where the
// reference stops) cannot occur on the grids the host accepts.
__device__ __forceinline__ void chardecomp(double gamma, const double* l, const double* r, double* rev,
                                           double* lev) {
  const double rero = 1.0e-12;
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
  lev[0] = 0.5 * (b2 + ugp * rcs);
  lev[1] = -0.5 * (b1 * u1 + g1 * rcs);
  lev[2] = -0.5 * (b1 * u2 + g2 * rcs);
  lev[3] = -0.5 * (b1 * u3 + g3 * rcs);
  lev[4] = 0.5 * b1;
  lev[5] = 1.0 - b2;
  lev[6] = b1 * u1;
  lev[7] = b1 * u2;
  lev[8] = b1 * u3;
  lev[9] = -b1;
  rev[0] = 1.0; rev[5] = u1 - Css * g1; rev[10] = u2 - Css * g2; rev[15] = u3 - Css * g3;
  rev[20] = HRoe - ugp * Css;
  rev[1] = 1.0; rev[6] = u1; rev[11] = u2; rev[16] = u3; rev[21] = HRoe - 1.0 / b1;
  if (fabs(var1) > rero) {
    const double rgp = 1.0 / g1;
    lev[10] = (ugp * g2 - u2) * rgp; lev[11] = -g2; lev[12] = (1.0 - g2 * g2) * rgp;
    lev[13] = -(g2 * g3 * rgp); lev[14] = 0.0;
    lev[15] = (ugp * g3 - u3) * rgp; lev[16] = -g3; lev[17] = -(g2 * g3 * rgp);
    lev[18] = (1.0 - g3 * g3) * rgp; lev[19] = 0.0;
    rev[2] = 0.0; rev[7] = -g2; rev[12] = g1; rev[17] = 0.0; rev[22] = u2 * g1 - u1 * g2;
    rev[3] = 0.0; rev[8] = -g3; rev[13] = 0.0; rev[18] = g1; rev[23] = u3 * g1 - u1 * g3;
  } else if (fabs(var2) > rero) {
    const double rgp = 1.0 / g2;
    lev[10] = (ugp * g1 - u1) * rgp; lev[11] = (1.0 - g1 * g1) * rgp; lev[12] = -g1;
    lev[13] = -(g1 * g3 * rgp); lev[14] = 0.0;
    lev[15] = (ugp * g3 - u3) * rgp; lev[16] = -(g1 * g3 * rgp); lev[17] = -g3;
    lev[18] = (1.0 - g3 * g3) * rgp; lev[19] = 0.0;
    rev[2] = 0.0; rev[7] = g2; rev[12] = -g1; rev[17] = 0.0; rev[22] = u1 * g2 - u2 * g1;
    rev[3] = 0.0; rev[8] = 0.0; rev[13] = -g3; rev[18] = g2; rev[23] = u3 * g2 - u2 * g3;
  } else {
    const double rgp = 1.0 / g3;
    lev[10] = (ugp * g1 - u1) * rgp; lev[11] = (1.0 - g1 * g1) * rgp; lev[12] = -(g1 * g2 * rgp);
    lev[13] = -g1; lev[14] = 0.0;
    lev[15] = (ugp * g2 - u2) * rgp; lev[16] = -(g1 * g2 * rgp); lev[17] = (1.0 - g2 * g2) * rgp;
    lev[18] = -g2; lev[19] = 0.0;
    rev[2] = 0.0; rev[7] = g3; rev[12] = 0.0; rev[17] = -g1; rev[22] = u1 * g3 - u3 * g1;
    rev[3] = 0.0; rev[8] = 0.0; rev[13] = g3; rev[18] = -g2; rev[23] = u2 * g3 - u3 * g2;
  }
  lev[20] = 0.5 * (b2 - ugp * rcs);
  lev[21] = -0.5 * (b1 * u1 - g1 * rcs);
  lev[22] = -0.5 * (b1 * u2 - g2 * rcs);
  lev[23] = -0.5 * (b1 * u3 - g3 * rcs);
  lev[24] = 0.5 * b1;
  rev[4] = 1.0; rev[9] = u1 + Css * g1; rev[14] = u2 + Css * g2; rev[19] = u3 + Css * g3;
  rev[24] = HRoe + ugp * Css;
}

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------

// offset of the point of line L of direction dir at row r (0..n)
__device__ __forceinline__ long long line_base(const FlowArgs& a, int dir, long long L, int* c3) {
  if (dir == 0) {
    c3[1] = (int)(L % (a.n[1] + 1));
    c3[2] = (int)(L / (a.n[1] + 1));
    c3[0] = 0;
  } else {
    c3[0] = (int)(L % (a.n[0] + 1));
    const int o = (int)(L / (a.n[0] + 1));
    c3[1] = dir == 1 ? 0 : o;
    c3[2] = dir == 1 ? o : 0;
  }
  return (long long)c3[2] * a.plane + (long long)c3[1] * a.pitch + c3[0];
}
__device__ __forceinline__ long long dir_stride(const FlowArgs& a, int dir) {
  return dir == 0 ? 1 : (dir == 1 ? (long long)a.pitch : a.plane);
}
// the halo-source row index of the periodic copy of row rs (1..4 or n-4..n-1)
__device__ __forceinline__ int hrow_of(int rs, int n) { return rs <= 5 ? rs - 1 : rs - (n - 5) + 5; }

// q of the halo copy (row r outside 0..n) of line (c3) of direction dir:
// the kept halo source where qswap filled the halo before a later step
// averaged the source (x-lines on j or k boundary planes, y-lines on k
// boundary planes), the current q of the source point otherwise
__device__ __forceinline__ void halo_q_at(const FlowArgs& a, int dir, const int* c3, int rs, long long os,
                                          double* q) {
  const int ex = dir < 2 ? halo_src_index(a.n, a.ndir, dir, c3[0], c3[1], c3[2], hrow_of(rs, a.n[dir])) : -1;
  if (ex >= 0) {
    const double* src = dir == 0 ? a.hx : a.hy;
    const long long hs = dir == 0 ? a.hxs : a.hys;
#pragma unroll
    for (int m = 0; m < 5; m++) q[m] = src[m * hs + ex];
  } else {
#pragma unroll
    for (int m = 0; m < 5; m++) q[m] = a.q[m * a.fs + os];
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// The Steger-Warming split of row rr (0..n+8, the line's row rr-4) of line
// L of direction DIR: F+ in out[0..4], F- in out[5..9] and, for the
// characteristic projection, the state the face kernels read in
// out[10..18] (rho, u, v, w, p, the energy q_4 and the metric row)
template <int DIR>
__device__ __forceinline__ void split_point(const FlowArgs& a, const UwArgs& u, long long L, int rr,
                                            double (&out)[kUwFields + 9]) {
  constexpr int dir = DIR;
  const int n = a.n[dir];
  const int r = rr - 4;
  int c3[3];
  const long long o0 = line_base(a, dir, L, c3);
  const bool halo = r < 0 || r > n;
  const int rs = r < 0 ? r + n : (r > n ? r - n : r);
  const long long o = o0 + rs * dir_stride(a, dir);
  const long long fs = a.fs;
  double q[5], pv[6];
  if (halo) {
    halo_q_at(a, dir, c3, rs, o, q);
    const Prim P = prim_of(a, q[0], q[1], q[2], q[3], q[4]);
    pv[0] = P.r; pv[1] = P.u; pv[2] = P.v; pv[3] = P.w; pv[4] = P.p; pv[5] = P.t;
  } else {
#pragma unroll
    for (int m = 0; m < 5; m++) q[m] = a.q[m * fs + o];
#pragma unroll
    for (int m = 0; m < 6; m++) pv[m] = a.pr[m * fs + o];
  }
  // a periodic halo copy carries its source point's metrics
  double dx[3], jb;
  if (a.met) {
    jb = a.met[o];
#pragma unroll
    for (int nn = 0; nn < 3; nn++) dx[nn] = a.met[(1 + dir + 3 * nn) * fs + o];
  } else {
    jb = a.jac;
#pragma unroll
    for (int nn = 0; nn < 3; nn++) dx[nn] = nn == dir ? a.inv_d[dir] : 0.0;
  }
  double fp[5], fm[5];
  sw_split(u.gamma, u.mach, jb, dx, q, pv, fp, fm);
#pragma unroll
  for (int m = 0; m < 5; m++) {
    out[m] = fp[m];
    out[5 + m] = fm[m];
  }
  out[10] = pv[0]; out[11] = pv[1]; out[12] = pv[2]; out[13] = pv[3]; out[14] = pv[4];
  out[15] = q[4]; out[16] = dx[0]; out[17] = dx[1]; out[18] = dx[2];
}

template <int DIR>
__global__ __launch_bounds__(256) void k_uw_split(FlowArgs a, UwArgs u) {
  constexpr int dir = DIR;
  const int n = a.n[dir];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, dir);
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= NL * NR) return;
  long long L;
  int rr;
  if (dir == 0) {
    // x-lines row-fastest: for k_uw_face_x's lanes, and for k_cuw_line,
    // which loads whole lines (the compact scheme's sweeps take
    // k_uw_split_xt's order instead)
    L = idx / NR;
    rr = (int)(idx - L * NR);
  } else {
    // y and z: lines fastest
    rr = (int)(idx / NL);
    L = idx - (long long)rr * NL;
  }
  double out[kUwFields + 9];
  split_point<DIR>(a, u, L, rr, out);
  double* sw = u.sw + idx;
  const int nf = u.chardec ? kUwFields + 9 : kUwFields;
#pragma unroll
  for (int m = 0; m < kUwFields + 9; m++)
    if (m < nf) sw[m * u.sws] = out[m];
}

// The x split in the lines-fastest order of the compact scheme's sweeps
// (row rr of line L at rr*NL + L).  A workgroup takes 16
// rows of 64 x-lines: it reads q with the rows fastest (each line's 16
// points are contiguous in memory), stages the 19 outputs in LDS and writes
// them with the lines fastest (64 lines of a row are contiguous in the split
// buffer; shorter runs, which start mid cache line since NL is odd, leave
// partial lines that other workgroups complete).  The LDS row stride of 66
// doubles puts the 32 lanes of either half-wave of the row-fastest write (16
// rows x 2 lines) on distinct banks.
constexpr int kXtRows = 16, kXtLines = 64, kXtPad = kXtLines + 2, kXtThreads = kXtRows * kXtLines;
// the outputs pass through LDS in two halves (fields 0..9, then, with the
// projection, 10..18)
constexpr int kXtHalf = kUwFields;
__global__ __launch_bounds__(kXtThreads) void k_uw_split_xt(FlowArgs a, UwArgs u) {
  __shared__ double sb[kXtHalf][kXtRows * kXtPad];
  const int NR = a.n[0] + 9;
  const long long NL = uw_lines(a.n, 0);
  const int nrt = (NR + kXtRows - 1) / kXtRows;
  const int rt = (int)(blockIdx.x % nrt);
  const long long lt = blockIdx.x / nrt;
  // compute in the row-fastest order
  const int ri = threadIdx.x % kXtRows, li = threadIdx.x / kXtRows;
  const int rr = rt * kXtRows + ri;
  const long long L = lt * kXtLines + li;
  const bool in = rr < NR && L < NL;
  double out[kUwFields + 9];
  if (in) split_point<0>(a, u, L, rr, out);
  // store in the lines-fastest order
  const int lo = threadIdx.x % kXtLines, ro = threadIdx.x / kXtLines;
  const int rro = rt * kXtRows + ro;
  const long long Lo = lt * kXtLines + lo;
  const bool ino = rro < NR && Lo < NL;
  double* sw = u.sw + (long long)rro * NL + Lo;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (h && !u.chardec) break;  // uniform: the buffer holds 10 fields
    const int m0 = h ? kXtHalf : 0, m1 = h ? kUwFields + 9 : kXtHalf;
    if (h) __syncthreads();  // the first half's reads are done
    if (in) {
#pragma unroll
      for (int m = m0; m < m1; m++) sb[m - m0][ri * kXtPad + li] = out[m];
    }
    __syncthreads();
    if (ino) {
#pragma unroll
      for (int m = m0; m < m1; m++) sw[m * u.sws] = sb[m - m0][ro * kXtPad + lo];
    }
  }
}

// the shock flag of interface i (solver.F90:630-644) on line (o0, stride)
__device__ __forceinline__ bool face_shock(const UwArgs& u, long long o0, long long st, int i, int n) {
  if (!u.sensor) return true;
  if (i < 0) return u.lsh[o0] != 0;
  if (i + 1 > n) return u.lsh[o0 + (long long)n * st] != 0;
  return u.lsh[o0 + (long long)i * st] != 0 || u.lsh[o0 + (long long)(i + 1) * st] != 0;
}

// Fh of the interface between rows i and i+1; c0 = the split-buffer index
// of row i, rs = the buffer stride between rows
template <int SCH, bool CD>
__device__ __forceinline__ void face_flux(const UwArgs& u, long long c0, long long rs, bool lsh, double* Fh) {
  const double* sw = u.sw;
  const long long S = u.sws;
  double Fhc[5];
  if constexpr (CD) {
    if (lsh) {
      double sl[9], sr[9], lev[25], rev[25];
#pragma unroll
      for (int m = 0; m < 9; m++) {
        sl[m] = sw[(kUwFields + m) * S + c0];
        sr[m] = sw[(kUwFields + m) * S + c0 + rs];
      }
      chardecomp(u.gamma, sl, sr, rev, lev);
#pragma unroll
      for (int m = 0; m < 5; m++) {
        double wp[8], wm[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          // iwind8: '+' element k is row i+k-3, '-' element k is row i+4-k
          const long long ip = c0 + (long long)(k - 3) * rs, im = c0 + (long long)(4 - k) * rs;
          wp[k] = lev[5 * m] * sw[ip] + lev[5 * m + 1] * sw[S + ip] + lev[5 * m + 2] * sw[2 * S + ip] +
                  lev[5 * m + 3] * sw[3 * S + ip] + lev[5 * m + 4] * sw[4 * S + ip];
          wm[k] = lev[5 * m] * sw[5 * S + im] + lev[5 * m + 1] * sw[6 * S + im] + lev[5 * m + 2] * sw[7 * S + im] +
                  lev[5 * m + 3] * sw[8 * S + im] + lev[5 * m + 4] * sw[9 * S + im];
        }
        const double v1 = recons<SCH>(wp, u.bfac, lsh);
        const double v2 = recons<SCH>(wm, u.bfac, lsh);
        Fhc[m] = v1 + v2;
      }
#pragma unroll
      for (int m = 0; m < 5; m++)
        Fh[m] = rev[5 * m] * Fhc[0] + rev[5 * m + 1] * Fhc[1] + rev[5 * m + 2] * Fhc[2] + rev[5 * m + 3] * Fhc[3] +
                rev[5 * m + 4] * Fhc[4];
      return;
    }
  }
#pragma unroll
  for (int m = 0; m < 5; m++) {
    double wp[8], wm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      wp[k] = sw[m * S + c0 + (long long)(k - 3) * rs];
      wm[k] = sw[(5 + m) * S + c0 + (long long)(4 - k) * rs];
    }
    const double v1 = recons<SCH>(wp, u.bfac, lsh);
    const double v2 = recons<SCH>(wm, u.bfac, lsh);
    Fh[m] = v1 + v2;
  }
}

// qrhs(t) = qrhs(t) + Fh(t) - Fh(t-1) (solver.F90:789-795) of component m;
// the first direction starts from rhscal's zero, the last one negates (:242)
__device__ __forceinline__ void face_update1(const FlowArgs& a, int m, long long o, double cur, double prev,
                                             bool first, bool last) {
  double* dst = a.qrhs + m * a.fs + o;
  const double old = first ? 0.0 : *dst;
  const double v = old + cur - prev;
  *dst = last ? -v : v;
}

// x-lines: lane l of a wavefront takes interface i0 - 1 + l of segment
// (L, sg); lanes 1..63 update points i0..i0+62
template <int SCH, bool CD>
__global__ __launch_bounds__(256) void k_uw_face_x(FlowArgs a, UwArgs u, int first, int last) {
  const int n = a.n[0];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, 0);
  const int segs = (n + 1 + 62) / 63;
  const int lane = threadIdx.x & 63;
  const long long w = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long L = w / segs;
  if (L >= NL) return;  // whole wavefronts only
  const int sg = (int)(w - L * segs);
  const int i = sg * 63 - 1 + lane;
  const bool ok = i <= n;
  int c3[3];
  const long long o0 = line_base(a, 0, L, c3);
  double Fh[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (ok) face_flux<SCH, CD>(u, L * NR + (i + 4), 1, face_shock(u, o0, 1, i, n), Fh);
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const double prev = __shfl_up(Fh[m], 1, 64);
    if (ok && lane >= 1) face_update1(a, m, o0 + i, Fh[m], prev, first, last);
  }
}

// y- and z-lines without the characteristic projection: one lane per line
// and a run of kRun points; per component the 8-row windows of F+ and F-
// slide along the run in registers (one new row of each per point)
template <int SCH, int DIR>
__global__ __launch_bounds__(256) void k_uw_face_yz(FlowArgs a, UwArgs u, int first, int last) {
  constexpr int dir = DIR;
  const int n = a.n[dir];
  const long long NL = uw_lines(a.n, dir);
  const int nruns = (n + 1 + kRun - 1) / kRun;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= NL * nruns) return;
  const int run = (int)(idx / NL);
  const long long L = idx - (long long)run * NL;
  const int r0 = run * kRun, r1 = min(r0 + kRun - 1, n);
  int c3[3];
  const long long o0 = line_base(a, dir, L, c3);
  const long long st = dir_stride(a, dir);
  const double* sw = u.sw;
  const long long S = u.sws;
  bool shp = face_shock(u, o0, st, r0 - 1, n);
#pragma unroll 1
  for (int m = 0; m < 5; m++) {
    // rows r0-4 .. r0+3: the window of interface r0-1
    double wp[8], wr[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const long long c = (long long)(r0 - 4 + k + 4) * NL + L;
      wp[k] = sw[m * S + c];
      wr[k] = sw[(5 + m) * S + c];
    }
    double wm[8];
#pragma unroll
    for (int k = 0; k < 8; k++) wm[k] = wr[7 - k];
    double prev = recons<SCH>(wp, u.bfac, shp) + recons<SCH>(wm, u.bfac, shp);
#pragma unroll 1
    for (int t = r0; t <= r1; t++) {
#pragma unroll
      for (int k = 0; k < 7; k++) {
        wp[k] = wp[k + 1];
        wr[k] = wr[k + 1];
      }
      const long long c = (long long)(t + 4 + 4) * NL + L;
      wp[7] = sw[m * S + c];
      wr[7] = sw[(5 + m) * S + c];
#pragma unroll
      for (int k = 0; k < 8; k++) wm[k] = wr[7 - k];
      const bool sh = face_shock(u, o0, st, t, n);
      const double cur = recons<SCH>(wp, u.bfac, sh) + recons<SCH>(wm, u.bfac, sh);
      face_update1(a, m, o0 + (long long)t * st, cur, prev, first, last);
      prev = cur;
    }
  }
}

// y- and z-lines with the characteristic projection: one lane per line
// and a run of kRun points, the five components of each interface together
template <int SCH, int DIR>
__global__ __launch_bounds__(256) void k_uw_face_yz_cd(FlowArgs a, UwArgs u, int first, int last) {
  constexpr int dir = DIR;
  const int n = a.n[dir];
  const long long NL = uw_lines(a.n, dir);
  const int nruns = (n + 1 + kRun - 1) / kRun;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= NL * nruns) return;
  const int run = (int)(idx / NL);
  const long long L = idx - (long long)run * NL;
  const int r0 = run * kRun, r1 = min(r0 + kRun - 1, n);
  int c3[3];
  const long long o0 = line_base(a, dir, L, c3);
  const long long st = dir_stride(a, dir);
  double prev[5], cur[5];
  face_flux<SCH, true>(u, (long long)(r0 - 1 + 4) * NL + L, NL, face_shock(u, o0, st, r0 - 1, n), prev);
#pragma unroll 1
  for (int t = r0; t <= r1; t++) {
    face_flux<SCH, true>(u, (long long)(t + 4) * NL + L, NL, face_shock(u, o0, st, t, n), cur);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      face_update1(a, m, o0 + (long long)t * st, cur[m], prev[m], first, last);
      prev[m] = cur[m];
    }
  }
}

// the stored pressure at the neighbour s = +-1 along d of point c3 (o), or
// the pressure q2fvar gives its halo copy
__device__ __forceinline__ double prs_nb(const FlowArgs& a, const int* c3, long long o, int d, int s) {
  const int n = a.n[d];
  const int c = c3[d] + s;
  const long long st = dir_stride(a, d);
  if (c >= 0 && c <= n) return a.pr[4 * a.fs + o + s * st];
  const int rs = c < 0 ? c + n : c - n;
  double q[5];
  halo_q_at(a, d, c3, rs, o + (long long)(rs - c3[d]) * st, q);
  return pressure_of(a, q[0], q[1], q[2], q[3], q[4]);
}

// ducrossensor's ratio (commcal.F90:226-267) at every computed point
__global__ __launch_bounds__(64) void k_ssf(FlowArgs a, UwArgs u) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i > a.n[0]) return;
  const int c3[3] = {i, (int)blockIdx.y, (int)blockIdx.z};
  const long long o = (long long)c3[2] * a.plane + (long long)c3[1] * a.pitch + i;
  const long long fs = a.fs;
  double g[3][3];
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int d = 0; d < 3; d++) g[m][d] = d < a.ndir ? a.dv[(m + 3 * d) * fs + o] : 0.0;
  const double dsum = g[0][0] + g[1][1] + g[2][2];
  const double div2 = dsum * dsum;
  const double vx = g[2][1] - g[1][2];
  const double vy = g[0][2] - g[2][0];
  const double vz = g[1][0] - g[0][1];
  const double vort = vx * vx + vy * vy + vz * vz;
  const double p0 = a.pr[4 * fs + o];
  double dmax = 0.0;
  for (int d = 0; d < a.ndir; d++) {
    const double pp = prs_nb(a, c3, o, d, 1), pm = prs_nb(a, c3, o, d, -1);
    const double r = fabs(pp - 2.0 * p0 + pm) / (pp + 2.0 * p0 + pm);
    dmax = d == 0 ? r : fmax(dmax, r);
  }
  u.ssf[o] = div2 / (div2 + vort + 1.0e-30) * dmax;
}

__device__ __forceinline__ int wrapn(int t, int n) { return t < 0 ? t + n : (t > n ? t - n : t); }

// the shock flags: ssf's maximum over offsets -4..5 along each direction
// (periodic copies) against shkcrt (commcal.F90:273-306)
__global__ __launch_bounds__(64) void k_lshock(FlowArgs a, UwArgs u) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i > a.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
  double mx = 0.0;
  for (int s = -4; s <= 5; s++) {
    mx = fmax(mx, u.ssf[o + (wrapn(i + s, a.n[0]) - i)]);
    mx = fmax(mx, u.ssf[o + (long long)(wrapn(j + s, a.n[1]) - j) * a.pitch]);
    if (a.ndir == 3) mx = fmax(mx, u.ssf[o + (long long)(wrapn(k + s, a.n[2]) - k) * a.plane]);
  }
  u.lsh[o] = mx > u.shkcrt ? 1 : 0;
}

// ---------------------------------------------------------------------------
// The compact upwind-biased scheme (conschm 543c, convrsdcmp,
// solver.F90:1271-1937) on periodic lines: the compact interface fluxes of
// every split component (k_cuw_solve), the MP5-limited interface flux
// (k_cuw_face) and its difference into qrhs (k_cuw_diff).
// ---------------------------------------------------------------------------

// split-buffer index of row r (-4..n+4) of line L for the compact scheme's
// lane-per-line sweeps (with the characteristic projection), which keep
// every direction's lines fastest, so that one lane per line coalesces
template <int DIR>
__device__ __forceinline__ long long cuw_at(long long L, int r, int NR, long long NL) {
  (void)NR;
  return (long long)(r + 4) * NL + L;
}

// MP5 (flux.F90:434-495) as mplimiter calls it on a periodic line
// (flux.F90:363-379): u(1:5) with u(3) the upwind point, ul the compact
// scheme's interface value
__device__ __forceinline__ double mp5(const double* u, double ul, bool shock) {
  const double ump = u[2] + minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
  if (shock && (ul - u[2]) * (ul - ump) >= 1.0e-10) {
    const double w[6] = {0.0, u[0], u[1], u[2], u[3], u[4]};
    return mp_limit(w, ul);
  }
  return ul;
}

// flux_compact (flux.F90:125-150, compact_flux_rhs :163-254, the Thomas
// sweeps of tridiagonal_thomas_solver, commfunc.F90:790-813) of one line and
// one split component (0..4 F+ with the upwind scheme, 5..9 F- with the
// downwind one): rows -2..n+1, the end rows the explicit 6th-order interface
// value, rows -1..n the 5th-order compact one.  The forward sweep leaves d'
// in the output rows, the backward sweep overwrites them with the solution;
// rows -1..n are the interface fluxes fh(-1:n).  One lane per (component,
// line); for y and z the lanes of a wavefront take consecutive lines, so
// every access coalesces.
template <int DIR>
__global__ __launch_bounds__(256) void k_cuw_solve(FlowArgs a, UwArgs u) {
  const int n = a.n[DIR];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, DIR);
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= 10 * NL) return;
  const int f = (int)(idx / NL);
  const long long L = idx - (long long)f * NL;
  const bool plus = f < 5;
  const double* __restrict__ ac = u.cuw_ac + cuw_ac_offset(a.n, DIR, plus);
  const double* __restrict__ src = u.sw + f * u.sws;
  double* __restrict__ dst = u.fhc + f * u.sws;
  const double b = u.bfac;
  const double k118 = 1.0 / 18.0, k136 = 1.0 / 36.0, k1918 = 19.0 / 18.0, k59 = 5.0 / 9.0, k936 = 9.0 / 36.0;
  const double w0 = k118 - k136 * b, w1 = k1918 - k936 * b, w2 = k59 + k936 * b, w3 = k136 * b;
  const double c37 = 37.0 / 60.0, c2 = 2.0 / 15.0, c1 = 1.0 / 60.0;
  auto F = [&](int r) { return src[cuw_at<DIR>(L, r, NR, NL)]; };
  auto end_row = [&](int j) {
    const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
    return c37 * v1 - c2 * v2 + c1 * v3;
  };
  const int N = n + 4;
  double dp = end_row(-2);
  dst[cuw_at<DIR>(L, -2, NR, NL)] = dp;
  // interior rows j = -1..n: the window F(j-1..j+2) slides one row a step;
  // unrolled so that the loads of later rows issue ahead of the recurrence
  double f0 = F(-2), f1 = F(-1), f2 = F(0);
#pragma unroll 8
  for (int r = 1; r < N - 1; r++) {
    const int j = r - 2;
    const double f3 = F(j + 2);
    const double d = plus ? w0 * f0 + w1 * f1 + w2 * f2 + w3 * f3 : w0 * f3 + w1 * f2 + w2 * f1 + w3 * f0;
    dp = d * ac[3 * r + 1] - dp * ac[3 * r + 2];
    dst[cuw_at<DIR>(L, j, NR, NL)] = dp;
    f0 = f1;
    f1 = f2;
    f2 = f3;
  }
  dp = end_row(n + 1) * ac[3 * (N - 1) + 1] - dp * ac[3 * (N - 1) + 2];
  dst[cuw_at<DIR>(L, n + 1, NR, NL)] = dp;
  double x = dp;
#pragma unroll 8
  for (int r = N - 2; r >= 0; r--) {
    const int j = r - 2;
    const long long c = cuw_at<DIR>(L, j, NR, NL);
    x = dst[c] - ac[3 * r] * x;
    dst[c] = x;
  }
}

// The interface flux Fh(i), i = -1..n, of convrsdcmp (solver.F90:1366-1490):
// the windows Flcp(m, n) = fplus(i+n-3), Flcm(m, n) = fminus(i+4-n), n =
// 1..5 (iwind6) and the compact values, with CD all of them projected on the
// Roe-state characteristics of rows i, i+1, each limited by MP5 (no crinod:
// hdiss false), summed, projected back.  Fh(i) replaces the compact value of
// F+ component m at row i (this lane's own row, read before).
template <int DIR, bool CD>
__global__ __launch_bounds__(256) void k_cuw_face(FlowArgs a, UwArgs u) {
  const int n = a.n[DIR];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, DIR);
  const long long NI = n + 2;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= NI * NL) return;
  // consecutive lanes take consecutive lines, as the split buffer holds them
  const int i = (int)(idx / NL) - 1;
  const long long L = idx - (long long)(i + 1) * NL;
  int c3[3];
  const long long o0 = line_base(a, DIR, L, c3);
  const long long st = dir_stride(a, DIR);
  const bool lsh = face_shock(u, o0, st, i, n);
  const double* sw = u.sw;
  const long long S = u.sws;
  const long long ci = cuw_at<DIR>(L, i, NR, NL);
  auto at = [&](int r) { return cuw_at<DIR>(L, r, NR, NL); };
  double hp[5], hm[5], Fh[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    hp[m] = u.fhc[m * S + ci];
    hm[m] = u.fhc[(5 + m) * S + ci];
  }
  if constexpr (CD) {
    double sl[9], sr[9], lev[25], rev[25];
    const long long cr = at(i + 1);
#pragma unroll
    for (int m = 0; m < 9; m++) {
      sl[m] = sw[(kUwFields + m) * S + ci];
      sr[m] = sw[(kUwFields + m) * S + cr];
    }
    chardecomp(u.gamma, sl, sr, rev, lev);
    double Fhc[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      double wp[5], wm[5];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const long long ip = at(i + k - 2), im = at(i + 3 - k);
        wp[k] = lev[5 * m] * sw[ip] + lev[5 * m + 1] * sw[S + ip] + lev[5 * m + 2] * sw[2 * S + ip] +
                lev[5 * m + 3] * sw[3 * S + ip] + lev[5 * m + 4] * sw[4 * S + ip];
        wm[k] = lev[5 * m] * sw[5 * S + im] + lev[5 * m + 1] * sw[6 * S + im] + lev[5 * m + 2] * sw[7 * S + im] +
                lev[5 * m + 3] * sw[8 * S + im] + lev[5 * m + 4] * sw[9 * S + im];
      }
      const double lp = lev[5 * m] * hp[0] + lev[5 * m + 1] * hp[1] + lev[5 * m + 2] * hp[2] +
                        lev[5 * m + 3] * hp[3] + lev[5 * m + 4] * hp[4];
      const double lm = lev[5 * m] * hm[0] + lev[5 * m + 1] * hm[1] + lev[5 * m + 2] * hm[2] +
                        lev[5 * m + 3] * hm[3] + lev[5 * m + 4] * hm[4];
      const double v1 = mp5(wp, lp, lsh);
      const double v2 = mp5(wm, lm, lsh);
      Fhc[m] = v1 + v2;
    }
#pragma unroll
    for (int m = 0; m < 5; m++)
      Fh[m] = rev[5 * m] * Fhc[0] + rev[5 * m + 1] * Fhc[1] + rev[5 * m + 2] * Fhc[2] + rev[5 * m + 3] * Fhc[3] +
              rev[5 * m + 4] * Fhc[4];
  } else {
#pragma unroll
    for (int m = 0; m < 5; m++) {
      double wp[5], wm[5];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        wp[k] = sw[m * S + at(i + k - 2)];
        wm[k] = sw[(5 + m) * S + at(i + 3 - k)];
      }
      const double v1 = mp5(wp, hp[m], lsh);
      const double v2 = mp5(wm, hm[m], lsh);
      Fh[m] = v1 + v2;
    }
  }
#pragma unroll
  for (int m = 0; m < 5; m++) u.fhc[m * S + ci] = Fh[m];
}

// qrhs(t) = qrhs(t) + Fh(t) - Fh(t-1), t = 0..n (solver.F90:1492-1497)
template <int DIR>
__global__ __launch_bounds__(256) void k_cuw_diff(FlowArgs a, UwArgs u, int first, int last) {
  const int n = a.n[DIR];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, DIR);
  const long long NP = n + 1;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= NP * NL) return;
  long long L;
  int t;
  if (DIR == 0) {
    // x: consecutive lanes take consecutive points of a line, so that qrhs
    // coalesces
    L = idx / NP;
    t = (int)(idx - L * NP);
  } else {
    t = (int)(idx / NL);
    L = idx - (long long)t * NL;
  }
  int c3[3];
  const long long o0 = line_base(a, DIR, L, c3);
  const long long o = o0 + (long long)t * dir_stride(a, DIR);
  const long long c = cuw_at<DIR>(L, t, NR, NL), cp = cuw_at<DIR>(L, t - 1, NR, NL);
#pragma unroll
  for (int m = 0; m < 5; m++)
    face_update1(a, m, o, u.fhc[m * u.sws + c], u.fhc[m * u.sws + cp], first, last);
}

// k_cuw_diff along x through LDS: the interface fluxes are held with the
// lines fastest and qrhs with the points of a line contiguous.  A workgroup
// takes 16 points of 64 x-lines: it reads the flux rows t0-1 .. t0+15 with
// the lines fastest into LDS (row stride 66 doubles, so that the 32 lanes of
// a half-wave reading by points, 16 points x 2 lines, fall on distinct banks)
// and updates qrhs with the points fastest.
constexpr int kDtRows = 16, kDtLines = 64, kDtPad = kDtLines + 2, kDtThreads = kDtRows * kDtLines;
__global__ __launch_bounds__(kDtThreads) void k_cuw_diff_xt(FlowArgs a, UwArgs u, int first, int last) {
  __shared__ double sb[5][(kDtRows + 1) * kDtPad];
  const int n = a.n[0];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, 0);
  const int ntt = (n + 1 + kDtRows - 1) / kDtRows;
  const int t0 = (int)(blockIdx.x % ntt) * kDtRows;
  const long long L0 = (long long)(blockIdx.x / ntt) * kDtLines;
  for (int e = threadIdx.x; e < (kDtRows + 1) * kDtLines; e += kDtThreads) {
    const int rr = e / kDtLines, l = e % kDtLines;
    const int t = t0 - 1 + rr;  // -1 .. n: the rows fh(-1:n) holds
    if (t <= n && L0 + l < NL) {
      const long long c = cuw_at<0>(L0 + l, t, NR, NL);
#pragma unroll
      for (int m = 0; m < 5; m++) sb[m][rr * kDtPad + l] = u.fhc[m * u.sws + c];
    }
  }
  __syncthreads();
  const int ti = threadIdx.x % kDtRows, l = threadIdx.x / kDtRows;
  const int t = t0 + ti;
  if (t <= n && L0 + l < NL) {
    int c3[3];
    const long long o = line_base(a, 0, L0 + l, c3) + t;
#pragma unroll
    for (int m = 0; m < 5; m++)
      face_update1(a, m, o, sb[m][(ti + 1) * kDtPad + l], sb[m][ti * kDtPad + l], first, last);
  }
}

// ---------------------------------------------------------------------------
// The compact scheme without the characteristic projection, one pass per
// direction with every line in LDS: per split component, the upwind and the
// downwind compact systems of TL lines are solved by the line engine's
// chunked Thomas recurrences (lane p of a line owns rows p*R-2 ..
// p*R+R-3, the chunks' carries composed by affine_scan), MP5 turns each
// interface's compact values into Fh(i) against the 5-point windows, and
// qrhs takes Fh(t) - Fh(t-1).  Nothing between the split buffer and qrhs
// goes through HBM.  The recurrences are re-associated across lanes, as in
// the derivative and filter solves; the arithmetic of each row is the
// reference's.
// ---------------------------------------------------------------------------

// LU factors of one chunk of a compact flux system (a != c, so ac1 and ac3
// differ): c1 = ac1 (0 on the last row), c2 = ac2 (1 on the first), c3 = ac3
// (0 on the first), all 0 past the last row
template <int R>
struct CuwCoef {
  double c1[R], c2[R], c3[R];
};

template <int R>
__device__ __forceinline__ void load_cuw_coef(CuwCoef<R>& C, const double* __restrict__ ac, int M, int u0) {
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int v = u0 + k;
    const bool ok = v < M;
    const int vv = ok ? v : 0;
    C.c1[k] = ok ? ac[3 * vv] : 0.0;
    C.c2[k] = ok ? ac[3 * vv + 1] : 0.0;
    C.c3[k] = ok ? ac[3 * vv + 2] : 0.0;
  }
}

// flux_compact of one line in LDS: T holds f(-8 .. P*R+2) at offset
// kLineOff (f(-4..n+4) the split values, zeros elsewhere), X receives
// fh(-1..n) at X[0..n+1]
template <int P, int R, bool PLUS>
__device__ __forceinline__ void solve_line_cuw(const double* __restrict__ T, double* __restrict__ X, int n, int p,
                                               const CuwCoef<R>& C, const double (&wt)[4]) {
  constexpr int OFF = 2, H = 3, W = R + 2 * H;
  const int u0 = p * R;
  const double* __restrict__ Tw = T + kLineOff + (u0 - OFF - H);
  double wv[W];
#pragma unroll
  for (int k = 0; k < W; k++) wv[k] = Tw[k];
  const double c37 = 37.0 / 60.0, c2 = 2.0 / 15.0, c1 = 1.0 / 60.0;
  double y[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    const double* c = wv + k + H;  // c[s] = f(r + s), r = u0 + k - OFF
    const int r = u0 + k - OFF;
    double d;
    if (r == -2 || r == n + 1) {
      const double v1 = c[0] + c[1], v2 = c[-1] + c[2], v3 = c[-2] + c[3];
      d = c37 * v1 - c2 * v2 + c1 * v3;
    } else if (PLUS) {
      d = wt[0] * c[-1] + wt[1] * c[0] + wt[2] * c[1] + wt[3] * c[2];
    } else {
      d = wt[0] * c[2] + wt[1] * c[1] + wt[2] * c[0] + wt[3] * c[-1];
    }
    y[k] = d;
  }
  // forward recurrence d'(v) = d(v) ac2(v) - d'(v-1) ac3(v), zero carry-in
  double B = 0.0, A = 1.0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    B = y[k] * C.c2[k] - B * C.c3[k];
    A *= -C.c3[k];
    y[k] = B;
  }
  const double cin = affine_scan<P, true>(A, B);
  {
    double m = 1.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      m *= -C.c3[k];
      y[k] = fma(m, cin, y[k]);
    }
  }
  // backward recurrence x(v) = d'(v) - ac1(v) x(v+1), zero carry from the right
  B = 0.0;
  A = 1.0;
#pragma unroll
  for (int k = R - 1; k >= 0; k--) {
    B = y[k] - C.c1[k] * B;
    A *= -C.c1[k];
    y[k] = B;
  }
  const double xin = affine_scan<P, false>(A, B);
  {
    double g = 1.0;
#pragma unroll
    for (int k = R - 1; k >= 0; k--) {
      g *= -C.c1[k];
      y[k] = fma(g, xin, y[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int r = u0 + k - OFF;
    if (r >= -1 && r <= n) X[r + 1] = y[k];
  }
}

template <class C>
constexpr size_t cuw_line_lds() {
  return 4 * (size_t)C::TL * C::LT * sizeof(double);
}

template <class C, int DIR>
__global__ __launch_bounds__(C::NT) void k_cuw_line(FlowArgs a, UwArgs u, int first, int last) {
  extern __shared__ double smem[];
  constexpr int TL = C::TL, P = C::P, R = C::R, LT = C::LT, NT = C::NT;
  const int n = a.n[DIR];
  const int NR = n + 9;
  const long long NL = uw_lines(a.n, DIR);
  const long long L0 = (long long)blockIdx.x * TL;
  const long long S = u.sws;
  double* Tp = smem;            // F+ lines
  double* Tm = Tp + TL * LT;    // F- lines
  double* Xp = Tm + TL * LT;    // fh+ (then Fh) at rows -1..n -> 0..n+1
  double* Xm = Xp + TL * LT;    // fh-
  const int ln = threadIdx.x / P, p = threadIdx.x % P;
  CuwCoef<R> Cp, Cm;
  load_cuw_coef<R>(Cp, u.cuw_ac + cuw_ac_offset(a.n, DIR, true), n + 4, p * R);
  load_cuw_coef<R>(Cm, u.cuw_ac + cuw_ac_offset(a.n, DIR, false), n + 4, p * R);
  const double b = u.bfac;
  const double k118 = 1.0 / 18.0, k136 = 1.0 / 36.0, k1918 = 19.0 / 18.0, k59 = 5.0 / 9.0, k936 = 9.0 / 36.0;
  const double wt[4] = {k118 - k136 * b, k1918 - k936 * b, k59 + k936 * b, k136 * b};
  // the rows the split does not fill read as zeros (only the padding of the
  // chunks past the system's last row and the windows' outer edge touch them)
  for (int e = threadIdx.x; e < 2 * TL * LT; e += NT) smem[e] = 0.0;
  __syncthreads();
#pragma unroll 1
  for (int m = 0; m < 5; m++) {
    // rows -4..n+4 of F+_m and F-_m of the tile's TL lines, read in the
    // split buffer's order (x: each line's rows contiguous; y, z: the lines
    // side by side)
    for (int e = threadIdx.x; e < TL * NR; e += NT) {
      const int l = DIR == 0 ? e / NR : e % TL, rr = DIR == 0 ? e % NR : e / TL;
      const long long L = L0 + l;
      double vp = 0.0, vm = 0.0;
      if (L < NL) {
        const long long c = DIR == 0 ? L * NR + rr : (long long)rr * NL + L;
        vp = u.sw[m * S + c];
        vm = u.sw[(5 + m) * S + c];
      }
      Tp[l * LT + kLineOff + rr - 4] = vp;
      Tm[l * LT + kLineOff + rr - 4] = vm;
    }
    __syncthreads();
    solve_line_cuw<P, R, true>(Tp + ln * LT, Xp + ln * LT, n, p, Cp, wt);
    solve_line_cuw<P, R, false>(Tm + ln * LT, Xm + ln * LT, n, p, Cm, wt);
    __syncthreads();
    // Fh(i), i = -1..n: MP5 of the windows F+(i-2..i+2), F-(i+3..i-1)
    // against the compact values; every interface is flagged (no sensor)
    for (int e = threadIdx.x; e < TL * (n + 2); e += NT) {
      const int l = e % TL, i = e / TL - 1;
      const double* tp = Tp + l * LT + kLineOff;
      const double* tm = Tm + l * LT + kLineOff;
      double wp[5], wm[5];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        wp[k] = tp[i + k - 2];
        wm[k] = tm[i + 3 - k];
      }
      const double v1 = mp5(wp, Xp[l * LT + i + 1], true);
      const double v2 = mp5(wm, Xm[l * LT + i + 1], true);
      Xp[l * LT + i + 1] = v1 + v2;
    }
    __syncthreads();
    // qrhs(t) = qrhs(t) + Fh(t) - Fh(t-1), t = 0..n
    for (int e = threadIdx.x; e < TL * (n + 1); e += NT) {
      int l, t;
      if (DIR == 0) {
        l = e / (n + 1);
        t = e - l * (n + 1);
      } else {
        l = e % TL;
        t = e / TL;
      }
      const long long L = L0 + l;
      if (L >= NL) continue;
      int c3[3];
      const long long o0 = line_base(a, DIR, L, c3);
      face_update1(a, m, o0 + (long long)t * dir_stride(a, DIR), Xp[l * LT + t + 1], Xp[l * LT + t], first,
                   last);
    }
    __syncthreads();
  }
}

// launch shapes of k_cuw_line: the first whose chunks hold the n+4 rows
using CuwCfg16 = LineCfg<256, 16, 5, 5>;  // P = 16: n <= 76
using CuwCfg32 = LineCfg<256, 8, 5, 5>;   // P = 32: n <= 156
using CuwCfg64 = LineCfg<512, 8, 5, 5>;   // P = 64: n <= 316
using CuwCfg64b = LineCfg<512, 8, 9, 9>;  // P = 64, 9 rows per chunk: n <= 572

// The compact scheme along dir runs the three-kernel sweeps (one lane per
// line and component, the split buffer holding the lines fastest) rather
// than the LDS-resident pass: with the characteristic projection, or for
// lines longer than the LDS pass holds (the flow passes stop at 569 points,
// below its 572).  The x split and the face launch both follow it.
bool cuw_sweeps(const FlowArgs& a, const UwArgs& u, int dir) {
  return u.compact && (u.chardec || a.n[dir] + 4 > CuwCfg64b::P * CuwCfg64b::R);
}

template <class C, int DIR>
hipError_t cuw_line_launch(const FlowArgs& a, const UwArgs& u, bool first, bool last, hipStream_t st) {
  const long long NL = uw_lines(a.n, DIR);
  const dim3 g((unsigned)((NL + C::TL - 1) / C::TL));
  hipLaunchKernelGGL((k_cuw_line<C, DIR>), g, dim3(C::NT), cuw_line_lds<C>(), st, a, u, (int)first, (int)last);
  return hipGetLastError();
}

template <int DIR>
hipError_t cuw_launch_dir(const FlowArgs& a, const UwArgs& u, bool first, bool last, hipStream_t st) {
  if (!cuw_sweeps(a, u, DIR)) {
    const int n = a.n[DIR];
    if (n + 4 <= CuwCfg16::P * CuwCfg16::R) return cuw_line_launch<CuwCfg16, DIR>(a, u, first, last, st);
    if (n + 4 <= CuwCfg32::P * CuwCfg32::R) return cuw_line_launch<CuwCfg32, DIR>(a, u, first, last, st);
    if (n + 4 <= CuwCfg64::P * CuwCfg64::R) return cuw_line_launch<CuwCfg64, DIR>(a, u, first, last, st);
    if (n + 4 <= CuwCfg64b::P * CuwCfg64b::R) return cuw_line_launch<CuwCfg64b, DIR>(a, u, first, last, st);
  }
  const long long NL = uw_lines(a.n, DIR);
  hipLaunchKernelGGL(k_cuw_solve<DIR>, dim3((unsigned)((10 * NL + 255) / 256)), dim3(256), 0, st, a, u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 gf((unsigned)(((a.n[DIR] + 2) * NL + 255) / 256));
  if (u.chardec) hipLaunchKernelGGL((k_cuw_face<DIR, true>), gf, dim3(256), 0, st, a, u);
  else hipLaunchKernelGGL((k_cuw_face<DIR, false>), gf, dim3(256), 0, st, a, u);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (DIR == 0) {
    const long long nb = (long long)((a.n[0] + 1 + kDtRows - 1) / kDtRows) * ((NL + kDtLines - 1) / kDtLines);
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cuw_diff_xt, dim3((unsigned)nb), dim3(kDtThreads), 0, st, a, u, (int)first, (int)last);
    return hipGetLastError();
  }
  const dim3 gd((unsigned)(((a.n[DIR] + 1) * NL + 255) / 256));
  hipLaunchKernelGGL(k_cuw_diff<DIR>, gd, dim3(256), 0, st, a, u, (int)first, (int)last);
  return hipGetLastError();
}

}  // namespace

hipError_t uw_launch_sensor(const FlowArgs& a, const UwArgs& u, hipStream_t st) {
  const dim3 g((unsigned)((a.n[0] + 64) / 64), (unsigned)(a.n[1] + 1), (unsigned)(a.n[2] + 1));
  hipLaunchKernelGGL(k_ssf, g, dim3(64), 0, st, a, u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lshock, g, dim3(64), 0, st, a, u);
  return hipGetLastError();
}

hipError_t uw_launch_split(const FlowArgs& a, const UwArgs& u, int dir, hipStream_t st) {
  if (dir == 0 && cuw_sweeps(a, u, 0)) {
    const long long nb = (long long)((a.n[0] + 9 + kXtRows - 1) / kXtRows) *
                         ((uw_lines(a.n, 0) + kXtLines - 1) / kXtLines);
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_uw_split_xt, dim3((unsigned)nb), dim3(kXtThreads), 0, st, a, u);
    return hipGetLastError();
  }
  const dim3 g((unsigned)((uw_points(a.n, dir) + 255) / 256));
  if (dir == 0) hipLaunchKernelGGL(k_uw_split<0>, g, dim3(256), 0, st, a, u);
  else if (dir == 1) hipLaunchKernelGGL(k_uw_split<1>, g, dim3(256), 0, st, a, u);
  else hipLaunchKernelGGL(k_uw_split<2>, g, dim3(256), 0, st, a, u);
  return hipGetLastError();
}

template <int SCH>
hipError_t launch_face_s(const FlowArgs& a, const UwArgs& u, int dir, bool first, bool last, hipStream_t st) {
  const long long NL = uw_lines(a.n, dir);
  if (dir == 0) {
    const long long waves = NL * ((a.n[0] + 1 + 62) / 63);
    const dim3 g((unsigned)((waves + 3) / 4));
    if (u.chardec)
      hipLaunchKernelGGL((k_uw_face_x<SCH, true>), g, dim3(256), 0, st, a, u, (int)first, (int)last);
    else
      hipLaunchKernelGGL((k_uw_face_x<SCH, false>), g, dim3(256), 0, st, a, u, (int)first, (int)last);
  } else {
    const long long nt = NL * ((a.n[dir] + 1 + kRun - 1) / kRun);
    const dim3 g((unsigned)((nt + 255) / 256));
    const int f = first, l = last;
    if (u.chardec && dir == 1) hipLaunchKernelGGL((k_uw_face_yz_cd<SCH, 1>), g, dim3(256), 0, st, a, u, f, l);
    else if (u.chardec) hipLaunchKernelGGL((k_uw_face_yz_cd<SCH, 2>), g, dim3(256), 0, st, a, u, f, l);
    else if (dir == 1) hipLaunchKernelGGL((k_uw_face_yz<SCH, 1>), g, dim3(256), 0, st, a, u, f, l);
    else hipLaunchKernelGGL((k_uw_face_yz<SCH, 2>), g, dim3(256), 0, st, a, u, f, l);
  }
  return hipGetLastError();
}

hipError_t uw_launch_face(const FlowArgs& a, const UwArgs& u, int dir, bool first, bool last, hipStream_t st) {
  if (u.compact) {
    if (dir == 0) return cuw_launch_dir<0>(a, u, first, last, st);
    if (dir == 1) return cuw_launch_dir<1>(a, u, first, last, st);
    return cuw_launch_dir<2>(a, u, first, last, st);
  }
  switch (u.recon) {
    case -1: return launch_face_s<-1>(a, u, dir, first, last, st);
    case 0: return launch_face_s<0>(a, u, dir, first, last, st);
    case 1: return launch_face_s<1>(a, u, dir, first, last, st);
    case 2: return launch_face_s<2>(a, u, dir, first, last, st);
    case 3: return launch_face_s<3>(a, u, dir, first, last, st);
    case 5: return launch_face_s<5>(a, u, dir, first, last, st);
    default: return launch_face_s<6>(a, u, dir, first, last, st);
  }
}

}  // namespace astr
