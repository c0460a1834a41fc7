// chem.hip -- per-cell chemical source terms on gfx950 and their C ABI.
//
// Replaces the reference's per-cell Cantera calls on the hot path:
//   chemrate(den,tmp,spc,wi)   src/thermchem.F90:1553-1575  (setState_TRY +
//                              getNetProductionRates, wi = wdot*W)
//   cpeval(tmp,spc,cp=)        src/thermchem.F90:1342-1428  (cp_mass)
//   enthpy(tmp,hi)             src/thermchem.F90:1498-1541  (h_k = H/RT *
//                              (8314.2/W_k) * T)
// for a CTML mechanism (Cantera's XML format, e.g.
// chemMech/Jachimowski88_H2-9-19.xml): ideal gas, NASA 7-coefficient
// thermodynamics, elementary and three-body mass-action reactions with
// Arrhenius rate constants, reverse rates from the equilibrium constant.
// One thread evaluates one cell; all arithmetic is FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/astr_hip.h"

namespace {

thread_local std::string g_cerr;
int cfail(const std::string& m) {
  g_cerr = m;
  return -1;
}

#define CHIP_OK(x)                                                               \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) return cfail(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int NSMAX = 16;
constexpr int NRMAX = 48;
constexpr int NTERM = 3;  // species per side of a reaction

// gas constant (J/kmol/K) of Cantera 2.5 (CODATA 2018 Avogadro x Boltzmann),
// used inside Cantera's thermodynamics and kinetics
constexpr double kRgas = 8314.46261815324;
// the reference's own constant in enthpy/rgcmix (thermchem.F90:26,490,1513)
constexpr double kRref = 8314.2;

struct Mech {
  int ns, nr;
  double W[NSMAX];         // molecular weights, kg/kmol
  double tmid[NSMAX];
  double lo[NSMAX][7], hi[NSMAX][7];
  double lnA[NRMAX], b[NRMAX], EaR[NRMAX];
  int rev[NRMAX], tb[NRMAX];
  double eff[NRMAX][NSMAX];  // third-body efficiencies
  int nre[NRMAX], rsp[NRMAX][NTERM];
  double rnu[NRMAX][NTERM];
  int npr[NRMAX], psp[NRMAX][NTERM];
  double pnu[NRMAX][NTERM];
  double dnu[NRMAX];       // sum of product minus reactant coefficients
  double p0;               // standard-state pressure of the thermo data, Pa
};

// ---------------------------------------------------------------------------
// device evaluation
// ---------------------------------------------------------------------------
// Reaction orders other than 1 and 2 go through pow, kept out of line: the
// inlined pow's temporaries set the register peak of every kernel that
// evaluates rates, although orders 1 and 2 never reach it.
__device__ __attribute__((noinline)) double conc_pow_general(double c, double nu) {
  return pow(c, nu);
}

__device__ __forceinline__ double conc_pow(double c, double nu) {
  return nu == 1.0 ? c : (nu == 2.0 ? c * c : conc_pow_general(c, nu));
}

// Thermodynamics and kinetics of one state: molar concentrations from rho and
// Y; per species cp/R (cpr), h/RT (hrt); net molar production rates w
// (kmol/m^3/s).  GasKinetics: kf = A T^b exp(-E/RT), kr = kf / Kc with
// Kc = exp(-dG/RT) (p0/RT)^dnu, three-body [M] = sum eff_k C_k.
__device__ __forceinline__ void kinetics(const Mech* __restrict__ M, double T, double rho,
                                         const double* y, double* w, double* cpr, double* hrt) {
  const int ns = M->ns, nr = M->nr;
  double c[NSMAX], g[NSMAX];
  const double logT = log(T), invT = 1.0 / T;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  for (int k = 0; k < ns; k++) {
    c[k] = rho * y[k] / M->W[k];
    const double* a = T <= M->tmid[k] ? M->lo[k] : M->hi[k];
    // NASA 7-coefficient polynomials
    cpr[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
    hrt[k] = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 +
             a[5] * invT;
    const double sr = a[0] * logT + a[1] * T + a[2] * T2 / 2.0 + a[3] * T3 / 3.0 +
                      a[4] * T4 / 4.0 + a[6];
    g[k] = hrt[k] - sr;
    w[k] = 0.0;
  }
  const double lc0 = log(M->p0 / (kRgas * T));  // ln of the standard concentration
  for (int r = 0; r < nr; r++) {
    const double kf = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT);
    double fwd = kf;
    double dg = 0.0;
    for (int j = 0; j < M->nre[r]; j++) {
      fwd *= conc_pow(c[M->rsp[r][j]], M->rnu[r][j]);
      dg -= M->rnu[r][j] * g[M->rsp[r][j]];
    }
    double rv = 0.0;
    if (M->rev[r]) {
      double pp = 1.0;
      for (int j = 0; j < M->npr[r]; j++) {
        pp *= conc_pow(c[M->psp[r][j]], M->pnu[r][j]);
        dg += M->pnu[r][j] * g[M->psp[r][j]];
      }
      const double kr = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT + dg - M->dnu[r] * lc0);
      rv = kr * pp;
    }
    double q = fwd - rv;
    if (M->tb[r]) {
      double cm = 0.0;
      for (int k = 0; k < ns; k++) cm += M->eff[r][k] * c[k];
      q *= cm;
    }
    for (int j = 0; j < M->nre[r]; j++) w[M->rsp[r][j]] -= M->rnu[r][j] * q;
    for (int j = 0; j < M->npr[r]; j++) w[M->psp[r][j]] += M->pnu[r][j] * q;
  }
}

// The net molar production rates w (kmol/m^3/s) of one cell at (T, rho, y):
// k_chem's arithmetic, in the same order, without the mixture cp and h.
// k_chem keeps its own copy: with this function inlined in it, its register
// peak rose from 164 VGPRs to 256 and its occupancy fell from three waves per
// SIMD to one.  NSA as in k_chem.
template <int NSA>
__device__ __forceinline__ void cell_rates(const Mech* __restrict__ M, double T, double rho,
                                           const double* y, double* w) {
  const int ns = NSA == NSMAX ? M->ns : NSA, nr = M->nr;
  double c[NSA], g[NSA];
  const double logT = log(T), invT = 1.0 / T;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  for (int k = 0; k < ns; k++) {
    c[k] = rho * y[k] / M->W[k];
    double hrt, sr;
    {
      double hr[2], srr[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const double* a = q == 0 ? M->lo[k] : M->hi[k];
        hr[q] = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 +
                a[5] * invT;
        srr[q] = a[0] * logT + a[1] * T + a[2] * T2 / 2.0 + a[3] * T3 / 3.0 + a[4] * T4 / 4.0 + a[6];
      }
      const int q = T <= M->tmid[k] ? 0 : 1;
      hrt = q == 0 ? hr[0] : hr[1];
      sr = q == 0 ? srr[0] : srr[1];
    }
    g[k] = hrt - sr;
    w[k] = 0.0;
  }
  const double lc0 = log(M->p0 / (kRgas * T));  // ln of the standard concentration
  for (int r = 0; r < nr; r++) {
    const double kf = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT);
    double fwd = kf;
    double dg = 0.0;
    for (int j = 0; j < M->nre[r]; j++) {
      fwd *= conc_pow(c[M->rsp[r][j]], M->rnu[r][j]);
      dg -= M->rnu[r][j] * g[M->rsp[r][j]];
    }
    double rv = 0.0;
    if (M->rev[r]) {
      double pp = 1.0;
      for (int j = 0; j < M->npr[r]; j++) {
        pp *= conc_pow(c[M->psp[r][j]], M->pnu[r][j]);
        dg += M->pnu[r][j] * g[M->psp[r][j]];
      }
      // kr = kf / Kc, Kc = exp(-dG/RT) * (p0/RT)^dnu
      const double kr = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT + dg - M->dnu[r] * lc0);
      rv = kr * pp;
    }
    double q = fwd - rv;
    if (M->tb[r]) {
      double cm = 0.0;
      for (int k = 0; k < ns; k++) cm += M->eff[r][k] * c[k];
      q *= cm;
    }
    for (int j = 0; j < M->nre[r]; j++) w[M->rsp[r][j]] -= M->rnu[r][j] * q;
    for (int j = 0; j < M->npr[r]; j++) w[M->psp[r][j]] += M->pnu[r][j] * q;
  }
}

// NSA: the per-cell arrays' size.  NSA = NSMAX serves any mechanism (loops
// to M->ns); an instance with NSA = the mechanism's species count keeps four
// arrays of that size in registers instead of NSMAX, with the same arithmetic
// in the same order.
template <int NSA>
__global__ __launch_bounds__(256) void k_chem(const Mech* __restrict__ M, long long n,
                                              const double* __restrict__ Tin,
                                              const double* __restrict__ RPin, int input_is_p,
                                              const double* __restrict__ Y, double* __restrict__ wdot,
                                              double* __restrict__ hmix, double* __restrict__ cpmix,
                                              double* __restrict__ hk) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = NSA == NSMAX ? M->ns : NSA, nr = M->nr;
  const double T = Tin[i];
  double y[NSA], c[NSA], g[NSA], w[NSA];
  double winv = 0.0;
  for (int k = 0; k < ns; k++) {
    y[k] = Y[k * n + i];
    winv += y[k] / M->W[k];
  }
  const double rho = input_is_p ? RPin[i] / (kRgas * T * winv) : RPin[i];
  const double logT = log(T), invT = 1.0 / T;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  double cp = 0.0, h = 0.0;
  for (int k = 0; k < ns; k++) {
    c[k] = rho * y[k] / M->W[k];
    // NASA 7-coefficient polynomials, evaluated with both ranges' coefficients
    // (uniform, so they stay in scalar registers) and the cell's range picked
    double cpr, hrt, sr;
    {
      double cr[2], hr[2], srr[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const double* a = q == 0 ? M->lo[k] : M->hi[k];
        cr[q] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
        hr[q] = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 +
                a[5] * invT;
        srr[q] = a[0] * logT + a[1] * T + a[2] * T2 / 2.0 + a[3] * T3 / 3.0 + a[4] * T4 / 4.0 + a[6];
      }
      const int q = T <= M->tmid[k] ? 0 : 1;
      cpr = q == 0 ? cr[0] : cr[1];
      hrt = q == 0 ? hr[0] : hr[1];
      sr = q == 0 ? srr[0] : srr[1];
    }
    g[k] = hrt - sr;
    cp += y[k] * cpr / M->W[k];
    const double hks = hrt * (kRref / M->W[k]) * T;
    h += y[k] * hks;
    if (hk) hk[k * n + i] = hks;
    w[k] = 0.0;
  }
  const double lc0 = log(M->p0 / (kRgas * T));  // ln of the standard concentration
  for (int r = 0; r < nr; r++) {
    const double kf = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT);
    double fwd = kf;
    double dg = 0.0;
    for (int j = 0; j < M->nre[r]; j++) {
      fwd *= conc_pow(c[M->rsp[r][j]], M->rnu[r][j]);
      dg -= M->rnu[r][j] * g[M->rsp[r][j]];
    }
    double rv = 0.0;
    if (M->rev[r]) {
      double pp = 1.0;
      for (int j = 0; j < M->npr[r]; j++) {
        pp *= conc_pow(c[M->psp[r][j]], M->pnu[r][j]);
        dg += M->pnu[r][j] * g[M->psp[r][j]];
      }
      // kr = kf / Kc, Kc = exp(-dG/RT) * (p0/RT)^dnu
      const double kr = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT + dg - M->dnu[r] * lc0);
      rv = kr * pp;
    }
    double q = fwd - rv;
    if (M->tb[r]) {
      double cm = 0.0;
      for (int k = 0; k < ns; k++) cm += M->eff[r][k] * c[k];
      q *= cm;
    }
    for (int j = 0; j < M->nre[r]; j++) w[M->rsp[r][j]] -= M->rnu[r][j] * q;
    for (int j = 0; j < M->npr[r]; j++) w[M->psp[r][j]] += M->pnu[r][j] * q;
  }
  for (int k = 0; k < ns; k++) wdot[k * n + i] = w[k] * M->W[k];
  if (hmix) hmix[i] = h;
  if (cpmix) cpmix[i] = cp * kRgas;
}

// srccomb (solver.F90:500-537): at every computed point (i, j, k) of
// 0:im x 0:jm x 0:km, chemrate of (rho, tmp, spc) (wirate = wdot_k W_k,
// thermchem.F90:1553-1575) times jacob added to qrhs(i,j,k,6:5+ns).  rho,
// tmp, jacob: (-hm:im+hm, -hm:jm+hm, -hm:km+hm), spc: the same with a last
// species index, qrhs: (0:im, 0:jm, 0:km, 1:numq), all column-major as the
// reference allocates them (commarray.F90).
template <int NSA>
__global__ __launch_bounds__(256) void k_srccomb(const Mech* __restrict__ M, int im, int jm, int km,
                                                 int hm, const double* __restrict__ rho,
                                                 const double* __restrict__ tmp,
                                                 const double* __restrict__ spc,
                                                 const double* __restrict__ jacob,
                                                 double* __restrict__ qrhs) {
  const long long nx = im + 1, ny = jm + 1, ncell = nx * ny * (km + 1);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int ns = NSA == NSMAX ? M->ns : NSA;
  const long long i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
  const long long sx = im + 1 + 2 * hm, sy = jm + 1 + 2 * hm, sz = km + 1 + 2 * hm;
  const long long hp = ((k + hm) * sy + (j + hm)) * sx + (i + hm);
  const long long hs = sx * sy * sz;
  double y[NSA], w[NSA];
  for (int s = 0; s < ns; s++) y[s] = spc[s * hs + hp];
  cell_rates<NSA>(M, tmp[hp], rho[hp], y, w);
  const double jac = jacob[hp];
  for (int s = 0; s < ns; s++) {
    double* q = qrhs + (5 + s) * ncell + c;
    *q = *q + (w[s] * M->W[s]) * jac;
  }
}

// ---------------------------------------------------------------------------
// temperature_calc (thermchem.F90:1775-1786): T from the mass-specific
// internal energy e at fixed density and composition (Cantera setState_UV;
// for an ideal gas e does not depend on density).  Newton on
//   e(T) = sum_k Y_k (h_k(T) - R T / W_k),  de/dT = cv = cp - R sum Y_k/W_k
// with the NASA-7 polynomials and Cantera's R, from the guess in T.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void e_cv_of(const Mech* __restrict__ M, const double* y, double T,
                                        double& e, double& cv) {
  const int ns = M->ns;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  double h = 0.0, cp = 0.0, winv = 0.0;
  for (int k = 0; k < ns; k++) {
    const double* a = T <= M->tmid[k] ? M->lo[k] : M->hi[k];
    const double cpr = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
    const double hrt = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 +
                       a[5] / T;
    const double yw = y[k] / M->W[k];
    h += yw * hrt * T;
    cp += yw * cpr;
    winv += yw;
  }
  e = kRgas * (h - winv * T);
  cv = kRgas * (cp - winv);
}

__global__ __launch_bounds__(256) void k_temperature(const Mech* __restrict__ M, long long n,
                                                     const double* __restrict__ Y,
                                                     const double* __restrict__ E,
                                                     double* __restrict__ T, int* __restrict__ iters) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[NSMAX];
  for (int k = 0; k < M->ns; k++) y[k] = Y[k * n + i];
  const double et = E[i];
  double t = T[i];
  int it = 0;
  for (; it < 50; it++) {
    double e, cv;
    e_cv_of(M, y, t, e, cv);
    const double dt = (et - e) / cv;
    t += dt;
    if (fabs(dt) <= 1e-13 * fabs(t)) break;
  }
  T[i] = t;
  if (iters) iters[i] = it + 1;
}

// ---------------------------------------------------------------------------
// 0-D constant-pressure adiabatic reactor (no reference counterpart: the
// reference's 0dreactor is constant-volume).  State y = (T, Y_1..Y_ns):
//   dY_k/dt = wdot_k W_k / rho,  dT/dt = -sum_k h_k dY_k/dt / cp,
// rho = p / (R T sum Y_k/W_k), h_k = (h/RT)_k R T / W_k, cp = R sum Y_k cpr_k/W_k,
// so the mixture enthalpy is conserved.  Integrated by RODAS4, the stiffly
// accurate, L-stable Rosenbrock method of order 4 with an embedded order-3
// solution (Hairer and Wanner, Solving Ordinary Differential Equations II,
// 2nd ed. 1996, Sect. VI.4), in the transformed form of their RODAS code:
// with A = I - gamma h J (J by forward differences, column by column),
//   U_i = gamma h A^-1 ( f(y + sum_{j<i} a_ij U_j) + sum_{j<i} (c_ij/h) U_j ),
// stage 6's argument Y_6 = Y_5 + U_5, y_new = Y_6 + U_6, error estimate U_6;
// step control h_new = h min(6, max(0.2, 0.9 err^(-1/4))) on a scaled
// max-norm.
// ---------------------------------------------------------------------------
constexpr double kRoGamma = 0.25;
// a_ij of stages 2..5 (row i-2 holds a_i1..a_i(i-1)) and c_ij of stages 2..6
constexpr double kRoA[4][4] = {
    {1.544, 0.0, 0.0, 0.0},
    {0.9466785280815826, 0.2557011698983284, 0.0, 0.0},
    {3.314825187068521, 2.896124015972201, 0.9986419139977817, 0.0},
    {1.221224509226641, 6.019134481288629, 12.53708332932087, -0.6878860361058950}};
constexpr double kRoC[5][5] = {
    {-5.6688, 0.0, 0.0, 0.0, 0.0},
    {-2.430093356833875, -0.2063599157091915, 0.0, 0.0, 0.0},
    {-0.1073529058151375, -9.594562251023355, -20.47028614809616, 0.0, 0.0},
    {7.496443313967647, -10.24680431464352, -33.99990352819905, 11.70890893206160, 0.0},
    {8.083246795921522, -7.981132988064893, -31.52159432874371, 16.31930543123136, -6.058818238834054}};

// the step-size factor after a step of scaled error err
__device__ __forceinline__ double ro_factor(double err, bool accepted) {
  if (accepted) return fmin(6.0, fmax(0.2, 0.9 / sqrt(sqrt(fmax(err, 1e-16)))));
  return err < 1e300 ? fmax(0.2, fmin(0.9, 0.9 / sqrt(sqrt(err)))) : 0.2;
}

constexpr int NYMAX = NSMAX + 1;

__device__ void reactor_rhs(const Mech* __restrict__ M, double p, const double* yv, double* f) {
  const int ns = M->ns;
  const double T = yv[0];
  const double* Y = yv + 1;
  double w[NSMAX], cpr[NSMAX], hrt[NSMAX];
  double winv = 0.0;
  for (int k = 0; k < ns; k++) winv += Y[k] / M->W[k];
  const double rho = p / (kRgas * T * winv);
  kinetics(M, T, rho, Y, w, cpr, hrt);
  double cp = 0.0, dh = 0.0;
  for (int k = 0; k < ns; k++) {
    cp += Y[k] * cpr[k] / M->W[k];
    const double dY = w[k] * M->W[k] / rho;
    f[1 + k] = dY;
    dh += hrt[k] * T / M->W[k] * dY;  // (h_k / R) dY_k/dt
  }
  f[0] = -dh / cp;                    // R cancels between h_k and cp
}

// LU with partial pivoting of the n x n matrix a (row-major, stride NYMAX)
__device__ bool lu_factor(double* a, int* piv, int n) {
  for (int k = 0; k < n; k++) {
    int pk = k;
    double m = fabs(a[k * NYMAX + k]);
    for (int i = k + 1; i < n; i++) {
      const double v = fabs(a[i * NYMAX + k]);
      if (v > m) { m = v; pk = i; }
    }
    piv[k] = pk;
    if (!(m > 0.0)) return false;
    if (pk != k)
      for (int j = 0; j < n; j++) {
        const double t = a[k * NYMAX + j];
        a[k * NYMAX + j] = a[pk * NYMAX + j];
        a[pk * NYMAX + j] = t;
      }
    const double inv = 1.0 / a[k * NYMAX + k];
    for (int i = k + 1; i < n; i++) {
      const double l = a[i * NYMAX + k] * inv;
      a[i * NYMAX + k] = l;
      for (int j = k + 1; j < n; j++) a[i * NYMAX + j] -= l * a[k * NYMAX + j];
    }
  }
  return true;
}

// The factorization swaps whole rows (its stored multipliers included), so
// the solve applies every interchange to b first, then the unit lower and the
// upper triangular solves (as LAPACK's getrs does).
__device__ void lu_solve(const double* a, const int* piv, int n, double* b) {
  for (int k = 0; k < n; k++) {
    if (piv[k] != k) {
      const double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  }
  for (int k = 0; k < n; k++)
    for (int i = k + 1; i < n; i++) b[i] -= a[i * NYMAX + k] * b[k];
  for (int k = n - 1; k >= 0; k--) {
    double s = b[k];
    for (int j = k + 1; j < n; j++) s -= a[k * NYMAX + j] * b[j];
    b[k] = s / a[k * NYMAX + k];
  }
}

struct ReactorPool {
  double* y;     // NYMAX x cap: T then Y_k
  double* t;     // time reached
  double* h;     // next step size
  double* tend;
  double* p;
  int* status;   // 0 free, 1 running, 2 done, 3 failed
  int* nstep;    // accepted steps
  int* nrej;     // rejected steps
  int cap;
};

enum { RX_FREE = 0, RX_RUN = 1, RX_DONE = 2, RX_FAIL = 3 };

// Advance every running reactor by up to max_tries step attempts (or to its
// end time).  A reactor's steps depend only on its own state, so splitting a
// run into launches, or sharing a launch with other reactors, does not change
// its result.
__global__ __launch_bounds__(64) void k_reactor(const Mech* __restrict__ M, ReactorPool P,
                                                int max_tries, double rtol, double atol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.cap || P.status[i] != RX_RUN) return;
  const int n = M->ns + 1;
  const double gam = kRoGamma;
  double y[NYMAX], f0[NYMAX], f1[NYMAX], yt[NYMAX], y5[NYMAX], U[6][NYMAX];
  double a[NYMAX * NYMAX];
  int piv[NYMAX];
  for (int k = 0; k < n; k++) y[k] = P.y[(size_t)k * P.cap + i];
  double t = P.t[i], h = P.h[i];
  const double tend = P.tend[i], p = P.p[i];
  int nstep = P.nstep[i], nrej = P.nrej[i], st = RX_RUN;
  for (int it = 0; it < max_tries; it++) {
    if (t >= tend) { st = RX_DONE; break; }
    if (h > tend - t) h = tend - t;
    if (!(h > 1e-18)) { st = RX_FAIL; break; }
    reactor_rhs(M, p, y, f0);
    // forward-difference Jacobian, column by column
    for (int j = 0; j < n; j++) {
      for (int k = 0; k < n; k++) yt[k] = y[k];
      // floor 1e-6 for mass fractions: a smaller step on a trace species
      // leaves the difference quotient to rounding
      const double d = 1.4901161193847656e-08 * fmax(fabs(y[j]), j == 0 ? 1.0 : 1e-6);
      yt[j] = y[j] + d;
      reactor_rhs(M, p, yt, f1);
      for (int k = 0; k < n; k++) a[k * NYMAX + j] = -gam * h * (f1[k] - f0[k]) / d;
    }
    for (int k = 0; k < n; k++) a[k * NYMAX + k] += 1.0;
    bool ok = lu_factor(a, piv, n);
    double err = 2.0;
    if (ok) {
      const double gh = gam * h;
      for (int k = 0; k < n; k++) U[0][k] = f0[k];
      lu_solve(a, piv, n, U[0]);
      for (int k = 0; k < n; k++) U[0][k] = gh * U[0][k];
      for (int s = 1; s < 6; s++) {
        for (int k = 0; k < n; k++) {
          double v = y[k];
          if (s < 5) {
            for (int j = 0; j < s; j++) v += kRoA[s - 1][j] * U[j][k];
            if (s == 4) y5[k] = v;
          } else {
            v = y5[k] + U[4][k];
          }
          yt[k] = v;
        }
        reactor_rhs(M, p, yt, f1);
        for (int k = 0; k < n; k++) {
          double c = 0.0;
          for (int j = 0; j < s; j++) c += (kRoC[s - 1][j] / h) * U[j][k];
          U[s][k] = f1[k] + c;
        }
        lu_solve(a, piv, n, U[s]);
        for (int k = 0; k < n; k++) U[s][k] = gh * U[s][k];
      }
      err = 0.0;
      for (int k = 0; k < n; k++) {
        const double yn = yt[k] + U[5][k];
        const double sc = (k == 0 ? 1e-6 : atol) + rtol * fmax(fabs(y[k]), fabs(yn));
        const double e = fabs(U[5][k]) / sc;
        if (!(e <= err)) err = e;  // a NaN makes err NaN and rejects the step
        yt[k] = yn;
      }
    }
    if (err <= 1.0) {
      t += h;
      for (int k = 0; k < n; k++) y[k] = yt[k];
      nstep++;
      h *= ro_factor(err, true);
    } else {
      nrej++;
      h *= ro_factor(err, false);
    }
    if (t >= tend) { st = RX_DONE; break; }
  }
  for (int k = 0; k < n; k++) P.y[(size_t)k * P.cap + i] = y[k];
  P.t[i] = t;
  P.h[i] = h;
  P.nstep[i] = nstep;
  P.nrej[i] = nrej;
  P.status[i] = st;
}

// ---------------------------------------------------------------------------
// The same reactor for a mechanism of exactly NS species, with every species
// loop unrolled so the per-reactor state stays in registers (the generic
// kernel's arrays, indexed at run time, live in scratch memory).  Species
// picked by a reaction (uniform indices) are read and updated through select
// chains; the Rosenbrock matrix of each thread sits in LDS, element (i,j) of
// thread t at [(i*N + j)*64 + t].  Same method, same arithmetic order as
// k_reactor.
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ double pick(const double (&v)[NS], int idx) {
  double r = v[0];
#pragma unroll
  for (int k = 1; k < NS; k++) r = idx == k ? v[k] : r;
  return r;
}

template <int NS>
__device__ __forceinline__ void add_at(double (&v)[NS], int idx, double x) {
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (idx == k) v[k] += x;
}

// Lanes of a reactor team (one DPP row); the team kernel is further down.
constexpr int kTeam = 16;

// The sum of v over the 16 lanes of a team: a fixed tree into the team's
// lane 0 (the same association on every team), then lane 0's sum to every
// lane, so that all lanes of the team hold the same bits.
__device__ __forceinline__ double team_sum(double v) {
  v += __shfl_down(v, 8, kTeam);
  v += __shfl_down(v, 4, kTeam);
  v += __shfl_down(v, 2, kTeam);
  v += __shfl_down(v, 1, kTeam);
  return __shfl(v, 0, kTeam);
}

// Inlined at every call, so that the caller's state arrays stay in registers
// (a call would take them by address, in scratch memory).  LANES = kTeam:
// the team evaluates one right-hand side together, lane l taking reactions
// l, l + 16, ...; each species rate is then summed over the team (team_sum),
// and every lane ends with the same f.  LANES = 1: one thread, every
// reaction in order.
template <int NS, int LANES = 1>
__device__ __forceinline__ void reactor_rhs_n(const Mech* __restrict__ M, double p, const double (&yv)[NS + 1],
                                              double (&f)[NS + 1], int lane = 0) {
  const double T = yv[0];
  double c[NS], g[NS], w[NS], cpr[NS], hrt[NS];
  double winv = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) winv += yv[1 + k] / M->W[k];
  const double rho = p / (kRgas * T * winv);
  const double logT = log(T), invT = 1.0 / T;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    c[k] = rho * yv[1 + k] / M->W[k];
    const double* a = T <= M->tmid[k] ? M->lo[k] : M->hi[k];
    cpr[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
    hrt[k] = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 +
             a[5] * invT;
    const double sr = a[0] * logT + a[1] * T + a[2] * T2 / 2.0 + a[3] * T3 / 3.0 +
                      a[4] * T4 / 4.0 + a[6];
    g[k] = hrt[k] - sr;
    w[k] = 0.0;
  }
  const double lc0 = log(M->p0 / (kRgas * T));
  for (int r = LANES > 1 ? lane : 0; r < M->nr; r += LANES) {
    const double kf = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT);
    double fwd = kf;
    double dg = 0.0;
    for (int j = 0; j < M->nre[r]; j++) {
      const int sp = M->rsp[r][j];
      fwd *= conc_pow(pick<NS>(c, sp), M->rnu[r][j]);
      dg -= M->rnu[r][j] * pick<NS>(g, sp);
    }
    double rv = 0.0;
    if (M->rev[r]) {
      double pp = 1.0;
      for (int j = 0; j < M->npr[r]; j++) {
        const int sp = M->psp[r][j];
        pp *= conc_pow(pick<NS>(c, sp), M->pnu[r][j]);
        dg += M->pnu[r][j] * pick<NS>(g, sp);
      }
      const double kr = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT + dg - M->dnu[r] * lc0);
      rv = kr * pp;
    }
    double q = fwd - rv;
    if (M->tb[r]) {
      double cm = 0.0;
#pragma unroll
      for (int k = 0; k < NS; k++) cm += M->eff[r][k] * c[k];
      q *= cm;
    }
    for (int j = 0; j < M->nre[r]; j++) add_at<NS>(w, M->rsp[r][j], -M->rnu[r][j] * q);
    for (int j = 0; j < M->npr[r]; j++) add_at<NS>(w, M->psp[r][j], M->pnu[r][j] * q);
  }
  if constexpr (LANES > 1) {
#pragma unroll
    for (int k = 0; k < NS; k++) w[k] = team_sum(w[k]);
  }
  double cp = 0.0, dh = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    cp += yv[1 + k] * cpr[k] / M->W[k];
    const double dY = w[k] * M->W[k] / rho;
    f[1 + k] = dY;
    dh += hrt[k] * T / M->W[k] * dY;
  }
  f[0] = -dh / cp;
}

// c^(nu-1) nu: the derivative of c^nu
__device__ __forceinline__ double conc_pow_d(double c, double nu) {
  return nu == 1.0 ? 1.0 : (nu == 2.0 ? 2.0 * c : nu * conc_pow_general(c, nu - 1.0));
}

// Column j (0: T, 1 + m: Y_m) of the reactor's Jacobian J = d f / d(T, Y),
// analytic, by the chain rule through the concentrations c_k = rho Y_k / W_k:
//   dc_k/dY_m = rho/W_m delta_km - c_k / (W_m sum_i Y_i/W_i),  dc_k/dT = -c_k/T,
// so per reaction, with s_r = sum_i c_i dq_r/dc_i (mass action is
// homogeneous: s_r = o_f fwd - o_r rev, times [M] and plus [M] (fwd - rev)
// for a three-body reaction, o = the sum of the orders),
//   dq_r/dY_m = dq_r/dc_m rho/W_m - s_r / (W_m sum Y/W),
//   dq_r/dT   = [M] (fwd dln kf/dT - rev dln kr/dT) - s_r / T,
//   dln kf/dT = b/T + E/(R T^2),  dln kr/dT = dln kf/dT - dH_r/(R T^2) + dnu_r/T;
// then f_Y = w W / rho and f_T = -sum_k h_k f_Yk / cp give the column.  f0 is
// f(y).  Every column is computed alone, so a team evaluates them on separate
// lanes with the arithmetic of the one-thread kernel's sequential loop.
template <int NS>
__device__ __forceinline__ void reactor_jac_col(const Mech* __restrict__ M, double p, const double (&yv)[NS + 1],
                                                const double (&f0)[NS + 1], int j, double (&col)[NS + 1]) {
#pragma clang fp contract(off)
  const double T = yv[0];
  double c[NS], g[NS], cpr[NS], hrt[NS], dw[NS];
  double winv = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) winv += yv[1 + k] / M->W[k];
  const double rho = p / (kRgas * T * winv);
  const double logT = log(T), invT = 1.0 / T;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  double dcpm = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    c[k] = rho * yv[1 + k] / M->W[k];
    const double* a = T <= M->tmid[k] ? M->lo[k] : M->hi[k];
    cpr[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
    hrt[k] = a[0] + a[1] * T / 2.0 + a[2] * T2 / 3.0 + a[3] * T3 / 4.0 + a[4] * T4 / 5.0 + a[5] * invT;
    const double sr = a[0] * logT + a[1] * T + a[2] * T2 / 2.0 + a[3] * T3 / 3.0 + a[4] * T4 / 4.0 + a[6];
    g[k] = hrt[k] - sr;
    dw[k] = 0.0;
    if (j == 0) dcpm += yv[1 + k] * (a[1] + 2.0 * a[2] * T + 3.0 * a[3] * T2 + 4.0 * a[4] * T3) / M->W[k];
  }
  const int m = j - 1;
  const double cm_ = m >= 0 ? pick<NS>(c, m) : 0.0;
  const double Wm = m >= 0 ? M->W[m] : 1.0;
  const double lc0 = log(M->p0 / (kRgas * T));
  for (int r = 0; r < M->nr; r++) {
    const double kf = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT);
    double fwd = kf, ofw = 0.0, dfw = 0.0, dg = 0.0, dh = 0.0;
    for (int q = 0; q < M->nre[r]; q++) {
      const int sp = M->rsp[r][q];
      const double nu = M->rnu[r][q];
      fwd *= conc_pow(pick<NS>(c, sp), nu);
      ofw += nu;
      dg -= nu * pick<NS>(g, sp);
      dh -= nu * pick<NS>(hrt, sp);
      if (sp == m) {
        double d = kf * conc_pow_d(cm_, nu);
        for (int q2 = 0; q2 < M->nre[r]; q2++)
          if (q2 != q) d *= conc_pow(pick<NS>(c, M->rsp[r][q2]), M->rnu[r][q2]);
        dfw += d;
      }
    }
    double rv = 0.0, orv = 0.0, drv = 0.0, dlkr = 0.0;
    const double dlkf = M->b[r] * invT + M->EaR[r] * invT * invT;
    if (M->rev[r]) {
      for (int q = 0; q < M->npr[r]; q++) {
        const int sp = M->psp[r][q];
        dg += M->pnu[r][q] * pick<NS>(g, sp);
        dh += M->pnu[r][q] * pick<NS>(hrt, sp);
      }
      const double kr = exp(M->lnA[r] + M->b[r] * logT - M->EaR[r] * invT + dg - M->dnu[r] * lc0);
      rv = kr;
      for (int q = 0; q < M->npr[r]; q++) {
        const int sp = M->psp[r][q];
        const double nu = M->pnu[r][q];
        rv *= conc_pow(pick<NS>(c, sp), nu);
        orv += nu;
        if (sp == m) {
          double d = kr * conc_pow_d(cm_, nu);
          for (int q2 = 0; q2 < M->npr[r]; q2++)
            if (q2 != q) d *= conc_pow(pick<NS>(c, M->psp[r][q2]), M->pnu[r][q2]);
          drv += d;
        }
      }
      dlkr = dlkf - dh * invT + M->dnu[r] * invT;
    }
    double s, dq;
    if (M->tb[r]) {
      double cmb = 0.0;
#pragma unroll
      for (int k = 0; k < NS; k++) cmb += M->eff[r][k] * c[k];
      s = cmb * ((ofw + 1.0) * fwd - (orv + 1.0) * rv);
      dq = m >= 0 ? cmb * (dfw - drv) + (fwd - rv) * M->eff[r][m] : cmb * (fwd * dlkf - rv * dlkr);
    } else {
      s = ofw * fwd - orv * rv;
      dq = m >= 0 ? dfw - drv : fwd * dlkf - rv * dlkr;
    }
    const double dqy = m >= 0 ? dq * rho / Wm - s / (Wm * winv) : dq - s * invT;
    for (int q = 0; q < M->nre[r]; q++) add_at<NS>(dw, M->rsp[r][q], -M->rnu[r][q] * dqy);
    for (int q = 0; q < M->npr[r]; q++) add_at<NS>(dw, M->psp[r][q], M->pnu[r][q] * dqy);
  }
  double cp = 0.0, dH = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const double fY = f0[1 + k];
    const double v = M->W[k] / rho * dw[k] + (m >= 0 ? fY / (Wm * winv) : fY * invT);
    col[1 + k] = v;
    cp += yv[1 + k] * cpr[k] / M->W[k];
    dH += hrt[k] * T / M->W[k] * v;
    if (j == 0) dH += cpr[k] / M->W[k] * fY;
  }
  const double dcp = m >= 0 ? pick<NS>(cpr, m) / Wm : dcpm;
  col[0] = -(dH + f0[0] * dcp) / cp;
}

// solve with LU factors in LDS, element (r,c) at a[(r*N + c)*S]: S = 64 for
// the one-thread kernel's interleaved matrices, 1 for a team's own matrix
// The factorization swaps whole rows (its stored multipliers included), so
// every interchange is applied to b before the unit lower triangular solve.
template <int N, int S = 64>
__device__ __forceinline__ void lu_solve_n(const double* a, const int (&piv)[N], double (&b)[N]) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    // the row swap of step k
    const double bk = b[k];
    double bp = bk;
#pragma unroll
    for (int r = k + 1; r < N; r++) bp = piv[k] == r ? b[r] : bp;
#pragma unroll
    for (int r = k + 1; r < N; r++)
      if (piv[k] == r) b[r] = bk;
    b[k] = bp;
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int r = k + 1; r < N; r++) b[r] -= a[(r * N + k) * S] * b[k];
  }
#pragma unroll
  for (int k = N - 1; k >= 0; k--) {
    double sum = b[k];
#pragma unroll
    for (int c = k + 1; c < N; c++) sum -= a[(k * N + c) * S] * b[c];
    b[k] = sum / a[(k * N + k) * S];
  }
}

// RODAS4's six stages from the factored A = I - gamma h J (in LDS, stride S)
// and f0 = f(y): the new state into ynew, the scaled error as the result
template <int NS, int S>
__device__ __forceinline__ double ro_stages(const Mech* __restrict__ M, double p, double h, const double* a,
                                            const int (&piv)[NS + 1], const double (&y)[NS + 1],
                                            const double (&f0)[NS + 1], double (&ynew)[NS + 1], double rtol,
                                            double atol) {
  constexpr int N = NS + 1;
  double U[6][N], yt[N], y5[N], f1[N];
  const double gh = kRoGamma * h;
#pragma unroll
  for (int k = 0; k < N; k++) U[0][k] = f0[k];
  lu_solve_n<N, S>(a, piv, U[0]);
#pragma unroll
  for (int k = 0; k < N; k++) U[0][k] = gh * U[0][k];
#pragma unroll
  for (int s = 1; s < 6; s++) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double v = y[k];
      if (s < 5) {
#pragma unroll
        for (int j = 0; j < s; j++) v += kRoA[s - 1][j] * U[j][k];
        if (s == 4) y5[k] = v;
      } else {
        v = y5[k] + U[4][k];
      }
      yt[k] = v;
    }
    reactor_rhs_n<NS>(M, p, yt, f1);
#pragma unroll
    for (int k = 0; k < N; k++) {
      double c = 0.0;
#pragma unroll
      for (int j = 0; j < s; j++) c += (kRoC[s - 1][j] / h) * U[j][k];
      U[s][k] = f1[k] + c;
    }
    lu_solve_n<N, S>(a, piv, U[s]);
#pragma unroll
    for (int k = 0; k < N; k++) U[s][k] = gh * U[s][k];
  }
  double err = 0.0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    const double yn = yt[k] + U[5][k];
    const double sc = (k == 0 ? 1e-6 : atol) + rtol * fmax(fabs(y[k]), fabs(yn));
    const double e = fabs(U[5][k]) / sc;
    if (!(e <= err)) err = e;  // a NaN makes err NaN and rejects the step
    ynew[k] = yn;
  }
  return err;
}

template <int NS>
__global__ __launch_bounds__(64) void k_reactor_n(const Mech* __restrict__ M, ReactorPool P,
                                                  int max_tries, double rtol, double atol) {
  constexpr int N = NS + 1;
  __shared__ double sA[N * N * 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.cap || P.status[i] != RX_RUN) return;
  double* a = sA + threadIdx.x;
#define AIJ(r, c) a[((r) * N + (c)) * 64]
  const double gam = kRoGamma;
  double y[N], f0[N], f1[N], yt[N];
  int piv[N];
#pragma unroll
  for (int k = 0; k < N; k++) y[k] = P.y[(size_t)k * P.cap + i];
  double t = P.t[i], h = P.h[i];
  const double tend = P.tend[i], p = P.p[i];
  int nstep = P.nstep[i], nrej = P.nrej[i], st = RX_RUN;
  for (int it = 0; it < max_tries; it++) {
    if (t >= tend) { st = RX_DONE; break; }
    if (h > tend - t) h = tend - t;
    if (!(h > 1e-18)) { st = RX_FAIL; break; }
    reactor_rhs_n<NS>(M, p, y, f0);
    // the analytic Jacobian, column by column: A = I - gamma h J
#pragma unroll 1
    for (int j = 0; j < N; j++) {
      reactor_jac_col<NS>(M, p, y, f0, j, f1);
#pragma unroll
      for (int k = 0; k < N; k++) AIJ(k, j) = -gam * h * f1[k];
    }
#pragma unroll
    for (int k = 0; k < N; k++) AIJ(k, k) += 1.0;
    // LU with partial pivoting (rows swapped in LDS)
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; k++) {
      int pk = k;
      double m = fabs(AIJ(k, k));
#pragma unroll
      for (int r = k + 1; r < N; r++) {
        const double v = fabs(AIJ(r, k));
        if (v > m) { m = v; pk = r; }
      }
      piv[k] = pk;
      ok = ok && (m > 0.0);
      if (pk != k) {
        for (int c = 0; c < N; c++) {
          const double tmp = AIJ(k, c);
          AIJ(k, c) = AIJ(pk, c);
          AIJ(pk, c) = tmp;
        }
      }
      const double inv = 1.0 / AIJ(k, k);
#pragma unroll
      for (int r = k + 1; r < N; r++) {
        const double l = AIJ(r, k) * inv;
        AIJ(r, k) = l;
#pragma unroll
        for (int c = k + 1; c < N; c++) AIJ(r, c) -= l * AIJ(k, c);
      }
    }
    double err = 2.0;
    if (ok) err = ro_stages<NS, 64>(M, p, h, a, piv, y, f0, yt, rtol, atol);
    if (err <= 1.0) {
      t += h;
#pragma unroll
      for (int k = 0; k < N; k++) y[k] = yt[k];
      nstep++;
      h *= ro_factor(err, true);
    } else {
      nrej++;
      h *= ro_factor(err, false);
    }
    if (t >= tend) { st = RX_DONE; break; }
  }
#undef AIJ
#pragma unroll
  for (int k = 0; k < N; k++) P.y[(size_t)k * P.cap + i] = y[k];
  P.t[i] = t;
  P.h[i] = h;
  P.nstep[i] = nstep;
  P.nrej[i] = nrej;
  P.status[i] = st;
}


// The reactor's right-hand side and its analytic Jacobian (the integrator's,
// before the -gamma h scaling) at n states, one thread per state: y, f
// row-major N per state, jac N x N per state (row k, column j)
template <int NS>
__global__ __launch_bounds__(64) void k_reactor_eval(const Mech* __restrict__ M, int n, const double* p,
                                                     const double* y, double* f, double* jac) {
  constexpr int N = NS + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double yv[N], f0[N], yt[N], f1[N];
#pragma unroll
  for (int k = 0; k < N; k++) yv[k] = y[(size_t)i * N + k];
  reactor_rhs_n<NS>(M, p[i], yv, f0);
#pragma unroll
  for (int k = 0; k < N; k++) f[(size_t)i * N + k] = f0[k];
  if (!jac) return;
#pragma unroll 1
  for (int j = 0; j < N; j++) {
    reactor_jac_col<NS>(M, p[i], yv, f0, j, f1);
#pragma unroll
    for (int k = 0; k < N; k++) jac[((size_t)i * N + k) * N + j] = f1[k];
  }
  (void)yt;
}

// ---------------------------------------------------------------------------
// The same reactor integrated by a team of 16 lanes (one DPP row), four
// reactors per wavefront, for latency: each right-hand side is evaluated by
// the whole team (the reactions split over the lanes, the species rates
// summed over the team into identical values on every lane), the step's
// Jacobian columns, which do not depend on each other, by separate lanes at
// once (lane j the analytic column j), and the LU elimination runs one row
// per lane.  The pivot search, both triangular solves, the stage sums and
// the step control are done by every lane of the team, in the one-thread
// kernel's order; only the species-rate sums associate differently.
// A team's lanes hold identical scalars (t, h, err), so its control flow is
// uniform; teams of one wavefront may diverge, which is why the kernel uses
// no workgroup barrier.  The team's matrix sits in LDS; a team is part of one
// wavefront, whose LDS accesses complete in order, so a wave barrier (for
// the compiler) orders one lane's write before another lane's read.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ro_stages for a team: the same six stages in the same arithmetic, with the
// stage vectors U_s in the team's LDS (Us, 6 N doubles) instead of in each
// lane's registers.  Every lane computes each stage (the right-hand side and
// the triangular solves, as the one-thread kernel does); lane k < N then
// stores component k of U_s, and the next stage reads them all back.  The
// stage loop stays rolled, so that the right-hand side is inlined once.
template <int NS>
__device__ __forceinline__ double ro_stages_team(const Mech* __restrict__ M, double p, double h, const double* a,
                                                 const int (&piv)[NS + 1], const double (&y)[NS + 1],
                                                 const double (&f0)[NS + 1], double (&ynew)[NS + 1],
                                                 double rtol, double atol, double* Us, int lane) {
  constexpr int N = NS + 1;
  double u[N], yt[N], f1[N];
  const double gh = kRoGamma * h;
#pragma unroll
  for (int k = 0; k < N; k++) u[k] = f0[k];
  lu_solve_n<N, 1>(a, piv, u);
#pragma unroll
  for (int k = 0; k < N; k++) u[k] = gh * u[k];
  if (lane < N) Us[lane] = pick<N>(u, lane);
  team_sync();
#pragma unroll 1
  for (int s = 1; s < 6; s++) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double v = y[k];
      // Y_s = y + sum_j a_sj U_j (s < 5); Y_6 = Y_5 + U_5
      const int nj = s < 5 ? s : 4;
      const int ra = s < 5 ? s - 1 : 3;
#pragma unroll 1
      for (int j = 0; j < nj; j++) v += kRoA[ra][j] * Us[j * N + k];
      if (s == 5) v = v + Us[4 * N + k];
      yt[k] = v;
    }
    reactor_rhs_n<NS, kTeam>(M, p, yt, f1, lane);
#pragma unroll
    for (int k = 0; k < N; k++) {
      double c = 0.0;
#pragma unroll 1
      for (int j = 0; j < s; j++) c += (kRoC[s - 1][j] / h) * Us[j * N + k];
      u[k] = f1[k] + c;
    }
    lu_solve_n<N, 1>(a, piv, u);
#pragma unroll
    for (int k = 0; k < N; k++) u[k] = gh * u[k];
    if (s < 5) {
      if (lane < N) Us[s * N + lane] = pick<N>(u, lane);
      team_sync();
    }
  }
  double err = 0.0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    const double yn = yt[k] + u[k];
    const double sc = (k == 0 ? 1e-6 : atol) + rtol * fmax(fabs(y[k]), fabs(yn));
    const double e = fabs(u[k]) / sc;
    if (!(e <= err)) err = e;  // a NaN makes err NaN and rejects the step
    ynew[k] = yn;
  }
  return err;
}

template <int NS>
__global__ __launch_bounds__(64) void k_reactor_team(const Mech* __restrict__ M, ReactorPool P,
                                                     int max_tries, double rtol, double atol) {
  constexpr int N = NS + 1;
  constexpr int TPB = 64 / kTeam;
  static_assert(N < kTeam, "a team needs a lane per column and one for the unperturbed state");
  __shared__ double sA[TPB][N * N];
  __shared__ double sU[TPB][6 * N];
  const int team = threadIdx.x / kTeam, lane = threadIdx.x % kTeam;
  const int i = blockIdx.x * TPB + team;
  if (i >= P.cap || P.status[i] != RX_RUN) return;  // the whole team leaves
  double* a = sA[team];
#define AT(r, c) a[(r) * N + (c)]
  const double gam = kRoGamma;
  double y[N], f0[N], f1[N], yt[N];
  int piv[N];
#pragma unroll
  for (int k = 0; k < N; k++) y[k] = P.y[(size_t)k * P.cap + i];
  double t = P.t[i], h = P.h[i];
  const double tend = P.tend[i], p = P.p[i];
  int nstep = P.nstep[i], nrej = P.nrej[i], st = RX_RUN;
  for (int it = 0; it < max_tries; it++) {
    if (t >= tend) { st = RX_DONE; break; }
    if (h > tend - t) h = tend - t;
    if (!(h > 1e-18)) { st = RX_FAIL; break; }
    // f(y) by the whole team, then lane j < N: column j of the analytic Jacobian
    reactor_rhs_n<NS, kTeam>(M, p, y, f0, lane);
    if (lane < N) {
      reactor_jac_col<NS>(M, p, y, f0, lane, f1);
#pragma unroll
      for (int k = 0; k < N; k++) AT(k, lane) = -gam * h * f1[k];
      AT(lane, lane) += 1.0;
    }
    team_sync();
    // LU with partial pivoting: rows swapped one column per lane, then
    // eliminated one row per lane
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; k++) {
      int pk = k;
      double m = fabs(AT(k, k));
#pragma unroll
      for (int r = k + 1; r < N; r++) {
        const double v = fabs(AT(r, k));
        if (v > m) { m = v; pk = r; }
      }
      piv[k] = pk;
      ok = ok && (m > 0.0);
      team_sync();
      if (pk != k && lane < N) {
        const double tmp = AT(k, lane);
        AT(k, lane) = AT(pk, lane);
        AT(pk, lane) = tmp;
      }
      team_sync();
      const double inv = 1.0 / AT(k, k);
      if (lane > k && lane < N) {
        const double l = AT(lane, k) * inv;
        AT(lane, k) = l;
#pragma unroll
        for (int c = k + 1; c < N; c++) AT(lane, c) -= l * AT(k, c);
      }
      team_sync();
    }
    double err = 2.0;
    if (ok) err = ro_stages_team<NS>(M, p, h, a, piv, y, f0, yt, rtol, atol, sU[team], lane);
    // the next step's matrix writes must follow every lane's solve reads
    team_sync();
    if (err <= 1.0) {
      t += h;
#pragma unroll
      for (int k = 0; k < N; k++) y[k] = yt[k];
      nstep++;
      h *= ro_factor(err, true);
    } else {
      nrej++;
      h *= ro_factor(err, false);
    }
    if (t >= tend) { st = RX_DONE; break; }
  }
#undef AT
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) P.y[(size_t)k * P.cap + i] = y[k];
    P.t[i] = t;
    P.h[i] = h;
    P.nstep[i] = nstep;
    P.nrej[i] = nrej;
    P.status[i] = st;
  }
}

// ---------------------------------------------------------------------------
// CTML reading (the subset the reference's mechanisms use)
// ---------------------------------------------------------------------------
std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// text between <tag ...> and </tag> starting at pos; attributes in attr
bool element(const std::string& x, const std::string& tag, size_t& pos, std::string& attr,
             std::string& body) {
  for (;;) {
    size_t a = x.find("<" + tag, pos);
    if (a == std::string::npos) return false;
    const char nc = x[a + 1 + tag.size()];
    if (nc != ' ' && nc != '>' && nc != '\t' && nc != '\n' && nc != '/') {
      pos = a + 1;
      continue;
    }
    size_t e = x.find('>', a);
    if (e == std::string::npos) return false;
    attr = x.substr(a + 1 + tag.size(), e - a - 1 - tag.size());
    if (x[e - 1] == '/') {
      body.clear();
      pos = e + 1;
      return true;
    }
    size_t c = x.find("</" + tag + ">", e);
    if (c == std::string::npos) return false;
    body = x.substr(e + 1, c - e - 1);
    pos = c + tag.size() + 3;
    return true;
  }
}

std::string attr_of(const std::string& attrs, const std::string& key) {
  size_t p = attrs.find(key + "=\"");
  if (p == std::string::npos) return "";
  p += key.size() + 2;
  size_t e = attrs.find('"', p);
  return attrs.substr(p, e - p);
}

// "A:1.0 B:2" -> pairs
std::vector<std::pair<std::string, double>> pairs_of(const std::string& s) {
  std::vector<std::pair<std::string, double>> out;
  std::istringstream is(s);
  std::string tok;
  while (is >> tok) {
    size_t c = tok.find(':');
    if (c == std::string::npos) continue;
    out.push_back({tok.substr(0, c), std::stod(tok.substr(c + 1))});
  }
  return out;
}

std::vector<double> floats_of(const std::string& s) {
  std::vector<double> v;
  std::string t = s;
  for (char& ch : t)
    if (ch == ',') ch = ' ';
  std::istringstream is(t);
  double d;
  while (is >> d) v.push_back(d);
  return v;
}

// atomic weights (kg/kmol) of the elements the reference's mechanisms use
double atomic_weight(const std::string& e) {
  static const std::map<std::string, double> w = {
      {"H", 1.008}, {"O", 15.999}, {"N", 14.007}, {"C", 12.011}, {"AR", 39.95}, {"Ar", 39.95},
      {"HE", 4.002602}, {"He", 4.002602}};
  auto it = w.find(e);
  return it == w.end() ? -1.0 : it->second;
}

// factor turning an activation energy in the given units into E/R (K)
bool ea_factor(const std::string& units, double& f) {
  if (units == "cal/mol") f = 4184.0 / kRgas;
  else if (units == "kcal/mol") f = 4184000.0 / kRgas;
  else if (units == "J/mol") f = 1000.0 / kRgas;
  else if (units == "kJ/mol") f = 1.0e6 / kRgas;
  else if (units == "J/kmol") f = 1.0 / kRgas;
  else if (units == "K") f = 1.0;
  else return false;
  return true;
}

int parse_ctml(const std::string& path, const std::string& phase_id, Mech& m,
               std::vector<std::string>& names) {
  std::ifstream f(path);
  if (!f) return cfail("cannot open mechanism file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string x = ss.str();
  std::memset(&m, 0, sizeof(m));
  // phase: species order
  size_t pos = 0;
  std::string attr, body;
  bool found = false;
  while (element(x, "phase", pos, attr, body)) {
    if (attr_of(attr, "id") == phase_id) { found = true; break; }
  }
  if (!found) return cfail("phase '" + phase_id + "' not found in " + path);
  {
    size_t p2 = 0;
    std::string a2, b2;
    if (!element(body, "speciesArray", p2, a2, b2)) return cfail("phase without speciesArray");
    std::istringstream is(b2);
    std::string nm;
    while (is >> nm) names.push_back(nm);
    p2 = 0;
    if (element(body, "kinetics", p2, a2, b2) && attr_of(a2, "model") != "GasKinetics")
      return cfail("unsupported kinetics model " + attr_of(a2, "model"));
  }
  m.ns = (int)names.size();
  if (m.ns < 1 || m.ns > NSMAX) return cfail("species count out of range");
  std::map<std::string, int> idx;
  for (int k = 0; k < m.ns; k++) idx[names[k]] = k;
  // species data
  int nfound = 0;
  pos = 0;
  m.p0 = -1.0;
  while (element(x, "species", pos, attr, body)) {
    auto it = idx.find(attr_of(attr, "name"));
    if (it == idx.end()) continue;
    const int k = it->second;
    size_t p2 = 0;
    std::string a2, b2;
    if (!element(body, "atomArray", p2, a2, b2)) return cfail("species without atomArray");
    double W = 0.0;
    for (auto& pr : pairs_of(b2)) {
      const double aw = atomic_weight(pr.first);
      if (aw < 0) return cfail("unknown element " + pr.first);
      W += pr.second * aw;
    }
    m.W[k] = W;
    p2 = 0;
    int nr = 0;
    while (element(body, "NASA", p2, a2, b2)) {
      size_t p3 = 0;
      std::string a3, b3;
      if (!element(b2, "floatArray", p3, a3, b3)) return cfail("NASA block without coeffs");
      auto v = floats_of(b3);
      if (v.size() != 7) return cfail("NASA block needs 7 coefficients");
      const double tmin = std::stod(attr_of(a2, "Tmin")), tmax = std::stod(attr_of(a2, "Tmax"));
      const std::string sp0 = attr_of(a2, "P0");
      if (!sp0.empty()) m.p0 = std::stod(sp0);
      double* dst = nr == 0 ? m.lo[k] : m.hi[k];
      for (int j = 0; j < 7; j++) dst[j] = v[j];
      if (nr == 0) m.tmid[k] = tmax;
      else if (tmin < m.tmid[k]) {  // listed high range first: swap
        double tmp[7];
        std::memcpy(tmp, m.lo[k], sizeof(tmp));
        std::memcpy(m.lo[k], m.hi[k], sizeof(tmp));
        std::memcpy(m.hi[k], tmp, sizeof(tmp));
        m.tmid[k] = tmax;
      }
      nr++;
    }
    if (nr != 2) return cfail("species " + names[k] + " needs two NASA ranges");
    nfound++;
  }
  if (nfound != m.ns) return cfail("species data missing for some phase species");
  if (m.p0 <= 0) m.p0 = 101325.0;
  // reactions
  pos = 0;
  int r = 0;
  while (element(x, "reaction", pos, attr, body)) {
    if (r >= NRMAX) return cfail("too many reactions");
    const std::string type = attr_of(attr, "type");
    if (!type.empty() && type != "threeBody") return cfail("unsupported reaction type " + type);
    m.tb[r] = type == "threeBody";
    m.rev[r] = attr_of(attr, "reversible") != "no";
    size_t p2 = 0;
    std::string a2, b2;
    if (!element(body, "Arrhenius", p2, a2, b2)) return cfail("reaction without Arrhenius");
    {
      size_t p3 = 0;
      std::string a3, b3;
      if (!element(b2, "A", p3, a3, b3)) return cfail("Arrhenius without A");
      const double A = std::stod(trim(b3));
      p3 = 0;
      if (!element(b2, "b", p3, a3, b3)) return cfail("Arrhenius without b");
      m.b[r] = std::stod(trim(b3));
      p3 = 0;
      if (!element(b2, "E", p3, a3, b3)) return cfail("Arrhenius without E");
      std::string u = attr_of(a3, "units");
      if (u.empty()) u = "cal/mol";
      double fac = 0.0;
      if (!ea_factor(u, fac)) return cfail("unsupported energy units " + u);
      m.EaR[r] = std::stod(trim(b3)) * fac;
      if (A <= 0) return cfail("non-positive pre-exponential factor");
      m.lnA[r] = std::log(A);
    }
    for (int k = 0; k < m.ns; k++) m.eff[r][k] = 1.0;
    p2 = 0;
    if (element(body, "efficiencies", p2, a2, b2)) {
      const std::string d = attr_of(a2, "default");
      const double def = d.empty() ? 1.0 : std::stod(d);
      for (int k = 0; k < m.ns; k++) m.eff[r][k] = def;
      for (auto& pr : pairs_of(b2)) {
        auto it = idx.find(pr.first);
        if (it != idx.end()) m.eff[r][it->second] = pr.second;
      }
    }
    double sr = 0.0, sp = 0.0;
    p2 = 0;
    if (!element(body, "reactants", p2, a2, b2)) return cfail("reaction without reactants");
    for (auto& pr : pairs_of(b2)) {
      auto it = idx.find(pr.first);
      if (it == idx.end() || m.nre[r] >= NTERM) return cfail("bad reactant " + pr.first);
      m.rsp[r][m.nre[r]] = it->second;
      m.rnu[r][m.nre[r]++] = pr.second;
      sr += pr.second;
    }
    p2 = 0;
    if (!element(body, "products", p2, a2, b2)) return cfail("reaction without products");
    for (auto& pr : pairs_of(b2)) {
      auto it = idx.find(pr.first);
      if (it == idx.end() || m.npr[r] >= NTERM) return cfail("bad product " + pr.first);
      m.psp[r][m.npr[r]] = it->second;
      m.pnu[r][m.npr[r]++] = pr.second;
      sp += pr.second;
    }
    m.dnu[r] = sp - sr;
    r++;
  }
  m.nr = r;
  return 0;
}

}  // namespace

struct astr_chem {
  Mech host;
  Mech* dev = nullptr;
  std::vector<std::string> names;
  int device = 0;
};

struct astr_reactor_pool {
  astr_chem* chem = nullptr;
  ReactorPool P{};
  void* base = nullptr;
  double rtol = 1e-6, atol = 1e-12;
  hipStream_t stream = nullptr;
  int lanes = 1;  // lanes per reactor: 1, or kTeam for the team kernel
};

extern "C" {

const char* astr_hip_chem_last_error(void) { return g_cerr.c_str(); }

int astr_hip_chem_create(const char* ctml_path, const char* phase_id, astr_chem** out) {
  if (!ctml_path || !out) return cfail("null argument");
  *out = nullptr;
  astr_chem* c = new astr_chem();
  if (parse_ctml(ctml_path, phase_id ? phase_id : "gas", c->host, c->names)) {
    delete c;
    return -1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete c;
    return cfail("no HIP device");
  }
  (void)hipGetDevice(&c->device);
  if (hipMalloc(&c->dev, sizeof(Mech)) != hipSuccess) {
    delete c;
    return cfail("hipMalloc of mechanism failed");
  }
  if (hipMemcpy(c->dev, &c->host, sizeof(Mech), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c->dev);
    delete c;
    return cfail("copy of mechanism failed");
  }
  *out = c;
  return 0;
}

int astr_hip_chem_destroy(astr_chem* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipFree(c->dev);
  delete c;
  return 0;
}

int astr_hip_chem_info(astr_chem* c, int* nsp, int* nrx, double* mw, char* names, int name_len) {
  if (!c) return cfail("null mechanism");
  if (nsp) *nsp = c->host.ns;
  if (nrx) *nrx = c->host.nr;
  for (int k = 0; k < c->host.ns; k++) {
    if (mw) mw[k] = c->host.W[k];
    if (names && name_len > 0) {
      std::strncpy(names + (size_t)k * name_len, c->names[k].c_str(), name_len - 1);
      names[(size_t)k * name_len + name_len - 1] = 0;
    }
  }
  return 0;
}

int astr_hip_chem_rates_device(astr_chem* c, long long n, const double* T, const double* rho_or_p,
                               int input_is_p, const double* Y, double* wdot, double* h_mix,
                               double* cp_mix, double* h_k, void* stream) {
  if (!c) return cfail("null mechanism");
  if (n < 0) return cfail("negative cell count");
  if (n == 0) return 0;
  if (!T || !rho_or_p || !Y || !wdot) return cfail("null buffer");
  CHIP_OK(hipSetDevice(c->device));
  const int bs = 256;
  const dim3 grid((unsigned)((n + bs - 1) / bs));
  if (c->host.ns == 9)
    hipLaunchKernelGGL(k_chem<9>, grid, dim3(bs), 0, (hipStream_t)stream, c->dev, n, T, rho_or_p,
                       input_is_p, Y, wdot, h_mix, cp_mix, h_k);
  else
    hipLaunchKernelGGL(k_chem<NSMAX>, grid, dim3(bs), 0, (hipStream_t)stream, c->dev, n, T, rho_or_p,
                       input_is_p, Y, wdot, h_mix, cp_mix, h_k);
  CHIP_OK(hipGetLastError());
  return 0;
}

int astr_hip_chem_rates(astr_chem* c, long long n, const double* T, const double* rho_or_p,
                        int input_is_p, const double* Y, double* wdot, double* h_mix,
                        double* cp_mix, double* h_k) {
  if (!c) return cfail("null mechanism");
  if (n < 0) return cfail("negative cell count");
  if (n == 0) return 0;
  if (!T || !rho_or_p || !Y || !wdot) return cfail("null buffer");
  CHIP_OK(hipSetDevice(c->device));
  const int ns = c->host.ns;
  const size_t b1 = (size_t)n * sizeof(double);
  double* d = nullptr;
  const size_t nbuf = 2 + 2 * (size_t)ns + 2 + (h_k ? ns : 0);
  CHIP_OK(hipMalloc(&d, nbuf * b1));
  double* dT = d;
  double* dR = dT + n;
  double* dY = dR + n;
  double* dW = dY + (size_t)ns * n;
  double* dH = dW + (size_t)ns * n;
  double* dC = dH + n;
  double* dK = h_k ? dC + n : nullptr;
  int rc = 0;
  if (hipMemcpy(dT, T, b1, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dR, rho_or_p, b1, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dY, Y, b1 * ns, hipMemcpyHostToDevice) != hipSuccess) {
    rc = cfail("copy to device failed");
  }
  if (!rc) rc = astr_hip_chem_rates_device(c, n, dT, dR, input_is_p, dY, dW, dH, dC, dK, nullptr);
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = cfail("chemistry kernel failed");
  if (!rc) {
    if (hipMemcpy(wdot, dW, b1 * ns, hipMemcpyDeviceToHost) != hipSuccess ||
        (h_mix && hipMemcpy(h_mix, dH, b1, hipMemcpyDeviceToHost) != hipSuccess) ||
        (cp_mix && hipMemcpy(cp_mix, dC, b1, hipMemcpyDeviceToHost) != hipSuccess) ||
        (h_k && hipMemcpy(h_k, dK, b1 * ns, hipMemcpyDeviceToHost) != hipSuccess))
      rc = cfail("copy from device failed");
  }
  (void)hipFree(d);
  return rc;
}

static int srccomb_check(astr_chem* c, int im, int jm, int km, int hm, int numq, const void* rho,
                         const void* tmp, const void* spc, const void* jacob, const void* qrhs) {
  if (!c) return cfail("srccomb: null mechanism");
  if (im < 0 || jm < 0 || km < 0 || hm < 0) return cfail("srccomb: negative extent or halo");
  if (numq != 5 + c->host.ns)
    return cfail("srccomb: numq must be 5 + the species count (qrhs(6:numq) takes the rates)");
  if (!rho || !tmp || !spc || !jacob || !qrhs) return cfail("srccomb: null buffer");
  return 0;
}

int astr_hip_srccomb_device(astr_chem* c, int im, int jm, int km, int hm, int numq, const double* rho,
                            const double* tmp, const double* spc, const double* jacob, double* qrhs,
                            void* stream) {
  if (int rc = srccomb_check(c, im, jm, km, hm, numq, rho, tmp, spc, jacob, qrhs)) return rc;
  CHIP_OK(hipSetDevice(c->device));
  const long long ncell = (long long)(im + 1) * (jm + 1) * (km + 1);
  const int bs = 256;
  const dim3 grid((unsigned)((ncell + bs - 1) / bs));
  if (c->host.ns == 9)
    hipLaunchKernelGGL(k_srccomb<9>, grid, dim3(bs), 0, (hipStream_t)stream, c->dev, im, jm, km, hm,
                       rho, tmp, spc, jacob, qrhs);
  else
    hipLaunchKernelGGL(k_srccomb<NSMAX>, grid, dim3(bs), 0, (hipStream_t)stream, c->dev, im, jm, km,
                       hm, rho, tmp, spc, jacob, qrhs);
  CHIP_OK(hipGetLastError());
  return 0;
}

int astr_hip_srccomb(astr_chem* c, int im, int jm, int km, int hm, int numq, const double* rho,
                     const double* tmp, const double* spc, const double* jacob, double* qrhs) {
  if (int rc = srccomb_check(c, im, jm, km, hm, numq, rho, tmp, spc, jacob, qrhs)) return rc;
  CHIP_OK(hipSetDevice(c->device));
  const size_t nh = (size_t)(im + 1 + 2 * hm) * (jm + 1 + 2 * hm) * (km + 1 + 2 * hm);
  const size_t nq = (size_t)(im + 1) * (jm + 1) * (km + 1) * numq;
  const size_t ns = c->host.ns;
  double* d = nullptr;
  CHIP_OK(hipMalloc(&d, ((3 + ns) * nh + nq) * sizeof(double)));
  double *dr = d, *dt = dr + nh, *dj = dt + nh, *ds = dj + nh, *dq = ds + ns * nh;
  const size_t b = nh * sizeof(double);
  int rc = 0;
  if (hipMemcpy(dr, rho, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dt, tmp, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dj, jacob, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ds, spc, b * ns, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dq, qrhs, nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = cfail("srccomb: copy to device failed");
  if (!rc) rc = astr_hip_srccomb_device(c, im, jm, km, hm, numq, dr, dt, ds, dj, dq, nullptr);
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = cfail("srccomb: kernel failed");
  if (!rc && hipMemcpy(qrhs, dq, nq * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = cfail("srccomb: copy from device failed");
  (void)hipFree(d);
  return rc;
}

int astr_hip_reactor_rhs(astr_chem* chem, int n, const double* p, const double* y, double* f, double* jac) {
  if (!chem || !p || !y || !f || n < 1) return cfail("reactor_rhs: null argument or no states");
  if (chem->host.ns != 9) return cfail("reactor_rhs: 9-species mechanisms only (the unrolled kernel)");
  CHIP_OK(hipSetDevice(chem->device));
  constexpr int N = 10;
  double* d = nullptr;
  const size_t ny = (size_t)n * N, nj = jac ? (size_t)n * N * N : 0;
  CHIP_OK(hipMalloc(&d, (n + 2 * ny + nj) * sizeof(double)));
  double *dp = d, *dy = d + n, *df = dy + ny, *dj = jac ? df + ny : nullptr;
  int rc = 0;
  if (hipMemcpy(dp, p, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy, y, ny * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    rc = cfail("reactor_rhs: copy in failed");
  } else {
    hipLaunchKernelGGL(k_reactor_eval<9>, dim3((n + 63) / 64), dim3(64), 0, nullptr, chem->dev, n, dp, dy, df, dj);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(f, df, ny * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        (jac && hipMemcpy(jac, dj, nj * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
      rc = cfail("reactor_rhs: kernel or copy out failed");
  }
  (void)hipFree(d);
  return rc;
}

int astr_hip_reactor_method(double* coef, int n) {
  if (!coef || n < 26) return cfail("reactor_method needs 26 values");
  int k = 0;
  coef[k++] = kRoGamma;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j <= i; j++) coef[k++] = kRoA[i][j];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j <= i; j++) coef[k++] = kRoC[i][j];
  return 0;
}

int astr_hip_reactor_create(astr_chem* chem, int capacity, double rtol, double atol,
                            astr_reactor_pool** out) {
  if (!chem || !out || capacity < 1) return cfail("bad argument to astr_hip_reactor_create");
  if (!(rtol > 0.0) || !(atol > 0.0)) return cfail("tolerances must be positive");
  *out = nullptr;
  CHIP_OK(hipSetDevice(chem->device));
  astr_reactor_pool* r = new astr_reactor_pool();
  r->chem = chem;
  r->rtol = rtol;
  r->atol = atol;
  const size_t cap = (size_t)capacity;
  const size_t bytes = (NYMAX + 4) * cap * sizeof(double) + 4 * cap * sizeof(int);
  if (hipMalloc(&r->base, bytes) != hipSuccess) { delete r; return cfail("hipMalloc of reactor pool failed"); }
  double* d = (double*)r->base;
  r->P.y = d;
  r->P.t = d + NYMAX * cap;
  r->P.h = r->P.t + cap;
  r->P.tend = r->P.h + cap;
  r->P.p = r->P.tend + cap;
  int* iv = (int*)(r->P.p + cap);
  r->P.status = iv;
  r->P.nstep = iv + cap;
  r->P.nrej = iv + 2 * cap;
  r->P.cap = capacity;
  r->lanes = chem->host.ns == 9 ? kTeam : 1;
  if (hipMemset(r->base, 0, bytes) != hipSuccess || hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipFree(r->base);
    delete r;
    return cfail("reactor pool set-up failed");
  }
  *out = r;
  return 0;
}

int astr_hip_reactor_destroy(astr_reactor_pool* r) {
  if (!r) return 0;
  (void)hipSetDevice(r->chem->device);
  (void)hipStreamSynchronize(r->stream);
  (void)hipFree(r->base);
  (void)hipStreamDestroy(r->stream);
  delete r;
  return 0;
}

int astr_hip_reactor_set(astr_reactor_pool* r, int slot, double T0, double p, const double* Y0,
                         double t_end, double h0) {
  if (!r || !Y0) return cfail("null argument");
  if (slot < 0 || slot >= r->P.cap) return cfail("reactor slot out of range");
  if (!(T0 > 0.0) || !(p > 0.0) || !(t_end >= 0.0) || !(h0 > 0.0)) return cfail("bad reactor state");
  CHIP_OK(hipSetDevice(r->chem->device));
  const int ns = r->chem->host.ns;
  const size_t cap = (size_t)r->P.cap;
  double yv[NYMAX];
  yv[0] = T0;
  for (int k = 0; k < ns; k++) yv[1 + k] = Y0[k];
  for (int k = 0; k <= ns; k++)
    CHIP_OK(hipMemcpyAsync(r->P.y + k * cap + slot, &yv[k], sizeof(double), hipMemcpyHostToDevice, r->stream));
  const double sc[4] = {0.0, h0, t_end, p};
  double* dst[4] = {r->P.t, r->P.h, r->P.tend, r->P.p};
  for (int q = 0; q < 4; q++)
    CHIP_OK(hipMemcpyAsync(dst[q] + slot, &sc[q], sizeof(double), hipMemcpyHostToDevice, r->stream));
  const int iv[3] = {RX_RUN, 0, 0};
  int* idst[3] = {r->P.status, r->P.nstep, r->P.nrej};
  for (int q = 0; q < 3; q++)
    CHIP_OK(hipMemcpyAsync(idst[q] + slot, &iv[q], sizeof(int), hipMemcpyHostToDevice, r->stream));
  CHIP_OK(hipStreamSynchronize(r->stream));
  return 0;
}

int astr_hip_reactor_advance(astr_reactor_pool* r, int max_tries) {
  if (!r) return cfail("null pool");
  if (max_tries < 1) return cfail("max_tries must be positive");
  CHIP_OK(hipSetDevice(r->chem->device));
  const int bs = 64;
  const dim3 grid((r->P.cap + bs - 1) / bs);
  if (r->lanes == kTeam)
    hipLaunchKernelGGL(k_reactor_team<9>, dim3((r->P.cap + bs / kTeam - 1) / (bs / kTeam)), dim3(bs), 0,
                       r->stream, r->chem->dev, r->P, max_tries, r->rtol, r->atol);
  else if (r->chem->host.ns == 9)
    hipLaunchKernelGGL(k_reactor_n<9>, grid, dim3(bs), 0, r->stream, r->chem->dev, r->P, max_tries,
                       r->rtol, r->atol);
  else
    hipLaunchKernelGGL(k_reactor, grid, dim3(bs), 0, r->stream, r->chem->dev, r->P, max_tries,
                       r->rtol, r->atol);
  CHIP_OK(hipGetLastError());
  return 0;
}

int astr_hip_reactor_set_lanes(astr_reactor_pool* r, int lanes) {
  if (!r) return cfail("null pool");
  if (lanes == kTeam && r->chem->host.ns != 9)
    return cfail("the team kernel is built for mechanisms of 9 species");
  if (lanes != 1 && lanes != kTeam) return cfail("lanes per reactor must be 1 or 16");
  CHIP_OK(hipStreamSynchronize(r->stream));
  r->lanes = lanes;
  return 0;
}

int astr_hip_reactor_get_lanes(astr_reactor_pool* r, int* lanes) {
  if (!r || !lanes) return cfail("null argument");
  *lanes = r->lanes;
  return 0;
}

int astr_hip_reactor_query(astr_reactor_pool* r, int* idle) {
  if (!r || !idle) return cfail("null argument");
  CHIP_OK(hipSetDevice(r->chem->device));
  const hipError_t e = hipStreamQuery(r->stream);
  if (e == hipErrorNotReady) {
    *idle = 0;
    return 0;
  }
  CHIP_OK(e);
  *idle = 1;
  return 0;
}

int astr_hip_reactor_sync(astr_reactor_pool* r) {
  if (!r) return cfail("null pool");
  CHIP_OK(hipStreamSynchronize(r->stream));
  return 0;
}

int astr_hip_reactor_status(astr_reactor_pool* r, int* status) {
  if (!r || !status) return cfail("null argument");
  CHIP_OK(hipSetDevice(r->chem->device));
  CHIP_OK(hipMemcpyAsync(status, r->P.status, r->P.cap * sizeof(int), hipMemcpyDeviceToHost, r->stream));
  CHIP_OK(hipStreamSynchronize(r->stream));
  return 0;
}

int astr_hip_reactor_get(astr_reactor_pool* r, int slot, double* t, double* yv, int* nstep, int* nrej) {
  if (!r) return cfail("null pool");
  if (slot < 0 || slot >= r->P.cap) return cfail("reactor slot out of range");
  CHIP_OK(hipSetDevice(r->chem->device));
  const int n = r->chem->host.ns + 1;
  const size_t cap = (size_t)r->P.cap;
  if (yv)
    for (int k = 0; k < n; k++)
      CHIP_OK(hipMemcpyAsync(&yv[k], r->P.y + k * cap + slot, sizeof(double), hipMemcpyDeviceToHost, r->stream));
  if (t) CHIP_OK(hipMemcpyAsync(t, r->P.t + slot, sizeof(double), hipMemcpyDeviceToHost, r->stream));
  if (nstep) CHIP_OK(hipMemcpyAsync(nstep, r->P.nstep + slot, sizeof(int), hipMemcpyDeviceToHost, r->stream));
  if (nrej) CHIP_OK(hipMemcpyAsync(nrej, r->P.nrej + slot, sizeof(int), hipMemcpyDeviceToHost, r->stream));
  CHIP_OK(hipStreamSynchronize(r->stream));
  return 0;
}

int astr_hip_reactor_release(astr_reactor_pool* r, int slot) {
  if (!r) return cfail("null pool");
  if (slot < 0 || slot >= r->P.cap) return cfail("reactor slot out of range");
  CHIP_OK(hipSetDevice(r->chem->device));
  const int v = RX_FREE;
  CHIP_OK(hipMemcpyAsync(r->P.status + slot, &v, sizeof(int), hipMemcpyHostToDevice, r->stream));
  CHIP_OK(hipStreamSynchronize(r->stream));
  return 0;
}

int astr_hip_chem_temperature_device(astr_chem* c, long long n, const double* Y, const double* e,
                                     double* T, void* stream) {
  if (!c) return cfail("null mechanism");
  if (n < 0) return cfail("negative cell count");
  if (n == 0) return 0;
  if (!Y || !e || !T) return cfail("null buffer");
  CHIP_OK(hipSetDevice(c->device));
  const int bs = 256;
  hipLaunchKernelGGL(k_temperature, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0,
                     (hipStream_t)stream, c->dev, n, Y, e, T, nullptr);
  CHIP_OK(hipGetLastError());
  return 0;
}

int astr_hip_chem_temperature(astr_chem* c, long long n, const double* Y, const double* e, double* T) {
  if (!c) return cfail("null mechanism");
  if (n < 0) return cfail("negative cell count");
  if (n == 0) return 0;
  if (!Y || !e || !T) return cfail("null buffer");
  CHIP_OK(hipSetDevice(c->device));
  const int ns = c->host.ns;
  const size_t b1 = (size_t)n * sizeof(double);
  double* d = nullptr;
  CHIP_OK(hipMalloc(&d, (2 + (size_t)ns) * b1));
  double* dY = d;
  double* dE = dY + (size_t)ns * n;
  double* dT = dE + n;
  int rc = 0;
  if (hipMemcpy(dY, Y, b1 * ns, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dE, e, b1, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dT, T, b1, hipMemcpyHostToDevice) != hipSuccess)
    rc = cfail("copy to device failed");
  if (!rc) rc = astr_hip_chem_temperature_device(c, n, dY, dE, dT, nullptr);
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = cfail("temperature kernel failed");
  if (!rc && hipMemcpy(T, dT, b1, hipMemcpyDeviceToHost) != hipSuccess) rc = cfail("copy from device failed");
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
