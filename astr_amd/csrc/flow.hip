// flow.hip -- the per-stage passes of the compact-scheme flow solver on
// gfx950, and the C-ABI that exposes them (include/astr_hip.h).
//
// Semantics follow the reference's runnable TGV solver
// miniapps/tgv_solver_3d/tgvsolver.F90 on src/derivative.F90 and
// src/filter.F90 with ntype=3 closures (one block holding the whole periodic
// domain).  One RK3 stage (tgvsolver.F90:1037-1074) is executed as these
// line passes, each one HBM sweep over its inputs and outputs:
//
//   GC x, GC y      gradients of u,v,w,T and convective-flux derivatives
//                   along the pass direction (gradcal :450-537,
//                   convection :539-633), accumulated into qrhs
//   GC z (last)     same along z, then qrhs=-qrhs (:982) and the pointwise
//                   stress / heat flux (diffusion :659-694); KE and
//                   enstrophy partial sums on stage 1 (stacal :893-942)
//   V x, V y        viscous-flux derivatives (diffusion :700-787) into qrhs
//   V z (last)      same along z, then the RK3 update (:1053-1059) and the
//                   z sweep of filterq (:857-873), in place on q
//   F x, F y        the x and y sweeps of filterq (:805-847)
//
// The filter sweeps act on different indices and are linear, so applying
// them z,x,y instead of x,y,z gives the same field up to rounding.
// Primitive variables are recomputed from q inside every pass
// (q2fvar :92-121) instead of being stored.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/astr_hip.h"
#include "lines.hpp"

using namespace astr;

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

#define HIP_OK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(std::string(#x) + ": " + hipGetErrorString(e_));                 \
  } while (0)

// ---------------------------------------------------------------------------
// device-side description of one case
// ---------------------------------------------------------------------------
struct FlowArgs {
  int n[3];             // im, jm, km
  int ndir;             // 3, or 2 when km==0
  int pitch;            // doubles between rows (j)
  long long plane;      // doubles between planes (k)
  long long fs;         // doubles between fields
  double inv_d[3];      // 1/dx, 1/dy, 1/dz
  double const2, const5, const6;
  double reynolds, prandtl, suth, suth1;
  double rk_a, rk_b, rk_c, dt;
  int trash;            // a point index inside every field's padding
  double alfa, bhalo;   // filter alpha and halo-row beta (LHS of the filter rows)
  int stage;            // 0,1,2
  int dostats;          // stage-1 KE/enstrophy partial sums
  int store_last_grad;  // last GC pass also stores its gradients
  double* q;            // 5 fields
  double* qsave;        // 5 fields
  double* qrhs;         // 5 fields
  double* dv;           // 12 fields: dvel(m,d) at m+3d, dtmp(d) at 9+d
  double* sg;           // 9 fields: sigma 11,12,13,22,23,33, qflux 1,2,3
  double* stats;        // 2 doubles per tile
  const double* dac[3]; // derivative LU tables [ac1|ac2|ac3] per direction
  const double* fac[3]; // filter LU tables
  FilterCoef fc;
};

enum Prog { GC_FIRST = 0, GC_MID, GC_LAST, V_MID, V_LAST_RHS, V_LAST_RK, F_ONLY };

// ntiles of direction dir for lines-per-tile TL
__host__ __device__ inline int tiles_of(const int n[3], int dir, int TL) {
  const int ni = (n[0] + TL) / TL;  // ceil((im+1)/TL)
  if (dir == 0) return ((n[1] + 1) * (n[2] + 1) + TL - 1) / TL;
  if (dir == 1) return ni * (n[2] + 1);
  return ni * (n[1] + 1);
}

// Blocks are handed to the 8 XCDs round robin; give each XCD a contiguous
// range of tiles so tiles sharing cache lines share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b & 7, s = b >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + s;
}

template <class C>
struct Slots {
  int g[C::NPT];  // offset inside a field; the trash point for lines past the
                  // end of the lattice (partial last tile)
  int s[C::NPT];  // index of the point in a line buffer (l*LT + r)
  int h[C::NPT];  // line-buffer index of the slot's periodic halo copy
  int interior;   // bit t: i,j,k >= 1 (stats domain)
};

// LDS of a pass: the output lines X (TL x LT), then the input lines T
// (TL x LT), then a write-only dummy cell.  A slot holding no point (past the
// tile) or no halo copy points s/h at index TL*LT: its T writes land in the
// dummy cell (T + kLineOff + TL*LT) and its X reads return T's zero padding
// (X + TL*LT = row -8 of T's first line), so no slot needs a guard.
template <class C>
constexpr int lds_doubles() {
  return (2 * C::TL * C::LT > 2 * C::NT ? 2 * C::TL * C::LT : 2 * C::NT) + kLineOff + 8;
}

// Map the tile's points onto slots.  dir 0: lines are (j,k), rows i; dir 1:
// lines (i,k) rows j; dir 2: lines (i,j) rows k.
template <class C>
__device__ __forceinline__ void make_slots(Slots<C>& S, const FlowArgs& a, int dir, int tile) {
  const int n = a.n[dir];
  const int np = C::TL * (n + 1);
  const int ni = (a.n[0] + C::TL) / C::TL;
  const int dummy = C::TL * C::LT;
  S.interior = 0;
#pragma unroll
  for (int t = 0; t < C::NPT; t++) {
    const int e = threadIdx.x + t * C::NT;
    S.g[t] = a.trash;
    S.s[t] = dummy;
    S.h[t] = dummy;
    if (e < np) {
      int l, r, i, j, k;
      bool ok;
      if (dir == 0) {
        l = e / (n + 1);
        r = e - l * (n + 1);
        const int L = tile * C::TL + l;
        j = L % (a.n[1] + 1);
        k = L / (a.n[1] + 1);
        i = r;
        ok = L < (a.n[1] + 1) * (a.n[2] + 1);
      } else {
        l = e & (C::TL - 1);
        r = e / C::TL;
        const int it = tile % ni, o = tile / ni;
        i = it * C::TL + l;
        ok = i <= a.n[0];
        if (dir == 1) { j = r; k = o; } else { j = o; k = r; }
      }
      S.s[t] = l * C::LT + r;
      // rows n-5..n-1 are copied to the halo rows -5..-1, rows 1..5 to
      // n+1..n+5 (disjoint for n >= 11)
      if (r >= n - 5 && r <= n - 1) S.h[t] = S.s[t] - n;
      if (r >= 1 && r <= 5) S.h[t] = S.s[t] + n;
      if (ok) {
        S.g[t] = (int)((long long)k * a.plane + (long long)j * a.pitch + i);
        const bool in = i >= 1 && j >= 1 && (a.ndir == 2 || k >= 1);
        if (in) S.interior |= 1 << t;
      }
    }
  }
}

// store slot t's value into its input line with the periodic halo copies
// q(-h)=q(n-h), q(n+h)=q(h) (tgvsolver.F90:371-440)
template <class C>
__device__ __forceinline__ void put(double* T, const Slots<C>& S, int t, double v) {
  T[kLineOff + S.s[t]] = v;
  T[kLineOff + S.h[t]] = v;
}

// zero the padding rows -8..-6 and n+6..P*R+2 of every input line (never
// written by put; read by the chunks at the ends of a line)
template <class C>
__device__ __forceinline__ void zero_pad(double* T, int n) {
  const int npad = 3 + (C::P * C::R + 2 - (n + 6) + 1);
  for (int e = threadIdx.x; e < C::TL * npad; e += C::NT) {
    const int l = e / npad, k = e - l * npad;
    const int row = k < 3 ? k - 8 : n + 6 + (k - 3);
    T[l * C::LT + kLineOff + row] = 0.0;
  }
}

// q2fvar + thermal_scar (tgvsolver.F90:73-121).  Pointwise thermodynamics is
// evaluated without FMA contraction and in the reference's operation order:
// p ~ 1/(gamma Ma^2) is large against its gradient, so one rounding step in p
// becomes a visible error in grad p.
__device__ __forceinline__ double pressure_of(const FlowArgs& a, double q1, double q2, double q3,
                                              double q4, double q5) {
#pragma clang fp contract(off)
  const double u = q2 / q1, v = q3 / q1, w = q4 / q1;
  return (q5 - 0.5 * q1 * (u * u + v * v + w * w)) / a.const6;
}

struct Prim {
  double r, u, v, w, p, t;
};

__device__ __forceinline__ Prim prim_of(const FlowArgs& a, double q1, double q2, double q3,
                                        double q4, double q5) {
#pragma clang fp contract(off)
  Prim P;
  P.r = q1;
  P.u = q2 / q1;
  P.v = q3 / q1;
  P.w = q4 / q1;
  P.p = (q5 - 0.5 * q1 * (P.u * P.u + P.v * P.v + P.w * P.w)) / a.const6;
  P.t = P.p / q1 * a.const2;
  return P;
}

// convective flux component m along dir (tgvsolver.F90:559-564, 586-591,
// 613-618) from q_m, the velocity along dir and p
__device__ __forceinline__ double conv_flux(int m, int dir, double qm, double q5, double uu,
                                            double p) {
#pragma clang fp contract(off)
  if (m == 4) return (q5 + p) * uu;
  double f = qm * uu;
  if (m == dir + 1) f = f + p;
  return f;
}

// Sutherland (tgvsolver.F90:123-140)
__device__ __forceinline__ double miucal(const FlowArgs& a, double t) {
  return t * sqrt(t) * a.suth1 / (t + a.suth);
}

// q <- a*qsave + b*q + c*qrhs*dt in the reference's order (tgvsolver.F90:1055-1057)
__device__ __forceinline__ double rk_update(const FlowArgs& a, double qs, double qo, double qr) {
#pragma clang fp contract(off)
  return a.rk_a * qs + a.rk_b * qo + a.rk_c * qr * a.dt;
}

// stress, heat flux (tgvsolver.F90:663-691); g[m][d] = d u_m / d x_d
__device__ __forceinline__ void visc_flux(const FlowArgs& a, const double g[3][3],
                                          const double gt[3], const Prim& P, double out[9]) {
#pragma clang fp contract(off)
  const double s11 = g[0][0];
  const double s12 = 0.5 * (g[0][1] + g[1][0]);
  const double s13 = 0.5 * (g[0][2] + g[2][0]);
  const double s22 = g[1][1];
  const double s23 = 0.5 * (g[1][2] + g[2][1]);
  const double s33 = g[2][2];
  const double skk = (1.0 / 3.0) * (s11 + s22 + s33);
  const double miu = miucal(a, P.t) / a.reynolds;
  const double miu2 = 2.0 * miu;
  const double hcc = (miu / a.prandtl) / a.const5;
  out[0] = miu2 * (s11 - skk);
  out[1] = miu2 * s12;
  out[2] = miu2 * s13;
  out[3] = miu2 * (s22 - skk);
  out[4] = miu2 * s23;
  out[5] = miu2 * (s33 - skk);
  out[6] = hcc * gt[0] + out[0] * P.u + out[1] * P.v + out[2] * P.w;
  out[7] = hcc * gt[1] + out[1] * P.u + out[3] * P.v + out[4] * P.w;
  out[8] = hcc * gt[2] + out[2] * P.u + out[4] * P.v + out[5] * P.w;
}

// ---------------------------------------------------------------------------
// the pass kernel
// ---------------------------------------------------------------------------
#ifndef ASTR_MINB_GC
#define ASTR_MINB_GC 2
#endif
#ifndef ASTR_MINB_V
#define ASTR_MINB_V 1
#endif
#ifndef ASTR_MINB_F
#define ASTR_MINB_F 1
#endif
// workgroups per CU the register allocation is held to, per pass kind
template <int PROG>
constexpr int pass_minb() {
  return PROG <= GC_LAST ? ASTR_MINB_GC : (PROG == F_ONLY ? ASTR_MINB_F : ASTR_MINB_V);
}

template <class C, int PROG>
__global__ __launch_bounds__(C::NT, pass_minb<PROG>()) void k_pass(FlowArgs a, int dir) {
  extern __shared__ double smem[];
  constexpr int NPT = C::NPT;
  const int n = a.n[dir];
  double* X = smem;                 // output lines, TL x LT (rows at offset 0)
  double* T = smem + C::TL * C::LT; // input lines, TL x LT
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  Slots<C> S;
  make_slots<C>(S, a, dir, tile);
  zero_pad<C>(T, n);
  const int ln = threadIdx.x / C::P, p = threadIdx.x % C::P;
  double* Tl = T + ln * C::LT;
  double* Xl = X + ln * C::LT;
  const double idn = a.inv_d[dir];
  const long long fs = a.fs;

  if constexpr (PROG == GC_FIRST || PROG == GC_MID || PROG == GC_LAST) {
    constexpr bool LAST = PROG == GC_LAST;
    ChunkCoef<C::R, true> CD;
    load_chunk_coef<C::R, true>(CD, a.dac[dir], n + 3, p * C::R);
    double qv[5][NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++)
#pragma unroll
      for (int m = 0; m < 5; m++) qv[m][t] = a.q[m * fs + S.g[t]];

    // gradients of u, v, w, T along dir (gradcal, tgvsolver.F90:450-537)
    double gz[4][NPT];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        {
          double f = 0.0;
          {
            if (c < 3) {
              f = qv[c + 1][t] / qv[0][t];
            } else {
              const double pr = pressure_of(a, qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t]);
              f = pr / qv[0][t] * a.const2;
            }
          }
          put<C>(T, S, t, f);
        }
      }
      __syncthreads();
      solve_line<C::P, C::R, false>(Tl, Xl, n, p, CD, a.fc);
      __syncthreads();
      const int fld = c < 3 ? c + 3 * dir : 9 + dir;
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        {
          const double v = X[S.s[t]] * idn;
          if (!LAST || a.store_last_grad) a.dv[fld * fs + S.g[t]] = v;
          gz[c][t] = v;
        }
      }
    }

    if constexpr (LAST) {
      // stress and heat flux (diffusion, tgvsolver.F90:659-694) and the
      // stage-1 KE / enstrophy partial sums (stacal, :893-942)
      double sk = 0.0, se = 0.0;
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        const int o = S.g[t];
        double g[3][3], gt[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if (d == dir) {
            g[0][d] = gz[0][t]; g[1][d] = gz[1][t]; g[2][d] = gz[2][t]; gt[d] = gz[3][t];
          } else if (d < a.ndir) {
            g[0][d] = a.dv[(0 + 3 * d) * fs + o];
            g[1][d] = a.dv[(1 + 3 * d) * fs + o];
            g[2][d] = a.dv[(2 + 3 * d) * fs + o];
            gt[d] = a.dv[(9 + d) * fs + o];
          } else {
            g[0][d] = 0.0; g[1][d] = 0.0; g[2][d] = 0.0; gt[d] = 0.0;
          }
        }
        const Prim P = prim_of(a, qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t]);
        double f[9];
        visc_flux(a, g, gt, P, f);
#pragma unroll
        for (int m = 0; m < 9; m++) a.sg[m * fs + o] = f[m];
        if (a.dostats && ((S.interior >> t) & 1)) {
          sk += P.r * (P.u * P.u + P.v * P.v + P.w * P.w);
          const double o1 = g[2][1] - g[1][2];
          const double o2 = g[0][2] - g[2][0];
          const double o3 = g[1][0] - g[0][1];
          se += P.r * (o1 * o1 + o2 * o2 + o3 * o3);
        }
      }
      if (a.dostats) {
        // deterministic block reduction into this tile's slot
        __syncthreads();
        X[threadIdx.x] = sk;
        X[C::NT + threadIdx.x] = se;
        __syncthreads();
        for (int h = C::NT / 2; h > 0; h >>= 1) {
          if ((int)threadIdx.x < h) {
            X[threadIdx.x] += X[threadIdx.x + h];
            X[C::NT + threadIdx.x] += X[C::NT + threadIdx.x + h];
          }
          __syncthreads();
        }
        if (threadIdx.x == 0) {
          a.stats[2 * tile] = X[0];
          a.stats[2 * tile + 1] = X[C::NT];
        }
        __syncthreads();
      }
    }

    // convective flux derivatives along dir (convection, tgvsolver.F90:539-633)
    double ud[NPT], pp[NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++) {
      pp[t] = pressure_of(a, qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t]);
      ud[t] = (dir == 0 ? qv[1][t] : (dir == 1 ? qv[2][t] : qv[3][t])) / qv[0][t];
    }
#pragma unroll
    for (int m = 0; m < 5; m++) {
#pragma unroll
      for (int t = 0; t < NPT; t++)
        put<C>(T, S, t, conv_flux(m, dir, qv[m][t], qv[4][t], ud[t], pp[t]));
      double* dst = a.qrhs + m * fs;
      double qr[NPT];
      if (PROG != GC_FIRST) {
        // issued before the solve so the load latency overlaps it
#pragma unroll
        for (int t = 0; t < NPT; t++) qr[t] = dst[S.g[t]];
      }
      __syncthreads();
      solve_line<C::P, C::R, false>(Tl, Xl, n, p, CD, a.fc);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        {
          const double v = X[S.s[t]] * idn;
          if (PROG == GC_FIRST) dst[S.g[t]] = v;
          else if (PROG == GC_MID) dst[S.g[t]] = qr[t] + v;
          else dst[S.g[t]] = -(qr[t] + v);
        }
      }
    }
  } else if constexpr (PROG == V_MID || PROG == V_LAST_RHS || PROG == V_LAST_RK) {
    // viscous-flux derivatives (diffusion, tgvsolver.F90:700-787); columns
    // per direction: x (s11,s12,s13,q1), y (s12,s22,s23,q2), z (s13,s23,s33,q3)
    ChunkCoef<C::R, true> CD;
    load_chunk_coef<C::R, true>(CD, a.dac[dir], n + 3, p * C::R);
    const int c0 = dir == 0 ? 0 : (dir == 1 ? 1 : 2);
    const int c1 = dir == 0 ? 1 : (dir == 1 ? 3 : 4);
    const int c2 = dir == 0 ? 2 : (dir == 1 ? 4 : 5);
    double acc[4][NPT];
    double nx[NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++) nx[t] = a.sg[c0 * fs + S.g[t]];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
      for (int t = 0; t < NPT; t++)
        put<C>(T, S, t, nx[t]);
      // loads for this column's accumulation and the next column's input are
      // issued before the solve so their latency overlaps it
      double qr[NPT];
      double* dst = a.qrhs + (1 + c) * fs;
#pragma unroll
      for (int t = 0; t < NPT; t++) qr[t] = dst[S.g[t]];
      if (c < 3) {
        const int col = c == 0 ? c1 : (c == 1 ? c2 : 6 + dir);
        const double* src = a.sg + col * fs;
#pragma unroll
        for (int t = 0; t < NPT; t++) nx[t] = src[S.g[t]];
      }
      __syncthreads();
      solve_line<C::P, C::R, false>(Tl, Xl, n, p, CD, a.fc);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        {
          const double v = qr[t] + X[S.s[t]] * idn;
          if (PROG == V_LAST_RK) acc[c][t] = v;
          else dst[S.g[t]] = v;
        }
      }
    }
    if constexpr (PROG == V_LAST_RK) {
      // RK3 stage update (tgvsolver.F90:1043-1059), then the filter along dir
      double qn[5][NPT];
#pragma unroll
      for (int t = 0; t < NPT; t++) {
#pragma unroll
        for (int m = 0; m < 5; m++) qn[m][t] = 0.0;
        const int o = S.g[t];
#pragma unroll
        for (int m = 0; m < 5; m++) {
          const double qr = m == 0 ? a.qrhs[o] : acc[m - 1][t];
          const double qo = a.q[m * fs + o];
          double qs;
          if (a.stage == 0) {
            a.qsave[m * fs + o] = qo;
            qs = qo;
          } else {
            qs = a.qsave[m * fs + o];
          }
          qn[m][t] = rk_update(a, qs, qo, qr);
        }
      }
      ChunkCoef<C::R, false> CF;
      load_chunk_coef<C::R, false>(CF, a.fac[dir], n + 7, p * C::R);
#pragma unroll
      for (int m = 0; m < 5; m++) {
#pragma unroll
        for (int t = 0; t < NPT; t++)
          put<C>(T, S, t, qn[m][t]);
        __syncthreads();
        solve_line<C::P, C::R, true>(Tl, Xl, n, p, CF, a.fc);
        __syncthreads();
        double* dst = a.q + m * fs;
#pragma unroll
        for (int t = 0; t < NPT; t++)
          dst[S.g[t]] = X[S.s[t]];
      }
    }
  } else {  // F_ONLY: filter of q along dir, in place (filterq, tgvsolver.F90:791-883)
    ChunkCoef<C::R, false> CF;
    load_chunk_coef<C::R, false>(CF, a.fac[dir], n + 7, p * C::R);
    double nx[NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++) nx[t] = a.q[S.g[t]];
#pragma unroll
    for (int m = 0; m < 5; m++) {
#pragma unroll
      for (int t = 0; t < NPT; t++)
        put<C>(T, S, t, nx[t]);
      if (m < 4) {
#pragma unroll
        for (int t = 0; t < NPT; t++) nx[t] = a.q[(m + 1) * fs + S.g[t]];
      }
      __syncthreads();
      solve_line<C::P, C::R, true>(Tl, Xl, n, p, CF, a.fc);
      __syncthreads();
      double* dst = a.q + m * fs;
#pragma unroll
      for (int t = 0; t < NPT; t++)
        dst[S.g[t]] = X[S.s[t]];
    }
  }
}

// sum the per-tile stats in a fixed order (one block of 256 threads)
__global__ __launch_bounds__(256) void k_stats_reduce(const double* __restrict__ part, int nt,
                                                      double* __restrict__ out, double scale) {
  __shared__ double sa[256], sb[256];
  double x = 0.0, y = 0.0;
  for (int i = threadIdx.x; i < nt; i += 256) {
    x += part[2 * i];
    y += part[2 * i + 1];
  }
  sa[threadIdx.x] = x;
  sb[threadIdx.x] = y;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sa[threadIdx.x] += sa[threadIdx.x + h];
      sb[threadIdx.x] += sb[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * sa[0] * scale;
    out[1] = 0.5 * sb[0] * scale;
  }
}

// ---------------------------------------------------------------------------
// initial field and layout conversion kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
    const uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

// Philox4x32-10 (Salmon et al. 2011) keyed by seed, counter (i,j,k,c), then
// Box-Muller: a standard normal.
__device__ double gauss_of(uint64_t seed, uint32_t i, uint32_t j, uint32_t k, uint32_t c) {
  uint32_t x[4] = {i, j, k, c};
  philox10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double two53 = 9007199254740992.0;
  const double u1 = ((double)(((uint64_t)(x[0] >> 5) << 26) | (uint64_t)(x[1] >> 6)) + 0.5) / two53;
  const double u2 = ((double)(((uint64_t)(x[2] >> 5) << 26) | (uint64_t)(x[3] >> 6)) + 0.5) / two53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// TGV initial field (tgvsolver.F90:172-234) with optional seeded velocity noise
__global__ void k_init_tgv(FlowArgs a, double amp, uint64_t seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > a.n[0]) return;
  const int im = a.n[0], jm = a.n[1], km = a.n[2];
  const double pi = 3.141592653589793238462643383279502884;
  const double x1 = 2.0 * pi / (double)im * (double)i;
  const double x2 = 2.0 * pi / (double)jm * (double)j;
  const double x3 = km > 0 ? 2.0 * pi / (double)km * (double)k : 0.0;
  double u = sin(x1) * cos(x2) * cos(x3);
  double v = -cos(x1) * sin(x2) * cos(x3);
  double w = 0.0;
  if (amp != 0.0) {
    const uint32_t ii = i == im ? 0 : i, jj = j == jm ? 0 : j, kk = (km > 0 && k == km) ? 0 : k;
    u += amp * gauss_of(seed, ii, jj, kk, 0);
    v += amp * gauss_of(seed, ii, jj, kk, 1);
    if (km > 0) w += amp * gauss_of(seed, ii, jj, kk, 2);
  }
  const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
  {
#pragma clang fp contract(off)
    const double rho = 1.0;
    const double pinf = 1.0 / a.const2;
    const double p = pinf + 1.0 / 16.0 * (cos(2.0 * x1) + cos(2.0 * x2)) * (cos(2.0 * x3) + 2.0);
    const double var1 = 0.5 * (u * u + v * v + w * w);
    a.q[o] = rho;
    a.q[a.fs + o] = rho * u;
    a.q[2 * a.fs + o] = rho * v;
    a.q[3 * a.fs + o] = rho * w;
    a.q[4 * a.fs + o] = p * a.const6 + rho * var1;
  }
}

// periodic index of a host-layout halo cell: q(-h)=q(n-h), q(n+h)=q(h)
// (bchomovec, tgvsolver.F90:371-440); a direction with n = 0 (2-D) has one plane
__device__ __forceinline__ int wrap_index(int i, int n) {
  if (n == 0) return 0;
  return i < 0 ? i + n : (i > n ? i - n : i);
}

// dense Fortran-layout host buffer f(-hm:im+hm,-hm:jm+hm,-hm:km+hm,nf) <->
// padded device fields.  Upload reads the computed lattice 0..n only; download
// writes every cell, halo cells holding their periodic copies.  The grid spans
// the dense box (i fastest).
__global__ void k_pack(const double* __restrict__ src, double* __restrict__ dst, FlowArgs a,
                       int nf, int hm, int to_device) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  const long long di = a.n[0] + 1 + 2 * hm, dj = a.n[1] + 1 + 2 * hm;
  const long long dk = a.n[2] + 1 + 2 * hm;
  if (ii >= di) return;
  const int i = ii - hm, j = jj - hm, k = kk - hm;
  const bool inside = i >= 0 && i <= a.n[0] && j >= 0 && j <= a.n[1] && k >= 0 && k <= a.n[2];
  if (to_device && !inside) return;
  const long long dfs = di * dj * dk;
  const long long od = (long long)kk * di * dj + (long long)jj * di + ii;
  const long long o = (long long)wrap_index(k, a.n[2]) * a.plane +
                      (long long)wrap_index(j, a.n[1]) * a.pitch + wrap_index(i, a.n[0]);
  for (int m = 0; m < nf; m++) {
    if (to_device) dst[m * a.fs + o] = src[m * dfs + od];
    else dst[m * dfs + od] = src[m * a.fs + o];
  }
}

// primitives rho,u,v,w,p,T (q2fvar) into 6 dense fields, halo cells included
__global__ void k_prims(FlowArgs a, double* __restrict__ out, int hm) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  const long long di = a.n[0] + 1 + 2 * hm, dj = a.n[1] + 1 + 2 * hm;
  const long long dk = a.n[2] + 1 + 2 * hm;
  if (ii >= di) return;
  const long long dfs = di * dj * dk;
  const long long od = (long long)kk * di * dj + (long long)jj * di + ii;
  const long long o = (long long)wrap_index(kk - hm, a.n[2]) * a.plane +
                      (long long)wrap_index(jj - hm, a.n[1]) * a.pitch + wrap_index(ii - hm, a.n[0]);
  const Prim P = prim_of(a, a.q[o], a.q[a.fs + o], a.q[2 * a.fs + o], a.q[3 * a.fs + o],
                         a.q[4 * a.fs + o]);
  out[od] = P.r;
  out[dfs + od] = P.u;
  out[2 * dfs + od] = P.v;
  out[3 * dfs + od] = P.w;
  out[4 * dfs + od] = P.p;
  out[5 * dfs + od] = P.t;
}

// ---------------------------------------------------------------------------
// standalone line batch: op 0 = df_compact, 1 = compact_filter, on lines with
// explicit halos f(-5:n+5) (the reference's per-line interface)
// ---------------------------------------------------------------------------
template <class C, bool FILT>
__global__ __launch_bounds__(C::NT) void k_lines(const double* __restrict__ f, double* __restrict__ out,
                                                 int n, long long nlines, const double* __restrict__ ac,
                                                 FilterCoef fc, double alfa, double bhalo) {
  extern __shared__ double smem[];
  double* T = smem;
  double* X = smem + C::TL * C::LT;
  const long long L0 = (long long)blockIdx.x * C::TL;
  zero_pad<C>(T, n);
  for (int e = threadIdx.x; e < C::TL * (n + 11); e += C::NT) {
    const int l = e / (n + 11), r = e - l * (n + 11);
    T[l * C::LT + kLineOff - 5 + r] = (L0 + l < nlines) ? f[(L0 + l) * (n + 11) + r] : 0.0;
  }
  const int l = threadIdx.x / C::P, p = threadIdx.x % C::P;
  ChunkCoef<C::R, !FILT> CC;
  load_chunk_coef<C::R, !FILT>(CC, ac, n + (FILT ? 7 : 3), p * C::R);
  __syncthreads();
  solve_line<C::P, C::R, FILT>(T + l * C::LT, X + l * C::LT, n, p, CC, fc);
  __syncthreads();
  for (int e = threadIdx.x; e < C::TL * (n + 1); e += C::NT) {
    const int l2 = e / (n + 1), r = e - l2 * (n + 1);
    if (L0 + l2 < nlines) out[(L0 + l2) * (n + 1) + r] = X[l2 * C::LT + r];
  }
}

// ---------------------------------------------------------------------------
// launch-shape table
// ---------------------------------------------------------------------------
using CfgA = LineCfg<256, 4, 5, 5>;    // P=64, n <= 313
using CfgB = LineCfg<256, 4, 9, 9>;    // P=64, n <= 569
using CfgC = LineCfg<256, 16, 5, 5>;   // P=16, n <= 73
using CfgD = LineCfg<256, 8, 9, 9>;    // P=32, n <= 281
using CfgE = LineCfg<256, 8, 5, 5>;    // P=32, n <= 153

struct ShapeInfo {
  int id, NT, TL, P, R, NPT;
};
const ShapeInfo kShapes[] = {
    {0, 256, 4, 64, 5, 5}, {1, 256, 4, 64, 9, 9}, {2, 256, 16, 16, 5, 5},
    {3, 256, 8, 32, 9, 9}, {4, 256, 8, 32, 5, 5},
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

bool shape_fits(const ShapeInfo& s, int n) {
  return s.P * s.R >= n + 7 && s.NT * s.NPT >= s.TL * (n + 1) && n >= 8;
}

// default: the fitting shape with the fewest idle slots+chunks
int pick_shape(int n) {
  int best = -1;
  double bestw = 1e30;
  for (int i = 0; i < kNumShapes; i++) {
    const ShapeInfo& s = kShapes[i];
    if (!shape_fits(s, n)) continue;
    const double waste = (double)(s.NT * s.NPT) / (s.TL * (n + 1)) +
                         (double)(s.P * s.R) / (n + 7) + 0.05 * s.R;
    if (waste < bestw) { bestw = waste; best = i; }
  }
  return best;
}

template <class C>
hipError_t launch_pass_c(int prog, const FlowArgs& a, int dir, hipStream_t st) {
  const int nt = tiles_of(a.n, dir, C::TL);
  const size_t lds = (size_t)lds_doubles<C>() * sizeof(double);
  dim3 g(nt), b(C::NT);
  switch (prog) {
    case GC_FIRST: hipLaunchKernelGGL((k_pass<C, GC_FIRST>), g, b, lds, st, a, dir); break;
    case GC_MID: hipLaunchKernelGGL((k_pass<C, GC_MID>), g, b, lds, st, a, dir); break;
    case GC_LAST: hipLaunchKernelGGL((k_pass<C, GC_LAST>), g, b, lds, st, a, dir); break;
    case V_MID: hipLaunchKernelGGL((k_pass<C, V_MID>), g, b, lds, st, a, dir); break;
    case V_LAST_RHS: hipLaunchKernelGGL((k_pass<C, V_LAST_RHS>), g, b, lds, st, a, dir); break;
    case V_LAST_RK: hipLaunchKernelGGL((k_pass<C, V_LAST_RK>), g, b, lds, st, a, dir); break;
    default: hipLaunchKernelGGL((k_pass<C, F_ONLY>), g, b, lds, st, a, dir); break;
  }
  return hipGetLastError();
}

hipError_t launch_pass(int shape, int prog, const FlowArgs& a, int dir, hipStream_t st) {
  switch (shape) {
    case 0: return launch_pass_c<CfgA>(prog, a, dir, st);
    case 1: return launch_pass_c<CfgB>(prog, a, dir, st);
    case 2: return launch_pass_c<CfgC>(prog, a, dir, st);
    case 3: return launch_pass_c<CfgD>(prog, a, dir, st);
    default: return launch_pass_c<CfgE>(prog, a, dir, st);
  }
}

int shape_tl(int shape) { return kShapes[shape].TL; }

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct astr_hip_ctx {
  astr_hip_config cfg;
  FlowArgs a;
  int shape[3];
  int ntiles[3];
  double* base = nullptr;   // one allocation for all fields
  double* coef = nullptr;   // LU tables
  double* stats_part = nullptr;
  double* stats_out = nullptr;  // [2*max steps] device
  int stats_cap = 0;
  hipStream_t stream = nullptr;
  int own_stream = 0;
  int device = 0;
  double ke_scale;
  // optional per-launch event timing (astr_hip_profile_begin/end)
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
  int ev_used = 0;
  bool prof = false;
};

namespace {

// LU tables (commfunc.F90:752-774) of the ntype=3 systems
void thomas_prep(const std::vector<double>& av, const std::vector<double>& cv,
                 std::vector<double>& out) {
  const int m = (int)av.size();
  out.assign(3 * m, 0.0);
  double* a1 = out.data();
  double* a2 = a1 + m;
  double* a3 = a2 + m;
  a1[0] = cv[0];
  for (int i = 1; i < m; i++) {
    const double den = 1.0 - av[i] * a1[i - 1];
    a1[i] = cv[i] / den;
    a2[i] = 1.0 / den;
    a3[i] = av[i] / den;
  }
}

// derivative.F90:63-158 with scheme 643c, ntype=3
void deriv_tables(int n, std::vector<double>& out) {
  const int m = n + 3;
  std::vector<double> a(m, 1.0 / 3.0), c(m, 1.0 / 3.0);
  a[0] = c[0] = 0.0;
  a[m - 1] = c[m - 1] = 0.0;
  thomas_prep(a, c, out);
}

// filter.F90:31-100, ntype=3
void filter_tables(int n, double alfa, double bh, std::vector<double>& out) {
  const int m = n + 7;
  std::vector<double> a(m, alfa), c(m, alfa);
  a[0] = c[0] = bh;
  a[m - 1] = c[m - 1] = bh;
  thomas_prep(a, c, out);
}

// filter_coefficient_cal (filter.F90:299-432): the coefficients ntype=3 uses;
// c10 into fc, the 33 halo-row coefficients coefh(0:2,0:10) into h33
void filter_coefs(double alfa, double bh, FilterCoef& fc, double* h33) {
  double* c = fc.c10;
  c[0] = (193.0 + 126.0 * alfa) / 512.0;
  c[1] = (105.0 + 302.0 * alfa) / 512.0;
  c[2] = (-15.0 + 30.0 * alfa) / 128.0;
  c[3] = (45.0 - 90.0 * alfa) / 1024.0;
  c[4] = (-5.0 + 10.0 * alfa) / 512.0;
  c[5] = (1.0 - 2.0 * alfa) / 1024.0;
  const double h[33] = {
      (-1.0 + 1.0 * bh) / 1024.0, (5.0 - 5.0 * bh) / 512.0, (979.0 + 45.0 * bh) / 1024.0,
      (15.0 + 113.0 * bh) / 128.0, (-105.0 + 105.0 * bh) / 512.0, (63.0 - 63.0 * bh) / 256.0,
      (-105.0 + 105.0 * bh) / 512.0, (15.0 - 15.0 * bh) / 128.0, (-45.0 + 45.0 * bh) / 1024.0,
      (5.0 - 5.0 * bh) / 512.0, (-1.0 + 1.0 * bh) / 1024.0,
      (1.0 - 2.0 * alfa) / 1024.0, (-5.0 + 10.0 * alfa) / 512.0, (45.0 + 934.0 * alfa) / 1024.0,
      (113.0 + 30.0 * alfa) / 128.0, (105.0 + 302.0 * alfa) / 512.0, (-63.0 + 126.0 * alfa) / 256.0,
      (105.0 - 210.0 * alfa) / 512.0, (-15.0 + 30.0 * alfa) / 128.0, (45.0 - 90.0 * alfa) / 1024.0,
      (-5.0 + 10.0 * alfa) / 512.0, (1.0 - 2.0 * alfa) / 1024.0,
      (-1.0 + 2.0 * alfa) / 1024.0, (5.0 - 10.0 * alfa) / 512.0, (-45.0 + 90.0 * alfa) / 1024.0,
      (15.0 + 98.0 * alfa) / 128.0, (407.0 + 210.0 * alfa) / 512.0, (63.0 + 130.0 * alfa) / 256.0,
      (-105.0 + 210.0 * alfa) / 512.0, (15.0 - 30.0 * alfa) / 128.0, (-45.0 + 90.0 * alfa) / 1024.0,
      (5.0 - 10.0 * alfa) / 512.0, (-1.0 + 2.0 * alfa) / 1024.0};
  std::memcpy(h33, h, sizeof(h));
  fc.ch = nullptr;
}

int check_ctx(astr_hip_ctx* c) {
  if (!c) return fail("null context");
  HIP_OK(hipSetDevice(c->device));
  return 0;
}

dim3 pt_grid(const FlowArgs& a) { return dim3((a.n[0] + 64) / 64, a.n[1] + 1, a.n[2] + 1); }
// the dense host-layout box with halo hm
dim3 dense_grid(const FlowArgs& a, int hm) {
  return dim3((a.n[0] + 2 * hm + 64) / 64, a.n[1] + 1 + 2 * hm, a.n[2] + 1 + 2 * hm);
}

// launch one pass; with profiling on, bracket it with events on the stream
int pass(astr_hip_ctx* c, int prog, int d) {
  const bool rec = c->prof && c->ev_used < (int)c->ev_kind.size();
  if (rec) HIP_OK(hipEventRecord(c->ev[2 * c->ev_used], c->stream));
  HIP_OK(launch_pass(c->shape[d], prog, c->a, d, c->stream));
  if (rec) {
    HIP_OK(hipEventRecord(c->ev[2 * c->ev_used + 1], c->stream));
    c->ev_kind[c->ev_used++] = prog * 3 + d;
  }
  return 0;
}

int run_rhs(astr_hip_ctx* c, bool rk, bool stats, bool store_last_grad) {
  FlowArgs& a = c->a;
  const int nd = a.ndir;
  a.dostats = stats ? 1 : 0;
  a.store_last_grad = store_last_grad ? 1 : 0;
  for (int d = 0; d < nd; d++) {
    const int prog = d == nd - 1 ? GC_LAST : (d == 0 ? GC_FIRST : GC_MID);
    if (pass(c, prog, d)) return -1;
  }
  a.dostats = 0;
  for (int d = 0; d < nd; d++) {
    const int prog = d == nd - 1 ? (rk ? V_LAST_RK : V_LAST_RHS) : V_MID;
    if (pass(c, prog, d)) return -1;
  }
  return 0;
}

}  // namespace

extern "C" {

const char* astr_hip_last_error(void) { return g_err.c_str(); }

int astr_hip_create(const astr_hip_config* cfg, astr_hip_ctx** out) {
  if (!cfg || !out) return fail("null argument");
  *out = nullptr;
  if (cfg->im < 11 || cfg->jm < 11 || cfg->km < 0 || (cfg->km > 0 && cfg->km < 11))
    return fail("grid sizes must be >= 11 (km == 0 for two-dimensional cases)");
  if (cfg->semantics != ASTR_SEM_MINIAPP) return fail("unsupported semantics mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device");
  astr_hip_ctx* c = new astr_hip_ctx();
  c->cfg = *cfg;
  (void)hipGetDevice(&c->device);
  FlowArgs& a = c->a;
  std::memset(&a, 0, sizeof(a));
  a.n[0] = cfg->im; a.n[1] = cfg->jm; a.n[2] = cfg->km;
  a.ndir = cfg->km > 0 ? 3 : 2;
  a.pitch = ((cfg->im + 1) + 15) / 16 * 16;
  a.plane = (long long)a.pitch * (cfg->jm + 1);
  a.trash = (int)(a.plane * (cfg->km + 1));
  a.fs = (a.plane * (cfg->km + 1) + 1 + 63) / 64 * 64;
  const double g = cfg->gamma, ma = cfg->mach;
  a.const2 = g * ma * ma;
  a.const5 = (g - 1.0) * ma * ma;
  a.const6 = 1.0 / (g - 1.0);
  a.reynolds = cfg->reynolds;
  a.prandtl = cfg->prandtl;
  a.suth = cfg->sutherland / cfg->ref_tem;
  a.suth1 = 1.0 + a.suth;
  a.inv_d[0] = (double)cfg->im / cfg->lx;
  a.inv_d[1] = (double)cfg->jm / cfg->ly;
  a.inv_d[2] = cfg->km > 0 ? (double)cfg->km / cfg->lz : 0.0;
  a.dt = cfg->deltat;
  a.alfa = cfg->alfa_filter;
  a.bhalo = cfg->beta_halo;
  double h33[33];
  filter_coefs(cfg->alfa_filter, cfg->beta_halo, a.fc, h33);
  c->ke_scale = 1.0 / ((double)cfg->im * cfg->jm * (cfg->km > 0 ? cfg->km : 1));
  for (int d = 0; d < a.ndir; d++) {
    c->shape[d] = pick_shape(a.n[d]);
    if (c->shape[d] < 0) { delete c; return fail("no launch shape fits this line length"); }
    c->ntiles[d] = tiles_of(a.n, d, shape_tl(c->shape[d]));
  }
  // fields: q 5, qsave 5, qrhs 5, dv 12, sg 9
  const long long nfield = 36;
  size_t bytes = (size_t)(nfield * a.fs) * sizeof(double);
  if (hipMalloc(&c->base, bytes) != hipSuccess) { delete c; return fail("hipMalloc of fields failed"); }
  (void)hipMemset(c->base, 0, bytes);
  a.q = c->base;
  a.qsave = a.q + 5 * a.fs;
  a.qrhs = a.qsave + 5 * a.fs;
  a.dv = a.qrhs + 5 * a.fs;
  a.sg = a.dv + 12 * a.fs;
  // halo-row filter coefficients, then the LU tables
  std::vector<double> all(h33, h33 + 33);
  size_t off[6];
  for (int d = 0; d < a.ndir; d++) {
    std::vector<double> t;
    deriv_tables(a.n[d], t);
    off[d] = all.size();
    all.insert(all.end(), t.begin(), t.end());
    filter_tables(a.n[d], cfg->alfa_filter, cfg->beta_halo, t);
    off[3 + d] = all.size();
    all.insert(all.end(), t.begin(), t.end());
  }
  if (hipMalloc(&c->coef, all.size() * sizeof(double)) != hipSuccess) {
    (void)hipFree(c->base); delete c; return fail("hipMalloc of tables failed");
  }
  (void)hipMemcpy(c->coef, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice);
  a.fc.ch = c->coef;
  for (int d = 0; d < a.ndir; d++) {
    a.dac[d] = c->coef + off[d];
    a.fac[d] = c->coef + off[3 + d];
  }
  const int ntl = c->ntiles[a.ndir - 1];
  if (hipMalloc(&c->stats_part, 2 * (size_t)ntl * sizeof(double)) != hipSuccess) {
    (void)hipFree(c->base); (void)hipFree(c->coef); delete c; return fail("hipMalloc of stats failed");
  }
  a.stats = c->stats_part;
  c->stream = nullptr;
  *out = c;
  return 0;
}

int astr_hip_destroy(astr_hip_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->base);
  (void)hipFree(c->coef);
  (void)hipFree(c->stats_part);
  if (c->stats_out) (void)hipFree(c->stats_out);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int astr_hip_set_stream(astr_hip_ctx* c, void* stream) {
  if (check_ctx(c)) return -1;
  c->stream = (hipStream_t)stream;
  return 0;
}

int astr_hip_set_deltat(astr_hip_ctx* c, double dt) {
  if (check_ctx(c)) return -1;
  c->a.dt = dt;
  c->cfg.deltat = dt;
  return 0;
}

int astr_hip_init_tgv(astr_hip_ctx* c, double noise_amp, unsigned long long seed) {
  if (check_ctx(c)) return -1;
  hipLaunchKernelGGL(k_init_tgv, pt_grid(c->a), dim3(64), 0, c->stream, c->a, noise_amp,
                     (uint64_t)seed);
  HIP_OK(hipGetLastError());
  return 0;
}

// halo depth of a host buffer: 0..8 and at most the shortest direction, so a
// halo cell wraps once
static bool bad_halo(const FlowArgs& a, int hm) {
  if (hm < 0 || hm > 8) return true;
  for (int d = 0; d < a.ndir; d++)
    if (hm > a.n[d]) return true;
  return false;
}

static int xfer(astr_hip_ctx* c, double* host, int nf, double* dev, int hm, int to_dev) {
  const FlowArgs& a = c->a;
  const size_t dn = (size_t)(a.n[0] + 1 + 2 * hm) * (a.n[1] + 1 + 2 * hm) * (a.n[2] + 1 + 2 * hm);
  double* tmp = nullptr;
  HIP_OK(hipMalloc(&tmp, dn * nf * sizeof(double)));
  if (to_dev) HIP_OK(hipMemcpyAsync(tmp, host, dn * nf * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_pack, dense_grid(a, hm), dim3(64), 0, c->stream, to_dev ? tmp : dev,
                     to_dev ? dev : tmp, a, nf, hm, to_dev);
  HIP_OK(hipGetLastError());
  if (!to_dev) HIP_OK(hipMemcpyAsync(host, tmp, dn * nf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipFree(tmp));
  return 0;
}

int astr_hip_upload(astr_hip_ctx* c, const double* q, int hm) {
  if (check_ctx(c)) return -1;
  if (!q || bad_halo(c->a, hm)) return fail("bad argument to astr_hip_upload");
  return xfer(c, const_cast<double*>(q), 5, c->a.q, hm, 1);
}

int astr_hip_download(astr_hip_ctx* c, double* q, double* prims, int hm) {
  if (check_ctx(c)) return -1;
  if (bad_halo(c->a, hm)) return fail("bad argument to astr_hip_download");
  if (q && xfer(c, q, 5, c->a.q, hm, 0)) return -1;
  if (prims) {
    const FlowArgs& a = c->a;
    const size_t dn = (size_t)(a.n[0] + 1 + 2 * hm) * (a.n[1] + 1 + 2 * hm) * (a.n[2] + 1 + 2 * hm);
    double* tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, dn * 6 * sizeof(double)));
    hipLaunchKernelGGL(k_prims, dense_grid(a, hm), dim3(64), 0, c->stream, a, tmp, hm);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(prims, tmp, dn * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipFree(tmp));
  }
  return 0;
}

int astr_hip_filterq(astr_hip_ctx* c) {
  if (check_ctx(c)) return -1;
  for (int d = 0; d < c->a.ndir; d++)
    if (pass(c, F_ONLY, d)) return -1;
  return 0;
}

int astr_hip_gradcal(astr_hip_ctx* c, double* dvel, double* dtmp) {
  if (check_ctx(c)) return -1;
  if (run_rhs(c, false, false, true)) return -1;
  if (dvel && xfer(c, dvel, 9, c->a.dv, 0, 0)) return -1;
  if (dtmp && xfer(c, dtmp, 3, c->a.dv + 9 * c->a.fs, 0, 0)) return -1;
  return 0;
}

int astr_hip_rhscal(astr_hip_ctx* c, double* qrhs) {
  if (check_ctx(c)) return -1;
  if (run_rhs(c, false, false, true)) return -1;
  if (qrhs && xfer(c, qrhs, 5, c->a.qrhs, 0, 0)) return -1;
  return 0;
}

int astr_hip_rk3_steps(astr_hip_ctx* c, int nsteps, double* ke, double* ens) {
  if (check_ctx(c)) return -1;
  if (nsteps < 0) return fail("negative step count");
  static const double rk[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  const bool want = ke || ens;
  if (want && c->stats_cap < nsteps) {
    if (c->stats_out) (void)hipFree(c->stats_out);
    c->stats_out = nullptr;
    HIP_OK(hipMalloc(&c->stats_out, 2 * (size_t)nsteps * sizeof(double)));
    c->stats_cap = nsteps;
  }
  FlowArgs& a = c->a;
  const int ntl = c->ntiles[a.ndir - 1];
  for (int it = 0; it < nsteps; it++) {
    for (int s = 0; s < 3; s++) {
      a.stage = s;
      a.rk_a = rk[s][0]; a.rk_b = rk[s][1]; a.rk_c = rk[s][2];
      if (run_rhs(c, true, want && s == 0, false)) return -1;
      if (want && s == 0) {
        hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, c->stream, c->stats_part, ntl,
                           c->stats_out + 2 * it, c->ke_scale);
        HIP_OK(hipGetLastError());
      }
      const int nd = a.ndir;
      for (int d = 0; d < nd - 1; d++)
        if (pass(c, F_ONLY, d)) return -1;
    }
  }
  if (want) {
    std::vector<double> h(2 * (size_t)nsteps);
    HIP_OK(hipMemcpyAsync(h.data(), c->stats_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int it = 0; it < nsteps; it++) {
      if (ke) ke[it] = h[2 * it];
      if (ens) ens[it] = h[2 * it + 1];
    }
  }
  return 0;
}

int astr_hip_sync(astr_hip_ctx* c) {
  if (check_ctx(c)) return -1;
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int astr_hip_field_info(astr_hip_ctx* c, astr_hip_layout* lay) {
  if (check_ctx(c) || !lay) return fail("bad argument");
  lay->pitch = c->a.pitch;
  lay->plane = c->a.plane;
  lay->field_stride = c->a.fs;
  lay->q = c->a.q;
  lay->qrhs = c->a.qrhs;
  return 0;
}

int astr_hip_profile_begin(astr_hip_ctx* c, int max_launches) {
  if (check_ctx(c)) return -1;
  if (max_launches < 1) return fail("max_launches must be positive");
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  c->ev.assign(2 * (size_t)max_launches, nullptr);
  for (auto& e : c->ev) HIP_OK(hipEventCreate(&e));
  c->ev_kind.assign(max_launches, -1);
  c->ev_used = 0;
  c->prof = true;
  return 0;
}

int astr_hip_profile_end(astr_hip_ctx* c, double* ms, int* count) {
  if (check_ctx(c)) return -1;
  c->prof = false;
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < ASTR_PROF_SLOTS; k++) {
    if (ms) ms[k] = 0.0;
    if (count) count[k] = 0;
  }
  for (int i = 0; i < c->ev_used; i++) {
    float t = 0.f;
    HIP_OK(hipEventElapsedTime(&t, c->ev[2 * i], c->ev[2 * i + 1]));
    const int k = c->ev_kind[i];
    if (ms) ms[k] += t;
    if (count) count[k] += 1;
  }
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  c->ev.clear();
  c->ev_kind.clear();
  c->ev_used = 0;
  return 0;
}

int astr_hip_line_batch(int op, int n, long long nlines, const double* f, double* out,
                        double alfa_filter, double beta_halo) {
  if (n < 8 || nlines < 0 || (op != 0 && op != 1)) return fail("bad argument to astr_hip_line_batch");
  if (nlines == 0) return 0;
  if (!f || !out) return fail("null buffer");
  const int sh = pick_shape(n);
  if (sh < 0) return fail("no launch shape fits this line length");
  FilterCoef fc;
  double h33[33];
  filter_coefs(alfa_filter, beta_halo, fc, h33);
  std::vector<double> tab(h33, h33 + 33), t2;
  if (op == 0) deriv_tables(n, t2);
  else filter_tables(n, alfa_filter, beta_halo, t2);
  tab.insert(tab.end(), t2.begin(), t2.end());
  double *dt = nullptr, *df = nullptr, *dout = nullptr;
  HIP_OK(hipMalloc(&dt, tab.size() * sizeof(double)));
  HIP_OK(hipMalloc(&df, (size_t)nlines * (n + 11) * sizeof(double)));
  HIP_OK(hipMalloc(&dout, (size_t)nlines * (n + 1) * sizeof(double)));
  HIP_OK(hipMemcpy(dt, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(df, f, (size_t)nlines * (n + 11) * sizeof(double), hipMemcpyHostToDevice));
  const ShapeInfo& s = kShapes[sh];
  const long long nb = (nlines + s.TL - 1) / s.TL;
  const size_t lds = (size_t)2 * s.TL * ((s.P * s.R + 11) | 1) * sizeof(double);
  fc.ch = dt;
  const double* lu = dt + 33;
#define LB(Cf)                                                                            \
  if (op == 0) hipLaunchKernelGGL((k_lines<Cf, false>), dim3(nb), dim3(Cf::NT), lds, 0, df, dout, n, nlines, lu, fc, alfa_filter, beta_halo); \
  else hipLaunchKernelGGL((k_lines<Cf, true>), dim3(nb), dim3(Cf::NT), lds, 0, df, dout, n, nlines, lu, fc, alfa_filter, beta_halo);
  switch (sh) {
    case 0: LB(CfgA); break;
    case 1: LB(CfgB); break;
    case 2: LB(CfgC); break;
    case 3: LB(CfgD); break;
    default: LB(CfgE); break;
  }
#undef LB
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dout, (size_t)nlines * (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(dt);
  (void)hipFree(df);
  (void)hipFree(dout);
  return 0;
}

}  // extern "C"
