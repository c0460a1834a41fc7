// flow_dev.hpp -- device side of the per-stage line passes of flow.hip: the
// case description FlowArgs, the pointwise thermodynamics and fluxes, and the
// pass kernel k_pass templated on launch shape, pass program and register
// budget.  passes.hip instantiates it once per launch shape.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lines.hpp"

namespace astr {

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
  int dostats;          // stage-1 KE/enstrophy partial sums (1) or none (0)
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
  // Main-path semantics (src/mainloop.F90:396-482; passes instantiated with
  // MAIN = true).  The metrics of a uniform Cartesian grid: inv_d[d] holds
  // dxi(d,d) and jac the Jacobian; dxi(d,n) = 0 for d != n.
  int main_sem;         // 1: main path, 0: miniapp
  double jac;
  double* pr;           // 6 fields: the primitives rho, u, v, w, p, T as the
                        // last updatefvar / qswap left them
  const double* hx;     // halo sources of the x-lines on j or k boundary planes
  const double* hy;     // halo sources of the y-lines on k boundary planes
  long long hxs, hys;   // field strides of hx, hy
  // Main path, physical boundaries and non-uniform grids: the boundary bits
  // of each direction (bit 0: physical boundary at 0, bit 1: at n; 0 for a
  // periodic direction) and, for a grid that is not uniform, the metrics per
  // point as 10 fields: J, then dxi(d,n) at 1 + d + 3n (nullptr: the uniform
  // grid's jac and inv_d above, dxi(d,n) = 0 for d != n)
  int bnd[3];
  const double* met;
  // main path: the convective term is the upwind-biased one (upwind.hip),
  // which the gradient passes then leave out
  int uwd;
  // main path: the RK pass also applies the next stage's filterq sweep along
  // its direction to the new q, after updatefvar took the primitives
  int rk_filter;
};

// Halo sources kept by qswap for the lines whose halo the reference fills
// before a later direction averages the source points (qswap's i step copies
// the x-halo of the lines on the j and k boundary planes before the j and k
// steps average those planes; its j step copies the y-halo of the lines on
// the k boundary planes before the k step).  Ten rows per line: rows 1..5 and
// n-5..n-1 of the line, whose copies are the halo rows n+1..n+5 and -5..-1.
constexpr int kHaloSrcRows = 10;
__host__ __device__ inline long long hx_lines(const int n[3], int ndir) {
  return 2ll * (n[2] + 1) + (ndir == 3 ? 2ll * (n[1] + 1) : 0);
}
__host__ __device__ inline long long hy_lines(const int n[3], int ndir) {
  return ndir == 3 ? 2ll * (n[0] + 1) : 0;
}
// index of the halo source (line, hrow) of an x-line (j,k) or y-line (i,k)
// in hx / hy, or -1 when the line takes its halo from the current q
__host__ __device__ inline int halo_src_index(const int n[3], int ndir, int dir, int i, int j, int k,
                                              int hrow) {
  if (dir == 0) {
    if (j == 0 || j == n[1]) return ((j == 0 ? 0 : 1) * (n[2] + 1) + k) * kHaloSrcRows + hrow;
    if (ndir == 3 && (k == 0 || k == n[2]))
      return (2 * (n[2] + 1) + (k == 0 ? 0 : 1) * (n[1] + 1) + j) * kHaloSrcRows + hrow;
  } else if (dir == 1 && ndir == 3) {
    if (k == 0 || k == n[2]) return ((k == 0 ? 0 : 1) * (n[0] + 1) + i) * kHaloSrcRows + hrow;
  }
  return -1;
}

// Pass programs.  GL_GRAD + GL_CONV are GC_LAST split in two kernels (the
// gradient/stress half and the convective half), each with fewer registers.
// FGC_FIRST is GC_FIRST preceded by the previous stage's filter sweep along
// the same lines (x): it filters the q it loads, writes it back and goes on
// with the filtered values (the sweeps along different indices commute, so
// the x sweep may come last).
enum Prog { GC_FIRST = 0, GC_MID, GC_LAST, V_MID, V_LAST_RHS, V_LAST_RK, F_ONLY, GL_GRAD, GL_CONV,
            FGC_FIRST };
constexpr int kNumProgs = 10;

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
  unsigned g[C::NPT];  // byte offset inside a field; the trash point for lines
                       // past the end of the lattice (partial last tile)
  int s[C::NPT];  // index of the point in a line buffer (l*LT + r)
  int h[C::NPT];  // line-buffer index of the slot's periodic halo copy
  int ex[C::NPT]; // main path: index of the halo copy's value in hx / hy
                  // (halo_src_index), or -1: the copy takes the point's q
  int interior;   // bit t: i,j,k >= 1 (stats domain)
};

// FP64 load / store at a byte offset from a field's base pointer.  The base is
// the same for every lane and the offset a 32-bit per-lane value, so the
// access uses the scalar-base + vector-offset address mode and one offset
// register per point serves every field.
// The base passes through readfirstlane so the compiler keeps it a scalar
// value and does not fold the field index into the per-lane address.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}
// The accesses are made through global (address space 1) pointers so the
// compiler emits global, not flat, instructions.
typedef __attribute__((address_space(1))) const char gchar_c;
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ double ldg(const double* base, unsigned off) {
  gchar_c* b = (gchar_c*)uniform_ptr(base);
  return *(__attribute__((address_space(1))) const double*)(b + off);
}
__device__ __forceinline__ void stg(double* base, unsigned off, double v) {
  gchar* b = (gchar*)uniform_ptr(base);
  *(__attribute__((address_space(1))) double*)(b + off) = v;
}

// LDS of a pass: the output lines X (TL x LT), then the input lines T
// (TL x LT), then a write-only dummy cell.  A slot holding no point (past the
// tile) or no halo copy points s/h at index TL*LT: its T writes land in the
// dummy cell (T + kLineOff + TL*LT) and its X reads return T's zero padding
// (X + TL*LT = row -8 of T's first line), so no slot needs a guard.
// After them sit the filter's halo-row coefficients (3 rows x 11), staged
// once per tile so that no filter solve waits on a global load.
// The table also holds the rows of a physical boundary (boundary_row in
// lines.hpp): coefh (33), coefb (21), coef6i (4), coef8i (5).
constexpr int kHaloCoefs = kBndCoefs;
constexpr int kHaloSlot = 64;
template <class C>
constexpr int lds_line_doubles() {
  return (2 * C::TL * C::LT > 2 * C::NT ? 2 * C::TL * C::LT : 2 * C::NT) + kLineOff + 8;
}
// After the halo slot: the LU factors ac2, ac1 of every chunk row of the
// pass's line systems, staged once per tile.  Each solve reads its chunk's
// factors back just before it runs instead of holding them in registers for
// the whole pass; the registers freed raise the waves per SIMD of the last
// gradient pass, the viscous passes and the filter pass.
template <class C>
constexpr int lds_coef_offset() {
  return lds_line_doubles<C>() + kHaloSlot;
}
// Regions of 2 P R doubles a pass stages: one (the derivative's) in the last
// gradient pass and the viscous passes, one (the filter's) in the filter
// pass and in the filtering first gradient pass, two in the RK pass, none in the first two gradient passes, which
// keep the factors in registers (there the freed registers would not reach
// another wave per SIMD, and the reloads only cost time).
constexpr int coef_regions(int prog) {
  return prog == V_LAST_RK ? 2 : (prog == GC_FIRST || prog == GC_MID ? 0 : 1);
}
template <class C>
constexpr int lds_doubles(int prog) {
  return lds_coef_offset<C>() + coef_regions(prog) * 2 * C::P * C::R;
}

// Stage the chunk factors of a line of M rows into coefficient region `region`
// (2 P R doubles each, after the halo slot): c2 (ac2, 1 at row 0) at [u],
// c1 (ac1, 0 at row M-1) at [P*R + u], 0 past the rows.
template <class C>
__device__ __forceinline__ double* stage_chunk_coefs(const double* __restrict__ ac, int M, double* smem,
                                                     int region) {
  double* cs = smem + lds_coef_offset<C>() + region * 2 * C::P * C::R;
  for (int u = threadIdx.x; u < C::P * C::R; u += C::NT) {
    const bool ok = u < M;
    const int uu = ok ? u : 0;
    double a1 = ac[uu], a2 = ac[M + uu];
    if (u == 0) a2 = 1.0;
    if (u == M - 1) a1 = 0.0;
    cs[u] = ok ? a2 : 0.0;
    cs[C::P * C::R + u] = ok ? a1 : 0.0;
  }
  return cs;
}

template <class C>
__device__ __forceinline__ void lds_chunk_coef(ChunkCoef<C::R, false>& CD, const double* cs, int u0) {
#pragma unroll
  for (int k = 0; k < C::R; k++) {
    CD.c2[k] = cs[u0 + k];
    CD.c1[k] = cs[C::P * C::R + u0 + k];
  }
}

// the filter's: also its last-row exception (load_chunk_coef in lines.hpp)
template <class C>
__device__ __forceinline__ void lds_chunk_coef_f(ChunkCoef<C::R, true>& CF, const double* cs, int u0,
                                                 int M, double c3last) {
#pragma unroll
  for (int k = 0; k < C::R; k++) {
    CF.c2[k] = cs[u0 + k];
    CF.c1[k] = cs[C::P * C::R + u0 + k];
  }
  const int kl = M - 1 - u0;
  CF.klast = kl >= 0 && kl < C::R ? kl : -1;
  CF.c3last = c3last;
}

// A line solve of a pass: the periodic form, or (main path, general-grid
// kernels only) the closed form of a direction with physical ends
template <class C, bool FILT, bool GEN>
__device__ __forceinline__ void solve_pass_line(const double* Tl, double* Xl, int n, int p,
                                                const ChunkCoef<C::R, FILT>& CC, const FilterCoef& fc,
                                                int bnd) {
  if constexpr (GEN) {
    if (bnd) {
      solve_line<C::P, C::R, FILT, true>(Tl, Xl, n, p, CC, fc, bnd);
      return;
    }
  }
  solve_line<C::P, C::R, FILT>(Tl, Xl, n, p, CC, fc);
}

// Copy the filter's halo-row coefficients into the tile's LDS; the barrier
// before the first filter solve publishes them.  Every read of the previous
// tile's copy ended before that tile's last barrier.
template <class C>
__device__ __forceinline__ FilterCoef stage_halo_coefs(const FilterCoef& fc, double* smem) {
  double* ch = smem + lds_line_doubles<C>();
  if (threadIdx.x < kHaloCoefs) ch[threadIdx.x] = fc.ch[threadIdx.x];
  FilterCoef f = fc;
  f.ch = ch;
  return f;
}

// Map the tile's points onto slots.  dir 0: lines are (j,k), rows i; dir 1:
// lines (i,k) rows j; dir 2: lines (i,j) rows k.
template <class C, bool MAIN = false>
__device__ __forceinline__ void make_slots(Slots<C>& S, const FlowArgs& a, int dir, int tile) {
  const int n = a.n[dir];
  const int np = C::TL * (n + 1);
  const int ni = (a.n[0] + C::TL) / C::TL;
  const int dummy = C::TL * C::LT;
  S.interior = 0;
#pragma unroll
  for (int t = 0; t < C::NPT; t++) {
    const int e = threadIdx.x + t * C::NT;
    S.g[t] = (unsigned)a.trash * 8u;
    S.s[t] = dummy;
    S.h[t] = dummy;
    if (MAIN) S.ex[t] = -1;
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
      if (MAIN && ok && S.h[t] != dummy)
        S.ex[t] = halo_src_index(a.n, a.ndir, dir, i, j, k, r <= 5 ? r - 1 : r - (n - 5) + 5);
      if (ok) {
        S.g[t] = (unsigned)((long long)k * a.plane + (long long)j * a.pitch + i) * 8u;
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

// what changes from stage to stage (the same for every case of a batch)
struct StageArgs {
  int stage;            // 0,1,2
  int dostats;          // stage-1 KE/enstrophy partial sums (1) or none (0)
  int store_last_grad;  // last GC pass also stores its gradients
  double rk_a, rk_b, rk_c;
};

__device__ __forceinline__ StageArgs stage_of(const FlowArgs& a) {
  return StageArgs{a.stage, a.dostats, a.store_last_grad, a.rk_a, a.rk_b, a.rk_c};
}

// q <- a*qsave + b*q + c*qrhs*dt in the reference's order (tgvsolver.F90:1055-1057)
__device__ __forceinline__ double rk_update(const StageArgs& sa, double dt, double qs, double qo,
                                            double qr) {
#pragma clang fp contract(off)
  return sa.rk_a * qs + sa.rk_b * qo + sa.rk_c * qr * dt;
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

// Main path (convrsdcal6, src/solver.F90:2196-2207): flux component m along
// dir, J (q_m U + dxi(dir,m) p) with the contravariant velocity
// U = sum_n dxi(dir,n) u_n, for the metrics of a uniform Cartesian grid
// (dxi(dir,n) = 0 for n != dir, so those products add exact zeros)
__device__ __forceinline__ double conv_flux_main(double jac, double dx, int m, int dir, double qm,
                                                 double q5, double ud, double p) {
#pragma clang fp contract(off)
  const double uu = dx * ud;
  if (m == 0) return jac * qm * uu;
  if (m == 4) return jac * (q5 + p) * uu;
  if (m == dir + 1) return jac * (qm * uu + dx * p);
  return jac * (qm * uu);
}

// Main path: the Jacobian J and dxi(dir,dir) of a point on a uniform grid
struct Met {
  double j, d;
};
__device__ __forceinline__ Met met_of(const FlowArgs& a, int dir) { return Met{a.jac, a.inv_d[dir]}; }

// Main path on a grid that is not uniform: J and the row dxi(dir, 0..2) of a point
struct MetRow {
  double j, r[3];
};
__device__ __forceinline__ MetRow met_row(const FlowArgs& a, int dir, unsigned g) {
  MetRow M;
  M.j = ldg(a.met, g);
#pragma unroll
  for (int nn = 0; nn < 3; nn++) M.r[nn] = ldg(a.met + (1 + dir + 3 * nn) * a.fs, g);
  return M;
}
// the Jacobian of a point, either grid (GEN: a general-grid kernel, which
// alone reads per-point metrics)
template <bool GEN>
__device__ __forceinline__ double jac_of(const FlowArgs& a, unsigned g) {
  return GEN && a.met ? ldg(a.met, g) : a.jac;
}

// convrsdcal6 (src/solver.F90:2196-2207) with the full metric row:
// U = dxi(d,1) u + dxi(d,2) v + dxi(d,3) w, F_m = J (q_m U + dxi(d,m) p)
__device__ __forceinline__ double conv_flux_full(const MetRow& M, int m, double qm, double q5,
                                                 double u, double v, double w, double p) {
#pragma clang fp contract(off)
  const double uu = M.r[0] * u + M.r[1] * v + M.r[2] * w;
  if (m == 0) return M.j * qm * uu;
  if (m == 4) return M.j * (q5 + p) * uu;
  return M.j * (qm * uu + M.r[m - 1] * p);
}

// diffrsdcal6 (src/solver.F90:2623-2635) with the full metric row: column
// c of direction dir, (s_c1 dxi(d,1) + s_c2 dxi(d,2) + s_c3 dxi(d,3)) J with
// s_c = the stress row c (c < 3) or the heat flux (c = 3)
__device__ __forceinline__ double visc_col_full(const FlowArgs& a, int dir, int c, unsigned g) {
#pragma clang fp contract(off)
  const MetRow M = met_row(a, dir, g);
  const int k0 = c == 0 ? 0 : (c == 1 ? 1 : (c == 2 ? 2 : 6));
  const int k1 = c == 0 ? 1 : (c == 1 ? 3 : (c == 2 ? 4 : 7));
  const int k2 = c == 0 ? 2 : (c == 1 ? 4 : (c == 2 ? 5 : 8));
  const double s0 = ldg(a.sg + k0 * a.fs, g), s1 = ldg(a.sg + k1 * a.fs, g), s2 = ldg(a.sg + k2 * a.fs, g);
  return (s0 * M.r[0] + s1 * M.r[1] + s2 * M.r[2]) * M.j;
}

// element m of a 5-vector by selects: in a loop over m that the compiler
// keeps rolled, indexing a register array by m would move the array to
// scratch memory
__device__ __forceinline__ double pick5(double a0, double a1, double a2, double a3, double a4, int m) {
  return m == 0 ? a0 : (m == 1 ? a1 : (m == 2 ? a2 : (m == 3 ? a3 : a4)));
}
#define ASTR_QV(m, t) pick5(qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t], (m))
#define ASTR_QH(m) pick5(qh[0], qh[1], qh[2], qh[3], qh[4], (m))

// Main path: q of the halo copy of slot t.  qswap fills the halo of a line
// from the filtered q; the lines on boundary planes that a later qswap step
// averages take the values kept before that step (hx / hy), the others the
// current q of the source point, which the slot holds.
template <class C>
__device__ __forceinline__ void halo_q(const FlowArgs& a, const Slots<C>& S, int t, int dir,
                                       const double (&qv)[5][C::NPT], double (&qh)[5]) {
  if (S.ex[t] >= 0) {
    const double* src = dir == 0 ? a.hx : a.hy;
    const long long st = dir == 0 ? a.hxs : a.hys;
#pragma unroll
    for (int m = 0; m < 5; m++) qh[m] = src[m * st + S.ex[t]];
  } else {
#pragma unroll
    for (int m = 0; m < 5; m++) qh[m] = qv[m][t];
  }
}

// main path RK3 update (src/mainloop.F90:441-449):
// q <- (a qsave + b q J + c qrhs dt) / J, in the reference's order
__device__ __forceinline__ double rk_update_main(const StageArgs& sa, double dt, double jac, double qs,
                                                 double qo, double qr) {
#pragma clang fp contract(off)
  return (sa.rk_a * qs + sa.rk_b * qo * jac + sa.rk_c * qr * dt) / jac;
}

// ---------------------------------------------------------------------------
// the pass kernel
// ---------------------------------------------------------------------------
// MINB: workgroups per CU the register allocation is held to (each
// workgroup puts one wave on each SIMD, so MINB waves per SIMD)
// One tile of one pass.  `a` describes the case, `sa` the stage.
// MAIN: the main path's semantics (src/mainloop.F90:396-482): the gradients
// and fluxes take the stored primitives (pr) as the last updatefvar left them
// and their halo copies from the filtered q (qswap), the fluxes carry the
// metrics J and dxi (convrsdcal6, diffrsdcal6), and the RK pass ends with
// updatefvar instead of the filter.
template <class C, int PROG, bool MAIN = false, bool GEN = false>
__device__ __forceinline__ void pass_body(const FlowArgs& a, const StageArgs& sa, int dir, int tile,
                                          double* smem) {
  constexpr int NPT = C::NPT;
  constexpr int kDummy = C::TL * C::LT;
  const int n = a.n[dir];
  double* X = smem;                 // output lines, TL x LT (rows at offset 0)
  double* T = smem + C::TL * C::LT; // input lines, TL x LT
  Slots<C> S;
  make_slots<C, MAIN>(S, a, dir, tile);
  zero_pad<C>(T, n);
  const int ln = threadIdx.x / C::P, p = threadIdx.x % C::P;
  double* Tl = T + ln * C::LT;
  double* Xl = X + ln * C::LT;
  const double idn = a.inv_d[dir];
  const long long fs = a.fs;

  if constexpr (PROG == GC_FIRST || PROG == FGC_FIRST || PROG == GC_MID || PROG == GC_LAST ||
                PROG == GL_GRAD || PROG == GL_CONV) {
    constexpr bool FIRST = PROG == GC_FIRST || PROG == FGC_FIRST;
    constexpr bool LAST = PROG == GC_LAST || PROG == GL_GRAD || PROG == GL_CONV;
    constexpr bool DO_GRAD = PROG != GL_CONV;
    constexpr bool DO_CONV = PROG != GL_GRAD;
    // the last pass, which holds the most registers, reads the chunk factors
    // back from LDS before each solve; the others keep them in registers
    ChunkCoef<C::R, false> CDr;
    const double* cs = nullptr;
    if constexpr (LAST) cs = stage_chunk_coefs<C>(a.dac[dir], n + 3, smem, 0);
    else load_chunk_coef<C::R, false>(CDr, a.dac[dir], n + 3, p * C::R);
    double qv[5][NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++)
#pragma unroll
      for (int m = 0; m < 5; m++) qv[m][t] = ldg(a.q + m * fs, S.g[t]);
    if constexpr (PROG == FGC_FIRST) {
      // the previous stage's filter sweep along these lines (filterq,
      // tgvsolver.F90:805-825), as the filter pass does it, on the loaded q;
      // the filtered q is stored and the gradients and fluxes below take it
      const double* csf = stage_chunk_coefs<C>(a.fac[dir], n + 7, smem, 0);
      const double c3l = a.fac[dir][2 * (n + 7) + (n + 7) - 1];
      const FilterCoef fl = stage_halo_coefs<C>(a.fc, smem);
#pragma unroll
      for (int m = 0; m < 5; m++) {
#pragma unroll
        for (int t = 0; t < NPT; t++) put<C>(T, S, t, qv[m][t]);
        __syncthreads();
        {
          ChunkCoef<C::R, true> CF;
          lds_chunk_coef_f<C>(CF, csf, p * C::R, n + 7, c3l);
          solve_line<C::P, C::R, true>(Tl, Xl, n, p, CF, fl);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          qv[m][t] = X[S.s[t]];
          stg(a.q + m * fs, S.g[t], qv[m][t]);
        }
      }
    }

    // main path: the stored primitives the gradient solves take (pv) and
    // the convective part's u_dir and p are loaded with q, at the start, so
    // that every load of the pass is in flight at once (these passes run at
    // two waves per SIMD, whose register budget holds them).  The split last
    // pass's gradient half loads each just before its solve: the registers
    // would cost it its third wave per SIMD.
    constexpr bool PF = MAIN && DO_GRAD && PROG != GL_GRAD;
    double pv[4][NPT], ud[NPT], pp[NPT];
    if constexpr (PF) {
#pragma unroll
      for (int t = 0; t < NPT; t++)
#pragma unroll
        for (int c = 0; c < 4; c++) pv[c][t] = ldg(a.pr + (c < 3 ? 1 + c : 5) * fs, S.g[t]);
      if (DO_CONV && !a.uwd) {
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          pp[t] = ldg(a.pr + 4 * fs, S.g[t]);
          ud[t] = ldg(a.pr + (1 + dir) * fs, S.g[t]);
        }
      }
    }
    // main path, halo slots: the halo's pressure (hp) and u_dir (hu), kept
    // from the gradient part for the convective part (set on halo lanes only)
    double hp[NPT], hu[NPT];
    if constexpr (DO_GRAD) {
    // gradients of u, v, w, T along dir (gradcal, tgvsolver.F90:450-537)
    double gz[4][NPT];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        if constexpr (MAIN) {
          // the stored u, v, w or T (gradcal, src/comsolver.F90:280-300);
          // the halo copy from the primitives of the halo's q (qswap's q2fvar)
          T[kLineOff + S.s[t]] = PF ? pv[c][t] : ldg(a.pr + (c < 3 ? 1 + c : 5) * fs, S.g[t]);
          if (S.h[t] != kDummy) {
            double qh[5];
            halo_q<C>(a, S, t, dir, qv, qh);
            double f;
            if (c < 3) {
              f = qh[c + 1] / qh[0];
              if (c == dir) hu[t] = f;
            } else {
              const double pr = pressure_of(a, qh[0], qh[1], qh[2], qh[3], qh[4]);
              hp[t] = pr;
              f = pr / qh[0] * a.const2;
            }
            T[kLineOff + S.h[t]] = f;
          }
        } else {
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
      if constexpr (LAST) {
        ChunkCoef<C::R, false> CD;
        lds_chunk_coef<C>(CD, cs, p * C::R);
        solve_pass_line<C, false, GEN>(Tl, Xl, n, p, CD, a.fc, a.bnd[dir]);
      } else {
        solve_pass_line<C, false, GEN>(Tl, Xl, n, p, CDr, a.fc, a.bnd[dir]);
      }
      __syncthreads();
      const int fld = c < 3 ? c + 3 * dir : 9 + dir;
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        if (GEN && a.met) {
          // gradcal's accumulation dvel(m,n) += D_d(u_m) dxi(d,n) over every
          // n (src/comsolver.F90:300-321); the last pass adds its terms with
          // the stress below
          const double x = X[S.s[t]];
          gz[c][t] = x;
          if constexpr (!LAST) {
            const unsigned o = S.g[t];
#pragma unroll
            for (int nn = 0; nn < 3; nn++) {
#pragma clang fp contract(off)
              if (nn >= a.ndir) continue;
              const int f2 = c < 3 ? c + 3 * nn : 9 + nn;
              const double v = x * ldg(a.met + (1 + dir + 3 * nn) * fs, o);
              if (FIRST) stg(a.dv + f2 * fs, o, v);
              else stg(a.dv + f2 * fs, o, ldg(a.dv + f2 * fs, o) + v);
            }
          }
        } else {
          // dvel(m,n) = D_d(u_m) dxi(d,n): dxi(d,d) of the uniform grid
          const double v = X[S.s[t]] * (MAIN ? met_of(a, dir).d : idn);
          if (!LAST || sa.store_last_grad) stg(a.dv + fld * fs, S.g[t], v);
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
        const unsigned o = S.g[t];
        double g[3][3], gt[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if (GEN && a.met) {
#pragma clang fp contract(off)
            // the earlier passes' terms plus this direction's, D_dir(u_m) dxi(dir,d)
            if (d < a.ndir) {
              const double md = ldg(a.met + (1 + dir + 3 * d) * fs, o);
#pragma unroll
              for (int m = 0; m < 3; m++) g[m][d] = ldg(a.dv + (m + 3 * d) * fs, o) + gz[m][t] * md;
              gt[d] = ldg(a.dv + (9 + d) * fs, o) + gz[3][t] * md;
              if (sa.store_last_grad) {
#pragma unroll
                for (int m = 0; m < 3; m++) stg(a.dv + (m + 3 * d) * fs, o, g[m][d]);
                stg(a.dv + (9 + d) * fs, o, gt[d]);
              }
            } else {
              g[0][d] = 0.0; g[1][d] = 0.0; g[2][d] = 0.0; gt[d] = 0.0;
            }
          } else if (d == dir) {
            g[0][d] = gz[0][t]; g[1][d] = gz[1][t]; g[2][d] = gz[2][t]; gt[d] = gz[3][t];
          } else if (d < a.ndir) {
            g[0][d] = ldg(a.dv + (0 + 3 * d) * fs, o);
            g[1][d] = ldg(a.dv + (1 + 3 * d) * fs, o);
            g[2][d] = ldg(a.dv + (2 + 3 * d) * fs, o);
            gt[d] = ldg(a.dv + (9 + d) * fs, o);
          } else {
            g[0][d] = 0.0; g[1][d] = 0.0; g[2][d] = 0.0; gt[d] = 0.0;
          }
        }
        Prim P;
        if constexpr (MAIN) {
          // diffrsdcal6 (src/solver.F90:2433-2545) and the statistics
          // (statistic.F90:871-992) take the stored rho, vel, tmp
          P.r = sa.dostats ? ldg(a.pr, o) : 0.0;
          P.u = ldg(a.pr + fs, o);
          P.v = ldg(a.pr + 2 * fs, o);
          P.w = ldg(a.pr + 3 * fs, o);
          P.p = 0.0;
          P.t = ldg(a.pr + 5 * fs, o);
        } else {
          P = prim_of(a, qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t]);
        }
        double f[9];
        visc_flux(a, g, gt, P, f);
#pragma unroll
        for (int m = 0; m < 9; m++) stg(a.sg + m * fs, o, f[m]);
        if (sa.dostats && ((S.interior >> t) & 1)) {
          sk += P.r * (P.u * P.u + P.v * P.v + P.w * P.w);
          const double o1 = g[2][1] - g[1][2];
          const double o2 = g[0][2] - g[2][0];
          const double o3 = g[1][0] - g[0][1];
          se += P.r * (o1 * o1 + o2 * o2 + o3 * o3);
        }
      }
      if (sa.dostats) {
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
    }

    // the main path's upwind-biased convection replaces this part (uniform
    // over the workgroup, after the last barrier of the gradient part)
    if (MAIN && a.uwd) return;
    if constexpr (DO_CONV) {
    // convective flux derivatives along dir (convection, tgvsolver.F90:539-633)
#pragma unroll
    for (int t = 0; t < NPT; t++) {
      if constexpr (MAIN) {
        // convrsdcal6 takes the stored vel and prs (src/solver.F90:2196-2207),
        // loaded at the start where the pass has gradient solves
        if constexpr (!PF) {
          pp[t] = ldg(a.pr + 4 * fs, S.g[t]);
          ud[t] = ldg(a.pr + (1 + dir) * fs, S.g[t]);
        }
      } else {
        pp[t] = pressure_of(a, qv[0][t], qv[1][t], qv[2][t], qv[3][t], qv[4][t]);
        ud[t] = (dir == 0 ? qv[1][t] : (dir == 1 ? qv[2][t] : qv[3][t])) / qv[0][t];
      }
    }
    // the main path's flux inputs of component MM: one copy per component
    // behind a wave-uniform switch, so that q is indexed by constants while
    // the loop over components (and its solve) stays rolled
    auto stage_main = [&](auto mc) __attribute__((always_inline)) {
      constexpr int MM = decltype(mc)::value;
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        // a periodic halo copy carries the source point's metrics, which are this point's
        if (GEN && a.met) {
          const unsigned o = S.g[t];
          const MetRow M = met_row(a, dir, o);
          const double u = ldg(a.pr + fs, o), v = ldg(a.pr + 2 * fs, o), w = ldg(a.pr + 3 * fs, o);
          T[kLineOff + S.s[t]] = conv_flux_full(M, MM, qv[MM][t], qv[4][t], u, v, w, pp[t]);
          if (S.h[t] != kDummy) {
            double qh[5];
            halo_q<C>(a, S, t, dir, qv, qh);
            const double ph = pressure_of(a, qh[0], qh[1], qh[2], qh[3], qh[4]);
            T[kLineOff + S.h[t]] =
                conv_flux_full(M, MM, qh[MM], qh[4], qh[1] / qh[0], qh[2] / qh[0], qh[3] / qh[0], ph);
          }
        } else {
          const Met M = met_of(a, dir);
          T[kLineOff + S.s[t]] = conv_flux_main(M.j, M.d, MM, dir, qv[MM][t], qv[4][t], ud[t], pp[t]);
          if (S.h[t] != kDummy) {
            double qh[5];
            halo_q<C>(a, S, t, dir, qv, qh);
            double ph, uh;
            if constexpr (DO_GRAD) {
              ph = hp[t];
              uh = hu[t];
            } else {
              ph = pressure_of(a, qh[0], qh[1], qh[2], qh[3], qh[4]);
              uh = pick5(0.0, qh[1], qh[2], qh[3], 0.0, 1 + dir) / qh[0];
            }
            T[kLineOff + S.h[t]] = conv_flux_main(M.j, M.d, MM, dir, qh[MM], qh[4], uh, ph);
          }
        }
      }
    };
#pragma unroll
    for (int m = 0; m < 5; m++) {
      if constexpr (MAIN) {
        switch (m) {
          case 0: stage_main(std::integral_constant<int, 0>{}); break;
          case 1: stage_main(std::integral_constant<int, 1>{}); break;
          case 2: stage_main(std::integral_constant<int, 2>{}); break;
          case 3: stage_main(std::integral_constant<int, 3>{}); break;
          default: stage_main(std::integral_constant<int, 4>{}); break;
        }
      } else {
#pragma unroll
        for (int t = 0; t < NPT; t++) put<C>(T, S, t, conv_flux(m, dir, ASTR_QV(m, t), qv[4][t], ud[t], pp[t]));
      }
      double* dst = a.qrhs + m * fs;
      double qr[NPT];
      if (!FIRST) {
        // issued before the solve so the load latency overlaps it
#pragma unroll
        for (int t = 0; t < NPT; t++) qr[t] = ldg(dst, S.g[t]);
      }
      __syncthreads();
      if constexpr (LAST) {
        ChunkCoef<C::R, false> CD;
        lds_chunk_coef<C>(CD, cs, p * C::R);
        solve_pass_line<C, false, GEN>(Tl, Xl, n, p, CD, a.fc, a.bnd[dir]);
      } else {
        solve_pass_line<C, false, GEN>(Tl, Xl, n, p, CDr, a.fc, a.bnd[dir]);
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        {
          // the main path's fluxes carry the metrics: no 1/dx after the solve
          const double v = MAIN ? X[S.s[t]] : X[S.s[t]] * idn;
          if (FIRST) stg(dst, S.g[t], v);
          else if (PROG == GC_MID) stg(dst, S.g[t], qr[t] + v);
          else stg(dst, S.g[t], -(qr[t] + v));
        }
      }
    }
    }
  } else if constexpr (PROG == V_MID || PROG == V_LAST_RHS || PROG == V_LAST_RK) {
    // viscous-flux derivatives (diffusion, tgvsolver.F90:700-787); columns
    // per direction: x (s11,s12,s13,q1), y (s12,s22,s23,q2), z (s13,s23,s33,q3).
    // V_LAST_RK: each variable's RK3 stage update (tgvsolver.F90:1043-1059)
    // and its filter along dir follow its viscous sum at once, so one sum is
    // live at a time; density, which has no viscous term, goes first.
    constexpr bool RK = PROG == V_LAST_RK;
    // the chunk factors are read back from LDS before each solve, so they do
    // not hold registers across the pass
    const double* cs = stage_chunk_coefs<C>(a.dac[dir], n + 3, smem, 0);
    const int c0 = dir == 0 ? 0 : (dir == 1 ? 1 : 2);
    const int c1 = dir == 0 ? 1 : (dir == 1 ? 3 : 4);
    const int c2 = dir == 0 ? 2 : (dir == 1 ? 4 : 5);
    double nx[NPT];
#pragma unroll
    for (int t = 0; t < NPT; t++) nx[t] = ldg(a.sg + c0 * fs, S.g[t]);
    if constexpr (!RK || MAIN) {
      // MAIN && RK (time_integration_rk, src/mainloop.F90:430-482): each
      // variable's RK3 update follows its viscous sum, density (no viscous
      // term) first; updatefvar of the new q ends the pass
      constexpr bool RKM = RK && MAIN;
      double qn[5][NPT];
      if constexpr (RKM) {
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          const unsigned o = S.g[t];
          const double jb = jac_of<GEN>(a, o);
          const double q0 = ldg(a.q, o);
          double s0;
          if (sa.stage == 0) {
            s0 = q0 * jb;  // qsave = q J (:423-427)
            stg(a.qsave, o, s0);
          } else {
            s0 = ldg(a.qsave, o);
          }
          qn[0][t] = rk_update_main(sa, a.dt, jb, s0, q0, ldg(a.qrhs, o));
          stg(a.q, o, qn[0][t]);
        }
      }
#pragma unroll
      for (int c = 0; c < 4; c++) {
        // the main path's viscous fluxes carry the metrics:
        // (sigma . dxi(dir,:)) J (diffrsdcal6, src/solver.F90:2623-2635)
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          if constexpr (MAIN) {
            if (GEN && a.met) {
              put<C>(T, S, t, visc_col_full(a, dir, c, S.g[t]));
            } else {
              const Met M = met_of(a, dir);
              put<C>(T, S, t, nx[t] * M.d * M.j);
            }
          } else {
            put<C>(T, S, t, nx[t]);
          }
        }
        // loads for this column's accumulation and the next column's input are
        // issued before the solve so their latency overlaps it
        double qr[NPT], qo[NPT], qs[NPT];
        double* dst = a.qrhs + (1 + c) * fs;
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          qr[t] = ldg(dst, S.g[t]);
          if constexpr (RKM) {
            qo[t] = ldg(a.q + (1 + c) * fs, S.g[t]);
            qs[t] = sa.stage == 0 ? 0.0 : ldg(a.qsave + (1 + c) * fs, S.g[t]);
          }
        }
        if (c < 3) {
          const int col = c == 0 ? c1 : (c == 1 ? c2 : 6 + dir);
#pragma unroll
          for (int t = 0; t < NPT; t++) nx[t] = ldg(a.sg + col * fs, S.g[t]);
        }
        __syncthreads();
        {
          ChunkCoef<C::R, false> CD;
          lds_chunk_coef<C>(CD, cs, p * C::R);
          solve_pass_line<C, false, GEN>(Tl, Xl, n, p, CD, a.fc, a.bnd[dir]);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          const double v = qr[t] + (MAIN ? X[S.s[t]] : X[S.s[t]] * idn);
          if constexpr (RKM) {
            const double jb = jac_of<GEN>(a, S.g[t]);
            double s0 = qs[t];
            if (sa.stage == 0) {
              s0 = qo[t] * jb;
              stg(a.qsave + (1 + c) * fs, S.g[t], s0);
            }
            qn[1 + c][t] = rk_update_main(sa, a.dt, jb, s0, qo[t], v);
            stg(a.q + (1 + c) * fs, S.g[t], qn[1 + c][t]);
          } else {
            stg(dst, S.g[t], v);
          }
        }
      }
      if constexpr (RKM) {
        // updatefvar (fludyna.F90:191-242): rho, u, v, w, p, T of the new q
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          const Prim P = prim_of(a, qn[0][t], qn[1][t], qn[2][t], qn[3][t], qn[4][t]);
          const unsigned o = S.g[t];
          stg(a.pr, o, P.r);
          stg(a.pr + fs, o, P.u);
          stg(a.pr + 2 * fs, o, P.v);
          stg(a.pr + 3 * fs, o, P.w);
          stg(a.pr + 4 * fs, o, P.p);
          stg(a.pr + 5 * fs, o, P.t);
        }
        if (a.rk_filter) {
          // the next stage's filterq sweep along dir (comsolver.F90:514-632),
          // as the filter pass does it, on the new q held in registers; the
          // primitives above stay those of the unfiltered q, as after the
          // reference's filterq.  The barrier before the solve orders the
          // staging below after every read of the viscous part's LDS.
          const double* csf = stage_chunk_coefs<C>(a.fac[dir], n + 7, smem, 1);
          const double c3l = a.fac[dir][2 * (n + 7) + (n + 7) - 1];
          const FilterCoef fl = stage_halo_coefs<C>(a.fc, smem);
#pragma unroll
          for (int m = 0; m < 5; m++) {
#pragma unroll
            for (int t = 0; t < NPT; t++) put<C>(T, S, t, qn[m][t]);
            __syncthreads();
            {
              ChunkCoef<C::R, true> CF;
              lds_chunk_coef_f<C>(CF, csf, p * C::R, n + 7, c3l);
              solve_pass_line<C, true, GEN>(Tl, Xl, n, p, CF, fl, a.bnd[dir]);
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < NPT; t++) stg(a.q + m * fs, S.g[t], X[S.s[t]]);
          }
        }
      }
    } else {
      const double* csf = stage_chunk_coefs<C>(a.fac[dir], n + 7, smem, 1);
      const double c3l = a.fac[dir][2 * (n + 7) + (n + 7) - 1];
      const FilterCoef fl = stage_halo_coefs<C>(a.fc, smem);
      // Loads run two solves ahead: variable c+1's qrhs, q, qsave are issued
      // before variable c's filter solve, the next stress column before
      // variable c's derivative solve.
      double qr[NPT], qo[NPT], qs[NPT];
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        const unsigned o = S.g[t];
        const double q0 = ldg(a.q, o);
        double s0 = q0;
        if (sa.stage == 0) stg(a.qsave, o, q0);
        else s0 = ldg(a.qsave, o);
        put<C>(T, S, t, rk_update(sa, a.dt, s0, q0, ldg(a.qrhs, o)));
      }
#pragma unroll
      for (int t = 0; t < NPT; t++) {
        qr[t] = ldg(a.qrhs + fs, S.g[t]);
        qo[t] = ldg(a.q + fs, S.g[t]);
        qs[t] = sa.stage == 0 ? 0.0 : ldg(a.qsave + fs, S.g[t]);
      }
      __syncthreads();
      {
          ChunkCoef<C::R, true> CF;
          lds_chunk_coef_f<C>(CF, csf, p * C::R, n + 7, c3l);
          solve_line<C::P, C::R, true>(Tl, Xl, n, p, CF, fl);
        }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < NPT; t++) stg(a.q, S.g[t], X[S.s[t]]);
#pragma unroll
      for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int t = 0; t < NPT; t++)
          put<C>(T, S, t, nx[t]);
        if (c < 3) {
          const int col = c == 0 ? c1 : (c == 1 ? c2 : 6 + dir);
#pragma unroll
          for (int t = 0; t < NPT; t++) nx[t] = ldg(a.sg + col * fs, S.g[t]);
        }
        __syncthreads();
        {
          ChunkCoef<C::R, false> CD;
          lds_chunk_coef<C>(CD, cs, p * C::R);
          solve_line<C::P, C::R, false>(Tl, Xl, n, p, CD, a.fc);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NPT; t++) {
          const double v = qr[t] + X[S.s[t]] * idn;
          double s0 = qs[t];
          if (sa.stage == 0) {
            stg(a.qsave + (1 + c) * fs, S.g[t], qo[t]);
            s0 = qo[t];
          }
          put<C>(T, S, t, rk_update(sa, a.dt, s0, qo[t], v));
        }
        if (c < 3) {
#pragma unroll
          for (int t = 0; t < NPT; t++) {
            qr[t] = ldg(a.qrhs + (2 + c) * fs, S.g[t]);
            qo[t] = ldg(a.q + (2 + c) * fs, S.g[t]);
            qs[t] = sa.stage == 0 ? 0.0 : ldg(a.qsave + (2 + c) * fs, S.g[t]);
          }
        }
        __syncthreads();
        {
          ChunkCoef<C::R, true> CF;
          lds_chunk_coef_f<C>(CF, csf, p * C::R, n + 7, c3l);
          solve_line<C::P, C::R, true>(Tl, Xl, n, p, CF, fl);
        }
        __syncthreads();
        double* dq = a.q + (1 + c) * fs;
#pragma unroll
        for (int t = 0; t < NPT; t++) stg(dq, S.g[t], X[S.s[t]]);
      }
    }
  } else {  // F_ONLY: filter of q along dir, in place (filterq, tgvsolver.F90:791-883)
    const double* csf = stage_chunk_coefs<C>(a.fac[dir], n + 7, smem, 0);
    const double c3l = a.fac[dir][2 * (n + 7) + (n + 7) - 1];
    const FilterCoef fl = stage_halo_coefs<C>(a.fc, smem);
#pragma unroll
    for (int m = 0; m < 5; m++) {
#pragma unroll
      for (int t = 0; t < NPT; t++)
        put<C>(T, S, t, ldg(a.q + m * fs, S.g[t]));
      __syncthreads();
      {
        ChunkCoef<C::R, true> CF;
        lds_chunk_coef_f<C>(CF, csf, p * C::R, n + 7, c3l);
        solve_pass_line<C, true, GEN>(Tl, Xl, n, p, CF, fl, a.bnd[dir]);
      }
      __syncthreads();
      double* dst = a.q + m * fs;
#pragma unroll
      for (int t = 0; t < NPT; t++)
        stg(dst, S.g[t], X[S.s[t]]);
    }
  }
}

// one pass of one case: block b handles tile xcd_tile(b)
// GEN (main path only): the general-grid form, for a direction with
// physical ends or a grid with per-point metrics; the periodic, uniform-grid
// kernels leave that code out (it would multiply their code size several
// times over, past what the instruction cache holds)
template <class C, int PROG, int MINB, bool MAIN = false, bool GEN = false>
__global__ __launch_bounds__(C::NT, MINB) void k_pass(FlowArgs a, int dir) {
  extern __shared__ double smem[];
  pass_body<C, PROG, MAIN, GEN>(a, stage_of(a), dir, xcd_tile(blockIdx.x, gridDim.x), smem);
}

// The tiles of one pass over several cases: case k owns global tiles
// tstart[k]..tstart[k+1]-1.  Device arrays, built by the batch host code.
struct BatchLaunch {
  const FlowArgs* cases;
  const int* tstart;
  int ncases;
  int total;
};

// one pass over a batch of cases: persistent blocks (grid a multiple of 8);
// the blocks of XCD x walk the contiguous range [x*total/8, (x+1)*total/8) so
// neighbouring tiles share that XCD's L2
template <class C, int PROG, int MINB>
__global__ __launch_bounds__(C::NT, MINB) void k_pass_batch(BatchLaunch bl, int dir, StageArgs sa) {
  extern __shared__ double smem[];
  const int x = blockIdx.x & 7, per = gridDim.x >> 3;
  const int lo = (int)((long long)bl.total * x / 8), hi = (int)((long long)bl.total * (x + 1) / 8);
  for (int g = lo + (int)(blockIdx.x >> 3); g < hi; g += per) {
    int kl = 0, kh = bl.ncases;  // tstart[kl] <= g < tstart[kh]
    while (kh - kl > 1) {
      const int mid = (kl + kh) >> 1;
      if (bl.tstart[mid] <= g) kl = mid;
      else kh = mid;
    }
    pass_body<C, PROG>(bl.cases[kl], sa, dir, g - bl.tstart[kl], smem);
    __syncthreads();
  }
}

template <class C, int MINB>
hipError_t launch_batch_c(int prog, const BatchLaunch& bl, int dir, const StageArgs& sa, int nblocks,
                          hipStream_t st) {
  const size_t lds = (size_t)lds_doubles<C>(prog) * sizeof(double);
  dim3 g(nblocks), b(C::NT);
  switch (prog) {
    case GC_FIRST: hipLaunchKernelGGL((k_pass_batch<C, GC_FIRST, MINB>), g, b, lds, st, bl, dir, sa); break;
    case GC_MID: hipLaunchKernelGGL((k_pass_batch<C, GC_MID, MINB>), g, b, lds, st, bl, dir, sa); break;
    case GC_LAST: hipLaunchKernelGGL((k_pass_batch<C, GC_LAST, MINB>), g, b, lds, st, bl, dir, sa); break;
    case V_MID: hipLaunchKernelGGL((k_pass_batch<C, V_MID, MINB>), g, b, lds, st, bl, dir, sa); break;
    case V_LAST_RHS: hipLaunchKernelGGL((k_pass_batch<C, V_LAST_RHS, MINB>), g, b, lds, st, bl, dir, sa); break;
    case V_LAST_RK: hipLaunchKernelGGL((k_pass_batch<C, V_LAST_RK, MINB>), g, b, lds, st, bl, dir, sa); break;
    case GL_GRAD: hipLaunchKernelGGL((k_pass_batch<C, GL_GRAD, MINB>), g, b, lds, st, bl, dir, sa); break;
    case GL_CONV: hipLaunchKernelGGL((k_pass_batch<C, GL_CONV, MINB>), g, b, lds, st, bl, dir, sa); break;
    case FGC_FIRST: return hipErrorInvalidValue;  // single-case passes only
    default: hipLaunchKernelGGL((k_pass_batch<C, F_ONLY, MINB>), g, b, lds, st, bl, dir, sa); break;
  }
  return hipGetLastError();
}

template <class C, int MINB, bool MAIN, bool GEN>
void launch_pass_m(int prog, const FlowArgs& a, int dir, hipStream_t st) {
  const int nt = tiles_of(a.n, dir, C::TL);
  const size_t lds = (size_t)lds_doubles<C>(prog) * sizeof(double);
  dim3 g(nt), b(C::NT);
  switch (prog) {
    case GC_FIRST: hipLaunchKernelGGL((k_pass<C, GC_FIRST, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case GC_MID: hipLaunchKernelGGL((k_pass<C, GC_MID, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case GC_LAST: hipLaunchKernelGGL((k_pass<C, GC_LAST, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case V_MID: hipLaunchKernelGGL((k_pass<C, V_MID, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case V_LAST_RHS: hipLaunchKernelGGL((k_pass<C, V_LAST_RHS, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case V_LAST_RK: hipLaunchKernelGGL((k_pass<C, V_LAST_RK, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case GL_GRAD: hipLaunchKernelGGL((k_pass<C, GL_GRAD, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    case GL_CONV: hipLaunchKernelGGL((k_pass<C, GL_CONV, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
    // the filtering first pass exists for the miniapp's stage only (the main
    // path filters at the start of its stage); launch_pass_c refuses it there
    case FGC_FIRST:
      if constexpr (!MAIN) hipLaunchKernelGGL((k_pass<C, FGC_FIRST, MINB, false, false>), g, b, lds, st, a, dir);
      break;
    // the filter pass of the main path also takes physical boundaries
    default: hipLaunchKernelGGL((k_pass<C, F_ONLY, MINB, MAIN, GEN>), g, b, lds, st, a, dir); break;
  }
}

template <class C, int MINB>
hipError_t launch_pass_c(int prog, const FlowArgs& a, int dir, hipStream_t st) {
  if (prog == FGC_FIRST && (a.main_sem || dir != 0)) return hipErrorInvalidValue;
  if (a.main_sem && (a.met || a.bnd[dir])) launch_pass_m<C, MINB, true, true>(prog, a, dir, st);
  else if (a.main_sem) launch_pass_m<C, MINB, true, false>(prog, a, dir, st);
  else launch_pass_m<C, MINB, false, false>(prog, a, dir, st);
  return hipGetLastError();
}

// launch shapes (threads, lines per tile, chunk length, point slots per thread)
using CfgA = LineCfg<256, 4, 5, 5>;    // P=64, n <= 313
using CfgB = LineCfg<256, 4, 9, 9>;    // P=64, n <= 569
using CfgC = LineCfg<256, 16, 5, 5>;   // P=16, n <= 73
using CfgD = LineCfg<256, 8, 9, 9>;    // P=32, n <= 281
using CfgE = LineCfg<256, 8, 5, 5>;    // P=32, n <= 153
using CfgF = LineCfg<512, 8, 5, 5>;    // P=64, n <= 313: CfgA's lines, 64-byte row segments
using CfgG = LineCfg<128, 4, 9, 9>;    // P=32, n <= 281: CfgD's lines, two workgroups per CU at 1 wave/SIMD
using CfgH = LineCfg<128, 2, 5, 5>;    // P=64, n <= 313: CfgA's lines in half-size workgroups
using CfgI = LineCfg<64, 1, 5, 5>;     // P=64, n <= 313: one line per single-wavefront workgroup
using CfgJ = LineCfg<64, 2, 9, 9>;     // P=32, n <= 281: two lines per single-wavefront workgroup
using CfgK = LineCfg<256, 16, 9, 9>;   // P=16, n <= 137: 128-byte row segments for mid-size lines
using CfgL = LineCfg<512, 8, 9, 9>;    // P=64, n <= 569: CfgB's lines, 64-byte row segments
using CfgM = LineCfg<64, 1, 9, 9>;     // P=64, n <= 569: one long line per single-wavefront workgroup
constexpr int kNumShapes = 13;
constexpr int kMinbVariants = 2;       // MINB = 1, 2

// one per shape, each in its own translation unit (passes.hip)
#define ASTR_DECLARE_SHAPE(K)                                                                 \
  hipError_t launch_pass_s##K(int prog, int minb, const FlowArgs& a, int dir, hipStream_t st); \
  hipError_t launch_batch_s##K(int prog, int minb, const BatchLaunch& bl, int dir,             \
                               const StageArgs& sa, int nblocks, hipStream_t st);
ASTR_DECLARE_SHAPE(0)
ASTR_DECLARE_SHAPE(1)
ASTR_DECLARE_SHAPE(2)
ASTR_DECLARE_SHAPE(3)
ASTR_DECLARE_SHAPE(4)
ASTR_DECLARE_SHAPE(5)
ASTR_DECLARE_SHAPE(6)
ASTR_DECLARE_SHAPE(7)
ASTR_DECLARE_SHAPE(8)
ASTR_DECLARE_SHAPE(9)
ASTR_DECLARE_SHAPE(10)
ASTR_DECLARE_SHAPE(11)
ASTR_DECLARE_SHAPE(12)
#undef ASTR_DECLARE_SHAPE

}  // namespace astr
