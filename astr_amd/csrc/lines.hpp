// lines.hpp -- device engine for many independent compact line solves.
//
// A workgroup owns a tile of TL whole lines of one direction.  The tile's
// points are spread over the NT threads for the pointwise work (NPT slots per
// thread, coalesced along the fastest memory index), and the solve maps each
// line onto P = NT/TL consecutive lanes of one wavefront, lane p owning the
// chunk of R consecutive unknowns [p*R, p*R+R).
//
// The tridiagonal systems are the reference's (derivative.F90:63-158,
// filter.F90:31-100) with the reference's LU factors ac1,ac2,ac3
// (commfunc.F90:752-774).  The reference's Thomas recurrences
//     d'(u) = d(u)*ac2(u) - d'(u-1)*ac3(u)        (forward)
//     x(u)  = d'(u) - ac1(u)*x(u+1)                (backward)
// (commfunc.F90:790-813) are first-order linear recurrences; each lane runs
// them over its chunk with a zero carry, the carries across the P chunks are
// composed by a Kogge-Stone scan over lanes (DPP row shifts inside each row
// of 16 lanes; across rows DPP row broadcasts, with a wavefront shift first
// for the backward scan's row pairs, and v_readlane only for the backward
// scan's step between the 32-lane halves; the exclusive carries by DPP
// wavefront shifts), and each lane then adds
// carry * (running product of the recurrence coefficients).  This is the
// same linear map as the sequential sweep, re-associated.
//
// LDS line layout: the input line holds rows -8..P*R+2 at offset 8, i.e. the
// reference's halo f(-5:n+5) (written explicitly) plus zero padding, so every
// chunk reads its stencil window at compile-time offsets.
#pragma once
#include <hip/hip_runtime.h>

namespace astr {

// stencil coefficients of 643c (constdef.F90:17-19)
constexpr double k7d9 = 7.0 / 9.0;
constexpr double k1d36 = 1.0 / 36.0;
constexpr double k1d60 = 1.0 / 60.0;

constexpr int kLineOff = 8;  // LDS offset of row 0 in an input line

// filter RHS coefficients, filled by the host from filter_coefficient_cal
// (filter.F90:299-432): c10[0..5] interior rows by value, the halo rows'
// coefh(k,j) as a device table ch[k*11+j], k=0..2.
struct FilterCoef {
  double c10[6];
  const double* ch;
};

// ---------------------------------------------------------------------------
// cross-lane helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lo32(double v) { return (int)(__double_as_longlong(v) & 0xffffffffll); }
__device__ __forceinline__ int hi32(double v) { return (int)(__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double mk64(int lo, int hi) {
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// DPP move inside rows of 16 lanes; lanes whose source is outside the row
// receive `old`.
template <int CTRL>
__device__ __forceinline__ double dpp(double v, double old) {
  const int l = __builtin_amdgcn_update_dpp(lo32(old), lo32(v), CTRL, 0xf, 0xf, false);
  const int h = __builtin_amdgcn_update_dpp(hi32(old), hi32(v), CTRL, 0xf, 0xf, false);
  return mk64(l, h);
}

// DPP move inside rows of 16 lanes; lanes whose source is outside the row
// receive 0.
template <int CTRL>
__device__ __forceinline__ double dpp0(double v) {
  const int l = __builtin_amdgcn_mov_dpp(lo32(v), CTRL, 0xf, 0xf, true);
  const int h = __builtin_amdgcn_mov_dpp(hi32(v), CTRL, 0xf, 0xf, true);
  return mk64(l, h);
}

// DPP move with a row mask: rows outside ROWMASK receive `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dppm(double v, double old) {
  const int l = __builtin_amdgcn_update_dpp(lo32(old), lo32(v), CTRL, ROWMASK, 0xf, false);
  const int h = __builtin_amdgcn_update_dpp(hi32(old), hi32(v), CTRL, ROWMASK, 0xf, false);
  return mk64(l, h);
}

__device__ __forceinline__ double readlane(double v, int lane) {
  return mk64(__builtin_amdgcn_readlane(lo32(v), lane), __builtin_amdgcn_readlane(hi32(v), lane));
}

// Inclusive scan of affine maps x -> A*x + B over segments of P lanes, in the
// direction of increasing lane (FWD) or decreasing lane (!FWD).  On return
// (A,B) of lane p composes the maps of lanes segstart..p (FWD) or p..segend.
// The exclusive carry (B of the neighbour towards the segment start, 0 at
// the segment boundary) is returned.
template <int P, bool FWD>
__device__ __forceinline__ double affine_scan(double& A, double& B) {
  static_assert(P == 16 || P == 32 || P == 64, "segment must be 16, 32 or 64 lanes");
  const int lane = threadIdx.x & 63;
  const int row = lane >> 4;
#define ASTR_SCAN_STEP(S)                                                  \
  {                                                                        \
    const double A2 = dpp<(FWD ? 0x110 : 0x100) + S>(A, 1.0);              \
    const double B2 = dpp0<(FWD ? 0x110 : 0x100) + S>(B);                  \
    B = fma(A, B2, B);                                                     \
    A = A * A2;                                                            \
  }
  ASTR_SCAN_STEP(1)
  ASTR_SCAN_STEP(2)
  ASTR_SCAN_STEP(4)
  ASTR_SCAN_STEP(8)
#undef ASTR_SCAN_STEP
  if constexpr (P >= 32 && FWD) {
    // rows 1 and 3 compose the totals of lanes 15 and 47 (row_bcast:15)
    const double Ac = dppm<0x142, 0xA>(A, 1.0);
    const double Bc = dppm<0x142, 0xA>(B, 0.0);
    B = fma(A, Bc, B);
    A = A * Ac;
  } else if constexpr (P >= 32) {
    // rows 0 and 2 compose the totals of lanes 16 and 48: wave_shl:1 brings
    // them to lanes 15 and 47, row_newbcast:15 spreads those over rows 0, 2
    const double As = dpp0<0x130>(A), Bs = dpp0<0x130>(B);
    const double Ac = dppm<0x15F, 0x5>(As, 1.0);
    const double Bc = dppm<0x15F, 0x5>(Bs, 0.0);
    B = fma(A, Bc, B);
    A = A * Ac;
  }
  if constexpr (P == 64 && FWD) {
    // rows 2 and 3 compose the total of lane 31 (row_bcast:31)
    const double Aa = dppm<0x143, 0xC>(A, 1.0);
    const double Ba = dppm<0x143, 0xC>(B, 0.0);
    B = fma(A, Ba, B);
    A = A * Aa;
  } else if constexpr (P == 64) {
    const double Aa = readlane(A, FWD ? 31 : 32), Ba = readlane(B, FWD ? 31 : 32);
    const bool recv = FWD ? row >= 2 : row <= 1;
    B = fma(A, recv ? Ba : 0.0, B);
    A = A * (recv ? Aa : 1.0);
  }
  // exclusive carry: B of the neighbouring lane towards the segment start
  // (wave_shr:1 forward, wave_shl:1 backward, across rows), 0 at the
  // segment boundary
  (void)row;
  const double c = dpp0<FWD ? 0x138 : 0x130>(B);
  const bool start = FWD ? (lane & (P - 1)) == 0 : (lane & (P - 1)) == P - 1;
  return start ? 0.0 : c;
}

// LU factors of one chunk, held in registers, with encodings that make the
// recurrences branch-free: the first row, which the reference's solve leaves
// as is, has c2=1, c3=0; the last row has c1=0; rows past the end have all
// factors 0.  For rows with a = c the reference's factors (commfunc.F90:
// 752-774) ac1 = c/den and ac3 = a/den are equal, so one array serves both
// with two exceptions:
//   - the first row's c3 is never used: it multiplies the zero carry that
//     enters each line, and its product is never composed with anything to
//     its left;
//   - the filter's last row has a = beta_h, c = 0.  Its c3 is kept as one
//     value, c3last, at chunk position klast (-1 if the chunk does not hold
//     that row).
// The derivative's last row has a = c = 0, so there no exception is needed.
// FILT selects the filter's last-row exception.
template <int R, bool FILT>
struct ChunkCoef {
  double c2[R];
  double c1[R];
  int klast = -1;
  double c3last = 0.0;
  __device__ __forceinline__ double f3(int k) const {
    if constexpr (FILT) return k == klast ? c3last : c1[k];
    else return c1[k];
  }
};

template <int R, bool FILT>
__device__ __forceinline__ void load_chunk_coef(ChunkCoef<R, FILT>& C,
                                                const double* __restrict__ ac, int M, int u0) {
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int u = u0 + k;
    const bool ok = u < M;
    const int uu = ok ? u : 0;
    double a1 = ac[uu], a2 = ac[M + uu];
    if (u == 0) a2 = 1.0;
    if (u == M - 1) a1 = 0.0;
    C.c2[k] = ok ? a2 : 0.0;
    C.c1[k] = ok ? a1 : 0.0;
  }
  if constexpr (FILT) {
    const int kl = M - 1 - u0;
    C.klast = kl >= 0 && kl < R ? kl : -1;
    C.c3last = ac[2 * M + M - 1];
  }
}

// Right-hand-side rows of a line whose ends may be physical boundaries
// (bnd bit 0: physical boundary at r=0, bit 1: at r=n; ntype 1, 2, 4 of the
// reference), or v where row r is an interior row.  An end that is not
// physical is an interface and takes the periodic form's rows.  The system
// keeps the periodic form's unknowns -OFF..n+OFF; at a physical end the rows
// beyond the boundary are decoupled (a = c = 0, right-hand side 0, so they
// solve to 0) and the boundary row's sub/super-diagonal is the reference's
// (a = c, the value the reference's solve reads), so that the
// recurrences over rows 0..n are the reference's.
//   derivative (compact_fd_rhs, derivative.F90:229-289): row 0 the
//     3rd-order biased -5/2 f0 + 2 f1 + 1/2 f2, row 1 the 4th-order
//     compact 3/4 (f2 - f0), mirrored at n;
//   filter (compact_filter_rhs, filter.F90:175-262): rows 0..2 the one-sided
//     coefb(k, 0..6), row 3 the 6th-order and row 4 the 8th-order central
//     rows, mirrored at n; interface halo rows from coefh.
// fc.ch holds coefh (33), then coefb rows 0..2 (21), coef6i (4), coef8i (5).
constexpr int kBndCoefs = 63;
template <bool FILT>
__device__ double boundary_row(const double* __restrict__ Tl, int n, int r, int bnd,
                               const FilterCoef& fc, double v) {
  const bool bl = bnd & 1, br = bnd & 2;
  if constexpr (!FILT) {
    if (bl) {
      if (r < 0) return 0.0;
      if (r == 0) return -2.5 * Tl[0] + 2.0 * Tl[1] + 0.5 * Tl[2];
      if (r == 1) return 0.75 * (Tl[2] - Tl[0]);
    }
    if (br) {
      if (r > n) return 0.0;
      if (r == n) return 2.5 * Tl[n] - 2.0 * Tl[n - 1] - 0.5 * Tl[n - 2];
      if (r == n - 1) return 0.75 * (Tl[n] - Tl[n - 2]);
    }
    return v;
  } else {
    const double* __restrict__ ch = fc.ch;
    const double* __restrict__ cb = ch + 33;
    const double* __restrict__ c6 = ch + 54;
    const double* __restrict__ c8 = ch + 58;
    if (r < -3 || r > n + 3) return v;  // a chunk slot past the system's rows
    int j = -1;  // a 6th/8th-order row centred at j
    if (bl) {
      if (r < 0) return 0.0;
      if (r <= 2) {
        double s = 0.0;
        for (int t = 0; t <= 6; t++) s = s + cb[r * 7 + t] * Tl[t];
        return s;
      }
      if (r == 3 || r == 4) j = r;
    } else if (r < 0) {
      double s = 0.0;
      for (int t = 0; t <= 10; t++) s = s + ch[(r + 3) * 11 + t] * Tl[t - 5];
      return s;
    }
    if (br) {
      if (r > n) return 0.0;
      if (r >= n - 2) {
        double s = 0.0;
        for (int t = 0; t <= 6; t++) s = s + cb[(n - r) * 7 + t] * Tl[n - t];
        return s;
      }
      if (r == n - 3 || r == n - 4) j = r;
    } else if (r > n) {
      double s = 0.0;
      for (int t = 0; t <= 10; t++) s = s + ch[(n + 3 - r) * 11 + t] * Tl[n + 5 - t];
      return s;
    }
    if (j < 0) return v;
    const bool six = (bl && j == 3) || (br && j == n - 3);
    if (six)
      return c6[0] * (Tl[j] + Tl[j]) + c6[1] * (Tl[j + 1] + Tl[j - 1]) + c6[2] * (Tl[j + 2] + Tl[j - 2]) +
             c6[3] * (Tl[j + 3] + Tl[j - 3]);
    return c8[0] * (Tl[j] + Tl[j]) + c8[1] * (Tl[j + 1] + Tl[j - 1]) + c8[2] * (Tl[j + 2] + Tl[j - 2]) +
           c8[3] * (Tl[j + 3] + Tl[j - 3]) + c8[4] * (Tl[j + 4] + Tl[j - 4]);
  }
}

// Solve one line held in LDS.
//   T  : input line, row r at T[kLineOff + r]; rows -5..n+5 hold the line
//        and its halo, the padding rows hold 0
//   X  : output, x(r) at X[r], r = 0..n
// FILT=false: 643c derivative, unknowns -1..n+1 (M=n+3), explicit 6th-order
//             rows at -1 and n+1 (derivative.F90:250-304).
// FILT=true : 10th-order filter, unknowns -3..n+3 (M=n+7), halo rows
//             -3..-1, n+1..n+3 from coefh (filter.F90:206-283).
// BND=true: the ends carry the boundary types bnd (boundary_row); the filter
//             then also resets a physical end to its input value
//             (filter.F90:141-142).
template <int P, int R, bool FILT, bool BND = false>
__device__ __forceinline__ void solve_line(const double* __restrict__ T, double* __restrict__ X,
                                           int n, int p, const ChunkCoef<R, FILT>& C,
                                           const FilterCoef& fc, int bnd = 0) {
  constexpr int OFF = FILT ? 3 : 1;
  constexpr int H = FILT ? 5 : 3;
  constexpr int W = R + 2 * H;
  const int u0 = p * R;
  const double* __restrict__ Tw = T + kLineOff + (u0 - OFF - H);

  double w[W];
#pragma unroll
  for (int k = 0; k < W; k++) w[k] = Tw[k];

  double y[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    const double* c = w + k + H;  // c[s] = f(r+s), r = u0+k-OFF
    if (FILT) {
      const double var0 = c[0] + c[0];
      const double var1 = c[1] + c[-1];
      const double var2 = c[2] + c[-2];
      const double var3 = c[3] + c[-3];
      const double var4 = c[4] + c[-4];
      const double var5 = c[5] + c[-5];
      y[k] = fc.c10[0] * var0 + fc.c10[1] * var1 + fc.c10[2] * var2 + fc.c10[3] * var3 +
             fc.c10[4] * var4 + fc.c10[5] * var5;
    } else {
      const int r = u0 + k - OFF;
      const double var1 = c[1] - c[-1];
      const double var2 = c[2] - c[-2];
      const double var3 = c[3] - c[-3];
      const double dc = k7d9 * var1 + k1d36 * var2;
      const double de = 0.75 * var1 - 0.15 * var2 + k1d60 * var3;
      y[k] = (r == -1 || r == n + 1) ? de : dc;
    }
  }

  if constexpr (BND) {
#pragma unroll
    for (int k = 0; k < R; k++) y[k] = boundary_row<FILT>(T + kLineOff, n, u0 + k - OFF, bnd, fc, y[k]);
  } else if constexpr (FILT) {
    // the six halo rows: lane q<6 of each line evaluates row q (-3,-2,-1,
    // n+1,n+2,n+3) over f(-5..5) or f(n+5..n-5) and leaves it in X[n+1+q],
    // a row of the output line that nothing reads before this solve writes
    // X[0..n]; a line's lanes are one wavefront, whose LDS accesses complete
    // in order, so the owners read the six values back after a wave barrier
    double ev = 0.0;
    if (p < 6) {
      const int hr = p < 3 ? p : 5 - p;  // coefh row
      const double* __restrict__ Tl = T + kLineOff;
#pragma unroll
      for (int j = 0; j <= 10; j++) {
        const double f = p < 3 ? Tl[j - 5] : Tl[n + 5 - j];
        ev = ev + fc.ch[hr * 11 + j] * f;
      }
    }
    double* hx = X + n + 1;
    if (p < 6) hx[p] = ev;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double mine[6];
#pragma unroll
    for (int q = 0; q < 6; q++) mine[q] = hx[q];
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int r = u0 + k - OFF;
      if (k < 3) {
        y[k] = r == -3 ? mine[0] : (r == -2 ? mine[1] : (r == -1 ? mine[2] : y[k]));
      }
      const int q = r - n - 1;
      y[k] = q == 0 ? mine[3] : (q == 1 ? mine[4] : (q == 2 ? mine[5] : y[k]));
    }
  }

  // forward recurrence over the chunk with zero carry-in
  double B = 0.0, A = 1.0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    const double c3 = C.f3(k);
    B = y[k] * C.c2[k] - B * c3;
    A *= -c3;
    y[k] = B;
  }
  const double cin = affine_scan<P, true>(A, B);
  {
    double m = 1.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      m *= -C.f3(k);
      y[k] = fma(m, cin, y[k]);
    }
  }

  // backward recurrence with zero carry from the right
  B = 0.0;
  A = 1.0;
#pragma unroll
  for (int k = R - 1; k >= 0; k--) {
    const double c1 = C.c1[k];
    B = y[k] - c1 * B;
    A *= -c1;
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
    double v = y[k];
    if constexpr (BND && FILT) {
      if (((bnd & 1) && r == 0) || ((bnd & 2) && r == n)) v = T[kLineOff + r];
    }
    if (r >= 0 && r <= n) X[r] = v;
  }
}

// LDS stride of one line: at least rows -8..P*R+2, and = 32/TL (mod 32)
// doubles, so that the y/z passes' point slots (TL adjacent lines x 32/TL
// consecutive rows per half-wave) hit distinct banks; the solve's chunk reads
// (stride R, odd) stay conflict-free for any stride.
constexpr int line_stride(int P, int R, int TL) {
  int L = P * R + 11;
  const int want = (32 / TL) % 32;
  while (L % 32 != want) L++;
  return L;
}

// Launch shape of a line kernel.
template <int NT_, int TL_, int R_, int NPT_>
struct LineCfg {
  static constexpr int NT = NT_;   // threads per workgroup
  static constexpr int TL = TL_;   // lines per tile
  static constexpr int P = NT_ / TL_;  // lanes (chunks) per line
  static constexpr int R = R_;     // unknowns per chunk
  static constexpr int NPT = NPT_; // point slots per thread
  static constexpr int LT = line_stride(P, R_, TL_);  // line stride in LDS (rows -8..P*R+2)
  static_assert(P == 16 || P == 32 || P == 64, "a line's chunks must fill 16, 32 or 64 lanes");
  static_assert((TL_ & (TL_ - 1)) == 0 && TL_ <= 32, "TL must be a power of two, at most 32");
  static_assert(R_ >= 3, "chunk 0 must hold the three filter halo rows");
};

}  // namespace astr
