// upwind.hpp -- host interface of the upwind-biased convective pass
// (upwind.hip): the reference's explicit upwind-biased convection
// (conschm with an odd first digit and suffix 'e'; convrsduwd,
// src/solver.F90:548-1201) and its shock sensor (ducrossensor,
// src/commcal.F90:196-360), on the main path's stored primitives and halo
// sources, for directions that are periodic on one rank.
#pragma once
#include <hip/hip_runtime.h>

#include "flow_dev.hpp"

namespace astr {

// the per-case state of the upwind pass
struct UwArgs {
  int recon;            // recon_schem: -1, 0 SUW7, 1 WENO7, 2 WENO7-Z, 3 MP7, 5 MP7-LD, 6 ROUND
  int chardec;          // lchardecomp
  int sensor;           // the shock flags are computed (recon 5 or chardec)
  double bfac;          // bfacmpld (MP7-LD bandwidth factor)
  double shkcrt;        // shock-sensor threshold
  double gamma, mach;
  double* sw;           // split buffer: kUwFields (chardec: kUwFieldsCd) fields of sws doubles
  long long sws;
  double* ssf;          // Ducros sensor, field layout of FlowArgs (computed points)
  unsigned char* lsh;   // shock flags, field layout of FlowArgs
  // the compact upwind-biased scheme (conschm 543c): the LU factors of its
  // line systems (cuw_ac_offset) and the interface-flux buffer, 10 fields of
  // sws doubles in the split buffer's layout
  int compact;
  const double* cuw_ac;
  double* fhc;
};

// LU factors ac(1:3, -2:n+1) of the compact flux systems: for each direction
// the upwind ('+') then the downwind ('-') system, 3 (n+4) doubles each
__host__ __device__ inline long long cuw_ac_offset(const int n[3], int dir, bool plus) {
  long long o = 0;
  for (int d = 0; d < dir; d++) o += 6ll * (n[d] + 4);
  return o + (plus ? 0 : 3ll * (n[dir] + 4));
}
__host__ __device__ inline long long cuw_ac_doubles(const int n[3], int ndir) {
  return cuw_ac_offset(n, ndir, true);
}

// fields of the split buffer: F+ (5), F- (5), then with the characteristic
// decomposition rho, u, v, w, p, E and the metric row dxi(dir, 0..2)
constexpr int kUwFields = 10;
constexpr int kUwFieldsCd = 19;

// rows -4..n+4 of every line of direction dir
__host__ __device__ inline long long uw_lines(const int n[3], int dir) {
  if (dir == 0) return (long long)(n[1] + 1) * (n[2] + 1);
  if (dir == 1) return (long long)(n[0] + 1) * (n[2] + 1);
  return (long long)(n[0] + 1) * (n[1] + 1);
}
__host__ __device__ inline long long uw_points(const int n[3], int dir) {
  return uw_lines(n, dir) * (n[dir] + 9);
}

// the Ducros sensor and the shock flags (after the gradient passes stored
// the whole velocity gradient)
hipError_t uw_launch_sensor(const FlowArgs& a, const UwArgs& u, hipStream_t st);
// the Steger-Warming split of every line of direction dir into u.sw
hipError_t uw_launch_split(const FlowArgs& a, const UwArgs& u, int dir, hipStream_t st);
// the interface fluxes of direction dir from u.sw, into qrhs: the first
// direction stores, later ones accumulate, the last one also negates
// (rhscal's qrhs = -qrhs); explicit (recons_exp) or, with u.compact, the
// compact scheme's line solves, limited interface fluxes and difference
hipError_t uw_launch_face(const FlowArgs& a, const UwArgs& u, int dir, bool first, bool last,
                          hipStream_t st);

}  // namespace astr
