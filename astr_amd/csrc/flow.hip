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
//   F y, F x        the y and x sweeps of filterq (:805-847); optionally the
//                   x sweep runs at the start of the next stage's GC x pass
//
// The filter sweeps act on different indices and are linear, so applying
// them z,y,x instead of x,y,z gives the same field up to rounding.
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
#include "flow_dev.hpp"
#include "upwind.hpp"

using namespace astr;

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

// device scratch released on every return path (hipFree waits for the
// device, so no kernel still uses it)
struct DevBuf {
  double* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
};

// Failure injection for the tests of the allocation error paths: with the
// environment variable ASTR_HIP_TEST_FAIL_ALLOC set to 1, the lazily grown
// device buffers report a failed allocation instead of allocating.
bool test_fail_alloc() {
  const char* v = std::getenv("ASTR_HIP_TEST_FAIL_ALLOC");
  return v && v[0] == '1';
}

#define HIP_OK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(std::string(#x) + ": " + hipGetErrorString(e_));                 \
  } while (0)

// Per-tile partial sums left by the last gradient pass: rho|u|^2, rho|w|^2.
constexpr int kTileStats = 2;
// Statistics per step of a single case: KE, enstrophy, dissipation rate.
constexpr int kStepStats = 3;
// Statistics per step of a batched case: KE, enstrophy.
constexpr int kBatchStats = 2;

// Sum the per-tile stats in a fixed order (one block of 256 threads) into
// KE = sum rho|u|^2 / 2N and enstrophy = sum rho|w|^2 / 2N (stacal,
// tgvsolver.F90:930-932), N = im*jm*km, scale = 1/N.
__device__ __forceinline__ void stats_sum(const double* __restrict__ part, int nt,
                                          double* __restrict__ out, double scale) {
  __shared__ double sv[kTileStats][256];
  double x[kTileStats] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nt; i += 256)
#pragma unroll
    for (int q = 0; q < kTileStats; q++) x[q] += part[kTileStats * i + q];
#pragma unroll
  for (int q = 0; q < kTileStats; q++) sv[q][threadIdx.x] = x[q];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int q = 0; q < kTileStats; q++) sv[q][threadIdx.x] += sv[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * sv[0][0] * scale;
    out[1] = 0.5 * sv[1][0] * scale;
  }
}

// fixed-order sum of 256 per-thread values of a block into v[0]
__device__ __forceinline__ double block_sum256(double x) {
  __shared__ double sv[256];
  sv[threadIdx.x] = x;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sv[threadIdx.x] += sv[threadIdx.x + h];
    __syncthreads();
  }
  return sv[0];
}

// Dissipation rate (diss_rate_cal, statistic.F90:994-1043): partial sums of
// 2 mu (s11^2+s22^2+s33^2 + 2(s12^2+s13^2+s23^2) - div^2/3), mu =
// miucal(T)/Re, over the points i,j,k = 1..im,jm,km (k = 0 in 2-D), from the
// stored velocity gradients dvel(m,d) and the temperature of q.  Block b
// sums the points b*256+t, b*256+t+256*nb, ... per thread, then in a fixed
// tree; k_diss_sum adds the nb block sums in a fixed tree.
constexpr int kDissBlocks = 1024;

// one point's term, in the reference's operation order (statistic.F90:1015-1031)
__device__ __forceinline__ double diss_point(const FlowArgs& a, double t, const double (&g)[3][3]) {
#pragma clang fp contract(off)
  const double miu = miucal(a, t) / a.reynolds;
  const double s11 = g[0][0], s22 = g[1][1], s33 = g[2][2];
  const double s12 = 0.5 * (g[0][1] + g[1][0]);
  const double s13 = 0.5 * (g[0][2] + g[2][0]);
  const double s23 = 0.5 * (g[1][2] + g[2][1]);
  const double div = s11 + s22 + s33;
  return 2.0 * miu *
         (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23) -
          (1.0 / 3.0) * (div * div));
}

__global__ __launch_bounds__(256) void k_diss_part(FlowArgs a, double* __restrict__ part) {
  const int im = a.n[0], jm = a.n[1], km = a.n[2];
  const long long npts = (long long)im * jm * (km > 0 ? km : 1);
  const long long fs = a.fs;
  double acc = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < npts; e += (long long)gridDim.x * 256) {
    const int i = 1 + (int)(e % im);
    const long long r = e / im;
    const int j = 1 + (int)(r % jm);
    const int k = km > 0 ? 1 + (int)(r / jm) : 0;
    const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
    // the main path's diss_rate_cal reads the stored tmp (statistic.F90:1015)
    const double t = a.main_sem ? a.pr[5 * fs + o]
                                : prim_of(a, a.q[o], a.q[fs + o], a.q[2 * fs + o], a.q[3 * fs + o],
                                          a.q[4 * fs + o]).t;
    double g[3][3];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) g[m][d] = d < a.ndir ? a.dv[(m + 3 * d) * fs + o] : 0.0;
    acc += diss_point(a, t, g);
  }
  const double b = block_sum256(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ __launch_bounds__(256) void k_diss_sum(const double* __restrict__ part, int nb,
                                                  double* __restrict__ out, double scale) {
  double x = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) x += part[i];
  const double b = block_sum256(x);
  if (threadIdx.x == 0) out[0] = b * scale;
}

__global__ __launch_bounds__(256) void k_stats_reduce(const double* __restrict__ part, int nt,
                                                      double* __restrict__ out, double scale) {
  stats_sum(part, nt, out, scale);
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

// Isentropic vortex (2-D, build-defined; no reference counterpart): free
// stream rho=1, T=1, velocity (u0,v0); vortex of strength beta centred at
// (xc,yc) with unit radius:
//   du = -beta/(2 pi) (y-yc) e,  dv = beta/(2 pi) (x-xc) e,  e = exp((1-r^2)/2)
//   T  = 1 - (gamma-1) Ma^2 beta^2/(8 pi^2) e^2   (cyclostrophic balance)
//   rho = T^(1/(gamma-1)),  p = rho T/(gamma Ma^2)  (isentropic)
// evaluated at i mod im, j mod jm so the duplicate planes equal planes 0.
__global__ void k_init_vortex(FlowArgs a, double beta, double xc, double yc, double u0, double v0,
                              double lx, double ly, double gm1, double ma2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > a.n[0]) return;
  const double pi = 3.141592653589793238462643383279502884;
  const double x = lx * (double)(i % a.n[0]) / (double)a.n[0];
  const double y = ly * (double)(j % a.n[1]) / (double)a.n[1];
  const double dx = x - xc, dy = y - yc;
  const double e = exp(0.5 * (1.0 - (dx * dx + dy * dy)));
  const double u = u0 - beta / (2.0 * pi) * dy * e;
  const double v = v0 + beta / (2.0 * pi) * dx * e;
  const double t = 1.0 - gm1 * ma2 * beta * beta / (8.0 * pi * pi) * e * e;
  const double rho = pow(t, 1.0 / gm1);
  const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
  {
#pragma clang fp contract(off)
    const double p = rho * t / a.const2;
    a.q[o] = rho;
    a.q[a.fs + o] = rho * u;
    a.q[2 * a.fs + o] = rho * v;
    a.q[3 * a.fs + o] = 0.0;
    a.q[4 * a.fs + o] = p * a.const6 + rho * 0.5 * (u * u + v * v);
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
// explicit halos f(-5:n+5) (the reference's per-line interface); BND with
// the boundary types bnd (bit 0: physical boundary at 0, bit 1: at n)
// ---------------------------------------------------------------------------
template <class C, bool FILT, bool BND = false>
__global__ __launch_bounds__(C::NT) void k_lines(const double* __restrict__ f, double* __restrict__ out,
                                                 int n, long long nlines, const double* __restrict__ ac,
                                                 FilterCoef fc, int bnd) {
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
  ChunkCoef<C::R, FILT> CC;
  load_chunk_coef<C::R, FILT>(CC, ac, n + (FILT ? 7 : 3), p * C::R);
  __syncthreads();
  solve_line<C::P, C::R, FILT, BND>(T + l * C::LT, X + l * C::LT, n, p, CC, fc, bnd);
  __syncthreads();
  for (int e = threadIdx.x; e < C::TL * (n + 1); e += C::NT) {
    const int l2 = e / (n + 1), r = e - l2 * (n + 1);
    if (L0 + l2 < nlines) out[(L0 + l2) * (n + 1) + r] = X[l2 * C::LT + r];
  }
}

// ---------------------------------------------------------------------------
// launch-shape table
// ---------------------------------------------------------------------------
struct ShapeInfo {
  int id, NT, TL, P, R, NPT;
};
const ShapeInfo kShapes[kNumShapes] = {
    {0, 256, 4, 64, 5, 5}, {1, 256, 4, 64, 9, 9}, {2, 256, 16, 16, 5, 5},
    {3, 256, 8, 32, 9, 9}, {4, 256, 8, 32, 5, 5}, {5, 512, 8, 64, 5, 5},
    {6, 128, 4, 32, 9, 9}, {7, 128, 2, 64, 5, 5}, {8, 64, 1, 64, 5, 5}, {9, 64, 2, 32, 9, 9},
    {10, 256, 16, 16, 9, 9}, {11, 512, 8, 64, 9, 9},
    {12, 64, 1, 64, 9, 9},
};

bool shape_fits(const ShapeInfo& s, int n) {
  return s.P * s.R >= n + 7 && s.NT * s.NPT >= s.TL * (n + 1) && n >= 8;
}

// LDS bytes of a pass in a shape (lds_doubles in flow_dev.hpp)
size_t pass_lds_bytes(int shape, int prog) {
  const ShapeInfo& s = kShapes[shape];
  const long long tl = (long long)s.TL * line_stride(s.P, s.R, s.TL);
  const long long one = (2 * tl > 2 * s.NT ? 2 * tl : 2 * s.NT) + kLineOff + 8 + kHaloSlot +
                        coef_regions(prog) * 2 * s.P * s.R;
  return (size_t)one * sizeof(double);
}

// per-workgroup LDS of a device, read once
size_t lds_limit(int device) {
  static std::mutex mu;
  static size_t cache[64] = {0};
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0 || device >= 64) return 65536;
  if (!cache[device]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
      v = 65536;
    cache[device] = (size_t)v;
  }
  return cache[device];
}

// a selection a pass of line length n can launch with
bool selection_ok(int shape, int prog, int n, int device) {
  return shape_fits(kShapes[shape], n) && pass_lds_bytes(shape, prog) <= lds_limit(device);
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

hipError_t launch_pass(int shape, int prog, int minb, const FlowArgs& a, int dir, hipStream_t st) {
  switch (shape) {
    case 0: return launch_pass_s0(prog, minb, a, dir, st);
    case 1: return launch_pass_s1(prog, minb, a, dir, st);
    case 2: return launch_pass_s2(prog, minb, a, dir, st);
    case 3: return launch_pass_s3(prog, minb, a, dir, st);
    case 4: return launch_pass_s4(prog, minb, a, dir, st);
    case 5: return launch_pass_s5(prog, minb, a, dir, st);
    case 6: return launch_pass_s6(prog, minb, a, dir, st);
    case 7: return launch_pass_s7(prog, minb, a, dir, st);
    case 8: return launch_pass_s8(prog, minb, a, dir, st);
    case 9: return launch_pass_s9(prog, minb, a, dir, st);
    case 10: return launch_pass_s10(prog, minb, a, dir, st);
    case 11: return launch_pass_s11(prog, minb, a, dir, st);
    default: return launch_pass_s12(prog, minb, a, dir, st);
  }
}

// register budget a pass uses until tuned: two workgroups per CU for the
// gradient/convective passes, one for the others
int default_minb(int prog) { return prog <= GC_LAST || prog >= GL_GRAD ? 2 : 1; }

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct astr_hip_ctx {
  astr_hip_config cfg;
  FlowArgs a;
  // launch selection per (pass program, direction): shape index and register
  // budget MINB; ms = the autotuner's measured time per launch (0: untuned)
  int shape[kNumProgs][3];
  int minb[kNumProgs][3];
  float tuned_ms[kNumProgs][3];
  // the last direction's gradient/convective work in two kernels
  // (GL_GRAD + GL_CONV) instead of GC_LAST; chosen by the autotuner
  int split_last = 0;
  // rk3_run: each stage's x filter sweep runs in the next stage's first pass
  // (FGC_FIRST) inside one call (fuse_fx, off unless the autotuner finds the
  // fused pass faster than the two it replaces); fx_pending says that the
  // sweep of the stage just ended is still owed
  int fuse_fx = 0;
  bool fx_pending = false;
  // main path: each stage's RK pass also applies the next stage's filterq
  // sweep along its direction (the last one) inside one rk3_steps call
  int fuse_main_filter = 1;
  double* base = nullptr;   // one allocation for all fields
  double* coef = nullptr;   // LU tables
  double* stats_part = nullptr;
  double* diss_part = nullptr;   // inside stats_part's allocation
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
  // device address of this device's copy of `scratch` (nullptr when the
  // lookup failed; scratch_err then holds the reason)
  double* scratch = nullptr;
  std::string scratch_err;
  // main path: the halo sources qswap keeps (hx then hy, one allocation),
  // and whether they belong to the current q (cleared by every change of q
  // outside qswap)
  double* halo = nullptr;
  bool halo_valid = false;
  // main path on a grid that is not uniform: J and dxi(d,n) per point (10 fields)
  double* met = nullptr;
  // main path, upwind-biased convection (a.uwd): its state and one
  // allocation for the split buffer, the sensor and the shock flags
  UwArgs uw{};
  double* uwbuf = nullptr;
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

// boundary bits of a reference ntype (1: physical boundary at 0, 2: at n,
// 3: interfaces, 4: both physical); -1 for another value
int ntype_bits(int ntype) {
  switch (ntype) {
    case 1: return 1;
    case 2: return 2;
    case 3: return 0;
    case 4: return 3;
    default: return -1;
  }
}

// derivative.F90:106-136 with physical ends, in the periodic form's rows
// -1..n+1: a physical end's row beyond the boundary is decoupled (a = c = 0)
// and its boundary rows take a = c = 2 (row 0) and 1/4 (row 1)
void deriv_tables_nt(int n, int bits, std::vector<double>& out) {
  const int m = n + 3;
  std::vector<double> a(m, 1.0 / 3.0), c(m, 1.0 / 3.0);
  a[0] = c[0] = 0.0;
  a[m - 1] = c[m - 1] = 0.0;
  if (bits & 1) {
    a[1] = c[1] = 2.0;
    a[2] = c[2] = 0.25;
  }
  if (bits & 2) {
    a[m - 2] = c[m - 2] = 2.0;
    a[m - 3] = c[m - 3] = 0.25;
  }
  thomas_prep(a, c, out);
}

// filter.F90:31-100 with physical ends (row 0 / n take 0.98, :47-66), in the
// periodic form's rows -3..n+3; the three rows beyond a physical end are
// decoupled
void filter_tables_nt(int n, int bits, double alfa, double bh, std::vector<double>& out) {
  const int m = n + 7;
  std::vector<double> a(m, alfa), c(m, alfa);
  a[0] = c[0] = bh;
  a[m - 1] = c[m - 1] = bh;
  if (bits & 1) {
    for (int u = 0; u < 3; u++) a[u] = c[u] = 0.0;
    a[3] = c[3] = 0.98;
  }
  if (bits & 2) {
    for (int u = m - 3; u < m; u++) a[u] = c[u] = 0.0;
    a[m - 4] = c[m - 4] = 0.98;
  }
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

// filter_coefficient_cal (filter.F90:299-432): the rows a physical boundary
// uses, coefb(0:2,0:6) row-major, coef6i(0:3), coef8i(0:4), into b30
void filter_coefs_b(double alfa, double bb, double* b30) {
  const double b[30] = {
      (63.0 + 1.0 * bb) / 64.0, (3.0 + 29.0 * bb) / 32.0, (-15.0 + 15.0 * bb) / 64.0,
      (5.0 - 5.0 * bb) / 16.0, (-15.0 + 15.0 * bb) / 64.0, (3.0 - 3.0 * bb) / 32.0,
      (-1.0 + 1.0 * bb) / 64.0,
      (1.0 + 62.0 * alfa) / 64.0, (29.0 + 6.0 * alfa) / 32.0, (15.0 + 34.0 * alfa) / 64.0,
      (-5.0 + 10.0 * alfa) / 16.0, (15.0 - 30.0 * alfa) / 64.0, (-3.0 + 6.0 * alfa) / 32.0,
      (1.0 - 2.0 * alfa) / 64.0,
      (-1.0 + 2.0 * alfa) / 64.0, (3.0 + 26.0 * alfa) / 32.0, (49.0 + 30.0 * alfa) / 64.0,
      (5.0 + 6.0 * alfa) / 16.0, (-15.0 + 30.0 * alfa) / 64.0, (3.0 - 6.0 * alfa) / 32.0,
      (-1.0 + 2.0 * alfa) / 64.0,
      (11.0 + 10.0 * alfa) / 32.0, (15.0 + 34.0 * alfa) / 64.0, (-3.0 + 6.0 * alfa) / 32.0,
      (1.0 - 2.0 * alfa) / 64.0,
      (93.0 + 70.0 * alfa) / 256.0, (7.0 + 18.0 * alfa) / 32.0, (-7.0 + 14.0 * alfa) / 64.0,
      (1.0 - 2.0 * alfa) / 32.0, (-1.0 + 2.0 * alfa) / 256.0};
  std::memcpy(b30, b, sizeof(b));
}

// k_lines on stream st with the launch shape sh: op 0 derivative, 1 filter;
// bits the boundary bits (0: the periodic form's kernel)
hipError_t launch_lines(int sh, int op, int bits, const double* f, double* out, int n, long long nlines,
                        const double* lu, const FilterCoef& fc, hipStream_t st) {
  const ShapeInfo& s = kShapes[sh];
  const long long nb = (nlines + s.TL - 1) / s.TL;
  const size_t lds = (size_t)2 * s.TL * line_stride(s.P, s.R, s.TL) * sizeof(double);
#define LB(Cf)                                                                                          \
  if (op == 0 && !bits)                                                                                 \
    hipLaunchKernelGGL((k_lines<Cf, false>), dim3(nb), dim3(Cf::NT), lds, st, f, out, n, nlines, lu, fc, 0); \
  else if (op == 0)                                                                                     \
    hipLaunchKernelGGL((k_lines<Cf, false, true>), dim3(nb), dim3(Cf::NT), lds, st, f, out, n, nlines, lu, fc, bits); \
  else if (!bits)                                                                                       \
    hipLaunchKernelGGL((k_lines<Cf, true>), dim3(nb), dim3(Cf::NT), lds, st, f, out, n, nlines, lu, fc, 0); \
  else                                                                                                  \
    hipLaunchKernelGGL((k_lines<Cf, true, true>), dim3(nb), dim3(Cf::NT), lds, st, f, out, n, nlines, lu, fc, bits);
  switch (sh) {
    case 0: LB(CfgA); break;
    case 1: LB(CfgB); break;
    case 2: LB(CfgC); break;
    case 3: LB(CfgD); break;
    case 4: LB(CfgE); break;
    case 5: LB(CfgF); break;
    case 6: LB(CfgG); break;
    case 7: LB(CfgH); break;
    case 8: LB(CfgI); break;
    case 9: LB(CfgJ); break;
    case 10: LB(CfgK); break;
    case 11: LB(CfgL); break;
    default: LB(CfgM); break;
  }
#undef LB
  return hipGetLastError();
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
// one launch (or launch sequence) timed into profile slot `slot` when the
// context is profiling
template <class F>
int timed(astr_hip_ctx* c, int slot, F&& launch) {
  const bool rec = c->prof && c->ev_used < (int)c->ev_kind.size();
  if (rec) HIP_OK(hipEventRecord(c->ev[2 * c->ev_used], c->stream));
  HIP_OK(launch());
  if (rec) {
    HIP_OK(hipEventRecord(c->ev[2 * c->ev_used + 1], c->stream));
    c->ev_kind[c->ev_used++] = slot;
  }
  return 0;
}
constexpr int kSlotUwSplit = 27, kSlotUwFace = 30, kSlotUwSensor = 33, kSlotFgcFirst = 36;
static_assert(kSlotFgcFirst + 3 == ASTR_PROF_SLOTS, "profile slots");

// shape >= 0 overrides the pass's launch shape
int pass(astr_hip_ctx* c, int prog, int d, int shape = -1) {
  const bool rec = c->prof && c->ev_used < (int)c->ev_kind.size();
  if (rec) HIP_OK(hipEventRecord(c->ev[2 * c->ev_used], c->stream));
  HIP_OK(launch_pass(shape >= 0 ? shape : c->shape[prog][d], prog, c->minb[prog][d], c->a, d, c->stream));
  if (rec) {
    HIP_OK(hipEventRecord(c->ev[2 * c->ev_used + 1], c->stream));
    c->ev_kind[c->ev_used++] = prog < FGC_FIRST ? prog * 3 + d : kSlotFgcFirst + d;
  }
  return 0;
}

// stats: 0 none, 1 KE/enstrophy, 2 also the dissipation rate
// diss_out (device, nullable): the dissipation rate of the current state,
// evaluated after the gradient passes (which then store every gradient) and
// before the viscous passes update q
int run_rhs(astr_hip_ctx* c, bool rk, bool stats, bool store_last_grad, double* diss_out = nullptr) {
  FlowArgs& a = c->a;
  const int nd = a.ndir;
  a.dostats = stats ? 1 : 0;
  // the shock sensor reads the whole velocity gradient
  const bool sensor = a.uwd && c->uw.sensor;
  a.store_last_grad = store_last_grad || diss_out || sensor ? 1 : 0;
  for (int d = 0; d < nd; d++) {
    if (d == nd - 1 && c->split_last) {
      if (pass(c, GL_GRAD, d) || (!a.uwd && pass(c, GL_CONV, d))) return -1;
      continue;
    }
    const int prog = d == nd - 1 ? GC_LAST : (d == 0 ? (c->fx_pending ? FGC_FIRST : GC_FIRST) : GC_MID);
    if (pass(c, prog, d)) return -1;
    if (d == 0) c->fx_pending = false;
  }
  if (a.uwd) {
    // rhscal's upwind branches (solver.F90:219-233): ducrossensor when
    // recon_schem = 5 or lchardecomp (explicit) or lchardecomp (compact),
    // then convrsduwd or convrsdcmp along i, j, k
    if (sensor && timed(c, kSlotUwSensor, [&] { return uw_launch_sensor(a, c->uw, c->stream); })) return -1;
    for (int d = 0; d < nd; d++) {
      if (timed(c, kSlotUwSplit + d, [&] { return uw_launch_split(a, c->uw, d, c->stream); })) return -1;
      if (timed(c, kSlotUwFace + d, [&] { return uw_launch_face(a, c->uw, d, d == 0, d == nd - 1, c->stream); }))
        return -1;
    }
  }
  if (diss_out) {
    const long long npts = (long long)a.n[0] * a.n[1] * (a.n[2] > 0 ? a.n[2] : 1);
    const int nb = (int)std::min<long long>(kDissBlocks, (npts + 255) / 256);
    hipLaunchKernelGGL(k_diss_part, dim3(nb), dim3(256), 0, c->stream, a, c->diss_part);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(k_diss_sum, dim3(1), dim3(256), 0, c->stream, c->diss_part, nb, diss_out, c->ke_scale);
    HIP_OK(hipGetLastError());
  }
  a.dostats = 0;
  for (int d = 0; d < nd; d++) {
    const int prog = d == nd - 1 ? (rk ? V_LAST_RK : V_LAST_RHS) : V_MID;
    if (pass(c, prog, d)) return -1;
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// qswap's single-rank periodic branch (parallel.F90:4876-4884, 5018-5030,
// 5159-5170).  Its halo copies are the passes' index wrap; what it adds is
// the duplicate-plane average q(0) = q(n) = 0.5 (q(0) + q(n)) of direction
// dir, over all points of the two other directions (i, then j, then k).
// ---------------------------------------------------------------------------
namespace {
__global__ void k_qswap_dir(FlowArgs a, int dir) {
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  const long long n1 = a.n[d1] + 1, n2 = d2 < 3 ? (long long)a.n[d2] + 1 : 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 5 * n1 * n2) return;
  const int m = (int)(idx / (n1 * n2));
  const long long r = idx - m * n1 * n2;
  const long long st[3] = {1, a.pitch, a.plane};
  const long long o0 = (r % n1) * st[d1] + (r / n1) * st[d2];
  const long long on = o0 + (long long)a.n[dir] * st[dir];
  double* q = a.q + m * a.fs;
  const double v = 0.5 * (q[o0] + q[on]);
  q[o0] = v;
  q[on] = v;
}

// Main path: the halo sources of the lines whose halo qswap fills before a
// later step averages the source points (dir 0: the x-lines on the j and k
// boundary planes, sources i = 1..5, im-5..im-1; dir 1: the y-lines on the k
// boundary planes, sources j = 1..5, jm-5..jm-1), in halo_src_index's order.
__global__ void k_halo_src(FlowArgs a, int dir, double* __restrict__ dst, long long dst_stride) {
  const long long nl = dir == 0 ? hx_lines(a.n, a.ndir) : hy_lines(a.n, a.ndir);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nl * kHaloSrcRows) return;
  const int line = (int)(idx / kHaloSrcRows), hrow = (int)(idx % kHaloSrcRows);
  const int n = a.n[dir];
  const int r = hrow < 5 ? hrow + 1 : n - 5 + (hrow - 5);
  int i, j, k;
  if (dir == 0) {
    i = r;
    const int nA = 2 * (a.n[2] + 1);
    if (line < nA) {
      j = line / (a.n[2] + 1) == 0 ? 0 : a.n[1];
      k = line % (a.n[2] + 1);
    } else {
      const int l2 = line - nA;
      k = l2 / (a.n[1] + 1) == 0 ? 0 : a.n[2];
      j = l2 % (a.n[1] + 1);
    }
  } else {
    j = r;
    k = line / (a.n[0] + 1) == 0 ? 0 : a.n[2];
    i = line % (a.n[0] + 1);
  }
  const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
#pragma unroll
  for (int m = 0; m < 5; m++) dst[m * dst_stride + idx] = a.q[m * a.fs + o];
}

// qswap's q2fvar of the boundary slabs (parallel.F90:4893-4899, 5048-5054,
// 5189-5195) for the computed points: rho, u, v, w, p, T of the points on the
// planes dir = 0 and dir = n[dir] from the averaged q
__global__ void k_bprims(FlowArgs a, int dir) {
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  const long long n1 = a.n[d1] + 1, n2 = (long long)a.n[d2] + 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * n1 * n2) return;
  const int side = (int)(idx / (n1 * n2));
  const long long r = idx - side * n1 * n2;
  const long long st[3] = {1, a.pitch, a.plane};
  const long long o = (r % n1) * st[d1] + (r / n1) * st[d2] + (side ? (long long)a.n[dir] * st[dir] : 0);
  const long long fs = a.fs;
  const Prim P = prim_of(a, a.q[o], a.q[fs + o], a.q[2 * fs + o], a.q[3 * fs + o], a.q[4 * fs + o]);
  a.pr[o] = P.r;
  a.pr[fs + o] = P.u;
  a.pr[2 * fs + o] = P.v;
  a.pr[3 * fs + o] = P.w;
  a.pr[4 * fs + o] = P.p;
  a.pr[5 * fs + o] = P.t;
}

// updatefvar (fludyna.F90:191-242): rho, u, v, w, p, T of q at every computed
// point into the stored primitives
__global__ void k_updatefvar(FlowArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > a.n[0]) return;
  const long long o = (long long)k * a.plane + (long long)j * a.pitch + i;
  const long long fs = a.fs;
  const Prim P = prim_of(a, a.q[o], a.q[fs + o], a.q[2 * fs + o], a.q[3 * fs + o], a.q[4 * fs + o]);
  a.pr[o] = P.r;
  a.pr[fs + o] = P.u;
  a.pr[2 * fs + o] = P.v;
  a.pr[3 * fs + o] = P.w;
  a.pr[4 * fs + o] = P.p;
  a.pr[5 * fs + o] = P.t;
}
}  // namespace

// ---------------------------------------------------------------------------
// grid metrics (gridgeom, geom.F90:99-692) from the grid coordinates, on the
// device with the line engine: dx(m,d) = d x_m / d xi_d by the compact
// derivative (:128-161), the Jacobian (:350-381), the symmetric conservative
// form of J d xi_r / d x_m (:395-540) and its division by J (:664-668).
// Fields are dense (im+1)(jm+1)(km+1), i fastest; dx and dxi field (a-1) +
// 3(b-1) holds the reference's (a,b).
// ---------------------------------------------------------------------------
namespace {

struct GeomGrid {
  int n[3];
  long long st[3];
  long long fs;
};

// coordinate m at index t (-5..n+5) along dir from the point base: the
// single-rank periodic extension of gridsendrecv (parallel.F90:2859-2864),
// x(n+h) = x(n) + (x(h) - x(0)), x(-h) = x(0) - (x(n) - x(n-h))
__device__ __forceinline__ double geom_x(const double* __restrict__ X, const GeomGrid& G, int m, int dir,
                                         long long base, int t) {
  const double* f = X + m * G.fs + base;
  const int n = G.n[dir];
  const long long s = G.st[dir];
  if (t < 0) return f[0] - (f[n * s] - f[(n + t) * s]);
  if (t > n) return f[n * s] + (f[(t - n) * s] - f[0]);
  return f[t * s];
}

// dx(a,e) at index t along dir: the periodic halo copy of dataswap
__device__ __forceinline__ double geom_dx(const double* __restrict__ D, const GeomGrid& G, int a, int e,
                                          int dir, long long base, int t) {
  const int n = G.n[dir];
  const int tt = t < 0 ? n + t : (t > n ? t - n : t);
  return D[(a + 3 * e) * G.fs + base + tt * G.st[dir]];
}

// what a line holds: kind 0 the coordinate m; kind 1 the product
// phi = 0.5 (dx(a,e) x_b - dx(b,e) x_a) (0-based indices)
struct GeomLine {
  int kind, m, a, e, b;
};

__device__ __forceinline__ long long geom_line_base(const GeomGrid& G, int dir, long long l) {
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  return (l % (G.n[d1] + 1)) * G.st[d1] + (l / (G.n[d1] + 1)) * G.st[d2];
}

__global__ void k_geom_gather(const double* __restrict__ X, const double* __restrict__ D, GeomGrid G,
                              int dir, GeomLine L, long long nlines, double* __restrict__ lines) {
#pragma clang fp contract(off)
  const int n = G.n[dir], W = n + 11;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlines * W) return;
  const long long l = idx / W;
  const int t = (int)(idx - l * W) - 5;
  const long long base = geom_line_base(G, dir, l);
  double v;
  if (L.kind == 0) {
    v = geom_x(X, G, L.m, dir, base, t);
  } else {
    v = 0.5 * (geom_dx(D, G, L.a, L.e, dir, base, t) * geom_x(X, G, L.b, dir, base, t) -
               geom_dx(D, G, L.b, L.e, dir, base, t) * geom_x(X, G, L.a, dir, base, t));
  }
  lines[idx] = v;
}

__global__ void k_geom_scatter(const double* __restrict__ out, GeomGrid G, int dir, long long nlines,
                               double* __restrict__ dst, int accumulate) {
  const int n = G.n[dir];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlines * (n + 1)) return;
  const long long l = idx / (n + 1);
  const int t = (int)(idx - l * (n + 1));
  const long long o = geom_line_base(G, dir, l) + t * G.st[dir];
  dst[o] = accumulate ? dst[o] + out[idx] : out[idx];
}

// Jacobian, the 2-D metrics, and dxi / J
__global__ void k_geom_final(const double* __restrict__ D, double* __restrict__ XI, double* __restrict__ J,
                             GeomGrid G, int ndims) {
#pragma clang fp contract(off)
  const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= G.fs) return;
  const long long fs = G.fs;
#define DX(a, b) D[((a) - 1 + 3 * ((b) - 1)) * fs + o]
#define XI_(a, b) XI[((a) - 1 + 3 * ((b) - 1)) * fs + o]
  double jac;
  if (ndims == 3) {
    jac = DX(1, 1) * DX(2, 2) * DX(3, 3) + DX(1, 2) * DX(2, 3) * DX(3, 1) + DX(1, 3) * DX(2, 1) * DX(3, 2) -
          DX(1, 3) * DX(2, 2) * DX(3, 1) - DX(1, 2) * DX(2, 1) * DX(3, 3) - DX(1, 1) * DX(2, 3) * DX(3, 2);
  } else {
    jac = DX(1, 1) * DX(2, 2) - DX(1, 2) * DX(2, 1);
    const double x11 = DX(1, 1), x12 = DX(1, 2), x21 = DX(2, 1), x22 = DX(2, 2);
    for (int b = 1; b <= 3; b++)
      for (int a = 1; a <= 3; a++) XI_(a, b) = 0.0;
    XI_(1, 1) = x22;
    XI_(1, 2) = -x12;
    XI_(2, 1) = -x21;
    XI_(2, 2) = x11;
  }
  J[o] = jac;
  for (int b = 1; b <= 3; b++)
    for (int a = 1; a <= 3; a++) XI_(a, b) = XI_(a, b) / jac;
#undef DX
#undef XI_
}

// the 18 conservative terms: J dxi_r/dx_m += D_d [0.5 (dx(a,e) x_b - dx(b,e) x_a)],
// as (d, r, m, a, e, b), 1-based, in the reference's order (geom.F90:395-515)
const int kGeomTerms[18][6] = {
    {1, 2, 1, 2, 3, 3}, {1, 2, 2, 3, 3, 1}, {1, 2, 3, 1, 3, 2},
    {1, 3, 1, 3, 2, 2}, {1, 3, 2, 1, 2, 3}, {1, 3, 3, 2, 2, 1},
    {2, 1, 1, 3, 3, 2}, {2, 1, 2, 1, 3, 3}, {2, 1, 3, 2, 3, 1},
    {2, 3, 1, 2, 1, 3}, {2, 3, 2, 3, 1, 1}, {2, 3, 3, 1, 1, 2},
    {3, 1, 1, 2, 2, 3}, {3, 1, 2, 3, 2, 1}, {3, 1, 3, 1, 2, 2},
    {3, 2, 1, 3, 1, 2}, {3, 2, 2, 1, 1, 3}, {3, 2, 3, 2, 1, 1},
};

// derivative lines of length n already in device memory (f: n+11 per line,
// out: n+1 per line) on stream st, with the LU table lu of n and the
// boundary bits
hipError_t deriv_lines_device(int n, long long nlines, const double* f, double* out, const double* lu,
                              int bits, hipStream_t st) {
  const int sh = pick_shape(n);
  if (sh < 0) return hipErrorInvalidValue;
  FilterCoef fc{};
  return launch_lines(sh, 0, bits, f, out, n, nlines, lu, fc, st);
}

}  // namespace

// Statistics buffers of astr_hip_rk3_steps_stats_noalloc, for callers that
// cannot allocate device memory: the per-tile partial sums, the dissipation
// rate's block partials and the per-step results, packed in this order.
// Each device holds its own copy (8 MiB) until the process ends.  Host code
// never passes the array's name to a kernel: the name is a host-side stand-in,
// and kernels take the device address that hipGetSymbolAddress returns.
constexpr long long kScratchDoubles = 1ll << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device; nullptr with the HIP error in *err when the lookup fails.
double* scratch_address(int device, std::string* err) {
  static std::mutex mu;
  static double* addr[64] = {nullptr};
  if (device < 0 || device >= 64) {
    *err = "device index out of range for the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    addr[device] = static_cast<double*>(p);
  }
  return addr[device];
}

// Device buffers of the per-step statistics: the per-tile partial sums the
// last gradient pass leaves (2 per tile), the dissipation rate's block
// partials (kDissBlocks) and the per-step results (kStepStats per step).
struct StatsBufs {
  double* part;
  double* diss;
  double* out;
};

// The coefficient table of a context: the filter's halo rows and the rows of
// a physical boundary (kBndCoefs), then per direction the derivative's and
// the filter's LU factors, of the periodic form or (a.bnd[d] != 0) of the
// closed form; (re)allocated on the device and pointed at by a.fc, dac, fac
int build_tables(astr_hip_ctx* c) {
  FlowArgs& a = c->a;
  const astr_hip_config& cfg = c->cfg;
  FilterCoef fc;
  std::vector<double> all(kBndCoefs);
  filter_coefs(cfg.alfa_filter, cfg.beta_halo, fc, all.data());
  filter_coefs_b(cfg.alfa_filter, cfg.beta_bound, all.data() + 33);
  size_t off[6] = {0, 0, 0, 0, 0, 0};
  for (int d = 0; d < a.ndir; d++) {
    std::vector<double> t;
    if (a.bnd[d]) deriv_tables_nt(a.n[d], a.bnd[d], t);
    else deriv_tables(a.n[d], t);
    off[d] = all.size();
    all.insert(all.end(), t.begin(), t.end());
    if (a.bnd[d]) filter_tables_nt(a.n[d], a.bnd[d], cfg.alfa_filter, cfg.beta_halo, t);
    else filter_tables(a.n[d], cfg.alfa_filter, cfg.beta_halo, t);
    off[3 + d] = all.size();
    all.insert(all.end(), t.begin(), t.end());
  }
  double* tab = nullptr;
  if (hipMalloc(&tab, all.size() * sizeof(double)) != hipSuccess) return fail("hipMalloc of tables failed");
  if (hipMemcpy(tab, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(tab);
    return fail("copy of the tables failed");
  }
  if (c->coef) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->coef);
  }
  c->coef = tab;
  std::memcpy(a.fc.c10, fc.c10, sizeof(fc.c10));
  a.fc.ch = c->coef;
  for (int d = 0; d < a.ndir; d++) {
    a.dac[d] = c->coef + off[d];
    a.fac[d] = c->coef + off[3 + d];
  }
  return 0;
}

// updatefvar into the stored primitives (main path)
int updatefvar_dev(astr_hip_ctx* c) {
  hipLaunchKernelGGL(k_updatefvar, pt_grid(c->a), dim3(64), 0, c->stream, c->a);
  HIP_OK(hipGetLastError());
  c->halo_valid = false;
  return 0;
}

// Main path qswap, single rank, periodic (parallel.F90:4876-4920,
// 5018-5060, 5159-5200), for the device layout: the halo copies are the
// passes' index wrap, except for the lines whose halo the reference copies
// before a later step averages its source points, whose sources are kept
// here first; then the duplicate-plane averages along i, j, k and the
// primitives of the boundary planes.  average = false keeps the halo
// sources of the current q only (the passes then see the halo the
// reference's dataswap of q gives).
int qswap_main(astr_hip_ctx* c, bool average) {
  const FlowArgs& a = c->a;
  auto avg = [&](int d) -> int {
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    const long long cnt = 5LL * (a.n[d1] + 1) * (a.n[d2] + 1);
    hipLaunchKernelGGL(k_qswap_dir, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, a, d);
    HIP_OK(hipGetLastError());
    return 0;
  };
  const long long nx = hx_lines(a.n, a.ndir) * kHaloSrcRows;
  hipLaunchKernelGGL(k_halo_src, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, c->stream, a, 0,
                     const_cast<double*>(a.hx), a.hxs);
  HIP_OK(hipGetLastError());
  // only periodic directions are swapped (parallel.F90:4876, 5018, 5159)
  if (average && !a.bnd[0] && avg(0)) return -1;
  if (a.ndir == 3) {
    const long long ny = hy_lines(a.n, a.ndir) * kHaloSrcRows;
    hipLaunchKernelGGL(k_halo_src, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, c->stream, a, 1,
                       const_cast<double*>(a.hy), a.hys);
    HIP_OK(hipGetLastError());
  }
  if (average) {
    for (int d = 1; d < a.ndir; d++)
      if (!a.bnd[d] && avg(d)) return -1;
    for (int d = 0; d < a.ndir; d++) {
      if (a.bnd[d]) continue;
      const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
      const long long cnt = 2LL * (a.n[d1] + 1) * (a.n[d2] + 1);
      hipLaunchKernelGGL(k_bprims, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, a, d);
      HIP_OK(hipGetLastError());
    }
    // a 2-D case still takes the k step (parallel.F90:5159-5170, 5184-5196):
    // its slabs -hm..0 and 0..hm hold the plane k = 0, so q2fvar refreshes
    // the primitives of every computed point
    if (a.ndir == 2) {
      hipLaunchKernelGGL(k_updatefvar, pt_grid(a), dim3(64), 0, c->stream, a);
      HIP_OK(hipGetLastError());
    }
  }
  c->halo_valid = true;
  return 0;
}

// the main path's right-hand side passes read the kept halo sources
int ensure_halo(astr_hip_ctx* c) {
  if (c->a.main_sem && !c->halo_valid) return qswap_main(c, false);
  return 0;
}

// steploop/rk3 (tgvsolver.F90:944-960, :994-1082) with the statistics of
// each step's first stage in the buffers sb; every statistics value is
// written before it is read, so the buffers need no clearing.  The main
// path's stage (time_integration_rk, src/mainloop.F90:396-482) is filterq,
// qswap, gradcal, [statcal], rhscal, the RK3 update and updatefvar.
int rk3_run(astr_hip_ctx* c, int nsteps, const StatsBufs& sb, double* ke, double* ens, double* diss) {
  static const double rk[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  const bool want = ke || ens || diss;
  FlowArgs& a = c->a;
  double* const stats_keep = a.stats;
  double* const diss_keep = c->diss_part;
  a.stats = sb.part;
  c->diss_part = sb.diss;
  const int stats_prog = c->split_last ? GL_GRAD : GC_LAST;
  const int ntl = tiles_of(a.n, a.ndir - 1, kShapes[c->shape[stats_prog][a.ndir - 1]].TL);
  int rc = 0;
  // main path: the last direction's filter sweep of this stage was applied by
  // the previous stage's RK pass (a.rk_filter)
  bool last_filtered = false;
  for (int it = 0; it < nsteps && !rc; it++) {
    for (int s = 0; s < 3 && !rc; s++) {
      a.stage = s;
      a.rk_a = rk[s][0]; a.rk_b = rk[s][1]; a.rk_c = rk[s][2];
      if (a.main_sem) {
        // filterq's sweeps act on different indices and are linear, so they
        // run last direction first: from the previous RK pass, or here on
        // that pass's launch shape, which makes both round alike (a call
        // split in two gives the bits of the whole)
        const int dl = a.ndir - 1;
        if (!last_filtered && pass(c, F_ONLY, dl, c->shape[V_LAST_RK][dl])) rc = -1;
        for (int d = 0; d < dl && !rc; d++)
          if (pass(c, F_ONLY, d)) rc = -1;
        if (!rc && qswap_main(c, true)) rc = -1;
        if (rc) break;
      }
      a.rk_filter = a.main_sem && c->fuse_main_filter && !(it == nsteps - 1 && s == 2) ? 1 : 0;
      const int rhs_rc = run_rhs(c, true, want && s == 0, false,
                                 diss && s == 0 ? sb.out + kStepStats * it + 2 : nullptr);
      last_filtered = a.rk_filter != 0;
      a.rk_filter = 0;
      if (rhs_rc) {
        rc = -1;
        break;
      }
      if (want && s == 0) {
        hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, c->stream, sb.part, ntl,
                           sb.out + kStepStats * it, c->ke_scale);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(std::string("k_stats_reduce: ") + hipGetErrorString(e));
      }
      // filterq's other sweeps, y then x (z went with the RK update): inside
      // the call, the x sweep of every stage but the last is left to the next
      // stage's first pass, which loads the same lines
      const int nd = a.ndir;
      if (!rc && !a.main_sem) {
        for (int d = nd - 2; d >= 1 && !rc; d--)
          if (pass(c, F_ONLY, d)) rc = -1;
        const bool defer = c->fuse_fx && !a.uwd && !(it == nsteps - 1 && s == 2);
        if (!rc) {
          if (defer) c->fx_pending = true;
          else if (pass(c, F_ONLY, 0)) rc = -1;
        }
      }
    }
  }
  c->fx_pending = false;
  // the RK passes changed q after the last qswap
  c->halo_valid = false;
  a.stats = stats_keep;
  c->diss_part = diss_keep;
  if (rc) return rc;
  if (want && nsteps > 0) {
    // strided copies of each statistic straight into the caller's arrays
    const size_t sp = kStepStats * sizeof(double);
    double* dst[3] = {ke, ens, diss};
    for (int q = 0; q < 3; q++)
      if (dst[q])
        HIP_OK(hipMemcpy2DAsync(dst[q], sizeof(double), sb.out + q, sp, sizeof(double), (size_t)nsteps,
                                hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
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
  if (cfg->semantics != ASTR_SEM_MINIAPP && cfg->semantics != ASTR_SEM_MAIN)
    return fail("unsupported semantics mode");
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
  // point offsets inside a field are 32-bit byte offsets (flow_dev.hpp ldg/stg)
  if ((unsigned long long)a.fs * sizeof(double) >= (1ull << 32)) {
    delete c;
    return fail("grid too large: a field must stay below 4 GiB");
  }
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
  a.main_sem = cfg->semantics == ASTR_SEM_MAIN ? 1 : 0;
  // the Jacobian of the uniform grid until set_metrics gives gridgeom's
  a.jac = (cfg->lx / cfg->im) * (cfg->ly / cfg->jm) * (cfg->km > 0 ? cfg->lz / cfg->km : 1.0);
  a.alfa = cfg->alfa_filter;
  a.bhalo = cfg->beta_halo;
  double h33[33];
  filter_coefs(cfg->alfa_filter, cfg->beta_halo, a.fc, h33);
  c->ke_scale = 1.0 / ((double)cfg->im * cfg->jm * (cfg->km > 0 ? cfg->km : 1));
  for (int d = 0; d < a.ndir; d++) {
    const int sh = pick_shape(a.n[d]);
    if (sh < 0) { delete c; return fail("no launch shape fits this line length"); }
    for (int pr = 0; pr < kNumProgs; pr++) {
      int use = sh;
      for (int k = 0; k < kNumShapes && !selection_ok(use, pr, a.n[d], c->device); k++)
        use = k;
      if (!selection_ok(use, pr, a.n[d], c->device)) { delete c; return fail("no launch shape fits the LDS"); }
      c->shape[pr][d] = use;
      c->minb[pr][d] = default_minb(pr);
      c->tuned_ms[pr][d] = 0.f;
    }
  }
  // fields: q 5, qsave 5, qrhs 5, dv 12, sg 9; main path: primitives 6
  const long long nfield = a.main_sem ? 42 : 36;
  size_t bytes = (size_t)(nfield * a.fs) * sizeof(double);
  if (hipMalloc(&c->base, bytes) != hipSuccess) { delete c; return fail("hipMalloc of fields failed"); }
  (void)hipMemset(c->base, 0, bytes);
  a.q = c->base;
  a.qsave = a.q + 5 * a.fs;
  a.qrhs = a.qsave + 5 * a.fs;
  a.dv = a.qrhs + 5 * a.fs;
  a.sg = a.dv + 12 * a.fs;
  a.pr = a.main_sem ? a.sg + 9 * a.fs : nullptr;
  if (a.main_sem) {
    // halo sources kept by qswap: hx then hy, 5 fields each
    a.hxs = hx_lines(a.n, a.ndir) * kHaloSrcRows;
    a.hys = hy_lines(a.n, a.ndir) * kHaloSrcRows;
    if (hipMalloc(&c->halo, (size_t)(5 * (a.hxs + a.hys) + 1) * sizeof(double)) != hipSuccess) {
      (void)hipFree(c->base); delete c; return fail("hipMalloc of the halo sources failed");
    }
    (void)hipMemset(c->halo, 0, (size_t)(5 * (a.hxs + a.hys) + 1) * sizeof(double));
    a.hx = c->halo;
    a.hy = c->halo + 5 * a.hxs;
  }
  (void)h33;
  if (build_tables(c)) {
    (void)hipFree(c->base); if (c->halo) (void)hipFree(c->halo);
    const std::string e = g_err;
    delete c;
    return fail(e);
  }
  // per-tile stats slots for the smallest tile any shape uses
  // (then the dissipation rate's block partials)
  int tl_min = 32;
  for (const ShapeInfo& s : kShapes) tl_min = std::min(tl_min, s.TL);
  const int ntl = tiles_of(a.n, a.ndir - 1, tl_min);
  if (hipMalloc(&c->stats_part, (kTileStats * (size_t)ntl + kDissBlocks) * sizeof(double)) != hipSuccess) {
    (void)hipFree(c->base); (void)hipFree(c->coef); if (c->halo) (void)hipFree(c->halo); delete c;
    return fail("hipMalloc of stats failed");
  }
  a.stats = c->stats_part;
  c->diss_part = c->stats_part + kTileStats * (size_t)ntl;
  c->stream = nullptr;
  // the no-allocation entry point's buffers: this device's copy of scratch
  c->scratch = scratch_address(c->device, &c->scratch_err);
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
  if (c->halo) (void)hipFree(c->halo);
  if (c->met) (void)hipFree(c->met);
  if (c->uwbuf) (void)hipFree(c->uwbuf);
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
  c->halo_valid = false;
  // main path: flowinit ends with the primitives of the initial field
  return c->a.main_sem ? updatefvar_dev(c) : 0;
}

int astr_hip_init_vortex(astr_hip_ctx* c, double beta, double xc, double yc, double u0, double v0) {
  if (check_ctx(c)) return -1;
  if (c->a.ndir != 2) return fail("the isentropic vortex is a two-dimensional case (km = 0)");
  const double g = c->cfg.gamma, ma = c->cfg.mach;
  hipLaunchKernelGGL(k_init_vortex, pt_grid(c->a), dim3(64), 0, c->stream, c->a, beta, xc, yc, u0,
                     v0, c->cfg.lx, c->cfg.ly, g - 1.0, ma * ma);
  HIP_OK(hipGetLastError());
  c->halo_valid = false;
  return c->a.main_sem ? updatefvar_dev(c) : 0;
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
  DevBuf tmp;
  HIP_OK(tmp.alloc(dn * nf * sizeof(double)));
  if (to_dev) HIP_OK(hipMemcpyAsync(tmp.p, host, dn * nf * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_pack, dense_grid(a, hm), dim3(64), 0, c->stream, to_dev ? tmp.p : dev,
                     to_dev ? dev : tmp.p, a, nf, hm, to_dev);
  HIP_OK(hipGetLastError());
  if (!to_dev) HIP_OK(hipMemcpyAsync(host, tmp.p, dn * nf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int astr_hip_upload(astr_hip_ctx* c, const double* q, int hm) {
  if (check_ctx(c)) return -1;
  if (!q || bad_halo(c->a, hm)) return fail("bad argument to astr_hip_upload");
  if (xfer(c, const_cast<double*>(q), 5, c->a.q, hm, 1)) return -1;
  c->halo_valid = false;
  // main path: the primitives of the uploaded q, as updatefvar leaves them
  if (c->a.main_sem && updatefvar_dev(c)) return -1;
  if (c->a.main_sem) HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int astr_hip_download(astr_hip_ctx* c, double* q, double* prims, int hm) {
  if (check_ctx(c)) return -1;
  if (bad_halo(c->a, hm)) return fail("bad argument to astr_hip_download");
  if (q && xfer(c, q, 5, c->a.q, hm, 0)) return -1;
  // main path: the stored primitives, as the last updatefvar / qswap left them
  if (prims && c->a.main_sem) return xfer(c, prims, 6, c->a.pr, hm, 0);
  if (prims) {
    const FlowArgs& a = c->a;
    const size_t dn = (size_t)(a.n[0] + 1 + 2 * hm) * (a.n[1] + 1 + 2 * hm) * (a.n[2] + 1 + 2 * hm);
    DevBuf tmp;
    HIP_OK(tmp.alloc(dn * 6 * sizeof(double)));
    hipLaunchKernelGGL(k_prims, dense_grid(a, hm), dim3(64), 0, c->stream, a, tmp.p, hm);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(prims, tmp.p, dn * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int astr_hip_updatefvar(astr_hip_ctx* c, double* rho, double* vel, double* prs, double* tmp, int hm) {
  if (check_ctx(c)) return -1;
  if (bad_halo(c->a, hm)) return fail("bad argument to astr_hip_updatefvar");
  const FlowArgs& a = c->a;
  const size_t dn = (size_t)(a.n[0] + 1 + 2 * hm) * (a.n[1] + 1 + 2 * hm) * (a.n[2] + 1 + 2 * hm);
  if (a.main_sem) {
    // the stored primitives become those of the current q; each requested
    // array is packed from them with its periodic halo
    if (updatefvar_dev(c)) return -1;
    if (rho && xfer(c, rho, 1, a.pr, hm, 0)) return -1;
    if (vel && xfer(c, vel, 3, a.pr + a.fs, hm, 0)) return -1;
    if (prs && xfer(c, prs, 1, a.pr + 4 * a.fs, hm, 0)) return -1;
    if (tmp && xfer(c, tmp, 1, a.pr + 5 * a.fs, hm, 0)) return -1;
    return 0;
  }
  DevBuf pb;
  HIP_OK(pb.alloc(dn * 6 * sizeof(double)));
  hipLaunchKernelGGL(k_prims, dense_grid(a, hm), dim3(64), 0, c->stream, a, pb.p, hm);
  HIP_OK(hipGetLastError());
  // k_prims writes rho, u, v, w, p, T as consecutive fields: vel(...,1:3) is
  // fields 1..3 in the column-major layout
  const size_t fb = dn * sizeof(double);
  if (rho) HIP_OK(hipMemcpyAsync(rho, pb.p, fb, hipMemcpyDeviceToHost, c->stream));
  if (vel) HIP_OK(hipMemcpyAsync(vel, pb.p + dn, 3 * fb, hipMemcpyDeviceToHost, c->stream));
  if (prs) HIP_OK(hipMemcpyAsync(prs, pb.p + 4 * dn, fb, hipMemcpyDeviceToHost, c->stream));
  if (tmp) HIP_OK(hipMemcpyAsync(tmp, pb.p + 5 * dn, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int astr_hip_set_fvar(astr_hip_ctx* c, const double* rho, const double* vel, const double* prs,
                      const double* tmp, int hm) {
  if (check_ctx(c)) return -1;
  const FlowArgs& a = c->a;
  if (!a.main_sem) return fail("astr_hip_set_fvar: the stored primitives exist in ASTR_SEM_MAIN only");
  if (!rho || !vel || !prs || !tmp || bad_halo(a, hm)) return fail("bad argument to astr_hip_set_fvar");
  if (xfer(c, const_cast<double*>(rho), 1, a.pr, hm, 1) || xfer(c, const_cast<double*>(vel), 3, a.pr + a.fs, hm, 1) ||
      xfer(c, const_cast<double*>(prs), 1, a.pr + 4 * a.fs, hm, 1) ||
      xfer(c, const_cast<double*>(tmp), 1, a.pr + 5 * a.fs, hm, 1))
    return -1;
  return 0;
}

// ---------------------------------------------------------------------------
// checkpoint / restart (writechkpt / readcheckpoint, src/readwrite.F90:2390,
// 1381).  The reference writes ro, u1, u2, u3, p, t, nstep and time to an
// HDF5 flowfield file and rebuilds q from them at restart (updateq); without
// HDF5 this file is raw little-endian FP64 behind a fixed header, and it also
// holds q, so that a resumed run continues bit for bit.
//   bytes 0..7   "ASTRCHK1"
//   int32        header bytes (64), im, jm, km, semantics, nfields (11)
//   int64        nstep
//   double       time
//   (zero padding to 64 bytes)
//   nfields fields of (im+1)(jm+1)(km+1) doubles, i fastest, in the order
//   ro, u1, u2, u3, p, t (the stored primitives in the main semantics, the
//   primitives of q otherwise), q1..q5
// ---------------------------------------------------------------------------
namespace {
constexpr char kChkMagic[9] = "ASTRCHK1";
constexpr int kChkHeader = 64;
constexpr int kChkFields = 11;

struct ChkHeader {
  char magic[8];
  int32_t hbytes, im, jm, km, semantics, nfields;
  int64_t nstep;
  double time;
  char pad[kChkHeader - 8 - 6 * 4 - 8 - 8];
};
static_assert(sizeof(ChkHeader) == kChkHeader, "checkpoint header size");
}  // namespace

int astr_hip_write_checkpoint(astr_hip_ctx* c, const char* path, long long nstep, double time) {
  if (check_ctx(c)) return -1;
  if (!path) return fail("null checkpoint path");
  const FlowArgs& a = c->a;
  const size_t np = (size_t)(a.n[0] + 1) * (a.n[1] + 1) * (a.n[2] + 1);
  std::vector<double> buf((size_t)kChkFields * np);
  if (astr_hip_download(c, buf.data() + 6 * np, buf.data(), 0)) return -1;
  ChkHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kChkMagic, 8);
  h.hbytes = kChkHeader;
  h.im = a.n[0]; h.jm = a.n[1]; h.km = a.n[2];
  h.semantics = a.main_sem ? ASTR_SEM_MAIN : ASTR_SEM_MINIAPP;
  h.nfields = kChkFields;
  h.nstep = nstep;
  h.time = time;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot open checkpoint for writing: ") + path);
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
                  std::fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
  if (std::fclose(f) != 0 || !ok) return fail(std::string("checkpoint write failed: ") + path);
  return 0;
}

int astr_hip_read_checkpoint(astr_hip_ctx* c, const char* path, long long* nstep, double* time) {
  if (check_ctx(c)) return -1;
  if (!path) return fail("null checkpoint path");
  const FlowArgs& a = c->a;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open checkpoint: ") + path);
  ChkHeader h;
  const bool hok = std::fread(&h, sizeof(h), 1, f) == 1;
  auto bad = [&](const std::string& m) {
    std::fclose(f);
    return fail("checkpoint " + std::string(path) + ": " + m);
  };
  if (!hok || std::memcmp(h.magic, kChkMagic, 8) != 0 || h.hbytes != kChkHeader)
    return bad("not an astr_hip checkpoint");
  if (h.im != a.n[0] || h.jm != a.n[1] || h.km != a.n[2]) return bad("grid size differs from the context's");
  if (h.nfields != kChkFields) return bad("unexpected field count");
  if (h.semantics != (a.main_sem ? ASTR_SEM_MAIN : ASTR_SEM_MINIAPP))
    return bad("written under the other semantics mode");
  const size_t np = (size_t)(a.n[0] + 1) * (a.n[1] + 1) * (a.n[2] + 1);
  std::vector<double> buf((size_t)kChkFields * np);
  if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return bad("truncated");
  std::fclose(f);
  if (astr_hip_upload(c, buf.data() + 6 * np, 0)) return -1;
  if (a.main_sem) {
    // the stored primitives as written (they lag q by the filter)
    const double* p = buf.data();
    if (astr_hip_set_fvar(c, p, p + np, p + 4 * np, p + 5 * np, 0)) return -1;
  }
  if (nstep) *nstep = h.nstep;
  if (time) *time = h.time;
  return 0;
}

int astr_hip_filterq(astr_hip_ctx* c) {
  if (check_ctx(c)) return -1;
  for (int d = 0; d < c->a.ndir; d++)
    if (pass(c, F_ONLY, d)) return -1;
  c->halo_valid = false;
  return 0;
}

int astr_hip_gradcal(astr_hip_ctx* c, double* dvel, double* dtmp) {
  if (check_ctx(c)) return -1;
  if (ensure_halo(c) || run_rhs(c, false, false, true)) return -1;
  if (dvel && xfer(c, dvel, 9, c->a.dv, 0, 0)) return -1;
  if (dtmp && xfer(c, dtmp, 3, c->a.dv + 9 * c->a.fs, 0, 0)) return -1;
  return 0;
}

int astr_hip_rhscal(astr_hip_ctx* c, double* qrhs) {
  if (check_ctx(c)) return -1;
  if (ensure_halo(c) || run_rhs(c, false, false, true)) return -1;
  if (qrhs && xfer(c, qrhs, 5, c->a.qrhs, 0, 0)) return -1;
  return 0;
}

int astr_hip_rk3_steps(astr_hip_ctx* c, int nsteps, double* ke, double* ens) {
  return astr_hip_rk3_steps_stats(c, nsteps, ke, ens, nullptr);
}

int astr_hip_rk3_steps_stats(astr_hip_ctx* c, int nsteps, double* ke, double* ens, double* diss) {
  if (check_ctx(c)) return -1;
  if (nsteps < 0) return fail("negative step count");
  const bool want = ke || ens || diss;
  if (want && (c->stats_cap < nsteps || !c->stats_out)) {
    // the capacity is recorded only once the buffer exists: a failed
    // allocation leaves (nullptr, 0), and the next call allocates again
    if (c->stats_out) (void)hipFree(c->stats_out);
    c->stats_out = nullptr;
    c->stats_cap = 0;
    double* p = nullptr;
    if (test_fail_alloc()) return fail("hipMalloc of the step statistics failed (injected)");
    HIP_OK(hipMalloc(&p, kStepStats * (size_t)std::max(nsteps, 1) * sizeof(double)));
    c->stats_out = p;
    c->stats_cap = nsteps;
  }
  return rk3_run(c, nsteps, StatsBufs{c->stats_part, c->diss_part, c->stats_out}, ke, ens, diss);
}

int astr_hip_rk3_steps_stats_noalloc(astr_hip_ctx* c, int nsteps, double* ke, double* ens, double* diss) {
  if (check_ctx(c)) return -1;
  if (nsteps < 0) return fail("negative step count");
  if (!c->scratch) return fail("no scratch buffer on this device: " + c->scratch_err);
  const FlowArgs& a = c->a;
  const int stats_prog = c->split_last ? GL_GRAD : GC_LAST;
  const long long ntl = tiles_of(a.n, a.ndir - 1, kShapes[c->shape[stats_prog][a.ndir - 1]].TL);
  const long long need = kTileStats * ntl + kDissBlocks + kStepStats * (long long)std::max(nsteps, 1);
  if (need > kScratchDoubles)
    return fail("astr_hip_rk3_steps_stats_noalloc: " + std::to_string(need) +
                " doubles of statistics exceed the scratch buffer of " + std::to_string(kScratchDoubles) +
                " (fewer steps per call, or astr_hip_rk3_steps_stats)");
  double* s = c->scratch;
  return rk3_run(c, nsteps, StatsBufs{s, s + kTileStats * ntl, s + kTileStats * ntl + kDissBlocks}, ke, ens,
                 diss);
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

// pass programs a case of ndir directions launches, per direction
static bool pass_used(int prog, int d, int nd) {
  switch (prog) {
    case GC_FIRST:
    case FGC_FIRST: return d == 0;
    case GC_MID: return d > 0 && d < nd - 1;
    case GC_LAST:
    case GL_GRAD:
    case GL_CONV: return d == nd - 1;
    case V_MID: return d < nd - 1;
    case V_LAST_RHS:
    case V_LAST_RK: return d == nd - 1;
    default: return d < nd;
  }
}

int astr_hip_autotune(astr_hip_ctx* c, int reps) {
  if (check_ctx(c)) return -1;
  if (reps < 1 || reps > 1000) return fail("autotune reps must be 1..1000");
  FlowArgs& a = c->a;
  const FlowArgs keep = a;
  // the passes overwrite q and qsave (and the main path's stored
  // primitives): keep a copy and put it back afterwards
  const size_t qbytes = (size_t)10 * a.fs * sizeof(double);
  const size_t pbytes = a.main_sem ? (size_t)6 * a.fs * sizeof(double) : 0;
  DevBuf snapb;
  HIP_OK(snapb.alloc(qbytes + pbytes));
  double* snap = snapb.p;
  HIP_OK(hipMemcpyAsync(snap, a.q, qbytes, hipMemcpyDeviceToDevice, c->stream));
  if (pbytes) HIP_OK(hipMemcpyAsync(snap + 10 * a.fs, a.pr, pbytes, hipMemcpyDeviceToDevice, c->stream));
  if (ensure_halo(c)) return -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail("hipEventCreate failed");
  }
  a.stage = 0;
  a.rk_a = 1.0; a.rk_b = 0.0; a.rk_c = 1.0;
  a.dostats = 0;
  a.store_last_grad = 0;
  // the main path's RK pass is timed with the filter sweep it carries in
  // rk3 steps
  a.rk_filter = a.main_sem && c->fuse_main_filter ? 1 : 0;
  int rc = 0;
  for (int pr = 0; pr < kNumProgs && !rc; pr++) {
    for (int d = 0; d < a.ndir && !rc; d++) {
      if (!pass_used(pr, d, a.ndir) || (pr == FGC_FIRST && a.main_sem)) continue;
      float best = 1e30f;
      for (int sh = 0; sh < kNumShapes && !rc; sh++) {
        if (!selection_ok(sh, pr, a.n[d], c->device)) continue;
        for (int mb = 1; mb <= kMinbVariants && !rc; mb++) {
          hipError_t e = launch_pass(sh, pr, mb, a, d, c->stream);
          if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
          for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_pass(sh, pr, mb, a, d, c->stream);
          if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) {
            rc = fail(std::string("autotune launch: ") + hipGetErrorString(e));
            break;
          }
          ms /= (float)reps;
          if (ms < best) {
            best = ms;
            c->shape[pr][d] = sh;
            c->minb[pr][d] = mb;
            c->tuned_ms[pr][d] = ms;
          }
        }
      }
    }
  }
  if (!rc) {
    // the fused and the split last gradient pass, each timed inside whole
    // right-hand sides with the selections above (in a step the split form's
    // second kernel finds the first one's data in cache, which the per-pass
    // timings do not see)
    float form_ms[2] = {0.f, 0.f};
    for (int sp = 0; sp < 2 && !rc; sp++) {
      c->split_last = sp;
      if (run_rhs(c, false, false, false)) {
        rc = -1;
        break;
      }
      hipError_t e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && e == hipSuccess; r++)
        if (run_rhs(c, false, false, false)) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(&form_ms[sp], e0, e1);
      if (e != hipSuccess) rc = fail(std::string("autotune rhs: ") + hipGetErrorString(e));
    }
    if (!rc) c->split_last = form_ms[1] < form_ms[0];
    // the filtering first pass against the first pass plus the x filter pass
    if (!rc && !a.main_sem)
      c->fuse_fx = c->tuned_ms[FGC_FIRST][0] < c->tuned_ms[GC_FIRST][0] + c->tuned_ms[F_ONLY][0];
  }
  a = keep;
  hipError_t e = hipMemcpyAsync(a.q, snap, qbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess && pbytes)
    e = hipMemcpyAsync(a.pr, snap + 10 * a.fs, pbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc && e != hipSuccess) rc = fail(std::string("autotune restore: ") + hipGetErrorString(e));
  return rc;
}

int astr_hip_get_pass_config(astr_hip_ctx* c, int prog, int dir, int* shape, int* minb, float* ms) {
  if (check_ctx(c)) return -1;
  if (prog < 0 || prog >= kNumProgs || dir < 0 || dir >= c->a.ndir) return fail("bad pass or direction");
  if (shape) *shape = c->shape[prog][dir];
  if (minb) *minb = c->minb[prog][dir];
  if (ms) *ms = c->tuned_ms[prog][dir];
  return 0;
}

int astr_hip_set_pass_config(astr_hip_ctx* c, int prog, int dir, int shape, int minb) {
  if (check_ctx(c)) return -1;
  if (prog < 0 || prog >= kNumProgs || dir < 0 || dir >= c->a.ndir) return fail("bad pass or direction");
  if (shape < 0 || shape >= kNumShapes || !selection_ok(shape, prog, c->a.n[dir], c->device))
    return fail("launch shape does not fit this line length or the LDS");
  if (minb < 1 || minb > kMinbVariants) return fail("register budget must be 1 or 2");
  c->shape[prog][dir] = shape;
  c->minb[prog][dir] = minb;
  c->tuned_ms[prog][dir] = 0.f;
  return 0;
}

int astr_hip_set_last_split(astr_hip_ctx* c, int split) {
  if (check_ctx(c)) return -1;
  c->split_last = split ? 1 : 0;
  return 0;
}

int astr_hip_set_filter_fusion(astr_hip_ctx* c, int on) {
  if (check_ctx(c)) return -1;
  c->fuse_fx = on ? 1 : 0;
  c->fuse_main_filter = on ? 1 : 0;
  return 0;
}

int astr_hip_get_last_split(astr_hip_ctx* c, int* split) {
  if (check_ctx(c) || !split) return fail("bad argument");
  *split = c->split_last;
  return 0;
}

int astr_hip_line_batch_nt(int op, int ntype, int n, long long nlines, const double* f, double* out,
                           double alfa_filter, double beta_halo, double beta_bound) {
  const int bits = ntype_bits(ntype);
  if (n < 8 || nlines < 0 || (op != 0 && op != 1) || bits < 0)
    return fail("bad argument to astr_hip_line_batch_nt");
  if (nlines == 0) return 0;
  if (!f || !out) return fail("null buffer");
  const int sh = pick_shape(n);
  if (sh < 0) return fail("no launch shape fits this line length");
  FilterCoef fc;
  double h63[kBndCoefs];
  filter_coefs(alfa_filter, beta_halo, fc, h63);
  filter_coefs_b(alfa_filter, beta_bound, h63 + 33);
  std::vector<double> tab(h63, h63 + kBndCoefs), t2;
  if (op == 0) deriv_tables_nt(n, bits, t2);
  else filter_tables_nt(n, bits, alfa_filter, beta_halo, t2);
  tab.insert(tab.end(), t2.begin(), t2.end());
  DevBuf bt, bf, bo;
  HIP_OK(bt.alloc(tab.size() * sizeof(double)));
  HIP_OK(bf.alloc((size_t)nlines * (n + 11) * sizeof(double)));
  HIP_OK(bo.alloc((size_t)nlines * (n + 1) * sizeof(double)));
  HIP_OK(hipMemcpy(bt.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(bf.p, f, (size_t)nlines * (n + 11) * sizeof(double), hipMemcpyHostToDevice));
  fc.ch = bt.p;
  HIP_OK(launch_lines(sh, op, bits, bf.p, bo.p, n, nlines, bt.p + kBndCoefs, fc, 0));
  HIP_OK(hipMemcpy(out, bo.p, (size_t)nlines * (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int astr_hip_line_batch(int op, int n, long long nlines, const double* f, double* out,
                        double alfa_filter, double beta_halo) {
  return astr_hip_line_batch_nt(op, 3, n, nlines, f, out, alfa_filter, beta_halo, 0.98);
}

int astr_hip_gridgeom_nt(astr_hip_ctx* c, const int* npdc, const double* x, double* jacob, double* dxi) {
  if (check_ctx(c)) return -1;
  if (!x || !npdc) return fail("null coordinates or boundary types");
  HIP_OK(hipSetDevice(c->device));
  const int nd = c->a.ndir;
  int bits[3] = {0, 0, 0};
  for (int d = 0; d < nd; d++) {
    if (npdc[d] != 3 && npdc[d] != 4)
      return fail("gridgeom on one rank takes boundary type 3 (periodic) or 4 (physical at both ends)");
    bits[d] = ntype_bits(npdc[d]);
  }
  GeomGrid G;
  for (int d = 0; d < 3; d++) G.n[d] = c->a.n[d];
  G.st[0] = 1;
  G.st[1] = G.n[0] + 1;
  G.st[2] = (long long)(G.n[0] + 1) * (G.n[1] + 1);
  G.fs = G.st[2] * (G.n[2] + 1);
  long long maxl = 0, maxo = 0;
  for (int d = 0; d < nd; d++) {
    const long long nl = G.fs / (G.n[d] + 1);
    maxl = std::max(maxl, nl * (G.n[d] + 11));
    maxo = std::max(maxo, nl * (G.n[d] + 1));
  }
  DevBuf bx, bdx, bxi, bj, bl, bo, bt[3];
  HIP_OK(bx.alloc(3 * G.fs * sizeof(double)));
  HIP_OK(bdx.alloc(9 * G.fs * sizeof(double)));
  HIP_OK(bxi.alloc(9 * G.fs * sizeof(double)));
  HIP_OK(bj.alloc(G.fs * sizeof(double)));
  HIP_OK(bl.alloc(maxl * sizeof(double)));
  HIP_OK(bo.alloc(maxo * sizeof(double)));
  for (int d = 0; d < nd; d++) {
    std::vector<double> t;
    deriv_tables_nt(G.n[d], bits[d], t);
    HIP_OK(bt[d].alloc(t.size() * sizeof(double)));
    HIP_OK(hipMemcpy(bt[d].p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  hipStream_t st = c->stream;
  HIP_OK(hipMemcpyAsync(bx.p, x, 3 * G.fs * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemsetAsync(bdx.p, 0, 9 * G.fs * sizeof(double), st));
  HIP_OK(hipMemsetAsync(bxi.p, 0, 9 * G.fs * sizeof(double), st));
  auto run = [&](int dir, GeomLine L, double* dst, int acc) -> int {
    const long long nl = G.fs / (G.n[dir] + 1);
    const long long ng = nl * (G.n[dir] + 11), ns = nl * (G.n[dir] + 1);
    hipLaunchKernelGGL(k_geom_gather, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, bx.p, bdx.p, G,
                       dir, L, nl, bl.p);
    HIP_OK(hipGetLastError());
    HIP_OK(deriv_lines_device(G.n[dir], nl, bl.p, bo.p, bt[dir].p, bits[dir], st));
    hipLaunchKernelGGL(k_geom_scatter, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, bo.p, G, dir,
                       nl, dst, acc);
    HIP_OK(hipGetLastError());
    return 0;
  };
  // dx(m,d) = d x_m / d xi_d (geom.F90:128-161)
  for (int d = 0; d < nd; d++)
    for (int m = 0; m < 3; m++)
      if (run(d, GeomLine{0, m, 0, 0, 0}, bdx.p + (m + 3 * d) * G.fs, 0)) return -1;
  // J dxi (geom.F90:395-515), 3-D only; the 2-D form is pointwise
  if (nd == 3)
    for (const auto& T : kGeomTerms) {
      double* dst = bxi.p + ((T[1] - 1) + 3 * (T[2] - 1)) * G.fs;
      if (run(T[0] - 1, GeomLine{1, 0, T[3] - 1, T[4] - 1, T[5] - 1}, dst, 1)) return -1;
    }
  hipLaunchKernelGGL(k_geom_final, dim3((unsigned)((G.fs + 255) / 256)), dim3(256), 0, st, bdx.p, bxi.p, bj.p,
                     G, nd);
  HIP_OK(hipGetLastError());
  if (jacob) HIP_OK(hipMemcpyAsync(jacob, bj.p, G.fs * sizeof(double), hipMemcpyDeviceToHost, st));
  if (dxi) HIP_OK(hipMemcpyAsync(dxi, bxi.p, 9 * G.fs * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return 0;
}

int astr_hip_gridgeom(astr_hip_ctx* c, const double* x, double* jacob, double* dxi) {
  const int npdc[3] = {3, 3, 3};
  return astr_hip_gridgeom_nt(c, npdc, x, jacob, dxi);
}

int astr_hip_qswap(astr_hip_ctx* c) {
  if (check_ctx(c)) return -1;
  HIP_OK(hipSetDevice(c->device));
  if (c->a.main_sem) return qswap_main(c, true);
  const FlowArgs& a = c->a;
  for (int d = 0; d < a.ndir; d++) {
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    const long long cnt = 5LL * (a.n[d1] + 1) * (a.n[d2] + 1);
    hipLaunchKernelGGL(k_qswap_dir, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, a, d);
    HIP_OK(hipGetLastError());
  }
  return 0;
}

int astr_hip_set_metrics(astr_hip_ctx* c, const double* jacob, const double* dxi) {
  if (check_ctx(c)) return -1;
  if (!jacob || !dxi) return fail("null metrics");
  FlowArgs& a = c->a;
  const long long fs = (long long)(a.n[0] + 1) * (a.n[1] + 1) * (a.n[2] + 1);
  const int nd = a.ndir;
  double mean[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < nd; d++) {
    const double* f = dxi + (d + 3 * d) * fs;
    // long double: a plain running sum of ~1e4 equal values drifts by
    // ~1e-14 relative, which the RHS would see as a changed grid spacing
    long double s = 0.0L;
    for (long long o = 0; o < fs; o++) s += f[o];
    mean[d] = (double)(s / (long double)fs);
  }
  long double js = 0.0L;
  for (long long o = 0; o < fs; o++) js += jacob[o];
  const double jmean = (double)(js / (long double)fs);
  for (long long o = 0; o < fs; o++)
    if (!(jacob[o] > 0.0)) return fail("metrics: the Jacobian must be positive");
  // uniform Cartesian: dxi diagonal and constant, J constant, to 1e-12
  const double tol = 1e-12;
  bool uniform = true;
  double scale = 0.0;
  for (int d = 0; d < nd; d++) scale = std::max(scale, std::fabs(mean[d]));
  for (int q = 0; q < 3 && uniform; q++)
    for (int p = 0; p < 3 && uniform; p++) {
      const double* f = dxi + (p + 3 * q) * fs;
      const bool diag = p == q && p < nd;
      for (long long o = 0; o < fs && uniform; o++) {
        const double ref = diag ? mean[p] : 0.0;
        if (!(std::fabs(f[o] - ref) <= tol * scale)) uniform = false;
      }
    }
  for (long long o = 0; o < fs && uniform; o++)
    if (!(std::fabs(jacob[o] - jmean) <= tol * jmean)) uniform = false;
  if (uniform) {
    for (int d = 0; d < nd; d++)
      if (!(mean[d] > 0.0)) return fail("metrics: d xi/d x must be positive");
    double jm = 1.0;
    for (int d = 0; d < nd; d++) jm *= 1.0 / mean[d];
    if (!(std::fabs(jmean - jm) <= tol * jm)) return fail("metrics: Jacobian inconsistent with d xi/d x");
    for (int d = 0; d < nd; d++) a.inv_d[d] = mean[d];
    a.jac = jmean;
    if (c->met) {
      HIP_OK(hipStreamSynchronize(c->stream));
      (void)hipFree(c->met);
      c->met = nullptr;
    }
    a.met = nullptr;
    return 0;
  }
  // any other grid: the main path's passes read J and every dxi(d,n) per point
  if (!a.main_sem)
    return fail("metrics of a grid that is not uniform Cartesian: the miniapp semantics take a "
                "uniform grid only (ASTR_SEM_MAIN takes any grid)");
  std::vector<double> host((size_t)10 * fs, 0.0);
  std::memcpy(host.data(), jacob, fs * sizeof(double));
  std::memcpy(host.data() + fs, dxi, 9 * fs * sizeof(double));
  if (!c->met) {
    double* p = nullptr;
    HIP_OK(hipMalloc(&p, (size_t)10 * a.fs * sizeof(double)));
    c->met = p;
    HIP_OK(hipMemset(c->met, 0, (size_t)10 * a.fs * sizeof(double)));
  }
  if (xfer(c, host.data(), 10, c->met, 0, 1)) return -1;
  a.jac = jmean;
  a.met = c->met;
  return 0;
}

int astr_hip_set_boundary(astr_hip_ctx* c, const int* npdc) {
  if (check_ctx(c)) return -1;
  if (!npdc) return fail("null npdc");
  FlowArgs& a = c->a;
  if (!a.main_sem) return fail("physical boundaries are a main-path (ASTR_SEM_MAIN) feature");
  int bits[3] = {0, 0, 0};
  for (int d = 0; d < a.ndir; d++) {
    if (npdc[d] != 3 && npdc[d] != 4)
      return fail("npdc must be 3 (periodic) or 4 (physical at both ends) on one rank");
    bits[d] = ntype_bits(npdc[d]);
    if (bits[d] && a.uwd) return fail("the upwind-biased convection runs on periodic directions only");
  }
  const FlowArgs keep = a;
  for (int d = 0; d < 3; d++) a.bnd[d] = bits[d];
  if (build_tables(c)) {
    a = keep;
    return -1;
  }
  c->halo_valid = false;
  return 0;
}

int astr_hip_set_convection(astr_hip_ctx* c, int upwind, int recon_schem, int lchardecomp, double bfacmpld,
                            double shkcrt) {
  if (check_ctx(c)) return -1;
  FlowArgs& a = c->a;
  if (!upwind) {
    (void)hipStreamSynchronize(c->stream);
    if (c->uwbuf) (void)hipFree(c->uwbuf);
    c->uwbuf = nullptr;
    c->uw = UwArgs{};
    a.uwd = 0;
    return 0;
  }
  if (upwind != 1 && upwind != 2)
    return fail("upwind must be 0 (central), 1 (explicit upwind-biased) or 2 (compact upwind-biased)");
  if (!a.main_sem) return fail("the upwind-biased convection is a main-path (ASTR_SEM_MAIN) feature");
  const bool compact = upwind == 2;
  if (!compact && !(recon_schem == -1 || recon_schem == 0 || recon_schem == 1 || recon_schem == 2 ||
                    recon_schem == 3 || recon_schem == 5 || recon_schem == 6))
    return fail("recon_schem must be -1, 0, 1, 2, 3, 5 or 6 (recons_exp, flux.F90:305-335)");
  for (int d = 0; d < a.ndir; d++)
    if (a.bnd[d]) return fail("the upwind-biased convection runs on periodic directions only");
  const int nf = lchardecomp ? kUwFieldsCd : kUwFields;
  long long sws = 0;
  for (int d = 0; d < a.ndir; d++) sws = std::max(sws, uw_points(a.n, d));
  sws = (sws + 63) / 64 * 64;
  const long long lsh_doubles = (a.fs + 7) / 8;
  // compact: the interface-flux buffer and the LU factors of the line
  // systems (compact_flux_initiate, flux.F90:33-118, scheme 543, ntype 3)
  const long long nfh = compact ? kUwFields * sws : 0;
  const long long nac = compact ? (cuw_ac_doubles(a.n, a.ndir) + 63) / 64 * 64 : 0;
  std::vector<double> ac(nac, 0.0);
  if (compact) {
    const double one6 = 1.0 / 6.0;
    for (int d = 0; d < a.ndir; d++)
      for (int side = 0; side < 2; side++) {
        const bool plus = side == 0;
        const double av = plus ? 0.5 - one6 * bfacmpld : one6 + one6 * bfacmpld;
        const double cv = plus ? one6 + one6 * bfacmpld : 0.5 - one6 * bfacmpld;
        double* t = ac.data() + cuw_ac_offset(a.n, d, plus);
        const int N = a.n[d] + 4;
        // tridiagonal_thomas_proprocess (commfunc.F90:752-774); a = c = 0 on
        // the end rows, row 0 holds ac(1,1) = c(1) only
        t[0] = 0.0;
        t[1] = 1.0;
        t[2] = 0.0;
        for (int r = 1; r < N; r++) {
          const bool end = r == N - 1;
          const double aa = end ? 0.0 : av, cc = end ? 0.0 : cv;
          t[3 * r] = cc / (1.0 - aa * t[3 * (r - 1)]);
          t[3 * r + 1] = 1.0 / (1.0 - aa * t[3 * (r - 1)]);
          t[3 * r + 2] = aa / (1.0 - aa * t[3 * (r - 1)]);
        }
      }
  }
  const size_t bytes = (size_t)(nf * sws + a.fs + lsh_doubles + nfh + nac) * sizeof(double);
  (void)hipStreamSynchronize(c->stream);
  if (c->uwbuf) (void)hipFree(c->uwbuf);
  c->uwbuf = nullptr;
  a.uwd = 0;
  if (test_fail_alloc() || hipMalloc(&c->uwbuf, bytes) != hipSuccess) {
    c->uwbuf = nullptr;
    c->uw = UwArgs{};
    return fail("hipMalloc of the upwind split buffer failed");
  }
  HIP_OK(hipMemset(c->uwbuf, 0, bytes));
  UwArgs& u = c->uw;
  u.recon = recon_schem;
  u.chardec = lchardecomp ? 1 : 0;
  u.sensor = recon_schem == 5 || lchardecomp;
  u.bfac = bfacmpld;
  u.shkcrt = shkcrt;
  u.gamma = c->cfg.gamma;
  u.mach = c->cfg.mach;
  u.sw = c->uwbuf;
  u.sws = sws;
  u.ssf = c->uwbuf + nf * sws;
  u.lsh = reinterpret_cast<unsigned char*>(u.ssf + a.fs);
  u.compact = compact ? 1 : 0;
  if (compact) {
    // the explicit scheme's sensor runs for recon_schem 5 or lchardecomp,
    // the compact one's for lchardecomp only (solver.F90:223, 229)
    u.sensor = lchardecomp ? 1 : 0;
    u.fhc = u.ssf + a.fs + lsh_doubles;
    double* act = u.fhc + nfh;
    HIP_OK(hipMemcpy(act, ac.data(), nac * sizeof(double), hipMemcpyHostToDevice));
    u.cuw_ac = act;
  }
  a.uwd = 1;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batching: many independent cases advanced together.  Every pass of a stage
// is launched once per group of cases sharing (program, launch shape,
// register budget) as a persistent kernel over the groups' concatenated
// tiles; a whole RK3 step can be captured in a hipGraph and replayed.
// ---------------------------------------------------------------------------
namespace {

struct StatsJob {
  const double* part;
  int ntiles;
  double scale;
  double* out;  // kBatchStats doubles per step
};

// per-case statistics of the current step (index *step), one block per case
__global__ __launch_bounds__(256) void k_stats_batch(const StatsJob* __restrict__ jobs,
                                                     const int* __restrict__ step) {
  const StatsJob J = jobs[blockIdx.x];
  stats_sum(J.part, J.ntiles, J.out + kBatchStats * *step, J.scale);
}

__global__ void k_set_int(int* p, int v) { *p = v; }
__global__ void k_add_int(int* p, int v) { *p += v; }

hipError_t launch_batch(int shape, int prog, int minb, const BatchLaunch& bl, int dir,
                        const StageArgs& sa, int nblocks, hipStream_t st) {
  switch (shape) {
    case 0: return launch_batch_s0(prog, minb, bl, dir, sa, nblocks, st);
    case 1: return launch_batch_s1(prog, minb, bl, dir, sa, nblocks, st);
    case 2: return launch_batch_s2(prog, minb, bl, dir, sa, nblocks, st);
    case 3: return launch_batch_s3(prog, minb, bl, dir, sa, nblocks, st);
    case 4: return launch_batch_s4(prog, minb, bl, dir, sa, nblocks, st);
    case 5: return launch_batch_s5(prog, minb, bl, dir, sa, nblocks, st);
    case 6: return launch_batch_s6(prog, minb, bl, dir, sa, nblocks, st);
    case 7: return launch_batch_s7(prog, minb, bl, dir, sa, nblocks, st);
    case 8: return launch_batch_s8(prog, minb, bl, dir, sa, nblocks, st);
    case 9: return launch_batch_s9(prog, minb, bl, dir, sa, nblocks, st);
    case 10: return launch_batch_s10(prog, minb, bl, dir, sa, nblocks, st);
    case 11: return launch_batch_s11(prog, minb, bl, dir, sa, nblocks, st);
    default: return launch_batch_s12(prog, minb, bl, dir, sa, nblocks, st);
  }
}

// the eight pass slots of a stage: GC x,y,z; V x,y,z; F y,x (the filter
// sweeps in the single-case stage's order).  The program a case of nd
// directions runs in slot (kind, d), or -1.
int slot_prog(int kind, int d, int nd) {
  if (kind == 0) return d < nd ? (d == nd - 1 ? GC_LAST : (d == 0 ? GC_FIRST : GC_MID)) : -1;
  if (kind == 1) return d < nd ? (d == nd - 1 ? V_LAST_RK : V_MID) : -1;
  return d < nd - 1 ? F_ONLY : -1;
}
const int kSlotKind[8] = {0, 0, 0, 1, 1, 1, 2, 2};
const int kSlotDir[8] = {0, 1, 2, 0, 1, 2, 1, 0};

struct BatchGroup {
  int slot, prog, dir, shape, minb, nblocks;
  std::vector<int> members;
  FlowArgs* d_args = nullptr;
  int* d_tstart = nullptr;
  BatchLaunch bl{};
};

}  // namespace

struct astr_hip_batch {
  std::vector<astr_hip_ctx*> cases;
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  std::vector<BatchGroup> groups;  // in slot order
  StatsJob* d_jobs = nullptr;
  int* d_step = nullptr;
  double* d_stats = nullptr;
  int stats_cap = 0;
  hipGraphExec_t graph = nullptr;
  bool graph_valid = false;
  int pending = 0;  // steps of the last begin not yet collected by end
};

namespace {

void free_groups(astr_hip_batch* b) {
  for (auto& g : b->groups) {
    (void)hipFree(g.d_args);
    (void)hipFree(g.d_tstart);
  }
  b->groups.clear();
}

void drop_graph(astr_hip_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_valid = false;
}

// group the cases of every slot by (program, shape, budget) from each case's
// launch selection; upload the groups' case arrays and tile offsets
int build_groups(astr_hip_batch* b) {
  drop_graph(b);
  free_groups(b);
  for (int sl = 0; sl < 8; sl++) {
    const int kind = kSlotKind[sl], d = kSlotDir[sl];
    std::vector<BatchGroup> here;
    for (int i = 0; i < (int)b->cases.size(); i++) {
      astr_hip_ctx* c = b->cases[i];
      const int prog = slot_prog(kind, d, c->a.ndir);
      if (prog < 0) continue;
      const int sh = c->shape[prog][d], mb = c->minb[prog][d];
      BatchGroup* g = nullptr;
      for (auto& h : here)
        if (h.prog == prog && h.shape == sh && h.minb == mb) g = &h;
      if (!g) {
        here.emplace_back();
        g = &here.back();
        g->slot = sl; g->prog = prog; g->dir = d; g->shape = sh; g->minb = mb;
      }
      g->members.push_back(i);
    }
    for (auto& g : here) b->groups.push_back(g);
  }
  for (auto& g : b->groups) {
    const int nm = (int)g.members.size();
    std::vector<FlowArgs> args(nm);
    std::vector<int> ts(nm + 1, 0);
    for (int k = 0; k < nm; k++) {
      const FlowArgs& a = b->cases[g.members[k]]->a;
      args[k] = a;
      ts[k + 1] = ts[k] + tiles_of(a.n, g.dir, kShapes[g.shape].TL);
    }
    HIP_OK(hipMalloc(&g.d_args, nm * sizeof(FlowArgs)));
    HIP_OK(hipMalloc(&g.d_tstart, (nm + 1) * sizeof(int)));
    HIP_OK(hipMemcpy(g.d_args, args.data(), nm * sizeof(FlowArgs), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(g.d_tstart, ts.data(), (nm + 1) * sizeof(int), hipMemcpyHostToDevice));
    g.bl.cases = g.d_args;
    g.bl.tstart = g.d_tstart;
    g.bl.ncases = nm;
    g.bl.total = ts[nm];
    // persistent grid: a multiple of 8 blocks, at most as many blocks as the
    // budget keeps resident (minb per CU), never more than the tiles
    // (a 128-thread workgroup is half a 256-thread one: twice as many fit)
    const int cap = b->ncu * g.minb * (kShapes[g.shape].NT < 256 ? 256 / kShapes[g.shape].NT : 1);
    int nb = g.bl.total < cap ? g.bl.total : cap;
    g.nblocks = (nb + 7) / 8 * 8;
  }
  return 0;
}

// the cases' FlowArgs can change between calls (time step); refresh the copies
int refresh_args(astr_hip_batch* b) {
  for (auto& g : b->groups) {
    std::vector<FlowArgs> args(g.members.size());
    for (size_t k = 0; k < g.members.size(); k++) args[k] = b->cases[g.members[k]]->a;
    HIP_OK(hipMemcpyAsync(g.d_args, args.data(), args.size() * sizeof(FlowArgs),
                          hipMemcpyHostToDevice, b->stream));
  }
  HIP_OK(hipStreamSynchronize(b->stream));
  return 0;
}

// stats jobs for a step capacity of cap steps
int ensure_stats(astr_hip_batch* b, int cap) {
  if (cap <= b->stats_cap && b->d_jobs) return 0;
  drop_graph(b);
  (void)hipFree(b->d_stats);
  (void)hipFree(b->d_jobs);
  b->d_stats = nullptr;
  b->d_jobs = nullptr;
  const int nc = (int)b->cases.size();
  HIP_OK(hipMalloc(&b->d_stats, kBatchStats * (size_t)nc * cap * sizeof(double)));
  HIP_OK(hipMalloc(&b->d_jobs, nc * sizeof(StatsJob)));
  b->stats_cap = cap;
  return 0;
}

int upload_jobs(astr_hip_batch* b) {
  const int nc = (int)b->cases.size();
  std::vector<StatsJob> jobs(nc);
  for (int i = 0; i < nc; i++) {
    astr_hip_ctx* c = b->cases[i];
    const int nd = c->a.ndir;
    jobs[i].part = c->stats_part;
    jobs[i].ntiles = tiles_of(c->a.n, nd - 1, kShapes[c->shape[GC_LAST][nd - 1]].TL);
    jobs[i].scale = c->ke_scale;
    jobs[i].out = b->d_stats + kBatchStats * (size_t)i * b->stats_cap;
  }
  HIP_OK(hipMemcpy(b->d_jobs, jobs.data(), nc * sizeof(StatsJob), hipMemcpyHostToDevice));
  return 0;
}

// enqueue one RK3 step of every case on the batch stream
int enqueue_step(astr_hip_batch* b) {
  static const double rk[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
  for (int s = 0; s < 3; s++) {
    StageArgs sa{s, s == 0 ? 1 : 0, 0, rk[s][0], rk[s][1], rk[s][2]};
    StageArgs sv = sa;
    sv.dostats = 0;
    for (int sl = 0; sl < 8; sl++) {
      for (auto& g : b->groups) {
        if (g.slot != sl) continue;
        HIP_OK(launch_batch(g.shape, g.prog, g.minb, g.bl, g.dir, g.prog == GC_LAST ? sa : sv,
                            g.nblocks, b->stream));
      }
      if (sl == 2 && s == 0) {
        hipLaunchKernelGGL(k_stats_batch, dim3((unsigned)b->cases.size()), dim3(256), 0, b->stream,
                           b->d_jobs, b->d_step);
        HIP_OK(hipGetLastError());
      }
    }
  }
  hipLaunchKernelGGL(k_add_int, dim3(1), dim3(1), 0, b->stream, b->d_step, 1);
  HIP_OK(hipGetLastError());
  return 0;
}

int capture_step(astr_hip_batch* b) {
  drop_graph(b);
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(b);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return -1;
  }
  HIP_OK(e);
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_OK(e);
  b->graph_valid = true;
  return 0;
}

int batch_sync_cases(astr_hip_batch* b) {
  for (astr_hip_ctx* c : b->cases) HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace

extern "C" {

int astr_hip_batch_create(astr_hip_ctx* const* cases, int ncases, astr_hip_batch** out) {
  if (!cases || !out || ncases < 1) return fail("bad argument to astr_hip_batch_create");
  *out = nullptr;
  astr_hip_batch* b = new astr_hip_batch();
  for (int i = 0; i < ncases; i++) {
    if (!cases[i]) { delete b; return fail("null case"); }
    if (i && cases[i]->device != cases[0]->device) { delete b; return fail("cases on different devices"); }
    if (cases[i]->a.main_sem) { delete b; return fail("batches take miniapp-semantics cases only"); }
    b->cases.push_back(cases[i]);
  }
  b->device = cases[0]->device;
  int rc = 0;
  if (hipSetDevice(b->device) != hipSuccess) rc = fail("hipSetDevice failed");
  if (!rc && hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess)
    rc = fail("device attribute query failed");
  if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail("hipStreamCreate failed");
  if (!rc && hipMalloc(&b->d_step, sizeof(int)) != hipSuccess) rc = fail("hipMalloc failed");
  if (!rc) rc = build_groups(b);
  if (!rc) rc = ensure_stats(b, 16);
  if (!rc) rc = upload_jobs(b);
  if (rc) {
    astr_hip_batch_destroy(b);
    return -1;
  }
  *out = b;
  return 0;
}

int astr_hip_batch_destroy(astr_hip_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  free_groups(b);
  (void)hipFree(b->d_jobs);
  (void)hipFree(b->d_stats);
  (void)hipFree(b->d_step);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int astr_hip_batch_rk3_begin(astr_hip_batch* b, int nsteps, int use_graph) {
  if (!b) return fail("null batch");
  if (nsteps < 0) return fail("negative step count");
  HIP_OK(hipSetDevice(b->device));
  if (batch_sync_cases(b)) return -1;
  b->pending = 0;  // set once every step is enqueued
  if (nsteps == 0) return 0;
  if (ensure_stats(b, nsteps) || upload_jobs(b) || refresh_args(b)) return -1;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, b->stream, b->d_step, 0);
  HIP_OK(hipGetLastError());
  if (use_graph) {
    if (!b->graph_valid && capture_step(b)) return -1;
    for (int it = 0; it < nsteps; it++) HIP_OK(hipGraphLaunch(b->graph, b->stream));
  } else {
    for (int it = 0; it < nsteps; it++)
      if (enqueue_step(b)) return -1;
  }
  b->pending = nsteps;
  return 0;
}

int astr_hip_batch_query(astr_hip_batch* b, int* done) {
  if (!b || !done) return fail("null argument");
  HIP_OK(hipSetDevice(b->device));
  const hipError_t e = hipStreamQuery(b->stream);
  if (e == hipErrorNotReady) {
    *done = 0;
    return 0;
  }
  HIP_OK(e);
  *done = 1;
  return 0;
}

int astr_hip_batch_rk3_end(astr_hip_batch* b, double* ke, double* ens) {
  if (!b) return fail("null batch");
  HIP_OK(hipSetDevice(b->device));
  const int nsteps = b->pending;
  b->pending = 0;
  const int nc = (int)b->cases.size();
  if (nsteps > 0 && (ke || ens)) {
    std::vector<double> h(kBatchStats * (size_t)nc * b->stats_cap);
    HIP_OK(hipMemcpyAsync(h.data(), b->d_stats, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < nc; i++)
      for (int it = 0; it < nsteps; it++) {
        const double* v = &h[kBatchStats * ((size_t)i * b->stats_cap + it)];
        if (ke) ke[(size_t)i * nsteps + it] = v[0];
        if (ens) ens[(size_t)i * nsteps + it] = v[1];
      }
  }
  HIP_OK(hipStreamSynchronize(b->stream));
  return 0;
}

int astr_hip_batch_rk3_steps(astr_hip_batch* b, int nsteps, int use_graph, double* ke, double* ens) {
  if (astr_hip_batch_rk3_begin(b, nsteps, use_graph)) return -1;
  return astr_hip_batch_rk3_end(b, ke, ens);
}

int astr_hip_batch_autotune(astr_hip_batch* b, int reps) {
  if (!b) return fail("null batch");
  if (reps < 1 || reps > 1000) return fail("autotune reps must be 1..1000");
  HIP_OK(hipSetDevice(b->device));
  if (batch_sync_cases(b)) return -1;
  const int nc = (int)b->cases.size();
  // snapshots of q and qsave, put back afterwards
  std::vector<double*> snap(nc, nullptr);
  int rc = 0;
  for (int i = 0; i < nc && !rc; i++) {
    const FlowArgs& a = b->cases[i]->a;
    const size_t qb = (size_t)10 * a.fs * sizeof(double);
    if (hipMalloc(&snap[i], qb) != hipSuccess) rc = fail("hipMalloc failed");
    else if (hipMemcpyAsync(snap[i], a.q, qb, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
      rc = fail("snapshot copy failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail("hipEventCreate failed");
  const StageArgs sa{0, 0, 0, 1.0, 0.0, 1.0};
  for (int sl = 0; sl < 8 && !rc; sl++) {
    const int kind = kSlotKind[sl], d = kSlotDir[sl];
    // classes: cases of this slot with the same program and the same set of
    // fitting shapes; each class gets the fastest common selection
    std::vector<std::pair<std::pair<int, int>, std::vector<int>>> classes;
    for (int i = 0; i < nc; i++) {
      astr_hip_ctx* c = b->cases[i];
      const int prog = slot_prog(kind, d, c->a.ndir);
      if (prog < 0) continue;
      int mask = 0;
      for (int sh = 0; sh < kNumShapes; sh++)
        if (selection_ok(sh, prog, c->a.n[d], c->device)) mask |= 1 << sh;
      bool placed = false;
      for (auto& cl : classes)
        if (cl.first.first == prog && cl.first.second == mask) { cl.second.push_back(i); placed = true; }
      if (!placed) classes.push_back({{prog, mask}, {i}});
    }
    for (auto& cl : classes) {
      if (rc) break;
      const int prog = cl.first.first, mask = cl.first.second;
      const int nm = (int)cl.second.size();
      std::vector<FlowArgs> args(nm);
      for (int k = 0; k < nm; k++) args[k] = b->cases[cl.second[k]]->a;
      FlowArgs* dargs = nullptr;
      int* dts = nullptr;
      if (hipMalloc(&dargs, nm * sizeof(FlowArgs)) != hipSuccess ||
          hipMalloc(&dts, (nm + 1) * sizeof(int)) != hipSuccess ||
          hipMemcpy(dargs, args.data(), nm * sizeof(FlowArgs), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail("autotune set-up failed");
      }
      float best = 1e30f;
      int bsh = -1, bmb = 1;
      for (int sh = 0; sh < kNumShapes && !rc; sh++) {
        if (!((mask >> sh) & 1)) continue;
        std::vector<int> ts(nm + 1, 0);
        for (int k = 0; k < nm; k++) ts[k + 1] = ts[k] + tiles_of(args[k].n, d, kShapes[sh].TL);
        if (hipMemcpy(dts, ts.data(), (nm + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
          rc = fail("autotune copy failed");
          break;
        }
        BatchLaunch bl{dargs, dts, nm, ts[nm]};
        for (int mb = 1; mb <= kMinbVariants && !rc; mb++) {
          const int cap = b->ncu * mb * (kShapes[sh].NT < 256 ? 256 / kShapes[sh].NT : 1);
          const int nb = ((bl.total < cap ? bl.total : cap) + 7) / 8 * 8;
          hipError_t e = launch_batch(sh, prog, mb, bl, d, sa, nb, b->stream);
          if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
          for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_batch(sh, prog, mb, bl, d, sa, nb, b->stream);
          if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) {
            rc = fail(std::string("batch autotune launch: ") + hipGetErrorString(e));
            break;
          }
          if (ms < best) { best = ms; bsh = sh; bmb = mb; }
        }
      }
      (void)hipFree(dargs);
      (void)hipFree(dts);
      if (!rc && bsh >= 0)
        for (int i : cl.second) {
          b->cases[i]->shape[prog][d] = bsh;
          b->cases[i]->minb[prog][d] = bmb;
          b->cases[i]->tuned_ms[prog][d] = best / reps;
        }
    }
  }
  for (int i = 0; i < nc; i++) {
    if (!snap[i]) continue;
    const FlowArgs& a = b->cases[i]->a;
    const size_t qb = (size_t)10 * a.fs * sizeof(double);
    if (!rc && hipMemcpyAsync(a.q, snap[i], qb, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
      rc = fail("restore copy failed");
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess && !rc) rc = fail("autotune sync failed");
  for (double* p : snap) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (build_groups(b) || upload_jobs(b)) return -1;
  return 0;
}

int astr_hip_batch_info(astr_hip_batch* b, int* ngroups, int* launches_per_step) {
  if (!b) return fail("null batch");
  if (ngroups) *ngroups = (int)b->groups.size();
  if (launches_per_step) *launches_per_step = 3 * (int)b->groups.size() + 2;
  return 0;
}

}  // extern "C"
