// probe.hip -- the HBM streaming probe: a copy kernel whose bytes are known
// exactly (n doubles read, n doubles written), used as the achievable-
// bandwidth reference of this GPU and to calibrate the FETCH_SIZE /
// WRITE_SIZE counters.  No reference counterpart.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "astr_hip.h"

namespace {

thread_local std::string g_err;

int pfail(const std::string& m) {
  g_err = m;
  return 1;
}

// W doubles per lane per access, as one (W = 1), one 16-byte (W = 2) or two
// 16-byte (W = 4) loads; consecutive lanes take consecutive W-double
// segments, so a wave reads 64 * W * 8 contiguous bytes.  Every block covers
// kUnroll such wave-wide strips per lane, strided by the block's width.
constexpr int kBlock = 256;
constexpr int kUnroll = 4;
typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int W>
__global__ __launch_bounds__(kBlock) void k_stream_copy(const double* __restrict__ src,
                                                        double* __restrict__ dst, long long nvec) {
  using V = typename std::conditional<W == 1, double, dvec2>::type;
  constexpr int VPL = W == 4 ? 2 : 1;  // vectors per lane per strip
  const V* s = reinterpret_cast<const V*>(src);
  V* d = reinterpret_cast<V*>(dst);
  const long long base = (long long)blockIdx.x * kBlock * kUnroll * VPL;
  V r[kUnroll * VPL];
#pragma unroll
  for (int u = 0; u < kUnroll; u++)
#pragma unroll
    for (int v = 0; v < VPL; v++) {
      const long long e = base + ((long long)u * kBlock + threadIdx.x) * VPL + v;
      if (e < nvec) r[u * VPL + v] = __builtin_nontemporal_load(s + e);
    }
#pragma unroll
  for (int u = 0; u < kUnroll; u++)
#pragma unroll
    for (int v = 0; v < VPL; v++) {
      const long long e = base + ((long long)u * kBlock + threadIdx.x) * VPL + v;
      if (e < nvec) __builtin_nontemporal_store(r[u * VPL + v], d + e);
    }
}

}  // namespace

extern "C" {

const char* astr_hip_probe_last_error(void) { return g_err.c_str(); }

int astr_hip_stream_probe(const double* src, double* dst, long long n, int width, void* stream) {
  if (!src || !dst) return pfail("stream_probe: null buffer");
  if (n <= 0) return pfail("stream_probe: n must be positive");
  if (width != 1 && width != 2 && width != 4) return pfail("stream_probe: width must be 1, 2 or 4");
  if (width > 1 && (n % 2 != 0 || ((size_t)src | (size_t)dst) % 16 != 0))
    return pfail("stream_probe: widths 2 and 4 need an even n and 16-byte aligned buffers");
  const long long nvec = width == 1 ? n : n / 2;
  const long long per_block = (long long)kBlock * kUnroll * (width == 4 ? 2 : 1);
  const long long blocks = (nvec + per_block - 1) / per_block;
  if (blocks > 0x7fffffffLL) return pfail("stream_probe: n too large");
  const hipStream_t st = (hipStream_t)stream;
  if (width == 1)
    hipLaunchKernelGGL(k_stream_copy<1>, dim3((unsigned)blocks), dim3(kBlock), 0, st, src, dst, nvec);
  else if (width == 2)
    hipLaunchKernelGGL(k_stream_copy<2>, dim3((unsigned)blocks), dim3(kBlock), 0, st, src, dst, nvec);
  else
    hipLaunchKernelGGL(k_stream_copy<4>, dim3((unsigned)blocks), dim3(kBlock), 0, st, src, dst, nvec);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pfail(std::string("stream_probe: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
