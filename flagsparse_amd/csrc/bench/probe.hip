// Random-row gather probe (a measurement kernel of libflagsparse_bench.so,
// not of the operator library): n_gathers reads of whole
// rows of row_bytes from a table of n_rows rows, the row ids from a hash, so
// no index stream sits in the dependence chain.  The rate it reaches is the
// practical ceiling of a kernel whose traffic is such gathers -- the CSR SpMM,
// whose B-row reads (one row of N values per nonzero, at uniformly random
// columns) dominate its bytes.
//
// Lane layout: G = row_bytes / 16 lanes read one row with one 16-byte load
// each, 64 / G rows per wave instruction, 8 instructions in flight per lane.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "fsp_bench.h"

namespace fsp {

static thread_local char g_bench_error[256];

static int bench_fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_bench_error, sizeof(g_bench_error), fmt, ap);
  va_end(ap);
  return status;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 8;

template <int G>
__global__ __launch_bounds__(kProbeThreads) void gather_rows_probe_kernel(const float4* __restrict__ table,
                                                                         uint32_t n_rows, int64_t n_gathers,
                                                                         float* __restrict__ sink) {
  constexpr int S = 64 / G;
  const int lane = threadIdx.x & 63;
  const int g = lane % G;
  const int slot = lane / G;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kProbeThreads + threadIdx.x) >> 6;
  const int64_t n_waves = (static_cast<int64_t>(gridDim.x) * kProbeThreads) >> 6;
  float4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t base = wave * S * kProbeUnroll; base < n_gathers; base += n_waves * S * kProbeUnroll) {
    float4 b[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      const int64_t q = base + u * S + slot;
      const uint32_t r = mix32(static_cast<uint32_t>(q) ^ static_cast<uint32_t>(q >> 32)) % n_rows;
      b[u] = q < n_gathers ? table[static_cast<int64_t>(r) * G + g] : float4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      acc.x += b[u].x;
      acc.y += b[u].y;
      acc.z += b[u].z;
      acc.w += b[u].w;
    }
  }
  // written only for a sum no ordinary table produces: the loads stay live
  // without a store per thread
  if (acc.x == 1234567.0f && acc.y == -7654321.0f) sink[0] = acc.z + acc.w;
}

}  // namespace fsp

using namespace fsp;

extern "C" const char* fsp_bench_last_error(void) { return g_bench_error; }

extern "C" int fsp_gather_rows_probe(const void* table_dev, int64_t table_bytes, int64_t n_rows, int32_t row_bytes,
                                     int64_t n_gathers, void* sink_dev, void* stream) {
  if (n_rows <= 0 || n_rows > int64_t(0xffffffff) || n_gathers < 0) {
    return bench_fail(-1, "fsp_gather_rows_probe: n_rows must be in [1, 2^32) and n_gathers >= 0");
  }
  if (row_bytes < 16 || row_bytes > 1024 || (row_bytes & (row_bytes - 1)) != 0) {
    return bench_fail(-1, "fsp_gather_rows_probe: row_bytes %d must be a power of two in [16, 1024]", row_bytes);
  }
  if (table_bytes < n_rows * row_bytes) {
    return bench_fail(-1, "fsp_gather_rows_probe: n_rows x row_bytes = %lld exceeds table_bytes = %lld",
                      (long long)(n_rows * row_bytes), (long long)table_bytes);
  }
  if (table_dev == nullptr || sink_dev == nullptr || reinterpret_cast<uintptr_t>(table_dev) % 16 != 0) {
    return bench_fail(-1, "fsp_gather_rows_probe: NULL or misaligned (16-byte) table / sink");
  }
  if (n_gathers == 0) return 0;
  const int G = row_bytes / 16;
  const int64_t per_block = int64_t(kProbeThreads / 64) * (64 / G) * kProbeUnroll;
  int64_t blocks = (n_gathers + per_block - 1) / per_block;
  if (blocks > 2048) blocks = 2048;  // 8 waves per SIMD on 256 CUs; the loop strides the rest
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>(blocks)), block(kProbeThreads);
  const float4* t = static_cast<const float4*>(table_dev);
  const uint32_t n = static_cast<uint32_t>(n_rows);
  float* sink = static_cast<float*>(sink_dev);
  switch (G) {
    case 1: hipLaunchKernelGGL(gather_rows_probe_kernel<1>, grid, block, 0, s, t, n, n_gathers, sink); break;
    case 2: hipLaunchKernelGGL(gather_rows_probe_kernel<2>, grid, block, 0, s, t, n, n_gathers, sink); break;
    case 4: hipLaunchKernelGGL(gather_rows_probe_kernel<4>, grid, block, 0, s, t, n, n_gathers, sink); break;
    case 8: hipLaunchKernelGGL(gather_rows_probe_kernel<8>, grid, block, 0, s, t, n, n_gathers, sink); break;
    case 16: hipLaunchKernelGGL(gather_rows_probe_kernel<16>, grid, block, 0, s, t, n, n_gathers, sink); break;
    case 32: hipLaunchKernelGGL(gather_rows_probe_kernel<32>, grid, block, 0, s, t, n, n_gathers, sink); break;
    default: hipLaunchKernelGGL(gather_rows_probe_kernel<64>, grid, block, 0, s, t, n, n_gathers, sink); break;
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bench_fail(-3, "gather_rows_probe: %s", hipGetErrorString(e));
}
