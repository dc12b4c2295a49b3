// Streaming-read probe (a measurement kernel of libflagsparse_bench.so, not
// of the operator library): reads n_arrays (1 or 2) arrays of n_chunks
// 16-byte chunks each, in lock step -- chunk i of every array by the same
// lane, with the nontemporal hint, four chunks per array in flight per lane
// -- the access pattern of the vector-chunk SpMV's column-id and value
// streams without its row pointers, x gathers or y writes.  Workgroup b of
// the (persistent) grid reads either a contiguous range of chunks
// (contiguous = 1, as the SpMV's tile ranges) or chunks b, b + grid, ...
// (contiguous = 0).  Nothing is written except, for a sum no ordinary data
// produces, one float at sink.  bytes / time is the practical read rate of
// that pattern.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fsp_bench.h"

namespace fsp_bench {

constexpr int kStreamThreads = 1024;
constexpr int kStreamUnroll = 4;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int NA>
__global__ __launch_bounds__(kStreamThreads) void stream_read_kernel(const u32x4* __restrict__ a,
                                                                     const u32x4* __restrict__ b, int64_t n,
                                                                     int contiguous, float* __restrict__ sink) {
  const int64_t nthreads = static_cast<int64_t>(gridDim.x) * kStreamThreads;
  int64_t first, step, end;
  if (contiguous) {
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    first = static_cast<int64_t>(blockIdx.x) * per + threadIdx.x;
    end = first - threadIdx.x + per < n ? first - threadIdx.x + per : n;
    step = kStreamThreads;
  } else {
    first = static_cast<int64_t>(blockIdx.x) * kStreamThreads + threadIdx.x;
    end = n;
    step = nthreads;
  }
  uint32_t acc = 0;
  int64_t i = first;
  for (; i + (kStreamUnroll - 1) * step < end; i += kStreamUnroll * step) {
    u32x4 va[kStreamUnroll], vb[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
      va[u] = __builtin_nontemporal_load(a + i + u * step);
      if (NA == 2) vb[u] = __builtin_nontemporal_load(b + i + u * step);
    }
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
      acc ^= va[u].x ^ va[u].y ^ va[u].z ^ va[u].w;
      if (NA == 2) acc ^= vb[u].x ^ vb[u].y ^ vb[u].z ^ vb[u].w;
    }
  }
  for (; i < end; i += step) {
    const u32x4 v = __builtin_nontemporal_load(a + i);
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
    if (NA == 2) {
      const u32x4 w = __builtin_nontemporal_load(b + i);
      acc ^= w.x ^ w.y ^ w.z ^ w.w;
    }
  }
  if (acc == 0x9E3779B9u) sink[0] = static_cast<float>(acc);
}

}  // namespace fsp_bench

extern "C" int fsp_stream_read_probe(const void* a_dev, const void* b_dev, int64_t n_chunks, int32_t n_arrays,
                                     int32_t n_workgroups, int32_t contiguous, void* sink_dev, void* stream) {
  using namespace fsp_bench;
  if (a_dev == nullptr || sink_dev == nullptr || n_chunks < 0 || n_workgroups <= 0 ||
      (n_arrays != 1 && n_arrays != 2) || (n_arrays == 2 && b_dev == nullptr) ||
      reinterpret_cast<uintptr_t>(a_dev) % 16 != 0 || reinterpret_cast<uintptr_t>(b_dev) % 16 != 0) {
    return -1;
  }
  if (n_chunks == 0) return 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const u32x4* a = static_cast<const u32x4*>(a_dev);
  const u32x4* b = static_cast<const u32x4*>(b_dev);
  float* sink = static_cast<float*>(sink_dev);
  if (n_arrays == 1) {
    hipLaunchKernelGGL(stream_read_kernel<1>, dim3(n_workgroups), dim3(kStreamThreads), 0, s, a, a, n_chunks,
                       contiguous, sink);
  } else {
    hipLaunchKernelGGL(stream_read_kernel<2>, dim3(n_workgroups), dim3(kStreamThreads), 0, s, a, b, n_chunks,
                       contiguous, sink);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
