// Column-slab schedule probe of the CSR SpMM (a measurement kernel of
// libflagsparse_bench.so, not of the operator library).
//
// C = A B with B row-major [K, 64] float32 is computed in S column slabs of B
// (slab k holds columns [k K / S, (k+1) K / S) of A, i.e. those rows of B).
// Each lane group of 16 lanes (4 output columns per lane) owns RPG virtual
// rows -- a CSR row of at most 256 nonzeros, or a 256-nonzero segment of a
// longer one -- and keeps their accumulators in registers while it sweeps the
// slabs in order, walking each row's nonzeros of the current slab (the
// caller's per-row slab boundary table).  With sync, the workgroups that share
// blockIdx % 8 (one XCD under round-robin dispatch) pass a phase counter
// (vector atomics) after every slab, so the waves resident on one XCD read
// one slab of B -- K / S rows, 4 MiB at S = 16 for the 262144 x 64 headline
// operand -- together and that XCD's L2 can serve it.  The launch is
// cooperative (every workgroup resident) and every wait is bounded: past
// spin_limit polls a workgroup sets *flag and goes on, so the kernel always
// ends.  Regular rows are written to C, segments to partials.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fsp_bench.h"

namespace fsp_bench {

template <int RPG>
__global__ __launch_bounds__(1024) void slab_probe_kernel(
    const int32_t* __restrict__ vbeg, const uint16_t* __restrict__ slab_end, const int32_t* __restrict__ vslot,
    const int32_t* __restrict__ indices, const float* __restrict__ data, const float4* __restrict__ B,
    float4* __restrict__ C, float4* __restrict__ partials, int32_t n_rows, int32_t n_virtual, int32_t S,
    int32_t passes, int32_t* __restrict__ counters, int32_t spin_limit, int32_t* __restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int gl = lane & 15;
  const int g = blockIdx.x * 64 + wave * 4 + (lane >> 4);
  const int groups_total = gridDim.x * 64;
  const int grp = blockIdx.x & 7;
  const int wgs_in_grp = (gridDim.x - grp + 7) / 8;
  int phase = 0;
  for (int pass = 0; pass < passes; ++pass) {
    float4 acc[RPG];
    int cur[RPG], base[RPG], vr[RPG];
#pragma unroll
    for (int r = 0; r < RPG; ++r) {
      const int v = (pass * RPG + r) * groups_total + g;
      vr[r] = v < n_virtual ? v : -1;
      base[r] = v < n_virtual ? vbeg[v] : 0;
      cur[r] = base[r];
      acc[r] = float4{0.f, 0.f, 0.f, 0.f};
    }
    for (int k = 0; k < S; ++k) {
      int endk[RPG];
#pragma unroll
      for (int r = 0; r < RPG; ++r) {
        endk[r] = vr[r] >= 0 ? base[r] + static_cast<int>(slab_end[static_cast<int64_t>(vr[r]) * S + k]) : 0;
      }
      while (true) {  // one nonzero per row and round; exits together (wave vote)
        bool ok[RPG];
        int col[RPG];
        float val[RPG];
        bool any = false;
#pragma unroll
        for (int r = 0; r < RPG; ++r) {
          ok[r] = cur[r] < endk[r];
          any = any || ok[r];
          const int p = ok[r] ? cur[r] : 0;
          col[r] = indices[p];
          val[r] = data[p];
        }
        if (!__any(any)) break;
        float4 b[RPG];
#pragma unroll
        for (int r = 0; r < RPG; ++r) b[r] = B[static_cast<int64_t>(ok[r] ? col[r] : 0) * 16 + gl];
#pragma unroll
        for (int r = 0; r < RPG; ++r) {
          const float a = ok[r] ? val[r] : 0.f;
          acc[r].x += a * b[r].x;
          acc[r].y += a * b[r].y;
          acc[r].z += a * b[r].z;
          acc[r].w += a * b[r].w;
          cur[r] += ok[r] ? 1 : 0;
        }
      }
      if (spin_limit >= 0) {  // uniform: a kernel argument
        __syncthreads();
        ++phase;
        if (threadIdx.x == 0) {
          __hip_atomic_fetch_add(&counters[grp], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          const int target = phase * wgs_in_grp;
          int spins = 0;
          while (__hip_atomic_load(&counters[grp], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > spin_limit) {
              __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        __syncthreads();
      }
    }
#pragma unroll
    for (int r = 0; r < RPG; ++r) {
      if (vr[r] >= 0) {
        const int slot = vslot[vr[r]];
        if (slot < n_rows) {
          C[static_cast<int64_t>(slot) * 16 + gl] = acc[r];
        } else {
          partials[static_cast<int64_t>(slot - n_rows) * 16 + gl] = acc[r];
        }
      }
    }
  }
}

__global__ void zero_counters_kernel(int32_t* p) {
  if (threadIdx.x < 8) p[threadIdx.x] = 0;
}

}  // namespace fsp_bench

extern "C" int fsp_slab_spmm_probe(int32_t rows_per_group, const int32_t* vbeg_dev, const uint16_t* slab_end_dev,
                                   const int32_t* vslot_dev, const int32_t* indices_dev, const float* data_dev,
                                   const float* B_dev, float* C_dev, float* partials_dev, int32_t n_rows,
                                   int32_t n_virtual, int32_t n_slabs, int32_t n_workgroups, int32_t* counters_dev,
                                   int32_t spin_limit, int32_t* flag_dev, void* stream) {
  using namespace fsp_bench;
  if (n_rows <= 0 || n_virtual <= 0 || n_slabs <= 0 || n_workgroups <= 0 || counters_dev == nullptr ||
      flag_dev == nullptr || vbeg_dev == nullptr || slab_end_dev == nullptr || vslot_dev == nullptr ||
      indices_dev == nullptr || data_dev == nullptr || B_dev == nullptr || C_dev == nullptr ||
      (rows_per_group != 4 && rows_per_group != 8)) {
    return -1;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters_dev);
  const int32_t groups_total = n_workgroups * 64;
  const int32_t passes = (n_virtual + groups_total * rows_per_group - 1) / (groups_total * rows_per_group);
  const float4* B4 = reinterpret_cast<const float4*>(B_dev);
  float4* C4 = reinterpret_cast<float4*>(C_dev);
  float4* P4 = reinterpret_cast<float4*>(partials_dev);
  void* args[] = {(void*)&vbeg_dev, (void*)&slab_end_dev, (void*)&vslot_dev, (void*)&indices_dev, (void*)&data_dev,
                  (void*)&B4, (void*)&C4, (void*)&P4, (void*)&n_rows, (void*)&n_virtual, (void*)&n_slabs,
                  (void*)&passes, (void*)&counters_dev, (void*)&spin_limit, (void*)&flag_dev};
  const void* fn = rows_per_group == 4 ? reinterpret_cast<const void*>(slab_probe_kernel<4>)
                                       : reinterpret_cast<const void*>(slab_probe_kernel<8>);
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(n_workgroups), dim3(1024), args, 0, s);
  return e == hipSuccess ? 0 : -3;
}
