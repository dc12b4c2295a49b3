// Device-side building blocks shared by every flagsparse_amd kernel (gfx950).
//
// Wave64 throughout: a workgroup is FSP_WG_THREADS threads = FSP_WAVES waves of
// 64 lanes.  Cross-lane sums use DPP inside a 16-lane row (quad_perm, row
// half-mirror, row mirror), ds_swizzle across the two rows of a 32-lane half
// and ds_bpermute across the two halves.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define FSP_WG_THREADS 256
#define FSP_WAVES (FSP_WG_THREADS / 64)

namespace fsp {

// ---------------------------------------------------------------- value types
// Storage types and their accumulation types.  f16 / bf16 use clang's native
// _Float16 / __bf16 (hardware conversions on gfx950).
template <typename T> struct acc_of { using type = float; };
template <> struct acc_of<double> { using type = double; };

// SpMV of f32 accumulates in f64 (the reference promotes f32 SpMV to f64).
template <typename T> struct spmv_acc_of { using type = float; };
template <> struct spmv_acc_of<float> { using type = double; };
template <> struct spmv_acc_of<double> { using type = double; };

// One rounding per term: c + a * b as a fused multiply-add, written out so
// every loop variant accumulates identically whatever the contraction mode.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename A, typename T> __device__ __forceinline__ A to_acc(T v) {
  return static_cast<A>(v);
}
template <typename T, typename A> __device__ __forceinline__ T from_acc(A v) {
  return static_cast<T>(v);
}

// ------------------------------------------------------------ vector loads
// VEC consecutive elements of T loaded / stored as one aligned access of
// VEC*sizeof(T) bytes (<= 16).
template <typename T, int VEC> struct alignas(sizeof(T) * VEC) vec_t {
  T v[VEC];
};

template <typename T, int VEC>
__device__ __forceinline__ vec_t<T, VEC> load_vec(const T* p) {
  return *reinterpret_cast<const vec_t<T, VEC>*>(p);
}
template <typename T, int VEC>
__device__ __forceinline__ void store_vec(T* p, const vec_t<T, VEC>& v) {
  *reinterpret_cast<vec_t<T, VEC>*>(p) = v;
}

// ---------------------------------------------------------- cross-lane moves
// DPP control codes (GCN/CDNA ISA): quad_perm selectors, row mirrors.
constexpr int DPP_QUAD_XOR1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

template <int CTRL> __device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = dpp_i32<CTRL>(static_cast<int>(static_cast<uint32_t>(u)));
  const int hi = dpp_i32<CTRL>(static_cast<int>(static_cast<uint32_t>(u >> 32)));
  return __builtin_bit_cast(
      double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                  static_cast<uint32_t>(lo));
}

// Lane exchange with lane ^ mask (mask < 32) through ds_swizzle bit mode:
// offset = xor_mask << 10 | or_mask << 5 | and_mask, and_mask = 0x1F.
__device__ __forceinline__ int swz_xor_i32(int v, int mask) {
  // ds_swizzle needs an immediate pattern; dispatch the handful of masks used.
  switch (mask) {
    case 1: return __builtin_amdgcn_ds_swizzle(v, (1 << 10) | 0x1F);
    case 2: return __builtin_amdgcn_ds_swizzle(v, (2 << 10) | 0x1F);
    case 4: return __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1F);
    case 8: return __builtin_amdgcn_ds_swizzle(v, (8 << 10) | 0x1F);
    default: return __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1F);
  }
}
// Lane exchange with lane ^ mask for any mask < 64 through ds_bpermute.
__device__ __forceinline__ int bperm_xor_i32(int v, int mask) {
  const int lane = __lane_id();
  return __builtin_amdgcn_ds_bpermute((lane ^ mask) << 2, v);
}

__device__ __forceinline__ int xor_i32(int v, int mask) {
  return mask < 32 ? swz_xor_i32(v, mask) : bperm_xor_i32(v, mask);
}
__device__ __forceinline__ float shfl_xor(float v, int mask) {
  return __builtin_bit_cast(float, xor_i32(__builtin_bit_cast(int, v), mask));
}
__device__ __forceinline__ double shfl_xor(double v, int mask) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = xor_i32(static_cast<int>(static_cast<uint32_t>(u)), mask);
  const int hi = xor_i32(static_cast<int>(static_cast<uint32_t>(u >> 32)), mask);
  return __builtin_bit_cast(
      double, (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                  static_cast<uint32_t>(lo));
}

// Sum over aligned groups of `lanes` consecutive lanes (power of two, 1..64);
// every lane of a group receives the group total, bitwise identical across
// the group (each step adds two partial sums that are exchanged pairwise).
// `lanes` is wave-uniform, so the guards are uniform branches.
template <typename A> __device__ __forceinline__ A group_sum(A v, int lanes) {
  if (lanes >= 2) v += dpp<DPP_QUAD_XOR1>(v);
  if (lanes >= 4) v += dpp<DPP_QUAD_XOR2>(v);
  if (lanes >= 8) v += dpp<DPP_ROW_HALF_MIRROR>(v);
  if (lanes >= 16) v += dpp<DPP_ROW_MIRROR>(v);
  if (lanes >= 32) v += shfl_xor(v, 16);
  if (lanes >= 64) v += shfl_xor(v, 32);
  return v;
}

// Byte-offset addressing: 32-bit offsets from a uniform base when OFF32.
template <bool OFF32> struct addr;
// (char-pointer arithmetic keeps the global address space visible to the
// compiler, so the loads stay global_load with a scalar base.)
template <> struct addr<true> {
  using off_t = uint32_t;
  template <typename P> static __device__ __forceinline__ P* at(P* base, off_t byte_off) {
    using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
    return reinterpret_cast<P*>(reinterpret_cast<C*>(base) + static_cast<uint64_t>(byte_off));
  }
};
template <> struct addr<false> {
  using off_t = uint64_t;
  template <typename P> static __device__ __forceinline__ P* at(P* base, off_t byte_off) {
    using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
    return reinterpret_cast<P*>(reinterpret_cast<C*>(base) + byte_off);
  }
};

template <bool OFF32, typename P>
__device__ __forceinline__ const P& elem(const P* base, typename addr<OFF32>::off_t i) {
  return *addr<OFF32>::at(base, i * static_cast<typename addr<OFF32>::off_t>(sizeof(P)));
}

// ------------------------------------------------------------ index helpers
template <typename I> __device__ __forceinline__ int64_t ld_idx(const I* p, int64_t i) {
  return static_cast<int64_t>(p[i]);
}

}  // namespace fsp
