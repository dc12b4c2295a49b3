// BSR (16 x 16 blocks) SpMM on the gfx950 matrix cores, plus the CSR -> BSR
// conversion and the operand packing / FP8 quantisation it needs.
//
// Block storage (as the reference's BSR SpMV, spmv_bsr.py:451-452):
// data[nnzb][16][16] row-major inside a block, block columns ascending inside
// a block row, indptr over block rows.
//
// SpMM C[M, N] (f32) = A_bsr B: one workgroup (4 waves) per block row and
// 128-column chunk; each wave owns 32 of the columns.  B is first packed into
// the MFMA operand layout Bp[k / 8][n][k % 8] (8 consecutive B rows of one
// column = one 16-byte operand for bf16/f16, 8 bytes for fp8), so every
// operand is one contiguous load.  bf16 / f16: one v_mfma_f32_16x16x32 per
// block pair and 16-column tile (lanes 0-31 carry the first block, lanes
// 32-63 the second).  FP8 (OCP e4m3): two 16x16x32 fp8 MFMAs per pair, each
// with the other block's half of K zeroed, so each block's product gets its
// own absmax scale before it is accumulated.
#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

constexpr int kBs = 16;         // block size
constexpr int kNt = 8;          // 16-column MFMA tiles per workgroup (128 columns)
constexpr float kE4m3Max = 448.f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static inline int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

// Packed-B column count: whole 128-column chunks, so every tile the SpMM
// kernel reads exists (padding columns are zero).
static inline int64_t bsr_npad(int64_t N) { return ceil_div(N, kNt * 16) * (kNt * 16); }

static inline int grid_1d(int64_t n) {
  int64_t g = ceil_div(n, FSP_WG_THREADS);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return static_cast<int>(g);
}

// ----------------------------------------------------------- OCP E4M3 (fn)
// Round to nearest even; |x| is clamped to 448 (the largest finite value).
__device__ __forceinline__ uint8_t e4m3_encode(float x) {
  const uint8_t sign = signbit(x) ? 0x80 : 0x00;
  const float a = fminf(fabsf(x), kE4m3Max);
  int code;
  if (a < 0.015625f) {  // below 2^-6: subnormal, quantum 2^-9 (m == 8 is the first normal)
    code = static_cast<int>(rintf(a * 512.f));
  } else {
    int e;
    frexpf(a, &e);  // a = f * 2^e, f in [0.5, 1)
    e -= 1;         // floor(log2 a)
    float mant = rintf(ldexpf(a, 3 - e));  // 8 .. 16
    if (mant == 16.f) {
      mant = 8.f;
      e += 1;
    }
    code = ((e + 7) << 3) | (static_cast<int>(mant) - 8);
  }
  return static_cast<uint8_t>(sign | code);
}

// -------------------------------------------------------------- CSR -> BSR
template <typename PT, typename IT>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_keys_kernel(const PT* __restrict__ indptr,
                                                                  const IT* __restrict__ indices, int64_t n_rows,
                                                                  int64_t nbc, uint64_t* __restrict__ keys,
                                                                  int32_t* __restrict__ vals) {
  // one thread per row: key = block_row * nbc + block_col for each of its nonzeros
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; r < n_rows;
       r += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const uint64_t br = static_cast<uint64_t>(r / kBs);
    for (int64_t p = static_cast<int64_t>(indptr[r]); p < static_cast<int64_t>(indptr[r + 1]); ++p) {
      keys[p] = br * static_cast<uint64_t>(nbc) + static_cast<uint64_t>(indices[p] / kBs);
      vals[p] = static_cast<int32_t>(p);
    }
  }
}

// flag[q] = 1 if sorted key q starts a new block.
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_flags_kernel(const uint64_t* __restrict__ skeys, int64_t nnz,
                                                                   int32_t* __restrict__ flags) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q < nnz;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    flags[q] = (q == 0 || skeys[q] != skeys[q - 1]) ? 1 : 0;
  }
}

// nnzb_out = blocks (inclusive scan value at the last position).
__global__ void bsr_count_kernel(const int32_t* __restrict__ blk_of, int64_t nnz, int64_t* __restrict__ nnzb_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) nnzb_out[0] = nnz > 0 ? static_cast<int64_t>(blk_of[nnz - 1]) : 0;
}

// Block metadata from the sorted keys: block_col, and block_ptr via the
// block-row boundaries (block_ptr[b] = first block whose row >= b).
template <typename BP>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_meta_kernel(const uint64_t* __restrict__ skeys,
                                                                  const int32_t* __restrict__ blk_of, int64_t nnz,
                                                                  int64_t nbc, int64_t nbr, int64_t nnzb,
                                                                  int32_t* __restrict__ block_col,
                                                                  BP* __restrict__ block_ptr) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q <= nnz;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const bool start = q < nnz && (q == 0 || skeys[q] != skeys[q - 1]);
    if (start) {
      const int64_t b = blk_of[q] - 1;
      block_col[b] = static_cast<int32_t>(skeys[q] % static_cast<uint64_t>(nbc));
    }
    // block row boundaries: between the previous block start and this one
    if (q == nnz || start) {
      // clamped so that an out-of-range column (rejected by the caller) can
      // never address block_ptr beyond nbr
      const int64_t cur_row =
          q == nnz ? nbr : std::min<int64_t>(nbr, static_cast<int64_t>(skeys[q] / static_cast<uint64_t>(nbc)));
      const int64_t prev_row =
          q == 0 ? -1 : std::min<int64_t>(nbr, static_cast<int64_t>(skeys[q - 1] / static_cast<uint64_t>(nbc)));
      const int64_t bidx = q == nnz ? nnzb : static_cast<int64_t>(blk_of[q] - 1);
      for (int64_t r = prev_row + 1; r <= cur_row; ++r) block_ptr[r] = static_cast<BP>(bidx);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_zero_kernel(T* __restrict__ data, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    data[i] = T(0);
  }
}

// Scatter every nonzero into its block slot (canonical CSR: one writer per slot).
template <typename T, typename PT, typename IT>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_scatter_kernel(const PT* __restrict__ indptr,
                                                                     const IT* __restrict__ indices,
                                                                     const T* __restrict__ data,
                                                                     const int32_t* __restrict__ svals,
                                                                     const int32_t* __restrict__ blk_of,
                                                                     int64_t n_rows, int64_t nnz,
                                                                     T* __restrict__ block_data) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q < nnz;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t p = svals[q];
    // row of nonzero p: largest r with indptr[r] <= p
    int64_t lo = 0, hi = n_rows;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (static_cast<int64_t>(indptr[mid]) <= p) lo = mid; else hi = mid;
    }
    const int64_t b = blk_of[q] - 1;
    const int64_t c = static_cast<int64_t>(indices[p]);
    block_data[(b * kBs + (lo % kBs)) * kBs + (c % kBs)] = data[p];
  }
}

struct bsr_layout {
  int64_t keys, skeys, vals, svals, flags, blk_of, sort_tmp, sort_bytes, scan_tmp, scan_bytes, nnzb, total;
};

static int bsr_workspace_layout(int64_t nnz, int64_t nbr, int64_t nbc, bsr_layout* L) {
  size_t sort_bytes = 0, scan_bytes = 0;
  int key_bits = 1;
  const uint64_t max_key = static_cast<uint64_t>(nbr > 0 ? nbr : 1) * static_cast<uint64_t>(nbc > 0 ? nbc : 1);
  while (key_bits < 64 && (uint64_t(1) << key_bits) < max_key) ++key_bits;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr),
                                           static_cast<uint64_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                           static_cast<int32_t*>(nullptr), static_cast<size_t>(nnz), 0, key_bits,
                                           hipStream_t(0));
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "bsr sort size query: %s", hipGetErrorString(e));
  e = rocprim::inclusive_scan(nullptr, scan_bytes, static_cast<int32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                              static_cast<size_t>(nnz), rocprim::plus<int32_t>(), hipStream_t(0));
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "bsr scan size query: %s", hipGetErrorString(e));
  int64_t off = 0;
  L->keys = off;     off = align256(off + nnz * 8);
  L->skeys = off;    off = align256(off + nnz * 8);
  L->vals = off;     off = align256(off + nnz * 4);
  L->svals = off;    off = align256(off + nnz * 4);
  L->flags = off;    off = align256(off + nnz * 4);
  L->blk_of = off;   off = align256(off + nnz * 4);
  L->sort_tmp = off; off = align256(off + static_cast<int64_t>(sort_bytes));
  L->sort_bytes = static_cast<int64_t>(sort_bytes);
  L->scan_tmp = off; off = align256(off + static_cast<int64_t>(scan_bytes));
  L->scan_bytes = static_cast<int64_t>(scan_bytes);
  L->nnzb = off;     off = align256(off + 8);
  L->total = off + 256;
  return FSP_OK;
}

static int key_bits_for(int64_t nbr, int64_t nbc) {
  const uint64_t max_key = static_cast<uint64_t>(nbr > 0 ? nbr : 1) * static_cast<uint64_t>(nbc > 0 ? nbc : 1);
  int bits = 1;
  while (bits < 64 && (uint64_t(1) << bits) < max_key) ++bits;
  return bits;
}

// --------------------------------------------------------- operand packing
// Bp[k8][n][i] = B[k8 * 8 + i][n] (zero for rows >= K or columns >= N); the
// padded sizes are K8 = ceil(K / 16) * 2 row groups and Npad = bsr_npad(N).
template <typename TI, typename TO>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_pack_b_kernel(const TI* __restrict__ B, int64_t ldb, int64_t K,
                                                                    int64_t N, int64_t K8, int64_t Npad,
                                                                    TO* __restrict__ Bp) {
  const int64_t total = K8 * Npad;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t k8 = t / Npad, n = t - k8 * Npad;
    vec_t<TO, 8> o;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = k8 * 8 + i;
      o.v[i] = (k < K && n < N) ? static_cast<TO>(static_cast<float>(B[k * ldb + n])) : TO(0);
    }
    *reinterpret_cast<vec_t<TO, 8>*>(Bp + t * 8) = o;
  }
}

// FP8 packing: codes of B / b_scale where b_scale = absmax / 448 (read from device).
template <typename TI>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_pack_b_fp8_kernel(const TI* __restrict__ B, int64_t ldb,
                                                                        int64_t K, int64_t N, int64_t K8,
                                                                        int64_t Npad, const float* __restrict__ absmax,
                                                                        uint8_t* __restrict__ Bp) {
  const float am = absmax[0];
  const float inv = am > 0.f ? kE4m3Max / am : 1.f;
  const int64_t total = K8 * Npad;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t k8 = t / Npad, n = t - k8 * Npad;
    uint64_t packed = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = k8 * 8 + i;
      const float v = (k < K && n < N) ? static_cast<float>(B[k * ldb + n]) : 0.f;
      packed |= static_cast<uint64_t>(e4m3_encode(v * inv)) << (8 * i);
    }
    reinterpret_cast<uint64_t*>(Bp)[t] = packed;
  }
}

// |v| as float bits: non-negative floats order like their bit patterns, and
// +Inf (0x7f800000) and NaN (above it) order above every finite value, so an
// integer max over these bits is the absmax and carries a non-finite input
// through (fmaxf would drop a NaN).
template <typename TI>
__device__ __forceinline__ unsigned int abs_bits(TI v) {
  return __float_as_uint(static_cast<float>(v)) & 0x7fffffffu;
}

constexpr int kAbsmaxUnroll = 4;
constexpr int kAbsmaxMaxBlocks = 1024;  // 4 workgroups per CU on 256 CUs; the loops stride the rest

// absmax bits over a [rows, cols] (row stride ld) array: a bounded grid,
// kAbsmaxUnroll independent loads in flight per thread, wave then workgroup
// max in registers / LDS, ONE atomic per workgroup.
template <typename TI>
__global__ __launch_bounds__(FSP_WG_THREADS) void absmax_kernel(const TI* __restrict__ x, int64_t rows, int64_t cols,
                                                                int64_t ld, unsigned int* __restrict__ out) {
  __shared__ unsigned int wave_max[FSP_WAVES];
  unsigned int m = 0u;
  if (ld == cols) {  // one flat range
    const int64_t total = rows * cols;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS;
    int64_t t = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x;
    for (; t + (kAbsmaxUnroll - 1) * stride < total; t += kAbsmaxUnroll * stride) {
      unsigned int a[kAbsmaxUnroll];
#pragma unroll
      for (int u = 0; u < kAbsmaxUnroll; ++u) a[u] = abs_bits(x[t + u * stride]);
#pragma unroll
      for (int u = 0; u < kAbsmaxUnroll; ++u) m = a[u] > m ? a[u] : m;
    }
    for (; t < total; t += stride) {
      const unsigned int a = abs_bits(x[t]);
      m = a > m ? a : m;
    }
  } else {  // strided rows: workgroups over rows, threads over columns
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
      for (int64_t c = threadIdx.x; c < cols; c += FSP_WG_THREADS) {
        const unsigned int a = abs_bits(x[r * ld + c]);
        m = a > m ? a : m;
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned int o = static_cast<unsigned int>(__shfl_xor(static_cast<int>(m), off));
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 1; w < FSP_WAVES; ++w) m = wave_max[w] > m ? wave_max[w] : m;
    atomicMax(out, m);
  }
}

__global__ void zero_u32_kernel(unsigned int* p) { p[0] = 0u; }

// Per-block absmax scaling to E4M3: one wave per block (4 values per lane).
template <typename TI>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_quantize_fp8_kernel(const TI* __restrict__ data, int64_t nnzb,
                                                                          uint8_t* __restrict__ q,
                                                                          float* __restrict__ scales) {
  const int lane = threadIdx.x & 63;
  for (int64_t b = (static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x) >> 6; b < nnzb;
       b += (static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) >> 6) {
    float v[4];
    unsigned int mb = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = static_cast<float>(data[b * 256 + lane * 4 + i]);
      const unsigned int a = abs_bits(v[i]);
      mb = a > mb ? a : mb;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const unsigned int o = static_cast<unsigned int>(__shfl_xor(static_cast<int>(mb), off));
      mb = o > mb ? o : mb;
    }
    const float m = __uint_as_float(mb);
    // a block holding Inf / NaN gets a NaN scale: its products come out NaN
    const float scale = mb >= 0x7f800000u ? __builtin_nanf("") : (m > 0.f ? m / kE4m3Max : 1.f);
    uint32_t packed = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) packed |= static_cast<uint32_t>(e4m3_encode(v[i] / scale)) << (8 * i);
    reinterpret_cast<uint32_t*>(q + b * 256)[lane] = packed;
    if (lane == 0) scales[b] = scale;
  }
}

// ------------------------------------------------------------ BSR SpMM
// KIND 0 = bf16, 1 = f16, 2 = fp8 (codes + per-block scales).
// One workgroup per (block row, 128 columns); wave w owns the two 16-column
// tiles 2w, 2w+1 and walks every block pair of the block row (lanes 0-31 carry
// block j, 32-63 block j+1), so each C element is summed by one lane in block
// order (deterministic, no cross-wave reduction, no LDS).  The next pair's A
// and B fragments are loaded while the current pair's MFMAs run.
// WPG = waves per workgroup; each wave owns kNt / WPG tiles.  Every wave
// re-reads the block row's A fragments, so fewer, wider waves pay off when B
// is cheap to move (fp8: 2 waves x 64 columns; bf16 / f16: 4 x 32; measured
// on MI355X).
template <int KIND, typename BP, int WPG>
__global__ __launch_bounds__(64 * WPG) void bsr_spmm_kernel(
    const BP* __restrict__ block_ptr, const int32_t* __restrict__ block_col, const void* __restrict__ block_data,
    const float* __restrict__ block_scales, const void* __restrict__ Bp, const float* __restrict__ b_absmax,
    int64_t nbc, int64_t Npad, int64_t N, float* __restrict__ C, int64_t ldc, int64_t M) {
  using V = typename std::conditional<KIND == 0, bf16x8, f16x8>::type;
  using E = typename std::conditional<KIND == 0, __bf16, _Float16>::type;
  using AF = typename std::conditional<KIND == 2, uint64_t, V>::type;  // one lane's A (or B) fragment
  constexpr int kTilesPerWave = kNt / WPG;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t br = blockIdx.x;
  const int64_t nbase = static_cast<int64_t>(blockIdx.y) * (kNt * 16) + wave * (kTilesPerWave * 16);
  const int64_t beg = static_cast<int64_t>(block_ptr[br]);
  const int64_t end = static_cast<int64_t>(block_ptr[br + 1]);
  const int half = lane >> 5;      // which block of the pair this lane carries
  const int r = lane & 15;         // A row / B column inside the tile
  const int kh = (lane >> 4) & 1;  // which 8-wide half of the block's K
  const int64_t last = end > beg ? end - 1 : beg;  // a stored block (clamped loads)

  struct frag {
    AF a;
    AF b[kTilesPerWave];
    bool valid;
  };
  // Fragments of pair j (unconditional, clamped loads; invalid lanes zeroed).
  auto load = [&](int64_t j) -> frag {
    frag f;
    const int64_t blk = j + half;
    const int64_t bc = blk < end ? blk : last;
    const int64_t col = static_cast<int64_t>(block_col[bc]);
    f.valid = blk < end && col >= 0 && col < nbc;  // an out-of-range column reads nothing useful
    const int64_t k8 = (f.valid ? col : 0) * 2 + kh;
    if constexpr (KIND == 2) {
      const uint8_t* A8 = static_cast<const uint8_t*>(block_data);
      const uint64_t* B8 = static_cast<const uint64_t*>(Bp);
      const uint64_t a = *reinterpret_cast<const uint64_t*>(A8 + bc * 256 + r * 16 + kh * 8);
      f.a = f.valid ? a : 0ull;
#pragma unroll
      for (int t = 0; t < kTilesPerWave; ++t) {
        const uint64_t b = B8[k8 * Npad + nbase + t * 16 + r];
        f.b[t] = f.valid ? b : 0ull;
      }
    } else {
      const E* Ae = static_cast<const E*>(block_data);
      const E* Be = static_cast<const E*>(Bp);
      const V a = *reinterpret_cast<const V*>(Ae + bc * 256 + r * 16 + kh * 8);
      f.a = f.valid ? a : V{};
#pragma unroll
      for (int t = 0; t < kTilesPerWave; ++t) {
        const V b = *reinterpret_cast<const V*>(Be + (k8 * Npad + nbase + t * 16 + r) * 8);
        f.b[t] = f.valid ? b : V{};
      }
    }
    return f;
  };

  f32x4 acc[kTilesPerWave];
#pragma unroll
  for (int t = 0; t < kTilesPerWave; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (beg < end) {
    frag cur = load(beg);
    for (int64_t j = beg; j < end; j += 2) {  // wave-uniform
      const frag nxt = load(j + 2 < end ? j + 2 : j);  // prefetch (a repeat of j past the end)
      if constexpr (KIND == 2) {
        const float s0 = block_scales[j];
        const float s1 = (j + 1 < end) ? block_scales[j + 1] : 0.f;
        const long a_lo = half == 0 ? static_cast<long>(cur.a) : 0l;
        const long a_hi = half == 1 ? static_cast<long>(cur.a) : 0l;
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          const f32x4 z = {0.f, 0.f, 0.f, 0.f};
          const f32x4 p0 = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a_lo, static_cast<long>(cur.b[t]), z, 0, 0, 0);
          const f32x4 p1 = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a_hi, static_cast<long>(cur.b[t]), z, 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 4; ++i) acc[t][i] += s0 * p0[i] + s1 * p1[i];
        }
      } else {
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if constexpr (KIND == 0) {
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.a, cur.b[t], acc[t], 0, 0, 0);
          } else {
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(cur.a, cur.b[t], acc[t], 0, 0, 0);
          }
        }
      }
      cur = nxt;
    }
  }
  float bscale = 1.f;
  if constexpr (KIND == 2) {
    // B holding Inf / NaN (absmax not finite) makes the whole product NaN
    const float am = b_absmax[0];
    bscale = __builtin_isfinite(am) ? (am > 0.f ? am / kE4m3Max : 1.f) : __builtin_nanf("");
  }
  // D layout: lane l holds rows (l / 16) * 4 + i, column l % 16 of a tile.
#pragma unroll
  for (int t = 0; t < kTilesPerWave; ++t) {
    const int64_t gcol = nbase + t * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t grow = br * kBs + (lane >> 4) * 4 + i;
      if (grow < M && gcol < N) C[grow * ldc + gcol] = acc[t][i] * bscale;
    }
  }
}

}  // namespace fsp

using namespace fsp;

extern "C" {

int fsp_csr_to_bsr_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz, int64_t* bytes) {
  if (bytes == nullptr || n_rows < 0 || n_cols < 0 || nnz < 0) return fail(FSP_ERR_ARG, "fsp_csr_to_bsr_workspace: bad arguments");
  if (nnz > INT32_MAX) return fail(FSP_ERR_ARG, "fsp_csr_to_bsr_workspace: nnz exceeds int32 positions");
  bsr_layout L;
  const int st = bsr_workspace_layout(nnz, ceil_div(n_rows, kBs), ceil_div(n_cols, kBs), &L);
  if (st) return st;
  *bytes = L.total;
  return FSP_OK;
}

int fsp_csr_to_bsr_count(int32_t indptr_type, int32_t index_type, const void* indptr_dev, const void* indices_dev,
                         int64_t n_rows, int64_t n_cols, int64_t nnz, void* ws_dev, int64_t ws_bytes,
                         int64_t* nnzb_dev, void* stream) {
  if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnzb_dev == nullptr || (nnz > 0 && (indptr_dev == nullptr || indices_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_csr_to_bsr_count: bad arguments");
  }
  if (nnz > INT32_MAX) return fail(FSP_ERR_ARG, "fsp_csr_to_bsr_count: nnz exceeds int32 positions");
  const int64_t nbr = ceil_div(n_rows, kBs), nbc = ceil_div(n_cols, kBs);
  bsr_layout L;
  int st = bsr_workspace_layout(nnz, nbr, nbc, &L);
  if (st) return st;
  if (ws_dev == nullptr || ws_bytes < L.total) return fail(FSP_ERR_WORKSPACE, "fsp_csr_to_bsr_count: workspace too small");
  const hipStream_t s = as_stream(stream);
  char* base = reinterpret_cast<char*>(align256(reinterpret_cast<intptr_t>(ws_dev)));
  uint64_t* keys = reinterpret_cast<uint64_t*>(base + L.keys);
  uint64_t* skeys = reinterpret_cast<uint64_t*>(base + L.skeys);
  int32_t* vals = reinterpret_cast<int32_t*>(base + L.vals);
  int32_t* svals = reinterpret_cast<int32_t*>(base + L.svals);
  int32_t* flags = reinterpret_cast<int32_t*>(base + L.flags);
  int32_t* blk_of = reinterpret_cast<int32_t*>(base + L.blk_of);
  if (nnz == 0) {
    hipLaunchKernelGGL(bsr_zero_kernel<int64_t>, dim3(1), dim3(FSP_WG_THREADS), 0, s, nnzb_dev, int64_t(1));
    return check_launch("bsr_count_empty");
  }
  FSP_DISPATCH_INDEX(indptr_type, PT, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      hipLaunchKernelGGL((bsr_keys_kernel<PT, IT>), dim3(grid_1d(n_rows)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<const PT*>(indptr_dev), static_cast<const IT*>(indices_dev), n_rows, nbc, keys,
                         vals);
      st = check_launch("bsr_keys");
      if (st) return st;
    });
  });
  size_t tb = static_cast<size_t>(L.sort_bytes);
  hipError_t e = rocprim::radix_sort_pairs(base + L.sort_tmp, tb, keys, skeys, vals, svals, static_cast<size_t>(nnz),
                                           0, key_bits_for(nbr, nbc), s);
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "bsr sort: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(bsr_flags_kernel, dim3(grid_1d(nnz)), dim3(FSP_WG_THREADS), 0, s, skeys, nnz, flags);
  st = check_launch("bsr_flags");
  if (st) return st;
  size_t sb = static_cast<size_t>(L.scan_bytes);
  e = rocprim::inclusive_scan(base + L.scan_tmp, sb, flags, blk_of, static_cast<size_t>(nnz), rocprim::plus<int32_t>(), s);
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "bsr scan: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(bsr_count_kernel, dim3(1), dim3(64), 0, s, blk_of, nnz, nnzb_dev);
  return check_launch("bsr_count");
}

int fsp_csr_to_bsr_fill(int32_t value_type, int32_t indptr_type, int32_t index_type, const void* indptr_dev,
                        const void* indices_dev, const void* data_dev, int64_t n_rows, int64_t n_cols, int64_t nnz,
                        int64_t nnzb, int32_t block_ptr_type, void* block_ptr_dev, int32_t* block_col_dev,
                        void* block_data_dev, void* ws_dev, int64_t ws_bytes, void* stream) {
  if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnzb < 0 || block_ptr_dev == nullptr ||
      (nnzb > 0 && (block_col_dev == nullptr || block_data_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_csr_to_bsr_fill: bad arguments");
  }
  const int64_t nbr = ceil_div(n_rows, kBs), nbc = ceil_div(n_cols, kBs);
  bsr_layout L;
  int st = bsr_workspace_layout(nnz, nbr, nbc, &L);
  if (st) return st;
  if (ws_dev == nullptr || ws_bytes < L.total) return fail(FSP_ERR_WORKSPACE, "fsp_csr_to_bsr_fill: workspace too small");
  const hipStream_t s = as_stream(stream);
  char* base = reinterpret_cast<char*>(align256(reinterpret_cast<intptr_t>(ws_dev)));
  const uint64_t* skeys = reinterpret_cast<const uint64_t*>(base + L.skeys);
  const int32_t* svals = reinterpret_cast<const int32_t*>(base + L.svals);
  const int32_t* blk_of = reinterpret_cast<const int32_t*>(base + L.blk_of);
  FSP_DISPATCH_INDEX(block_ptr_type, BP, {
    hipLaunchKernelGGL((bsr_meta_kernel<BP>), dim3(grid_1d(nnz + 1)), dim3(FSP_WG_THREADS), 0, s, skeys, blk_of, nnz,
                       nbc, nbr, nnzb, block_col_dev, static_cast<BP*>(block_ptr_dev));
    st = check_launch("bsr_meta");
    if (st) return st;
  });
  if (nnzb == 0) return FSP_OK;
  FSP_DISPATCH_VALUE(value_type, T, {
    hipLaunchKernelGGL(bsr_zero_kernel<T>, dim3(grid_1d(nnzb * 256)), dim3(FSP_WG_THREADS), 0, s,
                       static_cast<T*>(block_data_dev), nnzb * 256);
    st = check_launch("bsr_zero");
    if (st) return st;
    FSP_DISPATCH_INDEX(indptr_type, PT, {
      FSP_DISPATCH_INDEX(index_type, IT, {
        hipLaunchKernelGGL((bsr_scatter_kernel<T, PT, IT>), dim3(grid_1d(nnz)), dim3(FSP_WG_THREADS), 0, s,
                           static_cast<const PT*>(indptr_dev), static_cast<const IT*>(indices_dev),
                           static_cast<const T*>(data_dev), svals, blk_of, n_rows, nnz, static_cast<T*>(block_data_dev));
        st = check_launch("bsr_scatter");
        if (st) return st;
      });
    });
  });
  return FSP_OK;
}

int fsp_absmax(int32_t value_type, const void* x_dev, int64_t rows, int64_t cols, int64_t ld, float* out_dev,
               void* stream) {
  if (rows < 0 || cols < 0 || ld < cols || out_dev == nullptr || (rows * cols > 0 && x_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_absmax: bad arguments");
  }
  const hipStream_t s = as_stream(stream);
  unsigned int* o = reinterpret_cast<unsigned int*>(out_dev);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, o);
  if (rows * cols == 0) return check_launch("absmax_zero");
  int64_t blocks = ld == cols ? ceil_div(rows * cols, int64_t(FSP_WG_THREADS) * kAbsmaxUnroll) : rows;
  if (blocks > kAbsmaxMaxBlocks) blocks = kAbsmaxMaxBlocks;
  FSP_DISPATCH_VALUE(value_type, T, {
    hipLaunchKernelGGL(absmax_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(FSP_WG_THREADS), 0, s,
                       static_cast<const T*>(x_dev), rows, cols, ld, o);
  });
  return check_launch("absmax");
}

int fsp_bsr_pack_b(int32_t in_type, int32_t out_type, const void* B_dev, int64_t ldb, int64_t K, int64_t N,
                   const float* absmax_dev, void* Bp_dev, void* stream) {
  if (K < 0 || N < 0 || ldb < N || Bp_dev == nullptr || (K * N > 0 && B_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_bsr_pack_b: bad arguments");
  }
  const int64_t K8 = ceil_div(K, kBs) * 2, Npad = bsr_npad(N);
  const hipStream_t s = as_stream(stream);
  const int64_t total = K8 * Npad;
  FSP_DISPATCH_VALUE(in_type, TI, {
    if (out_type == FSP_BF16) {
      hipLaunchKernelGGL((bsr_pack_b_kernel<TI, __bf16>), dim3(grid_1d(total)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<const TI*>(B_dev), ldb, K, N, K8, Npad, static_cast<__bf16*>(Bp_dev));
    } else if (out_type == FSP_F16) {
      hipLaunchKernelGGL((bsr_pack_b_kernel<TI, _Float16>), dim3(grid_1d(total)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<const TI*>(B_dev), ldb, K, N, K8, Npad, static_cast<_Float16*>(Bp_dev));
    } else if (out_type == FSP_F8E4M3) {
      if (absmax_dev == nullptr) return fail(FSP_ERR_ARG, "fsp_bsr_pack_b: fp8 packing needs absmax");
      hipLaunchKernelGGL((bsr_pack_b_fp8_kernel<TI>), dim3(grid_1d(total)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<const TI*>(B_dev), ldb, K, N, K8, Npad, absmax_dev,
                         static_cast<uint8_t*>(Bp_dev));
    } else {
      return fail(FSP_ERR_DTYPE, "fsp_bsr_pack_b: unsupported packed type %d", (int)out_type);
    }
  });
  return check_launch("bsr_pack_b");
}

int fsp_bsr_quantize_fp8(int32_t value_type, const void* block_data_dev, int64_t nnzb, uint8_t* q_dev,
                         float* scales_dev, void* stream) {
  if (nnzb < 0 || (nnzb > 0 && (block_data_dev == nullptr || q_dev == nullptr || scales_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_bsr_quantize_fp8: bad arguments");
  }
  if (nnzb == 0) return FSP_OK;
  const hipStream_t s = as_stream(stream);
  int64_t g = ceil_div(nnzb, FSP_WAVES);
  if (g > 65536) g = 65536;
  FSP_DISPATCH_VALUE(value_type, T, {
    hipLaunchKernelGGL(bsr_quantize_fp8_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                       static_cast<const T*>(block_data_dev), nnzb, q_dev, scales_dev);
  });
  return check_launch("bsr_quantize_fp8");
}

int fsp_bsr_spmm(int32_t value_type, int32_t block_ptr_type, const void* block_ptr_dev, const int32_t* block_col_dev,
                 const void* block_data_dev, const float* block_scales_dev, int64_t n_block_rows, int64_t nnzb,
                 const void* Bp_dev, const float* b_absmax_dev, int64_t K, int64_t N, float* C_dev, int64_t ldc,
                 int64_t M, void* stream) {
  if (n_block_rows < 0 || nnzb < 0 || K < 0 || N < 0 || M < 0 || ldc < N || block_ptr_dev == nullptr ||
      C_dev == nullptr || M > n_block_rows * kBs) {
    return fail(FSP_ERR_ARG, "fsp_bsr_spmm: bad arguments");
  }
  if (n_block_rows == 0 || N == 0) return FSP_OK;
  if (nnzb > 0 && (block_col_dev == nullptr || block_data_dev == nullptr || Bp_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_bsr_spmm: NULL operand");
  }
  if (value_type == FSP_F8E4M3 && (block_scales_dev == nullptr || b_absmax_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_bsr_spmm: fp8 needs block scales and the B absmax");
  }
  const int64_t Npad = bsr_npad(N);  // the layout fsp_bsr_pack_b wrote
  const int64_t nbc = ceil_div(K, kBs);
  const int64_t ychunks = ceil_div(N, kNt * 16);
  if (n_block_rows > INT32_MAX || ychunks > 65535) return fail(FSP_ERR_ARG, "fsp_bsr_spmm: grid too large");
  const hipStream_t s = as_stream(stream);
  const dim3 grid(static_cast<unsigned>(n_block_rows), static_cast<unsigned>(ychunks));
  FSP_DISPATCH_INDEX(block_ptr_type, BP, {
    const BP* bp = static_cast<const BP*>(block_ptr_dev);
    if (value_type == FSP_BF16) {
      hipLaunchKernelGGL((bsr_spmm_kernel<0, BP, 4>), grid, dim3(64 * 4), 0, s, bp, block_col_dev, block_data_dev,
                         nullptr, Bp_dev, nullptr, nbc, Npad, N, C_dev, ldc, M);
    } else if (value_type == FSP_F16) {
      hipLaunchKernelGGL((bsr_spmm_kernel<1, BP, 4>), grid, dim3(64 * 4), 0, s, bp, block_col_dev, block_data_dev,
                         nullptr, Bp_dev, nullptr, nbc, Npad, N, C_dev, ldc, M);
    } else if (value_type == FSP_F8E4M3) {
      hipLaunchKernelGGL((bsr_spmm_kernel<2, BP, 2>), grid, dim3(64 * 2), 0, s, bp, block_col_dev, block_data_dev,
                         block_scales_dev, Bp_dev, b_absmax_dev, nbc, Npad, N, C_dev, ldc, M);
    } else {
      return fail(FSP_ERR_DTYPE, "fsp_bsr_spmm: value type %d has no MFMA path", (int)value_type);
    }
  });
  return check_launch("bsr_spmm");
}

}  // extern "C"
