// Plan building for the CSR kernels: the library plumbing (error message,
// version, arch query), the stable row-length binning, and the CSR -> CSC
// permutation (stable radix sort of the column keys) with its value gather.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <mutex>
#include <string>

#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error.assign(buf);
}

static inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// ------------------------------------------------------------------ binning
constexpr int kBinItems = 4;                                  // rows per thread
constexpr int kBinRowsPerBlock = FSP_WG_THREADS * kBinItems;  // 1024

struct bin_bounds {
  int64_t upper[FSP_MAX_BINS];
};

__device__ __forceinline__ int bin_of(int64_t len, const bin_bounds& bb, int n_bins) {
  int b = 0;
  while (b < n_bins - 1 && len > bb.upper[b]) ++b;
  return b;
}

// counts[b * n_blocks + blk] = rows of block blk that fall in bin b.
template <typename PT>
__global__ __launch_bounds__(FSP_WG_THREADS) void bins_count_kernel(
    const PT* __restrict__ indptr, int64_t n_rows, int n_bins, bin_bounds bb,
    int64_t* __restrict__ counts, int64_t n_blocks) {
  __shared__ int s_cnt[FSP_MAX_BINS];
  if (threadIdx.x < FSP_MAX_BINS) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kBinRowsPerBlock;
  for (int k = 0; k < kBinItems; ++k) {
    const int64_t r = base + k * FSP_WG_THREADS + threadIdx.x;
    if (r < n_rows) {
      const int64_t len = static_cast<int64_t>(indptr[r + 1]) - static_cast<int64_t>(indptr[r]);
      atomicAdd(&s_cnt[bin_of(len, bb, n_bins)], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < n_bins) counts[threadIdx.x * n_blocks + blockIdx.x] = s_cnt[threadIdx.x];
}

// In-place exclusive scan of counts[0..n) by one workgroup; bin_offsets[b] is
// the scanned value at the start of bin b, bin_offsets[n_bins] = n_rows.
__global__ __launch_bounds__(FSP_WG_THREADS) void bins_scan_kernel(
    int64_t* __restrict__ counts, int64_t n_blocks, int n_bins, int64_t n_rows,
    int64_t* __restrict__ bin_offsets) {
  __shared__ int64_t s_part[FSP_WG_THREADS];
  __shared__ int64_t s_carry;
  const int64_t n = n_blocks * n_bins;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t chunk = 0; chunk < n; chunk += FSP_WG_THREADS) {
    const int64_t i = chunk + threadIdx.x;
    const int64_t v = i < n ? counts[i] : 0;
    s_part[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 entries.
    for (int off = 1; off < FSP_WG_THREADS; off <<= 1) {
      const int64_t add = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
      __syncthreads();
      s_part[threadIdx.x] += add;
      __syncthreads();
    }
    const int64_t carry = s_carry;
    const int64_t excl = carry + s_part[threadIdx.x] - v;
    if (i < n) {
      counts[i] = excl;
      if (i % n_blocks == 0) bin_offsets[i / n_blocks] = excl;
    }
    __syncthreads();
    if (threadIdx.x == FSP_WG_THREADS - 1) s_carry = carry + s_part[FSP_WG_THREADS - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) bin_offsets[n_bins] = n_rows;
}

// Stable scatter: rows keep ascending order inside each bin.
template <typename PT>
__global__ __launch_bounds__(FSP_WG_THREADS) void bins_scatter_kernel(
    const PT* __restrict__ indptr, int64_t n_rows, int n_bins, bin_bounds bb,
    const int64_t* __restrict__ offsets, int64_t n_blocks, int32_t* __restrict__ rows_out) {
  __shared__ int s_wave_cnt[FSP_WAVES][FSP_MAX_BINS];
  __shared__ int64_t s_base[FSP_MAX_BINS];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x < FSP_MAX_BINS) {
    s_base[threadIdx.x] =
        threadIdx.x < n_bins ? offsets[threadIdx.x * n_blocks + blockIdx.x] : 0;
  }
  __syncthreads();
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kBinRowsPerBlock;
  for (int k = 0; k < kBinItems; ++k) {
    const int64_t r = base + k * FSP_WG_THREADS + threadIdx.x;
    const bool valid = r < n_rows;
    int b = -1;
    if (valid) {
      const int64_t len = static_cast<int64_t>(indptr[r + 1]) - static_cast<int64_t>(indptr[r]);
      b = bin_of(len, bb, n_bins);
    }
    int my_rank = 0;
    for (int j = 0; j < n_bins; ++j) {
      const uint64_t m = __ballot(b == j);
      if (b == j) my_rank = __popcll(m & lt_mask);
      if (lane == 0) s_wave_cnt[wave][j] = __popcll(m);
    }
    __syncthreads();
    if (valid) {
      int64_t pos = s_base[b] + my_rank;
      for (int w = 0; w < wave; ++w) pos += s_wave_cnt[w][b];
      rows_out[pos] = static_cast<int32_t>(r);
    }
    __syncthreads();
    if (threadIdx.x < n_bins) {
      int tot = 0;
      for (int w = 0; w < FSP_WAVES; ++w) tot += s_wave_cnt[w][threadIdx.x];
      s_base[threadIdx.x] += tot;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- CSR -> CSC
template <typename IT, typename KT, typename QT>
__global__ __launch_bounds__(FSP_WG_THREADS) void csc_init_kernel(
    const IT* __restrict__ indices, int64_t nnz, KT* __restrict__ keys, QT* __restrict__ vals) {
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; p < nnz;
       p += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    keys[p] = static_cast<KT>(indices[p]);
    vals[p] = static_cast<QT>(p);
  }
}

// colptr[c] = first CSC position whose key >= c (keys sorted ascending).
template <typename KT, typename CT>
__global__ __launch_bounds__(FSP_WG_THREADS) void csc_colptr_kernel(
    const KT* __restrict__ keys, int64_t nnz, int64_t n_cols, CT* __restrict__ colptr) {
  const int64_t total = nnz + 1;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q < total;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t prev = q == 0 ? -1 : static_cast<int64_t>(keys[q - 1]);
    const int64_t cur = q == nnz ? n_cols : static_cast<int64_t>(keys[q]);
    for (int64_t c = prev + 1; c <= cur; ++c) colptr[c] = static_cast<CT>(q);
  }
}

// Largest r in [lo, hi) with indptr[r] <= p (indptr[lo] <= p guaranteed).
template <typename PT>
__device__ __forceinline__ int64_t row_search(const PT* __restrict__ indptr, int64_t lo, int64_t hi, int64_t p) {
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (static_cast<int64_t>(indptr[mid]) <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// row_of[p] = row r with indptr[r] <= p < indptr[r+1].  A workgroup covers
// kExpandNnz consecutive nonzeros: it bounds their rows with two searches over
// the whole indptr, then every thread searches only that (small, cached) range.
constexpr int kExpandNnz = 1024;
template <typename PT, typename RT>
__global__ __launch_bounds__(FSP_WG_THREADS) void csc_row_of_kernel(const PT* __restrict__ indptr, int64_t n_rows,
                                                                    int64_t nnz, RT* __restrict__ row_of) {
  __shared__ int64_t s_lo, s_hi;
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kExpandNnz;
  const int64_t b1 = b0 + kExpandNnz < nnz ? b0 + kExpandNnz : nnz;
  if (threadIdx.x == 0) {
    s_lo = row_search(indptr, 0, n_rows, b0);
    s_hi = row_search(indptr, 0, n_rows, b1 - 1) + 1;
  }
  __syncthreads();
  const int64_t lo = s_lo, hi = s_hi;
  for (int64_t p = b0 + threadIdx.x; p < b1; p += FSP_WG_THREADS) {
    row_of[p] = static_cast<RT>(row_search(indptr, lo, hi, p));
  }
}

// rowind[q] = row_of[perm[q]].
template <typename QT, typename RT>
__global__ __launch_bounds__(FSP_WG_THREADS) void csc_rowind_kernel(const RT* __restrict__ row_of,
                                                                    const QT* __restrict__ perm, int64_t nnz,
                                                                    RT* __restrict__ rowind) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q < nnz;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    rowind[q] = row_of[static_cast<int64_t>(perm[q])];
  }
}

__global__ __launch_bounds__(FSP_WG_THREADS) void fill_i64_kernel(int64_t* __restrict__ p, int64_t n,
                                                                 int64_t v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    p[i] = v;
  }
}

template <int BYTES> struct blob { alignas(BYTES < 16 ? BYTES : 16) unsigned char b[BYTES]; };

template <int BYTES, typename QT>
__global__ __launch_bounds__(FSP_WG_THREADS) void gather_kernel(
    const QT* __restrict__ perm, const blob<BYTES>* __restrict__ in, blob<BYTES>* __restrict__ out,
    int64_t n) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q < n;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    out[q] = in[static_cast<int64_t>(perm[q])];
  }
}

static inline int grid_for(int64_t n) {
  int64_t g = ceil_div(n, FSP_WG_THREADS);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return static_cast<int>(g);
}

static inline int key_bits(int64_t n_cols) {
  int bits = 0;
  while (bits < 63 && (int64_t(1) << bits) < n_cols) ++bits;
  return bits < 1 ? 1 : bits;
}

template <typename KT, typename QT>
static int csc_sort_bytes(int64_t nnz, int64_t n_cols, size_t* bytes) {
  *bytes = 0;
  const hipError_t e = rocprim::radix_sort_pairs(
      nullptr, *bytes, static_cast<KT*>(nullptr), static_cast<KT*>(nullptr),
      static_cast<QT*>(nullptr), static_cast<QT*>(nullptr), static_cast<size_t>(nnz), 0,
      key_bits(n_cols), hipStream_t(0));
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "radix sort size query: %s", hipGetErrorString(e));
  return FSP_OK;
}

struct csc_layout {
  int64_t keys_in, keys_out, vals_in, row_of, sort_tmp, sort_bytes, total;
};

template <typename KT, typename QT>
static int csc_workspace_layout(int64_t nnz, int64_t n_cols, csc_layout* L) {
  size_t sort_bytes = 0;
  int st = csc_sort_bytes<KT, QT>(nnz, n_cols, &sort_bytes);
  if (st != FSP_OK) return st;
  int64_t off = 0;
  L->keys_in = off;  off = align_up(off + nnz * (int64_t)sizeof(KT), 256);
  L->keys_out = off; off = align_up(off + nnz * (int64_t)sizeof(KT), 256);
  L->vals_in = off;  off = align_up(off + nnz * (int64_t)sizeof(QT), 256);
  L->row_of = off;   off = align_up(off + nnz * (int64_t)sizeof(int64_t), 256);
  L->sort_tmp = off; off = align_up(off + (int64_t)sort_bytes, 256);
  L->sort_bytes = (int64_t)sort_bytes;
  L->total = off + 256;
  return FSP_OK;
}

template <typename F>
static int with_key_perm(int64_t n_cols, int32_t perm_type, F&& f) {
  const bool wide_key = n_cols > (int64_t)UINT32_MAX;
  if (perm_type == FSP_I32) {
    return wide_key ? f(uint64_t{}, int32_t{}) : f(uint32_t{}, int32_t{});
  }
  if (perm_type == FSP_I64) {
    return wide_key ? f(uint64_t{}, int64_t{}) : f(uint32_t{}, int64_t{});
  }
  return fail(FSP_ERR_DTYPE, "unsupported perm type %d", (int)perm_type);
}

// ---------------------------------------------------------------- inspect
// Validation figures of a CSR in one pass, for one host readback:
// stats[0] = indptr[0], stats[1] = indptr[n_rows], stats[2] = number of rows
// with indptr[i+1] < indptr[i], stats[3] / stats[4] = smallest / largest
// column index (INT64_MAX / INT64_MIN without nonzeros), stats[5] = longest
// row.  min / max / counts are order-independent, so the result is
// deterministic although waves combine with atomics.
constexpr int64_t kStatsMin = INT64_MAX;
constexpr int64_t kStatsMax = INT64_MIN;

__global__ __launch_bounds__(64) void stats_init_kernel(int64_t* __restrict__ stats) {
  if (threadIdx.x == 0) {
    stats[2] = 0;
    stats[3] = kStatsMin;
    stats[4] = kStatsMax;
    stats[5] = 0;
  }
}

__device__ __forceinline__ int64_t wave_min(int64_t v) {
  for (int m = 32; m > 0; m >>= 1) {
    const int64_t o = __shfl_xor(v, m);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
  for (int m = 32; m > 0; m >>= 1) {
    const int64_t o = __shfl_xor(v, m);
    v = o > v ? o : v;
  }
  return v;
}

template <typename PT, typename IT>
__global__ __launch_bounds__(FSP_WG_THREADS) void csr_stats_kernel(const PT* __restrict__ indptr,
                                                                   const IT* __restrict__ indices, int64_t n_rows,
                                                                   int64_t nnz, int64_t* __restrict__ stats) {
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS;
  if (tid == 0) {
    stats[0] = static_cast<int64_t>(indptr[0]);
    stats[1] = static_cast<int64_t>(indptr[n_rows]);
  }
  int64_t dec = 0, longest = 0, lo = kStatsMin, hi = kStatsMax;
  for (int64_t i = tid; i < n_rows; i += stride) {
    const int64_t len = static_cast<int64_t>(indptr[i + 1]) - static_cast<int64_t>(indptr[i]);
    dec += len < 0 ? 1 : 0;
    longest = len > longest ? len : longest;
  }
  for (int64_t p = tid; p < nnz; p += stride) {
    const int64_t c = static_cast<int64_t>(indices[p]);
    lo = c < lo ? c : lo;
    hi = c > hi ? c : hi;
  }
  for (int m = 32; m > 0; m >>= 1) dec += __shfl_xor(dec, m);
  longest = wave_max(longest);
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) {
    if (dec) __hip_atomic_fetch_add(&stats[2], dec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lo != kStatsMin) __hip_atomic_fetch_min(&stats[3], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (hi != kStatsMax) __hip_atomic_fetch_max(&stats[4], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (longest) __hip_atomic_fetch_max(&stats[5], longest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename QT>
static int launch_gather(int32_t elem_bytes, const QT* perm, const void* in, void* out, int64_t n, hipStream_t s) {
  const dim3 grid(grid_for(n)), block(FSP_WG_THREADS);
  switch (elem_bytes) {
    case 1: hipLaunchKernelGGL((gather_kernel<1, QT>), grid, block, 0, s, perm, static_cast<const blob<1>*>(in), static_cast<blob<1>*>(out), n); break;
    case 2: hipLaunchKernelGGL((gather_kernel<2, QT>), grid, block, 0, s, perm, static_cast<const blob<2>*>(in), static_cast<blob<2>*>(out), n); break;
    case 4: hipLaunchKernelGGL((gather_kernel<4, QT>), grid, block, 0, s, perm, static_cast<const blob<4>*>(in), static_cast<blob<4>*>(out), n); break;
    case 8: hipLaunchKernelGGL((gather_kernel<8, QT>), grid, block, 0, s, perm, static_cast<const blob<8>*>(in), static_cast<blob<8>*>(out), n); break;
    case 16: hipLaunchKernelGGL((gather_kernel<16, QT>), grid, block, 0, s, perm, static_cast<const blob<16>*>(in), static_cast<blob<16>*>(out), n); break;
    default: return fail(FSP_ERR_ARG, "fsp_gather: unsupported element size %d", (int)elem_bytes);
  }
  return check_launch("gather");
}

}  // namespace fsp

using namespace fsp;

extern "C" {

int fsp_abi_version(void) { return FSP_ABI_VERSION; }

const char* fsp_last_error(void) { return g_last_error.c_str(); }

int fsp_device_arch(char* buf, int64_t buf_len) {
  if (buf == nullptr || buf_len <= 0) return fail(FSP_ERR_ARG, "fsp_device_arch: empty buffer");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "hipGetDevice: %s", hipGetErrorString(e));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  snprintf(buf, static_cast<size_t>(buf_len), "%s", prop.gcnArchName);
  return FSP_OK;
}

int fsp_csr_inspect(int32_t indptr_type, const void* indptr_dev, int32_t index_type, const void* indices_dev,
                    int64_t n_rows, int64_t nnz, int64_t* stats_dev, void* stream) {
  if (n_rows < 0 || nnz < 0 || indptr_dev == nullptr || stats_dev == nullptr ||
      (nnz > 0 && indices_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_csr_inspect: bad arguments");
  }
  const hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(stats_init_kernel, dim3(1), dim3(64), 0, s, stats_dev);
  int st = check_launch("csr_stats_init");
  if (st) return st;
  const int64_t work = n_rows > nnz ? n_rows : nnz;
  int64_t g = ceil_div(work, FSP_WG_THREADS * 8);
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  FSP_DISPATCH_INDEX(indptr_type, PT, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      hipLaunchKernelGGL((csr_stats_kernel<PT, IT>), dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<const PT*>(indptr_dev), static_cast<const IT*>(indices_dev), n_rows, nnz,
                         stats_dev);
      return check_launch("csr_stats");
    });
  });
  return FSP_OK;
}

int fsp_csr_row_bins_workspace(int64_t n_rows, int32_t n_bins, int64_t* bytes) {
  if (bytes == nullptr || n_rows < 0 || n_bins < 1 || n_bins > FSP_MAX_BINS) {
    return fail(FSP_ERR_ARG, "fsp_csr_row_bins_workspace: bad arguments");
  }
  const int64_t n_blocks = ceil_div(n_rows, kBinRowsPerBlock) > 0 ? ceil_div(n_rows, kBinRowsPerBlock) : 1;
  *bytes = align_up(n_blocks * n_bins * (int64_t)sizeof(int64_t), 256);
  return FSP_OK;
}

int fsp_csr_row_bins(int32_t indptr_type, const void* indptr_dev, int64_t n_rows, int32_t n_bins,
                     const int64_t* upper, int32_t* rows_out_dev, int64_t* bin_offsets_out_dev,
                     void* ws_dev, int64_t ws_bytes, void* stream) {
  if (n_rows < 0 || n_bins < 1 || n_bins > FSP_MAX_BINS || upper == nullptr ||
      bin_offsets_out_dev == nullptr || (n_rows > 0 && (indptr_dev == nullptr || rows_out_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_csr_row_bins: bad arguments");
  }
  if (n_rows > (int64_t)INT32_MAX) return fail(FSP_ERR_ARG, "fsp_csr_row_bins: n_rows exceeds int32 row ids");
  int64_t need = 0;
  fsp_csr_row_bins_workspace(n_rows, n_bins, &need);
  if (ws_bytes < need || ws_dev == nullptr) {
    return fail(FSP_ERR_WORKSPACE, "fsp_csr_row_bins: workspace %lld < %lld", (long long)ws_bytes, (long long)need);
  }
  bin_bounds bb;
  for (int b = 0; b < FSP_MAX_BINS; ++b) bb.upper[b] = INT64_MAX;
  for (int b = 0; b + 1 < n_bins; ++b) bb.upper[b] = upper[b];
  const hipStream_t s = as_stream(stream);
  const int64_t n_blocks = ceil_div(n_rows, kBinRowsPerBlock) > 0 ? ceil_div(n_rows, kBinRowsPerBlock) : 1;
  int64_t* counts = static_cast<int64_t*>(ws_dev);
  if (n_rows == 0) {
    hipLaunchKernelGGL(fill_i64_kernel, dim3(1), dim3(FSP_WG_THREADS), 0, s, bin_offsets_out_dev,
                       int64_t(n_bins + 1), int64_t(0));
    return check_launch("bins_empty");
  }
  FSP_DISPATCH_INDEX(indptr_type, PT, {
    const PT* ip = static_cast<const PT*>(indptr_dev);
    hipLaunchKernelGGL(bins_count_kernel<PT>, dim3(n_blocks), dim3(FSP_WG_THREADS), 0, s, ip, n_rows,
                       n_bins, bb, counts, n_blocks);
    int st = check_launch("bins_count");
    if (st) return st;
    hipLaunchKernelGGL(bins_scan_kernel, dim3(1), dim3(FSP_WG_THREADS), 0, s, counts, n_blocks, n_bins,
                       n_rows, bin_offsets_out_dev);
    st = check_launch("bins_scan");
    if (st) return st;
    hipLaunchKernelGGL(bins_scatter_kernel<PT>, dim3(n_blocks), dim3(FSP_WG_THREADS), 0, s, ip, n_rows,
                       n_bins, bb, counts, n_blocks, rows_out_dev);
    st = check_launch("bins_scatter");
    if (st) return st;
  });
  return FSP_OK;
}

int fsp_csr_to_csc_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz, int32_t perm_type,
                             int64_t* bytes) {
  (void)n_rows;
  if (bytes == nullptr || nnz < 0 || n_cols < 0) return fail(FSP_ERR_ARG, "fsp_csr_to_csc_workspace: bad arguments");
  return with_key_perm(n_cols, perm_type, [&](auto k, auto q) -> int {
    using KT = decltype(k);
    using QT = decltype(q);
    csc_layout L;
    int st = csc_workspace_layout<KT, QT>(nnz, n_cols, &L);
    if (st) return st;
    *bytes = L.total;
    return FSP_OK;
  });
}

int fsp_csr_to_csc(int32_t indptr_type, int32_t index_type, const void* indptr_dev, const void* indices_dev,
                   int64_t n_rows, int64_t n_cols, int64_t nnz, int32_t colptr_type, void* colptr_dev,
                   int32_t rowind_type, void* rowind_dev, int32_t perm_type, void* perm_dev, void* ws_dev,
                   int64_t ws_bytes, void* stream) {
  if (n_rows < 0 || n_cols < 0 || nnz < 0 || colptr_dev == nullptr ||
      (nnz > 0 && (indptr_dev == nullptr || indices_dev == nullptr || rowind_dev == nullptr || perm_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_csr_to_csc: bad arguments");
  }
  if (perm_type == FSP_I32 && nnz > (int64_t)INT32_MAX) return fail(FSP_ERR_ARG, "fsp_csr_to_csc: int32 perm cannot address nnz");
  const hipStream_t s = as_stream(stream);
  return with_key_perm(n_cols, perm_type, [&](auto k, auto q) -> int {
    using KT = decltype(k);
    using QT = decltype(q);
    csc_layout L;
    int st = csc_workspace_layout<KT, QT>(nnz, n_cols, &L);
    if (st) return st;
    if (ws_bytes < L.total || ws_dev == nullptr) {
      return fail(FSP_ERR_WORKSPACE, "fsp_csr_to_csc: workspace %lld < %lld", (long long)ws_bytes, (long long)L.total);
    }
    char* base = reinterpret_cast<char*>(align_up(reinterpret_cast<intptr_t>(ws_dev), 256));
    KT* keys_in = reinterpret_cast<KT*>(base + L.keys_in);
    KT* keys_out = reinterpret_cast<KT*>(base + L.keys_out);
    QT* vals_in = reinterpret_cast<QT*>(base + L.vals_in);
    void* sort_tmp = base + L.sort_tmp;
    QT* perm = static_cast<QT*>(perm_dev);
    FSP_DISPATCH_INDEX(index_type, IT, {
      if (nnz > 0) {
        hipLaunchKernelGGL((csc_init_kernel<IT, KT, QT>), dim3(grid_for(nnz)), dim3(FSP_WG_THREADS), 0, s,
                           static_cast<const IT*>(indices_dev), nnz, keys_in, vals_in);
        st = check_launch("csc_init");
        if (st) return st;
        size_t tb = static_cast<size_t>(L.sort_bytes);
        const hipError_t e = rocprim::radix_sort_pairs(sort_tmp, tb, keys_in, keys_out, vals_in, perm,
                                                       static_cast<size_t>(nnz), 0, key_bits(n_cols), s);
        if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "radix sort: %s", hipGetErrorString(e));
      }
    });
    FSP_DISPATCH_INDEX(colptr_type, CT, {
      hipLaunchKernelGGL((csc_colptr_kernel<KT, CT>), dim3(grid_for(nnz + 1)), dim3(FSP_WG_THREADS), 0, s,
                         keys_out, nnz, n_cols, static_cast<CT*>(colptr_dev));
      st = check_launch("csc_colptr");
      if (st) return st;
    });
    if (nnz > 0) {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        FSP_DISPATCH_INDEX(rowind_type, RT, {
          RT* row_of = reinterpret_cast<RT*>(base + L.row_of);
          const int64_t nblk = ceil_div(nnz, kExpandNnz);
          if (nblk > INT32_MAX) return fail(FSP_ERR_ARG, "fsp_csr_to_csc: nnz too large");
          hipLaunchKernelGGL((csc_row_of_kernel<PT, RT>), dim3(static_cast<unsigned>(nblk)), dim3(FSP_WG_THREADS),
                             0, s, static_cast<const PT*>(indptr_dev), n_rows, nnz, row_of);
          st = check_launch("csc_row_of");
          if (st) return st;
          hipLaunchKernelGGL((csc_rowind_kernel<QT, RT>), dim3(grid_for(nnz)), dim3(FSP_WG_THREADS), 0, s, row_of,
                             perm, nnz, static_cast<RT*>(rowind_dev));
          st = check_launch("csc_rowind");
          if (st) return st;
        });
      });
    }
    return FSP_OK;
  });
}

int fsp_gather(int32_t elem_bytes, int32_t perm_type, const void* perm_dev, const void* in_dev, void* out_dev,
               int64_t nnz, void* stream) {
  if (nnz < 0 || (nnz > 0 && (perm_dev == nullptr || in_dev == nullptr || out_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_gather: bad arguments");
  }
  if (nnz == 0) return FSP_OK;
  const hipStream_t s = as_stream(stream);
  if (perm_type == FSP_I32) return launch_gather(elem_bytes, static_cast<const int32_t*>(perm_dev), in_dev, out_dev, nnz, s);
  if (perm_type == FSP_I64) return launch_gather(elem_bytes, static_cast<const int64_t*>(perm_dev), in_dev, out_dev, nnz, s);
  return fail(FSP_ERR_DTYPE, "fsp_gather: unsupported perm type %d", (int)perm_type);
}

}  // extern "C"
