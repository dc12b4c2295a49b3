// COO -> CSR on the GPU: a stable LSD radix sort of the (row * n_cols + col)
// keys with the nonzero positions as values, then the row pointer from the
// sorted keys (first position of each row) and the column of every sorted
// key.  Duplicate (row, col) entries are kept, in input order (the sort is
// stable).  The caller gathers the values through the returned permutation.
// Row / column ids outside the shape set bits of *err (1: row, 2: column);
// every write stays inside the outputs whatever the ids are.
#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

constexpr unsigned kCooErrRow = 1u;
constexpr unsigned kCooErrCol = 2u;

static inline int64_t coo_align(int64_t v) { return (v + 255) / 256 * 256; }

static inline int coo_grid(int64_t n) {
  int64_t g = ceil_div(n, FSP_WG_THREADS);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return static_cast<int>(g);
}

static inline int coo_key_bits(int64_t n_rows, int64_t n_cols) {
  const uint64_t span = static_cast<uint64_t>(n_rows > 0 ? n_rows : 1) * static_cast<uint64_t>(n_cols > 0 ? n_cols : 1);
  int bits = 1;
  while (bits < 64 && (uint64_t(1) << bits) < span) ++bits;
  return bits;
}

__global__ void coo_zero_err_kernel(unsigned int* err) { err[0] = 0u; }

// key = row * n_cols + col (ids out of range are clamped into the key space
// and reported), value = position.
template <typename IT, typename QT>
__global__ __launch_bounds__(FSP_WG_THREADS) void coo_keys_kernel(const IT* __restrict__ row, const IT* __restrict__ col,
                                                                  int64_t nnz, int64_t n_rows, int64_t n_cols,
                                                                  uint64_t* __restrict__ keys, QT* __restrict__ vals,
                                                                  unsigned int* __restrict__ err) {
  unsigned int bad = 0u;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; p < nnz;
       p += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    int64_t r = static_cast<int64_t>(row[p]);
    int64_t c = static_cast<int64_t>(col[p]);
    if (r < 0 || r >= n_rows) {
      bad |= kCooErrRow;
      r = r < 0 ? 0 : n_rows - 1;
    }
    if (c < 0 || c >= n_cols) {
      bad |= kCooErrCol;
      c = c < 0 ? 0 : n_cols - 1;
    }
    keys[p] = static_cast<uint64_t>(r) * static_cast<uint64_t>(n_cols) + static_cast<uint64_t>(c);
    vals[p] = static_cast<QT>(p);
  }
  if (bad) atomicOr(err, bad);
}

// indptr[r] = first sorted position whose row >= r; cols[q] = column of key q.
template <typename PT, typename OT>
__global__ __launch_bounds__(FSP_WG_THREADS) void coo_rows_kernel(const uint64_t* __restrict__ skeys, int64_t nnz,
                                                                  int64_t n_rows, int64_t n_cols,
                                                                  PT* __restrict__ indptr, OT* __restrict__ cols) {
  const uint64_t nc = static_cast<uint64_t>(n_cols);
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; q <= nnz;
       q += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t prev = q == 0 ? -1 : std::min<int64_t>(n_rows, static_cast<int64_t>(skeys[q - 1] / nc));
    const int64_t cur = q == nnz ? n_rows : std::min<int64_t>(n_rows, static_cast<int64_t>(skeys[q] / nc));
    for (int64_t r = prev + 1; r <= cur; ++r) indptr[r] = static_cast<PT>(q);
    if (q < nnz) cols[q] = static_cast<OT>(skeys[q] % nc);
  }
}

struct coo_layout {
  int64_t keys, skeys, vals, sort_tmp, sort_bytes, total;
};

template <typename QT>
static int coo_workspace_layout(int64_t nnz, int64_t n_rows, int64_t n_cols, coo_layout* L) {
  size_t sort_bytes = 0;
  const hipError_t e = rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr),
                                                 static_cast<uint64_t*>(nullptr), static_cast<QT*>(nullptr),
                                                 static_cast<QT*>(nullptr), static_cast<size_t>(nnz), 0,
                                                 coo_key_bits(n_rows, n_cols), hipStream_t(0));
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "coo sort size query: %s", hipGetErrorString(e));
  int64_t off = 0;
  L->keys = off;     off = coo_align(off + nnz * 8);
  L->skeys = off;    off = coo_align(off + nnz * 8);
  L->vals = off;     off = coo_align(off + nnz * static_cast<int64_t>(sizeof(QT)));
  L->sort_tmp = off; off = coo_align(off + static_cast<int64_t>(sort_bytes));
  L->sort_bytes = static_cast<int64_t>(sort_bytes);
  L->total = off + 256;
  return FSP_OK;
}

template <typename F>
static int with_perm_type(int32_t perm_type, F&& f) {
  if (perm_type == FSP_I32) return f(int32_t{});
  if (perm_type == FSP_I64) return f(int64_t{});
  return fail(FSP_ERR_DTYPE, "unsupported perm type %d", (int)perm_type);
}

}  // namespace fsp

using namespace fsp;

extern "C" {

int fsp_coo_to_csr_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz, int32_t perm_type, int64_t* bytes) {
  if (bytes == nullptr || n_rows < 0 || n_cols < 0 || nnz < 0) {
    return fail(FSP_ERR_ARG, "fsp_coo_to_csr_workspace: bad arguments");
  }
  return with_perm_type(perm_type, [&](auto q) -> int {
    coo_layout L;
    const int st = coo_workspace_layout<decltype(q)>(nnz, n_rows, n_cols, &L);
    if (st) return st;
    *bytes = L.total;
    return FSP_OK;
  });
}

int fsp_coo_to_csr(int32_t index_type, const void* row_dev, const void* col_dev, int64_t n_rows, int64_t n_cols,
                   int64_t nnz, int32_t indptr_type, void* indptr_dev, int32_t out_index_type, void* cols_dev,
                   int32_t perm_type, void* perm_dev, uint32_t* err_dev, void* ws_dev, int64_t ws_bytes,
                   void* stream) {
  if (n_rows < 0 || n_cols < 0 || nnz < 0 || indptr_dev == nullptr || err_dev == nullptr ||
      (nnz > 0 && (row_dev == nullptr || col_dev == nullptr || cols_dev == nullptr || perm_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_coo_to_csr: bad arguments");
  }
  if (nnz > 0 && (n_rows == 0 || n_cols == 0)) return fail(FSP_ERR_ARG, "fsp_coo_to_csr: nonzeros in an empty shape");
  if (perm_type == FSP_I32 && nnz > INT32_MAX) return fail(FSP_ERR_ARG, "fsp_coo_to_csr: int32 perm cannot address nnz");
  if (static_cast<double>(n_rows) * static_cast<double>(n_cols) > 1.8e19) {
    return fail(FSP_ERR_ARG, "fsp_coo_to_csr: n_rows * n_cols exceeds 64-bit keys");
  }
  const hipStream_t s = as_stream(stream);
  unsigned int* err = reinterpret_cast<unsigned int*>(err_dev);
  hipLaunchKernelGGL(coo_zero_err_kernel, dim3(1), dim3(1), 0, s, err);
  int st = check_launch("coo_zero_err");
  if (st) return st;
  return with_perm_type(perm_type, [&](auto qtag) -> int {
    using QT = decltype(qtag);
    coo_layout L;
    int st2 = coo_workspace_layout<QT>(nnz, n_rows, n_cols, &L);
    if (st2) return st2;
    if (ws_dev == nullptr || ws_bytes < L.total) return fail(FSP_ERR_WORKSPACE, "fsp_coo_to_csr: workspace too small");
    char* base = reinterpret_cast<char*>(coo_align(reinterpret_cast<intptr_t>(ws_dev)));
    uint64_t* keys = reinterpret_cast<uint64_t*>(base + L.keys);
    uint64_t* skeys = reinterpret_cast<uint64_t*>(base + L.skeys);
    QT* vals = reinterpret_cast<QT*>(base + L.vals);
    QT* perm = static_cast<QT*>(perm_dev);
    if (nnz > 0) {
      FSP_DISPATCH_INDEX(index_type, IT, {
        hipLaunchKernelGGL((coo_keys_kernel<IT, QT>), dim3(coo_grid(nnz)), dim3(FSP_WG_THREADS), 0, s,
                           static_cast<const IT*>(row_dev), static_cast<const IT*>(col_dev), nnz, n_rows, n_cols,
                           keys, vals, err);
        st2 = check_launch("coo_keys");
        if (st2) return st2;
      });
      size_t tb = static_cast<size_t>(L.sort_bytes);
      const hipError_t e = rocprim::radix_sort_pairs(base + L.sort_tmp, tb, keys, skeys, vals, perm,
                                                     static_cast<size_t>(nnz), 0, coo_key_bits(n_rows, n_cols), s);
      if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "coo sort: %s", hipGetErrorString(e));
    }
    FSP_DISPATCH_INDEX(indptr_type, PT, {
      FSP_DISPATCH_INDEX(out_index_type, OT, {
        hipLaunchKernelGGL((coo_rows_kernel<PT, OT>), dim3(coo_grid(nnz + 1)), dim3(FSP_WG_THREADS), 0, s, skeys,
                           nnz, n_rows, n_cols, static_cast<PT*>(indptr_dev), static_cast<OT*>(cols_dev));
        return check_launch("coo_rows");
      });
    });
    return FSP_OK;
  });
}

}  // extern "C"
