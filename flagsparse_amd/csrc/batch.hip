// Packing of a batch of CSR matrices into one block-diagonal CSR.
//
// A batch arrives as a packed offset table over concatenated arrays: matrix m
// owns rows [row_off[m], row_off[m+1]) and columns [col_off[m], col_off[m+1])
// of the batch, nonzeros [nnz_off[m], nnz_off[m+1]) of the concatenated
// indices / values, and its own row pointer (n_m + 1 local offsets starting
// at 0) at local_indptr[row_off[m] + m ...].  The packed matrix has global row
// pointers (nnz_off[m] + local offset) and global column ids (col_off[m] +
// local column); the values are shared, not copied.  A single launch of the
// row-bin kernels over the packed matrix then runs the whole batch, and every
// row is reduced in the same order as in a launch over its own matrix (row
// bins depend on the row's length only).
//
// The same pass validates every matrix (row pointer starts at 0, ends at the
// matrix's nonzero count, never decreases; columns inside the matrix) and
// reports violations as bits of *err.
#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

constexpr unsigned kBatchErrBounds = 1u;    // indptr does not start at 0 / end at nnz_m
constexpr unsigned kBatchErrDecrease = 2u;  // indptr decreases
constexpr unsigned kBatchErrColumn = 4u;    // column outside [0, n_cols_m)

// Largest m in [0, M) with off[m] <= v (off ascending, off[0] = 0 <= v).
__device__ __forceinline__ int64_t batch_owner(const int64_t* __restrict__ off, int64_t M, int64_t v) {
  int64_t lo = 0, hi = M;  // off[lo] <= v, answer in [lo, hi)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void batch_zero_err_kernel(unsigned int* err) { err[0] = 0u; }

// Thread t: matrix t (bounds checks), row t (global row pointer), nonzero t
// (global column id).
template <typename LP, typename LI, typename GP, typename GI>
__global__ __launch_bounds__(FSP_WG_THREADS) void batch_pack_kernel(
    int64_t M, const int64_t* __restrict__ row_off, const int64_t* __restrict__ col_off,
    const int64_t* __restrict__ nnz_off, const LP* __restrict__ lindptr, const LI* __restrict__ lind,
    int64_t total_rows, int64_t total_nnz, GP* __restrict__ gindptr, GI* __restrict__ gind,
    unsigned int* __restrict__ err) {
  const int64_t span = total_nnz > total_rows + 1 ? total_nnz : total_rows + 1;
  const int64_t n = span > M ? span : M;
  unsigned int bad = 0u;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    if (t < M) {
      const int64_t first = row_off[t] + t;        // local indptr of matrix t
      const int64_t last = row_off[t + 1] + t;
      if (static_cast<int64_t>(lindptr[first]) != 0 ||
          static_cast<int64_t>(lindptr[last]) != nnz_off[t + 1] - nnz_off[t]) {
        bad |= kBatchErrBounds;
      }
    }
    if (t < total_rows) {
      const int64_t m = batch_owner(row_off, M, t);
      const int64_t li = t + m;                    // local indptr slot of row t
      const int64_t lp = static_cast<int64_t>(lindptr[li]);
      if (static_cast<int64_t>(lindptr[li + 1]) < lp) bad |= kBatchErrDecrease;
      gindptr[t] = static_cast<GP>(nnz_off[m] + lp);
    } else if (t == total_rows) {
      gindptr[t] = static_cast<GP>(total_nnz);
    }
    if (t < total_nnz) {
      const int64_t m = batch_owner(nnz_off, M, t);
      const int64_t c = static_cast<int64_t>(lind[t]);
      if (c < 0 || c >= col_off[m + 1] - col_off[m]) bad |= kBatchErrColumn;
      gind[t] = static_cast<GI>(col_off[m] + c);
    }
  }
  if (bad) atomicOr(err, bad);
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_batch_pack(int32_t indptr_type, int32_t index_type, int64_t n_mats, const int64_t* row_off_dev,
                              const int64_t* col_off_dev, const int64_t* nnz_off_dev, const void* local_indptr_dev,
                              const void* local_indices_dev, int64_t total_rows, int64_t total_nnz,
                              int32_t out_indptr_type, void* out_indptr_dev, int32_t out_index_type,
                              void* out_indices_dev, uint32_t* err_dev, void* stream) {
  if (n_mats < 1 || total_rows < 0 || total_nnz < 0 || row_off_dev == nullptr || col_off_dev == nullptr ||
      nnz_off_dev == nullptr || local_indptr_dev == nullptr || out_indptr_dev == nullptr || err_dev == nullptr ||
      (total_nnz > 0 && (local_indices_dev == nullptr || out_indices_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_batch_pack: bad arguments");
  }
  if (out_indptr_type == FSP_I32 && total_nnz > INT32_MAX) {
    return fail(FSP_ERR_ARG, "fsp_batch_pack: %lld nonzeros need an int64 row pointer", (long long)total_nnz);
  }
  const hipStream_t s = as_stream(stream);
  unsigned int* err = reinterpret_cast<unsigned int*>(err_dev);
  hipLaunchKernelGGL(batch_zero_err_kernel, dim3(1), dim3(1), 0, s, err);
  int st = check_launch("batch_zero_err");
  if (st) return st;
  int64_t span = total_nnz > total_rows + 1 ? total_nnz : total_rows + 1;
  if (n_mats > span) span = n_mats;
  int64_t g = ceil_div(span, FSP_WG_THREADS);
  if (g > 65536) g = 65536;
  FSP_DISPATCH_INDEX(indptr_type, LP, {
    FSP_DISPATCH_INDEX(index_type, LI, {
      FSP_DISPATCH_INDEX(out_indptr_type, GP, {
        FSP_DISPATCH_INDEX(out_index_type, GI, {
          hipLaunchKernelGGL((batch_pack_kernel<LP, LI, GP, GI>), dim3(static_cast<unsigned>(g)),
                             dim3(FSP_WG_THREADS), 0, s, n_mats, row_off_dev, col_off_dev, nnz_off_dev,
                             static_cast<const LP*>(local_indptr_dev), static_cast<const LI*>(local_indices_dev),
                             total_rows, total_nnz, static_cast<GP*>(out_indptr_dev),
                             static_cast<GI*>(out_indices_dev), err);
          return check_launch("batch_pack");
        });
      });
    });
  });
  return FSP_OK;
}
