// BSR SpMV y = A x (and A^T x) on gfx950, any block size d in [2, 32].
//
// One wave per output block row.  Lanes are laid out as [P blocks][R rows][dp
// columns] with dp = pow2ceil(d) lanes along a block row (lanes c >= d idle),
// R = min(pow2ceil(d), 64 / dp) block rows handled per pass (T = ceil(d / R)
// passes, one accumulator each) and P = 64 / (dp * R) blocks of the block row
// in flight at once (small blocks).  Each lane accumulates its (row, column)
// slot over the blocks it visits, in block order; the dp column lanes are then
// summed with DPP / swizzle and the P block lanes with lane-xor exchanges, so
// the result is deterministic and atomic-free.  A d x d block is read as d
// runs of d consecutive values; x as d consecutive values per block.
//
// Transposed (A^T x): the launch runs over the block CSC of A (block column
// pointer, block rows, and the permutation of block ids, as fsp_csr_to_csc
// builds them from the block row pointer / block columns); element (i, c) of
// a block of A^T is element (c, i) of the stored block.
#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

constexpr int kBsrSpmvMaxDim = 32;

struct bsr_spmv_shape {
  int d, dp, R, P, T;
};

static inline int pow2ceil_int(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

static inline bsr_spmv_shape bsr_spmv_layout(int d) {
  bsr_spmv_shape s;
  s.d = d;
  s.dp = pow2ceil_int(d);
  const int rows_cap = 64 / s.dp;
  s.R = s.dp < rows_cap ? s.dp : rows_cap;
  s.P = 64 / (s.dp * s.R);
  s.T = (d + s.R - 1) / s.R;
  return s;
}

// NT = the largest pass count the launch may use (1 for d <= 8, 4 for
// d <= 16, 16 for d <= 32), so small blocks keep few registers.
template <typename T, typename PT, typename IT, typename QT, bool TRANS, int NT>
__global__ __launch_bounds__(FSP_WG_THREADS) void bsr_spmv_kernel(const PT* __restrict__ ptr,
                                                                  const IT* __restrict__ bidx,
                                                                  const QT* __restrict__ perm,
                                                                  const T* __restrict__ data,
                                                                  const T* __restrict__ x, T* __restrict__ y,
                                                                  int64_t n_out_blocks, bsr_spmv_shape sh) {
  using A = typename spmv_acc_of<T>::type;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int d = sh.d, dp = sh.dp, R = sh.R, P = sh.P, npass = sh.T;
  const int c = lane & (dp - 1);
  const int r = (lane / dp) & (R - 1);
  const int jj = lane / (dp * R);
  const int64_t D = static_cast<int64_t>(d) * d;
  const bool c_ok = c < d;
  for (int64_t br = static_cast<int64_t>(blockIdx.x) * FSP_WAVES + wave; br < n_out_blocks;
       br += static_cast<int64_t>(gridDim.x) * FSP_WAVES) {  // wave-uniform
    const int64_t beg = static_cast<int64_t>(ptr[br]);
    const int64_t end = static_cast<int64_t>(ptr[br + 1]);
    A acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = A(0);
    for (int64_t q = beg + jj; q < end; q += P) {
      const int64_t blk = TRANS ? static_cast<int64_t>(perm[q]) : q;
      const int64_t col = static_cast<int64_t>(bidx[q]);
      const A xv = c_ok ? static_cast<A>(x[col * d + c]) : A(0);
      const T* base = data + blk * D;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (t < npass) {
          const int i = t * R + r;
          if (c_ok && i < d) {
            const T v = TRANS ? base[static_cast<int64_t>(c) * d + i] : base[static_cast<int64_t>(i) * d + c];
            acc[t] += static_cast<A>(v) * xv;
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < npass) {  // wave-uniform
        A v = group_sum(acc[t], dp);
        for (int m = dp * R; m < 64; m <<= 1) v += shfl_xor(v, m);
        const int i = t * R + r;
        if (c == 0 && jj == 0 && i < d) y[br * d + i] = static_cast<T>(v);
      }
    }
  }
}

template <typename T, typename PT, typename IT, typename QT, bool TRANS>
static int launch_bsr_spmv(int nt, int64_t g, hipStream_t s, const void* ptr, const void* bidx, const void* perm,
                           const void* data, const void* x, void* y, int64_t n_out_blocks, bsr_spmv_shape sh) {
  const dim3 grid(static_cast<unsigned>(g)), block(FSP_WG_THREADS);
  const PT* p = static_cast<const PT*>(ptr);
  const IT* b = static_cast<const IT*>(bidx);
  const QT* q = static_cast<const QT*>(perm);
  const T* d = static_cast<const T*>(data);
  const T* xv = static_cast<const T*>(x);
  T* yv = static_cast<T*>(y);
  if (nt == 1) {
    hipLaunchKernelGGL((bsr_spmv_kernel<T, PT, IT, QT, TRANS, 1>), grid, block, 0, s, p, b, q, d, xv, yv, n_out_blocks, sh);
  } else if (nt == 4) {
    hipLaunchKernelGGL((bsr_spmv_kernel<T, PT, IT, QT, TRANS, 4>), grid, block, 0, s, p, b, q, d, xv, yv, n_out_blocks, sh);
  } else {
    hipLaunchKernelGGL((bsr_spmv_kernel<T, PT, IT, QT, TRANS, 16>), grid, block, 0, s, p, b, q, d, xv, yv, n_out_blocks, sh);
  }
  return check_launch(TRANS ? "bsr_spmv_trans" : "bsr_spmv");
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_bsr_spmv(int32_t value_type, int32_t ptr_type, int32_t index_type, const void* ptr_dev,
                            const void* bidx_dev, int32_t perm_type, const void* perm_dev, const void* data_dev,
                            int32_t block_dim, const void* x_dev, void* y_dev, int64_t n_out_blocks, void* stream) {
  if (block_dim < 1 || block_dim > kBsrSpmvMaxDim || n_out_blocks < 0 || ptr_dev == nullptr ||
      (n_out_blocks > 0 && (y_dev == nullptr || x_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_bsr_spmv: bad arguments (block_dim must be 1..%d)", kBsrSpmvMaxDim);
  }
  if (n_out_blocks == 0) return FSP_OK;
  const bsr_spmv_shape sh = bsr_spmv_layout(block_dim);
  int64_t g = ceil_div(n_out_blocks, FSP_WAVES);
  if (g > 65536) g = 65536;
  const hipStream_t s = as_stream(stream);
  const int nt = block_dim <= 8 ? 1 : (block_dim <= 16 ? 4 : 16);
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(ptr_type, PT, {
      FSP_DISPATCH_INDEX(index_type, IT, {
        if (perm_dev == nullptr) {
          return launch_bsr_spmv<T, PT, IT, int32_t, false>(nt, g, s, ptr_dev, bidx_dev, nullptr, data_dev, x_dev,
                                                            y_dev, n_out_blocks, sh);
        }
        FSP_DISPATCH_INDEX(perm_type, QT, {
          return launch_bsr_spmv<T, PT, IT, QT, true>(nt, g, s, ptr_dev, bidx_dev, perm_dev, data_dev, x_dev, y_dev,
                                                      n_out_blocks, sh);
        });
      });
    });
  });
  return FSP_OK;
}
