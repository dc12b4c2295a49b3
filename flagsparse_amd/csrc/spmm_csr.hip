// CSR SpMM C = A B on gfx950 (B, C row-major).
//
// Lane layout of a wave: [rows of the wave][nonzero slots][G lanes along N].
// The G lanes of a slot each own VEC consecutive columns (one 16-byte access
// of a B row when VEC*sizeof(T) == 16), so a slot streams whole B rows with
// coalesced loads.  The S = 64/(G*rows) slots of a row take nonzeros
// begin+slot, begin+slot+S, ..., UNROLL of them per trip (UNROLL B-row loads in
// flight per lane), and are summed by a lane-xor exchange at the end.  The
// row bins pick S by row length so that short rows share a wave and long rows
// get all 64/G slots; rows longer than the last bin are cut into segments whose
// partial rows a second pass adds in segment order (atomic-free).
// Over a CSC with perm != NULL the same kernel computes A^T B, reading the
// values as data[perm[q]] (no transposed copy of A).
#include "fsp_device.hpp"
#include "fsp_host.hpp"
#include "fsp_plan.hpp"

namespace fsp {

template <typename T, typename IT, typename PT, typename QT, int VEC, int UNROLL, bool PERM>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmm_bins_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ data,
    const QT* __restrict__ perm, const T* __restrict__ B, int64_t ldb, T* __restrict__ C, int64_t ldc,
    int64_t n_dense, int g_log2, typename acc_of<T>::type* __restrict__ partials, launch_table tab) {
  using A = typename acc_of<T>::type;
  const int k = find_bin(tab, blockIdx.x);
  const launch_bin& lb = tab.b[k];
  const int s_log2 = lb.lanes_log2;
  const int G = 1 << g_log2;
  const int S = 1 << s_log2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int gl = lane & (G - 1);
  const int slot = (lane >> g_log2) & (S - 1);
  const int sub = lane >> (g_log2 + s_log2);
  const int64_t per_wave = 64 >> (g_log2 + s_log2);
  const int64_t per_block = per_wave * FSP_WAVES;
  const int64_t col0 = (static_cast<int64_t>(blockIdx.y) * G + gl) * VEC;
  const bool col_ok = col0 < n_dense;
  for (int64_t tile = blockIdx.x - lb.block_begin; tile * per_block < lb.count; tile += lb.nblocks) {
    const int64_t rr = tile * per_block + wave * per_wave + sub;
    const bool active = rr < lb.count;
    A acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = A(0);
    work_item w{0, 0, 0, 0};
    if (active) w = load_item(tab, lb, indptr, rr);
    if (active && col_ok) {
      const T* Bc = B + col0;
      for (int64_t p = w.begin + slot; p < w.end; p += S * UNROLL) {
        int64_t c[UNROLL];
        A a[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          const int64_t q = p + u * S;
          const bool ok = q < w.end;
          const int64_t qc = ok ? q : p;
          c[u] = static_cast<int64_t>(indices[qc]);
          const T av = data[PERM ? static_cast<int64_t>(perm[qc]) : qc];
          a[u] = ok ? static_cast<A>(av) : A(0);
        }
        vec_t<T, VEC> b[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          if (p + u * S < w.end) {
            b[u] = load_vec<T, VEC>(Bc + c[u] * ldb);
          } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) b[u].v[v] = T(0);
          }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
          for (int v = 0; v < VEC; ++v) acc[v] += a[u] * static_cast<A>(b[u].v[v]);
        }
      }
    }
    // Sum the S slots of a row: lanes that differ only in the slot bits.
    for (int j = 0; j < s_log2; ++j) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc[v] += shfl_xor(acc[v], G << j);
    }
    if (active && col_ok && slot == 0) {
      if (lb.is_segment) {
        A* dst = partials + w.out * n_dense + col0;
#pragma unroll
        for (int v = 0; v < VEC; ++v) dst[v] = acc[v];
      } else {
        vec_t<T, VEC> o;
#pragma unroll
        for (int v = 0; v < VEC; ++v) o.v[v] = static_cast<T>(acc[v]);
        store_vec<T, VEC>(C + w.out * ldc + col0, o);
      }
    }
  }
}

// C[long_rows[i], :] = sum of the row's segment partials, in segment order.
template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmm_long_reduce_kernel(
    const typename acc_of<T>::type* __restrict__ partials, const int32_t* __restrict__ long_rows,
    const int64_t* __restrict__ seg_ptr, int64_t n_long, int64_t n_dense, T* __restrict__ C, int64_t ldc) {
  using A = typename acc_of<T>::type;
  for (int64_t i = blockIdx.y; i < n_long; i += gridDim.y) {
    const int64_t s0 = seg_ptr[i], s1 = seg_ptr[i + 1];
    const int64_t row = long_rows[i];
    for (int64_t col = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; col < n_dense;
         col += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
      A acc = A(0);
      for (int64_t sg = s0; sg < s1; ++sg) acc += partials[sg * n_dense + col];
      C[row * ldc + col] = static_cast<T>(acc);
    }
  }
}

template <typename T, typename IT, typename PT, typename QT, int VEC, bool PERM>
static int launch_spmm(const void* indptr, const void* indices, const void* data, const void* perm,
                       const void* B, int64_t ldb, void* C, int64_t ldc, int64_t n_dense, int lanes_n,
                       const fsp_csr_plan* plan, hipStream_t s) {
  using A = typename acc_of<T>::type;
  const int64_t vcols = ceil_div(n_dense, VEC);
  int G = lanes_n;
  if (G <= 0) {
    G = 1;
    while (G < 64 && G < vcols) G <<= 1;
  }
  if (G > 64 || !is_pow2(G)) return fail(FSP_ERR_ARG, "fsp_spmm_csr: lanes_n %d must be a power of two <= 64", G);
  const int g_log2 = ilog2(G);
  const int max_s_log2 = 6 - g_log2;
  launch_table tab;
  const int64_t blocks = build_launch_table(
      plan, max_s_log2, max_s_log2,
      [g_log2](int sl, int) { return int64_t(FSP_WAVES) * (int64_t(64) >> (g_log2 + sl)); }, &tab);
  if (plan->n_segments > 0 &&
      (plan->partials_dev == nullptr || plan->partials_bytes < plan->n_segments * n_dense * (int64_t)sizeof(A))) {
    return fail(FSP_ERR_WORKSPACE, "fsp_spmm_csr: partials need %lld bytes",
                (long long)(plan->n_segments * n_dense * (int64_t)sizeof(A)));
  }
  A* partials = static_cast<A*>(plan->partials_dev);
  const int64_t ytiles = ceil_div(vcols, G);
  if (ytiles > 65535) return fail(FSP_ERR_ARG, "fsp_spmm_csr: n_dense %lld too large", (long long)n_dense);
  if (blocks > 0) {
    const dim3 grid(static_cast<unsigned>(blocks), static_cast<unsigned>(ytiles));
#define FSP_SPMM_LAUNCH(U)                                                                                 \
  hipLaunchKernelGGL((spmm_bins_kernel<T, IT, PT, QT, VEC, U, PERM>), grid, dim3(FSP_WG_THREADS), 0, s,     \
                     static_cast<const PT*>(indptr), static_cast<const IT*>(indices),                       \
                     static_cast<const T*>(data), static_cast<const QT*>(perm), static_cast<const T*>(B), ldb, \
                     static_cast<T*>(C), ldc, n_dense, g_log2, partials, tab)
    if (plan->unroll <= 2) {
      FSP_SPMM_LAUNCH(2);
    } else if (plan->unroll <= 4) {
      FSP_SPMM_LAUNCH(4);
    } else {
      FSP_SPMM_LAUNCH(8);
    }
#undef FSP_SPMM_LAUNCH
    const int st = check_launch("spmm_bins");
    if (st) return st;
  }
  if (plan->n_long_rows > 0) {
    const int64_t gx = ceil_div(n_dense, FSP_WG_THREADS);
    int64_t gy = plan->n_long_rows;
    if (gy > 65535) gy = 65535;
    hipLaunchKernelGGL(spmm_long_reduce_kernel<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)),
                       dim3(FSP_WG_THREADS), 0, s, partials, plan->long_rows_dev, plan->long_seg_ptr_dev,
                       plan->n_long_rows, n_dense, static_cast<T*>(C), ldc);
    return check_launch("spmm_long_reduce");
  }
  return FSP_OK;
}

template <typename T>
static bool wide_ok(const void* B, int64_t ldb, const void* C, int64_t ldc, int64_t n_dense) {
  constexpr int V = 16 / sizeof(T);
  const bool aligned = (reinterpret_cast<uintptr_t>(B) % 16 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0);
  return aligned && n_dense % V == 0 && ldb % V == 0 && ldc % V == 0;
}

template <typename T, typename IT, typename PT, typename QT, bool PERM>
static int launch_spmm_vec(const void* indptr, const void* indices, const void* data, const void* perm,
                           const void* B, int64_t ldb, void* C, int64_t ldc, int64_t n_dense, int lanes_n,
                           const fsp_csr_plan* plan, hipStream_t s) {
  constexpr int V = 16 / sizeof(T);
  if (wide_ok<T>(B, ldb, C, ldc, n_dense)) {
    return launch_spmm<T, IT, PT, QT, V, PERM>(indptr, indices, data, perm, B, ldb, C, ldc, n_dense, lanes_n,
                                               plan, s);
  }
  return launch_spmm<T, IT, PT, QT, 1, PERM>(indptr, indices, data, perm, B, ldb, C, ldc, n_dense,
                                             0, plan, s);
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_spmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type, const void* indptr_dev,
                            const void* indices_dev, const void* data_dev, int32_t perm_type, const void* perm_dev,
                            const void* B_dev, int64_t ldb, void* C_dev, int64_t ldc, int64_t n_rows,
                            int64_t n_dense, int32_t lanes_n, const fsp_csr_plan* plan, void* stream) {
  if (n_rows < 0 || n_dense < 0) return fail(FSP_ERR_ARG, "fsp_spmm_csr: negative size");
  int st = check_plan(plan, n_rows);
  if (st) return st;
  if (n_rows == 0 || n_dense == 0) return FSP_OK;
  if (indptr_dev == nullptr || B_dev == nullptr || C_dev == nullptr) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr: NULL pointer");
  }
  if (ldb < n_dense || ldc < n_dense) return fail(FSP_ERR_ARG, "fsp_spmm_csr: leading dimension < n_dense");
  const hipStream_t s = as_stream(stream);
  const bool use_perm = perm_dev != nullptr;
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        if (!use_perm) {
          return launch_spmm_vec<T, IT, PT, int32_t, false>(indptr_dev, indices_dev, data_dev, nullptr, B_dev, ldb,
                                                            C_dev, ldc, n_dense, lanes_n, plan, s);
        }
        FSP_DISPATCH_INDEX(perm_type, QT, {
          return launch_spmm_vec<T, IT, PT, QT, true>(indptr_dev, indices_dev, data_dev, perm_dev, B_dev, ldb,
                                                      C_dev, ldc, n_dense, lanes_n, plan, s);
        });
      });
    });
  });
  return FSP_OK;
}
