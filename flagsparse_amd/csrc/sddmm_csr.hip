// CSR SDDMM out[p] = alpha * <X[row(p)], Y[col(p)]> + beta * data[p] on gfx950.
//
// Row-grouped: the lane layout matches the SpMM kernel ([rows][slots][G lanes
// along K]); a row's X slice is loaded once into registers and reused for all
// of the row's nonzeros, while each slot streams the Y rows of its nonzeros
// (UNROLL in flight) and reduces every dot product with a DPP group sum.
// Long rows are cut into segments; no reduction across segments is needed
// because every nonzero owns its output.
#include "fsp_device.hpp"
#include "fsp_host.hpp"
#include "fsp_plan.hpp"

namespace fsp {

template <typename T, typename IT, typename PT, int VEC, int UNROLL, bool ONE_CHUNK>
__global__ __launch_bounds__(FSP_WG_THREADS) void sddmm_bins_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ X, int64_t ldx,
    const T* __restrict__ Y, int64_t ldy, const T* __restrict__ data_in, T* __restrict__ out, int64_t k,
    typename acc_of<T>::type alpha, typename acc_of<T>::type beta, int g_log2, launch_table tab) {
  using A = typename acc_of<T>::type;
  const int kb = find_bin(tab, blockIdx.x);
  const launch_bin& lb = tab.b[kb];
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
  const int64_t kk0 = static_cast<int64_t>(gl) * VEC;
  const int64_t kstep = static_cast<int64_t>(G) * VEC;
  for (int64_t tile = blockIdx.x - lb.block_begin; tile * per_block < lb.count; tile += lb.nblocks) {
    const int64_t rr = tile * per_block + wave * per_wave + sub;
    if (rr >= lb.count) continue;  // whole groups leave together
    const work_item w = load_item(tab, lb, indptr, rr);
    const T* xrow = X + w.row * ldx;
    A xr[VEC];
    if (ONE_CHUNK) {
      if (kk0 < k) {
        const vec_t<T, VEC> xv = load_vec<T, VEC>(xrow + kk0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) xr[v] = static_cast<A>(xv.v[v]);
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) xr[v] = A(0);
      }
    }
    for (int64_t p = w.begin + slot; p < w.end; p += S * UNROLL) {
      A dot[UNROLL];
      int64_t c[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int64_t q = p + u * S;
        c[u] = static_cast<int64_t>(indices[q < w.end ? q : p]);
        dot[u] = A(0);
      }
      if (ONE_CHUNK) {
        if (kk0 < k) {
          vec_t<T, VEC> yv[UNROLL];
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) yv[u] = load_vec<T, VEC>(Y + c[u] * ldy + kk0);
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dot[u] += xr[v] * static_cast<A>(yv[u].v[v]);
          }
        }
      } else {
        for (int64_t kk = kk0; kk < k; kk += kstep) {
          const vec_t<T, VEC> xv = load_vec<T, VEC>(xrow + kk);
          vec_t<T, VEC> yv[UNROLL];
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) yv[u] = load_vec<T, VEC>(Y + c[u] * ldy + kk);
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dot[u] += static_cast<A>(xv.v[v]) * static_cast<A>(yv[u].v[v]);
          }
        }
      }
      if (UNROLL == 4 && G >= 4) {
        // Transposed reduction of the 4 dot products over the G lanes: two
        // exchange steps halve the vector each time, then one value per lane
        // is summed over the aligned sub-group of G/4 lanes.  The total of
        // dot u ends in lane (u >> 1) * G/2 + (u & 1) * G/4 of the group.
        const int h = G >> 1, qg = G >> 2;
        const bool hi = (gl & h) != 0;
        const A k0 = hi ? dot[2] : dot[0], k1 = hi ? dot[3] : dot[1];
        const A s0 = hi ? dot[0] : dot[2], s1 = hi ? dot[1] : dot[3];
        const A e0 = k0 + shfl_xor(s0, h);
        const A e1 = k1 + shfl_xor(s1, h);
        const bool hi2 = (gl & qg) != 0;
        A f = (hi2 ? e1 : e0) + shfl_xor(hi2 ? e0 : e1, qg);
        f = group_sum(f, qg);
        if ((gl & (qg - 1)) == 0) {
          const int u = (hi ? 2 : 0) + (hi2 ? 1 : 0);
          const int64_t q = p + static_cast<int64_t>(u) * S;
          if (q < w.end) {
            A r = alpha * f;
            if (data_in != nullptr) r += beta * static_cast<A>(data_in[q]);
            out[q] = static_cast<T>(r);
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) dot[u] = group_sum(dot[u], G);
        if (gl == 0) {
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
            const int64_t q = p + u * S;
            if (q < w.end) {
              A r = alpha * dot[u];
              if (data_in != nullptr) r += beta * static_cast<A>(data_in[q]);
              out[q] = static_cast<T>(r);
            }
          }
        }
      }
    }
  }
}

template <typename T, typename IT, typename PT, int VEC>
static int launch_sddmm(const void* indptr, const void* indices, const void* X, int64_t ldx, const void* Y,
                        int64_t ldy, const void* data_in, void* out, int64_t k, double alpha, double beta,
                        int lanes_k, const fsp_csr_plan* plan, hipStream_t s) {
  using A = typename acc_of<T>::type;
  const int64_t vk = ceil_div(k, VEC);
  int G = lanes_k;
  if (G <= 0) {
    G = 1;
    while (G < 64 && G < vk) G <<= 1;
  }
  if (G > 64 || !is_pow2(G)) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: lanes_k %d must be a power of two <= 64", G);
  const int g_log2 = ilog2(G);
  const int max_s_log2 = 6 - g_log2;
  launch_table tab;
  const int64_t blocks = build_launch_table(
      plan, max_s_log2, max_s_log2,
      [g_log2](int sl, int) { return int64_t(FSP_WAVES) * (int64_t(64) >> (g_log2 + sl)); }, &tab);
  if (blocks == 0) return FSP_OK;
  const bool one_chunk = vk <= G;
#define FSP_SDDMM_LAUNCH(ONE)                                                                             \
  hipLaunchKernelGGL((sddmm_bins_kernel<T, IT, PT, VEC, 4, ONE>), dim3(blocks), dim3(FSP_WG_THREADS), 0, s, \
                     static_cast<const PT*>(indptr), static_cast<const IT*>(indices),                      \
                     static_cast<const T*>(X), ldx, static_cast<const T*>(Y), ldy,                         \
                     static_cast<const T*>(data_in), static_cast<T*>(out), k, static_cast<A>(alpha),       \
                     static_cast<A>(beta), g_log2, tab)
  if (one_chunk) {
    FSP_SDDMM_LAUNCH(true);
  } else {
    FSP_SDDMM_LAUNCH(false);
  }
#undef FSP_SDDMM_LAUNCH
  return check_launch("sddmm_bins");
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_sddmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type, const void* indptr_dev,
                             const void* indices_dev, const void* X_dev, int64_t ldx, const void* Y_dev,
                             int64_t ldy, const void* data_in_dev, void* out_dev, int64_t n_rows, int64_t k,
                             double alpha, double beta, int32_t lanes_k, const fsp_csr_plan* plan, void* stream) {
  if (n_rows < 0 || k < 0) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: negative size");
  int st = check_plan(plan, n_rows);
  if (st) return st;
  if (n_rows == 0) return FSP_OK;
  if (indptr_dev == nullptr) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: NULL indptr");
  if (k > 0 && (ldx < k || ldy < k || X_dev == nullptr || Y_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_sddmm_csr: bad dense operands");
  }
  const hipStream_t s = as_stream(stream);
  FSP_DISPATCH_VALUE(value_type, T, {
    if (k == 0) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: k must be positive");
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        constexpr int V = 16 / sizeof(T);
        const bool wide = (reinterpret_cast<uintptr_t>(X_dev) % 16 == 0) &&
                          (reinterpret_cast<uintptr_t>(Y_dev) % 16 == 0) && k % V == 0 && ldx % V == 0 &&
                          ldy % V == 0;
        if (wide) {
          return launch_sddmm<T, IT, PT, V>(indptr_dev, indices_dev, X_dev, ldx, Y_dev, ldy, data_in_dev, out_dev,
                                            k, alpha, beta, lanes_k, plan, s);
        }
        return launch_sddmm<T, IT, PT, 1>(indptr_dev, indices_dev, X_dev, ldx, Y_dev, ldy, data_in_dev, out_dev,
                                          k, alpha, beta, 0, plan, s);
      });
    });
  });
  return FSP_OK;
}
