// CSR SDDMM out[p] = alpha * <X[row(p)], Y[col(p)]> + beta * data[p] on gfx950.
//
// Row-grouped: the lane layout matches the SpMM kernel ([rows][slots][G lanes
// along K]); a row's X slice is loaded once into registers and reused for all
// of the row's nonzeros, while each slot streams the Y rows of its nonzeros
// (UNROLL in flight).  The UNROLL dot products of a slot are reduced together
// by a transposed lane exchange.  Long rows are cut into segments; no
// reduction across segments is needed because every nonzero owns its output.
//
// When K fits one pass of the G lanes (K <= G*VEC) the loop is
// software-pipelined like the SpMM kernel: while the Y rows of one trip are in
// flight a lane loads the column ids of its next trip (or of its next row, with
// that row's X slice), and every load is unconditional (clamped indices,
// masked results) so the compiler waits only for what the FMAs need.
#include "fsp_device.hpp"
#include "fsp_host.hpp"
#include "fsp_plan.hpp"

namespace fsp {

// Sum UNROLL == 4 dot products over the G lanes of a slot group.  Two
// exchange steps halve the vector each time, then one value per lane is
// summed over the aligned sub-group of G/4 lanes; the total of dot u ends in
// lane (u >> 1) * G/2 + (u & 1) * G/4 of the group.  Returns that lane's
// total and sets `u_out` to the dot it holds (-1 for lanes that hold none).
template <typename A>
__device__ __forceinline__ A transposed_sum4(const A (&dot)[4], int gl, int G, int& u_out) {
  const int h = G >> 1, qg = G >> 2;
  const bool hi = (gl & h) != 0;
  const A k0 = hi ? dot[2] : dot[0], k1 = hi ? dot[3] : dot[1];
  const A s0 = hi ? dot[0] : dot[2], s1 = hi ? dot[1] : dot[3];
  const A e0 = k0 + shfl_xor(s0, h);
  const A e1 = k1 + shfl_xor(s1, h);
  const bool hi2 = (gl & qg) != 0;
  A f = (hi2 ? e1 : e0) + shfl_xor(hi2 ? e0 : e1, qg);
  f = group_sum(f, qg);
  u_out = (gl & (qg - 1)) == 0 ? (hi ? 2 : 0) + (hi2 ? 1 : 0) : -1;
  return f;
}

template <typename T, typename IT, typename PT, int VEC, int UNROLL, bool ONE_CHUNK, bool OFF32>
__global__ __launch_bounds__(FSP_WG_THREADS) void sddmm_bins_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ X, int64_t ldx,
    const T* __restrict__ Y, int64_t ldy, const T* __restrict__ data_in, T* __restrict__ out, int64_t k,
    typename acc_of<T>::type alpha, typename acc_of<T>::type beta, int g_log2, launch_table tab) {
  using A = typename acc_of<T>::type;
  using O = typename addr<OFF32>::off_t;
  using P = typename std::conditional<OFF32, int32_t, int64_t>::type;
  using V = vec_t<T, VEC>;
  const int kb = find_bin(tab, blockIdx.x);
  const launch_bin lb = tab.b[kb];  // a copy: the fields stay in registers
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
  // this workgroup's tiles: first, first + step, ... < end
  const tile_range tr = bin_tiles(lb, per_block);
  const int64_t rr_step = tr.step * per_block;
  const O kk0 = static_cast<O>(gl) * VEC;
  const bool k_ok = static_cast<int64_t>(kk0) < k;
  const O kk_ld = k_ok ? kk0 : O(0);  // lanes past K read column 0 (masked)
  const O ldx_o = static_cast<O>(ldx), ldy_o = static_cast<O>(ldy);
  const int64_t last_item = lb.count - 1;
  const bool use_transposed = UNROLL == 4 && G >= 4;

  auto item = [&](int64_t r, O& row, P& beg, P& end) {
    const bool live = r < lb.count;
    const int64_t rc = live ? r : last_item;
    P b, e;
    if (lb.is_segment) {
      row = static_cast<O>(tab.seg_row[rc]);
      b = static_cast<P>(tab.seg_begin[rc]);
      e = static_cast<P>(tab.seg_end[rc]);
    } else {
      const int64_t rw = lb.rows ? static_cast<int64_t>(lb.rows[rc]) : rc;
      row = static_cast<O>(rw);
      b = static_cast<P>(indptr[rw]);
      e = static_cast<P>(indptr[rw + 1]);
    }
    beg = live ? b : P(0);
    end = live ? e : P(0);
  };
  auto trip = [&](P p, P end, IT (&c)[UNROLL]) {
    const P last = end > 0 ? end - 1 : P(0);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const P q = p + static_cast<P>(u) * S;
      c[u] = elem<OFF32>(indices, static_cast<O>(q < end ? q : last));
    }
  };
  auto emit = [&](P q, A dot) {
    A r = alpha * dot;
    if (data_in != nullptr) r += beta * static_cast<A>(elem<OFF32>(data_in, static_cast<O>(q)));
    *addr<OFF32>::at(out, static_cast<O>(q) * static_cast<O>(sizeof(T))) = static_cast<T>(r);
  };
  auto reduce_emit = [&](P p, P end, A (&dot)[UNROLL]) {
    if (use_transposed) {
      int u = -1;
      const A f = transposed_sum4(reinterpret_cast<const A(&)[4]>(dot), gl, G, u);
      if (u >= 0) {
        const P q = p + static_cast<P>(u) * S;
        if (q < end) emit(q, f);
      }
    } else {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) dot[u] = group_sum(dot[u], G);
      if (gl == 0) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          const P q = p + static_cast<P>(u) * S;
          if (q < end) emit(q, dot[u]);
        }
      }
    }
  };

  if constexpr (ONE_CHUNK) {
    int64_t rr = tr.first * per_block + wave * per_wave + sub;
    O row;
    P beg, end;
    item(rr, row, beg, end);
    V xv = *addr<OFF32>::at(reinterpret_cast<const V*>(X), (row * ldx_o + kk_ld) * static_cast<O>(sizeof(T)));
    IT c[UNROLL];
    trip(beg + slot, end, c);
    for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {  // wave-uniform
      const int64_t rr_next = rr + rr_step;
      // the next item is prefetched only when this workgroup will process it
      const bool has_next = tile + tr.step < tr.end;  // wave-uniform
      O row_n = 0;
      P beg_n = 0, end_n = 0;
      V xv_n{};
      if (has_next) {
        item(rr_next, row_n, beg_n, end_n);
        xv_n = *addr<OFF32>::at(reinterpret_cast<const V*>(X), (row_n * ldx_o + kk_ld) * static_cast<O>(sizeof(T)));
      }
      A xr[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) xr[v] = k_ok ? static_cast<A>(xv.v[v]) : A(0);
      bool prefetched = false;
      for (P p = beg + slot; p < end; p += static_cast<P>(S) * UNROLL) {
        V yv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          yv[u] = *addr<OFF32>::at(reinterpret_cast<const V*>(Y),
                                   (static_cast<O>(c[u]) * ldy_o + kk_ld) * static_cast<O>(sizeof(T)));
        }
        const P pn = p + static_cast<P>(S) * UNROLL;
        const bool same = pn < end;
        prefetched = prefetched || !same;
        IT cn[UNROLL];
        if (same || has_next) trip(same ? pn : beg_n + slot, same ? end : end_n, cn);
        A dot[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          dot[u] = A(0);
#pragma unroll
          for (int v = 0; v < VEC; ++v) dot[u] += xr[v] * static_cast<A>(yv[u].v[v]);
          dot[u] = k_ok ? dot[u] : A(0);  // lanes past K read column 0: drop it exactly
        }
        reduce_emit(p, end, dot);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) c[u] = cn[u];
      }
      if (!prefetched && has_next) trip(beg_n + slot, end_n, c);
      xv = xv_n;
      row = row_n;
      beg = beg_n;
      end = end_n;
      rr = rr_next;
    }
  } else {
    // K spans several passes of the G lanes: the X slice is re-read per pass.
    const O kstep = static_cast<O>(G) * VEC;
    for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {
      const int64_t rr = tile * per_block + wave * per_wave + sub;
      O row;
      P beg, end;
      item(rr, row, beg, end);
      for (P p = beg + slot; p < end; p += static_cast<P>(S) * UNROLL) {
        IT c[UNROLL];
        trip(p, end, c);
        A dot[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) dot[u] = A(0);
        for (O kk = kk0; static_cast<int64_t>(kk) < k; kk += kstep) {
          const V xv = *addr<OFF32>::at(reinterpret_cast<const V*>(X), (row * ldx_o + kk) * static_cast<O>(sizeof(T)));
          V yv[UNROLL];
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
            yv[u] = *addr<OFF32>::at(reinterpret_cast<const V*>(Y),
                                     (static_cast<O>(c[u]) * ldy_o + kk) * static_cast<O>(sizeof(T)));
          }
#pragma unroll
          for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) dot[u] += static_cast<A>(xv.v[v]) * static_cast<A>(yv[u].v[v]);
          }
        }
        reduce_emit(p, end, dot);
      }
    }
  }
}

template <typename T, typename IT, typename PT, int VEC>
static int launch_sddmm(const void* indptr, const void* indices, const void* X, int64_t ldx, const void* Y,
                        int64_t ldy, const void* data_in, void* out, int64_t n_rows, int64_t k, int64_t nnz,
                        int64_t y_rows, double alpha, double beta, int lanes_k, const fsp_csr_plan* plan,
                        hipStream_t s) {
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
  const int64_t lim = int64_t(1) << 32;
  const bool off32 = std::is_same<PT, int32_t>::value && nnz * (int64_t)sizeof(T) < lim &&
                     nnz * (int64_t)sizeof(IT) < lim && n_rows * ldx * (int64_t)sizeof(T) < lim &&
                     y_rows * ldy * (int64_t)sizeof(T) < lim;
  const dim3 grid(static_cast<unsigned>(blocks)), block(FSP_WG_THREADS);
#define FSP_SDDMM_LAUNCH(ONE, O32)                                                                         \
  hipLaunchKernelGGL((sddmm_bins_kernel<T, IT, PT, VEC, 4, ONE, O32>), grid, block, 0, s,                    \
                     static_cast<const PT*>(indptr), static_cast<const IT*>(indices),                       \
                     static_cast<const T*>(X), ldx, static_cast<const T*>(Y), ldy,                          \
                     static_cast<const T*>(data_in), static_cast<T*>(out), k, static_cast<A>(alpha),        \
                     static_cast<A>(beta), g_log2, tab)
  if (one_chunk) {
    if (off32) {
      FSP_SDDMM_LAUNCH(true, true);
    } else {
      FSP_SDDMM_LAUNCH(true, false);
    }
  } else {
    if (off32) {
      FSP_SDDMM_LAUNCH(false, true);
    } else {
      FSP_SDDMM_LAUNCH(false, false);
    }
  }
#undef FSP_SDDMM_LAUNCH
  return check_launch("sddmm_bins");
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_sddmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type, const void* indptr_dev,
                             const void* indices_dev, const void* X_dev, int64_t ldx, const void* Y_dev,
                             int64_t ldy, const void* data_in_dev, void* out_dev, int64_t n_rows, int64_t k,
                             int64_t nnz, int64_t y_rows, double alpha, double beta, int32_t lanes_k,
                             const fsp_csr_plan* plan, void* stream) {
  if (n_rows < 0 || k < 0 || nnz < 0 || y_rows < 0) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: negative size");
  int st = check_plan(plan, n_rows);
  if (st) return st;
  if (n_rows == 0 || nnz == 0) return FSP_OK;
  if (indptr_dev == nullptr || indices_dev == nullptr || out_dev == nullptr) {
    return fail(FSP_ERR_ARG, "fsp_sddmm_csr: NULL pattern or output");
  }
  if (k == 0) return fail(FSP_ERR_ARG, "fsp_sddmm_csr: k must be positive");
  if (ldx < k || ldy < k || X_dev == nullptr || Y_dev == nullptr || y_rows == 0) {
    return fail(FSP_ERR_ARG, "fsp_sddmm_csr: bad dense operands");
  }
  const hipStream_t s = as_stream(stream);
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        constexpr int V = 16 / sizeof(T);
        const bool wide = (reinterpret_cast<uintptr_t>(X_dev) % 16 == 0) &&
                          (reinterpret_cast<uintptr_t>(Y_dev) % 16 == 0) && k % V == 0 && ldx % V == 0 &&
                          ldy % V == 0;
        if (wide) {
          return launch_sddmm<T, IT, PT, V>(indptr_dev, indices_dev, X_dev, ldx, Y_dev, ldy, data_in_dev, out_dev,
                                            n_rows, k, nnz, y_rows, alpha, beta, lanes_k, plan, s);
        }
        return launch_sddmm<T, IT, PT, 1>(indptr_dev, indices_dev, X_dev, ldx, Y_dev, ldy, data_in_dev, out_dev,
                                          n_rows, k, nnz, y_rows, alpha, beta, 0, plan, s);
      });
    });
  });
  return FSP_OK;
}
