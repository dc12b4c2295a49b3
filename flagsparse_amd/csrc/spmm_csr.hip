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
//
// SDD (the differentiable SpMM's backward, over A's CSC): while a slot
// accumulates dB[j] += a_q dC[i] it also has the gathered row dC[i] in
// registers, so the same pass writes dvalues at the nonzero's CSR position,
// dvalues[pos[q]] = <dC[i], B[j]> (B the forward's dense operand, its row j
// loaded once per item).  One gather of dC rows serves both gradients.
//
// The loop is software-pipelined: while the B-row loads of one trip are in
// flight a lane already loads the column ids / values of its next trip -- of
// the same row, or the first trip of its next row -- so the index stream's
// latency hides behind the gathers.  When every operand fits in 4 GiB
// (OFF32), addresses are 32-bit byte offsets from the (uniform) base
// pointers: one VGPR per address, which keeps occupancy at 8 waves/SIMD.
#include <mutex>

#include "fsp_device.hpp"
#include "fsp_host.hpp"
#include "fsp_plan.hpp"

namespace fsp {

// <x, y> over a lane's VEC columns in the accumulation type.  16-bit values
// go through v_dot2_f32_{bf16,f16}: two products and the running sum per
// instruction, where converting each element and an FMA would take four.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
template <typename A, typename T, int VEC>
__device__ __forceinline__ A row_dot(const vec_t<T, VEC>& x, const vec_t<T, VEC>& y) {
  if constexpr (std::is_same<T, __bf16>::value && VEC % 2 == 0) {
    float t = 0.0f;
#pragma unroll
    for (int v = 0; v < VEC; v += 2) {
      const bf16x2_t a = {x.v[v], x.v[v + 1]};
      const bf16x2_t c = {y.v[v], y.v[v + 1]};
      t = __builtin_amdgcn_fdot2_f32_bf16(a, c, t, false);
    }
    return t;
  } else if constexpr (std::is_same<T, _Float16>::value && VEC % 2 == 0) {
    float t = 0.0f;
#pragma unroll
    for (int v = 0; v < VEC; v += 2) {
      const f16x2_t a = {x.v[v], x.v[v + 1]};
      const f16x2_t c = {y.v[v], y.v[v + 1]};
      t = __builtin_amdgcn_fdot2(a, c, t, false);
    }
    return t;
  } else {
    A t = A(0);
#pragma unroll
    for (int v = 0; v < VEC; ++v) t += static_cast<A>(x.v[v]) * static_cast<A>(y.v[v]);
    return t;
  }
}

// The backward's extra operands (SDD): the forward's B (row j of the item),
// each nonzero's CSR position, and the dvalues output in CSR order.
template <typename T, typename QT>
struct sdd_args {
  const T* Bj;
  int64_t ldbj;
  const QT* pos;
  T* dvals;
};

template <typename T, typename IT, typename PT, typename QT, int VEC, int UNROLL, bool PERM, bool OFF32, bool SDD>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmm_bins_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ data,
    const QT* __restrict__ perm, const T* __restrict__ B, int64_t ldb, T* __restrict__ C, int64_t ldc,
    int64_t n_dense, int g_log2, typename acc_of<T>::type* __restrict__ partials, launch_table tab,
    sdd_args<T, QT> sd) {
  using A = typename acc_of<T>::type;
  using O = typename addr<OFF32>::off_t;  // element / byte offsets
  using P = typename std::conditional<OFF32, int32_t, int64_t>::type;  // nonzero positions
  const int k = find_bin(tab, blockIdx.x);
  const launch_bin lb = tab.b[k];  // a copy: the fields stay in registers
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
  const O col0 = static_cast<O>((static_cast<int64_t>(blockIdx.y) * G + gl) * VEC);
  const bool col_ok = static_cast<int64_t>(col0) < n_dense;
  // this workgroup's tiles: first, first + step, ... < end
  const tile_range tr = bin_tiles(lb, per_block);
  const int64_t rr_step = tr.step * per_block;
  const O ldb_o = static_cast<O>(ldb);
  const O ldc_o = static_cast<O>(ldc);
  const O nd_o = static_cast<O>(n_dense);

  // Every load below is unconditional (indices clamped into range, results
  // masked by selects) so the compiler can count outstanding loads exactly
  // and wait only for the B rows before the FMAs, not for the prefetch.
  const O col_ld = col_ok ? col0 : O(0);  // lanes past n_dense read column 0
  const int64_t last_item = lb.count - 1;  // bins in the table are non-empty

  // Item (row or segment) of this lane: output slot, nonzero range, and the
  // row it belongs to (its B row under SDD).  Under SDD the lanes past
  // n_dense run their slot's nonzeros too (reading column 0, contributing 0
  // to the dots): a dot's lane-group sum needs every lane of the group.
  auto item = [&](int64_t r, O& out, P& beg, P& end, O& jrow) {
    const bool live = r < lb.count && (SDD || col_ok);
    const int64_t rc = r < lb.count ? r : last_item;
    O o, jr;
    P b, e;
    if (lb.is_segment) {
      o = static_cast<O>(rc);
      jr = SDD ? static_cast<O>(tab.seg_row[rc]) : O(0);
      b = static_cast<P>(tab.seg_begin[rc]);
      e = static_cast<P>(tab.seg_end[rc]);
    } else {
      const int64_t row = lb.rows ? static_cast<int64_t>(lb.rows[rc]) : rc;
      o = static_cast<O>(row);
      jr = o;
      b = static_cast<P>(indptr[row]);
      e = static_cast<P>(indptr[row + 1]);
    }
    out = o;
    jrow = jr;
    beg = live ? b : P(0);
    end = live ? e : P(0);
  };
  // Column ids and values of nonzeros p, p+S, ...; past `end` the loads
  // repeat the item's last nonzero (a valid column of the same row).  They
  // stay raw: masking or converting them here would make the compiler wait
  // for every load in flight (the B rows included) right after the prefetch,
  // so the mask and the conversion are applied where they are consumed.
  auto trip = [&](P p, P end, IT (&c)[UNROLL], T (&a)[UNROLL], QT (&pq)[UNROLL]) {
    const P last = end > 0 ? end - 1 : P(0);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const P q = p + static_cast<P>(u) * S;
      const O qc = static_cast<O>(q < end ? q : last);
      c[u] = elem<OFF32>(indices, qc);
      const O qv = PERM ? static_cast<O>(elem<OFF32>(perm, qc)) : qc;
      a[u] = elem<OFF32>(data, qv);
      if constexpr (SDD) pq[u] = elem<OFF32>(sd.pos, qc);
    }
  };
  // Row j of the forward's B for this lane's columns (zero past n_dense).
  auto b_row = [&](O j) {
    vec_t<T, VEC> r;
    if constexpr (SDD) {
      r = *addr<OFF32>::at(reinterpret_cast<const vec_t<T, VEC>*>(sd.Bj),
                           (j * static_cast<O>(sd.ldbj) + col_ld) * static_cast<O>(sizeof(T)));
    }
    return r;
  };

  int64_t rr = tr.first * per_block + wave * per_wave + sub;
  O out, jrow;
  P beg, end;
  item(rr, out, beg, end, jrow);
  IT c[UNROLL];
  T a[UNROLL];
  QT pq[UNROLL];
  trip(beg + slot, end, c, a, pq);

  for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {  // wave-uniform
    const bool active = rr < lb.count && col_ok;
    const int64_t rr_next = rr + rr_step;
    // the next item is prefetched only when this workgroup will process it
    const bool has_next = tile + tr.step < tr.end;  // wave-uniform
    O out_n = 0, jrow_n = 0;
    P beg_n = 0, end_n = 0;
    if (has_next) item(rr_next, out_n, beg_n, end_n, jrow_n);
    vec_t<T, VEC> bj;
    if constexpr (SDD) bj = b_row(jrow);
    A acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = A(0);
    bool prefetched = false;
    for (P p = beg + slot; p < end; p += static_cast<P>(S) * UNROLL) {
      vec_t<T, VEC> b[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const O boff = (static_cast<O>(c[u]) * ldb_o + col_ld) * static_cast<O>(sizeof(T));
        b[u] = *addr<OFF32>::at(reinterpret_cast<const vec_t<T, VEC>*>(B), boff);
      }
      // prefetch: the next trip of this item, else the first trip of the next
      const P pn = p + static_cast<P>(S) * UNROLL;
      const bool same = pn < end;
      prefetched = prefetched || !same;
      T an[UNROLL];
      IT cn[UNROLL];
      QT pqn[UNROLL];
      if constexpr (SDD) {
        // unconditional (with nothing next it reloads this trip's nonzeros,
        // unused): with loads and stores under branches the compiler cannot
        // count what is outstanding and waits for all of it, the dvalue
        // stores included, every trip.  Without SDD the branch stays: there
        // the B loads issue as each pair of column ids arrives (5 % faster on
        // the 4096-matrix batch's short rows).
        trip(same ? pn : (has_next ? beg_n + slot : p), same ? end : (has_next ? end_n : end), cn, an, pqn);
      } else if (same || has_next) {
        trip(same ? pn : beg_n + slot, same ? end : end_n, cn, an, pqn);
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const bool ok = p + static_cast<P>(u) * S < end;
        const A av = ok ? static_cast<A>(a[u]) : A(0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const A bv = ok ? static_cast<A>(b[u].v[v]) : A(0);
          acc[v] += av * bv;
        }
      }
      if constexpr (SDD) {
        // dvalues of this trip's nonzeros: a lane's VEC products, then the
        // sum over the slot's G lanes (all of them run this trip)
        // Every lane of the slot stores the slot's total, unconditionally (a
        // store under a branch has the same effect on the waits as a load).
        // Past the item's end a lane's loads were clamped to the item's last
        // nonzero, so it writes that nonzero's own value to its own place.
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          const A d = group_sum(col_ok ? row_dot<A>(b[u], bj) : A(0), G);
          *addr<OFF32>::at(sd.dvals, static_cast<O>(pq[u]) * static_cast<O>(sizeof(T))) = static_cast<T>(d);
        }
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        c[u] = cn[u];
        a[u] = an[u];
        if constexpr (SDD) pq[u] = pqn[u];
      }
    }
    if (!prefetched && has_next) trip(beg_n + slot, end_n, c, a, pq);
    // Sum the S slots of a row: lanes that differ only in the slot bits.
    for (int j = 0; j < s_log2; ++j) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc[v] += shfl_xor(acc[v], G << j);
    }
    if (active && slot == 0) {
      if (lb.is_segment) {
        A* dst = addr<OFF32>::at(partials, (out * nd_o + col0) * static_cast<O>(sizeof(A)));
#pragma unroll
        for (int v = 0; v < VEC; ++v) dst[v] = acc[v];
      } else {
        vec_t<T, VEC> o;
#pragma unroll
        for (int v = 0; v < VEC; ++v) o.v[v] = static_cast<T>(acc[v]);
        *addr<OFF32>::at(reinterpret_cast<vec_t<T, VEC>*>(C), (out * ldc_o + col0) * static_cast<O>(sizeof(T))) = o;
      }
    }
    out = out_n;
    jrow = jrow_n;
    beg = beg_n;
    end = end_n;
    rr = rr_next;
  }
}

// C[long_rows[i], :] = sum of the row's segment partials, in segment order.
// One wave per long row (a grid-stride loop over rows): the row's segment
// range and output row are wave-uniform loads, lane l takes columns l,
// l + 64, ..., and each segment's partial row is one coalesced read whose
// address does not depend on the previous one.
template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmm_long_reduce_kernel(
    const typename acc_of<T>::type* __restrict__ partials, const int32_t* __restrict__ long_rows,
    const int64_t* __restrict__ seg_ptr, int64_t n_long, int64_t n_dense, T* __restrict__ C, int64_t ldc) {
  using A = typename acc_of<T>::type;
  const int lane = threadIdx.x & 63;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * FSP_WAVES + (threadIdx.x >> 6); i < n_long;
       i += static_cast<int64_t>(gridDim.x) * FSP_WAVES) {
    const int64_t s0 = seg_ptr[i], s1 = seg_ptr[i + 1];
    T* __restrict__ c_row = C + static_cast<int64_t>(long_rows[i]) * ldc;
    for (int64_t col = lane; col < n_dense; col += 64) {
      const A* __restrict__ p = partials + s0 * n_dense + col;
      A acc = A(0);
#pragma unroll 4
      for (int64_t sg = s0; sg < s1; ++sg, p += n_dense) acc += *p;
      c_row[col] = static_cast<T>(acc);
    }
  }
}

// C[r, :n_dense] = 0 for every row (the matrix has no nonzeros).
template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void fill_zero_rows_kernel(T* __restrict__ C, int64_t ldc,
                                                                        int64_t n_rows, int64_t n_dense) {
  const int64_t total = n_rows * n_dense;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    const int64_t r = t / n_dense;
    C[r * ldc + (t - r * n_dense)] = T(0);
  }
}

struct spmm_args {
  const void* indptr;
  const void* indices;
  const void* data;
  const void* perm;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  int64_t n_rows;
  int64_t n_dense;
  int lanes_n;
  int64_t nnz;
  int64_t b_rows;
  // the fused backward (SDD): forward B, CSR position per nonzero, dvalues
  const void* Bj = nullptr;
  int64_t ldbj = 0;
  const void* pos = nullptr;
  void* dvals = nullptr;
};

template <typename T, typename IT, typename PT, typename QT, int VEC, int UNROLL, bool PERM, bool OFF32, bool SDD>
static void launch_bins(const spmm_args& x, dim3 grid, int g_log2, typename acc_of<T>::type* partials,
                        const launch_table& tab, hipStream_t s) {
  const sdd_args<T, QT> sd{static_cast<const T*>(x.Bj), x.ldbj, static_cast<const QT*>(x.pos),
                           static_cast<T*>(x.dvals)};
  hipLaunchKernelGGL((spmm_bins_kernel<T, IT, PT, QT, VEC, UNROLL, PERM, OFF32, SDD>), grid, dim3(FSP_WG_THREADS), 0,
                     s, static_cast<const PT*>(x.indptr), static_cast<const IT*>(x.indices),
                     static_cast<const T*>(x.data), static_cast<const QT*>(x.perm), static_cast<const T*>(x.B),
                     x.ldb, static_cast<T*>(x.C), x.ldc, x.n_dense, g_log2, partials, tab, sd);
}

template <typename T, typename IT, typename PT, typename QT, int VEC, bool PERM, bool SDD>
static int launch_spmm(const spmm_args& x, const fsp_csr_plan* plan, hipStream_t s) {
  using A = typename acc_of<T>::type;
  const int64_t vcols = ceil_div(x.n_dense, VEC);
  int G = x.lanes_n;
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
  const int64_t partial_bytes = plan->n_segments * x.n_dense * (int64_t)sizeof(A);
  if (plan->n_segments > 0 && (plan->partials_dev == nullptr || plan->partials_bytes < partial_bytes)) {
    return fail(FSP_ERR_WORKSPACE, "fsp_spmm_csr: partials need %lld bytes", (long long)partial_bytes);
  }
  A* partials = static_cast<A*>(plan->partials_dev);
  const int64_t ytiles = ceil_div(vcols, G);
  if (ytiles > 65535) return fail(FSP_ERR_ARG, "fsp_spmm_csr: n_dense %lld too large", (long long)x.n_dense);
  if (SDD && ytiles != 1) {  // a dot needs all of a row's columns in one lane group
    return fail(FSP_ERR_ARG, "fsp_spmm_csr_trans_sddmm: n_dense %lld exceeds one lane group (%d columns)",
                (long long)x.n_dense, G * VEC);
  }
  if (blocks > 0) {
    const dim3 grid(static_cast<unsigned>(blocks), static_cast<unsigned>(ytiles));
    // 32-bit byte offsets when every array the kernel addresses fits in 4 GiB.
    const int64_t lim = int64_t(1) << 32;
    const bool off32 = std::is_same<PT, int32_t>::value &&
                       (!(PERM || SDD) || std::is_same<QT, int32_t>::value) &&
                       x.nnz * (int64_t)sizeof(T) < lim && x.nnz * (int64_t)sizeof(IT) < lim &&
                       x.nnz * (int64_t)sizeof(QT) < lim && x.b_rows * x.ldb * (int64_t)sizeof(T) < lim &&
                       x.n_rows * x.ldc * (int64_t)sizeof(T) < lim && partial_bytes < lim &&
                       (!SDD || x.n_rows * x.ldbj * (int64_t)sizeof(T) < lim);
    if (off32) {
      using P32 = int32_t;
      using Q32 = int32_t;
      if (plan->unroll >= 8) {
        launch_bins<T, IT, P32, Q32, VEC, 8, PERM, true, SDD>(x, grid, g_log2, partials, tab, s);
      } else {
        launch_bins<T, IT, P32, Q32, VEC, 4, PERM, true, SDD>(x, grid, g_log2, partials, tab, s);
      }
    } else {
      launch_bins<T, IT, PT, QT, VEC, 4, PERM, false, SDD>(x, grid, g_log2, partials, tab, s);
    }
    const int st = check_launch("spmm_bins");
    if (st) return st;
  }
  if (plan->n_long_rows > 0) {
    int64_t g = ceil_div(plan->n_long_rows, FSP_WAVES);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(spmm_long_reduce_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                       partials, plan->long_rows_dev, plan->long_seg_ptr_dev, plan->n_long_rows, x.n_dense,
                       static_cast<T*>(x.C), x.ldc);
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

template <typename T, typename IT, typename PT, typename QT, bool PERM, bool SDD = false>
static int launch_spmm_vec(spmm_args x, const fsp_csr_plan* plan, hipStream_t s) {
  constexpr int V = 16 / sizeof(T);
  const bool wide = wide_ok<T>(x.B, x.ldb, x.C, x.ldc, x.n_dense) &&
                    (!SDD || (reinterpret_cast<uintptr_t>(x.Bj) % 16 == 0 && x.ldbj % V == 0));
  if (wide) return launch_spmm<T, IT, PT, QT, V, PERM, SDD>(x, plan, s);
  x.lanes_n = 0;
  return launch_spmm<T, IT, PT, QT, 1, PERM, SDD>(x, plan, s);
}

}  // namespace fsp

// Scratch of the no-allocation SpMM entry: the long-row partials live in
// this file-scope device array (one copy per GPU, held until the process
// ends).  Host code never passes the array's name to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, resolved once per device
// with that device current and kept in scratch_dev.
static __device__ __align__(256) unsigned char scratch[FSP_NOALLOC_SCRATCH_BYTES];

namespace fsp {

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxScratchDevices];

// Device address of `scratch` on the current device (resolved on first use).
static int current_scratch(void* stream, void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FSP_ERR_LAUNCH, "fsp_spmm_csr_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(FSP_ERR_ARG, "fsp_spmm_csr_noalloc: device %d", dev);
  if (stream != nullptr) {
    int sdev = -1;
    e = hipStreamGetDevice(as_stream(stream), &sdev);
    if (e != hipSuccess) {
      return fail(FSP_ERR_LAUNCH, "fsp_spmm_csr_noalloc: hipStreamGetDevice: %s", hipGetErrorString(e));
    }
    if (sdev != dev) {
      return fail(FSP_ERR_ARG, "fsp_spmm_csr_noalloc: stream belongs to device %d, current device is %d", sdev, dev);
    }
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      return fail(FSP_ERR_LAUNCH, "fsp_spmm_csr_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return FSP_OK;
}

static int spmm_entry(int32_t value_type, int32_t index_type, int32_t indptr_type, const void* indptr_dev,
                      const void* indices_dev, const void* data_dev, int32_t perm_type, const void* perm_dev,
                      const void* B_dev, int64_t ldb, void* C_dev, int64_t ldc, int64_t n_rows, int64_t n_dense,
                      int64_t nnz, int64_t b_rows, int32_t lanes_n, const fsp_csr_plan* plan, void* stream) {  if (n_rows < 0 || n_dense < 0 || nnz < 0 || b_rows < 0) return fail(FSP_ERR_ARG, "fsp_spmm_csr: negative size");
  int st = check_plan(plan, n_rows);
  if (st) return st;
  if (n_rows == 0 || n_dense == 0) return FSP_OK;
  if (indptr_dev == nullptr || B_dev == nullptr || C_dev == nullptr) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr: NULL pointer");
  }
  if (ldb < n_dense || ldc < n_dense) return fail(FSP_ERR_ARG, "fsp_spmm_csr: leading dimension < n_dense");
  const hipStream_t s = as_stream(stream);
  const bool use_perm = perm_dev != nullptr;
  spmm_args x{indptr_dev, indices_dev, data_dev, perm_dev, B_dev, ldb, C_dev, ldc, n_rows, n_dense, lanes_n, nnz, b_rows};
  if (nnz == 0) {
    // The kernels read clamped in-range indices; with no nonzeros there are
    // none, so the all-zero result is written directly.
    FSP_DISPATCH_VALUE(value_type, T, {
      int64_t g = ceil_div(n_rows * n_dense, FSP_WG_THREADS);
      if (g > 65536) g = 65536;
      hipLaunchKernelGGL(fill_zero_rows_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<T*>(C_dev), ldc, n_rows, n_dense);
      return check_launch("fill_zero_rows");
    });
  }
  if (indices_dev == nullptr || data_dev == nullptr || b_rows == 0) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr: NULL indices/data or empty B with nonzeros");
  }
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        if (!use_perm) {
          return launch_spmm_vec<T, IT, PT, int32_t, false>(x, plan, s);
        }
        FSP_DISPATCH_INDEX(perm_type, QT, {
          return launch_spmm_vec<T, IT, PT, QT, true>(x, plan, s);
        });
      });
    });
  });
  return FSP_OK;
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_spmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type, const void* indptr_dev,
                            const void* indices_dev, const void* data_dev, int32_t perm_type, const void* perm_dev,
                            const void* B_dev, int64_t ldb, void* C_dev, int64_t ldc, int64_t n_rows,
                            int64_t n_dense, int64_t nnz, int64_t b_rows, int32_t lanes_n,
                            const fsp_csr_plan* plan, void* stream) {
  return spmm_entry(value_type, index_type, indptr_type, indptr_dev, indices_dev, data_dev, perm_type, perm_dev,
                    B_dev, ldb, C_dev, ldc, n_rows, n_dense, nnz, b_rows, lanes_n, plan, stream);
}

extern "C" int fsp_spmm_csr_trans_sddmm(int32_t value_type, int32_t index_type, int32_t indptr_type,
                                        const void* colptr_dev, const void* rowind_dev, const void* data_csc_dev,
                                        int32_t pos_type, const void* pos_dev, const void* dC_dev, int64_t lddc,
                                        const void* B_dev, int64_t ldb, void* dB_dev, int64_t lddb,
                                        void* dvalues_dev, int64_t n_cols, int64_t n_dense, int64_t nnz,
                                        int64_t n_rows, int32_t lanes_n, const fsp_csr_plan* plan, void* stream) {
  if (n_cols < 0 || n_dense < 0 || nnz < 0 || n_rows < 0) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr_trans_sddmm: negative size");
  }
  int st = check_plan(plan, n_cols);
  if (st) return st;
  if (colptr_dev == nullptr || (nnz > 0 && (rowind_dev == nullptr || data_csc_dev == nullptr || pos_dev == nullptr ||
                                            dvalues_dev == nullptr || dC_dev == nullptr || B_dev == nullptr)) ||
      (n_cols > 0 && n_dense > 0 && dB_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr_trans_sddmm: NULL pointer");
  }
  if (lddc < n_dense || ldb < n_dense || lddb < n_dense) {
    return fail(FSP_ERR_ARG, "fsp_spmm_csr_trans_sddmm: leading dimension < n_dense");
  }
  const hipStream_t s = as_stream(stream);
  if (nnz == 0 || n_dense == 0) {
    // no nonzeros: dB = 0; no dense columns: every dvalue is an empty dot, 0
    FSP_DISPATCH_VALUE(value_type, T, {
      if (n_cols > 0 && n_dense > 0) {
        int64_t g = ceil_div(n_cols * n_dense, FSP_WG_THREADS);
        if (g > 65536) g = 65536;
        hipLaunchKernelGGL(fill_zero_rows_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                           static_cast<T*>(dB_dev), lddb, n_cols, n_dense);
        st = check_launch("fill_zero_rows");
        if (st) return st;
      }
      if (nnz > 0) {
        int64_t g = ceil_div(nnz, FSP_WG_THREADS);
        if (g > 65536) g = 65536;
        hipLaunchKernelGGL(fill_zero_rows_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                           static_cast<T*>(dvalues_dev), int64_t(1), nnz, int64_t(1));
        return check_launch("fill_zero_rows");
      }
      return FSP_OK;
    });
  }
  if (n_cols == 0) return FSP_OK;  // nnz > 0 needs columns: unreachable for a valid CSC
  spmm_args x{colptr_dev, rowind_dev, data_csc_dev, nullptr, dC_dev, lddc, dB_dev, lddb, n_cols, n_dense, lanes_n,
              nnz, n_rows};
  x.Bj = B_dev;
  x.ldbj = ldb;
  x.pos = pos_dev;
  x.dvals = dvalues_dev;
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        FSP_DISPATCH_INDEX(pos_type, QT, {
          return launch_spmm_vec<T, IT, PT, QT, false, true>(x, plan, s);
        });
      });
    });
  });
  return FSP_OK;
}

extern "C" int fsp_spmm_csr_noalloc(int32_t value_type, int32_t index_type, int32_t indptr_type,
                                    const void* indptr_dev, const void* indices_dev, const void* data_dev,
                                    int32_t perm_type, const void* perm_dev, const void* B_dev, int64_t ldb,
                                    void* C_dev, int64_t ldc, int64_t n_rows, int64_t n_dense, int64_t nnz,
                                    int64_t b_rows, int32_t lanes_n, const fsp_csr_plan* plan, void* stream) {
  if (plan == nullptr) return fail(FSP_ERR_ARG, "fsp_spmm_csr_noalloc: plan is NULL");
  if (n_dense < 0 || plan->n_segments < 0) return fail(FSP_ERR_ARG, "fsp_spmm_csr_noalloc: negative size");
  const int64_t acc_bytes = value_type == FSP_F64 ? 8 : 4;
  // n_segments * n_dense * acc_bytes <= the scratch size, without overflow
  if (plan->n_segments > 0 && n_dense > 0 &&
      plan->n_segments > FSP_NOALLOC_SCRATCH_BYTES / acc_bytes / n_dense) {
    return fail(FSP_ERR_WORKSPACE,
                "fsp_spmm_csr_noalloc: %lld long-row segments x %lld dense columns x %lld bytes exceed the %lld-byte "
                "static scratch; use fsp_spmm_csr with a caller-supplied partials buffer",
                (long long)plan->n_segments, (long long)n_dense, (long long)acc_bytes,
                (long long)FSP_NOALLOC_SCRATCH_BYTES);
  }
  fsp_csr_plan p = *plan;
  p.partials_dev = nullptr;
  p.partials_bytes = 0;
  if (p.n_segments > 0 && n_rows > 0 && n_dense > 0 && nnz > 0) {
    const int st = current_scratch(stream, &p.partials_dev);
    if (st) return st;
    p.partials_bytes = FSP_NOALLOC_SCRATCH_BYTES;
  }
  return spmm_entry(value_type, index_type, indptr_type, indptr_dev, indices_dev, data_dev, perm_type, perm_dev,
                    B_dev, ldb, C_dev, ldc, n_rows, n_dense, nnz, b_rows, lanes_n, &p, stream);
}
