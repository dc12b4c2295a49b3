// CSR SpMV y = A x on gfx950.
//
// Rows are binned by length; a bin with lanes L puts 64/L rows on one wave and
// splits each row's nonzeros over L lanes, then a DPP group sum.  Rows longer
// than the last bin are cut into segments; one wave reduces each segment into
// a partial and a second pass adds a row's partials in segment order (no
// atomics: the result is deterministic for a plan).  One fused launch runs
// every bin and the segments with the loop plan.variant selects:
//   0  spmv_bins_kernel: lane l takes nonzeros l, l+L, ..., UNROLL loads in
//      flight, software-pipelined (the next trip's column ids / values load
//      while this trip's x gathers are in flight);
//   1  spmv_simple_kernel: the same order, plain loop;
//   2  spmv_simple_kernel and 3 spmv_bins_kernel with x staged in LDS by
//      1024-thread workgroups;
//   4  spmv_vec_kernel: the index / value streams read as 4-element chunks,
//      x in LDS; a plan whose rows all have one length (plan.row_len) runs
//      its own instantiation that computes the row bounds and, when a row
//      fits one trip, software-pipelines the row tiles.
// Variants 0-3 are bitwise equal, and so are variant 4's two forms.  spmv_parts_kernel (fsp_spmv_csr_parts)
// runs row-range parts, each with its own column window of x in LDS (the
// batched SpMV).
#include <atomic>

#include "fsp_device.hpp"
#include "fsp_host.hpp"
#include "fsp_plan.hpp"

namespace fsp {

// x staged in LDS (plan.variant 2 and 3), for x of at most kLdsxBytes: each
// 1024-thread workgroup copies x into LDS once and the loop then gathers x
// from LDS.  Through the vector L1 every gathered element is a separate
// cache-line lookup, and for short rows those lookups, not HBM, bound the
// loop (the x reads are as many as the nonzeros); from LDS the index / value
// stream runs near the device copy rate (measured on MI355X).  The grid is
// persistent (plan.max_blocks, default two workgroups per CU) so x is loaded
// once per workgroup.  Per lane the arithmetic is unchanged: results are
// bitwise equal to the loops that read x from global memory.
constexpr int kLdsxWaves = 16;
constexpr int kLdsxThreads = 64 * kLdsxWaves;
constexpr int64_t kLdsxBytes = 65536;

// Vector of BYTES bytes as 32-bit words (one global / LDS access).
template <int BYTES>
struct chunk_words;
template <>
struct chunk_words<8> {
  typedef uint32_t type __attribute__((ext_vector_type(2)));
};
template <>
struct chunk_words<16> {
  typedef uint32_t type __attribute__((ext_vector_type(4)));
};
template <>
struct chunk_words<32> {
  typedef uint32_t type __attribute__((ext_vector_type(8)));
};

// Unsigned integer of BYTES bytes (element-wise copies move raw words).
template <int BYTES>
struct word_of;
template <>
struct word_of<2> {
  typedef uint16_t type;
};
template <>
struct word_of<4> {
  typedef uint32_t type;
};
template <>
struct word_of<8> {
  typedef uint64_t type;
};

// An empty asm that reads and writes v: the load producing v completes here
// (the compiler waits for it before the asm) and cannot sink below it.
template <typename V>
__device__ __forceinline__ void pin(V& v) {
  asm volatile("" : "+v"(v));
}

// dst[0, n) = src[0, n) by a 1024-thread workgroup, for n * sizeof(E) <=
// kLdsxBytes (dst in LDS, 16-byte aligned).  Every load of a thread is issued
// before its first LDS write, so the copy costs one memory round trip: a
// plain strided loop waits for each load before its write, one round trip per
// 4 KiB pass (16 for a 64 KiB window).  The loads are unconditional, with
// indices clamped to the last element, and pinned before the guarded writes:
// the compiler would otherwise sink a load into its write's guard, where it
// waits for every outstanding load.  The writes stay guarded (lanes past the
// end writing one LDS address would conflict).  A 16-byte aligned source
// moves in 16-byte vectors (at most four per thread) plus a tail of fewer
// than one vector; an unaligned source moves element by element, at most
// eight loads per thread at a time.  Indices are 32-bit (n <= 32768), so an
// address is one register over the uniform base: the copy must not raise
// the loop kernels' registers past what two workgroups per CU allow.
// No barrier: the caller's __syncthreads() publishes dst.
template <typename E>
__device__ __forceinline__ void copy_to_lds(const E* __restrict__ src, E* dst, int64_t n64) {
  if (n64 <= 0) return;
  const int n = static_cast<int>(n64);
  const int t = static_cast<int>(threadIdx.x);
  if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    using W = typename chunk_words<16>::type;
    constexpr int kPer = static_cast<int>(kLdsxBytes / (16 * kLdsxThreads));
    constexpr int kElems = 16 / static_cast<int>(sizeof(E));
    const int nv = n / kElems;
    if (nv > 0) {
      const W* s = reinterpret_cast<const W*>(src);
      W* d = reinterpret_cast<W*>(dst);
      W v[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = t + k * kLdsxThreads;
        v[k] = s[i < nv ? i : nv - 1];
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) pin(v[k]);
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = t + k * kLdsxThreads;
        if (i < nv) d[i] = v[k];
      }
    }
    const int done = nv * kElems;
    if (t < n - done) dst[done + t] = src[done + t];  // fewer than kElems elements
    return;
  }
  using U = typename word_of<sizeof(E)>::type;
  constexpr int kPer = 8;
  constexpr int kStep = kPer * kLdsxThreads;
  const U* s = reinterpret_cast<const U*>(src);
  U* d = reinterpret_cast<U*>(dst);
  for (int base = 0; base < n; base += kStep) {  // uniform
    U v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = base + t + k * kLdsxThreads;
      v[k] = s[i < n ? i : n - 1];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) pin(v[k]);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = base + t + k * kLdsxThreads;
      if (i < n) d[i] = v[k];
    }
  }
}

// Pointer to x for the loop: LDS (after the workgroup's copy) or global.
template <typename T, bool LDSX>
__device__ __forceinline__ const T* stage_x(const T* __restrict__ x, int64_t x_len) {
  if constexpr (LDSX) {
    extern __shared__ __align__(16) unsigned char fsp_ldsx_mem[];
    T* xs = reinterpret_cast<T*>(fsp_ldsx_mem);
    copy_to_lds(x, xs, x_len);
    __syncthreads();
    return xs;
  } else {
    return x;
  }
}

template <typename T, typename IT, typename PT, typename QT, int UNROLL, bool PERM, int WAVES, bool LDSX>
__global__ __launch_bounds__(64 * WAVES) void spmv_bins_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ data,
    const QT* __restrict__ perm, const T* __restrict__ x_in, T* __restrict__ y,
    typename spmv_acc_of<T>::type* __restrict__ partials, launch_table tab, int64_t x_len) {
  using A = typename spmv_acc_of<T>::type;
  static_assert(!LDSX || WAVES == kLdsxWaves, "the LDS copy of x runs on 1024-thread workgroups");
  const T* __restrict__ x = stage_x<T, LDSX>(x_in, x_len);
  const int k = find_bin(tab, blockIdx.x);
  const launch_bin lb = tab.b[k];  // a copy: the fields stay in registers
  const int ll = lb.lanes_log2;
  const int L = 1 << ll;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane >> ll;
  const int l = lane & (L - 1);
  const int64_t per_wave = 64 >> ll;
  const int64_t per_block = per_wave * WAVES;
  const tile_range tr = bin_tiles(lb, per_block);
  const int64_t rr_step = tr.step * per_block;
  const int64_t last_item = lb.count - 1;  // bins in the table are non-empty

  // Item of this lane: output slot and nonzero range (empty past the bin).
  auto item = [&](int64_t r, int64_t& out, int64_t& beg, int64_t& end) {
    const bool live = r < lb.count;
    const work_item w = load_item(tab, lb, indptr, live ? r : last_item);
    out = w.out;
    beg = live ? w.begin : 0;
    end = live ? w.end : 0;
  };
  // Column ids and values of nonzeros p, p+L, ...; past `end` the loads repeat
  // the item's last nonzero.  Raw: the mask is applied at the FMA, so the
  // prefetch does not make the compiler wait for the x gathers in flight.
  auto trip = [&](int64_t p, int64_t end, IT (&c)[UNROLL], T (&a)[UNROLL]) {
    const int64_t lastnz = end > 0 ? end - 1 : 0;
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const int64_t q = p + static_cast<int64_t>(u) * L;
      const int64_t qc = q < end ? q : lastnz;
      c[u] = indices[qc];
      a[u] = data[PERM ? static_cast<int64_t>(perm[qc]) : qc];
    }
  };

  int64_t rr = tr.first * per_block + wave * per_wave + sub;
  int64_t out, beg, end;
  item(rr, out, beg, end);
  IT c[UNROLL];
  T a[UNROLL];
  trip(beg + l, end, c, a);
  for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {  // wave-uniform
    const bool active = rr < lb.count;
    const int64_t rr_next = rr + rr_step;
    // the next item is prefetched only when this workgroup will process it
    const bool has_next = tile + tr.step < tr.end;  // wave-uniform
    int64_t out_n = 0, beg_n = 0, end_n = 0;
    if (has_next) item(rr_next, out_n, beg_n, end_n);
    A acc = A(0);
    bool prefetched = false;
    for (int64_t p = beg + l; p < end; p += static_cast<int64_t>(L) * UNROLL) {
      T xv[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) xv[u] = x[static_cast<int64_t>(c[u])];
      // prefetch: the next trip of this item, else the first trip of the next
      const int64_t pn = p + static_cast<int64_t>(L) * UNROLL;
      const bool same = pn < end;
      prefetched = prefetched || !same;
      IT cn[UNROLL];
      T an[UNROLL];
      if (same || has_next) trip(same ? pn : beg_n + l, same ? end : end_n, cn, an);
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const bool ok = p + static_cast<int64_t>(u) * L < end;
        const A t = fmadd(static_cast<A>(a[u]), static_cast<A>(xv[u]), acc);
        acc = ok ? t : acc;
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        c[u] = cn[u];
        a[u] = an[u];
      }
    }
    if (!prefetched && has_next) trip(beg_n + l, end_n, c, a);
    acc = group_sum(acc, L);
    if (active && l == 0) {
      if (lb.is_segment) {
        partials[out] = acc;
      } else {
        y[out] = static_cast<T>(acc);
      }
    }
    rr = rr_next;
    out = out_n;
    beg = beg_n;
    end = end_n;
  }
}

// The plain loop (plan.variant = 1): every trip loads its column ids / values
// and then gathers x, with no prefetch across trips.  For short uniform rows
// under a persistent grid it streams the index / value arrays faster than the
// pipelined loop (measured on MI355X); the autotuner picks between the two.
template <typename T, typename IT, typename PT, typename QT, int UNROLL, bool PERM, int WAVES, bool LDSX>
__global__ __launch_bounds__(64 * WAVES) void spmv_simple_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ data,
    const QT* __restrict__ perm, const T* __restrict__ x_in, T* __restrict__ y,
    typename spmv_acc_of<T>::type* __restrict__ partials, launch_table tab, int64_t x_len) {
  using A = typename spmv_acc_of<T>::type;
  static_assert(!LDSX || WAVES == kLdsxWaves, "the LDS copy of x runs on 1024-thread workgroups");
  const T* __restrict__ x = stage_x<T, LDSX>(x_in, x_len);
  const int k = find_bin(tab, blockIdx.x);
  const launch_bin lb = tab.b[k];  // a copy: the fields stay in registers
  const int ll = lb.lanes_log2;
  const int L = 1 << ll;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane >> ll;
  const int l = lane & (L - 1);
  const int64_t per_wave = 64 >> ll;
  const int64_t per_block = per_wave * WAVES;
  const tile_range tr = bin_tiles(lb, per_block);
  for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {
    const int64_t rr = tile * per_block + wave * per_wave + sub;
    const bool active = rr < lb.count;
    A acc = A(0);
    work_item w{0, 0, 0, 0};
    if (active) {
      w = load_item(tab, lb, indptr, rr);
      int64_t p = w.begin + l;
      for (; p + (UNROLL - 1) * L < w.end; p += UNROLL * L) {
        int64_t c[UNROLL];
        T a[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          const int64_t q = p + u * L;
          c[u] = static_cast<int64_t>(indices[q]);
          a[u] = data[PERM ? static_cast<int64_t>(perm[q]) : q];
        }
        T xv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) acc = fmadd(static_cast<A>(a[u]), static_cast<A>(xv[u]), acc);
      }
      for (; p < w.end; p += L) {
        const int64_t q = PERM ? static_cast<int64_t>(perm[p]) : p;
        acc = fmadd(static_cast<A>(data[q]), static_cast<A>(x[static_cast<int64_t>(indices[p])]), acc);
      }
    }
    acc = group_sum(acc, L);
    if (active && l == 0) {
      if (lb.is_segment) {
        partials[w.out] = acc;
      } else {
        y[w.out] = static_cast<T>(acc);
      }
    }
  }
}

// Vector-chunk loop with x in LDS (plan.variant = 4).  The index / value
// (and, for a CSC launch, permutation) streams are read as aligned chunks of
// four elements -- one 16-byte load per lane for 32-bit data -- with the
// nontemporal hint: they are read once, so they should not displace
// anything from the caches.  The L lanes of a row take chunks l, l+L, ... of
// the aligned span covering [begin, end); elements outside the row are
// masked at the FMA.  With whole cache lines per load instruction the stream
// runs near the device's read bandwidth (2^24 x 16 SpMV: 5.95 TB/s in a probe
// against 4.98 TB/s for 4-byte lane loads, measured on MI355X).  A chunk
// that would reach past the arrays' end falls back to guarded element loads.
// Within a lane, terms are added chunk by chunk (element order inside a
// chunk): deterministic, but not the order of variants 0-3.
template <typename E>
struct chunk4 {
  E v[4];
};

// Four consecutive elements at p (p aligned to 4 * sizeof(E)), streamed.
template <typename E>
__device__ __forceinline__ chunk4<E> load_chunk_nt(const E* p) {
  using W = typename chunk_words<4 * sizeof(E)>::type;
  const W w = __builtin_nontemporal_load(reinterpret_cast<const W*>(p));
  return __builtin_bit_cast(chunk4<E>, w);
}

// Elements q .. q+3 of p with those at or past n read as zero.
template <typename E>
__device__ __forceinline__ chunk4<E> load_chunk_tail(const E* p, int64_t q, int64_t n) {
  chunk4<E> c;
#pragma unroll
  for (int j = 0; j < 4; ++j) c.v[j] = q + j < n ? p[q + j] : E(0);
  return c;
}

// Rows of one bin, L lanes per row (a compile-time constant, so the row's
// group sum is a fixed DPP sequence and the chunk strides are immediates).
// REG: an identity bin of a constant-length plan (row_len > 0): the row
// bounds are computed, so the first wait of a tile is for its chunks.  A
// separate instantiation because a load that may or may not have happened
// before the chunk loads makes the compiler wait for every outstanding
// access there, the previous tile's y store included: a second round trip
// per tile.
template <int L, typename T, typename IT, typename PT, typename QT, int CHUNKS, bool PERM, bool REG>
__device__ __forceinline__ void spmv_vec_rows(const launch_bin& lb, const launch_table& tab,
                                              const PT* __restrict__ indptr, const IT* __restrict__ indices,
                                              const T* __restrict__ data, const QT* __restrict__ perm,
                                              const T* __restrict__ x, T* __restrict__ y,
                                              typename spmv_acc_of<T>::type* __restrict__ partials, int64_t nnz) {
  using A = typename spmv_acc_of<T>::type;
  constexpr int64_t per_wave = 64 / L;
  constexpr int64_t per_block = per_wave * kLdsxWaves;
  constexpr int64_t stride = 4 * L * CHUNKS;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / L;
  const int l = lane % L;
  const tile_range tr = bin_tiles(lb, per_block);
  for (int64_t tile = tr.first; tile < tr.end; tile += tr.step) {
    const int64_t rr = tile * per_block + wave * per_wave + sub;
    const bool active = rr < lb.count;
    A acc = A(0);
    work_item w{0, 0, 0, 0};
    if (REG) {
      w.row = rr;
      w.begin = tab.indptr_base + rr * tab.row_len;
      w.end = w.begin + tab.row_len;
      w.out = rr;
    } else if (active) {
      w = load_item(tab, lb, indptr, rr);
    }
    // no lane's chunks reach past the arrays' end: whole-chunk loads only (a
    // wave-uniform choice).  A lane's chunks start below w.end, 4-aligned, and
    // a trip reaches 4 * L * (CHUNKS - 1) elements further.
    const bool whole = __all(!active || w.end + 3 + 4 * L * (CHUNKS - 1) <= nnz);
    if (active) {
      const int64_t a0 = w.begin & ~int64_t(3);
      for (int64_t q = a0 + 4 * l; q < w.end; q += stride) {
        chunk4<IT> c[CHUNKS];
        chunk4<T> v[CHUNKS];
#pragma unroll
        for (int u = 0; u < CHUNKS; ++u) {
          const int64_t qu = q + 4 * L * u;
          if (whole || qu + 4 <= nnz) {
            c[u] = load_chunk_nt(indices + qu);
            if (PERM) {
              const chunk4<QT> pq = load_chunk_nt(perm + qu);
#pragma unroll
              for (int j = 0; j < 4; ++j) v[u].v[j] = data[static_cast<int64_t>(pq.v[j])];
            } else {
              v[u] = load_chunk_nt(data + qu);
            }
          } else {
            c[u] = load_chunk_tail(indices, qu, nnz);
            if (PERM) {
              const chunk4<QT> pq = load_chunk_tail(perm, qu, nnz);
#pragma unroll
              for (int j = 0; j < 4; ++j) v[u].v[j] = qu + j < nnz ? data[static_cast<int64_t>(pq.v[j])] : T(0);
            } else {
              v[u] = load_chunk_tail(data, qu, nnz);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < CHUNKS; ++u) {
          const int64_t qu = q + 4 * L * u;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool ok = qu + j >= w.begin && qu + j < w.end;
            const A t = fmadd(static_cast<A>(v[u].v[j]), static_cast<A>(x[static_cast<int64_t>(c[u].v[j])]), acc);
            acc = ok ? t : acc;
          }
        }
      }
    }
    acc = group_sum(acc, L);
    if (active && l == 0) {
      if (lb.is_segment) {
        partials[w.out] = acc;
      } else {
        y[w.out] = static_cast<T>(acc);
      }
    }
  }
}

// Identity bin of a constant-length plan whose rows each fit one trip of
// the lane group (indptr_base and row_len multiples of 4, row_len <=
// 4 * L * CHUNKS, every row inside the arrays): software-pipelined over row
// tiles.  The next tile's chunks load before this tile's are summed, from
// addresses that need no load, so a wave keeps one tile in flight while it
// works on another.  As in the pipelined parts loop, the waits come out
// exact only because (1) the loads are unconditional (a finished or short
// lane reloads a chunk it already holds, a row past the bin's end reloads
// the last row, and the FMA masks them); (2) the y store is unconditional
// (every lane of a group holds the group total, bitwise identical, and a row
// past the end stores the last row's own total to the last row); (3) two
// register buffers alternate, with the first tile peeled, so no loaded
// register is copied and every edge into the loop ends with a store; an
// empty asm with a memory clobber keeps the next tile's loads ahead of this
// tile's x reads.  Per lane the terms are added in the order of
// spmv_vec_rows (bitwise equal).
template <typename T, typename IT, int CHUNKS>
struct vec_tile {
  chunk4<IT> c[CHUNKS];
  chunk4<T> v[CHUNKS];
};

template <int L, typename T, typename IT, int CHUNKS>
__device__ __forceinline__ void spmv_vec_rows_pipe(const launch_bin& lb, const launch_table& tab,
                                                   const IT* __restrict__ indices, const T* __restrict__ data,
                                                   const T* __restrict__ x, T* __restrict__ y) {
  using A = typename spmv_acc_of<T>::type;
  constexpr int64_t per_wave = 64 / L;
  constexpr int64_t per_block = per_wave * kLdsxWaves;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / L;
  const int l = lane % L;
  const int64_t len = tab.row_len;
  const int64_t last = lb.count - 1;
  const tile_range tr = bin_tiles(lb, per_block);
  if (tr.first >= tr.end) return;  // uniform over the workgroup
  int64_t off[CHUNKS];  // this lane's chunk offsets inside a row (clamped to 0 past its end)
#pragma unroll
  for (int u = 0; u < CHUNKS; ++u) {
    const int64_t o = 4 * l + 4 * L * u;
    off[u] = o < len ? o : 0;
  }
  auto row_of = [&](int64_t tile) {
    const int64_t rr = tile * per_block + wave * per_wave + sub;
    return rr < last ? rr : last;
  };
  auto load = [&](int64_t tile, vec_tile<T, IT, CHUNKS>& f) {
    const int64_t b = tab.indptr_base + row_of(tile) * len;
#pragma unroll
    for (int u = 0; u < CHUNKS; ++u) {
      f.c[u] = load_chunk_nt(indices + b + off[u]);
      f.v[u] = load_chunk_nt(data + b + off[u]);
    }
  };
  // sums tile `tile` from f after loading tile `next` into g
  auto pass = [&](int64_t tile, int64_t next, const vec_tile<T, IT, CHUNKS>& f, vec_tile<T, IT, CHUNKS>& g) {
    load(next, g);
    asm volatile("" ::: "memory");
    A acc = A(0);
#pragma unroll
    for (int u = 0; u < CHUNKS; ++u) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool ok = 4 * l + 4 * L * u + j < len;
        const A t = fmadd(static_cast<A>(f.v[u].v[j]), static_cast<A>(x[static_cast<int64_t>(f.c[u].v[j])]), acc);
        acc = ok ? t : acc;
      }
    }
    acc = group_sum(acc, L);
    y[row_of(tile)] = static_cast<T>(acc);
  };
  auto after = [&](int64_t tile) { return tile + tr.step < tr.end ? tile + tr.step : tile; };
  vec_tile<T, IT, CHUNKS> f0, f1;
  load(tr.first, f0);
  int64_t tile = tr.first;
  pass(tile, after(tile), f0, f1);  // peeled
  tile += tr.step;
  while (tile < tr.end) {  // uniform over the workgroup
    pass(tile, after(tile), f1, f0);
    tile += tr.step;
    if (tile >= tr.end) break;
    pass(tile, after(tile), f0, f1);
    tile += tr.step;
  }
}

// REG: a launch whose bins are all identity bins of a constant-length plan
// (tab.row_len > 0, no segments, no permutation; the host checks).  Its own
// kernel, so the general one keeps its register budget: the computed-bounds
// paths raise the scalar registers past what two 16-wave workgroups per CU
// allow.
template <typename T, typename IT, typename PT, typename QT, int CHUNKS, bool PERM, bool REG>
__global__ __launch_bounds__(kLdsxThreads) void spmv_vec_kernel(
    const PT* __restrict__ indptr, const IT* __restrict__ indices, const T* __restrict__ data,
    const QT* __restrict__ perm, const T* __restrict__ x_in, T* __restrict__ y,
    typename spmv_acc_of<T>::type* __restrict__ partials, launch_table tab, int64_t x_len, int64_t nnz) {
  const T* __restrict__ x = stage_x<T, true>(x_in, x_len);
  const launch_bin lb = tab.b[find_bin(tab, blockIdx.x)];
  // one trip per row, every row inside the arrays: the pipelined loop
  const bool pipe = REG && (tab.indptr_base & 3) == 0 && (tab.row_len & 3) == 0 &&
                    tab.indptr_base + lb.count * tab.row_len <= nnz;  // uniform over the launch
#define FSP_VEC_ROWS(LN)                                                                                     \
  if constexpr (REG) {                                                                                     \
    if (pipe && tab.row_len <= 4 * (LN) * CHUNKS) {                                                        \
      spmv_vec_rows_pipe<LN, T, IT, CHUNKS>(lb, tab, indices, data, x, y);                                  \
    } else {                                                                                               \
      spmv_vec_rows<LN, T, IT, PT, QT, CHUNKS, PERM, true>(lb, tab, indptr, indices, data, perm, x, y,        \
                                                           partials, nnz);                                 \
    }                                                                                                      \
  } else {                                                                                                 \
    spmv_vec_rows<LN, T, IT, PT, QT, CHUNKS, PERM, false>(lb, tab, indptr, indices, data, perm, x, y,         \
                                                          partials, nnz);                                  \
  }
  switch (lb.lanes_log2) {  // uniform over the workgroup
    case 0: FSP_VEC_ROWS(1); break;
    case 1: FSP_VEC_ROWS(2); break;
    case 2: FSP_VEC_ROWS(4); break;
    case 3: FSP_VEC_ROWS(8); break;
    case 4: FSP_VEC_ROWS(16); break;
    case 5: FSP_VEC_ROWS(32); break;
    default: FSP_VEC_ROWS(64); break;
  }
#undef FSP_VEC_ROWS
}

// Row-range parts (fsp_spmv_csr_parts): workgroup p runs rows
// [row_begin, row_end) of part p, whose nonzeros use columns [col_lo, col_hi)
// only.  It copies x[col_lo, col_hi) into LDS once -- one barrier per part --
// when the window fits kLdsxBytes, else gathers from global memory, and runs
// the rows with the vector-chunk loop (L lanes per row).  A batch of
// matrices maps each matrix's rows (cut into parts of bounded nonzero count)
// to parts whose window is the matrix's column range.  Row bounds (BND):
// kBndGlobal loads them from indptr; kBndComputed (row_len > 0: every row
// holds row_len nonzeros from indptr_base + r * row_len) computes them;
// kBndLds reads the part's slice of indptr that the workgroup copied into
// LDS with x.  The last two leave no global load ahead of a pass's chunk
// loads, so a pass's first wait is for its chunks (see spmv_vec_rows).
constexpr int kBndGlobal = 0, kBndComputed = 1, kBndLds = 2;
template <int L, typename T, typename IT, typename PT, int BND>
__device__ __forceinline__ void spmv_part_rows(const PT* __restrict__ indptr, const IT* __restrict__ indices,
                                               const T* __restrict__ data, const T* __restrict__ xb,
                                               int64_t xoff, T* __restrict__ y, int64_t row_begin,
                                               int64_t row_end, int64_t nnz, int64_t row_len,
                                               int64_t indptr_base, const int64_t* __restrict__ bnd) {
  using A = typename spmv_acc_of<T>::type;
  constexpr int64_t per_wave = 64 / L;
  constexpr int64_t per_pass = per_wave * kLdsxWaves;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / L;
  const int l = lane % L;
  for (int64_t base = row_begin; base < row_end; base += per_pass) {  // uniform over the workgroup
    const int64_t r = base + wave * per_wave + sub;
    const bool active = r < row_end;
    int64_t b = 0, e = 0;
    if (BND == kBndComputed) {
      b = indptr_base + r * row_len;
      e = b + row_len;
    } else if (BND == kBndLds) {
      const int64_t ri = active ? r - row_begin : 0;  // inside the staged slice
      b = bnd[ri];
      e = bnd[ri + 1];
    } else if (active) {
      b = static_cast<int64_t>(indptr[r]);
      e = static_cast<int64_t>(indptr[r + 1]);
    }
    const bool whole = __all(!active || e + 3 <= nnz);
    A acc = A(0);
    if (active) {
      for (int64_t q = (b & ~int64_t(3)) + 4 * l; q < e; q += 4 * L) {
        chunk4<IT> c;
        chunk4<T> v;
        if (whole || q + 4 <= nnz) {
          c = load_chunk_nt(indices + q);
          v = load_chunk_nt(data + q);
        } else {
          c = load_chunk_tail(indices, q, nnz);
          v = load_chunk_tail(data, q, nnz);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool ok = q + j >= b && q + j < e;
          const int64_t col = ok ? static_cast<int64_t>(c.v[j]) : xoff;
          const A t = fmadd(static_cast<A>(v.v[j]), static_cast<A>(xb[col - xoff]), acc);
          acc = ok ? t : acc;
        }
      }
    }
    acc = group_sum(acc, L);
    if (active && l == 0) y[r] = static_cast<T>(acc);
  }
}

// Parts whose row pointers are staged in LDS, whose nonzeros all lie below
// nnz rounded down to a whole chunk and whose rows mostly fit one trip (the
// kernel's conditions below), software-pipelined over passes: the row bounds
// need no global load, so each pass first loads the next pass's first-trip
// chunks and a wave keeps two passes in flight.  What keeps the
// waits exact (the compiler's counters are in order and merge paths
// pessimistically):
//  * the loads are unconditional: a lane past its row's end, or a row past
//    the part's end (which takes the part's last row), loads a chunk clamped
//    into the arrays and masks it at the FMA;
//  * the y store is unconditional: every lane of a group holds the group
//    total, bitwise identical, and a clamped row stores the last row's total
//    to the last row, the value that row's own group stores;
//  * two register buffers alternate (a copy of a loaded register would wait
//    for its load), and an empty asm with a memory clobber keeps the next
//    pass's loads ahead of this pass's LDS reads.
// Further trips of a row longer than one trip load in place.  Per lane the
// chunks are summed in the order of spmv_part_rows (bitwise equal).
template <typename T, typename IT>
struct part_buf {
  int64_t b, e;
  chunk4<IT> c;
  chunk4<T> v;
};

template <int L, typename T, typename IT>
__device__ __forceinline__ void part_load(const IT* __restrict__ indices, const T* __restrict__ data,
                                          const int64_t* __restrict__ bnd, int64_t ri, int64_t last,
                                          int64_t max_q, part_buf<T, IT>& f) {
  const int64_t rc = ri < last ? ri : last;
  f.b = bnd[rc];
  f.e = bnd[rc + 1];
  const int64_t q = (f.b & ~int64_t(3)) + 4 * static_cast<int64_t>(threadIdx.x % L);
  const int64_t qc = q < max_q ? q : max_q;
  f.c = load_chunk_nt(indices + qc);
  f.v = load_chunk_nt(data + qc);
}

template <int L, typename T, typename IT>
__device__ __forceinline__ void part_pass(const IT* __restrict__ indices, const T* __restrict__ data,
                                          const T* __restrict__ xb, int64_t xoff, T* __restrict__ y,
                                          const int64_t* __restrict__ bnd, int64_t row_begin, int64_t last,
                                          int64_t max_q, int64_t ri, int64_t ri_next, const part_buf<T, IT>& f,
                                          part_buf<T, IT>& g) {
  using A = typename spmv_acc_of<T>::type;
  part_load<L>(indices, data, bnd, ri_next, last, max_q, g);
  asm volatile("" ::: "memory");
  const int l = threadIdx.x % L;
  const int64_t q0 = (f.b & ~int64_t(3)) + 4 * l;
  A acc = A(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool ok = q0 + j >= f.b && q0 + j < f.e;
    const int64_t col = ok ? static_cast<int64_t>(f.c.v[j]) : xoff;
    const A t = fmadd(static_cast<A>(f.v.v[j]), static_cast<A>(xb[col - xoff]), acc);
    acc = ok ? t : acc;
  }
  for (int64_t q = q0 + 4 * L; q < f.e; q += 4 * L) {  // rows longer than one trip (q + 4 <= e <= max_q + 4)
    const chunk4<IT> c2 = load_chunk_nt(indices + q);
    const chunk4<T> v2 = load_chunk_nt(data + q);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool ok = q + j < f.e;
      const int64_t col = ok ? static_cast<int64_t>(c2.v[j]) : xoff;
      const A t = fmadd(static_cast<A>(v2.v[j]), static_cast<A>(xb[col - xoff]), acc);
      acc = ok ? t : acc;
    }
  }
  acc = group_sum(acc, L);
  y[row_begin + (ri < last ? ri : last)] = static_cast<T>(acc);
}

template <int L, typename T, typename IT>
__device__ __forceinline__ void spmv_part_rows_pipe(const IT* __restrict__ indices, const T* __restrict__ data,
                                                    const T* __restrict__ xb, int64_t xoff, T* __restrict__ y,
                                                    int64_t row_begin, int64_t row_end, int64_t max_q,
                                                    const int64_t* __restrict__ bnd) {
  constexpr int64_t per_wave = 64 / L;
  constexpr int64_t per_pass = per_wave * kLdsxWaves;
  const int64_t last = row_end - 1 - row_begin;
  const int64_t n_pass = (last + per_pass) / per_pass;  // >= 1
  const int64_t r0 = (threadIdx.x >> 6) * per_wave + (threadIdx.x & 63) / L;  // row of this lane in a pass
  auto next = [n_pass](int64_t k) { return k + 1 < n_pass ? k + 1 : k; };
  part_buf<T, IT> f0, f1;
  part_load<L>(indices, data, bnd, r0, last, max_q, f0);
  // pass 0 peeled: both edges into the loop then end with a y store, so the
  // waits inside count it exactly
  part_pass<L>(indices, data, xb, xoff, y, bnd, row_begin, last, max_q, r0, next(0) * per_pass + r0, f0, f1);
  for (int64_t k = 1; k < n_pass;) {  // uniform over the workgroup
    part_pass<L>(indices, data, xb, xoff, y, bnd, row_begin, last, max_q, k * per_pass + r0,
                 next(k) * per_pass + r0, f1, f0);
    if (++k >= n_pass) break;
    part_pass<L>(indices, data, xb, xoff, y, bnd, row_begin, last, max_q, k * per_pass + r0,
                 next(k) * per_pass + r0, f0, f1);
    ++k;
  }
}

template <typename T, typename IT, typename PT>
__global__ __launch_bounds__(kLdsxThreads) void spmv_parts_kernel(const PT* __restrict__ indptr,
                                                                  const IT* __restrict__ indices,
                                                                  const T* __restrict__ data,
                                                                  const T* __restrict__ x, T* __restrict__ y,
                                                                  const int64_t* __restrict__ parts, int lanes_log2,
                                                                  int64_t nnz, int64_t row_len, int64_t indptr_base) {
  extern __shared__ __align__(16) unsigned char fsp_part_mem[];
  T* xs = reinterpret_cast<T*>(fsp_part_mem);
  const int64_t* pt = parts + 4 * static_cast<int64_t>(blockIdx.x);
  const int64_t row_begin = pt[0], row_end = pt[1], lo = pt[2], span = pt[3] - pt[2];
  const bool fits = span >= 0 && span * static_cast<int64_t>(sizeof(T)) <= kLdsxBytes;  // uniform
  // the part's indptr slice goes into LDS after x when both fit (uniform)
  const int64_t boff = (span * static_cast<int64_t>(sizeof(T)) + 15) & ~int64_t(15);
  const int64_t n_bnd = row_end - row_begin + 1;
  const bool stage_b = fits && row_len == 0 && n_bnd >= 1 && boff + n_bnd * 8 <= kLdsxBytes;
  int64_t* bnd = reinterpret_cast<int64_t*>(fsp_part_mem + (stage_b ? boff : 0));
  if (fits) copy_to_lds(x + lo, xs, span);
  if (stage_b) {  // n_bnd * 8 <= kLdsxBytes: at most kPer entries per thread (loads as in copy_to_lds)
    constexpr int kPer = static_cast<int>(kLdsxBytes / (8 * kLdsxThreads));
    const PT* src = indptr + row_begin;
    const int nb = static_cast<int>(n_bnd);  // <= 8192: 32-bit indices (see copy_to_lds)
    PT v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = static_cast<int>(threadIdx.x) + k * kLdsxThreads;
      v[k] = src[i < nb ? i : nb - 1];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) pin(v[k]);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = static_cast<int>(threadIdx.x) + k * kLdsxThreads;
      if (i < nb) bnd[i] = static_cast<int64_t>(v[k]);
    }
  }
  __syncthreads();
  // The pipelined loop (uniform conditions): row pointers in LDS, every
  // nonzero of the part inside the whole chunks of the arrays (the clamped
  // loads stay in bounds), and a mean row length + 3 of at most 4 L, one
  // trip (a row longer than one trip loads its further trips in place, and
  // then waits for the next pass's loads as well: at 2-3 trips a row the
  // plain loop measured faster).  8-byte values with 8-byte column ids keep
  // the plain loop: the two buffers would take the kernel past 64 vector
  // registers.  Constant-length rows keep their computed-bounds loop: a
  // pipelined form of it gained nothing and its paths raised this kernel's
  // scalar registers past what two 16-wave workgroups per CU allow (84
  // against 72), which slowed every part.
  const int64_t max_q = (nnz & ~int64_t(3)) - 4;
  const int64_t n_rows_p = row_end - row_begin;
  constexpr bool kPipeTypes = sizeof(T) + sizeof(IT) <= 12;
  const bool pipe_ok = kPipeTypes && stage_b && n_rows_p > 0 && max_q >= 0 && bnd[0] >= 0 &&
                       bnd[n_bnd - 1] <= max_q + 4;
  const int64_t trip_nnz = pipe_ok ? (bnd[n_bnd - 1] - bnd[0]) + 3 * n_rows_p : 0;
#define FSP_PART_ROWS(LN)                                                                                   \
  if (fits && row_len > 0) {                                                                              \
    spmv_part_rows<LN, T, IT, PT, kBndComputed>(indptr, indices, data, xs, lo, y, row_begin, row_end, nnz,  \
                                                row_len, indptr_base, nullptr);                           \
  } else if (pipe_ok && trip_nnz <= 4 * (LN) * n_rows_p) {                                                \
    spmv_part_rows_pipe<LN, T, IT>(indices, data, xs, lo, y, row_begin, row_end, max_q, bnd);              \
  } else if (stage_b) {                                                                                   \
    spmv_part_rows<LN, T, IT, PT, kBndLds>(indptr, indices, data, xs, lo, y, row_begin, row_end, nnz, 0, 0, \
                                           bnd);                                                          \
  } else if (fits) {                                                                                      \
    spmv_part_rows<LN, T, IT, PT, kBndGlobal>(indptr, indices, data, xs, lo, y, row_begin, row_end, nnz, 0,  \
                                              0, nullptr);                                                \
  } else {                                                                                                \
    spmv_part_rows<LN, T, IT, PT, kBndGlobal>(indptr, indices, data, x, 0, y, row_begin, row_end, nnz, 0, 0, \
                                              nullptr);                                                   \
  }
  switch (lanes_log2) {  // uniform
    case 2: FSP_PART_ROWS(4); break;
    case 3: FSP_PART_ROWS(8); break;
    default: FSP_PART_ROWS(16); break;
  }
#undef FSP_PART_ROWS
}

// Compute units of the current device (cached per device).
static int device_cu_count() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
  if (dev < 64) {
    const int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (dev < 64) cache[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmv_zero_kernel(T* __restrict__ y, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    y[i] = T(0);
  }
}

// y[long_rows[i]] = sum of the row's segment partials, in segment order.
template <typename T>
__global__ __launch_bounds__(FSP_WG_THREADS) void spmv_long_reduce_kernel(
    const typename spmv_acc_of<T>::type* __restrict__ partials, const int32_t* __restrict__ long_rows,
    const int64_t* __restrict__ seg_ptr, int64_t n_long, T* __restrict__ y) {
  using A = typename spmv_acc_of<T>::type;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * FSP_WG_THREADS + threadIdx.x; i < n_long;
       i += static_cast<int64_t>(gridDim.x) * FSP_WG_THREADS) {
    A acc = A(0);
    for (int64_t s = seg_ptr[i]; s < seg_ptr[i + 1]; ++s) acc += partials[s];
    y[long_rows[i]] = static_cast<T>(acc);
  }
}

template <typename T, typename IT, typename PT, typename QT, bool PERM>
static int launch_spmv(const void* indptr, const void* indices, const void* data, const void* perm,
                       const void* x, void* y, int64_t x_len, int64_t nnz, const fsp_csr_plan* plan,
                       hipStream_t s) {
  using A = typename spmv_acc_of<T>::type;
  launch_table tab;
  const bool vec = plan->variant == 4;
  const bool ldsx = plan->variant >= 2;
  const int64_t x_bytes = x_len * static_cast<int64_t>(sizeof(T));
  if (ldsx && (x_len < 1 || x_bytes > kLdsxBytes)) {
    return fail(FSP_ERR_ARG, "fsp_spmv_csr: variants 2-4 stage x in LDS and need 1 <= x_len * %d <= %lld bytes",
                (int)sizeof(T), (long long)kLdsxBytes);
  }
  if (vec) {
    const bool aligned = reinterpret_cast<uintptr_t>(indices) % (4 * sizeof(IT)) == 0 &&
                         (PERM ? reinterpret_cast<uintptr_t>(perm) % (4 * sizeof(QT)) == 0
                               : reinterpret_cast<uintptr_t>(data) % (4 * sizeof(T)) == 0);
    if (!aligned) {
      return fail(FSP_ERR_ARG, "fsp_spmv_csr: variant 4 reads 4-element chunks and needs indices and %s aligned to "
                  "4 elements", PERM ? "perm" : "data");
    }
    if (plan->unroll != 4 && plan->unroll != 8) {
      return fail(FSP_ERR_ARG, "fsp_spmv_csr: variant 4 needs unroll 4 or 8 (one or two chunks per lane)");
    }
  }
  fsp_csr_plan lplan = *plan;
  if (ldsx && lplan.max_blocks == 0) {
    const int cus = device_cu_count();
    lplan.max_blocks = 2 * static_cast<int64_t>(cus > 0 ? cus : 256);
  }
  const int64_t waves = ldsx ? kLdsxWaves : FSP_WAVES;
  const int64_t blocks = build_launch_table(
      &lplan, /*seg_lanes_log2=*/6, /*max_lanes_log2=*/6,
      [waves](int ll, int) { return waves * (int64_t(64) >> ll); }, &tab);
  if (plan->n_segments > 0 &&
      (plan->partials_dev == nullptr || plan->partials_bytes < plan->n_segments * (int64_t)sizeof(A))) {
    return fail(FSP_ERR_WORKSPACE, "fsp_spmv_csr: partials need %lld bytes",
                (long long)(plan->n_segments * (int64_t)sizeof(A)));
  }
  A* partials = static_cast<A*>(plan->partials_dev);
  if (blocks > 0) {
    const int unroll = plan->unroll;
#define FSP_SPMV_KERNEL(NAME, U, W, LX)                                                              \
  hipLaunchKernelGGL((NAME<T, IT, PT, QT, U, PERM, W, LX>), dim3(blocks), dim3(64 * (W)),                \
                     LX ? static_cast<size_t>(x_bytes) : 0, s, static_cast<const PT*>(indptr),          \
                     static_cast<const IT*>(indices), static_cast<const T*>(data),                        \
                     static_cast<const QT*>(perm), static_cast<const T*>(x), static_cast<T*>(y), partials, \
                     tab, x_len)
#define FSP_SPMV_LAUNCH(U)                                                 \
  switch (plan->variant) {                                                 \
    case 1: FSP_SPMV_KERNEL(spmv_simple_kernel, U, FSP_WAVES, false); break; \
    case 2: FSP_SPMV_KERNEL(spmv_simple_kernel, U, kLdsxWaves, true); break; \
    case 3: FSP_SPMV_KERNEL(spmv_bins_kernel, U, kLdsxWaves, true); break;   \
    default: FSP_SPMV_KERNEL(spmv_bins_kernel, U, FSP_WAVES, false); break;  \
  }
#define FSP_SPMV_VEC(CH, REG)                                                                          \
  hipLaunchKernelGGL((spmv_vec_kernel<T, IT, PT, QT, CH, PERM, REG>), dim3(blocks), dim3(kLdsxThreads),   \
                     static_cast<size_t>(x_bytes), s, static_cast<const PT*>(indptr),                    \
                     static_cast<const IT*>(indices), static_cast<const T*>(data),                        \
                     static_cast<const QT*>(perm), static_cast<const T*>(x), static_cast<T*>(y), partials, \
                     tab, x_len, nnz)
    if (vec) {
      // constant-length plan, identity bins only, no segments: row bounds computed
      bool reg = !PERM && plan->row_len > 0 && plan->n_segments == 0;
      for (int i = 0; i < plan->n_bins; ++i) reg = reg && (plan->bins[i].count == 0 || plan->bins[i].rows == nullptr);
      if constexpr (!PERM) {
        if (reg) {
          if (unroll == 4) {
            FSP_SPMV_VEC(1, true);
          } else {
            FSP_SPMV_VEC(2, true);
          }
        }
      }
      if (!reg) {
        if (unroll == 4) {
          FSP_SPMV_VEC(1, false);
        } else {
          FSP_SPMV_VEC(2, false);
        }
      }
    } else if (unroll <= 2) {
      FSP_SPMV_LAUNCH(2);
    } else if (unroll <= 4) {
      FSP_SPMV_LAUNCH(4);
    } else {
      FSP_SPMV_LAUNCH(8);
    }
#undef FSP_SPMV_VEC
#undef FSP_SPMV_LAUNCH
#undef FSP_SPMV_KERNEL
    const int st = check_launch("spmv_bins");
    if (st) return st;
  }
  if (plan->n_long_rows > 0) {
    int64_t g = ceil_div(plan->n_long_rows, FSP_WG_THREADS);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(spmv_long_reduce_kernel<T>, dim3(g), dim3(FSP_WG_THREADS), 0, s, partials,
                       plan->long_rows_dev, plan->long_seg_ptr_dev, plan->n_long_rows, static_cast<T*>(y));
    return check_launch("spmv_long_reduce");
  }
  return FSP_OK;
}

}  // namespace fsp

using namespace fsp;

extern "C" int fsp_spmv_csr(int32_t value_type, int32_t index_type, int32_t indptr_type,
                            const void* indptr_dev, const void* indices_dev, const void* data_dev,
                            int32_t perm_type, const void* perm_dev, const void* x_dev, void* y_dev,
                            int64_t n_rows, int64_t nnz, int64_t x_len, const fsp_csr_plan* plan, void* stream) {
  if (n_rows < 0 || nnz < 0 || x_len < 0) return fail(FSP_ERR_ARG, "fsp_spmv_csr: negative n_rows, nnz or x_len");
  int st = check_plan(plan, n_rows);
  if (st) return st;
  if (n_rows == 0) return FSP_OK;
  if (indptr_dev == nullptr || y_dev == nullptr) return fail(FSP_ERR_ARG, "fsp_spmv_csr: NULL indptr or y");
  if (nnz > 0 && (indices_dev == nullptr || data_dev == nullptr || x_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "fsp_spmv_csr: NULL indices, data or x");
  }
  const hipStream_t s = as_stream(stream);
  if (nnz == 0) {  // y = 0: the row kernels read the first nonzero unconditionally
    FSP_DISPATCH_VALUE(value_type, T, {
      int64_t g = ceil_div(n_rows, FSP_WG_THREADS);
      if (g > 65536) g = 65536;
      hipLaunchKernelGGL(spmv_zero_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(FSP_WG_THREADS), 0, s,
                         static_cast<T*>(y_dev), n_rows);
      return check_launch("spmv_zero");
    });
  }
  const bool use_perm = perm_dev != nullptr;
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        if (!use_perm) {
          return launch_spmv<T, IT, PT, int32_t, false>(indptr_dev, indices_dev, data_dev, nullptr, x_dev,
                                                        y_dev, x_len, nnz, plan, s);
        }
        FSP_DISPATCH_INDEX(perm_type, QT, {
          return launch_spmv<T, IT, PT, QT, true>(indptr_dev, indices_dev, data_dev, perm_dev, x_dev, y_dev,
                                                  x_len, nnz, plan, s);
        });
      });
    });
  });
  return FSP_OK;
}

extern "C" int fsp_spmv_csr_parts(int32_t value_type, int32_t index_type, int32_t indptr_type,
                                  const void* indptr_dev, const void* indices_dev, const void* data_dev,
                                  const void* x_dev, void* y_dev, int64_t nnz, int64_t n_parts,
                                  const int64_t* parts_dev, int32_t lanes, int64_t row_len,
                                  int64_t indptr_base, void* stream) {
  if (nnz < 0 || n_parts < 0 || (lanes != 4 && lanes != 8 && lanes != 16)) {
    return fail(FSP_ERR_ARG, "fsp_spmv_csr_parts: bad arguments (lanes must be 4, 8 or 16)");
  }
  if (row_len < 0 || indptr_base < 0) {
    return fail(FSP_ERR_ARG, "fsp_spmv_csr_parts: row_len %lld / indptr_base %lld negative", (long long)row_len,
                (long long)indptr_base);
  }
  if (n_parts == 0) return FSP_OK;
  if (n_parts > 0x7fffffff || parts_dev == nullptr || indptr_dev == nullptr || y_dev == nullptr ||
      (nnz > 0 && (indices_dev == nullptr || data_dev == nullptr || x_dev == nullptr))) {
    return fail(FSP_ERR_ARG, "fsp_spmv_csr_parts: NULL array or too many parts");
  }
  const hipStream_t s = as_stream(stream);
  const int ll = lanes == 4 ? 2 : (lanes == 8 ? 3 : 4);
  FSP_DISPATCH_VALUE(value_type, T, {
    FSP_DISPATCH_INDEX(index_type, IT, {
      FSP_DISPATCH_INDEX(indptr_type, PT, {
        if (reinterpret_cast<uintptr_t>(indices_dev) % (4 * sizeof(IT)) != 0 ||
            reinterpret_cast<uintptr_t>(data_dev) % (4 * sizeof(T)) != 0) {
          return fail(FSP_ERR_ARG, "fsp_spmv_csr_parts: indices and data must be aligned to 4 elements");
        }
        hipLaunchKernelGGL((spmv_parts_kernel<T, IT, PT>), dim3(static_cast<unsigned>(n_parts)), dim3(kLdsxThreads),
                           static_cast<size_t>(kLdsxBytes), s, static_cast<const PT*>(indptr_dev),
                           static_cast<const IT*>(indices_dev), static_cast<const T*>(data_dev),
                           static_cast<const T*>(x_dev), static_cast<T*>(y_dev), parts_dev, ll, nnz, row_len,
                           indptr_base);
        return check_launch("spmv_parts");
      });
    });
  });
  return FSP_OK;
}
