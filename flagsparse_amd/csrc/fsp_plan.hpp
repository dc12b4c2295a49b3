// Launch table of the fused row-bin kernels.
//
// One launch covers every row bin of a plan plus the long-row segments: the
// table (passed by value as a kernel argument) gives each bin its range of
// workgroups; a workgroup finds its bin with a uniform scan of <= 9 entries.
// Inside a bin, workgroup j processes the contiguous row tiles
// [j * chunk, (j + 1) * chunk).  chunk is 1 unless the plan caps the grid
// (max_blocks, a persistent launch sized to the chip): then each workgroup
// walks a contiguous run of rows.
//
// With plan.xcd_groups = X > 1 the tiles of a bin are instead cut into X
// contiguous regions and the workgroups b with b % X == x stride through
// region x together (tile = start_x + b / X, + the group's size, ...).
// Workgroups are dispatched round-robin over the XCDs (8 on MI355X), so with
// X = 8 the workgroups resident on one XCD at any moment work on neighbouring
// rows: the dense rows those rows gather (one matrix of a packed batch, say)
// stay in that XCD's L2 instead of being pulled into all eight.
#pragma once

#include "fsp_device.hpp"
#include "fsp_host.hpp"

namespace fsp {

constexpr int kMaxLaunchBins = FSP_MAX_BINS + 1;  // bins + the segment pseudo-bin
constexpr int64_t kMaxBlocksPerBin = 1 << 20;

struct launch_bin {
  const int32_t* rows;   // NULL = identity (bins) ; unused for segments
  int64_t count;         // rows (or segments) in the bin
  int64_t block_begin;   // first workgroup of this bin
  int64_t nblocks;       // workgroups of this bin
  int64_t chunk;         // consecutive tiles per workgroup (xcds <= 1)
  int64_t xcds;          // > 1: XCD-region mapping over this many groups
  int32_t lanes_log2;    // SpMV: lanes per row; SpMM/SDDMM: nonzero slots per row
  int32_t is_segment;    // 1 = items are long-row segments
};

struct launch_table {
  launch_bin b[kMaxLaunchBins];
  int32_t n;
  int32_t pad;
  const int32_t* seg_row;
  const int64_t* seg_begin;
  const int64_t* seg_end;
  int64_t row_len;      // > 0: rows of exactly this length (the SpMV vector-chunk loop computes bounds)
  int64_t indptr_base;  // indptr[0] when row_len > 0
};

// Tiles of one workgroup: first, first + step, ... while < end.
struct tile_range {
  int64_t first, step, end;
};

__device__ __forceinline__ tile_range bin_tiles(const launch_bin& lb, int64_t per_block) {
  const int64_t n_tiles = (lb.count + per_block - 1) / per_block;
  const int64_t b = static_cast<int64_t>(blockIdx.x) - lb.block_begin;
  tile_range r;
  if (lb.xcds <= 1) {
    r.first = b * lb.chunk;
    r.step = 1;
    r.end = r.first + lb.chunk < n_tiles ? r.first + lb.chunk : n_tiles;
  } else {
    const int64_t X = lb.xcds;
    const int64_t x = b % X;
    r.first = n_tiles * x / X + b / X;
    r.step = (lb.nblocks - x + X - 1) / X;  // workgroups of group x (>= 1: x < X <= nblocks)
    r.end = n_tiles * (x + 1) / X;
  }
  return r;
}

// Work item = (row, first nonzero, one past the last nonzero, output slot).
struct work_item {
  int64_t row;
  int64_t begin;
  int64_t end;
  int64_t out;  // row for bins, segment index for segments
};

__device__ __forceinline__ int find_bin(const launch_table& t, int64_t block) {
  int k = 0;
  for (int i = 1; i < t.n; ++i) {
    if (block >= t.b[i].block_begin) k = i;
  }
  return k;
}

template <typename PT>
__device__ __forceinline__ work_item load_item(const launch_table& t, const launch_bin& lb,
                                               const PT* __restrict__ indptr, int64_t rr) {
  work_item w;
  if (lb.is_segment) {
    w.row = t.seg_row[rr];
    w.begin = t.seg_begin[rr];
    w.end = t.seg_end[rr];
    w.out = rr;
  } else {
    w.row = lb.rows ? static_cast<int64_t>(lb.rows[rr]) : rr;
    w.begin = static_cast<int64_t>(indptr[w.row]);
    w.end = static_cast<int64_t>(indptr[w.row + 1]);
    w.out = w.row;
  }
  return w;
}

static inline int ilog2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << (l + 1)) <= v) ++l;
  return l;
}

// Builds the launch table.  `items_per_block(lanes_log2, is_segment)` gives
// the rows (segments) one workgroup covers per tile.  `seg_lanes_log2` is the
// lanes_log2 used for segments.  plan->max_blocks > 0 caps the total grid:
// each bin then gets a share of the workgroups proportional to its tiles
// (at least one) and every workgroup a contiguous chunk of tiles.  Returns
// the total number of workgroups.
template <typename F>
static int64_t build_launch_table(const fsp_csr_plan* plan, int seg_lanes_log2, int max_lanes_log2,
                                  F&& items_per_block, launch_table* t) {
  int64_t tiles[kMaxLaunchBins];
  int64_t total_tiles = 0;
  t->n = 0;
  // Longest work first: the long-row segments, then the bins from the
  // longest rows down.  Workgroups dispatch in block order, so the items
  // that take longest start first and the short rows fill in behind them.
  if (plan->n_segments > 0) {
    launch_bin& lb = t->b[t->n];
    lb.rows = nullptr;
    lb.count = plan->n_segments;
    lb.lanes_log2 = seg_lanes_log2;
    lb.is_segment = 1;
    tiles[t->n] = ceil_div(plan->n_segments, items_per_block(seg_lanes_log2, 1));
    total_tiles += tiles[t->n];
    ++t->n;
  }
  for (int i = (plan->n_bins < FSP_MAX_BINS ? plan->n_bins : FSP_MAX_BINS) - 1; i >= 0; --i) {
    const fsp_bin& b = plan->bins[i];
    if (b.count <= 0) continue;
    launch_bin& lb = t->b[t->n];
    int ll = ilog2(b.lanes > 0 ? b.lanes : 1);
    if (ll > max_lanes_log2) ll = max_lanes_log2;
    lb.rows = b.rows;
    lb.count = b.count;
    lb.lanes_log2 = ll;
    lb.is_segment = 0;
    tiles[t->n] = ceil_div(b.count, items_per_block(ll, 0));
    total_tiles += tiles[t->n];
    ++t->n;
  }
  const int64_t cap = plan->max_blocks;
  int64_t blocks = 0;
  for (int i = 0; i < t->n; ++i) {
    launch_bin& lb = t->b[i];
    int64_t nb = tiles[i];
    if (cap > 0 && total_tiles > cap) {
      nb = tiles[i] * cap / total_tiles;  // tiles <= 2^40, cap <= 2^20: no overflow
      if (nb < 1) nb = 1;
    }
    if (nb > kMaxBlocksPerBin) nb = kMaxBlocksPerBin;
    const int64_t groups = plan->xcd_groups < nb ? plan->xcd_groups : nb;
    if (groups > 1) {  // region mapping needs two or more workgroups
      lb.chunk = 1;
      lb.nblocks = nb;
      lb.xcds = groups;
    } else {
      lb.chunk = ceil_div(tiles[i], nb);
      lb.nblocks = ceil_div(tiles[i], lb.chunk);  // no workgroup without tiles
      lb.xcds = 1;
    }
    lb.block_begin = blocks;
    blocks += lb.nblocks;
  }
  t->seg_row = plan->seg_row_dev;
  t->seg_begin = plan->seg_begin_dev;
  t->seg_end = plan->seg_end_dev;
  t->row_len = plan->row_len;
  t->indptr_base = plan->indptr_base;
  return blocks;
}

static inline int check_plan(const fsp_csr_plan* plan, int64_t n_rows) {
  if (plan == nullptr) return fail(FSP_ERR_ARG, "plan is NULL");
  if (plan->n_bins < 0 || plan->n_bins > FSP_MAX_BINS) return fail(FSP_ERR_ARG, "plan: bad n_bins %d", plan->n_bins);
  int64_t total = 0;
  for (int i = 0; i < plan->n_bins; ++i) {
    const fsp_bin& b = plan->bins[i];
    if (b.count < 0 || b.lanes < 1 || b.lanes > 64 || !is_pow2(b.lanes)) {
      return fail(FSP_ERR_ARG, "plan: bin %d has count %lld lanes %d", i, (long long)b.count, b.lanes);
    }
    if (b.rows == nullptr && b.count > n_rows) return fail(FSP_ERR_ARG, "plan: identity bin larger than n_rows");
    total += b.count;
  }
  if (total + plan->n_long_rows != n_rows) {
    return fail(FSP_ERR_ARG, "plan: bins cover %lld rows + %lld long rows, matrix has %lld",
                (long long)total, (long long)plan->n_long_rows, (long long)n_rows);
  }
  if (plan->n_segments < 0 || plan->n_long_rows < 0) return fail(FSP_ERR_ARG, "plan: negative segment counts");
  if (plan->row_len < 0 || plan->indptr_base < 0) {
    return fail(FSP_ERR_ARG, "plan: row_len %lld / indptr_base %lld negative", (long long)plan->row_len,
                (long long)plan->indptr_base);
  }
  if (plan->variant < 0 || plan->variant > 4) return fail(FSP_ERR_ARG, "plan: unknown variant %lld", (long long)plan->variant);
  if (plan->xcd_groups < 0 || plan->xcd_groups > 64) {
    return fail(FSP_ERR_ARG, "plan: xcd_groups %lld outside [0, 64]", (long long)plan->xcd_groups);
  }
  if (plan->max_blocks < 0 || plan->max_blocks > kMaxBlocksPerBin) {
    return fail(FSP_ERR_ARG, "plan: max_blocks %lld outside [0, %lld]", (long long)plan->max_blocks,
                (long long)kMaxBlocksPerBin);
  }
  if (plan->n_segments > 0 && (plan->seg_row_dev == nullptr || plan->seg_begin_dev == nullptr ||
                               plan->seg_end_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "plan: segment arrays missing");
  }
  if (plan->n_long_rows > 0 && (plan->long_rows_dev == nullptr || plan->long_seg_ptr_dev == nullptr)) {
    return fail(FSP_ERR_ARG, "plan: long-row arrays missing");
  }
  return FSP_OK;
}

}  // namespace fsp
