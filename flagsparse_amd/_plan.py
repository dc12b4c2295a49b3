"""Row-length work plans for the CSR kernels (prepare-time, cached).

A plan bins the rows of a CSR (or CSC) index structure by length on the GPU
(``fsp_csr_row_bins``: a stable multi-split, rows stay ascending inside a
bin) and cuts rows longer than the last bin into nnz-balanced segments.  The
kernels then run every bin and every segment in one launch; rows that were
split get a second, deterministic pass that adds their segment partials in
order.  Building a plan costs one device->host copy of the bin boundaries and
of the long rows' extents; running with a plan costs no host sync.

Bin layouts per kernel family (``unroll`` = nonzeros per lane per trip):

* SpMV: lanes per row L in {4, 8, 16, 32, 64}; a row of length n goes to the
  smallest L with n <= L * unroll, L = 64 up to ``long_threshold``.
* SpMM / SDDMM with G lanes along the dense dimension: nonzero slots per row
  S in {1, 2, ..., 64 / G}; smallest S with n <= S * unroll * slot_scale,
  S = 64 / G up to ``long_threshold``.  SDDMM uses half the lanes of a full
  K slice (``sddmm_lanes``) and walks K in two passes.

The reference rebuilds its buckets with ``torch.nonzero`` on every call
(spmv_csr.py:371-420, 910-911) or with atomic count / compact kernels
(spmm_csr.py:3022-3072, 3114-3115).
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

# Defaults, overridable per call (the autotuner picks per shape class).
SPMV_UNROLL = 4
SPMV_LONG_THRESHOLD = 4096
SPMV_SEG_LEN = 4096
SPMM_UNROLL = 4
SPMM_LONG_THRESHOLD = 256
SPMM_SEG_LEN = 256
SDDMM_UNROLL = 4
SDDMM_LONG_THRESHOLD = 256
SDDMM_SEG_LEN = 256
# Rows per wave over slots per row: measured best on MI355X for power-law rows
# (mean 14-24 nonzeros) in f32 and bf16 (SpMM); SDDMM is best with one slot
# per row up to the long threshold.
SLOT_SCALE = 4
SDDMM_SLOT_SCALE = 1


@dataclass(frozen=True)
class PlanConfig:
    """Row-bin thresholds of one plan (hashable: it keys the plan cache)."""

    lanes: tuple  # lanes (SpMV) or slots (SpMM/SDDMM) of each regular bin
    uppers: tuple  # inclusive upper row length of each regular bin
    seg_len: int  # segment length of rows above the last upper bound
    unroll: int


# SpMV variant 2 stages x in LDS: x of at most this many bytes, and only for
# matrices large enough that every workgroup's one copy of x (a persistent
# grid of two 1024-thread workgroups per CU) is small beside the nonzeros
LDSX_MAX_BYTES = 65536
LDSX_MIN_NNZ = 1 << 20


def chunk_aligned(*tensors):
    """True when every tensor's storage starts on a 4-element boundary (the
    vector-chunk SpMV loop reads 4-element chunks)."""
    return all(t.data_ptr() % (4 * t.element_size()) == 0 for t in tensors)


def spmv_default_variant(x_len, elem_bytes, nnz, chunk_ok=False):
    """SpMV loop used without tuning: when x fits in LDS and the matrix is
    large, 4 (vector chunks, needs 4-element aligned index / value arrays and
    unroll 4 or 8) or else 2 (x in LDS); otherwise 0 (software-pipelined)."""
    if 0 < x_len * elem_bytes <= LDSX_MAX_BYTES and nnz >= LDSX_MIN_NNZ:
        return 4 if chunk_ok else 2
    return 0


def spmv_config(unroll=SPMV_UNROLL, long_threshold=SPMV_LONG_THRESHOLD, seg_len=SPMV_SEG_LEN):
    lanes = (4, 8, 16, 32, 64)
    uppers = [min(l * unroll, long_threshold) for l in lanes[:-1]] + [long_threshold]
    return _dedupe(lanes, uppers, seg_len, unroll)


def rows_config(lanes_dense, unroll, long_threshold, seg_len, slot_scale=1):
    """SpMM / SDDMM bins for `lanes_dense` lanes along the dense dimension.

    A row of length n gets the fewest slots S with n <= S * unroll *
    slot_scale: slot_scale > 1 packs more rows into a wave (more independent
    rows in flight, more trips per row)."""
    max_slots = max(1, 64 // int(lanes_dense))
    lanes = []
    s = 1
    while s <= max_slots:
        lanes.append(s)
        s *= 2
    uppers = [min(l * unroll * slot_scale, long_threshold) for l in lanes[:-1]] + [long_threshold]
    return _dedupe(tuple(lanes), uppers, seg_len, unroll)


def spmm_config(lanes_dense, unroll=SPMM_UNROLL, long_threshold=SPMM_LONG_THRESHOLD, seg_len=SPMM_SEG_LEN,
                slot_scale=SLOT_SCALE):
    return rows_config(lanes_dense, unroll, long_threshold, seg_len, slot_scale)


def sddmm_config(lanes_dense, unroll=SDDMM_UNROLL, long_threshold=SDDMM_LONG_THRESHOLD, seg_len=SDDMM_SEG_LEN,
                 slot_scale=SDDMM_SLOT_SCALE):
    return rows_config(lanes_dense, unroll, long_threshold, seg_len, slot_scale)


def _dedupe(lanes, uppers, seg_len, unroll):
    # Drop bins made empty by a small long_threshold (upper not increasing).
    keep_l, keep_u = [], []
    prev = -1
    for l, u in zip(lanes, uppers):
        if u > prev:
            keep_l.append(l)
            keep_u.append(u)
            prev = u
    return PlanConfig(tuple(keep_l), tuple(keep_u), int(max(1, seg_len)), int(unroll))


def sddmm_lanes(k, elem_bytes):
    """SDDMM lanes along K: half of a full 16-byte-per-lane slice once that
    slice spans 8 lanes or more (two passes over K per nonzero keep twice the
    nonzero slots in a wave; measured 1.6-1.7x faster on power-law graphs
    with K = 64 / 128 in f32 and bf16 on MI355X)."""
    g = lanes_for_dense(k, elem_bytes)
    return g // 2 if g >= 8 else g


def lanes_for_dense(n_dense, elem_bytes):
    """Lanes along a dense dimension of n_dense elements read 16 bytes at a time."""
    vec = max(1, 16 // int(elem_bytes))
    vcols = -(-int(n_dense) // vec)
    g = 1
    while g < 64 and g < vcols:
        g *= 2
    return g


def _index_code(t):
    return _lib.index_code(t.dtype)


@dataclass
class CsrPlan:
    """Device work plan of one index structure (see module docstring)."""

    config: PlanConfig
    n_rows: int
    rows: torch.Tensor  # int32 [n_rows], regular bins then the long bin
    bin_counts: list  # host counts of the regular bins
    identity: bool  # the single regular bin holds rows 0..n_rows-1
    seg_row: torch.Tensor
    seg_begin: torch.Tensor
    seg_end: torch.Tensor
    long_rows: torch.Tensor
    long_seg_ptr: torch.Tensor
    n_segments: int
    n_long_rows: int
    row_len: int = 0  # > 0: every row has exactly this many nonzeros
    indptr_base: int = 0  # indptr[0] when row_len > 0

    @property
    def bin_sizes(self):
        return list(self.bin_counts) + [self.n_long_rows]

    def partials(self, n_dense, acc_dtype, device):
        """Scratch of n_segments x n_dense partial sums for ONE launch.

        A fresh buffer from torch's stream-ordered caching allocator on the
        current stream (the stream the launch goes to), so calls that share a
        plan on different streams never share partials; under graph capture it
        comes from the graph's private pool."""
        if self.n_segments == 0:
            return None
        return torch.empty(self.n_segments * max(1, int(n_dense)), dtype=acc_dtype, device=device)

    def struct(self, partials=None, max_blocks=0, xcd_groups=0, variant=0):
        """The C plan.  max_blocks > 0 makes the launch persistent: at most
        that many workgroups, each walking a contiguous run of row tiles;
        xcd_groups > 1 cuts each bin's tiles into that many contiguous regions
        served by the workgroups b with b % xcd_groups == region (8 = one
        region per XCD).  variant selects the SpMV loop (0 pipelined, 1 plain,
        2 plain and 3 pipelined with x staged in LDS, 4 vector chunks with x
        in LDS)."""
        st = _lib.FspCsrPlan()
        st.n_bins = len(self.bin_counts)
        st.unroll = self.config.unroll
        st.max_blocks = int(max_blocks)
        st.xcd_groups = int(xcd_groups)
        st.variant = int(variant)
        off = 0
        for i, (lanes, count) in enumerate(zip(self.config.lanes, self.bin_counts)):
            st.bins[i].lanes = int(lanes)
            st.bins[i].count = int(count)
            if self.identity or count == 0:
                st.bins[i].rows = None
            else:
                st.bins[i].rows = self.rows.data_ptr() + 4 * off
            off += count
        st.n_segments = self.n_segments
        st.n_long_rows = self.n_long_rows
        st.row_len = self.row_len
        st.indptr_base = self.indptr_base
        if self.n_segments:
            st.seg_row_dev = self.seg_row.data_ptr()
            st.seg_begin_dev = self.seg_begin.data_ptr()
            st.seg_end_dev = self.seg_end.data_ptr()
            st.long_rows_dev = self.long_rows.data_ptr()
            st.long_seg_ptr_dev = self.long_seg_ptr.data_ptr()
        if partials is not None:
            st.partials_dev = partials.data_ptr()
            st.partials_bytes = partials.numel() * partials.element_size()
        return st


def _launch_row_bins(indptr, n_rows, config, offsets):
    """Launch the row binning of `indptr` (device tensor, int32/int64) into
    `offsets` (device int64[n_bins + 1]); returns the row-id buffer."""
    device = indptr.device
    nb = len(config.uppers) + 1  # regular bins + the long bin
    lib = _lib.load()
    rows = torch.empty(max(1, n_rows), dtype=torch.int32, device=device)
    ws_bytes = ctypes.c_int64(0)
    _lib.check(lib.fsp_csr_row_bins_workspace(n_rows, nb, ctypes.byref(ws_bytes)), "fsp_csr_row_bins_workspace")
    ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=device)
    uppers = (ctypes.c_int64 * nb)(*[int(u) for u in config.uppers], 0)
    _lib.check(
        lib.fsp_csr_row_bins(
            _index_code(indptr), _lib.ptr(indptr), n_rows, nb, uppers, rows.data_ptr(),
            offsets.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(device),
        ),
        "fsp_csr_row_bins",
    )
    return rows


def _row_len_if_constant(n_rows, first, last, max_len):
    """The common row length when every row has max_len nonzeros (n_rows
    rows of at most max_len summing to last - first), else 0."""
    if n_rows > 0 and max_len > 0 and max_len * n_rows == last - first:
        return int(max_len)
    return 0


def build_plan(indptr, n_rows, config):
    """Bin the rows described by `indptr` (device tensor, int32/int64)."""
    n_rows = int(n_rows)
    nb = len(config.uppers) + 1
    meta = torch.empty(nb + 1 + 3, dtype=torch.int64, device=indptr.device)
    rows = _launch_row_bins(indptr, n_rows, config, meta[: nb + 1])
    if n_rows > 0:
        ip = indptr[: n_rows + 1]
        meta[nb + 1] = ip[0]
        meta[nb + 2] = ip[n_rows]
        meta[nb + 3] = (ip[1:] - ip[:-1]).max()
    else:
        meta[nb + 1:] = 0
    host = meta.cpu().tolist()  # host sync (prepare time only)
    first, last, max_len = host[nb + 1:]
    return _finish_plan(indptr, n_rows, config, rows, host[: nb + 1],
                        _row_len_if_constant(n_rows, first, last, max_len), first)


def inspect_csr(indptr, indices, n_rows, nnz, config=None):
    """Validation figures of a CSR and, with `config`, its row-bin plan, for
    ONE host readback.  Returns (stats, plan): stats has indptr_first,
    indptr_last, n_decreasing, min_col, max_col (None without nonzeros) and
    max_row; plan is None without `config` or when the row pointer is not a
    valid CSR (the caller raises on the stats first)."""
    device = indptr.device
    n_rows, nnz = int(n_rows), int(nnz)
    nb = len(config.uppers) + 1 if config is not None else 0
    meta = torch.empty(6 + (nb + 1 if config is not None else 0), dtype=torch.int64, device=device)
    _lib.check(
        _lib.load().fsp_csr_inspect(
            _index_code(indptr), _lib.ptr(indptr), _index_code(indices), _lib.ptr(indices) if nnz else None,
            n_rows, nnz, meta.data_ptr(), _lib.stream_ptr(device),
        ),
        "fsp_csr_inspect",
    )
    rows = _launch_row_bins(indptr, n_rows, config, meta[6:]) if config is not None else None
    host = meta.cpu().tolist()  # the one host sync
    stats = {
        "indptr_first": host[0],
        "indptr_last": host[1],
        "n_decreasing": host[2],
        "min_col": host[3] if nnz else None,
        "max_col": host[4] if nnz else None,
        "max_row": host[5] if n_rows else 0,
    }
    plan = None
    if config is not None and host[0] == 0 and host[1] == nnz and host[2] == 0:
        plan = _finish_plan(indptr, n_rows, config, rows, host[6:],
                            _row_len_if_constant(n_rows, host[0], host[1], stats["max_row"]), host[0])
    return stats, plan


def _finish_plan(indptr, n_rows, config, rows, off, row_len=0, indptr_base=0):
    """The plan from the bin offsets `off` (host list) of a valid CSR.
    row_len > 0 (every row that long, indptr[i] = indptr_base + i * row_len)
    is kept for a single-bin plan: its kernels then compute row bounds."""
    device = indptr.device
    nb = len(config.uppers) + 1
    counts = [off[i + 1] - off[i] for i in range(nb)]
    regular = counts[:-1]
    n_long = counts[-1]
    identity = n_long == 0 and sum(1 for c in regular if c) <= 1
    if identity and n_rows:
        # one non-empty regular bin holding every row, in order: drop the list
        regular = [n_rows if c else 0 for c in regular]
    empty_i32 = torch.empty(0, dtype=torch.int32, device=device)
    empty_i64 = torch.empty(0, dtype=torch.int64, device=device)
    seg_row = seg_begin = seg_end = long_seg_ptr = empty_i64
    long_rows = empty_i32
    n_seg = 0
    if n_long:
        long_rows = rows[off[nb - 1]: off[nb]].clone()
        lr = long_rows.to(torch.int64)
        ip = indptr.to(torch.int64)
        host = torch.stack([lr, ip[lr], ip[lr + 1]]).cpu().numpy()  # one readback
        host_rows, starts, ends = host[0].astype(np.int32), host[1], host[2]
        lens = ends - starts
        seg_len = config.seg_len
        nseg = (lens + seg_len - 1) // seg_len
        ptr = np.zeros(n_long + 1, dtype=np.int64)
        np.cumsum(nseg, out=ptr[1:])
        n_seg = int(ptr[-1])
        owner = np.repeat(np.arange(n_long, dtype=np.int64), nseg)
        j = np.arange(n_seg, dtype=np.int64) - ptr[owner]
        # even split: segment j of a row covers [len*j/n, len*(j+1)/n)
        b = starts[owner] + (lens[owner] * j) // nseg[owner]
        e = starts[owner] + (lens[owner] * (j + 1)) // nseg[owner]
        seg_row = torch.from_numpy(host_rows[owner].copy()).to(device)
        seg_begin = torch.from_numpy(b).to(device)
        seg_end = torch.from_numpy(e).to(device)
        long_seg_ptr = torch.from_numpy(ptr).to(device)
    return CsrPlan(
        config=config,
        n_rows=n_rows,
        rows=rows,
        bin_counts=regular,
        identity=bool(identity),
        seg_row=seg_row if n_seg else empty_i32,
        seg_begin=seg_begin,
        seg_end=seg_end,
        long_rows=long_rows,
        long_seg_ptr=long_seg_ptr,
        n_segments=n_seg,
        n_long_rows=int(n_long),
        row_len=int(row_len) if identity else 0,
        indptr_base=int(indptr_base) if identity and row_len else 0,
    )
