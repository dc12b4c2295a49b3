"""Batches of CSR matrices: SpMV / SpMM of every matrix in one launch.

No counterpart in the reference, which runs one matrix per call.  A batch is
a packed offset table over concatenated arrays (``CsrBatch``): matrix m owns
rows ``row_offsets[m]:row_offsets[m+1]``, columns
``col_offsets[m]:col_offsets[m+1]`` and nonzeros
``nnz_offsets[m]:nnz_offsets[m+1]`` of the concatenated ``indices`` /
``data``; its own row pointer (n_m + 1 local offsets from 0) is
``indptr[row_offsets[m] + m : row_offsets[m+1] + m + 1]``.  Dense operands and
results are packed the same way: x / B rows by column offsets, y / C rows by
row offsets.

At construction the batch is packed once on the GPU into its block-diagonal
CSR (global row pointer and column ids; the values are shared, so they may
change between calls) and validated.  ``spmv`` / ``spmm`` are then ONE
launch over that matrix -- the row-bin kernels (persistent when the
autotuner picks a capped grid) or, for the SpMV, the row-range parts kernel
that keeps each matrix's column range of x in LDS -- and ``graph_spmv`` /
``graph_spmm`` capture it in a HIP graph.  Each row is reduced in the same
order as in a launch of the same kernel over its own matrix, so the batched
result is bitwise equal to ``spmv_per_matrix`` / ``spmm_per_matrix``, the
per-matrix launch loop kept as the comparator.
"""

import numpy as np
import torch

from . import _lib, _plan, autotune
from ._matrix import CsrHandle

_VALUE_DTYPES = (torch.float32, torch.float64, torch.float16, torch.bfloat16)
_ERR_MESSAGES = (
    (1, "every matrix's indptr must start at 0 and end at its nnz"),
    (2, "indptr must be nondecreasing"),
    (4, "indices out of range for their matrix's n_cols"),
)


def _host_offsets(v, name, n=None):
    a = np.asarray(v.cpu() if torch.is_tensor(v) else v, dtype=np.int64).reshape(-1)
    if a.size < 2 or (n is not None and a.size != n):
        raise ValueError(f"{name} must have n_mats + 1 >= 2 entries")
    if a[0] != 0 or np.any(np.diff(a) < 0):
        raise ValueError(f"{name} must start at 0 and be nondecreasing")
    return a


def parts_table(row_offsets, col_offsets, nnz_offsets, part_nnz):
    """Row-range parts of a batch for fsp_spmv_csr_parts (host arrays).

    Matrix m's rows are cut evenly into min(rows_m, ceil(nnz_m / part_nnz))
    ranges (at least one for a matrix with rows); each part's window is the
    matrix's column range.  Returns (table int64 [P, 4] of (row_begin,
    row_end, col_lo, col_hi), start int64 [n_mats + 1]: matrix m owns parts
    start[m] .. start[m+1]-1)."""
    rows = np.asarray(row_offsets, dtype=np.int64)
    cols = np.asarray(col_offsets, dtype=np.int64)
    nnz = np.asarray(nnz_offsets, dtype=np.int64)
    n_mats = rows.size - 1
    n_rows = rows[1:] - rows[:-1]
    n_parts = np.where(n_rows > 0, np.maximum(1, -(-(nnz[1:] - nnz[:-1]) // max(1, int(part_nnz)))), 0)
    n_parts = np.minimum(n_parts, n_rows)
    start = np.concatenate([[0], np.cumsum(n_parts)]).astype(np.int64)
    owner = np.repeat(np.arange(n_mats, dtype=np.int64), n_parts)
    k = np.arange(int(start[-1]), dtype=np.int64) - start[owner]
    rb = rows[owner] + n_rows[owner] * k // n_parts[owner]
    re = rows[owner] + n_rows[owner] * (k + 1) // n_parts[owner]
    table = np.stack([rb, re, cols[owner], cols[owner + 1]], 1).astype(np.int64).reshape(-1, 4)
    return table, start


class CsrBatch:
    """A packed batch of CSR matrices (see the module docstring)."""

    def __init__(self, data, indices, indptr, row_offsets, col_offsets=None, nnz_offsets=None):
        if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
            raise TypeError("data, indices, indptr must all be torch.Tensor")
        if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
            raise ValueError("data, indices, and indptr must be 1D tensors")
        if not all(t.is_cuda for t in (data, indices, indptr)):
            raise ValueError("data, indices, and indptr must be CUDA tensors")
        if not all(t.device == data.device for t in (indices, indptr)):
            raise ValueError("data, indices, and indptr must be on the same CUDA device")
        if data.dtype not in _VALUE_DTYPES:
            raise TypeError("data dtype must be one of: float16, bfloat16, float32, float64")
        if indices.dtype not in (torch.int32, torch.int64) or indptr.dtype not in (torch.int32, torch.int64):
            raise TypeError("indices and indptr dtype must be torch.int32 or torch.int64")
        if data.numel() != indices.numel():
            raise ValueError("data and indices must have the same length (nnz)")
        rows = _host_offsets(row_offsets, "row_offsets")
        n_mats = rows.size - 1
        cols = rows if col_offsets is None else _host_offsets(col_offsets, "col_offsets", n_mats + 1)
        if indptr.numel() != rows[-1] + n_mats:
            raise ValueError(f"indptr length must be total_rows + n_mats = {rows[-1] + n_mats}, got {indptr.numel()}")
        if nnz_offsets is None:
            # each matrix's nnz is the last entry of its own row pointer
            ends = torch.as_tensor(rows[1:] + np.arange(n_mats), device=indptr.device)
            nnz = np.concatenate([[0], np.cumsum(indptr[ends].cpu().numpy().astype(np.int64))])
            nnz = _host_offsets(nnz, "per-matrix nnz (last indptr entries)")
        else:
            nnz = _host_offsets(nnz_offsets, "nnz_offsets", n_mats + 1)
        if nnz[-1] != data.numel():
            raise ValueError(f"nnz_offsets[-1] must equal nnz={data.numel()}, got {nnz[-1]}")
        self.device = data.device
        self.data = data.contiguous()
        self.local_indices = indices.contiguous()
        self.local_indptr = indptr.contiguous()
        self.row_offsets, self.col_offsets, self.nnz_offsets = rows, cols, nnz
        self.n_mats = int(n_mats)
        self.total_rows, self.total_cols, self.nnz = int(rows[-1]), int(cols[-1]), int(nnz[-1])
        self._pack()
        self._params = {}
        self._parts = {}
        self._per_matrix = None

    @classmethod
    def from_matrices(cls, matrices):
        """Pack a list of (data, indices, indptr, shape) CSR matrices (copies)."""
        if not matrices:
            raise ValueError("a batch needs at least one matrix")
        dtypes = {m[0].dtype for m in matrices}
        if len(dtypes) != 1:
            raise TypeError(f"every matrix of a batch needs the same value dtype, got {sorted(map(str, dtypes))}")
        shapes = [(int(s[0]), int(s[1])) for _, _, _, s in matrices]
        rows = np.concatenate([[0], np.cumsum([s[0] for s in shapes])])
        cols = np.concatenate([[0], np.cumsum([s[1] for s in shapes])])
        nnz = np.concatenate([[0], np.cumsum([int(m[0].numel()) for m in matrices])])
        data = torch.cat([m[0] for m in matrices])
        idx_dtype = torch.int64 if any(m[1].dtype == torch.int64 for m in matrices) else torch.int32
        ptr_dtype = torch.int64 if any(m[2].dtype == torch.int64 for m in matrices) else torch.int32
        indices = torch.cat([m[1].to(idx_dtype) for m in matrices])
        indptr = torch.cat([m[2].to(ptr_dtype) for m in matrices])
        return cls(data, indices, indptr, rows, cols, nnz)

    # ------------------------------------------------------------- packing
    def _pack(self):
        dev = self.device
        ptr_t = torch.int32 if self.nnz <= np.iinfo(np.int32).max else torch.int64
        idx_t = torch.int32 if self.total_cols <= np.iinfo(np.int32).max else torch.int64
        self.indptr = torch.empty(self.total_rows + 1, dtype=ptr_t, device=dev)
        self.indices = torch.empty(self.nnz, dtype=idx_t, device=dev)
        offs = torch.from_numpy(np.stack([self.row_offsets, self.col_offsets, self.nnz_offsets])).to(dev)
        err = torch.empty(1, dtype=torch.int32, device=dev)
        _lib.check(
            _lib.load().fsp_batch_pack(
                _lib.index_code(self.local_indptr.dtype), _lib.index_code(self.local_indices.dtype), self.n_mats,
                offs[0].data_ptr(), offs[1].data_ptr(), offs[2].data_ptr(), self.local_indptr.data_ptr(),
                _lib.ptr(self.local_indices), self.total_rows, self.nnz, _lib.index_code(ptr_t),
                self.indptr.data_ptr(), _lib.index_code(idx_t), _lib.ptr(self.indices), err.data_ptr(),
                _lib.stream_ptr(dev),
            ),
            "fsp_batch_pack",
        )
        bits = int(err.item())  # host sync, construction time only
        for bit, msg in _ERR_MESSAGES:
            if bits & bit:
                if bit == 4:
                    raise IndexError(msg)
                raise ValueError(msg)
        self.handle = CsrHandle(self.indices, self.indptr, (self.total_rows, self.total_cols))
        lengths = self.indptr[1:] - self.indptr[:-1] if self.total_rows else self.indptr[:0]
        self.max_row = int(lengths.max().item()) if self.total_rows else 0
        # every row of every matrix the same length: the parts loop computes
        # row bounds (the packed indptr starts at 0)
        self.row_len = _plan._row_len_if_constant(self.total_rows, 0, self.nnz, self.max_row)

    # ------------------------------------------------------------- tuning
    def params(self, kind, n_dense=1, dtype=None, tune=True, data=None):
        """Kernel parameters of this batch for `kind` ("spmv" / "spmm") run
        with the values `data` (default: the batch's own): the cached winner
        of its shape class, else tuned now (tune=True) or the defaults.  The
        SpMV class includes whether the call's values sit on 4-element
        boundaries (the vector-chunk and parts loops need it), so a winner
        tuned on aligned values is never served to unaligned ones."""
        data = self.data if data is None else data
        dtype = dtype or data.dtype
        if data.dtype != dtype:
            data = data.to(dtype)
        cus = autotune.compute_units(self.device)
        tag = ""
        if kind == "spmv":
            aligned = _plan.chunk_aligned(self.indices, data)
            tag = autotune.spmv_class_tag(self.total_cols, self.indices, data) + ("|a4" if aligned else "|u")
        key = autotune.shape_class(kind, dtype, self.indices.dtype, self.indptr.dtype, n_dense, self.total_rows,
                                   self.nnz, self.max_row, cus, extra=tag)
        p = self._params.get(key) or autotune.lookup(key)
        if p is None:
            if not tune:
                return autotune.TuneParams()
            p = self._tune(kind, key, n_dense, data)
        self._params[key] = p
        return p

    def _tune(self, kind, key, n_dense, data):
        dtype = data.dtype
        elem = data.element_size()
        chunk_ok = _plan.chunk_aligned(self.indices, data)
        if kind == "spmv":
            cands = autotune.batch_spmv_candidates(elem, self.total_cols, chunk_ok)
        else:
            cands = autotune.candidates(kind, n_dense, elem, self.total_cols, chunk_ok)
        if kind == "spmv":
            x = torch.ones(self.total_cols, dtype=dtype, device=self.device)
            y = torch.empty(self.total_rows, dtype=dtype, device=self.device)
            for p in cands:  # plans / parts are built (host syncs) before any timing
                if p.variant == 5:
                    self.parts(p.part_nnz)
                else:
                    self.handle.plan(p.spmv_config())

            def run(p):
                self.spmv(x, out=y, data=data, params=p)
        else:
            B = torch.ones((self.total_cols, n_dense), dtype=dtype, device=self.device)
            C = torch.empty((self.total_rows, n_dense), dtype=dtype, device=self.device)
            for p in cands:
                self.handle.plan(p.spmm_config(self._lanes(p, n_dense, elem)))

            def run(p):
                self.handle.spmm(data, B, out=C, lanes_n=self._lanes(p, n_dense, elem),
                                 config=p.spmm_config(self._lanes(p, n_dense, elem)),
                                 max_blocks=p.max_blocks(self.device), xcd_groups=p.xcd_groups)
        best, self.tune_results = autotune.tune(key, cands, run)
        return best

    @staticmethod
    def _lanes(p, n_dense, elem):
        return p.lanes_n or _plan.lanes_for_dense(n_dense, elem)

    # ------------------------------------------------------------ kernels
    def _values(self, data):
        """The values a call runs with: the batch's own (None) or a replacement
        of the same pattern, one per nonzero on the batch's device."""
        if data is None:
            return self.data
        if not torch.is_tensor(data) or not data.is_cuda or data.device != self.device:
            raise ValueError("data must be a CUDA tensor on the batch's device")
        if data.ndim != 1 or data.numel() != self.nnz:
            raise ValueError(f"data must be 1D with nnz={self.nnz} entries, got shape {tuple(data.shape)}")
        if data.dtype not in _VALUE_DTYPES:
            raise TypeError("data dtype must be one of: float16, bfloat16, float32, float64")
        return data.contiguous()

    def _check_dense(self, t, n, what):
        if not torch.is_tensor(t) or not t.is_cuda or t.device != self.device:
            raise ValueError(f"{what} must be a CUDA tensor on the batch's device")
        if t.shape[0] != n:
            raise ValueError(f"{what}.shape[0] must be {n}, got {t.shape[0]}")

    def _check_out(self, out, shape, dtype):
        """`out` (written in place by the kernels): a contiguous tensor of the
        result's shape and dtype on the batch's device, or None."""
        if out is None:
            return
        if (not torch.is_tensor(out) or out.device != self.device or tuple(out.shape) != tuple(shape)
                or out.dtype != dtype or not out.is_contiguous()):
            raise ValueError(f"out must be a contiguous {dtype} tensor of shape {tuple(shape)} on the batch's device")

    def spmv(self, x, out=None, data=None, params=None):
        """y (packed by rows) = blockdiag(A_m) x (packed by columns): one launch."""
        data = self._values(data)
        self._check_dense(x, self.total_cols, "x")
        if x.ndim != 1 or x.dtype != data.dtype:
            raise ValueError("x must be 1D with the values' dtype")
        self._check_out(out, (self.total_rows,), data.dtype)
        p = params or self.params("spmv", 1, data.dtype, data=data)
        if p.variant in (4, 5) and not _plan.chunk_aligned(self.indices, data):
            # the chunk loops read 4-element chunks: unaligned values take the
            # pipelined loop (same row order, results within rounding)
            p = autotune.TuneParams()
        if p.variant == 5:
            y = out if out is not None else torch.empty(self.total_rows, dtype=data.dtype, device=self.device)
            parts, _ = self.parts(p.part_nnz)
            self._spmv_parts(data, x.contiguous(), y, parts, p.lanes_n)
            return y
        return self.handle.spmv(data, x.contiguous(), out=out, config=p.spmv_config(),
                                max_blocks=p.max_blocks(self.device), xcd_groups=p.xcd_groups, variant=p.variant)

    # ----------------------------------------------- row-range parts
    def parts(self, part_nnz):
        """(device int64 [P, 4] parts, host start index of each matrix's parts)
        for fsp_spmv_csr_parts: every matrix's rows cut evenly into
        ceil(nnz_m / part_nnz) row ranges, each with the matrix's column range
        as its window (built once per part size)."""
        key = int(part_nnz)
        hit = self._parts.get(key)
        if hit is not None:
            return hit
        table, start = parts_table(self.row_offsets, self.col_offsets, self.nnz_offsets, key)
        dev = torch.from_numpy(np.ascontiguousarray(table)).to(self.device)
        self._parts[key] = (dev, start)
        return self._parts[key]

    def _spmv_parts(self, data, x, y, parts, lanes):
        if parts.shape[0] == 0:
            return
        _lib.check(
            _lib.load().fsp_spmv_csr_parts(
                _lib.value_code(data.dtype), _lib.index_code(self.indices.dtype), _lib.index_code(self.indptr.dtype),
                self.indptr.data_ptr(), _lib.ptr(self.indices), _lib.ptr(data), _lib.ptr(x), y.data_ptr(),
                int(data.numel()), int(parts.shape[0]), parts.data_ptr(), int(lanes), self.row_len, 0,
                _lib.stream_ptr(self.device),
            ),
            "fsp_spmv_csr_parts",
        )

    def spmm(self, B, out=None, data=None, params=None):
        """C (packed by rows) = blockdiag(A_m) B (packed by rows): one launch."""
        data = self._values(data)
        self._check_dense(B, self.total_cols, "B")
        if B.ndim != 2 or B.dtype != data.dtype:
            raise ValueError("B must be 2D with the values' dtype")
        n_dense = int(B.shape[1])
        self._check_out(out, (self.total_rows, n_dense), data.dtype)
        p = params or self.params("spmm", n_dense, data.dtype)
        lanes = self._lanes(p, n_dense, data.element_size())
        return self.handle.spmm(data, B.contiguous(), out=out, lanes_n=lanes, config=p.spmm_config(lanes),
                                max_blocks=p.max_blocks(self.device), xcd_groups=p.xcd_groups)

    # ----------------------------------------------------- graph capture
    def _capture(self, fn):
        fn()  # plans and parameters exist before capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        return g

    def graph_spmv(self, x, y, data=None):
        """HIP graph of spmv(x, out=y); replay() reruns it on the current x."""
        if y is None:
            raise ValueError("graph_spmv needs the output tensor y: the graph writes into it")
        return self._capture(lambda: self.spmv(x, out=y, data=data))

    def graph_spmm(self, B, C, data=None):
        """HIP graph of spmm(B, out=C); replay() reruns it on the current B."""
        if C is None:
            raise ValueError("graph_spmm needs the output tensor C: the graph writes into it")
        return self._capture(lambda: self.spmm(B, out=C, data=data))

    # ------------------------------------------- per-matrix comparator
    def _matrices(self):
        if self._per_matrix is None:
            mats = []
            for m in range(self.n_mats):
                r0, r1 = int(self.row_offsets[m]), int(self.row_offsets[m + 1])
                n0, n1 = int(self.nnz_offsets[m]), int(self.nnz_offsets[m + 1])
                c0, c1 = int(self.col_offsets[m]), int(self.col_offsets[m + 1])
                h = CsrHandle(self.local_indices[n0:n1], self.local_indptr[r0 + m: r1 + m + 1], (r1 - r0, c1 - c0))
                # the matrix's rows of the packed block-diagonal CSR: a row
                # range of the packed row pointer over the packed column ids
                # (global columns, the packed arrays' own alignment)
                hg = CsrHandle(self.indices, self.indptr[r0: r1 + 1], (r1 - r0, self.total_cols))
                mats.append((h, hg, r0, r1, n0, n1, c0, c1))
            self._per_matrix = mats
        return self._per_matrix

    def spmv_per_matrix(self, x, out=None, data=None, params=None):
        """The same product as spmv, one launch per matrix (the comparator)."""
        data = self._values(data)
        self._check_dense(x, self.total_cols, "x")
        if x.ndim != 1 or x.dtype != data.dtype:
            raise ValueError("x must be 1D with the values' dtype")
        x = x.contiguous()
        self._check_out(out, (self.total_rows,), data.dtype)
        p = params or self.params("spmv", 1, data.dtype, data=data)
        if p.variant in (4, 5) and not _plan.chunk_aligned(self.indices, data):
            p = autotune.TuneParams()  # as in spmv: no chunk loops on unaligned values
        y = out if out is not None else torch.empty(self.total_rows, dtype=data.dtype, device=self.device)
        if p.variant == 5:  # each matrix's own parts, one launch per matrix
            parts, start = self.parts(p.part_nnz)
            for m in range(self.n_mats):
                self._spmv_parts(data, x, y, parts[int(start[m]):int(start[m + 1])], p.lanes_n)
            return y
        cfg = p.spmv_config()
        for _, hg, r0, r1, _n0, _n1, _c0, _c1 in self._matrices():
            # one launch per matrix over the packed arrays: the same chunk
            # boundaries and summation order as the batched launch
            hg.spmv(data, x, out=y[r0:r1], config=cfg, variant=p.variant)
        return y

    def spmm_per_matrix(self, B, out=None, data=None, params=None):
        """The same product as spmm, one launch per matrix (the comparator)."""
        data = self._values(data)
        self._check_dense(B, self.total_cols, "B")
        if B.ndim != 2 or B.dtype != data.dtype:
            raise ValueError("B must be 2D with the values' dtype")
        B = B.contiguous()
        n_dense = int(B.shape[1])
        self._check_out(out, (self.total_rows, n_dense), data.dtype)
        p = params or self.params("spmm", n_dense, data.dtype)
        lanes = self._lanes(p, n_dense, data.element_size())
        cfg = p.spmm_config(lanes)
        C = out if out is not None else torch.empty((self.total_rows, n_dense), dtype=data.dtype,
                                                    device=self.device)
        for h, _hg, r0, r1, n0, n1, c0, c1 in self._matrices():
            h.spmm(data[n0:n1], B[c0:c1], out=C[r0:r1], lanes_n=lanes, config=cfg)
        return C

    def split(self, packed):
        """Per-matrix views of a packed result (rows by row offsets)."""
        return [packed[int(self.row_offsets[m]): int(self.row_offsets[m + 1])] for m in range(self.n_mats)]
