"""BSR SpMV y = op(A) x on gfx950 (reference: spmv_bsr.py).

Same public names, arguments, validation order and messages as the
reference's ``prepare_spmv_bsr`` / ``flagsparse_spmv_bsr`` /
``PreparedBsrSpmv`` (spmv_bsr.py:109-181, 448-573, 895-1024): square blocks
``data[nnzb, d, d]`` row-major, block columns per block row, ``indptr`` over
block rows; ``x`` of logical or padded length; the result has the padded
length (n_block_rows * d, or n_block_cols * d for trans / conj); meta keys
of ``return_meta``.

The compute is one HIP kernel (``fsp_bsr_spmv``) for both of the reference's
algorithms ("base" and "blockrow_reduce"; their argument checks are kept):
one wave per output block row, deterministic and atomic-free.  op trans /
conj runs the same kernel over the block CSC of A (built once per prepared
object with the CSR -> CSC radix sort over the block structure), so no
atomics and no transposed copy.  Differences: block_dim is limited to 32;
complex dtypes pass validation but have no HIP kernel (TypeError at launch);
f32 accumulates in f64 as the CSR SpMV does.
"""

import time

import torch

from . import _lib
from ._common import normalize_op, op_to_name, op_transposes, require_kernel_dtype
from ._csc import csr_to_csc_perm

SUPPORTED_SPMV_BSR_VALUE_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)
SPMV_BSR_OP_NON, SPMV_BSR_OP_TRANS, SPMV_BSR_OP_CONJ_TRANS = 0, 1, 2
SPMV_BSR_SUPPORTED_OP_NAMES = ("non", "trans", "conj")
SPMV_BSR_ALG_BASE = "base"
SPMV_BSR_ALG_BLOCKROW_REDUCE = "blockrow_reduce"
SPMV_BSR_SUPPORTED_ALGORITHMS = (SPMV_BSR_ALG_BASE, SPMV_BSR_ALG_BLOCKROW_REDUCE)
MAX_BLOCK_DIM = 32


def _spmv_bsr_dtype_error_message():
    return "BSR SpMV supports float32, float64, complex64, and complex128"


def _normalize_spmv_bsr_algorithm(use_opt=False):
    """spmv_bsr.py:83-99."""
    if isinstance(use_opt, str):
        token = use_opt.strip().lower().replace("-", "_")
        if token in ("false", "0", "no", "base", "spmv_bsr_base"):
            return SPMV_BSR_ALG_BASE
        if token in ("true", "1", "yes", "opt", "auto", "blockrow_reduce", "spmv_bsr_blockrow_reduce"):
            return SPMV_BSR_ALG_BLOCKROW_REDUCE
        raise ValueError("use_opt must be False, True, 'base', 'auto', or 'blockrow_reduce'")
    return SPMV_BSR_ALG_BLOCKROW_REDUCE if bool(use_opt) else SPMV_BSR_ALG_BASE


def _normalize_spmv_bsr_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


class PreparedBsrSpmv:
    """Prepared BSR metadata for repeated SpMV calls (spmv_bsr.py:109-180),
    plus the cached block CSC for trans / conj."""

    __slots__ = (
        "data", "kernel_indices", "kernel_indptr", "shape", "n_rows", "n_cols", "padded_n_rows", "padded_n_cols",
        "block_dim", "n_block_rows", "n_block_cols", "nnzb", "stored_nnz", "block_row_lengths",
        "max_block_row_nnz", "block_nnz", "max_segments", "op", "transpose", "index_fallback_policy",
        "index_fallback_applied", "index_fallback_reason", "_block_csc",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, block_dim, n_block_rows, n_block_cols, block_nnz,
                 max_segments, max_block_row_nnz, block_row_lengths=None, op=None, transpose=False,
                 index_fallback_policy="auto", index_fallback_applied=False, index_fallback_reason=None):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.block_dim = int(block_dim)
        self.n_block_rows = int(n_block_rows)
        self.n_block_cols = int(n_block_cols)
        self.padded_n_rows = self.n_block_rows * self.block_dim
        self.padded_n_cols = self.n_block_cols * self.block_dim
        self.nnzb = int(data.shape[0])
        self.stored_nnz = int(data.numel())
        if block_row_lengths is None:
            block_row_lengths = kernel_indptr[1:] - kernel_indptr[:-1]
        self.block_row_lengths = block_row_lengths
        self.max_block_row_nnz = int(max_block_row_nnz)
        self.block_nnz = int(block_nnz)
        self.max_segments = int(max_segments)
        self.op = normalize_op(op, transpose=transpose)
        self.transpose = op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self._block_csc = None

    def block_csc(self):
        """(block colptr, block rows, block-id permutation) of A, cached."""
        if self._block_csc is None:
            self._block_csc = csr_to_csc_perm(self.kernel_indptr, self.kernel_indices, self.n_block_rows,
                                              self.n_block_cols)
        return self._block_csc


def _prepare_spmv_bsr_matrix(data, indices, indptr, shape, block_dim):
    """Validation contract of spmv_bsr.py:448-507 (same order and messages)."""
    if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
        raise TypeError("data, indices, indptr must all be torch.Tensor")
    if data.ndim != 3:
        raise ValueError("data must have shape (nnzb, block_dim, block_dim)")
    if indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("indices and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    block_dim = int(block_dim)
    if block_dim <= 1:
        raise ValueError("block_dim must be greater than 1 for BSR SpMV")
    if block_dim > MAX_BLOCK_DIM:
        raise ValueError(f"block_dim must be at most {MAX_BLOCK_DIM} for the HIP BSR SpMV")
    if data.shape[1] != block_dim or data.shape[2] != block_dim:
        raise ValueError("data block dimensions must match block_dim")
    n_block_rows = (n_rows + block_dim - 1) // block_dim
    n_block_cols = (n_cols + block_dim - 1) // block_dim
    if indptr.numel() != n_block_rows + 1:
        raise ValueError(f"indptr length must be n_block_rows+1={n_block_rows + 1}, got {indptr.numel()}")
    if data.shape[0] != indices.numel():
        raise ValueError("data.shape[0] and indices length must both equal nnzb")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMV_BSR_VALUE_DTYPES:
        raise TypeError(_spmv_bsr_dtype_error_message())
    if indices.dtype not in (torch.int32, torch.int64):
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    if int(indptr[0].item()) != 0:
        raise ValueError("indptr must start at zero")
    if int(indptr[-1].item()) != data.shape[0]:
        raise ValueError("indptr[-1] must equal nnzb")
    if indptr.numel() > 1 and torch.any(indptr[1:] < indptr[:-1]).item():
        raise ValueError("indptr must be non-decreasing")
    if indices.numel() > 0:
        min_index = int(indices.min().item())
        max_index = int(indices.max().item())
        if min_index < 0 or max_index >= n_block_cols:
            raise IndexError("indices out of range for n_block_cols")
    block_row_lengths = indptr[1:] - indptr[:-1]
    max_block_row_nnz = int(block_row_lengths.max().item()) if n_block_rows > 0 else 0
    return (data, indices, indptr, n_rows, n_cols, n_block_rows, n_block_cols, block_row_lengths,
            max_block_row_nnz)


def prepare_spmv_bsr(data, indices, indptr, shape, block_dim, block_nnz=128, max_segments=None, transpose=False,
                     op=None, use_opt=False, index_fallback_policy="auto"):
    """spmv_bsr.py:510-571."""
    index_fallback_policy = _normalize_spmv_bsr_index_fallback_policy(index_fallback_policy)
    op_code = normalize_op(op, transpose=transpose)
    algorithm = _normalize_spmv_bsr_algorithm(use_opt)
    if algorithm == SPMV_BSR_ALG_BLOCKROW_REDUCE and op_code != SPMV_BSR_OP_NON:
        raise ValueError("spmv_bsr_blockrow_reduce only supports op='non'")
    if algorithm == SPMV_BSR_ALG_BLOCKROW_REDUCE and int(block_dim) not in (2, 4, 8, 16):
        raise ValueError("spmv_bsr_blockrow_reduce supports block_dim values 2, 4, 8, and 16")
    (data, indices, indptr, n_rows, n_cols, n_block_rows, n_block_cols, block_row_lengths,
     max_block_row_nnz) = _prepare_spmv_bsr_matrix(data, indices, indptr, shape, block_dim)
    block_nnz_use = int(block_nnz)
    if block_nnz_use <= 0:
        raise ValueError("block_nnz must be positive")
    if max_segments is None:
        max_segments_use = max((max_block_row_nnz + block_nnz_use - 1) // block_nnz_use, 1)
        while max_segments_use > 2048 and block_nnz_use < 65536:
            block_nnz_use *= 2
            max_segments_use = max((max_block_row_nnz + block_nnz_use - 1) // block_nnz_use, 1)
    else:
        max_segments_use = max(1, int(max_segments))
    return PreparedBsrSpmv(data=data, kernel_indices=indices, kernel_indptr=indptr, shape=shape, block_dim=block_dim,
                           n_block_rows=n_block_rows, n_block_cols=n_block_cols, block_nnz=block_nnz_use,
                           max_segments=max_segments_use, max_block_row_nnz=max_block_row_nnz,
                           block_row_lengths=block_row_lengths, op=op_code,
                           index_fallback_policy=index_fallback_policy)


def _validate_spmv_bsr_x(x, prepared, op_code):
    """spmv_bsr.py:574-600: logical or padded length; padded with zeros."""
    if x is None or not torch.is_tensor(x):
        raise TypeError("x must be a torch.Tensor")
    if x.ndim != 1:
        raise ValueError("x must be a 1D tensor")
    if not x.is_cuda:
        raise ValueError("x must be a CUDA tensor")
    if x.dtype != prepared.data.dtype:
        raise TypeError("x dtype must match sparse matrix dtype")
    trans = op_transposes(op_code)
    logical_expected = prepared.n_rows if trans else prepared.n_cols
    padded_expected = prepared.padded_n_rows if trans else prepared.padded_n_cols
    if x.numel() not in (logical_expected, padded_expected):
        raise ValueError(f"x length must be {logical_expected} or padded length {padded_expected}, got {x.numel()}")
    if x.device != prepared.data.device:
        raise ValueError("x must be on the same device as sparse matrix data")
    x = x.contiguous()
    if x.numel() == padded_expected:
        return x
    padded = torch.zeros(padded_expected, dtype=x.dtype, device=x.device)
    padded[: x.numel()].copy_(x)
    return padded


def _hip_spmv_bsr(prepared, x, op_code, out=None):
    """The launch: y (padded) = op(A) x, one kernel over block rows (non) or
    block columns through the block CSC (trans / conj)."""
    dtype = prepared.data.dtype
    require_kernel_dtype(dtype)
    trans = op_transposes(op_code)
    n_out_blocks = prepared.n_block_cols if trans else prepared.n_block_rows
    y = out if out is not None else torch.empty(n_out_blocks * prepared.block_dim, dtype=dtype, device=x.device)
    if n_out_blocks == 0:
        return y
    if trans:
        ptr, bidx, perm = prepared.block_csc()
    else:
        ptr, bidx, perm = prepared.kernel_indptr, prepared.kernel_indices, None
    _lib.check(
        _lib.load().fsp_bsr_spmv(
            _lib.value_code(dtype), _lib.index_code(ptr.dtype), _lib.index_code(bidx.dtype), ptr.data_ptr(),
            _lib.ptr(bidx), _lib.index_code(perm.dtype) if perm is not None else 0, _lib.ptr(perm),
            _lib.ptr(prepared.data), prepared.block_dim, x.data_ptr(), y.data_ptr(), n_out_blocks,
            _lib.stream_ptr(x.device),
        ),
        "fsp_bsr_spmv",
    )
    return y


def flagsparse_spmv_bsr(data=None, indices=None, indptr=None, x=None, shape=None, block_dim=None, block_nnz=128,
                        max_segments=None, out=None, return_time=False, return_meta=False, prepared=None,
                        transpose=None, op=None, use_opt=False, index_fallback_policy="auto"):
    """BSR SpMV (spmv_bsr.py:895-1024); the result has the padded length."""
    op_explicit = op is not None
    op_code = normalize_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    algorithm = _normalize_spmv_bsr_algorithm(use_opt)
    if algorithm == SPMV_BSR_ALG_BLOCKROW_REDUCE and op_code != SPMV_BSR_OP_NON:
        raise ValueError("spmv_bsr_blockrow_reduce only supports op='non'")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape, block_dim)):
            raise ValueError("data, indices, indptr, shape, and block_dim are required when prepared is not provided")
        prepared = prepare_spmv_bsr(data, indices, indptr, shape, block_dim, block_nnz=block_nnz,
                                    max_segments=max_segments, op=op_code, use_opt=algorithm,
                                    index_fallback_policy=index_fallback_policy)
    else:
        if op_explicit and op_code != prepared.op:
            raise ValueError(f"op={op_to_name(op_code)} does not match prepared.op={op_to_name(prepared.op)}")
        if not op_explicit and transpose is not None and bool(transpose) != prepared.transpose:
            raise ValueError(f"transpose={bool(transpose)} does not match prepared.transpose={prepared.transpose}")
        if not op_explicit:
            op_code = prepared.op
    if algorithm == SPMV_BSR_ALG_BLOCKROW_REDUCE and op_code != SPMV_BSR_OP_NON:
        raise ValueError("spmv_bsr_blockrow_reduce only supports op='non'")
    if algorithm == SPMV_BSR_ALG_BLOCKROW_REDUCE and prepared.block_dim not in (2, 4, 8, 16):
        raise ValueError("spmv_bsr_blockrow_reduce supports block_dim values 2, 4, 8, and 16")
    x = _validate_spmv_bsr_x(x, prepared, op_code)
    out_len = (prepared.padded_n_cols if op_transposes(op_code) else prepared.padded_n_rows)
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != x.device:
            raise ValueError("out must be on the same CUDA device as the result")
        if tuple(out.shape) != (out_len,) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    do_timing = bool(return_time or return_meta)
    if do_timing:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    target = out if (out is not None and out.is_contiguous()) else None
    y = _hip_spmv_bsr(prepared, x, op_code, out=target)
    if out is not None and target is None:
        out.copy_(y)
        y = out
    op_total_ms = None
    if do_timing:
        torch.cuda.synchronize()
        op_total_ms = (time.perf_counter() - t0) * 1000.0
    if return_meta:
        meta = {
            "algorithm": algorithm,
            "op": op_to_name(op_code),
            "block_dim": prepared.block_dim,
            "logical_shape": prepared.shape,
            "padded_shape": (prepared.padded_n_rows, prepared.padded_n_cols),
            "n_block_rows": prepared.n_block_rows,
            "n_block_cols": prepared.n_block_cols,
            "nnzb": prepared.nnzb,
            "stored_nnz": prepared.stored_nnz,
            "symbolic_ms": 0.0,
            "process_cpu_ms": 0.0,
            "process_gpu_ms": 0.0,
            "compute_ms": op_total_ms,
            "op_total_ms": op_total_ms,
            "bucket_counts": None,
            "launch_configs": [{"kernel": "bsr_spmv", "waves_per_block_row": 1, "block_dim": prepared.block_dim}],
            "index_fallback_applied": prepared.index_fallback_applied,
            "index_fallback_reason": prepared.index_fallback_reason,
        }
        if return_time:
            return y, op_total_ms, meta
        return y, meta
    if return_time:
        return y, op_total_ms
    return y


__all__ = [
    "PreparedBsrSpmv",
    "prepare_spmv_bsr",
    "flagsparse_spmv_bsr",
    "SUPPORTED_SPMV_BSR_VALUE_DTYPES",
    "SPMV_BSR_SUPPORTED_ALGORITHMS",
]
