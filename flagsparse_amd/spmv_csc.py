"""CSC SpMV: y = op(A) x for A (n_rows x n_cols) stored by columns
(reference: spmv_csc.py).

The CSC arrays of A -- column pointer, row ids, values -- are the CSR arrays
of A^T (n_cols x n_rows).  So ``op="trans"`` (and ``"conj"``, the same for
real values) is the CSR row SpMV over the given arrays, and ``op="non"`` is
the CSR SpMV of A, whose row arrays the prepare step builds once through
the stable CSC permutation (bit-exact value moves, no atomics; the
reference's column kernel scatters with atomics instead).  Both run
``fsp_spmv_csr`` through a ``PreparedCsrSpmv`` held by the prepared object.
Validation order, messages, the raw-or-prepared contract, ``out`` and the
meta keys follow spmv_csc.py:258-600.  Complex values pass the dtype check
and then raise TypeError (no HIP kernel); float16 / bfloat16 are rejected
as in the reference.
"""

import time

import torch

from . import _plan
from ._common import SUPPORTED_INDEX_DTYPES
from .spmv_csr import _run_spmv_prepared_with_fallback, prepare_spmv_csr

SUPPORTED_SPMV_CSC_VALUE_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)

SPMV_CSC_OP_NON = 0
SPMV_CSC_OP_TRANS = 1
SPMV_CSC_OP_CONJ_TRANS = 2
SPMV_CSC_OP_NAMES = {SPMV_CSC_OP_NON: "non", SPMV_CSC_OP_TRANS: "trans", SPMV_CSC_OP_CONJ_TRANS: "conj"}
_SPMV_CSC_OP_NAME_TO_CODE = {name: code for code, name in SPMV_CSC_OP_NAMES.items()}
_OP_ERROR = "op must be one of: 0=non, 1=trans, 2=conj"


def _spmv_csc_dtype_error_message():
    return "CSC SpMV supports float32, float64, complex64, and complex128"


def _normalize_spmv_csc_op(op=None, transpose=False):
    """spmv_csc.py:44-58: codes 0/1/2 or their names (case-insensitive)."""
    if op is None:
        return SPMV_CSC_OP_TRANS if bool(transpose) else SPMV_CSC_OP_NON
    if isinstance(op, str):
        token = op.strip().lower()
        if token not in _SPMV_CSC_OP_NAME_TO_CODE:
            raise ValueError(_OP_ERROR)
        return _SPMV_CSC_OP_NAME_TO_CODE[token]
    try:
        code = int(op)
    except (TypeError, ValueError) as exc:
        raise ValueError(_OP_ERROR) from exc
    if code not in SPMV_CSC_OP_NAMES:
        raise ValueError(_OP_ERROR)
    return code


def _spmv_csc_op_to_name(op):
    return SPMV_CSC_OP_NAMES[_normalize_spmv_csc_op(op)]


def _spmv_csc_op_transposes(op):
    return _normalize_spmv_csc_op(op) in (SPMV_CSC_OP_TRANS, SPMV_CSC_OP_CONJ_TRANS)


def _normalize_spmv_csc_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


class PreparedCscSpmv:
    """Prepared CSC metadata for repeated SpMV calls (spmv_csc.py:79-137);
    ``_csr`` is the prepared CSR SpMV of op(A) that the kernels run."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "nnz",
        "block_nnz",
        "max_segments",
        "col_lengths",
        "max_col_nnz",
        "op",
        "transpose",
        "index_fallback_policy",
        "index_fallback_applied",
        "index_fallback_reason",
        "_csr",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, block_nnz, max_segments,
                 max_col_nnz, col_lengths=None, op=None, transpose=False, index_fallback_policy="auto",
                 index_fallback_applied=False, index_fallback_reason=None, csr=None):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.nnz = int(data.numel())
        self.block_nnz = int(block_nnz)
        self.max_segments = int(max_segments)
        if col_lengths is None:
            col_lengths = kernel_indptr[1:] - kernel_indptr[:-1]
        self.col_lengths = col_lengths
        self.max_col_nnz = int(max_col_nnz)
        self.op = _normalize_spmv_csc_op(op, transpose=transpose)
        self.transpose = _spmv_csc_op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self._csr = csr


def _prepare_spmv_csc_matrix(data, indices, indptr, shape):
    """spmv_csc.py:258-297 (same checks, order and messages); the pointer and
    index checks come from one device inspection and one host readback."""
    if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
        raise TypeError("data, indices, indptr must all be torch.Tensor")
    if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("data, indices, indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if indptr.numel() != n_cols + 1:
        raise ValueError(f"indptr length must be n_cols+1={n_cols + 1}, got {indptr.numel()}")
    if data.numel() != indices.numel():
        raise ValueError("data and indices must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMV_CSC_VALUE_DTYPES:
        raise TypeError(_spmv_csc_dtype_error_message())
    if indices.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    nnz = int(data.numel())
    stats, _ = _plan.inspect_csr(indptr, indices, n_cols, nnz)
    if stats["indptr_first"] != 0:
        raise ValueError("indptr must start at zero")
    if stats["indptr_last"] != nnz:
        raise ValueError("indptr[-1] must equal nnz")
    if stats["n_decreasing"] != 0:
        raise ValueError("indptr must be non-decreasing")
    if nnz > 0 and (stats["min_col"] < 0 or stats["max_col"] >= n_rows):
        raise IndexError("indices out of range for n_rows")
    col_lengths = indptr[1:] - indptr[:-1]
    return data, indices, indptr, n_rows, n_cols, col_lengths, int(stats["max_row"])


def prepare_spmv_csc(data, indices, indptr, shape, block_nnz=256, max_segments=None, transpose=False, op=None,
                     index_fallback_policy="auto"):
    """spmv_csc.py:300-346, plus the CSR SpMV of op(A) the kernels run."""
    index_fallback_policy = _normalize_spmv_csc_index_fallback_policy(index_fallback_policy)
    op_code = _normalize_spmv_csc_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_CSC_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    data, indices, indptr, n_rows, n_cols, col_lengths, max_col_nnz = _prepare_spmv_csc_matrix(
        data, indices, indptr, shape)
    block_nnz_use = int(block_nnz)
    if block_nnz_use <= 0:
        raise ValueError("block_nnz must be positive")
    if max_segments is None:
        max_segments_use = max((max_col_nnz + block_nnz_use - 1) // block_nnz_use, 1)
        while max_segments_use > 2048 and block_nnz_use < 65536:
            block_nnz_use *= 2
            max_segments_use = max((max_col_nnz + block_nnz_use - 1) // block_nnz_use, 1)
    else:
        max_segments_use = max(1, int(max_segments))
    # the given arrays are A^T's CSR: op(A) = A^T runs them as they are,
    # op(A) = A = (A^T)^T transposes them once here
    csr_op = "non" if _spmv_csc_op_transposes(op_code) else "trans"
    csr = prepare_spmv_csr(data, indices, indptr, (n_cols, n_rows), op=csr_op,
                           index_fallback_policy=index_fallback_policy)
    return PreparedCscSpmv(data=data, kernel_indices=indices, kernel_indptr=indptr, shape=shape, n_rows=n_rows,
                           n_cols=n_cols, block_nnz=block_nnz_use, max_segments=max_segments_use,
                           max_col_nnz=max_col_nnz, col_lengths=col_lengths, op=op_code,
                           index_fallback_policy=index_fallback_policy, csr=csr)


def _validate_spmv_csc_x(x, prepared, op_code):
    """spmv_csc.py:349-363."""
    if x is None or not torch.is_tensor(x):
        raise TypeError("x must be a torch.Tensor")
    if x.ndim != 1:
        raise ValueError("x must be a 1D tensor")
    if not x.is_cuda:
        raise ValueError("x must be a CUDA tensor")
    if x.dtype != prepared.data.dtype:
        raise TypeError("x dtype must match sparse matrix dtype")
    expected = prepared.n_rows if _spmv_csc_op_transposes(op_code) else prepared.n_cols
    if x.numel() != expected:
        raise ValueError(f"x length must be {expected}, got {x.numel()}")
    if x.device != prepared.data.device:
        raise ValueError("x must be on the same device as sparse matrix data")
    return x.contiguous()


def flagsparse_spmv_csc(data=None, indices=None, indptr=None, x=None, shape=None, block_nnz=256, max_segments=None,
                        out=None, return_time=False, return_meta=False, prepared=None, transpose=None, op=None,
                        index_fallback_policy="auto"):
    """CSC SpMV, y = op(A) x (spmv_csc.py:487-600)."""
    op_explicit = op is not None
    op_code = _normalize_spmv_csc_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != _spmv_csc_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape)):
            raise ValueError("data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmv_csc(data, indices, indptr, shape, block_nnz=block_nnz, max_segments=max_segments,
                                    op=op_code, index_fallback_policy=index_fallback_policy)
    else:
        if not isinstance(prepared, PreparedCscSpmv):
            raise TypeError("prepared must be a PreparedCscSpmv instance")
        if op_explicit and op_code != prepared.op:
            raise ValueError(f"op={_spmv_csc_op_to_name(op_code)} does not match "
                             f"prepared.op={_spmv_csc_op_to_name(prepared.op)}")
        if not op_explicit and transpose is not None and bool(transpose) != prepared.transpose:
            raise ValueError(f"transpose={bool(transpose)} does not match prepared.transpose={prepared.transpose}")
        if not op_explicit:
            op_code = prepared.op
    x = _validate_spmv_csc_x(x, prepared, op_code)
    out_len = prepared.n_cols if _spmv_csc_op_transposes(op_code) else prepared.n_rows
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as the result")
        if tuple(out.shape) != (out_len,) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    do_timing = bool(return_time or return_meta)
    if do_timing:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    y = _run_spmv_prepared_with_fallback(prepared._csr, x)
    if do_timing:
        torch.cuda.synchronize()
        compute_ms = (time.perf_counter() - t0) * 1000.0
        op_total_ms = compute_ms
    else:
        compute_ms = op_total_ms = None
    if out is not None:
        out.copy_(y)
        y = out
    if return_meta:
        meta = {
            "op": _spmv_csc_op_to_name(op_code),
            "symbolic_ms": 0.0 if do_timing else None,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
            "index_fallback_applied": prepared.index_fallback_applied,
            "index_fallback_reason": prepared.index_fallback_reason,
        }
        if return_time:
            return y, op_total_ms, meta
        return y, meta
    if return_time:
        return y, op_total_ms
    return y
