"""COO SpMV: through a GPU COO -> CSR conversion (reference:
spmv_csr.py:948-1091, ``coo_to_csr_for_spmv``, ``prepare_spmv_coo_tocsr``,
``flagsparse_spmv_coo_tocsr``), and the native COO entry points
(spmv_coo.py: ``prepare_spmv_coo``, ``PreparedCoo``, ``flagsparse_spmv_coo``)
described at the end of this file.

Same names, arguments, validation order and messages, and the same CSR it
hands to the CSR SpMV: nonzeros ordered by (row, col), int32 column indices,
int64 row pointer, duplicates kept.  The conversion is a stable radix sort in
the HIP library (``fsp_coo_to_csr``), so duplicates keep their input order
(the reference's ``torch.argsort`` leaves their order unspecified).
``assume_sorted`` is accepted; the sort always runs, and a sorted input comes
back unchanged.  Row ids outside [0, n_rows) raise IndexError (the reference
fails inside ``bincount``); out-of-range columns raise the CSR validator's
IndexError.
"""

import ctypes
import time

import torch

from . import _lib
from ._csc import gather_values
from .spmv_csr import (
    SUPPORTED_SPMV_VALUE_DTYPES,
    _hip_spmv_csr_impl_prepared,
    _spmv_dtype_error_message,
    flagsparse_spmv_csr,
    prepare_spmv_csr,
)


def coo_to_csr_for_spmv(data, row, col, shape, assume_sorted=False):
    """COO -> CSR triple (data, csr_col_indices int32, indptr int64) on the GPU."""
    del assume_sorted  # the stable sort returns sorted input unchanged
    if row.numel() != col.numel() or data.numel() != row.numel():
        raise ValueError("data, row, col must have the same length")
    if not all(t.is_cuda and t.device == data.device for t in (data, row, col)):
        raise ValueError("data, row, col must all be CUDA tensors on one device")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    device = data.device
    nnz = int(row.numel())
    if nnz == 0:
        indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=device)
        return data, col.to(torch.int32), indptr
    index_dtype = torch.int64 if torch.int64 in (row.dtype, col.dtype) else torch.int32
    row_i = row.to(index_dtype).contiguous()
    col_i = col.to(index_dtype).contiguous()
    perm_dtype = torch.int32 if nnz <= 2**31 - 1 else torch.int64
    indptr = torch.empty(n_rows + 1, dtype=torch.int64, device=device)
    indices = torch.empty(nnz, dtype=torch.int32 if n_cols <= 2**31 else torch.int64, device=device)
    perm = torch.empty(nnz, dtype=perm_dtype, device=device)
    err = torch.empty(1, dtype=torch.int32, device=device)
    lib = _lib.load()
    ws_bytes = ctypes.c_int64(0)
    _lib.check(lib.fsp_coo_to_csr_workspace(n_rows, n_cols, nnz, _lib.index_code(perm_dtype), ctypes.byref(ws_bytes)),
               "fsp_coo_to_csr_workspace")
    ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=device)
    _lib.check(
        lib.fsp_coo_to_csr(_lib.index_code(index_dtype), row_i.data_ptr(), col_i.data_ptr(), n_rows, n_cols, nnz,
                           _lib.index_code(torch.int64), indptr.data_ptr(), _lib.index_code(indices.dtype),
                           indices.data_ptr(), _lib.index_code(perm_dtype), perm.data_ptr(), err.data_ptr(),
                           ws.data_ptr(), ws.numel(), _lib.stream_ptr(device)),
        "fsp_coo_to_csr",
    )
    bits = int(err.item())  # conversion time only
    if bits & 1:
        raise IndexError("row indices out of range for n_rows")
    if bits & 2:
        raise IndexError("indices out of range for n_cols")
    return gather_values(data.contiguous(), perm), indices, indptr


def prepare_spmv_coo_tocsr(data, row, col, shape, block_nnz=256, max_segments=None, assume_sorted=False):
    """One-time COO -> CSR + row-bin plan; use with
    ``flagsparse_spmv_coo_tocsr(..., prepared=p)`` (spmv_csr.py:984-1013)."""
    if not all(torch.is_tensor(t) for t in (data, row, col)):
        raise TypeError("data, row, col must all be torch.Tensor")
    if not all(t.is_cuda for t in (data, row, col)):
        raise ValueError("data, row, col must all be CUDA tensors")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, col must all be 1D tensors")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if row.numel() != col.numel() or data.numel() != row.numel():
        raise ValueError("data, row, col must have the same length")
    data_s, indices, indptr = coo_to_csr_for_spmv(data, row, col, shape, assume_sorted=assume_sorted)
    return prepare_spmv_csr(data_s, indices, indptr, shape, block_nnz=block_nnz, max_segments=max_segments)


def flagsparse_spmv_coo_tocsr(data=None, row=None, col=None, x=None, shape=None, block_nnz=256, max_segments=None,
                              out=None, return_time=False, use_opt=True, prepared=None, assume_sorted=False):
    """COO SpMV via CSR conversion: y = A @ x (spmv_csr.py:1016-1091)."""
    if prepared is not None:
        if x is None:
            raise TypeError("x is required")
        if shape is None:
            shape = prepared.shape
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
        return flagsparse_spmv_csr(x=x, shape=shape, block_nnz=block_nnz, max_segments=max_segments, out=out,
                                   return_time=return_time, use_opt=use_opt, prepared=prepared)
    if not all(torch.is_tensor(t) for t in (data, row, col, x)):
        raise TypeError("data, row, col, x must all be torch.Tensor")
    if not all(t.is_cuda for t in (data, row, col, x)):
        raise ValueError("data, row, col, x must all be CUDA tensors")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1 or x.ndim != 1:
        raise ValueError("data, row, col, x must all be 1D tensors")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if x.dtype != data.dtype:
        raise TypeError("x dtype must match data dtype")
    data_s, indices, indptr = coo_to_csr_for_spmv(data, row, col, shape, assume_sorted=assume_sorted)
    return flagsparse_spmv_csr(data_s, indices, indptr, x, shape, block_nnz=block_nnz, max_segments=max_segments,
                               out=out, return_time=return_time, use_opt=use_opt)


# ---------------------------------------------------------------- native COO SpMV
# flagsparse_spmv_coo / prepare_spmv_coo / PreparedCoo (reference spmv_coo.py):
# y = op(A) x for COO input with no indptr in the interface.  The reference
# lex-sorts (row, col), stores row-run offsets and reduces each run in one
# program (or scatters with atomics when sort_by_row=False).  Here the same
# (row, col) order comes from the stable COO -> CSR conversion above, once for
# A and once for A^T at prepare time, and every op runs the CSR row SpMV
# (fsp_spmv_csr) over one of the two: no atomics, deterministic, duplicates
# summed.  sort_by_row=False keeps the given order in the data / row / col
# slots; the kernels still run over the sorted CSR.

SUPPORTED_SPMV_COO_VALUE_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)
SPMV_COO_OP_NON = 0
SPMV_COO_OP_TRANS = 1
SPMV_COO_OP_CONJ_TRANS = 2
SPMV_COO_OP_NAMES = {SPMV_COO_OP_NON: "non", SPMV_COO_OP_TRANS: "trans", SPMV_COO_OP_CONJ_TRANS: "conj"}
_SPMV_COO_OP_NAME_TO_CODE = {name: code for code, name in SPMV_COO_OP_NAMES.items()}
_INT32_MAX = 2**31 - 1


def _spmv_coo_dtype_error_message():
    return "COO SpMV supports float32, float64, complex64, and complex128"


def _normalize_spmv_coo_op(op=None, transpose=False):
    """spmv_coo.py:54-68."""
    if op is None:
        return SPMV_COO_OP_TRANS if bool(transpose) else SPMV_COO_OP_NON
    if isinstance(op, str):
        token = op.strip().lower()
        if token not in _SPMV_COO_OP_NAME_TO_CODE:
            raise ValueError("op must be one of: 0=non, 1=trans, 2=conj")
        return _SPMV_COO_OP_NAME_TO_CODE[token]
    try:
        code = int(op)
    except (TypeError, ValueError) as exc:
        raise ValueError("op must be one of: 0=non, 1=trans, 2=conj") from exc
    if code not in SPMV_COO_OP_NAMES:
        raise ValueError("op must be one of: 0=non, 1=trans, 2=conj")
    return code


def _spmv_coo_op_transposes(op):
    return _normalize_spmv_coo_op(op) in (SPMV_COO_OP_TRANS, SPMV_COO_OP_CONJ_TRANS)


def _normalize_spmv_coo_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


class PreparedCoo:
    """Prepared COO structure that serves non / trans / conj at run time
    (spmv_coo.py:90-148); ``_csr_non`` / ``_csr_trans`` are the CSR SpMV
    preparations of A and A^T the kernels run."""

    __slots__ = (
        "data_non", "row_non", "col_non", "seg_starts_non", "data_trans", "row_trans", "col_trans",
        "seg_starts_trans", "shape", "n_rows", "n_cols", "nnz", "sort_by_row", "op", "transpose",
        "index_fallback_policy", "index_fallback_applied", "index_fallback_reason", "_csr_non", "_csr_trans",
    )

    def __init__(self, data_non, row_non, col_non, shape, seg_starts_non=None, data_trans=None, row_trans=None,
                 col_trans=None, seg_starts_trans=None, sort_by_row=True, transpose=False, op=None,
                 index_fallback_policy="auto", index_fallback_applied=False, index_fallback_reason=None,
                 csr_non=None, csr_trans=None):
        self.data_non = data_non
        self.row_non = row_non
        self.col_non = col_non
        self.seg_starts_non = seg_starts_non
        self.data_trans = data_trans if data_trans is not None else data_non
        self.row_trans = row_trans if row_trans is not None else col_non
        self.col_trans = col_trans if col_trans is not None else row_non
        self.seg_starts_trans = seg_starts_trans
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.nnz = int(data_non.numel())
        self.sort_by_row = bool(sort_by_row)
        self.op = _normalize_spmv_coo_op(op, transpose=transpose)
        self.transpose = _spmv_coo_op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self._csr_non = csr_non
        self._csr_trans = csr_trans


class _CooLaunch:
    """One op's launch view: the COO arrays of op(A), its shape and the CSR
    SpMV preparation the kernel runs (spmv_coo.py:151-202)."""

    __slots__ = ("data", "row", "col", "shape", "n_rows", "n_cols", "nnz", "seg_starts", "op",
                 "index_fallback_policy", "index_fallback_applied", "index_fallback_reason", "csr")

    def __init__(self, data, row, col, shape, seg_starts, op, index_fallback_policy, csr,
                 index_fallback_applied=False, index_fallback_reason=None):
        self.data, self.row, self.col = data, row, col
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.nnz = int(data.numel())
        self.seg_starts = seg_starts
        self.op = op
        self.index_fallback_policy = index_fallback_policy
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self.csr = csr


def _seg_starts(indptr, nnz, index_dtype):
    """Starts of the constant-row runs of a row-sorted COO plus nnz (the
    reference's seg_starts), from its CSR row pointer."""
    if nnz == 0:
        return None
    if nnz > _INT32_MAX:
        index_dtype = torch.int64
    starts = indptr[:-1][indptr[1:] > indptr[:-1]]
    return torch.cat([starts, indptr[-1:]]).to(index_dtype)


def _coo_sorted(data, row, col, shape, index_dtype):
    """(CSR SpMV preparation, sorted data, row, col, seg_starts) of a COO
    matrix; out-of-range ids raise the reference's COO messages."""
    try:
        data_s, indices, indptr = coo_to_csr_for_spmv(data, row, col, shape)
    except IndexError as exc:
        raise IndexError("row indices out of range" if "row" in str(exc) else "col indices out of range") from None
    nnz = int(data_s.numel())
    rows = torch.repeat_interleave(torch.arange(int(shape[0]), device=data.device, dtype=index_dtype),
                                   indptr[1:] - indptr[:-1], output_size=nnz)
    csr = prepare_spmv_csr(data_s, indices, indptr, shape)
    return csr, data_s, rows, indices.to(index_dtype), _seg_starts(indptr, nnz, index_dtype)


def prepare_spmv_coo(data, row, col, shape, sort_by_row=True, transpose=False, op=None,
                     index_fallback_policy="auto"):
    """Validate and sort a COO matrix for both orientations (spmv_coo.py:407-491)."""
    index_fallback_policy = _normalize_spmv_coo_index_fallback_policy(index_fallback_policy)
    op_code = _normalize_spmv_coo_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_COO_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    shape = (int(shape[0]), int(shape[1]))
    if not all(torch.is_tensor(t) for t in (data, row, col)):
        raise TypeError("data, row, col must all be torch.Tensor")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, col must be 1D")
    if not all(t.is_cuda for t in (data, row, col)):
        raise ValueError("data, row, col must be CUDA tensors")
    if data.numel() != row.numel() or data.numel() != col.numel():
        raise ValueError("data, row, col must have the same length")
    if data.dtype not in SUPPORTED_SPMV_COO_VALUE_DTYPES:
        raise TypeError(_spmv_coo_dtype_error_message())
    if row.dtype not in (torch.int32, torch.int64):
        raise TypeError("row dtype must be torch.int32 or torch.int64")
    if col.dtype not in (torch.int32, torch.int64):
        raise TypeError("col dtype must be torch.int32 or torch.int64")
    index_dtype = torch.promote_types(row.dtype, col.dtype)
    csr_non, d_n, r_n, c_n, s_n = _coo_sorted(data, row, col, shape, index_dtype)
    csr_trans, d_t, r_t, c_t, s_t = _coo_sorted(data, col, row, (shape[1], shape[0]), index_dtype)
    if not sort_by_row:  # the slots keep the caller's order
        d_n, r_n, c_n, s_n = data.contiguous(), row.to(index_dtype).contiguous(), col.to(index_dtype).contiguous(), None
        d_t, r_t, c_t, s_t = d_n, c_n, r_n, None
    return PreparedCoo(d_n, r_n, c_n, shape, seg_starts_non=s_n, data_trans=d_t, row_trans=r_t, col_trans=c_t,
                       seg_starts_trans=s_t, sort_by_row=sort_by_row, transpose=transpose, op=op_code,
                       index_fallback_policy=index_fallback_policy, csr_non=csr_non, csr_trans=csr_trans)


def _resolve_spmv_coo_launch(prepared, op):
    """spmv_coo.py:542-571 (real values: conj = trans)."""
    op_code = _normalize_spmv_coo_op(op)
    if op_code == SPMV_COO_OP_NON:
        return _CooLaunch(prepared.data_non, prepared.row_non, prepared.col_non, prepared.shape,
                          prepared.seg_starts_non, op_code, prepared.index_fallback_policy, prepared._csr_non,
                          prepared.index_fallback_applied, prepared.index_fallback_reason)
    return _CooLaunch(prepared.data_trans, prepared.row_trans, prepared.col_trans,
                      (prepared.n_cols, prepared.n_rows), prepared.seg_starts_trans, op_code,
                      prepared.index_fallback_policy, prepared._csr_trans, prepared.index_fallback_applied,
                      prepared.index_fallback_reason)


def _validate_x_coo(x, launch):
    """spmv_coo.py:574-587."""
    if x is None or not torch.is_tensor(x):
        raise TypeError("x must be a torch.Tensor")
    if x.ndim != 1:
        raise ValueError("x must be a 1D tensor")
    if not x.is_cuda:
        raise ValueError("x must be a CUDA tensor")
    if x.dtype != launch.data.dtype:
        raise TypeError("x dtype must match sparse matrix dtype")
    if x.numel() != launch.n_cols:
        raise ValueError(f"x length must be n_cols={launch.n_cols}, got {x.numel()}")
    if x.device != launch.data.device:
        raise ValueError("x must be on the same device as sparse matrix data")
    return x.contiguous()


def _hip_spmv_coo_kernel(launch, x, block_size, num_warps, block_inner):
    """y = op(A) x: the CSR row SpMV over the launch's sorted CSR (replaces
    _triton_spmv_coo_kernel, spmv_coo.py:590-657; the Triton tile arguments
    have no counterpart)."""
    del block_size, num_warps, block_inner
    return _hip_spmv_csr_impl_prepared(launch.csr, x)


def _spmv_coo_uses_int64_indices(launch):
    return (launch.row.dtype == torch.int64 or launch.col.dtype == torch.int64
            or (launch.seg_starts is not None and launch.seg_starts.dtype == torch.int64))


def _spmv_coo_int32_fallback_blocker(launch):
    """spmv_coo.py:670-686."""
    for name, tensor in (("row", launch.row), ("col", launch.col)):
        if tensor.numel() > 0:
            lo, hi = int(tensor.min().item()), int(tensor.max().item())
            if lo < 0 or hi > _INT32_MAX:
                return f"{name} index range [{lo}, {hi}] cannot fit int32 for shape={launch.shape}"
    if launch.seg_starts is not None and launch.seg_starts.numel() > 0:
        max_offset = int(launch.seg_starts[-1].item())
        if max_offset > _INT32_MAX:
            return f"COO nnz offset {max_offset} cannot fit int32 for shape={launch.shape}"
    if launch.n_rows > _INT32_MAX or launch.n_cols > _INT32_MAX:
        return f"shape {launch.shape} cannot fit int32 row/col metadata"
    return None


def _spmv_coo_prepared_with_int32_indices(launch, reason):
    """spmv_coo.py:689-710: the launch over int32 indices (its CSR too)."""
    blocker = _spmv_coo_int32_fallback_blocker(launch)
    if blocker is not None:
        raise RuntimeError(f"native int64 COO SpMV failed and int32 fallback is unsafe: {blocker}")
    c = launch.csr
    csr32 = prepare_spmv_csr(c.data, c.kernel_indices.to(torch.int32), c.kernel_indptr.to(torch.int32), c.shape)
    seg = None if launch.seg_starts is None else launch.seg_starts.to(torch.int32).contiguous()
    return _CooLaunch(launch.data, launch.row.to(torch.int32).contiguous(), launch.col.to(torch.int32).contiguous(),
                      launch.shape, seg, launch.op, launch.index_fallback_policy, csr32, True, str(reason))


def _run_spmv_coo_prepared_with_fallback(launch, x, block_size, num_warps, block_inner):
    """spmv_coo.py:713-737."""
    try:
        return _hip_spmv_coo_kernel(launch, x, block_size, num_warps, block_inner)
    except Exception as exc:
        if launch.index_fallback_policy != "auto" or not _spmv_coo_uses_int64_indices(launch):
            raise
        fallback = _spmv_coo_prepared_with_int32_indices(launch, exc)
        return _hip_spmv_coo_kernel(fallback, x, block_size, num_warps, block_inner)


def flagsparse_spmv_coo(data=None, row=None, col=None, x=None, shape=None, out=None, return_time=False,
                        prepared=None, sort_by_row=True, block_size=256, num_warps=4, block_inner=128,
                        transpose=None, op=None, index_fallback_policy="auto"):
    """COO SpMV y = op(A) x with no indptr in the interface (spmv_coo.py:740-829)."""
    transpose_flag = False if transpose is None else bool(transpose)
    op_explicit = op is not None
    op_code = _normalize_spmv_coo_op(op, transpose=transpose_flag)
    if op_explicit and transpose is not None and bool(transpose) != _spmv_coo_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if prepared is None:
        if any(a is None for a in (data, row, col, x, shape)):
            raise ValueError("data, row, col, x, shape required when prepared is None")
        prepared = prepare_spmv_coo(data, row, col, shape, sort_by_row=sort_by_row, transpose=transpose_flag, op=op,
                                    index_fallback_policy=index_fallback_policy)
    else:
        if not isinstance(prepared, PreparedCoo):
            raise TypeError("prepared must be a PreparedCoo instance")
        if x is None:
            raise TypeError("x is required when prepared is set")
        if shape is None:
            shape = prepared.shape
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
        if not op_explicit and transpose is None:
            op_code = prepared.op
        elif not op_explicit and transpose is not None:
            op_code = _normalize_spmv_coo_op(None, transpose=transpose_flag)
    launch = _resolve_spmv_coo_launch(prepared, op_code)
    x = _validate_x_coo(x, launch)
    if num_warps not in (1, 2, 4, 8, 16, 32):
        raise ValueError("num_warps must be a power of 2 in [1, 32]")
    if block_inner <= 0 or (block_inner & (block_inner - 1)) != 0:
        raise ValueError("block_inner must be a positive power of 2")
    if out is not None and (tuple(out.shape) != (launch.n_rows,) or out.dtype != launch.data.dtype
                            or not out.is_cuda or out.device != launch.data.device):
        raise ValueError("out shape/dtype must match result")
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    y = _run_spmv_coo_prepared_with_fallback(launch, x, block_size, num_warps, block_inner)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None:
        out.copy_(y)
        y = out
    if return_time:
        return y, elapsed_ms
    return y
