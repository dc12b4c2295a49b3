"""CSR SpMV y = op(A) x on MI355X (HIP kernels, no Triton).

Drop-in for the reference's spmv_csr operator (src/flagsparse/
sparse_operations/spmv_csr.py): same entry points, argument conventions,
validation order, exception types and messages, prepared-object semantics and
int64 -> int32 index fallback.  The compute runs in ``fsp_spmv_csr`` (row-bin
plan built once per prepared matrix, one fused launch per call, f64
accumulation for f32 inputs as the reference does at spmv_csr.py:138-143).
"""

import torch

from . import _lib, _plan
from ._common import (
    OP_CONJ_TRANS,
    OP_NAMES,
    OP_NON,
    OP_TRANS,
    SUPPORTED_INDEX_DTYPES,
    _INDEX_LIMIT_INT32,
    _Timer,
    _is_complex_dtype,
    normalize_op,
    op_to_name,
    op_transposes,
    require_kernel_dtype,
)
from ._csc import transpose_csr

SUPPORTED_SPMV_VALUE_DTYPES = (
    torch.float16,
    torch.bfloat16,
    torch.float32,
    torch.float64,
    torch.complex64,
    torch.complex128,
)

SPMV_OP_NON = OP_NON
SPMV_OP_TRANS = OP_TRANS
SPMV_OP_CONJ_TRANS = OP_CONJ_TRANS
SPMV_OP_NAMES = dict(OP_NAMES)

_normalize_spmv_op = normalize_op
_spmv_op_to_name = op_to_name
_spmv_op_transposes = op_transposes


class PreparedCsrSpmv:
    """Cached CSR metadata + HIP work plan for repeated SpMV calls
    (reference: spmv_csr.py:69-144)."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "block_nnz",
        "max_segments",
        "opt_max_segments",
        "row_lengths",
        "max_row_nnz",
        "opt_buckets",
        "supports_opt",
        "transpose",
        "op",
        "index_fallback_policy",
        "index_fallback_applied",
        "index_fallback_reason",
        "_plan",
    )

    def __init__(
        self,
        data,
        kernel_indices,
        kernel_indptr,
        shape,
        n_rows,
        n_cols,
        block_nnz,
        max_segments,
        max_row_nnz,
        opt_buckets=None,
        opt_max_segments=None,
        row_lengths=None,
        transpose=False,
        op=None,
        index_fallback_policy="auto",
        index_fallback_applied=False,
        index_fallback_reason=None,
    ):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = n_rows
        self.n_cols = n_cols
        self.block_nnz = block_nnz
        self.max_segments = max_segments
        self.opt_max_segments = opt_max_segments
        if row_lengths is None:
            row_lengths = kernel_indptr[1:] - kernel_indptr[:-1]
        self.row_lengths = row_lengths
        self.max_row_nnz = max_row_nnz
        self.opt_buckets = [] if opt_buckets is None else opt_buckets
        self.supports_opt = (
            data.dtype in (torch.float32, torch.float64) and kernel_indices.dtype == torch.int32
        )
        self.op = _normalize_spmv_op(op, transpose=transpose)
        self.transpose = _spmv_op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self._plan = None

    def plan(self, config=None):
        """The row-bin plan of this matrix (built on first use, then cached)."""
        if self._plan is None or (config is not None and self._plan.config != config):
            self._plan = _plan.build_plan(self.kernel_indptr, self.n_rows, config or _plan.spmv_config())
        return self._plan


def _normalize_spmv_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


def _spmv_dtype_error_message():
    return "data dtype must be one of: " + ", ".join(
        str(dtype).replace("torch.", "") for dtype in SUPPORTED_SPMV_VALUE_DTYPES
    )


def _transpose_csr_for_spmv(data, indices, indptr, shape):
    """Transposed CSR via the HIP stable CSC permutation (reference
    spmv_csr.py:498-536 uses a stable torch.argsort)."""
    return transpose_csr(data, indices, indptr, shape)


def _prepare_spmv_csr_matrix(data, indices, indptr, shape, index_fallback_policy="auto"):
    """Validation contract of spmv_csr.py:539-595 (same order and messages)."""
    return _inspect_spmv_csr_matrix(data, indices, indptr, shape, index_fallback_policy)[0]


def _inspect_spmv_csr_matrix(data, indices, indptr, shape, index_fallback_policy="auto", plan_config=None):
    """_prepare_spmv_csr_matrix plus, with `plan_config`, the row-bin plan,
    from one device inspection and ONE host readback (the reference makes a
    host sync per check).  Returns (validated tuple, plan or None)."""
    _normalize_spmv_index_fallback_policy(index_fallback_policy)
    if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
        raise TypeError("data, indices, indptr must all be torch.Tensor")
    if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("data, indices, indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if data.numel() != indices.numel():
        raise ValueError("data and indices must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if indices.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")

    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    nnz = data.numel()
    stats, plan = _plan.inspect_csr(indptr, indices, n_rows, nnz, plan_config)
    if stats["indptr_first"] != 0:
        raise ValueError("indptr must start at zero")
    if stats["indptr_last"] != nnz:
        raise ValueError("indptr[-1] must equal nnz")
    if stats["n_decreasing"]:
        raise ValueError("indptr must be non-decreasing")
    if nnz > 0 and (stats["min_col"] < 0 or stats["max_col"] >= n_cols):
        raise IndexError("indices out of range for n_cols")
    row_lengths = indptr[1:] - indptr[:-1]
    return (data, indices, indptr, n_rows, n_cols, row_lengths, stats["max_row"]), plan


def _validate_spmv_x(x, prepared):
    """spmv_csr.py:598-611."""
    if x is None or not torch.is_tensor(x):
        raise TypeError("x must be a torch.Tensor")
    if x.ndim != 1:
        raise ValueError("x must be a 1D tensor")
    if not x.is_cuda:
        raise ValueError("x must be a CUDA tensor")
    if x.dtype != prepared.data.dtype:
        raise TypeError("x dtype must match sparse matrix dtype")
    if x.numel() != prepared.n_cols:
        raise ValueError(f"x length must be n_cols={prepared.n_cols}, got {x.numel()}")
    if x.device != prepared.data.device:
        raise ValueError("x must be on the same device as sparse matrix data")
    return x.contiguous()


def prepare_spmv_csr(
    data,
    indices,
    indptr,
    shape,
    block_nnz=256,
    max_segments=None,
    transpose=False,
    op=None,
    index_fallback_policy="auto",
):
    """One-time validation, transpose (trans/conj) and row-bin plan
    (reference: spmv_csr.py:614-687)."""
    index_fallback_policy = _normalize_spmv_index_fallback_policy(index_fallback_policy)
    op_code = _normalize_spmv_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    transpose = _spmv_op_transposes(op_code)
    if transpose:
        data, indices, indptr, *_ = _prepare_spmv_csr_matrix(
            data, indices, indptr, shape, index_fallback_policy=index_fallback_policy
        )
        if op_code == SPMV_OP_CONJ_TRANS and _is_complex_dtype(data.dtype):
            data = data.conj().resolve_conj()
        data, indices, indptr, shape = _transpose_csr_for_spmv(data, indices, indptr, shape)
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz), plan = _inspect_spmv_csr_matrix(
        data, indices, indptr, shape, index_fallback_policy=index_fallback_policy, plan_config=_plan.spmv_config()
    )
    # block_nnz / max_segments keep the caller's launch hints in the prepared
    # object's public slots; the HIP kernel takes its split from the plan
    prepared = PreparedCsrSpmv(
        data=data,
        kernel_indices=kernel_indices,
        kernel_indptr=kernel_indptr,
        shape=shape,
        n_rows=n_rows,
        n_cols=n_cols,
        block_nnz=block_nnz,
        max_segments=max_segments,
        opt_max_segments=max_segments,
        row_lengths=row_lengths,
        max_row_nnz=max_row_nnz,
        opt_buckets=None,
        transpose=transpose,
        op=op_code,
        index_fallback_policy=index_fallback_policy,
    )
    prepared._plan = plan  # built with the validation readback
    return prepared


def _hip_spmv_csr_impl_prepared(prepared, x, out=None):
    """Launch fsp_spmv_csr over the prepared plan; returns y (or `out`, a
    contiguous validated result tensor the kernel writes directly).
    Replaces _triton_spmv_csr_impl_prepared (spmv_csr.py:702-746)."""
    dtype = prepared.data.dtype
    require_kernel_dtype(dtype)
    device = prepared.data.device
    y = out if out is not None else torch.empty(prepared.n_rows, dtype=dtype, device=device)
    if prepared.n_rows == 0:
        return y
    plan = prepared.plan()
    acc = torch.float64 if dtype in (torch.float32, torch.float64) else torch.float32
    partials = plan.partials(1, acc, device)
    nnz = int(prepared.kernel_indices.numel())
    variant = _plan.spmv_default_variant(int(x.shape[0]), x.element_size(), nnz,
                                         _plan.chunk_aligned(prepared.kernel_indices, prepared.data))
    st = plan.struct(partials, variant=variant)
    _lib.check(
        _lib.load().fsp_spmv_csr(
            _lib.value_code(dtype),
            _lib.index_code(prepared.kernel_indices.dtype),
            _lib.index_code(prepared.kernel_indptr.dtype),
            _lib.ptr(prepared.kernel_indptr),
            _lib.ptr(prepared.kernel_indices),
            _lib.ptr(prepared.data),
            0,
            None,
            _lib.ptr(x),
            y.data_ptr(),
            prepared.n_rows,
            nnz,
            int(x.shape[0]),
            st,
            _lib.stream_ptr(device),
        ),
        "fsp_spmv_csr",
    )
    return y


def _spmv_uses_int64_indices(prepared):
    return prepared.kernel_indices.dtype == torch.int64 or prepared.kernel_indptr.dtype == torch.int64


def _spmv_int32_fallback_blocker(prepared):
    """spmv_csr.py:756-786."""
    if prepared.kernel_indices.dtype == torch.int64 and prepared.kernel_indices.numel() > 0:
        min_index = int(prepared.kernel_indices.min().item())
        max_index = int(prepared.kernel_indices.max().item())
        if min_index < 0 or max_index > _INDEX_LIMIT_INT32:
            return f"column index range [{min_index}, {max_index}] cannot fit int32 for shape={prepared.shape}"
    if prepared.kernel_indptr.dtype == torch.int64 and prepared.kernel_indptr.numel() > 0:
        max_offset = int(prepared.kernel_indptr[-1].item())
        if max_offset > _INDEX_LIMIT_INT32:
            return f"CSR nnz offset {max_offset} cannot fit int32 for shape={prepared.shape}"
    if prepared.n_rows > _INDEX_LIMIT_INT32:
        return f"row count {prepared.n_rows} cannot fit int32 row metadata"
    return None


def _spmv_prepared_with_int32_indices(prepared, reason):
    blocker = _spmv_int32_fallback_blocker(prepared)
    if blocker is not None:
        raise RuntimeError(f"native int64 CSR SpMV failed and int32 fallback is unsafe: {blocker}")
    kernel_indices = prepared.kernel_indices.to(torch.int32)
    kernel_indptr = prepared.kernel_indptr.to(torch.int32)
    row_lengths = kernel_indptr[1:] - kernel_indptr[:-1]
    max_row_nnz = int(row_lengths.max().item()) if prepared.n_rows > 0 else 0
    return PreparedCsrSpmv(
        data=prepared.data,
        kernel_indices=kernel_indices,
        kernel_indptr=kernel_indptr,
        shape=prepared.shape,
        n_rows=prepared.n_rows,
        n_cols=prepared.n_cols,
        block_nnz=prepared.block_nnz,
        max_segments=prepared.max_segments,
        opt_max_segments=prepared.opt_max_segments,
        row_lengths=row_lengths,
        max_row_nnz=max_row_nnz,
        opt_buckets=None,
        transpose=prepared.transpose,
        op=prepared.op,
        index_fallback_policy=prepared.index_fallback_policy,
        index_fallback_applied=True,
        index_fallback_reason=str(reason),
    )


def _run_spmv_prepared(prepared, x, use_opt=False, opt_buckets=None, out=None):
    # The HIP path is always the row-binned one; use_opt keeps its meaning of
    # "report the bucket count" (there is no separate unbucketed kernel).
    return _hip_spmv_csr_impl_prepared(prepared, x, out=out)


def _run_spmv_prepared_with_fallback(prepared, x, use_opt=False, opt_buckets=None, out=None):
    """spmv_csr.py:817-834."""
    try:
        return _run_spmv_prepared(prepared, x, use_opt=use_opt, opt_buckets=opt_buckets, out=out)
    except Exception as exc:
        if prepared.index_fallback_policy != "auto" or not _spmv_uses_int64_indices(prepared):
            raise
        fallback_prepared = _spmv_prepared_with_int32_indices(prepared, exc)
        return _run_spmv_prepared(fallback_prepared, x, use_opt=use_opt, out=out)


def flagsparse_spmv_csr(
    data=None,
    indices=None,
    indptr=None,
    x=None,
    shape=None,
    block_nnz=256,
    max_segments=None,
    out=None,
    return_time=False,
    return_meta=False,
    use_opt=False,
    prepared=None,
    transpose=None,
    op=None,
    index_fallback_policy="auto",
):
    """CSR SpMV on MI355X.

    data, indices, indptr: CSR arrays; x: dense vector; shape: (n_rows, n_cols).
    prepared: cached CSR metadata from prepare_spmv_csr for steady-state runs.
    op: 0/'non' for A @ x, 1/'trans' for A.T @ x, 2/'conj' for A.conj().T @ x.
    Reference: spmv_csr.py:837-945.
    """
    op_explicit = op is not None
    op_code = _normalize_spmv_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != _spmv_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape)):
            raise ValueError("data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmv_csr(
            data,
            indices,
            indptr,
            shape,
            block_nnz=block_nnz,
            max_segments=max_segments,
            op=op_code,
            index_fallback_policy=index_fallback_policy,
        )
    else:
        if op_explicit and op_code != prepared.op:
            raise ValueError(
                f"op={_spmv_op_to_name(op_code)} does not match prepared.op={_spmv_op_to_name(prepared.op)}"
            )
        if not op_explicit and transpose is not None and bool(transpose) != prepared.transpose:
            raise ValueError(
                f"transpose={bool(transpose)} does not match prepared.transpose={prepared.transpose}"
            )
    x = _validate_spmv_x(x, prepared)
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as the result")
        if tuple(out.shape) != (prepared.n_rows,) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    timer = _Timer(return_time or return_meta)
    timer.mark()
    bucket_count = 0
    if use_opt and prepared.supports_opt:
        bucket_count = sum(1 for c in prepared.plan().bin_sizes if c)
    symbolic_ms = timer.mark()
    if symbolic_ms is not None and not (use_opt and prepared.supports_opt):
        symbolic_ms = 0.0
    # a contiguous `out` is written by the kernel itself; a strided one gets a copy
    direct = out if out is not None and out.is_contiguous() else None
    y = _run_spmv_prepared_with_fallback(prepared, x, use_opt=use_opt, out=direct)
    compute_ms = timer.mark()
    op_total_ms = None if compute_ms is None else symbolic_ms + compute_ms
    if out is not None:
        if out.data_ptr() != y.data_ptr():
            out.copy_(y)
        y = out
    if return_meta:
        meta = {
            "symbolic_ms": symbolic_ms,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
            "bucket_count": int(bucket_count),
        }
        if return_time:
            return y, op_total_ms, meta
        return y, meta
    if return_time:
        return y, op_total_ms
    return y
