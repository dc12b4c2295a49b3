"""CSR SpMM C = op(A) B on MI355X (HIP kernels, no Triton).

Drop-in for the reference's spmm_csr operator (src/flagsparse/
sparse_operations/spmm_csr.py): ``flagsparse_spmm_csr`` (:3981-4137), the
opt / alg1 prepared API (:2720-3253, :4140-4301), the opt-alg2 prepared API
(spmm_csr_opt_alg2.py:39-76, :916-1241) and the route API
(``SPMM_CSR_ALGORITHMS``, ``prepare_spmm_csr_route``,
``flagsparse_spmm_csr_run``; :619-842, :2452-2717).  Validation order,
exception types and messages follow the reference.  Every route runs the same
HIP kernel family (``fsp_spmm_csr``): row bins + nnz-balanced long-row
segments, f32 accumulation for f16/bf16/f32 (no upcast copies), f64 for f64.
op=trans/conj runs over a stable CSC permutation with values read through it
(no transposed copy of the values).
"""

import math
from dataclasses import dataclass

import torch

from . import _plan, autotune
from ._common import (
    OP_CONJ_TRANS,
    OP_NAMES,
    OP_NON,
    OP_TRANS,
    KERNEL_VALUE_DTYPES,
    SUPPORTED_INDEX_DTYPES,
    SUPPORTED_VALUE_DTYPES,
    _INDEX_LIMIT_INT32,
    _Timer,
    _is_complex_dtype,
    _tolerance_for_dtype,
    error_metrics,
    normalize_op,
    op_to_name,
    op_transposes,
    relative_threshold,
    require_kernel_dtype,
    validation_metrics,
)
from ._csc import transpose_csr
from ._matrix import CsrHandle

SUPPORTED_SPMM_VALUE_DTYPES = SUPPORTED_VALUE_DTYPES

SPMM_OP_NON = OP_NON
SPMM_OP_TRANS = OP_TRANS
SPMM_OP_CONJ_TRANS = OP_CONJ_TRANS
SPMM_OP_NAMES = dict(OP_NAMES)

_normalize_spmm_op = normalize_op
_spmm_op_to_name = op_to_name
_spmm_op_transposes = op_transposes
_spmm_relative_threshold = relative_threshold
_spmm_error_metrics = error_metrics
_spmm_validation_metrics = validation_metrics


# ------------------------------------------------------------------ validation
def _prepare_spmm_csr_matrix(data, indices, indptr, shape):
    """spmm_csr.py:342-404 (same checks and messages).  Unlike the reference,
    index dtypes are kept native: the HIP kernels read int32 and int64
    indices directly, so no narrowed copy is made."""
    return _inspect_spmm_csr_matrix(data, indices, indptr, shape)[0]


def _inspect_spmm_csr_matrix(data, indices, indptr, shape, plan_config=None):
    """_prepare_spmm_csr_matrix plus, with `plan_config`, the row-bin plan,
    from one device inspection and ONE host readback.  Returns (validated
    tuple, plan or None)."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple: (n_rows, n_cols)")
    if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("data, indices, and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if data.numel() != indices.numel():
        raise ValueError("data and indices must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, and indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, and indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMM_VALUE_DTYPES:
        raise TypeError("data dtype must be one of: float16, bfloat16, float32, float64, complex64, complex128")
    if indices.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    nnz = data.numel()
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    stats, plan = _plan.inspect_csr(indptr, indices, n_rows, nnz, plan_config)
    if stats["indptr_first"] != 0:
        raise ValueError("indptr[0] must be 0")
    if stats["indptr_last"] != nnz:
        raise ValueError(f"indptr[-1] must equal nnz={nnz}, got {stats['indptr_last']}")
    if stats["n_decreasing"]:
        raise ValueError("indptr must be nondecreasing")
    if nnz > 0:
        if stats["min_col"] < 0 or stats["max_col"] >= n_cols:
            raise IndexError("indices out of range for n_cols")
        if stats["max_col"] > _INDEX_LIMIT_INT32:
            raise ValueError("column indices exceed the int32 range supported by the SpMM kernel")
    row_lengths = indptr[1:] - indptr[:-1] if n_rows > 0 else indptr.new_empty((0,))
    return (data, indices, indptr, n_rows, n_cols, row_lengths, stats["max_row"]), plan


def _transpose_csr_for_spmm(data, indices, indptr, shape):
    """Transposed CSR via the HIP stable CSC permutation (spmm_csr.py:407-445)."""
    return transpose_csr(data, indices, indptr, shape)


def _materialize_spmm_csr_op(data, indices, indptr, shape, op_code):
    """spmm_csr.py:448-456."""
    if op_code == SPMM_OP_NON:
        return data, indices, indptr, shape
    data_op = data
    if op_code == SPMM_OP_CONJ_TRANS and _is_complex_dtype(data.dtype):
        data_op = data.conj().resolve_conj()
    return _transpose_csr_for_spmm(data_op, indices, indptr, shape)


def _resolve_spmm_alg1_launch_config(n_dense_cols, max_row_nnz, block_n=None, block_nnz=None, max_segments=None):
    """Argument contract of spmm_csr.py:491-526 (positivity and the explicit
    max_segments check).  block_n / block_nnz become tuning hints of the HIP
    plan: lanes along N and nonzeros per lane per trip."""
    if block_n is not None and block_n <= 0 or block_nnz is not None and block_nnz <= 0:
        raise ValueError("block_n and block_nnz must be positive when provided")
    if max_segments is not None and max_segments <= 0:
        raise ValueError("max_segments must be positive when provided")
    bnnz = 32 if block_nnz is None else int(block_nnz)
    required_segments = math.ceil(max_row_nnz / bnnz) if max_row_nnz > 0 else 0
    if max_segments is not None and required_segments > int(max_segments):
        raise ValueError(
            "row nnz requires more CSR segments than the explicit max_segments override allows: "
            f"required {required_segments}, provided {int(max_segments)}"
        )
    return {"block_n": block_n, "block_nnz": block_nnz, "required_segments": int(required_segments)}


def _hint_lanes(block_n, n_dense, elem_bytes):
    """Lanes along N for an explicit block_n (columns per program)."""
    if block_n is None:
        return None
    vec = max(1, 16 // int(elem_bytes))
    want = max(1, -(-int(block_n) // vec))
    g = 1
    while g < 64 and g < want:
        g *= 2
    return min(g, _plan.lanes_for_dense(n_dense, elem_bytes)) or 1


def _hint_config(lanes, block_nnz):
    if block_nnz is None:
        return None
    unroll = 2 if block_nnz <= 8 else (4 if block_nnz <= 32 else 8)
    return _plan.spmm_config(lanes, unroll=unroll)


def _base_lanes(n_dense, elem_bytes, block_n):
    return _hint_lanes(block_n, n_dense, elem_bytes) or _plan.lanes_for_dense(n_dense, elem_bytes)


def _base_plan_config(n_dense, elem_bytes, block_n, block_nnz):
    """The plan configuration _hip_spmm_csr_impl launches with."""
    lanes = _base_lanes(n_dense, elem_bytes, block_n)
    return _hint_config(lanes, block_nnz) or _plan.spmm_config(lanes)


def _hip_spmm_csr_impl(handle, data, B, transpose=False, out=None, block_n=None, block_nnz=None):
    """Run fsp_spmm_csr (replaces _triton_spmm_csr_impl, spmm_csr.py:3373-3466)."""
    require_kernel_dtype(data.dtype)
    n_dense = int(B.shape[1])
    lanes = _base_lanes(n_dense, data.element_size(), block_n)
    return handle.spmm(data, B, transpose=transpose, out=out, lanes_n=lanes,
                       config=_base_plan_config(n_dense, data.element_size(), block_n, block_nnz))


# ------------------------------------------------------------------ public API
def flagsparse_spmm_csr(
    data,
    indices,
    indptr,
    B,
    shape,
    block_n=None,
    block_nnz=None,
    max_segments=None,
    out=None,
    return_time=False,
    transpose=None,
    op=None,
    return_meta=False,
):
    """CSR SpMM on MI355X.

    op: 0/'non' for A @ B, 1/'trans' for A.T @ B, 2/'conj' for A.conj().T @ B.
    Reference: spmm_csr.py:3981-4137.
    """
    op_explicit = op is not None
    op_code = _normalize_spmm_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != _spmm_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if block_n is not None and block_n <= 0:
        raise ValueError("block_n must be positive when provided")
    if block_nnz is not None and block_nnz <= 0:
        raise ValueError("block_nnz must be positive when provided")
    if max_segments is not None and max_segments <= 0:
        raise ValueError("max_segments must be positive when provided")
    transposes = _spmm_op_transposes(op_code)
    # the row-bin plan of the kernel is built with the validation readback
    # (op non only: trans runs over the CSC and plans that)
    plan_cfg = None
    if not transposes and torch.is_tensor(B) and B.ndim == 2 and data.dtype in KERNEL_VALUE_DTYPES:
        plan_cfg = _base_plan_config(int(B.shape[1]), data.element_size(), block_n, block_nnz)
    (data, indices, indptr, n_rows, n_cols, row_lengths, max_row_nnz), plan = _inspect_spmm_csr_matrix(
        data, indices, indptr, shape, plan_cfg
    )
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    eff_rows, eff_cols = (n_cols, n_rows) if transposes else (n_rows, n_cols)
    if B.shape[0] != eff_cols:
        raise ValueError(f"B.shape[0] must be n_cols={eff_cols}, got {B.shape[0]}")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != data.device:
        raise ValueError("B must be on the same CUDA device as the sparse matrix")
    if B.dtype != data.dtype:
        raise TypeError("B dtype must match data dtype")
    B = B.contiguous()
    n_dense = int(B.shape[1])
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != data.device:
            raise ValueError("out must be on the same CUDA device as the inputs")
        if tuple(out.shape) != (eff_rows, n_dense) or out.dtype != data.dtype:
            raise ValueError("out shape/dtype must match result")
    require_kernel_dtype(data.dtype)
    timer = _Timer(return_time or return_meta)
    timer.mark()
    handle = CsrHandle(indices, indptr, (n_rows, n_cols))
    if plan is not None:
        handle._plans[plan.config] = plan
    values = data
    if transposes:
        handle.csc()  # the transposed structure: one radix sort, values stay in place
        if op_code == SPMM_OP_CONJ_TRANS and _is_complex_dtype(data.dtype):
            values = data.conj().resolve_conj()
    symbolic_ms = timer.mark()
    if symbolic_ms is not None and not transposes:
        symbolic_ms = 0.0
    max_row_nnz_op = max_row_nnz
    if transposes and max_segments is not None and eff_rows > 0:
        # rows of op(A) are the columns of A
        max_row_nnz_op = int(torch.diff(handle.csc()[0]).max().item())
    _resolve_spmm_alg1_launch_config(n_dense, max_row_nnz_op, block_n, block_nnz, max_segments)
    direct_out = out if (out is not None and out.is_contiguous()) else None
    C = _hip_spmm_csr_impl(handle, values, B, transpose=transposes, out=direct_out, block_n=block_n,
                           block_nnz=block_nnz)
    compute_ms = timer.mark()
    op_total_ms = None if compute_ms is None else symbolic_ms + compute_ms
    if out is not None:
        if C.data_ptr() != out.data_ptr():
            out.copy_(C)
        C = out
    if return_meta:
        meta = {
            "symbolic_ms": symbolic_ms,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
            "op": _spmm_op_to_name(op_code),
        }
        if return_time:
            return C, op_total_ms, meta
        return C, meta
    if return_time:
        return C, op_total_ms
    return C


# ---------------------------------------------------------- opt / alg1 API
class PreparedCsrSpmmOpt:
    """Cached CSR metadata for repeated SpMM-opt calls (reference:
    spmm_csr.py:2720-2782).  ``row_buckets`` describes the HIP plan bins."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "row_lengths",
        "max_row_nnz",
        "row_buckets",
        "supports_opt",
        "long_part_rows",
        "long_part_starts",
        "long_part_ends",
        "long_row_ids",
        "long_row_part_ptr",
        "long_row_fallback_only",
        "_handle",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths, max_row_nnz):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self.supports_opt = data.dtype in (torch.float32, torch.float64)
        self._handle = CsrHandle(kernel_indices, kernel_indptr, self.shape)
        self.row_buckets = []
        empty = torch.empty(0, dtype=torch.int64, device=data.device)
        self.long_part_rows = empty
        self.long_part_starts = empty
        self.long_part_ends = empty
        self.long_row_ids = empty
        self.long_row_part_ptr = torch.zeros(1, dtype=torch.int64, device=data.device)
        self.long_row_fallback_only = False

    def _describe(self, plan):
        """Fill the reference's bucket / split fields from a HIP plan."""
        self.row_buckets = [
            {"kind": "rows", "slots": int(l), "max_row_nnz": int(u), "rows": int(c)}
            for l, u, c in zip(plan.config.lanes, plan.config.uppers, plan.bin_counts)
            if c
        ]
        if plan.n_segments:
            self.row_buckets.append({"kind": "split", "rows": plan.n_long_rows, "parts": plan.n_segments})
            self.long_part_rows = plan.seg_row
            self.long_part_starts = plan.seg_begin
            self.long_part_ends = plan.seg_end
            self.long_row_ids = plan.long_rows
            self.long_row_part_ptr = plan.long_seg_ptr


def prepare_spmm_csr_opt_alg1(data, indices, indptr, shape):
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz) = _prepare_spmm_csr_matrix(
        data, indices, indptr, shape
    )
    return PreparedCsrSpmmOpt(data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths, max_row_nnz)


def prepare_spmm_csr_opt(data, indices, indptr, shape):
    return prepare_spmm_csr_opt_alg1(data, indices, indptr, shape)


def prepare_spmm_csr_opt_alg1_preprocess(data, indices, indptr, shape):
    return prepare_spmm_csr_opt_alg1(data, indices, indptr, shape)


def _flagsparse_spmm_csr_opt_alg1_impl(
    data=None,
    indices=None,
    indptr=None,
    B=None,
    shape=None,
    prepared=None,
    out=None,
    return_time=False,
    return_meta=False,
    api_name="flagsparse_spmm_csr_opt_alg1",
):
    """spmm_csr.py:4140-4222 (argument contract and meta keys)."""
    if prepared is not None and not isinstance(prepared, PreparedCsrSpmmOpt):
        raise TypeError("prepared must be a PreparedCsrSpmmOpt instance")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape)):
            raise ValueError("data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmm_csr_opt_alg1(data, indices, indptr, shape)
    elif shape is not None:
        resolved_shape = (int(shape[0]), int(shape[1]))
        if resolved_shape != prepared.shape:
            raise ValueError(f"shape {resolved_shape} does not match prepared.shape {prepared.shape}")
    if B is None:
        raise ValueError("B is required")
    if not prepared.supports_opt:
        raise TypeError(f"{api_name} only supports float32 and float64")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as sparse matrix data")
        if tuple(out.shape) != (prepared.n_rows, int(B.shape[1])) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if B.shape[0] != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    B = B.contiguous()
    timer = _Timer(return_time or return_meta)
    timer.mark()
    n_dense = int(B.shape[1])
    lanes = _plan.lanes_for_dense(n_dense, B.element_size())
    cfg = _plan.spmm_config(lanes)
    plan = prepared._handle.plan(cfg) if prepared.n_rows else None
    if plan is not None:
        prepared._describe(plan)
    symbolic_ms = timer.mark()
    direct_out = out if (out is not None and out.is_contiguous()) else None
    C = prepared._handle.spmm(prepared.data, B, out=direct_out, lanes_n=lanes, config=cfg)
    compute_ms = timer.mark()
    op_total_ms = None if compute_ms is None else symbolic_ms + compute_ms
    if out is not None:
        if C.data_ptr() != out.data_ptr():
            out.copy_(C)
        C = out
    if return_meta:
        meta = {
            "symbolic_ms": symbolic_ms,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
            "long_part_count": int(plan.n_segments) if plan is not None else 0,
            "long_row_count": int(plan.n_long_rows) if plan is not None else 0,
            "bucket_count": int(len(prepared.row_buckets)),
        }
        if return_time:
            return C, op_total_ms, meta
        return C, meta
    if return_time:
        return C, op_total_ms
    return C


def flagsparse_spmm_csr_opt_alg1(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None,
                                 out=None, return_time=False, return_meta=False):
    """CSR SpMM-opt alg1: float32/float64 row-binned path for CSR @ dense."""
    return _flagsparse_spmm_csr_opt_alg1_impl(data, indices, indptr, B, shape, prepared, out, return_time,
                                              return_meta, api_name="flagsparse_spmm_csr_opt_alg1")


def flagsparse_spmm_csr_opt_alg1_preprocess(data=None, indices=None, indptr=None, B=None, shape=None,
                                            prepared=None, out=None, return_time=False, return_meta=False):
    """Compatibility alias for CSR SpMM-opt alg1 with runtime preprocessing."""
    return _flagsparse_spmm_csr_opt_alg1_impl(data, indices, indptr, B, shape, prepared, out, return_time,
                                              return_meta, api_name="flagsparse_spmm_csr_opt_alg1_preprocess")


def flagsparse_spmm_csr_opt(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None, out=None,
                            return_time=False, return_meta=False):
    """Compatibility alias for CSR SpMM-opt alg1."""
    return flagsparse_spmm_csr_opt_alg1(data=data, indices=indices, indptr=indptr, B=B, shape=shape,
                                        prepared=prepared, out=out, return_time=return_time,
                                        return_meta=return_meta)


# ------------------------------------------------------------ opt_alg2 API
SUPPORTED_SPMM_OPT_ALG2_DTYPES = (torch.float32, torch.float64)


class PreparedCsrSpmmOptAlg2:
    """Cached CSR metadata for repeated opt-alg2 SpMM calls (reference:
    spmm_csr_opt_alg2.py:39-76).  opt-alg2 is the device-aware route: its
    launch parameters are the autotuner's winner for the matrix's shape class
    on this GPU (row-length histogram, dense width, compute units), tuned on
    the first call of a class and cached.  ``opt_buckets`` describes the row
    bins of the tuned plan."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "row_lengths",
        "max_row_nnz",
        "opt_buckets",
        "supports_opt",
        "_handle",
        "_params",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths, max_row_nnz,
                 opt_buckets=None):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self.opt_buckets = list(opt_buckets or [])
        self.supports_opt = data.dtype in SUPPORTED_SPMM_OPT_ALG2_DTYPES
        self._handle = CsrHandle(kernel_indices, kernel_indptr, self.shape)
        self._params = {}  # n_dense -> TuneParams


def prepare_spmm_csr_opt_alg2(data, indices, indptr, shape):
    """spmm_csr_opt_alg2.py:916-936."""
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz) = _prepare_spmm_csr_matrix(
        data, indices, indptr, shape
    )
    return PreparedCsrSpmmOptAlg2(data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths,
                                  max_row_nnz)


def prepare_spmm_csr_opt_alg2_preprocess(data, indices, indptr, shape):
    return prepare_spmm_csr_opt_alg2(data, indices, indptr, shape)


def _validate_spmm_opt_alg2_runtime_inputs(prepared, B, out):
    """Checks and messages in the order of spmm_csr_opt_alg2.py:943-978."""
    if not prepared.supports_opt:
        raise TypeError("flagsparse_spmm_csr_opt_alg2 only supports float32 and float64")
    if B is None:
        raise ValueError("B is required")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if B.shape[0] != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as sparse matrix data")
        if tuple(out.shape) != (prepared.n_rows, int(B.shape[1])) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")


def flagsparse_spmm_csr_opt_alg2(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None,
                                 out=None, return_time=False, return_meta=False):
    """CSR SpMM opt-alg2: the autotuned row-bin path for CSR @ dense
    (argument contract, returns and meta keys of spmm_csr_opt_alg2.py:1136-1215)."""
    if prepared is not None and not isinstance(prepared, PreparedCsrSpmmOptAlg2):
        raise TypeError("prepared must be a PreparedCsrSpmmOptAlg2 instance")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape)):
            raise ValueError("data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmm_csr_opt_alg2(data, indices, indptr, shape)
    elif shape is not None:
        resolved_shape = (int(shape[0]), int(shape[1]))
        if resolved_shape != prepared.shape:
            raise ValueError(f"shape {resolved_shape} does not match prepared.shape {prepared.shape}")
    _validate_spmm_opt_alg2_runtime_inputs(prepared, B, out)
    B = B.contiguous()
    n_dense = int(B.shape[1])
    device = prepared.data.device
    timer = _Timer(return_time or return_meta)
    timer.mark()
    params = prepared._params.get(n_dense)
    if params is None:
        params = autotune.tune_spmm(prepared._handle, prepared.data, n_dense)
        prepared._params[n_dense] = params
    lanes = params.lanes_n or _plan.lanes_for_dense(n_dense, B.element_size())
    cfg = params.spmm_config(lanes)
    plan = prepared._handle.plan(cfg) if prepared.n_rows else None
    if plan is not None:
        prepared.opt_buckets = [
            {"kind": "rows", "slots": int(l), "max_row_nnz": int(u), "rows": int(c)}
            for l, u, c in zip(plan.config.lanes, plan.config.uppers, plan.bin_counts)
            if c
        ]
        if plan.n_segments:
            prepared.opt_buckets.append({"kind": "split", "rows": plan.n_long_rows, "parts": plan.n_segments})
    symbolic_ms = timer.mark()
    direct_out = out if (out is not None and out.is_contiguous()) else None
    C = prepared._handle.spmm(prepared.data, B, out=direct_out, lanes_n=lanes, config=cfg,
                              max_blocks=params.max_blocks(device) if prepared.n_rows else 0,
                              xcd_groups=params.xcd_groups)
    compute_ms = timer.mark()
    op_total_ms = None if compute_ms is None else symbolic_ms + compute_ms
    if out is not None:
        if C.data_ptr() != out.data_ptr():
            out.copy_(C)
        C = out
    meta = None
    if return_meta:
        props = torch.cuda.get_device_properties(device)
        meta = {
            "device_name": props.name,
            "sm_count": int(props.multi_processor_count),
            "warp_size": int(getattr(props, "warp_size", 64)),
            "bucket_hits": {f"{b['kind']}{i}": int(b["rows"]) for i, b in enumerate(prepared.opt_buckets)},
            "launch_configs": {"unroll": int(params.unroll), "slot_scale": int(params.slot_scale),
                               "lanes_n": int(lanes), "blocks_per_cu": int(params.blocks_per_cu),
                               "xcd_groups": int(params.xcd_groups)},
            "max_row_nnz": prepared.max_row_nnz,
            "n_dense_cols": n_dense,
            "symbolic_ms": symbolic_ms,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
        }
    if return_time and return_meta:
        return C, op_total_ms, meta
    if return_time:
        return C, op_total_ms
    if return_meta:
        return C, meta
    return C


def flagsparse_spmm_csr_opt_alg2_preprocess(data=None, indices=None, indptr=None, B=None, shape=None,
                                            prepared=None, out=None, return_time=False, return_meta=False):
    """Compatibility alias for CSR SpMM opt-alg2 (spmm_csr_opt_alg2.py:1218-1241)."""
    return flagsparse_spmm_csr_opt_alg2(data=data, indices=indices, indptr=indptr, B=B, shape=shape,
                                        prepared=prepared, out=out, return_time=return_time,
                                        return_meta=return_meta)


# ------------------------------------------------------------------- routes
@dataclass(frozen=True)
class SpmmCsrAlgorithm:
    """Registered CSR SpMM route (reference: spmm_csr.py:619-627)."""

    name: str
    display_name: str
    supported_ops: tuple
    supported_dtypes: tuple
    run: object


class SpmmCsrAlgorithmUnavailable(RuntimeError):
    """Raised when a registered SpMM CSR route is unavailable in this runtime."""


class PreparedCsrSpmmRoute:
    """Matrix-level CSR SpMM route preparation (reference: spmm_csr.py:633-676)."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "row_lengths",
        "max_row_nnz",
        "nnz",
        "avg_nnz_per_row",
        "op",
        "alg",
        "_handle",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths, max_row_nnz, op,
                 alg):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self.nnz = int(data.numel())
        self.avg_nnz_per_row = float(self.nnz) / float(max(1, self.n_rows))
        self.op = str(op)
        self.alg = str(alg)
        self._handle = CsrHandle(kernel_indices, kernel_indptr, self.shape)


def _normalize_spmm_csr_alg(alg):
    """spmm_csr.py:679-699."""
    token = "auto" if alg is None else str(alg).strip().lower()
    if token in ("base", "csr", "csr_base"):
        return "csr_base"
    if token in ("base_accuracy", "csr_base_accuracy"):
        return "csr_base_accuracy"
    if token in ("alg1", "csr_alg1", "spmm_csr_alg1"):
        return "spmm_csr_alg1"
    if token in ("alg2", "csr_alg2", "spmm_csr_alg2"):
        return "spmm_csr_alg2"
    if token in ("alg2_accuracy", "csr_alg2_accuracy", "spmm_csr_alg2_accuracy"):
        return "spmm_csr_alg2_accuracy"
    if token in ("alg2_accuracy_hp", "csr_alg2_accuracy_hp", "spmm_csr_alg2_accuracy_hp"):
        return "spmm_csr_alg2_accuracy_hp"
    if token == "auto":
        return "auto"
    return token


def _normalize_dense_layout(layout):
    """spmm_csr.py:702-722."""
    token = "row" if layout is None else str(layout).strip().lower()
    if token in ("auto", "default"):
        return "row"
    if token in ("row", "row_major", "row-major", "c", "c_order"):
        return "row"
    if token in ("col", "column", "col_major", "column_major", "col-major", "column-major", "f", "fortran"):
        return "col"
    raise ValueError("dense_layout must be one of: auto, row, col")


def _is_col_major_2d(tensor):
    return (
        torch.is_tensor(tensor)
        and tensor.ndim == 2
        and tensor.stride(0) == 1
        and tensor.stride(1) >= max(1, int(tensor.shape[0]))
    )


def _dense_layout_name(tensor):
    if not torch.is_tensor(tensor) or tensor.ndim != 2:
        return "unknown"
    if tensor.is_contiguous():
        return "row"
    if _is_col_major_2d(tensor):
        return "col"
    return "strided"


def _empty_dense_layout(shape, dtype, device, layout):
    layout = _normalize_dense_layout(layout)
    rows, cols = int(shape[0]), int(shape[1])
    if layout == "col":
        return torch.empty_strided((rows, cols), (1, max(1, rows)), dtype=dtype, device=device)
    return torch.empty((rows, cols), dtype=dtype, device=device)


def _route_runner(route_name, display_name, accuracy):
    def run(prepared, B, *, op_name="non", timing=False, diagnostics=False, dense_layout="row"):
        dense_layout_n = _normalize_dense_layout(dense_layout)
        transposes = op_name in ("trans", "conj")
        n_out = prepared.n_cols if transposes else prepared.n_rows
        # the kernel streams row-major B rows; a col-major B is transposed once
        B_row = B.contiguous()
        compute_ms = None
        if timing:
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
        values = prepared.data
        if op_name == "conj" and _is_complex_dtype(values.dtype):
            values = values.conj().resolve_conj()
        require_kernel_dtype(values.dtype)
        C_row = prepared._handle.spmm(values, B_row, transpose=transposes)
        if dense_layout_n == "col":
            C = _empty_dense_layout((n_out, int(B.shape[1])), values.dtype, values.device, "col")
            C.copy_(C_row)
        else:
            C = C_row
        if timing:
            end.record()
            torch.cuda.synchronize()
            compute_ms = start.elapsed_time(end)
        meta = {
            "alg": route_name,
            "display_name": display_name,
            "op": op_name,
            "process_cpu_ms": 0.0,
            "process_gpu_ms": 0.0 if timing else None,
            "compute_ms": compute_ms,
            "dense_layout": dense_layout_n,
            "b_stride": tuple(int(v) for v in B.stride()),
            "c_stride": tuple(int(v) for v in C.stride()),
            "output_layout": _dense_layout_name(C),
        }
        if diagnostics:
            handle = prepared._handle
            lanes = _plan.lanes_for_dense(int(B.shape[1]), B.element_size())
            cfg = _plan.spmm_config(lanes)
            plan = handle.csc_plan(cfg) if transposes else handle.plan(cfg)
            meta["diagnostics"] = {
                "launch_config_scope": "matrix",
                "launch_config_count": 1,
                "bucket_count": sum(1 for c in plan.bin_counts if c),
                "long_row_count": plan.n_long_rows,
                "long_part_count": plan.n_segments,
                "launch_version": "hip_rows_v1",
                "lanes_n": lanes,
                "bin_slots": list(plan.config.lanes),
                "bin_uppers": list(plan.config.uppers),
                "bin_rows": list(plan.bin_counts),
                "unroll": plan.config.unroll,
                "warp_size": 64,
                "dense_layout": dense_layout_n,
                "b_stride": tuple(int(v) for v in B.stride()),
                "c_stride": tuple(int(v) for v in C.stride()),
                "output_layout": _dense_layout_name(C),
            }
        return C, meta

    return run


_REAL = (torch.float32, torch.float64)
_ALL_REAL = (torch.float16, torch.bfloat16, torch.float32, torch.float64)

SPMM_CSR_ALGORITHMS = {
    "csr_base": SpmmCsrAlgorithm("csr_base", "Base", tuple(SPMM_OP_NAMES.values()), _ALL_REAL,
                                 _route_runner("csr_base", "Base", False)),
    "csr_base_accuracy": SpmmCsrAlgorithm("csr_base_accuracy", "BaseAccuracy", tuple(SPMM_OP_NAMES.values()),
                                          _REAL, _route_runner("csr_base_accuracy", "BaseAccuracy", True)),
    "alpha_alg1_tle_opt": SpmmCsrAlgorithm("alpha_alg1_tle_opt", "TLEOpt", tuple(SPMM_OP_NAMES.values()), _REAL,
                                           _route_runner("alpha_alg1_tle_opt", "TLEOpt", False)),
    "alpha_alg1_tle_opt2": SpmmCsrAlgorithm("alpha_alg1_tle_opt2", "TLEOpt2", tuple(SPMM_OP_NAMES.values()),
                                            _REAL, _route_runner("alpha_alg1_tle_opt2", "TLEOpt2", False)),
    "spmm_csr_alg1": SpmmCsrAlgorithm("spmm_csr_alg1", "Alg1", tuple(SPMM_OP_NAMES.values()), _REAL,
                                      _route_runner("spmm_csr_alg1", "Alg1", False)),
    "spmm_csr_alg2": SpmmCsrAlgorithm("spmm_csr_alg2", "Alg2", tuple(SPMM_OP_NAMES.values()), _REAL,
                                      _route_runner("spmm_csr_alg2", "Alg2", False)),
    "spmm_csr_alg2_accuracy": SpmmCsrAlgorithm("spmm_csr_alg2_accuracy", "Alg2Accuracy",
                                               tuple(SPMM_OP_NAMES.values()), _REAL,
                                               _route_runner("spmm_csr_alg2_accuracy", "Alg2Accuracy", True)),
    "spmm_csr_alg2_accuracy_hp": SpmmCsrAlgorithm("spmm_csr_alg2_accuracy_hp", "Alg2AccuracyHP",
                                                  tuple(SPMM_OP_NAMES.values()), _REAL,
                                                  _route_runner("spmm_csr_alg2_accuracy_hp", "Alg2AccuracyHP",
                                                                True)),
}


def resolve_spmm_csr_algorithm(alg, op, dtype):
    """spmm_csr.py:2547-2562."""
    token = _normalize_spmm_csr_alg(alg)
    if token == "auto":
        token = "csr_base"
    if token not in SPMM_CSR_ALGORITHMS:
        supported = ", ".join(sorted(SPMM_CSR_ALGORITHMS))
        raise ValueError(f"unsupported CSR SpMM algorithm {alg!r}; supported: auto, {supported}")
    algorithm = SPMM_CSR_ALGORITHMS[token]
    op_name = _spmm_op_to_name(op)
    if op_name not in algorithm.supported_ops:
        raise ValueError(f"algorithm {token!r} does not support op {op_name!r}")
    if dtype not in algorithm.supported_dtypes:
        raise TypeError(f"algorithm {token!r} does not support dtype {dtype}")
    return algorithm


def list_spmm_csr_algorithms(op=None, dtype=None):
    op_name = None if op is None else _spmm_op_to_name(op)
    names = []
    for name, algorithm in SPMM_CSR_ALGORITHMS.items():
        if op_name is not None and op_name not in algorithm.supported_ops:
            continue
        if dtype is not None and dtype not in algorithm.supported_dtypes:
            continue
        names.append(name)
    return tuple(names)


def prepare_spmm_csr_route(data, indices, indptr, shape, *, op="non", alg="auto"):
    """Prepare matrix-level CSR SpMM metadata for the route-based run API."""
    op_code = _normalize_spmm_op(op)
    op_name = _spmm_op_to_name(op_code)
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz) = _prepare_spmm_csr_matrix(
        data, indices, indptr, shape
    )
    resolved_alg = _normalize_spmm_csr_alg(alg)
    if resolved_alg != "auto":
        resolve_spmm_csr_algorithm(resolved_alg, op_name, data.dtype)
    return PreparedCsrSpmmRoute(data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths,
                                max_row_nnz, op_name, resolved_alg)


def flagsparse_spmm_csr_run(
    prepared,
    B,
    *,
    alg=None,
    op=None,
    dense_layout="auto",
    return_time=False,
    return_meta=False,
    timing=False,
    diagnostics=False,
):
    """Run a registered CSR SpMM route (reference: spmm_csr.py:2607-2717)."""
    if not isinstance(prepared, PreparedCsrSpmmRoute):
        raise TypeError("prepared must be a PreparedCsrSpmmRoute instance")
    op_name = prepared.op if op is None else _spmm_op_to_name(op)
    alg_name = prepared.alg if alg is None else _normalize_spmm_csr_alg(alg)
    algorithm = resolve_spmm_csr_algorithm(alg_name, op_name, prepared.data.dtype)
    dense_layout = _normalize_dense_layout(dense_layout)
    if B is None or not torch.is_tensor(B):
        raise TypeError("B must be a torch.Tensor")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    eff_cols = prepared.n_rows if op_name in ("trans", "conj") else prepared.n_cols
    if B.shape[0] != eff_cols:
        raise ValueError(f"B.shape[0] must be n_cols={eff_cols}, got {B.shape[0]}")
    timed = return_time or return_meta
    if timed:
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
    C, route_meta = algorithm.run(prepared, B, op_name=op_name, timing=bool(timing), diagnostics=bool(diagnostics),
                                  dense_layout=dense_layout)
    gpu_ms = None
    if timed:
        end.record()
        torch.cuda.synchronize()
        gpu_ms = start.elapsed_time(end)
    process_cpu_ms = float(route_meta.get("process_cpu_ms", 0.0) or 0.0)
    operator_ms = (process_cpu_ms + float(gpu_ms)) if gpu_ms is not None else None
    meta = None
    if return_meta:
        meta = {
            "alg": algorithm.name,
            "display_name": algorithm.display_name,
            "op": op_name,
            "operator_ms": operator_ms,
            "gpu_ms": gpu_ms,
            "process_cpu_ms": process_cpu_ms,
            "dense_layout": route_meta.get("dense_layout", dense_layout),
            "b_stride": route_meta.get("b_stride"),
            "c_stride": route_meta.get("c_stride"),
            "output_layout": route_meta.get("output_layout"),
        }
        if timing:
            meta["process_gpu_ms"] = float(route_meta.get("process_gpu_ms") or 0.0)
            meta["compute_ms"] = route_meta.get("compute_ms")
        if diagnostics and "diagnostics" in route_meta:
            meta["diagnostics"] = route_meta["diagnostics"]
    if return_time and return_meta:
        return C, operator_ms, meta
    if return_time:
        return C, operator_ms
    if return_meta:
        return C, meta
    return C
