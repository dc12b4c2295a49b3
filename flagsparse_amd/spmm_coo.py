"""COO SpMM: C = op(A) @ B for COO input (reference: spmm_coo.py).

The reference canonicalises the COO (duplicates summed, lex (row, col)
order, row-run offsets) and runs a row-run, an atomic or a bucketed Triton
kernel, registered as the routes ``coo_rowrun`` / ``coo_atomic`` /
``spmm_coo_alg1``.  Its numerics: float32 inputs are computed in float64,
float16 / bfloat16 in float32, and the result is cast to the input dtype.

Here the canonical order comes from the stable HIP COO -> CSR conversion
(``fsp_coo_to_csr``, values moved bit-exactly) and every route runs the HIP
row-bin SpMM (``fsp_spmm_csr``) over that CSR in the same compute dtype, so
no atomics are involved and results are deterministic.  Duplicate (row, col)
entries stay as separate nonzeros in (row, col) order and the kernel sums
them (the reference sums them first; the two agree up to rounding).  Names,
arguments, validation order, messages, returns, meta keys and the dense
layouts (row / col) follow spmm_coo.py:28-1817.
"""

import time
from dataclasses import dataclass

import torch

from ._common import OP_NAMES, SUPPORTED_INDEX_DTYPES, _INDEX_LIMIT_INT32, _is_complex_dtype, normalize_op
from ._common import require_kernel_dtype
from ._matrix import CsrHandle
from .spmm_csr import _dense_layout_name, _empty_dense_layout, _is_col_major_2d, _normalize_dense_layout
from .spmv_coo import coo_to_csr_for_spmv

SUPPORTED_SPMM_VALUE_DTYPES = (torch.float16, torch.bfloat16, torch.float32, torch.float64, torch.complex64,
                               torch.complex128)
_DTYPE_ERROR = "data dtype must be one of: float16, bfloat16, float32, float64, complex64, complex128"


def _spmm_coo_compute_dtype(value_dtype):
    """spmm_coo.py:146-153: float32 in float64, 16-bit types in float32."""
    if _is_complex_dtype(value_dtype):
        return torch.complex128 if value_dtype == torch.complex64 else value_dtype
    if value_dtype in (torch.float16, torch.bfloat16):
        return torch.float32
    if value_dtype == torch.float32:
        return torch.float64
    return value_dtype


def _materialize_spmm_coo_op(data, row, col, shape, op_code):
    """op(A) as COO triples: trans / conj swap row and col (spmm_coo.py:64-72)."""
    if op_code == 0:
        return data, row, col, shape
    if op_code == 2 and _is_complex_dtype(data.dtype):
        data = data.conj().resolve_conj()
    return data, col, row, (int(shape[1]), int(shape[0]))


def _materialize_dense_layout(tensor, layout):
    """spmm_coo.py:133-143."""
    layout = _normalize_dense_layout(layout)
    if tensor.ndim != 2:
        raise ValueError("dense layout materialization expects a 2D tensor")
    if layout == "row":
        return tensor.contiguous()
    if _is_col_major_2d(tensor):
        return tensor
    out = _empty_dense_layout(tensor.shape, tensor.dtype, tensor.device, layout)
    out.copy_(tensor)
    return out


def _check_coo_ranges(row, col, n_rows, n_cols, nnz):
    """The nnz / index checks of spmm_coo.py:700-718 (same order and messages)."""
    if nnz > _INDEX_LIMIT_INT32:
        raise ValueError("nnz exceeds the int32 range supported by the Triton COO kernel")
    if nnz > 0:
        lo_r, hi_r = int(row.min().item()), int(row.max().item())
        lo_c, hi_c = int(col.min().item()), int(col.max().item())
        if lo_r < 0 or hi_r >= n_rows:
            raise IndexError("row indices out of range for n_rows")
        if lo_c < 0 or hi_c >= n_cols:
            raise IndexError("col indices out of range for n_cols")
        if hi_r > _INDEX_LIMIT_INT32:
            raise ValueError("row indices exceed the int32 range supported by the Triton kernel")
        if hi_c > _INDEX_LIMIT_INT32:
            raise ValueError("column indices exceed the int32 range supported by the Triton kernel")


def _prepare_spmm_coo_matrix(data, row, col, shape):
    """spmm_coo.py:1078-1132 (same checks, order and messages)."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple: (n_rows, n_cols)")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, and col must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if data.numel() != row.numel() or data.numel() != col.numel():
        raise ValueError("data, row, and col must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, row, col)):
        raise ValueError("data, row, and col must be CUDA tensors")
    if not all(t.device == data.device for t in (row, col)):
        raise ValueError("data, row, and col must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMM_VALUE_DTYPES:
        raise TypeError(_DTYPE_ERROR)
    if row.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("row dtype must be torch.int32 or torch.int64")
    if col.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("col dtype must be torch.int32 or torch.int64")
    _check_coo_ranges(row, col, n_rows, n_cols, int(data.numel()))
    return data.contiguous(), row.contiguous(), col.contiguous(), (n_rows, n_cols)


def _prepare_spmm_coo_inputs(data, row, col, B, shape, dense_layout="row"):
    """spmm_coo.py:666-728: the matrix checks interleaved with B's."""
    dense_layout = _normalize_dense_layout(dense_layout)
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple: (n_rows, n_cols)")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, and col must be 1D tensors")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if data.numel() != row.numel() or data.numel() != col.numel():
        raise ValueError("data, row, and col must have the same length (nnz)")
    if B.shape[0] != n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={n_cols}, got {B.shape[0]}")
    if not all(t.is_cuda for t in (data, row, col, B)):
        raise ValueError("data, row, col, and B must be CUDA tensors")
    if not all(t.device == data.device for t in (row, col, B)):
        raise ValueError("data, row, col, and B must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMM_VALUE_DTYPES:
        raise TypeError(_DTYPE_ERROR)
    if B.dtype != data.dtype:
        raise TypeError("B dtype must match data dtype")
    if row.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("row dtype must be torch.int32 or torch.int64")
    if col.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("col dtype must be torch.int32 or torch.int64")
    _check_coo_ranges(row, col, n_rows, n_cols, int(data.numel()))
    return data.contiguous(), row.contiguous(), col.contiguous(), _materialize_dense_layout(B, dense_layout), \
        n_rows, n_cols, int(B.shape[1])


class _CanonicalCoo:
    """op(A) in (row, col) order: sorted COO triples, row runs and the CSR
    handle the SpMM kernel runs, values in the compute dtype."""

    __slots__ = ("data", "row", "col", "indptr", "seg_starts", "row_lengths", "shape", "handle")

    def __init__(self, data, row, col, shape, compute_dtype):
        n_rows, n_cols = int(shape[0]), int(shape[1])
        if compute_dtype in (torch.float16, torch.bfloat16, torch.float32, torch.float64):
            values, indices, indptr = coo_to_csr_for_spmv(data.to(compute_dtype), row, col, (n_rows, n_cols))
        else:  # complex: the validators accept it, the kernels do not
            require_kernel_dtype(compute_dtype)
        nnz = int(values.numel())
        self.data = values
        self.col = indices
        self.indptr = indptr
        self.row = torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int32, device=data.device),
                                           indptr[1:] - indptr[:-1], output_size=nnz)
        runs = indptr[:-1][indptr[1:] > indptr[:-1]]
        self.seg_starts = torch.cat([runs, indptr[-1:]]).to(torch.int32) if nnz else None
        self.row_lengths = ((self.seg_starts[1:] - self.seg_starts[:-1]).contiguous() if nnz
                            else torch.empty(0, dtype=torch.int32, device=data.device))
        self.shape = (n_rows, n_cols)
        self.handle = CsrHandle(indices, indptr, self.shape)

    def spmm(self, B_compute, output_dtype, dense_layout, out=None):
        """C = op(A) B in the requested layout and dtype (row-major kernel)."""
        n_rows = self.shape[0]
        n_dense = int(B_compute.shape[1])
        C_acc = self.handle.spmm(self.data, B_compute.contiguous())
        if out is not None:
            out.copy_(C_acc)
            return out
        if dense_layout == "col":
            C = _empty_dense_layout((n_rows, n_dense), output_dtype, C_acc.device, "col")
            C.copy_(C_acc)
            return C
        return C_acc if C_acc.dtype == output_dtype else C_acc.to(output_dtype)


@dataclass(frozen=True)
class SpmmCooAlgorithm:
    """Registered COO SpMM route for the route-based run API (spmm_coo.py:308-315)."""

    name: str
    display_name: str
    supported_ops: tuple
    supported_dtypes: tuple
    run: object


class SpmmCooAlgorithmUnavailable(RuntimeError):
    """Raised when a registered COO SpMM algorithm is unavailable."""


class PreparedCooSpmmRoute:
    """Matrix-level COO SpMM preparation shared by the registered routes
    (spmm_coo.py:322-373): op(A) in (row, col) order in the compute dtype,
    plus the CSR handle of the HIP kernel."""

    __slots__ = (
        "data", "row", "col", "shape", "n_rows", "n_cols", "seg_starts", "row_lengths", "n_segs", "nnz",
        "max_row_nnz", "avg_nnz_per_row", "output_dtype", "compute_dtype", "op", "alg", "_canonical",
    )

    def __init__(self, canonical, output_dtype, compute_dtype, op, alg):
        self._canonical = canonical
        self.data = canonical.data
        self.row = canonical.row
        self.col = canonical.col
        self.shape = canonical.shape
        self.n_rows, self.n_cols = canonical.shape
        self.seg_starts = canonical.seg_starts
        self.row_lengths = canonical.row_lengths
        self.n_segs = int(self.row_lengths.numel())
        self.nnz = int(self.data.numel())
        self.max_row_nnz = int(self.row_lengths.max().item()) if self.row_lengths.numel() else 0
        self.avg_nnz_per_row = float(self.nnz) / float(max(1, self.n_rows))
        self.output_dtype = output_dtype
        self.compute_dtype = compute_dtype
        self.op = str(op)
        self.alg = str(alg)


def _validate_spmm_coo_route_runtime_inputs(prepared, B, dense_layout):
    """spmm_coo.py:1135-1155."""
    if B is None or not torch.is_tensor(B):
        raise TypeError("B must be a torch.Tensor")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if B.dtype != prepared.output_dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if int(B.shape[0]) != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    B_compute = B if prepared.compute_dtype == prepared.output_dtype else B.to(prepared.compute_dtype)
    return _materialize_dense_layout(B_compute, dense_layout)


def _route_runner(name, display_name):
    def run(prepared, B, *, timing=False, diagnostics=False, dense_layout="row"):
        dense_layout = _normalize_dense_layout(dense_layout)
        t_cpu = time.perf_counter()
        B_compute = _validate_spmm_coo_route_runtime_inputs(prepared, B, dense_layout)
        process_cpu_ms = (time.perf_counter() - t_cpu) * 1000.0
        compute_ms = None
        if timing:
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
        C = prepared._canonical.spmm(B_compute, prepared.output_dtype, dense_layout)
        if timing:
            end.record()
            torch.cuda.synchronize()
            compute_ms = start.elapsed_time(end)
        meta = {
            "alg": name,
            "display_name": display_name,
            "process_cpu_ms": process_cpu_ms,
            "process_gpu_ms": 0.0 if timing else None,
            "compute_ms": compute_ms,
            "dense_layout": dense_layout,
            "b_stride": tuple(int(v) for v in B_compute.stride()),
            "c_stride": tuple(int(v) for v in C.stride()),
            "output_layout": _dense_layout_name(C),
        }
        if diagnostics:
            meta["diagnostics"] = {
                "launch_version": "hip_rows_v1",
                "n_segs": prepared.n_segs,
                "max_row_nnz": prepared.max_row_nnz,
                "avg_nnz_per_row": prepared.avg_nnz_per_row,
                "compute_dtype": str(prepared.compute_dtype).replace("torch.", ""),
                "warp_size": 64,
            }
        return C, meta

    return run


_ALL_OPS = tuple(OP_NAMES.values())
SPMM_COO_ALGORITHMS = {
    "coo_rowrun": SpmmCooAlgorithm("coo_rowrun", "COORowRun", _ALL_OPS, SUPPORTED_SPMM_VALUE_DTYPES,
                                   _route_runner("coo_rowrun", "COORowRun")),
    "coo_atomic": SpmmCooAlgorithm("coo_atomic", "COOAtomic", _ALL_OPS, SUPPORTED_SPMM_VALUE_DTYPES,
                                   _route_runner("coo_atomic", "COOAtomic")),
    "spmm_coo_alg1": SpmmCooAlgorithm("spmm_coo_alg1", "COOAlg1", _ALL_OPS, (torch.float32, torch.float64),
                                      _route_runner("spmm_coo_alg1", "COOAlg1")),
}
_ALIASES = {"rowrun": "coo_rowrun", "coo": "coo_rowrun", "base": "coo_rowrun", "coo_base": "coo_rowrun",
            "atomic": "coo_atomic", "alg1": "spmm_coo_alg1", "coo_alg1": "spmm_coo_alg1"}


def _normalize_spmm_coo_alg(alg):
    """spmm_coo.py:1062-1075."""
    token = "auto" if alg is None else str(alg).strip().lower()
    if token == "auto":
        return "auto"
    return _ALIASES.get(token, token)


def resolve_spmm_coo_algorithm(alg, op, dtype):
    """spmm_coo.py:1470-1485."""
    token = _normalize_spmm_coo_alg(alg)
    if token == "auto":
        token = "coo_rowrun"
    if token not in SPMM_COO_ALGORITHMS:
        supported = ", ".join(sorted(SPMM_COO_ALGORITHMS))
        raise ValueError(f"unsupported COO SpMM algorithm {alg!r}; supported: auto, {supported}")
    algorithm = SPMM_COO_ALGORITHMS[token]
    op_name = OP_NAMES[normalize_op(op)]
    if op_name not in algorithm.supported_ops:
        raise ValueError(f"algorithm {token!r} does not support op {op_name!r}")
    if dtype not in algorithm.supported_dtypes:
        raise TypeError(f"algorithm {token!r} does not support dtype {dtype}")
    return algorithm


def list_spmm_coo_algorithms(op=None, dtype=None):
    """spmm_coo.py:1488-1497."""
    op_name = None if op is None else OP_NAMES[normalize_op(op)]
    return tuple(name for name, a in SPMM_COO_ALGORITHMS.items()
                 if (op_name is None or op_name in a.supported_ops) and (dtype is None or dtype in a.supported_dtypes))


def prepare_spmm_coo_route(data, row, col, shape, *, op="non", alg="auto"):
    """Matrix-level COO preparation for the registered routes (spmm_coo.py:1500-1541)."""
    op_code = normalize_op(op)
    op_name = OP_NAMES[op_code]
    data, row, col, shape = _materialize_spmm_coo_op(data, row, col, shape, op_code)
    output_dtype = data.dtype
    compute_dtype = _spmm_coo_compute_dtype(output_dtype)
    data, row, col, shape = _prepare_spmm_coo_matrix(data, row, col, shape)
    resolved_alg = _normalize_spmm_coo_alg(alg)
    if resolved_alg != "auto":
        resolve_spmm_coo_algorithm(resolved_alg, op_name, output_dtype)
    canonical = _CanonicalCoo(data, row, col, shape, compute_dtype)
    return PreparedCooSpmmRoute(canonical, output_dtype, compute_dtype, op_name, resolved_alg)


def flagsparse_spmm_coo_run(prepared, B, *, alg=None, dense_layout="auto", return_time=False, return_meta=False,
                            timing=False, diagnostics=False):
    """Run a registered COO SpMM route (spmm_coo.py:1544-1608)."""
    if not isinstance(prepared, PreparedCooSpmmRoute):
        raise TypeError("prepared must be a PreparedCooSpmmRoute instance")
    alg_name = prepared.alg if alg is None else _normalize_spmm_coo_alg(alg)
    algorithm = resolve_spmm_coo_algorithm(alg_name, prepared.op, prepared.output_dtype)
    dense_layout = _normalize_dense_layout(dense_layout)
    timed = bool(return_time or return_meta)
    if timed:
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
    C, route_meta = algorithm.run(prepared, B, timing=bool(timing), diagnostics=bool(diagnostics),
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
            "op": prepared.op,
            "operator_ms": operator_ms,
            "gpu_ms": gpu_ms,
            "process_cpu_ms": process_cpu_ms,
            "dense_layout": route_meta.get("dense_layout", dense_layout),
            "b_stride": route_meta.get("b_stride"),
            "c_stride": route_meta.get("c_stride"),
            "output_layout": route_meta.get("output_layout"),
        }
        if timing:
            meta["process_gpu_ms"] = route_meta.get("process_gpu_ms")
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


def flagsparse_spmm_coo(data, row, col, B, shape, block_n=None, block_nnz=256, out=None, return_time=False,
                        transpose=None, op=None, return_meta=False, dense_layout="auto"):
    """COO SpMM, C = op(A) @ B; op 0/'non', 1/'trans', 2/'conj'
    (spmm_coo.py:1780-1815 through _run_spmm_coo_route, :1666-1777)."""
    dense_layout = _normalize_dense_layout(dense_layout)
    op_explicit = op is not None
    op_code = normalize_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != (op_code != 0):
        raise ValueError("transpose conflicts with op")
    op_name = OP_NAMES[op_code]
    if block_n is not None and block_n <= 0:
        raise ValueError("block_n must be positive when provided")
    if block_nnz is not None and block_nnz <= 0:
        raise ValueError("block_nnz must be positive when provided")
    do_timing = bool(return_time or return_meta)
    symbolic_ms = 0.0 if do_timing else None
    if do_timing:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    data, row, col, shape = _materialize_spmm_coo_op(data, row, col, shape, op_code)
    data, row, col, B, n_rows, n_cols, n_dense = _prepare_spmm_coo_inputs(data, row, col, B, shape, dense_layout)
    output_dtype = data.dtype
    compute_dtype = _spmm_coo_compute_dtype(output_dtype)
    canonical = _CanonicalCoo(data, row, col, (n_rows, n_cols), compute_dtype)
    if do_timing:
        torch.cuda.synchronize()
        symbolic_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != data.device:
            raise ValueError("out must be on the same CUDA device as the inputs")
        if tuple(out.shape) != (n_rows, n_dense) or out.dtype != output_dtype:
            raise ValueError("out shape/dtype must match result")
    B_compute = B if compute_dtype == output_dtype else B.to(compute_dtype)
    if do_timing:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    C = canonical.spmm(B_compute, output_dtype, dense_layout, out=out)
    compute_ms = op_total_ms = None
    if do_timing:
        torch.cuda.synchronize()
        compute_ms = (time.perf_counter() - t1) * 1000.0
        op_total_ms = symbolic_ms + compute_ms
    meta = None
    if return_meta:
        meta = {
            "op": op_name,
            "route": "rowrun",
            "symbolic_ms": symbolic_ms,
            "compute_ms": compute_ms,
            "op_total_ms": op_total_ms,
            "dense_layout": dense_layout,
            "b_stride": tuple(int(v) for v in B_compute.stride()),
            "c_stride": tuple(int(v) for v in C.stride()),
            "output_layout": _dense_layout_name(C),
        }
    if return_time and return_meta:
        return C, op_total_ms, meta
    if return_time:
        return C, op_total_ms
    if return_meta:
        return C, meta
    return C
