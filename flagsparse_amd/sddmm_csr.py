"""CSR SDDMM out[p] = alpha <x[row p], y[col p]> + beta data[p] on MI355X.

Drop-in for the reference's sddmm_csr operator (src/flagsparse/
sparse_operations/sddmm_csr.py:389-468): same keyword-first signature,
prepared object, validation order, exception types and messages
(int32 column indices, float32/float64 values, data required iff beta != 0,
K = 0 -> beta*data or zeros).  The compute runs in ``fsp_sddmm_csr``
(row-grouped: x[row] kept in registers across the row's nonzeros).
"""

import time

import torch

from . import _plan
from ._matrix import CsrHandle

SUPPORTED_SDDMM_VALUE_DTYPES = (torch.float32, torch.float64)
# waves per workgroup of the SDDMM kernel (256-thread workgroups)
_SDDMM_WAVES = 4


class SDDMMPrepared:
    """Prepared CSR pattern for SDDMM (reference type: sddmm_csr.py:23-47):
    the validated pattern and its handle, which caches the row-bin plans."""

    __slots__ = ("indices", "indptr", "shape", "n_rows", "n_cols", "nnz", "_handle")

    def __init__(self, indices, indptr, shape):
        self.indices = indices
        self.indptr = indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = self.shape[0]
        self.n_cols = self.shape[1]
        self.nnz = int(indices.numel())
        self._handle = CsrHandle(indices, indptr, self.shape)


def _sddmm_launch_meta(k, dtype):
    """The meta keys the reference reports for its launch (block_k,
    num_warps), filled with this kernel's: lanes along K and waves per
    workgroup."""
    elem = torch.empty(0, dtype=dtype).element_size()
    return {"block_k": _plan.sddmm_lanes(max(int(k), 1), elem), "num_warps": _SDDMM_WAVES}


def _prepare_sddmm_csr_pattern(indices, indptr, shape):
    """sddmm_csr.py:60-93."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple")
    if indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("indices and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if not indices.is_cuda or not indptr.is_cuda:
        raise ValueError("indices and indptr must be CUDA tensors")
    if indices.dtype != torch.int32:
        raise TypeError("indices dtype must be torch.int32")
    if indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    indptr64 = indptr.to(torch.int64).contiguous()
    indices = indices.contiguous()
    nnz = int(indices.numel())
    # every check from one device inspection and one host readback
    stats, _ = _plan.inspect_csr(indptr64, indices, n_rows, nnz)
    if stats["indptr_first"] != 0:
        raise ValueError("indptr[0] must be 0")
    if stats["indptr_last"] != nnz:
        raise ValueError(f"indptr[-1] must equal nnz={nnz}")
    if stats["n_decreasing"]:
        raise ValueError("indptr must be nondecreasing")
    if nnz > 0 and (stats["min_col"] < 0 or stats["max_col"] >= n_cols):
        raise IndexError("indices out of range for shape[1]")
    return indices, indptr64, (n_rows, n_cols)


def prepare_sddmm_csr(indices, indptr, shape, k_hint=64):
    """sddmm_csr.py:107-118.  ``k_hint`` is part of the reference's signature;
    the HIP launch picks its lanes from K at call time, so it is unused."""
    indices, indptr, shape = _prepare_sddmm_csr_pattern(indices, indptr, shape)
    return SDDMMPrepared(indices=indices, indptr=indptr, shape=shape)


def _validate_sddmm_dense_inputs(data, prepared, x, y):
    """sddmm_csr.py:219-244."""
    if x.ndim != 2 or y.ndim != 2:
        raise ValueError("x and y must be 2D dense tensors")
    if not x.is_cuda or not y.is_cuda:
        raise ValueError("x and y must be CUDA tensors")
    if x.device != y.device or x.device != prepared.indices.device:
        raise ValueError("x, y, and sparse pattern must be on the same CUDA device")
    if x.dtype not in SUPPORTED_SDDMM_VALUE_DTYPES:
        raise TypeError("x dtype must be torch.float32 or torch.float64")
    if y.dtype != x.dtype:
        raise TypeError("y dtype must match x dtype")
    if data is not None and data.dtype != x.dtype:
        raise TypeError("data dtype must match x/y dtype")
    if data is not None and (not torch.is_tensor(data) or not data.is_cuda or data.device != x.device):
        raise ValueError("data must be a CUDA tensor on the same device as x")
    if x.shape[0] != prepared.n_rows:
        raise ValueError(f"x.shape[0] must be n_rows={prepared.n_rows}, got {x.shape[0]}")
    if y.shape[0] != prepared.n_cols:
        raise ValueError(f"y.shape[0] must be n_cols={prepared.n_cols}, got {y.shape[0]}")
    if x.shape[1] != y.shape[1]:
        raise ValueError("x and y must have the same K dimension")
    if data is not None and data.numel() != prepared.nnz:
        raise ValueError("data length must equal nnz of sparse pattern")
    return int(x.shape[1])


def _prepare_validated_sddmm_out(prepared, x, out, out_dtype=None):
    """sddmm_csr.py:247-258."""
    nnz = prepared.nnz
    target_dtype = x.dtype if out_dtype is None else out_dtype
    if out is None:
        return torch.empty(nnz, dtype=target_dtype, device=x.device)
    if out.ndim != 1 or out.numel() != nnz:
        raise ValueError("out must be a 1D tensor with length nnz")
    if not out.is_cuda or out.device != x.device:
        raise ValueError("out must be a CUDA tensor on the same device as x")
    if out.dtype != target_dtype:
        raise TypeError("out dtype must match the requested output dtype")
    return out


def _run_sddmm_prepared(prepared, x, y, data, alpha, beta, out, allow_fallback=False):
    """Launch fsp_sddmm_csr (replaces sddmm_csr.py:291-386).  The HIP path has
    no silent fallback: `allow_fallback` is accepted for interface parity and
    a launch failure always raises."""
    out = _prepare_validated_sddmm_out(prepared, x, out)
    k_dim = int(x.shape[1])
    meta = {
        **_sddmm_launch_meta(k_dim, x.dtype),
        "fallback_used": False,
        "variant": "baseline",
        "acc_dtype": "float64" if x.dtype == torch.float64 else "float32",
        "out_dtype": str(out.dtype).replace("torch.", ""),
    }
    if prepared.nnz == 0:
        return out, meta
    # the kernel writes nnz consecutive elements: a strided `out` receives a copy
    target = out if out.is_contiguous() else torch.empty(prepared.nnz, dtype=out.dtype, device=out.device)
    prepared._handle.sddmm(x, y, alpha=alpha, beta=beta, data_in=data, out=target)
    if target is not out:
        out.copy_(target)
    return out, meta


def flagsparse_sddmm_csr(
    data=None,
    indices=None,
    indptr=None,
    x=None,
    y=None,
    shape=None,
    alpha=1.0,
    beta=0.0,
    prepared=None,
    out=None,
    return_time=False,
    return_meta=False,
    allow_fallback=False,
):
    """CSR SDDMM: out[p] = alpha * dot(x[row(p)], y[col(p)]) + beta * data[p]."""
    prepare_ms = 0.0
    if prepared is None:
        if any(v is None for v in (indices, indptr, shape)):
            raise ValueError("indices, indptr, and shape are required when prepared is not provided")
        t_prepare0 = time.perf_counter()
        prepared = prepare_sddmm_csr(indices, indptr, shape)
        prepare_ms = (time.perf_counter() - t_prepare0) * 1000.0
    elif not isinstance(prepared, SDDMMPrepared):
        raise TypeError("prepared must be a SDDMMPrepared instance")
    if x is None or y is None:
        raise ValueError("x and y are required")
    if data is None and float(beta) != 0.0:
        raise ValueError("data is required when beta is non-zero")
    k_dim = _validate_sddmm_dense_inputs(data, prepared, x, y)
    if k_dim == 0:
        out = _prepare_validated_sddmm_out(prepared, x, out)
        if beta == 0.0 or data is None:
            out.zero_()
        else:
            out.copy_(data * beta)
        meta = {"prepare_ms": prepare_ms, **_sddmm_launch_meta(0, x.dtype), "fallback_used": False}
        if return_time and return_meta:
            return out, 0.0, meta
        if return_time:
            return out, 0.0
        if return_meta:
            return out, meta
        return out
    timed = return_time or return_meta
    if timed:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    out_tensor, launch_meta = _run_sddmm_prepared(
        prepared, x.contiguous(), y.contiguous(), data.contiguous() if data is not None else None, alpha, beta, out,
        allow_fallback=allow_fallback,
    )
    if timed:
        torch.cuda.synchronize()
    elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if return_time and return_meta:
        return out_tensor, elapsed_ms, {"prepare_ms": prepare_ms, **launch_meta}
    if return_time:
        return out_tensor, elapsed_ms
    if return_meta:
        return out_tensor, {"prepare_ms": prepare_ms, **launch_meta}
    return out_tensor
