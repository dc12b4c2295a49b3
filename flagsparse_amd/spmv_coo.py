"""COO SpMV through a GPU COO -> CSR conversion (reference:
spmv_csr.py:948-1091, ``coo_to_csr_for_spmv``, ``prepare_spmv_coo_tocsr``,
``flagsparse_spmv_coo_tocsr``).

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

import torch

from . import _lib
from ._csc import gather_values
from .spmv_csr import (
    SUPPORTED_SPMV_VALUE_DTYPES,
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
