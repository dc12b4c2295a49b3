"""CSR -> CSC permutation on the GPU (stable by column, rows ascending).

``csr_to_csc_perm`` returns (colptr, rowind, perm) with CSC position q holding
CSR nonzero perm[q].  ``transpose_csr`` materialises the transposed CSR triple
with the output dtypes of the reference's _transpose_csr_for_spmv
(spmv_csr.py:498-536) / _transpose_csr_for_spmm (spmm_csr.py:407-445):
indices keep their dtype unless n_rows > 2**31-1 (then int64), indptr keeps its
dtype unless nnz > 2**31-1 (then int64).  Values are moved bit-exactly.
"""

import ctypes

import torch

from . import _lib
from ._common import _INDEX_LIMIT_INT32


def _perm_dtype(nnz):
    return torch.int32 if nnz <= _INDEX_LIMIT_INT32 else torch.int64


def transposed_index_dtypes(indices_dtype, indptr_dtype, n_rows, nnz):
    out_index = indices_dtype if n_rows <= _INDEX_LIMIT_INT32 else torch.int64
    out_indptr = indptr_dtype if nnz <= _INDEX_LIMIT_INT32 else torch.int64
    return out_index, out_indptr


def csr_to_csc_perm(indptr, indices, n_rows, n_cols, rowind_dtype=None, colptr_dtype=None):
    """Stable CSC permutation of a CSR index structure (one radix sort)."""
    device = indices.device
    nnz = int(indices.numel())
    n_rows, n_cols = int(n_rows), int(n_cols)
    if rowind_dtype is None or colptr_dtype is None:
        ri, cp = transposed_index_dtypes(indices.dtype, indptr.dtype, n_rows, nnz)
        rowind_dtype = rowind_dtype or ri
        colptr_dtype = colptr_dtype or cp
    perm_dtype = _perm_dtype(nnz)
    colptr = torch.empty(n_cols + 1, dtype=colptr_dtype, device=device)
    rowind = torch.empty(nnz, dtype=rowind_dtype, device=device)
    perm = torch.empty(nnz, dtype=perm_dtype, device=device)
    lib = _lib.load()
    ws_bytes = ctypes.c_int64(0)
    _lib.check(
        lib.fsp_csr_to_csc_workspace(n_rows, n_cols, nnz, _lib.index_code(perm_dtype), ctypes.byref(ws_bytes)),
        "fsp_csr_to_csc_workspace",
    )
    ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=device)
    _lib.check(
        lib.fsp_csr_to_csc(
            _lib.index_code(indptr.dtype), _lib.index_code(indices.dtype), _lib.ptr(indptr), _lib.ptr(indices),
            n_rows, n_cols, nnz, _lib.index_code(colptr_dtype), colptr.data_ptr(), _lib.index_code(rowind_dtype),
            _lib.ptr(rowind), _lib.index_code(perm_dtype), _lib.ptr(perm), ws.data_ptr(), ws.numel(),
            _lib.stream_ptr(device),
        ),
        "fsp_csr_to_csc",
    )
    return colptr, rowind, perm


def gather_values(values, perm):
    """out[q] = values[perm[q]] on the GPU (bit-exact copy)."""
    out = torch.empty(perm.numel(), dtype=values.dtype, device=values.device)
    if perm.numel() == 0:
        return out
    _lib.check(
        _lib.load().fsp_gather(
            values.element_size(), _lib.index_code(perm.dtype), perm.data_ptr(), values.data_ptr(),
            out.data_ptr(), perm.numel(), _lib.stream_ptr(values.device),
        ),
        "fsp_gather",
    )
    return out


def transpose_csr(data, indices, indptr, shape):
    """Transposed CSR triple (data_t, indices_t, indptr_t, (n_cols, n_rows))."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    nnz = int(data.numel())
    device = data.device
    ri, cp = transposed_index_dtypes(indices.dtype, indptr.dtype, n_rows, nnz)
    if nnz == 0:
        return (
            data,
            torch.empty(0, dtype=ri, device=device),
            torch.zeros(n_cols + 1, dtype=cp, device=device),
            (n_cols, n_rows),
        )
    colptr, rowind, perm = csr_to_csc_perm(indptr, indices, n_rows, n_cols, ri, cp)
    data_t = gather_values(data.contiguous(), perm)
    return data_t, rowind, colptr, (n_cols, n_rows)
