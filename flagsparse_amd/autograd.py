"""Differentiable CSR SpMM with a hand-written backward.

C = A B with A given by (values, indices, indptr, shape).  Backward:

* dB = A^T dC: the transposed SpMM kernel over the handle's cached stable CSC
  permutation (values gathered into CSC order per call; no transposed copy
  of A's indices, no atomics);
* dA.values = SDDMM(dC, B) on A's pattern: dvalues[p] = <dC[row p], B[col p]>.

When both are needed they come from one kernel pass over the CSC
(``CsrHandle.spmm_trans_sddmm``): a slot that gathers dC[i] for dB[j] also
forms <dC[i], B[j]>, so dC's rows are gathered once instead of once per
gradient.  Operands too wide for one lane group of that kernel take the two
kernels.

The reference has no autograd (no backward anywhere in its sources); its
forward is flagsparse_spmm_csr (spmm_csr.py:3981) and the gradient pieces map
onto its trans-SpMM and flagsparse_sddmm_csr (sddmm_csr.py:389).  Values and
B may be float64, float32, bfloat16 or float16 (f32 accumulation for the
16-bit types, outputs in the storage dtype).
"""

import torch

from ._common import require_kernel_dtype
from ._matrix import CsrHandle
from .spmm_csr import _prepare_spmm_csr_matrix


class CsrMatrix:
    """A CSR sparsity pattern prepared for repeated differentiable SpMM.

    Holds the index arrays and caches the row-bin plans and the CSC
    permutation used by the backward.  Values are passed per call so they
    can be a trainable parameter.
    """

    def __init__(self, indices, indptr, shape, validate=True):
        if validate:
            dummy = torch.empty(indices.numel(), dtype=torch.float32, device=indices.device)
            _, indices, indptr, *_ = _prepare_spmm_csr_matrix(dummy, indices, indptr, shape)
        self.handle = CsrHandle(indices.contiguous(), indptr.contiguous(), shape)
        self.shape = self.handle.shape

    @property
    def indices(self):
        return self.handle.indices

    @property
    def indptr(self):
        return self.handle.indptr

    def prepare_backward(self):
        """Build the CSC permutation now (otherwise on the first backward)."""
        self.handle.csc()
        return self

    def spmm(self, values, B):
        return spmm_csr_autograd(values, B, self)


class _SpmmCsrFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, B, matrix):
        require_kernel_dtype(values.dtype, "values")
        if B.dtype != values.dtype:
            raise TypeError("B dtype must match values dtype")
        if B.ndim != 2 or B.shape[0] != matrix.shape[1]:
            raise ValueError(f"B must be 2D with B.shape[0] == n_cols={matrix.shape[1]}")
        if values.ndim != 1 or values.numel() != matrix.handle.nnz:
            raise ValueError(f"values must be 1D with one entry per nonzero (nnz={matrix.handle.nnz})")
        device = matrix.indices.device
        if not values.is_cuda or not B.is_cuda or values.device != device or B.device != device:
            raise ValueError("values and B must be CUDA tensors on the matrix's device")
        values_c = values.contiguous()
        B_c = B.contiguous()
        ctx.matrix = matrix
        ctx.save_for_backward(values_c, B_c)
        return matrix.handle.spmm(values_c, B_c)

    @staticmethod
    def backward(ctx, dC):
        values, B = ctx.saved_tensors
        matrix = ctx.matrix
        dC = dC.contiguous()
        if dC.dtype != values.dtype:
            dC = dC.to(values.dtype)
        d_values = d_B = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            fused = matrix.handle.spmm_trans_sddmm(values, dC, B)
            if fused is not None:
                d_B, d_values = fused
                return d_values, d_B, None
        if ctx.needs_input_grad[0]:
            d_values = matrix.handle.sddmm(dC, B, alpha=1.0, beta=0.0)
        if ctx.needs_input_grad[1]:
            d_B = matrix.handle.spmm(values, dC, transpose=True)
        return d_values, d_B, None


def spmm_csr_autograd(values, B, matrix):
    """C = A B, differentiable in `values` and `B` (A's pattern from `matrix`)."""
    if not isinstance(matrix, CsrMatrix):
        raise TypeError("matrix must be a CsrMatrix")
    return _SpmmCsrFunction.apply(values, B, matrix)


def flagsparse_spmm_csr_autograd(data, indices, indptr, B, shape, matrix=None):
    """Functional form: builds (or reuses) the CsrMatrix and applies SpMM."""
    if matrix is None:
        matrix = CsrMatrix(indices, indptr, shape)
    return spmm_csr_autograd(data, B, matrix)
