"""Internal handle of one CSR index structure and its cached GPU metadata.

A ``CsrHandle`` owns (references to) indptr / indices and caches, per kernel
configuration, the row-bin plans of the CSR and -- on first transposed use --
the stable CSC permutation (colptr, rowind, perm) with its own plans.  Only
the index structure is cached: a transposed SpMM gathers the values into CSC
order on every call (``fsp_gather`` into a buffer of the caching allocator)
and runs the plain kernel over the CSC, so a handle stays valid while the
values change (training).  The transposed SpMV reads them through ``perm``
(data[perm[q]]) inside its kernel.  All launches go through the C ABI of
libflagsparse_hip.so.
"""

import torch

from . import _lib, _plan
from ._csc import csr_to_csc_perm, gather_values


def _acc_dtype(dtype):
    return torch.float64 if dtype == torch.float64 else torch.float32


class CsrHandle:
    __slots__ = ("indices", "indptr", "shape", "n_rows", "n_cols", "nnz", "_plans", "_csc", "_csc_plans")

    def __init__(self, indices, indptr, shape):
        self.indices = indices
        self.indptr = indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.nnz = int(indices.numel())
        self._plans = {}
        self._csc = None
        self._csc_plans = {}

    # ------------------------------------------------------------- metadata
    def plan(self, config):
        p = self._plans.get(config)
        if p is None:
            p = _plan.build_plan(self.indptr, self.n_rows, config)
            self._plans[config] = p
        return p

    def csc(self):
        """(colptr, rowind, perm) of the stable CSC permutation (cached)."""
        if self._csc is None:
            self._csc = csr_to_csc_perm(self.indptr, self.indices, self.n_rows, self.n_cols)
        return self._csc

    def csc_plan(self, config):
        p = self._csc_plans.get(config)
        if p is None:
            colptr, _, _ = self.csc()
            p = _plan.build_plan(colptr, self.n_cols, config)
            self._csc_plans[config] = p
        return p

    # --------------------------------------------------------------- kernels
    def spmm(self, data, B, transpose=False, out=None, lanes_n=None, config=None, max_blocks=0, xcd_groups=0,
             values_via_perm=False):
        """C = A B (transpose=False) or A^T B (transpose=True); B row-major.
        max_blocks > 0: persistent launch of at most that many workgroups;
        xcd_groups > 1: XCD-region tile mapping (see _plan.CsrPlan.struct).
        transpose: the values are first gathered into CSC order (one pass of
        nnz random reads); the kernel then streams them like a plain CSR.  A
        kernel that loads data[perm[q]] itself puts that random read in every
        nonzero's dependent chain and measures 10-19 % slower (262144^2 power
        law at N = 64 and the 200000-node graph at N = 128, f32 and bf16).
        values_via_perm=True keeps that form; both are bitwise equal."""
        n_out = self.n_cols if transpose else self.n_rows
        n_dense = int(B.shape[1])
        device = B.device
        C = out if out is not None else torch.empty((n_out, n_dense), dtype=data.dtype, device=device)
        if n_out == 0 or n_dense == 0:
            return C
        if (self.n_rows if transpose else self.n_cols) == 0 or self.nnz == 0:
            C.zero_()
            return C
        G = int(lanes_n) if lanes_n else _plan.lanes_for_dense(n_dense, data.element_size())
        cfg = config or _plan.spmm_config(G)
        if transpose:
            colptr, rowind, perm = self.csc()
            plan = self.csc_plan(cfg)
            ptr_t, idx_t, perm_t = colptr, rowind, perm
            if not values_via_perm:
                data, perm_t = gather_values(data, perm), None
        else:
            plan = self.plan(cfg)
            ptr_t, idx_t, perm_t = self.indptr, self.indices, None
        partials = plan.partials(n_dense, _acc_dtype(data.dtype), device)
        st = plan.struct(partials, max_blocks, xcd_groups)
        _lib.check(
            _lib.load().fsp_spmm_csr(
                _lib.value_code(data.dtype),
                _lib.index_code(idx_t.dtype),
                _lib.index_code(ptr_t.dtype),
                _lib.ptr(ptr_t),
                _lib.ptr(idx_t),
                _lib.ptr(data),
                _lib.index_code(perm_t.dtype) if perm_t is not None else 0,
                _lib.ptr(perm_t),
                B.data_ptr(),
                int(B.stride(0)),
                C.data_ptr(),
                int(C.stride(0)),
                n_out,
                n_dense,
                self.nnz,
                int(B.shape[0]),
                G,
                st,
                _lib.stream_ptr(device),
            ),
            "fsp_spmm_csr",
        )
        return C

    def spmm_trans_sddmm(self, data, dC, B, dB=None, dvalues=None):
        """The backward of C = A B in one pass over the CSC:
        dB = A^T dC and dvalues[p] = <dC[row p], B[col p]> (CSR order), from
        one gather of dC's rows (fsp_spmm_csr_trans_sddmm).  dB is bitwise
        equal to spmm(data, dC, transpose=True).  Returns None when n_dense
        does not fit one lane group of the kernel (the caller then runs the
        transposed SpMM and the SDDMM)."""
        n_dense = int(dC.shape[1])
        elem = data.element_size()
        device = dC.device
        dB = dB if dB is not None else torch.empty((self.n_cols, n_dense), dtype=data.dtype, device=device)
        dvalues = dvalues if dvalues is not None else torch.empty(self.nnz, dtype=data.dtype, device=device)
        vec = 16 // elem
        wide = all(t.data_ptr() % 16 == 0 and int(t.stride(0)) % vec == 0 for t in (dC, B, dB)) and n_dense % vec == 0
        if n_dense > 64 * (vec if wide else 1):
            return None
        G = _plan.lanes_for_dense(n_dense, elem)
        cfg = _plan.spmm_config(G)
        colptr, rowind, perm = self.csc()
        plan = self.csc_plan(cfg)
        data_t = gather_values(data, perm) if self.nnz else data
        partials = plan.partials(n_dense, _acc_dtype(data.dtype), device)
        st = plan.struct(partials)
        _lib.check(
            _lib.load().fsp_spmm_csr_trans_sddmm(
                _lib.value_code(data.dtype), _lib.index_code(rowind.dtype), _lib.index_code(colptr.dtype),
                _lib.ptr(colptr), _lib.ptr(rowind), _lib.ptr(data_t), _lib.index_code(perm.dtype), _lib.ptr(perm),
                dC.data_ptr(), int(dC.stride(0)), B.data_ptr(), int(B.stride(0)), dB.data_ptr(), int(dB.stride(0)),
                dvalues.data_ptr(), self.n_cols, n_dense, self.nnz, self.n_rows, G, st, _lib.stream_ptr(device),
            ),
            "fsp_spmm_csr_trans_sddmm",
        )
        return dB, dvalues

    def spmv(self, data, x, transpose=False, out=None, config=None, max_blocks=0, xcd_groups=0, variant=0,
             values_via_perm=True):
        """y = A x or A^T x.  transpose with values_via_perm=False gathers the
        values into CSC order first, as the transposed SpMM does."""
        n_out = self.n_cols if transpose else self.n_rows
        y = out if out is not None else torch.empty(n_out, dtype=data.dtype, device=x.device)
        if n_out == 0:
            return y
        cfg = config or _plan.spmv_config()
        if transpose:
            colptr, rowind, perm = self.csc()
            plan = self.csc_plan(cfg)
            ptr_t, idx_t, perm_t = colptr, rowind, perm
            if not values_via_perm and self.nnz:
                data, perm_t = gather_values(data, perm), None
        else:
            plan = self.plan(cfg)
            ptr_t, idx_t, perm_t = self.indptr, self.indices, None
        acc = torch.float64 if data.dtype in (torch.float32, torch.float64) else torch.float32
        partials = plan.partials(1, acc, x.device)
        st = plan.struct(partials, max_blocks, xcd_groups, variant)
        _lib.check(
            _lib.load().fsp_spmv_csr(
                _lib.value_code(data.dtype),
                _lib.index_code(idx_t.dtype),
                _lib.index_code(ptr_t.dtype),
                _lib.ptr(ptr_t),
                _lib.ptr(idx_t),
                _lib.ptr(data),
                _lib.index_code(perm_t.dtype) if perm_t is not None else 0,
                _lib.ptr(perm_t),
                _lib.ptr(x),
                y.data_ptr(),
                n_out,
                self.nnz,
                int(x.shape[0]),
                st,
                _lib.stream_ptr(x.device),
            ),
            "fsp_spmv_csr",
        )
        return y

    def sddmm(self, X, Y, alpha=1.0, beta=0.0, data_in=None, out=None, lanes_k=None, config=None, max_blocks=0,
              xcd_groups=0):
        """out[p] = alpha <X[row p], Y[col p]> + beta data_in[p] on this pattern."""
        k = int(X.shape[1])
        device = X.device
        out = out if out is not None else torch.empty(self.nnz, dtype=X.dtype, device=device)
        if self.nnz == 0:
            return out
        if k == 0:
            if data_in is None or beta == 0.0:
                out.zero_()
            else:
                out.copy_(data_in * beta)
            return out
        G = int(lanes_k) if lanes_k else _plan.sddmm_lanes(k, X.element_size())
        cfg = config or _plan.sddmm_config(G)
        plan = self.plan(cfg)
        st = plan.struct(None, max_blocks, xcd_groups)
        _lib.check(
            _lib.load().fsp_sddmm_csr(
                _lib.value_code(X.dtype),
                _lib.index_code(self.indices.dtype),
                _lib.index_code(self.indptr.dtype),
                _lib.ptr(self.indptr),
                _lib.ptr(self.indices),
                X.data_ptr(),
                int(X.stride(0)),
                Y.data_ptr(),
                int(Y.stride(0)),
                _lib.ptr(data_in),
                out.data_ptr(),
                self.n_rows,
                k,
                self.nnz,
                int(Y.shape[0]),
                float(alpha),
                float(beta),
                G,
                st,
                _lib.stream_ptr(device),
            ),
            "fsp_sddmm_csr",
        )
        return out
