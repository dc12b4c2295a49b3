"""fsp_spmm_csr_noalloc (the SpMM entry whose long-row partials live in a
static device array) against fsp_spmm_csr (partials from the caller), called
through the C ABI: two calls in a row, the second on a smaller input, each
bit-equal to the existing entry point and within the float32 / float64 bar of
the oracle."""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from flagsparse_amd import _lib, _plan
from flagsparse_amd._matrix import CsrHandle
from oracle import oracle

from ._helpers import REL_TOL, assert_rel, device_csr, long_row_matrix

pytestmark = pytest.mark.gpu


def _call_noalloc(handle, plan, data, B, C, lanes, transpose=False):
    if transpose:
        colptr, rowind, perm = handle.csc()
        ptr_t, idx_t, perm_t = colptr, rowind, perm
        n_out = handle.n_cols
    else:
        ptr_t, idx_t, perm_t = handle.indptr, handle.indices, None
        n_out = handle.n_rows
    st = plan.struct(None)  # no partials: the entry point uses its own scratch
    return _lib.load().fsp_spmm_csr_noalloc(
        _lib.value_code(data.dtype), _lib.index_code(idx_t.dtype), _lib.index_code(ptr_t.dtype),
        _lib.ptr(ptr_t), _lib.ptr(idx_t), _lib.ptr(data),
        _lib.index_code(perm_t.dtype) if perm_t is not None else 0, _lib.ptr(perm_t),
        B.data_ptr(), int(B.stride(0)), C.data_ptr(), int(C.stride(0)), n_out, int(B.shape[1]), handle.nnz,
        int(B.shape[0]), lanes, st, _lib.stream_ptr(B.device))


def _case(device, n_rows, n_cols, n_dense, long_rows, long_len, seed, dtype, transpose):
    if transpose:  # long columns, so that A^T has long rows
        ip0, ix0, v0 = long_row_matrix(n_cols, n_rows, long_rows, long_len, 12, seed)
        At = sp.csr_matrix((v0, ix0, ip0), shape=(n_cols, n_rows)).T.tocsr()
        At.sort_indices()
        ip, ix, v = At.indptr.astype(np.int64), At.indices.astype(np.int64), At.data
    else:
        ip, ix, v = long_row_matrix(n_rows, n_cols, long_rows, long_len, 12, seed)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    rng = np.random.default_rng(seed + 1)
    b_rows = n_rows if transpose else n_cols
    Bn = rng.standard_normal((b_rows, n_dense))
    B = torch.tensor(Bn, dtype=dtype, device=device)
    handle = CsrHandle(idx, ptr, (n_rows, n_cols))
    lanes = _plan.lanes_for_dense(n_dense, data.element_size())
    cfg = _plan.spmm_config(lanes)
    plan = handle.csc_plan(cfg) if transpose else handle.plan(cfg)
    assert plan.n_segments > 0  # the case exercises the scratch
    ref = handle.spmm(data, B, transpose=transpose, lanes_n=lanes, config=cfg)
    n_out = n_cols if transpose else n_rows
    C = torch.full((n_out, n_dense), float("nan"), dtype=dtype, device=device)
    status = _call_noalloc(handle, plan, data, B, C, lanes, transpose)
    assert status == 0, _lib.load().fsp_last_error()
    torch.cuda.synchronize()
    assert torch.equal(C, ref)
    vals = v.astype(np.float32) if dtype == torch.float32 else v
    want = oracle.spmm(ip, ix, vals, Bn.astype(np.float32) if dtype == torch.float32 else Bn, (n_rows, n_cols),
                       op="trans" if transpose else "non")
    assert_rel(C, want, dtype, "noalloc vs oracle")
    return plan.n_segments


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("transpose", [False, True])
def test_noalloc_matches_spmm_twice_second_smaller(device, dtype, transpose):
    big = _case(device, 3000, 2600, 64, [5, 900, 2599], 2500, 31, dtype, transpose)
    small = _case(device, 700, 650, 32, [3], 600, 32, dtype, transpose)
    assert small < big
    assert REL_TOL[dtype] > 0


def test_noalloc_rejects_oversized_partials_before_launch(device):
    ip, ix, v = long_row_matrix(400, 400, [1], 399, 4, 5)
    data, idx, ptr = device_csr(ip, ix, v, device)
    handle = CsrHandle(idx, ptr, (400, 400))
    plan = handle.plan(_plan.spmm_config(16))
    st = plan.struct(None)
    # claim enough segments that n_segments x 64 x 4 bytes exceed the scratch
    st.n_segments = _lib.NOALLOC_SCRATCH_BYTES // (64 * 4) + 1
    B = torch.zeros((400, 64), device=device)
    C = torch.zeros((400, 64), device=device)
    status = _lib.load().fsp_spmm_csr_noalloc(
        _lib.FSP_F32, _lib.FSP_I32, _lib.FSP_I32, ptr.data_ptr(), idx.data_ptr(), data.data_ptr(), 0, None,
        B.data_ptr(), 64, C.data_ptr(), 64, 400, 64, handle.nnz, 400, 16, st, _lib.stream_ptr(device))
    assert status == -4
    assert b"static scratch" in _lib.load().fsp_last_error()
