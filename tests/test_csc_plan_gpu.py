"""Index work on the GPU, bit-exact: the stable CSC permutation, the
transposed CSR triple (values and output dtypes of the reference's
_transpose_csr_for_spmv / _for_spmm) and the row-bin plan."""

import ctypes

import numpy as np
import pytest
import torch

from flagsparse_amd import _lib, _plan, synthetic
from flagsparse_amd._csc import csr_to_csc_perm, transpose_csr
from oracle import oracle

from ._helpers import device_csr, long_row_matrix

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize(
    "m,n,density,seed", [(40, 33, 0.2, 1), (1, 100, 0.5, 2), (100, 1, 0.5, 3), (17, 64, 0.0, 4), (64, 64, 0.9, 5),
                         (3000, 70000, 0.0003, 6)]
)
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_csc_permutation_bit_exact(device, m, n, density, seed, index_dtype):
    ip, ix, v = synthetic.random_csr(m, n, density, seed, ensure_nonempty=False)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float64, index_dtype)
    colptr, rowind, perm = csr_to_csc_perm(ptr, idx, m, n)
    c_ref, r_ref, p_ref = oracle.csr_to_csc(ip, ix, (m, n))
    np.testing.assert_array_equal(colptr.cpu().numpy(), c_ref)
    np.testing.assert_array_equal(rowind.cpu().numpy(), r_ref)
    np.testing.assert_array_equal(perm.cpu().numpy(), p_ref)


def test_csc_power_law_full_size(device):
    ip, ix, v, _ = synthetic.power_law_rows(seed=1, n_dense=0)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    colptr, rowind, perm = csr_to_csc_perm(ptr, idx, 262144, 262144)
    c_ref, r_ref, p_ref = oracle.csr_to_csc(ip, ix, (262144, 262144))
    assert np.array_equal(colptr.cpu().numpy(), c_ref)
    assert np.array_equal(rowind.cpu().numpy(), r_ref)
    assert np.array_equal(perm.cpu().numpy(), p_ref)


@pytest.mark.parametrize("index_dtype,indptr_dtype", [(torch.int32, torch.int32), (torch.int64, torch.int32),
                                                      (torch.int32, torch.int64), (torch.int64, torch.int64)])
def test_transpose_csr_values_and_dtypes(device, index_dtype, indptr_dtype):
    ip, ix, v = synthetic.random_csr(90, 45, 0.1, seed=7)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32, index_dtype, indptr_dtype)
    dt, it, pt, shape_t = transpose_csr(data, idx, ptr, (90, 45))
    assert shape_t == (45, 90) and it.dtype == index_dtype and pt.dtype == indptr_dtype
    c_ref, r_ref, p_ref = oracle.csr_to_csc(ip, ix, (90, 45))
    assert np.array_equal(pt.cpu().numpy(), c_ref) and np.array_equal(it.cpu().numpy(), r_ref)
    assert torch.equal(dt.cpu(), data.cpu()[torch.from_numpy(p_ref)])  # bit-exact move


def test_transpose_csr_empty(device):
    data = torch.empty(0, device=device)
    dt, it, pt, shape_t = transpose_csr(data, torch.empty(0, dtype=torch.int32, device=device),
                                        torch.zeros(4, dtype=torch.int64, device=device), (3, 6))
    assert shape_t == (6, 3) and pt.shape == (7,) and pt.dtype == torch.int64 and int(pt.abs().sum()) == 0


def test_csc_workspace_query_on_device():
    n = ctypes.c_int64(0)
    assert _lib.load().fsp_csr_to_csc_workspace(1000, 1000, 5000, _lib.FSP_I32, ctypes.byref(n)) == 0
    assert n.value >= 5000 * 12


@pytest.mark.parametrize("indptr_dtype", [torch.int32, torch.int64])
def test_row_bins_partition_is_stable_and_complete(device, indptr_dtype):
    rng = np.random.Generator(np.random.PCG64(11))
    lengths = rng.integers(0, 600, 5000)
    lengths[[7, 4000]] = 9000  # long rows
    ip, ix = synthetic.distinct_sorted_columns(rng, lengths, 10000)
    ptr = torch.tensor(ip, dtype=indptr_dtype, device=device)
    cfg = _plan.spmm_config(16)
    plan = _plan.build_plan(ptr, 5000, cfg)
    rows = plan.rows.cpu().numpy()
    assert sorted(rows.tolist()) == list(range(5000))
    off = 0
    prev_upper = -1
    for upper, count in zip(cfg.uppers, plan.bin_counts):
        seg = rows[off: off + count]
        assert np.all(np.diff(seg) > 0)  # stable (ascending)
        assert np.all((lengths[seg] > prev_upper) & (lengths[seg] <= upper))
        off += count
        prev_upper = upper
    long_rows = plan.long_rows.cpu().numpy()
    assert long_rows.tolist() == np.flatnonzero(lengths > cfg.uppers[-1]).tolist()
    assert 7 in long_rows and 4000 in long_rows
    sb, se, sr = plan.seg_begin.cpu().numpy(), plan.seg_end.cpu().numpy(), plan.seg_row.cpu().numpy()
    ptr_seg = plan.long_seg_ptr.cpu().numpy()
    for i, r in enumerate(long_rows):
        s0, s1 = ptr_seg[i], ptr_seg[i + 1]
        assert sb[s0] == ip[r] and se[s1 - 1] == ip[r + 1] and np.all(sb[s0 + 1:s1] == se[s0:s1 - 1])
        assert np.all(sr[s0:s1] == r) and np.all(se[s0:s1] - sb[s0:s1] <= cfg.seg_len)


def test_identity_plan_for_uniform_rows(device):
    ip, ix, v, _ = synthetic.uniform_rows(4096, 4096, 16, seed=0)
    ptr = torch.tensor(ip, dtype=torch.int32, device=device)
    plan = _plan.build_plan(ptr, 4096, _plan.spmv_config())
    assert plan.identity and plan.bin_counts[0] == 4096 and plan.n_segments == 0


def test_plan_long_rows_only(device):
    ip, ix, v = long_row_matrix(3, 20000, [0, 1, 2], 10000, 0, seed=1)
    ptr = torch.tensor(ip, dtype=torch.int64, device=device)
    plan = _plan.build_plan(ptr, 3, _plan.spmm_config(16))
    assert sum(plan.bin_counts) == 0 and plan.n_long_rows == 3
    assert plan.n_segments == 3 * (-(-10000 // plan.config.seg_len))
