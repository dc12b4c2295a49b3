"""COO SpMV through the GPU COO -> CSR conversion: the CSR triple bit-exact
against the oracle's stable (row, col) ordering, SpMV against the FP64
oracle, the prepared path and the validation errors."""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from oracle import oracle as O

from ._helpers import REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def _coo(m, n, k, seed, dup=0.2):
    rng = np.random.default_rng(seed)
    row, col = rng.integers(0, m, k), rng.integers(0, n, k)
    d = int(k * dup)
    if d:
        row[:d], col[:d] = row[k - d:], col[k - d:]
    return row, col, rng.uniform(-1, 1, k)


@pytest.mark.parametrize("m,n,k", [(50, 40, 300), (1, 1000, 500), (2000, 3, 4000), (7, 9, 1), (300000, 5000, 200000)])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64])
def test_coo_to_csr_bit_exact(device, m, n, k, index_dtype):
    row, col, v = _coo(m, n, k, seed=m + k)
    d = torch.from_numpy(v).to(device, torch.float32)
    r = torch.from_numpy(row).to(device, index_dtype)
    c = torch.from_numpy(col).to(device, index_dtype)
    data_s, indices, indptr = fs.coo_to_csr_for_spmv(d, r, c, (m, n))
    ip, ix, dv = O.coo_to_csr(row, col, d.cpu().numpy(), (m, n))
    assert indices.dtype == torch.int32 and indptr.dtype == torch.int64
    np.testing.assert_array_equal(indptr.cpu().numpy(), ip)
    np.testing.assert_array_equal(indices.cpu().numpy(), ix)
    np.testing.assert_array_equal(data_s.cpu().numpy(), dv)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_coo_spmv_matches_oracle(device, dtype):
    m, n = 700, 500
    row, col, v = _coo(m, n, 9000, seed=3)
    x = np.random.default_rng(4).uniform(-1, 1, n)
    args = (torch.from_numpy(v).to(device, dtype), torch.from_numpy(row).to(device), torch.from_numpy(col).to(device))
    xt = torch.from_numpy(x).to(device, dtype)
    y = fs.flagsparse_spmv_coo_tocsr(*args, xt, (m, n))
    ip, ix, dv = O.coo_to_csr(row, col, v, (m, n))
    want = O.spmv(ip, ix, dv, x, (m, n))
    assert rel_err(y, want) <= REL_TOL[dtype]
    p = fs.prepare_spmv_coo_tocsr(*args, (m, n))
    y2 = fs.flagsparse_spmv_coo_tocsr(x=xt, prepared=p)
    assert torch.equal(y, y2)


def test_coo_empty_and_errors(device):
    e = torch.empty(0, device=device)
    ei = torch.empty(0, dtype=torch.int64, device=device)
    y = fs.flagsparse_spmv_coo_tocsr(e, ei, ei, torch.ones(4, device=device), (3, 4))
    assert y.tolist() == [0.0, 0.0, 0.0]
    d = torch.ones(2, device=device)
    with pytest.raises(IndexError, match="row indices out of range"):
        fs.coo_to_csr_for_spmv(d, torch.tensor([0, 3], device=device), torch.tensor([0, 1], device=device), (3, 4))
    with pytest.raises(IndexError, match="out of range for n_cols"):
        fs.coo_to_csr_for_spmv(d, torch.tensor([0, 1], device=device), torch.tensor([0, -1], device=device), (3, 4))
    with pytest.raises(ValueError, match="same length"):
        fs.prepare_spmv_coo_tocsr(d, torch.tensor([0], device=device), torch.tensor([0, 1], device=device), (3, 4))
    with pytest.raises(TypeError, match="x dtype"):
        fs.flagsparse_spmv_coo_tocsr(d, torch.tensor([0, 1], device=device), torch.tensor([0, 1], device=device),
                                     torch.ones(4, dtype=torch.float64, device=device), (3, 4))
