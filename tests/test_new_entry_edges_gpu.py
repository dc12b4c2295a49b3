"""Empty and degenerate shapes through the CSC / COO SpMV, COO SpMM and
AlphaSparse-ALG1 entry points: zero rows, zero columns, no nonzeros and a
dense operand of zero width all return correctly shaped zeros (or empty
tensors) without launching on empty arrays."""

import pytest
import torch

import flagsparse_amd as fs

pytestmark = pytest.mark.gpu

I32 = torch.int32


def _e(device, dtype=torch.float32):
    return torch.empty(0, dtype=dtype, device=device)


@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("shape", [(0, 5), (5, 0), (0, 0), (4, 6)])
def test_csc_spmv_empty(device, op, shape):
    M, N = shape
    ptr = torch.zeros(N + 1, dtype=I32, device=device)
    x = torch.randn(N if op == "non" else M, device=device)
    y = fs.flagsparse_spmv_csc(_e(device), _e(device, I32), ptr, x, shape, op=op)
    assert y.shape == ((M,) if op == "non" else (N,)) and torch.all(y == 0)


@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("shape", [(0, 5), (5, 0), (0, 0), (4, 6)])
def test_coo_spmv_empty(device, op, shape):
    M, N = shape
    x = torch.randn(N if op == "non" else M, device=device)
    y = fs.flagsparse_spmv_coo(_e(device), _e(device, I32), _e(device, I32), x, shape, op=op)
    assert y.shape == ((M,) if op == "non" else (N,)) and torch.all(y == 0)


@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("shape,n_dense", [((0, 5), 3), ((5, 0), 3), ((4, 6), 0), ((4, 6), 3)])
def test_coo_spmm_empty(device, op, shape, n_dense):
    M, K = shape
    B = torch.randn(K if op == "non" else M, n_dense, device=device)
    C = fs.flagsparse_spmm_coo(_e(device), _e(device, I32), _e(device, I32), B, shape, op=op)
    assert C.shape == ((M if op == "non" else K), n_dense) and torch.all(C == 0)
    p = fs.prepare_spmm_coo_route(_e(device), _e(device, I32), _e(device, I32), shape, op=op)
    C2 = fs.flagsparse_spmm_coo_run(p, B, dense_layout="col")
    assert C2.shape == C.shape and torch.all(C2 == 0)


def test_coo_spmm_zero_width_with_nonzeros(device):
    row = torch.tensor([0, 2, 2], dtype=I32, device=device)
    col = torch.tensor([1, 0, 3], dtype=I32, device=device)
    data = torch.tensor([1.0, 2.0, 3.0], device=device)
    C = fs.flagsparse_spmm_coo(data, row, col, torch.empty(4, 0, device=device), (3, 4))
    assert C.shape == (3, 0)


@pytest.mark.parametrize("fn", [fs.flagsparse_alpha_spmm_alg1, fs.flagsparse_alpha_spmm_alg1_tle_opt2])
def test_alpha_empty_and_zero_width(device, fn):
    ptr = torch.zeros(5, dtype=I32, device=device)
    C = fn(_e(device), _e(device, I32), ptr, torch.randn(6, 3, device=device), (4, 6))
    assert C.shape == (4, 3) and torch.all(C == 0)
    data = torch.tensor([1.0, -2.0], device=device)
    idx = torch.tensor([0, 5], dtype=I32, device=device)
    ptr = torch.tensor([0, 1, 1, 2, 2], dtype=I32, device=device)
    C = fn(data, idx, ptr, torch.empty(6, 0, device=device), (4, 6))
    assert C.shape == (4, 0)
    C = fn(_e(device), _e(device, I32), torch.zeros(1, dtype=I32, device=device), torch.randn(6, 3, device=device),
           (0, 6))
    assert C.shape == (0, 3)
