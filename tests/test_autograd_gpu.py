"""Differentiable SpMM: FP64 gradcheck (finite differences) on a 512-node
power-law graph, gradients vs the oracle's trans-SpMM + SDDMM composition, and
BF16 forward+backward within 1e-2 of the FP64 oracle."""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import synthetic
from oracle import oracle

from ._helpers import assert_rel, device_csr, np_dtype_round

pytestmark = pytest.mark.gpu


def _graph(n, seed=3):
    return synthetic.power_law_graph(n, 2.3, 14, seed=seed)


def test_gradcheck_fp64_512_nodes(device):
    ip, ix, v = _graph(512)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float64)
    mat = fs.CsrMatrix(idx, ptr, (512, 512))
    vals = data.clone().requires_grad_(True)
    B = torch.randn(512, 8, dtype=torch.float64, device=device, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: fs.spmm_csr_autograd(a, b, mat), (vals, B), eps=1e-6, atol=1e-8,
                                    rtol=1e-6)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_gradients_match_oracle_composition(device, dtype):
    n = 4000
    ip, ix, v = _graph(n, seed=5)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    rng = np.random.default_rng(5)
    Bn, dCn = rng.standard_normal((n, 128)), rng.standard_normal((n, 128))
    mat = fs.CsrMatrix(idx, ptr, (n, n)).prepare_backward()
    vals = data.clone().requires_grad_(True)
    B = torch.tensor(Bn, dtype=dtype, device=device, requires_grad=True)
    C = fs.spmm_csr_autograd(vals, B, mat)
    C.backward(torch.tensor(dCn, dtype=dtype, device=device))
    vr, Br, dCr = np_dtype_round(v, dtype), np_dtype_round(Bn, dtype), np_dtype_round(dCn, dtype)
    assert_rel(C, oracle.spmm(ip, ix, vr, Br, (n, n)), dtype, "fwd")
    assert_rel(B.grad, oracle.spmm(ip, ix, vr, dCr, (n, n), "trans"), dtype, "dB")
    assert_rel(vals.grad, oracle.sddmm(ip, ix, dCr, Br), dtype, "dvalues")


def test_bf16_forward_backward_vs_fp64(device):
    n = 20000
    ip, ix, v = _graph(n, seed=3)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.bfloat16)
    rng = np.random.default_rng(3)
    Bn, dCn = rng.standard_normal((n, 128)), rng.standard_normal((n, 128))
    mat = fs.CsrMatrix(idx, ptr, (n, n))
    vals = data.clone().requires_grad_(True)
    B = torch.tensor(Bn, dtype=torch.bfloat16, device=device, requires_grad=True)
    C = fs.spmm_csr_autograd(vals, B, mat)
    C.backward(torch.tensor(dCn, dtype=torch.bfloat16, device=device))
    assert C.dtype == B.grad.dtype == vals.grad.dtype == torch.bfloat16
    assert_rel(C, oracle.spmm(ip, ix, v, Bn, (n, n)), torch.bfloat16, "bf16 fwd")
    assert_rel(B.grad, oracle.spmm(ip, ix, v, dCn, (n, n), "trans"), torch.bfloat16, "bf16 dB")
    assert_rel(vals.grad, oracle.sddmm(ip, ix, dCn, Bn), torch.bfloat16, "bf16 dvalues")


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16], ids=["f32", "bf16"])
def test_full_size_graph_forward_backward(device, dtype):
    """BASELINE config 4 at its size: the 200000-node power-law graph (exponent
    2.3, mean degree 14, D^-1/2 A D^-1/2 values, seed 3), B and dC of
    200000 x 128.  float32: forward, dB = A^T dC and dvalues = SDDMM(dC, B)
    within 1e-5 of the oracle on the float32-rounded inputs; bfloat16: within
    1e-2 of the FP64 oracle on the unrounded inputs."""
    n = 200000
    ip, ix, v = _graph(n, seed=3)
    rng = np.random.Generator(np.random.PCG64(3))
    Bn, dCn = rng.standard_normal((n, 128)), rng.standard_normal((n, 128))
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    mat = fs.CsrMatrix(idx, ptr, (n, n)).prepare_backward()
    vals = data.clone().requires_grad_(True)
    B = torch.tensor(Bn, dtype=dtype, device=device, requires_grad=True)
    C = fs.spmm_csr_autograd(vals, B, mat)
    C.backward(torch.tensor(dCn, dtype=dtype, device=device))
    if dtype == torch.float32:
        vr, Br, dCr = np_dtype_round(v, dtype), np_dtype_round(Bn, dtype), np_dtype_round(dCn, dtype)
    else:
        vr, Br, dCr = v, Bn, dCn
    assert_rel(C, oracle.spmm(ip, ix, vr, Br, (n, n)), dtype, "fwd")
    assert_rel(B.grad, oracle.spmm(ip, ix, vr, dCr, (n, n), "trans"), dtype, "dB")
    assert_rel(vals.grad, oracle.sddmm(ip, ix, dCr, Br), dtype, "dvalues")


def test_only_requested_gradients(device):
    ip, ix, v = _graph(300, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    mat = fs.CsrMatrix(idx, ptr, (300, 300))
    B = torch.randn(300, 16, device=device, requires_grad=True)
    fs.spmm_csr_autograd(data, B, mat).sum().backward()
    assert B.grad is not None
    assert mat.handle._csc is not None  # dB ran over the CSC permutation
    mat2 = fs.CsrMatrix(idx, ptr, (300, 300))
    vals = data.clone().requires_grad_(True)
    fs.spmm_csr_autograd(vals, torch.randn(300, 16, device=device), mat2).sum().backward()
    assert vals.grad is not None and mat2.handle._csc is None  # no CSC needed without dB


def test_rejects_mismatched_values_and_devices(device):
    ip, ix, v = _graph(200, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    mat = fs.CsrMatrix(idx, ptr, (200, 200))
    B = torch.randn(200, 8, device=device)
    with pytest.raises(ValueError, match="one entry per nonzero"):
        fs.spmm_csr_autograd(data[:-1], B, mat)
    with pytest.raises(ValueError, match="one entry per nonzero"):
        fs.spmm_csr_autograd(data.reshape(1, -1), B, mat)
    with pytest.raises(ValueError, match="matrix's device"):
        fs.spmm_csr_autograd(data, B.cpu(), mat)
    with pytest.raises(ValueError, match="matrix's device"):
        fs.spmm_csr_autograd(data.cpu(), B, mat)
