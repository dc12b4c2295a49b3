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


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("n_dense", [1, 30, 64, 128])
def test_fused_backward_matches_two_kernels(device, dtype, index_dtype, n_dense):
    """dB and dvalues from one pass over the CSC: dB bitwise equal to the
    transposed SpMM, dvalues within tolerance of the SDDMM kernel and of the
    oracle.  A^T has long rows (columns 7, 900 and 1999 of A hold 2500
    nonzeros: CSC segments); 30 columns miss the 16-byte path."""
    import scipy.sparse as sp

    from ._helpers import long_row_matrix

    ip_t, ix_t, v_t = long_row_matrix(2000, 3000, [7, 900, 1999], 2500, 5, seed=n_dense)
    A = sp.csr_matrix((v_t, ix_t, ip_t), shape=(2000, 3000)).T.tocsr()
    A.sort_indices()
    ip, ix, v = A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    m, n = A.shape
    data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
    h = fs.CsrMatrix(idx, ptr, (m, n)).handle
    rng = np.random.default_rng(n_dense)
    Bn, dCn = rng.standard_normal((n, n_dense)), rng.standard_normal((m, n_dense))
    B = torch.tensor(Bn, dtype=dtype, device=device)
    dC = torch.tensor(dCn, dtype=dtype, device=device)
    fused = h.spmm_trans_sddmm(data, dC, B)
    assert fused is not None
    dB, dv = fused
    assert torch.equal(dB, h.spmm(data, dC, transpose=True))
    two = h.sddmm(dC, B)
    vr, Br, dCr = np_dtype_round(v, dtype), np_dtype_round(Bn, dtype), np_dtype_round(dCn, dtype)
    ref = oracle.sddmm(ip, ix, dCr, Br)
    assert_rel(dv, ref, dtype, "fused dvalues")
    assert_rel(two, ref, dtype, "sddmm dvalues")
    assert_rel(dB, oracle.spmm(ip, ix, vr, dCr, (m, n), "trans"), dtype, "fused dB")


def test_fused_backward_width_limit_takes_two_kernels(device):
    ip, ix, v = _graph(500, seed=4)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    mat = fs.CsrMatrix(idx, ptr, (500, 500))
    # 300 float32 columns exceed one lane group (64 lanes x 4): the two kernels run
    assert mat.handle.spmm_trans_sddmm(data, torch.randn(500, 300, device=device),
                                       torch.randn(500, 300, device=device)) is None
    vals = data.clone().requires_grad_(True)
    B = torch.randn(500, 300, device=device, requires_grad=True)
    dCn = np.random.default_rng(4).standard_normal((500, 300))
    fs.spmm_csr_autograd(vals, B, mat).backward(torch.tensor(dCn, dtype=torch.float32, device=device))
    Br = B.detach().double().cpu().numpy()
    assert_rel(vals.grad, oracle.sddmm(ip, ix, np_dtype_round(dCn, torch.float32), Br), torch.float32, "dvalues")
    assert_rel(B.grad, oracle.spmm(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(dCn, torch.float32),
                                   (500, 500), "trans"), torch.float32, "dB")
