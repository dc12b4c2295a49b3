"""SDDMM parity on the GPU vs the C oracle.

Mirrors tests/pytest/test_spgemm_sddmm_accuracy.py (SDDMM half) of the
reference: (M,N,K) in {(8,10,8),(64,96,16)}, alpha=1.25, beta=0.5, indptr
int32/int64, dtype rejections; adds K widths that miss the vector path or
need several K chunks, long rows, K=0 and the bf16 handle path used by the
SpMM gradient.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import synthetic
from flagsparse_amd._matrix import CsrHandle
from oracle import oracle

from ._helpers import assert_rel, device_csr, long_row_matrix, np_dtype_round

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(8, 10, 8), (64, 96, 16)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["float32", "float64"])
@pytest.mark.parametrize("indptr_dtype", [torch.int32, torch.int64], ids=["ptr32", "ptr64"])
def test_sddmm_matches_sampled_dense_reference(device, M, N, K, dtype, indptr_dtype):
    ip, ix, v = synthetic.random_csr(M, N, 0.3, seed=M + N + K)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype, torch.int32, indptr_dtype)
    rng = np.random.default_rng(K)
    X = rng.standard_normal((M, K)) * 0.125
    Y = rng.standard_normal((N, K)) * 0.125
    got = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr, x=torch.tensor(X, dtype=dtype, device=device),
                                  y=torch.tensor(Y, dtype=dtype, device=device), shape=(M, N), alpha=1.25, beta=0.5)
    ref = oracle.sddmm(ip, ix, np_dtype_round(X, dtype), np_dtype_round(Y, dtype), np_dtype_round(v, dtype), 1.25,
                       0.5)
    assert_rel(got, ref, dtype)


@pytest.mark.parametrize("K", [1, 3, 4, 5, 64, 128, 256, 257, 600, 1024])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_sddmm_k_widths(device, K, dtype):
    ip, ix, v = synthetic.random_csr(50, 40, 0.2, seed=K)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    rng = np.random.default_rng(K)
    X, Y = rng.standard_normal((50, K)), rng.standard_normal((40, K))
    got = fs.flagsparse_sddmm_csr(indices=idx, indptr=ptr, x=torch.tensor(X, dtype=dtype, device=device),
                                  y=torch.tensor(Y, dtype=dtype, device=device), shape=(50, 40))
    assert_rel(got, oracle.sddmm(ip, ix, np_dtype_round(X, dtype), np_dtype_round(Y, dtype)), dtype, f"K={K}")


def test_sddmm_long_rows(device):
    ip, ix, v = long_row_matrix(40, 9000, [3, 39], 8000, 5, seed=3)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    rng = np.random.default_rng(3)
    X, Y = rng.standard_normal((40, 128)), rng.standard_normal((9000, 128))
    got = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr,
                                  x=torch.tensor(X, dtype=torch.float32, device=device),
                                  y=torch.tensor(Y, dtype=torch.float32, device=device), shape=(40, 9000), alpha=2.0,
                                  beta=-1.0)
    ref = oracle.sddmm(ip, ix, np_dtype_round(X, torch.float32), np_dtype_round(Y, torch.float32),
                       np_dtype_round(v, torch.float32), 2.0, -1.0)
    assert_rel(got, ref, torch.float32)


def test_sddmm_k_zero_and_empty(device):
    ip, ix, v = synthetic.random_csr(8, 10, 0.3, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    out = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr, x=torch.empty(8, 0, device=device),
                                  y=torch.empty(10, 0, device=device), shape=(8, 10), beta=0.5)
    assert torch.allclose(out, data * 0.5)
    e = fs.flagsparse_sddmm_csr(indices=torch.empty(0, dtype=torch.int32, device=device),
                                indptr=torch.zeros(4, dtype=torch.int32, device=device),
                                x=torch.randn(3, 5, device=device), y=torch.randn(2, 5, device=device), shape=(3, 2))
    assert e.shape == (0,)


def test_sddmm_rejects_unsupported_index_and_value_dtypes(device):
    ip, ix, v = synthetic.random_csr(8, 10, 0.3, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    x = torch.randn(8, 4, device=device)
    y = torch.randn(10, 4, device=device)
    with pytest.raises(TypeError, match="indices dtype must be torch.int32"):
        fs.flagsparse_sddmm_csr(data=data, indices=idx.long(), indptr=ptr, x=x, y=y, shape=(8, 10))
    with pytest.raises(TypeError, match="x dtype must be torch.float32 or torch.float64"):
        fs.flagsparse_sddmm_csr(data=data.to(torch.complex64), indices=idx, indptr=ptr, x=x.to(torch.complex64),
                                y=y.to(torch.complex64), shape=(8, 10))
    with pytest.raises(ValueError, match="data is required when beta is non-zero"):
        fs.flagsparse_sddmm_csr(indices=idx, indptr=ptr, x=x, y=y, shape=(8, 10), beta=1.0)
    with pytest.raises(ValueError, match="x and y must have the same K dimension"):
        fs.flagsparse_sddmm_csr(indices=idx, indptr=ptr, x=x, y=torch.randn(10, 5, device=device), shape=(8, 10))


def test_sddmm_prepared_reuse_and_meta(device):
    ip, ix, v = synthetic.random_csr(64, 96, 0.2, seed=5)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = fs.prepare_sddmm_csr(idx, ptr, (64, 96), k_hint=16)
    x, y = torch.randn(64, 16, device=device), torch.randn(96, 16, device=device)
    a, ms, meta = fs.flagsparse_sddmm_csr(x=x, y=y, prepared=prepared, return_time=True, return_meta=True)
    b = fs.flagsparse_sddmm_csr(x=x, y=y, prepared=prepared)
    assert torch.equal(a, b) and ms >= 0 and meta["fallback_used"] is False


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16], ids=["bf16", "f16"])
def test_sddmm_half_storage_via_handle(device, dtype):
    ip, ix, v = synthetic.power_law_graph(3000, 2.3, 14, seed=3)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    rng = np.random.default_rng(4)
    X, Y = rng.standard_normal((3000, 128)), rng.standard_normal((3000, 128))
    h = CsrHandle(idx, ptr, (3000, 3000))
    got = h.sddmm(torch.tensor(X, dtype=dtype, device=device), torch.tensor(Y, dtype=dtype, device=device))
    assert got.dtype == dtype
    assert_rel(got, oracle.sddmm(ip, ix, X, Y), dtype)


def test_sddmm_strided_out_and_host_data(device):
    """A strided `out` receives the result without touching the elements
    between its own; values on the host are rejected before any launch."""
    ip, ix, v = synthetic.random_csr(40, 30, 0.2, seed=21)
    data, idx, ptr = device_csr(ip, ix, v, device)
    rng = np.random.default_rng(22)
    X = torch.tensor(rng.standard_normal((40, 8)), dtype=torch.float32, device=device)
    Y = torch.tensor(rng.standard_normal((30, 8)), dtype=torch.float32, device=device)
    want = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr, x=X, y=Y, shape=(40, 30), beta=0.5)
    buf = torch.full((2 * data.numel(),), 7.0, device=device)
    got = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr, x=X, y=Y, shape=(40, 30), beta=0.5,
                                  out=buf[::2])
    assert got.data_ptr() == buf.data_ptr() and torch.equal(buf[::2], want)
    assert torch.all(buf[1::2] == 7.0)
    with pytest.raises(ValueError, match="data must be a CUDA tensor"):
        fs.flagsparse_sddmm_csr(data=data.cpu(), indices=idx, indptr=ptr, x=X, y=Y, shape=(40, 30), beta=0.5)
