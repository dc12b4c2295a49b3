"""CSC SpMV on the GPU vs the C oracle.

Mirrors tests/pytest/test_spmv_csc_accuracy.py of the reference: op in
{non, trans, conj} x float32 / float64 x int32 / int64 indices on its shape
grid, the prepared path, op / transpose mismatches against a prepared
object and the int64 -> int32 fallback policy (auto retries, strict
raises); adds a long column, validation messages and the return forms.
The CSC arrays of A are the CSR arrays of A^T, so the oracle's CSR SpMV of
(n_cols x n_rows) with the opposite op is the expected result.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import spmv_csr as spmv_mod
from flagsparse_amd import synthetic
from oracle import oracle

from ._helpers import assert_rel, device_csr, long_row_matrix, np_dtype_round

pytestmark = pytest.mark.gpu

SHAPES = [(1, 32), (160, 1024), (128, 256), (700, 500)]
DTYPES = [torch.float32, torch.float64]


def _csc(M, N, dtype, index_dtype, seed, device):
    """CSC arrays of a random M x N matrix (= CSR arrays of its N x M transpose)."""
    cp, ri, v = synthetic.random_csr(N, M, 0.1, seed=seed)
    return (cp, ri, v), device_csr(cp, ri, v, device, dtype, index_dtype, index_dtype)


def _expected(host, x, M, N, op, dtype):
    cp, ri, v = host
    csr_op = "non" if op in ("trans", "conj") else "trans"
    return oracle.spmv(cp, ri, np_dtype_round(v, dtype), np_dtype_round(x, dtype), (N, M), op=csr_op)


@pytest.mark.parametrize("M,N", SHAPES)
@pytest.mark.parametrize("dtype", DTYPES, ids=["float32", "float64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmv_csc_matches_oracle(device, M, N, dtype, index_dtype, op):
    host, (data, idx, ptr) = _csc(M, N, dtype, index_dtype, M + N, device)
    x = np.random.default_rng(M * N).standard_normal(M if op != "non" else N)
    y = fs.flagsparse_spmv_csc(data, idx, ptr, torch.tensor(x, dtype=dtype, device=device), (M, N), op=op)
    assert y.shape == ((N,) if op != "non" else (M,)) and y.dtype == dtype
    assert_rel(y, _expected(host, x, M, N, op, dtype), dtype, f"{op} {M}x{N}")


@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmv_csc_prepared_path(device, op):
    M, N = 96, 80
    host, (data, idx, ptr) = _csc(M, N, torch.float32, torch.int32, 7, device)
    prepared = fs.prepare_spmv_csc(data, idx, ptr, (M, N), op=op)
    assert prepared.op == {"non": 0, "trans": 1, "conj": 2}[op] and prepared.transpose == (op != "non")
    for seed in (1, 2):  # the prepared object serves repeated calls
        x = np.random.default_rng(seed).standard_normal(M if op != "non" else N)
        y = fs.flagsparse_spmv_csc(x=torch.tensor(x, dtype=torch.float32, device=device), prepared=prepared)
        assert_rel(y, _expected(host, x, M, N, op, torch.float32), torch.float32)


def test_spmv_csc_prepared_mismatches_rejected(device):
    _, (data, idx, ptr) = _csc(20, 16, torch.float32, torch.int32, 3, device)
    prepared = fs.prepare_spmv_csc(data, idx, ptr, (20, 16), op="non")
    x = torch.randn(16, device=device)
    with pytest.raises(ValueError, match="does not match prepared.transpose"):
        fs.flagsparse_spmv_csc(x=x, prepared=prepared, transpose=True)
    with pytest.raises(ValueError, match="does not match prepared.op"):
        fs.flagsparse_spmv_csc(x=x, prepared=prepared, op="trans")
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmv_csc(x=x, prepared=prepared, op="non", transpose=True)
    with pytest.raises(ValueError, match="transpose=True conflicts with op=non"):
        fs.prepare_spmv_csc(data, idx, ptr, (20, 16), op="non", transpose=True)
    with pytest.raises(TypeError, match="PreparedCscSpmv"):
        fs.flagsparse_spmv_csc(x=x, prepared=object())


def test_spmv_csc_validation(device):
    _, (data, idx, ptr) = _csc(20, 16, torch.float32, torch.int32, 4, device)
    x = torch.randn(16, device=device)
    with pytest.raises(ValueError, match="indptr length must be n_cols\\+1=17"):
        fs.flagsparse_spmv_csc(data, idx, ptr[:-1], x, (20, 16))
    bad = idx.clone()
    bad[0] = 20
    with pytest.raises(IndexError, match="indices out of range for n_rows"):
        fs.flagsparse_spmv_csc(data, bad, ptr, x, (20, 16))
    shifted = ptr.clone()
    shifted[0] = 1
    with pytest.raises(ValueError, match="indptr must start at zero"):
        fs.flagsparse_spmv_csc(data, idx, shifted, x, (20, 16))
    with pytest.raises(TypeError, match="CSC SpMV supports float32, float64, complex64, and complex128"):
        fs.flagsparse_spmv_csc(data.half(), idx, ptr, x.half(), (20, 16))
    with pytest.raises(ValueError, match="x length must be 16, got 20"):
        fs.flagsparse_spmv_csc(data, idx, ptr, torch.randn(20, device=device), (20, 16))
    with pytest.raises(TypeError, match="x dtype must match"):
        fs.flagsparse_spmv_csc(data, idx, ptr, x.double(), (20, 16))
    with pytest.raises(ValueError, match="op must be one of"):
        fs.flagsparse_spmv_csc(data, idx, ptr, x, (20, 16), op="sideways")
    with pytest.raises(ValueError, match="index_fallback_policy"):
        fs.flagsparse_spmv_csc(data, idx, ptr, x, (20, 16), index_fallback_policy="sometimes")
    with pytest.raises(TypeError):  # complex values: no HIP kernel
        fs.flagsparse_spmv_csc(data.to(torch.complex64), idx, ptr, torch.randn(20, device=device).to(torch.complex64),
                               (20, 16), op="trans")


def test_spmv_csc_long_column_and_empty(device):
    # a 9000-nonzero column of the 10000 x 64 matrix (a row of its transpose
    # longer than the 4096-nonzero segment length)
    cp, ri, v = long_row_matrix(64, 10000, long_rows=(5,), long_len=9000, short_len=6, seed=8)
    data, idx, ptr = device_csr(cp, ri, v, device)
    host = (cp, ri, v)
    for op, n_x in (("non", 64), ("trans", 10000)):
        x = np.random.default_rng(9).standard_normal(n_x)
        y = fs.flagsparse_spmv_csc(data, idx, ptr, torch.tensor(x, dtype=torch.float32, device=device),
                                   (10000, 64), op=op)
        assert_rel(y, _expected(host, x, 10000, 64, op, torch.float32), torch.float32, f"long column {op}")
    empty_ptr = torch.zeros(9, dtype=torch.int32, device=device)
    e = torch.empty(0, device=device)
    ei = torch.empty(0, dtype=torch.int32, device=device)
    y = fs.flagsparse_spmv_csc(e, ei, empty_ptr, torch.randn(8, device=device), (5, 8))
    assert y.shape == (5,) and torch.all(y == 0)


def test_spmv_csc_returns_and_out(device):
    host, (data, idx, ptr) = _csc(50, 40, torch.float64, torch.int32, 5, device)
    x = np.random.default_rng(6).standard_normal(40)
    xt = torch.tensor(x, device=device)
    y, ms, meta = fs.flagsparse_spmv_csc(data, idx, ptr, xt, (50, 40), return_time=True, return_meta=True)
    assert ms >= 0 and set(meta) == {"op", "symbolic_ms", "compute_ms", "op_total_ms", "index_fallback_applied",
                                     "index_fallback_reason"}
    assert meta["op"] == "non" and meta["symbolic_ms"] == 0.0 and meta["index_fallback_applied"] is False
    assert_rel(y, _expected(host, x, 50, 40, "non", torch.float64), torch.float64)
    out = torch.empty(50, dtype=torch.float64, device=device)
    got = fs.flagsparse_spmv_csc(data, idx, ptr, xt, (50, 40), out=out)
    assert got is out and torch.equal(out, y)
    with pytest.raises(ValueError, match="out shape/dtype must match result"):
        fs.flagsparse_spmv_csc(data, idx, ptr, xt, (50, 40), out=torch.empty(40, dtype=torch.float64, device=device))


def test_spmv_csc_int64_auto_fallback_to_int32(device, monkeypatch):
    host, (data, idx, ptr) = _csc(12, 9, torch.float32, torch.int64, 4, device)
    x = np.random.default_rng(4).standard_normal(12)
    state = {"forced_once": False}
    original = spmv_mod._hip_spmv_csr_impl_prepared

    def fail_int64_once(prepared, x_in, out=None):
        if prepared.kernel_indices.dtype == torch.int64 and not state["forced_once"]:
            state["forced_once"] = True
            raise RuntimeError("forced int64 launch failure")
        return original(prepared, x_in, out)

    monkeypatch.setattr(spmv_mod, "_hip_spmv_csr_impl_prepared", fail_int64_once)
    y = fs.flagsparse_spmv_csc(data, idx, ptr, torch.tensor(x, dtype=torch.float32, device=device), (12, 9),
                               op="trans", index_fallback_policy="auto")
    assert state["forced_once"]
    assert_rel(y, _expected(host, x, 12, 9, "trans", torch.float32), torch.float32)


def test_spmv_csc_int64_strict_no_fallback(device, monkeypatch):
    _, (data, idx, ptr) = _csc(12, 9, torch.float32, torch.int64, 4, device)

    def fail_int64(prepared, x_in, out=None):
        raise RuntimeError("forced int64 launch failure")

    monkeypatch.setattr(spmv_mod, "_hip_spmv_csr_impl_prepared", fail_int64)
    with pytest.raises(RuntimeError, match="forced int64 launch failure"):
        fs.flagsparse_spmv_csc(data, idx, ptr, torch.randn(12, device=device), (12, 9), op="trans",
                               index_fallback_policy="strict")
