"""Native COO SpMV (flagsparse_spmv_coo / prepare_spmv_coo) on the GPU vs the
C oracle.

Mirrors tests/pytest/test_spmv_coo_accuracy.py of the reference: op in
{non, trans, conj} x float32 / float64 x int32 / int64 ids on its shape grid,
one prepared object serving every op, the prepared transpose conflict and
the int64 -> int32 fallback policy; adds unsorted input with duplicates
(summed), sort_by_row=False, the row-run offsets, validation messages and
``out``.  The oracle converts the COO to CSR (stable, duplicates kept) and
runs its CSR SpMV.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import spmv_coo as coo_mod
from oracle import oracle

from ._helpers import assert_rel, np_dtype_round

pytestmark = pytest.mark.gpu

SHAPES = [(1, 32), (160, 1024), (128, 256), (700, 500)]
DTYPES = [torch.float32, torch.float64]


def _coo(M, N, nnz, seed, dup=True):
    """Shuffled COO triples with some repeated (row, col) pairs."""
    rng = np.random.default_rng(seed)
    row = rng.integers(0, M, nnz)
    col = rng.integers(0, N, nnz)
    k = nnz // 4
    if dup and k > 0:
        row[:k] = row[k: 2 * k]
        col[:k] = col[k: 2 * k]
    perm = rng.permutation(nnz)
    return row[perm], col[perm], rng.standard_normal(nnz)


def _dev(row, col, v, dtype, index_dtype, device):
    return (torch.tensor(v, dtype=dtype, device=device), torch.tensor(row, dtype=index_dtype, device=device),
            torch.tensor(col, dtype=index_dtype, device=device))


def _expected(row, col, v, x, M, N, op, dtype):
    ip, ix, vs = oracle.coo_to_csr(row, col, np_dtype_round(v, dtype), (M, N))
    return oracle.spmv(ip, ix, vs, np_dtype_round(x, dtype), (M, N), op="non" if op == "non" else "trans")


@pytest.mark.parametrize("M,N", SHAPES)
@pytest.mark.parametrize("dtype", DTYPES, ids=["float32", "float64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmv_coo_matches_oracle(device, M, N, dtype, index_dtype, op):
    row, col, v = _coo(M, N, max(8, M * N // 12), seed=M + N)
    data, r, c = _dev(row, col, v, dtype, index_dtype, device)
    x = np.random.default_rng(M * N).standard_normal(N if op == "non" else M)
    y = fs.flagsparse_spmv_coo(data, r, c, torch.tensor(x, dtype=dtype, device=device), (M, N), op=op)
    assert y.shape == ((M,) if op == "non" else (N,)) and y.dtype == dtype
    assert_rel(y, _expected(row, col, v, x, M, N, op, dtype), dtype, f"{op} {M}x{N}")


def test_prepared_serves_every_op(device):
    M, N = 90, 70
    row, col, v = _coo(M, N, 900, seed=3)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    prepared = fs.prepare_spmv_coo(data, r, c, (M, N))
    assert prepared.op == 0 and not prepared.transpose and prepared.nnz == 900
    for op in ("non", "trans", "conj"):
        x = np.random.default_rng(len(op)).standard_normal(N if op == "non" else M)
        y = fs.flagsparse_spmv_coo(x=torch.tensor(x, dtype=torch.float32, device=device), prepared=prepared, op=op)
        assert_rel(y, _expected(row, col, v, x, M, N, op, torch.float32), torch.float32, op)
    # transpose=True without op selects trans at run time
    xt = np.random.default_rng(9).standard_normal(M)
    y = fs.flagsparse_spmv_coo(x=torch.tensor(xt, dtype=torch.float32, device=device), prepared=prepared,
                               transpose=True)
    assert_rel(y, _expected(row, col, v, xt, M, N, "trans", torch.float32), torch.float32)


def test_prepared_explicit_transpose_conflict_rejected(device):
    row, col, v = _coo(8, 10, 20, seed=4)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    prepared = fs.prepare_spmv_coo(data, r, c, (8, 10), transpose=True)
    assert prepared.op == 1
    x = torch.randn(10, device=device)
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmv_coo(x=x, prepared=prepared, op="non", transpose=True)
    with pytest.raises(ValueError, match="does not match prepared.shape"):
        fs.flagsparse_spmv_coo(x=x, prepared=prepared, shape=(10, 8), op="non")
    with pytest.raises(TypeError, match="x is required when prepared is set"):
        fs.flagsparse_spmv_coo(prepared=prepared)
    with pytest.raises(ValueError, match="transpose=True conflicts with op=non"):
        fs.prepare_spmv_coo(data, r, c, (8, 10), transpose=True, op="non")


def test_sorted_slots_and_row_runs(device):
    M, N = 40, 30
    row, col, v = _coo(M, N, 300, seed=5)
    data, r, c = _dev(row, col, v, torch.float64, torch.int64, device)
    p = fs.prepare_spmv_coo(data, r, c, (M, N))
    order = np.lexsort((col, row))  # stable (row, col) order
    assert np.array_equal(p.row_non.cpu().numpy(), row[order])
    assert np.array_equal(p.col_non.cpu().numpy(), col[order])
    assert np.array_equal(p.data_non.cpu().numpy(), v[order])  # values moved, never recomputed
    starts = np.concatenate([[0], np.nonzero(np.diff(row[order]))[0] + 1, [row.size]])
    assert np.array_equal(p.seg_starts_non.cpu().numpy(), starts)
    order_t = np.lexsort((row, col))
    assert np.array_equal(p.row_trans.cpu().numpy(), col[order_t])
    assert p.row_non.dtype == torch.int64
    unsorted = fs.prepare_spmv_coo(data, r, c, (M, N), sort_by_row=False)
    assert torch.equal(unsorted.row_non, r) and unsorted.seg_starts_non is None
    x = torch.randn(N, dtype=torch.float64, device=device)
    assert torch.equal(fs.flagsparse_spmv_coo(x=x, prepared=unsorted), fs.flagsparse_spmv_coo(x=x, prepared=p))


def test_validation(device):
    row, col, v = _coo(20, 16, 60, seed=6)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    x = torch.randn(16, device=device)
    bad_r = r.clone()
    bad_r[3] = 20
    with pytest.raises(IndexError, match="row indices out of range"):
        fs.flagsparse_spmv_coo(data, bad_r, c, x, (20, 16))
    bad_c = c.clone()
    bad_c[5] = -1
    with pytest.raises(IndexError, match="col indices out of range"):
        fs.flagsparse_spmv_coo(data, r, bad_c, x, (20, 16))
    with pytest.raises(TypeError, match="COO SpMV supports float32, float64, complex64, and complex128"):
        fs.flagsparse_spmv_coo(data.half(), r, c, x.half(), (20, 16))
    with pytest.raises(ValueError, match="same length"):
        fs.flagsparse_spmv_coo(data[:-1], r, c, x, (20, 16))
    with pytest.raises(TypeError, match="row dtype must be"):
        fs.flagsparse_spmv_coo(data, r.to(torch.int16), c, x, (20, 16))
    with pytest.raises(ValueError, match="x length must be n_cols=16, got 20"):
        fs.flagsparse_spmv_coo(data, r, c, torch.randn(20, device=device), (20, 16))
    with pytest.raises(ValueError, match="num_warps must be a power of 2"):
        fs.flagsparse_spmv_coo(data, r, c, x, (20, 16), num_warps=3)
    with pytest.raises(ValueError, match="block_inner must be a positive power of 2"):
        fs.flagsparse_spmv_coo(data, r, c, x, (20, 16), block_inner=96)
    with pytest.raises(ValueError, match="required when prepared is None"):
        fs.flagsparse_spmv_coo(data, r, c, None, (20, 16))
    with pytest.raises(ValueError, match="out shape/dtype must match result"):
        fs.flagsparse_spmv_coo(data, r, c, x, (20, 16), out=torch.empty(16, device=device))
    with pytest.raises(TypeError):  # complex values: no HIP kernel
        fs.flagsparse_spmv_coo(data.to(torch.complex64), r, c, x.to(torch.complex64), (20, 16))


def test_out_time_and_empty(device):
    row, col, v = _coo(50, 40, 400, seed=7)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    x = np.random.default_rng(8).standard_normal(40)
    xt = torch.tensor(x, dtype=torch.float32, device=device)
    out = torch.empty(50, device=device)
    got, ms = fs.flagsparse_spmv_coo(data, r, c, xt, (50, 40), out=out, return_time=True)
    assert got is out and ms >= 0
    assert_rel(out, _expected(row, col, v, x, 50, 40, "non", torch.float32), torch.float32)
    e = torch.empty(0, device=device)
    ei = torch.empty(0, dtype=torch.int32, device=device)
    y = fs.flagsparse_spmv_coo(e, ei, ei, torch.randn(6, device=device), (4, 6))
    assert y.shape == (4,) and torch.all(y == 0)


def test_int64_auto_fallback_to_int32(device, monkeypatch):
    row, col, v = _coo(12, 9, 40, seed=10)
    data, r, c = _dev(row, col, v, torch.float32, torch.int64, device)
    x = np.random.default_rng(11).standard_normal(9)
    state = {"forced_once": False}
    original = coo_mod._hip_spmv_coo_kernel

    def fail_int64_once(launch, x_in, block_size, num_warps, block_inner):
        if launch.row.dtype == torch.int64 and not state["forced_once"]:
            state["forced_once"] = True
            raise RuntimeError("forced int64 launch failure")
        return original(launch, x_in, block_size, num_warps, block_inner)

    monkeypatch.setattr(coo_mod, "_hip_spmv_coo_kernel", fail_int64_once)
    y = fs.flagsparse_spmv_coo(data, r, c, torch.tensor(x, dtype=torch.float32, device=device), (12, 9),
                               index_fallback_policy="auto")
    assert state["forced_once"]
    assert_rel(y, _expected(row, col, v, x, 12, 9, "non", torch.float32), torch.float32)


def test_int64_strict_no_fallback(device, monkeypatch):
    row, col, v = _coo(12, 9, 40, seed=12)
    data, r, c = _dev(row, col, v, torch.float32, torch.int64, device)

    def fail(launch, x_in, block_size, num_warps, block_inner):
        raise RuntimeError("forced int64 launch failure")

    monkeypatch.setattr(coo_mod, "_hip_spmv_coo_kernel", fail)
    with pytest.raises(RuntimeError, match="forced int64 launch failure"):
        fs.flagsparse_spmv_coo(data, r, c, torch.randn(9, device=device), (12, 9), index_fallback_policy="strict")
