"""COO SpMM (flagsparse_spmm_coo and its route API) on the GPU vs the C oracle.

Mirrors tests/pytest/test_spmm_coo_accuracy.py of the reference: op in
{non, trans, conj} x value dtypes x int32 / int64 ids against the float64
product, the registered routes (coo_rowrun, coo_atomic, spmm_coo_alg1) with
row- and col-major dense layouts, plus duplicates (summed), validation
messages, ``out`` and the empty matrix.  The oracle orders the COO into a
CSR (stable, duplicates kept) and runs its CSR SpMM.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from oracle import oracle

from ._helpers import assert_rel, np_dtype_round

pytestmark = pytest.mark.gpu

SHAPES = [(8, 4, 16), (160, 200, 33), (300, 120, 64)]  # (M, K, N)
DTYPES = [torch.float32, torch.float64, torch.float16, torch.bfloat16]
ROUTES = ("coo_rowrun", "coo_atomic", "spmm_coo_alg1")


def _coo(M, K, nnz, seed):
    """Shuffled COO triples of an M x K matrix with some repeated pairs."""
    rng = np.random.default_rng(seed)
    row = rng.integers(0, M, nnz)
    col = rng.integers(0, K, nnz)
    k = nnz // 5
    row[:k], col[:k] = row[k: 2 * k], col[k: 2 * k]
    perm = rng.permutation(nnz)
    return row[perm], col[perm], rng.standard_normal(nnz) * 0.5


def _expected(row, col, v, b, M, K, op, dtype):
    ip, ix, vs = oracle.coo_to_csr(row, col, np_dtype_round(v, dtype), (M, K))
    return oracle.spmm(ip, ix, vs, np_dtype_round(b, dtype), (M, K), op="non" if op == "non" else "trans")


def _dev(row, col, v, dtype, index_dtype, device):
    return (torch.tensor(v, dtype=dtype, device=device), torch.tensor(row, dtype=index_dtype, device=device),
            torch.tensor(col, dtype=index_dtype, device=device))


@pytest.mark.parametrize("M,K,N", SHAPES)
@pytest.mark.parametrize("dtype", DTYPES, ids=["f32", "f64", "f16", "bf16"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmm_coo_matches_oracle(device, M, K, N, dtype, index_dtype, op):
    row, col, v = _coo(M, K, max(10, M * K // 8), seed=M + K + N)
    data, r, c = _dev(row, col, v, dtype, index_dtype, device)
    b = np.random.default_rng(N).standard_normal((K if op == "non" else M, N))
    C = fs.flagsparse_spmm_coo(data, r, c, torch.tensor(b, dtype=dtype, device=device), (M, K), op=op)
    assert C.shape == ((M if op == "non" else K), N) and C.dtype == dtype
    assert_rel(C, _expected(row, col, v, b, M, K, op, dtype), dtype, f"{op} {M}x{K} n{N}")


@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_route_api(device, op, layout):
    M, K, N = 120, 90, 24
    row, col, v = _coo(M, K, 1500, seed=5)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    b = np.random.default_rng(6).standard_normal((K if op == "non" else M, N))
    B = torch.tensor(b, dtype=torch.float32, device=device)
    ref = _expected(row, col, v, b, M, K, op, torch.float32)
    results = []
    for alg in ("auto",) + ROUTES + ("atomic", "alg1"):
        p = fs.prepare_spmm_coo_route(data, r, c, (M, K), op=op, alg=alg)
        assert p.op == op and p.output_dtype == torch.float32 and p.compute_dtype == torch.float64
        C, ms, meta = fs.flagsparse_spmm_coo_run(p, B, dense_layout=layout, return_time=True, return_meta=True,
                                                 timing=True, diagnostics=True)
        assert_rel(C, ref, torch.float32, f"{alg} {op} {layout}")
        assert meta["output_layout"] == layout and meta["dense_layout"] == layout and ms > 0
        assert meta["alg"] in ROUTES and meta["op"] == op and "diagnostics" in meta
        assert {"operator_ms", "gpu_ms", "process_cpu_ms", "process_gpu_ms", "compute_ms"} <= set(meta)
        results.append(C.contiguous())
    assert all(torch.equal(results[0], x) for x in results[1:])  # every route runs the same kernel
    p = fs.prepare_spmm_coo_route(data, r, c, (M, K), op=op)
    assert p.nnz == 1500 and p.n_segs == int(p.row_lengths.numel()) and int(p.row_lengths.sum()) == 1500
    assert torch.equal(fs.flagsparse_spmm_coo_run(p, B, alg="coo_atomic").contiguous(), results[0])


def test_registry_and_errors(device):
    assert fs.list_spmm_coo_algorithms() == ROUTES
    assert fs.list_spmm_coo_algorithms(dtype=torch.float16) == ("coo_rowrun", "coo_atomic")
    assert fs.resolve_spmm_coo_algorithm("auto", "non", torch.float32).name == "coo_rowrun"
    with pytest.raises(ValueError, match="unsupported COO SpMM algorithm"):
        fs.resolve_spmm_coo_algorithm("nope", "non", torch.float32)
    with pytest.raises(TypeError, match="does not support dtype"):
        fs.resolve_spmm_coo_algorithm("alg1", "non", torch.float16)
    row, col, v = _coo(30, 20, 100, seed=7)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    B = torch.randn(20, 8, device=device)
    bad = r.clone()
    bad[0] = 30
    with pytest.raises(IndexError, match="row indices out of range for n_rows"):
        fs.flagsparse_spmm_coo(data, bad, c, B, (30, 20))
    bad = c.clone()
    bad[1] = -2
    with pytest.raises(IndexError, match="col indices out of range for n_cols"):
        fs.flagsparse_spmm_coo(data, r, bad, B, (30, 20))
    with pytest.raises(ValueError, match=r"B.shape\[0\] must be n_cols=20, got 21"):
        fs.flagsparse_spmm_coo(data, r, c, torch.randn(21, 8, device=device), (30, 20))
    with pytest.raises(TypeError, match="B dtype must match data dtype"):
        fs.flagsparse_spmm_coo(data, r, c, B.double(), (30, 20))
    with pytest.raises(ValueError, match="block_n must be positive"):
        fs.flagsparse_spmm_coo(data, r, c, B, (30, 20), block_n=0)
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmm_coo(data, r, c, B, (30, 20), op="non", transpose=True)
    with pytest.raises(ValueError, match="op must be one of"):
        fs.flagsparse_spmm_coo(data, r, c, B, (30, 20), op="diag")
    with pytest.raises(ValueError, match="dense_layout must be one of"):
        fs.flagsparse_spmm_coo(data, r, c, B, (30, 20), dense_layout="diagonal")
    with pytest.raises(ValueError, match="out shape/dtype must match result"):
        fs.flagsparse_spmm_coo(data, r, c, B, (30, 20), out=torch.empty(30, 9, device=device))
    with pytest.raises(TypeError, match="data dtype must be one of"):
        fs.flagsparse_spmm_coo(data.to(torch.int32), r, c, B.to(torch.int32), (30, 20))
    with pytest.raises(TypeError, match="PreparedCooSpmmRoute"):
        fs.flagsparse_spmm_coo_run(object(), B)
    p = fs.prepare_spmm_coo_route(data, r, c, (30, 20))
    with pytest.raises(ValueError, match=r"B.shape\[0\] must be n_cols=20"):
        fs.flagsparse_spmm_coo_run(p, B[:5])
    with pytest.raises(TypeError, match="B dtype must match sparse matrix dtype"):
        fs.flagsparse_spmm_coo_run(p, B.half())
    with pytest.raises(TypeError):  # complex values: no HIP kernel
        fs.flagsparse_spmm_coo(data.to(torch.complex64), r, c, B.to(torch.complex64), (30, 20))


def test_out_meta_and_empty(device):
    row, col, v = _coo(50, 40, 300, seed=8)
    data, r, c = _dev(row, col, v, torch.float32, torch.int64, device)
    b = np.random.default_rng(9).standard_normal((40, 12))
    B = torch.tensor(b, dtype=torch.float32, device=device)
    out = torch.empty(50, 12, device=device)
    C, ms, meta = fs.flagsparse_spmm_coo(data, r, c, B, (50, 40), out=out, return_time=True, return_meta=True)
    assert C is out and ms >= 0 and meta["route"] == "rowrun" and meta["op"] == "non"
    assert set(meta) == {"op", "route", "symbolic_ms", "compute_ms", "op_total_ms", "dense_layout", "b_stride",
                         "c_stride", "output_layout"}
    assert_rel(out, _expected(row, col, v, b, 50, 40, "non", torch.float32), torch.float32)
    col_major = fs.flagsparse_spmm_coo(data, r, c, B, (50, 40), dense_layout="col")
    assert col_major.stride() == (1, 50) and torch.equal(col_major, out)
    e = torch.empty(0, device=device)
    ei = torch.empty(0, dtype=torch.int32, device=device)
    Z = fs.flagsparse_spmm_coo(e, ei, ei, torch.randn(6, 3, device=device), (4, 6))
    assert Z.shape == (4, 3) and torch.all(Z == 0)


def test_float32_is_computed_in_float64(device):
    """The reference computes float32 COO SpMM in float64 and rounds once
    (spmm_coo.py:146-153): every element is then within one float32 ulp of
    the float64 product rounded to float32 (float32 accumulation over these
    ~40-term rows would not be)."""
    M, K, N = 400, 300, 32
    row, col, v = _coo(M, K, 16000, seed=21)
    data, r, c = _dev(row, col, v, torch.float32, torch.int32, device)
    b = np.random.default_rng(22).standard_normal((K, N))
    C = fs.flagsparse_spmm_coo(data, r, c, torch.tensor(b, dtype=torch.float32, device=device), (M, K))
    ref32 = _expected(row, col, v, b, M, K, "non", torch.float32).astype(np.float32)
    got = C.cpu().numpy()
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    assert np.all(np.abs(got.astype(np.float64) - ref32.astype(np.float64)) <= ulp)
