"""SpMM parity on the GPU: HIP kernels vs the C oracle (FP64 golden).

Mirrors tests/pytest/test_spmm_csr_accuracy.py of the reference ((M,N,K) in
{(8,4,16),(64,16,96)}, values x0.125, op x index-dtype grid, return_meta
keys, validation messages) with fixed seeds, and adds dense widths that miss
the 16-byte vector path, long rows (segment split), the 262144^2 power-law
workload at full size (non and trans), the opt and route APIs, col-major
layout, bf16/f16 storage and bitwise determinism.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import flagsparse_amd as fs
from flagsparse_amd import synthetic
from flagsparse_amd._matrix import CsrHandle
from oracle import oracle

from ._helpers import REL_TOL, assert_rel, device_csr, long_row_matrix, np_dtype_round

pytestmark = pytest.mark.gpu

MNK = [(8, 4, 16), (64, 16, 96)]


def _case(M, K, seed, density=None):
    d = density if density is not None else min(0.25, max(0.06, 32.0 / max(M * K, 1)))
    return synthetic.random_csr(M, K, d, seed=seed)


@pytest.mark.parametrize("M,N,K", MNK)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["float32", "float64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["int32", "int64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmm_op_matches_oracle(device, M, N, K, dtype, index_dtype, op):
    ip, ix, v = _case(M, K, seed=M * 7 + N * 3 + K)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
    eff = M if op != "non" else K
    B = np.random.default_rng(M + K).standard_normal((eff, N)) * 0.125
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(B, dtype=dtype, device=device), (M, K), op=op)
    assert C.shape == ((K if op != "non" else M), N) and C.dtype == dtype
    ref = oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(B, dtype), (M, K), op)
    assert_rel(C, ref, dtype, f"spmm {op}")


@pytest.mark.parametrize("N", [1, 2, 3, 5, 17, 31, 64, 100, 128, 256, 300, 513])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
def test_spmm_dense_widths(device, N, dtype):
    ip, ix, v = _case(97, 83, seed=N, density=0.1)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    B = np.random.default_rng(N).standard_normal((83, N))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(B, dtype=dtype, device=device), (97, 83))
    ref = oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(B, dtype), (97, 83))
    assert_rel(C, ref, dtype, f"N={N}")


def test_spmm_strided_dense_operand(device):
    ip, ix, v = _case(50, 40, seed=3, density=0.2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    Bfull = torch.randn(40, 70, device=device)
    B = Bfull[:, 3:67]  # non-contiguous view: made contiguous by the API
    C = fs.flagsparse_spmm_csr(data, idx, ptr, B, (50, 40))
    ref = oracle.spmm(ip, ix, np_dtype_round(v, torch.float32), B.double().cpu().numpy(), (50, 40))
    assert_rel(C, ref, torch.float32)


@pytest.mark.parametrize("long_len", [257, 1536, 5000])
@pytest.mark.parametrize("op", ["non", "trans"])
def test_spmm_long_rows(device, long_len, op):
    ip, ix, v = long_row_matrix(300, 6000, [0, 17, 299], long_len, 3, seed=long_len)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    eff = 300 if op == "trans" else 6000
    B = np.random.default_rng(long_len).standard_normal((eff, 64))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(B, dtype=torch.float32, device=device), (300, 6000),
                               op=op)
    ref = oracle.spmm(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(B, torch.float32), (300, 6000), op)
    assert_rel(C, ref, torch.float32, f"long {long_len}")


def test_spmm_single_long_row_1536(device):
    # the reference's long-row API case: one row with 1536 nonzeros
    ip, ix, v = long_row_matrix(1, 4096, [0], 1536, 0, seed=9)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    B = np.random.default_rng(9).standard_normal((4096, 32))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(B, dtype=torch.float32, device=device), (1, 4096))
    assert_rel(C, oracle.spmm(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(B, torch.float32),
                              (1, 4096)), torch.float32)


@pytest.mark.slow
@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("dtype,index_dtype", [(torch.float32, torch.int32), (torch.float32, torch.int64),
                                               (torch.float64, torch.int32)], ids=["f32i32", "f32i64", "f64i32"])
def test_spmm_power_law_262144_full_size(device, op, dtype, index_dtype):
    ip, ix, v, B = synthetic.power_law_rows(seed=1)
    n = 262144
    data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
    Bt = torch.tensor(B, dtype=dtype, device=device)
    C = fs.flagsparse_spmm_csr(data, idx, ptr, Bt, (n, n), op=op)
    ref = oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(B, dtype), (n, n), op)
    assert_rel(C, ref, dtype, f"power-law {op}")


def test_spmm_deterministic_and_plan_reuse(device):
    ip, ix, v, _ = synthetic.power_law_rows(30000, 30000, 2.1, 24, 4096, n_dense=0, seed=5)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = fs.prepare_spmm_csr_opt(data, idx, ptr, (30000, 30000))
    B = torch.randn(30000, 64, device=device)
    C1 = fs.flagsparse_spmm_csr_opt(B=B, prepared=prepared)
    C2 = fs.flagsparse_spmm_csr_opt(B=B, prepared=prepared)
    C3 = fs.flagsparse_spmm_csr(data, idx, ptr, B, (30000, 30000))
    assert torch.equal(C1, C2) and torch.equal(C1, C3)


def test_spmm_empty_and_zero_width(device):
    ptr = torch.zeros(4, dtype=torch.int32, device=device)
    idx = torch.empty(0, dtype=torch.int32, device=device)
    data = torch.empty(0, device=device)
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.randn(5, 4, device=device), (3, 5))
    assert torch.equal(C, torch.zeros(3, 4, device=device))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.randn(3, 4, device=device), (3, 5), op="trans")
    assert torch.equal(C, torch.zeros(5, 4, device=device))
    ip, ix, v = _case(6, 5, seed=1)
    d2, i2, p2 = device_csr(ip, ix, v, device, torch.float32)
    C = fs.flagsparse_spmm_csr(d2, i2, p2, torch.randn(5, 0, device=device), (6, 5))
    assert C.shape == (6, 0)


def test_spmm_return_meta_times_transpose_path(device):
    ip, ix, v = _case(8, 12, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32, torch.int64)
    B = torch.randn(8, 4, device=device)
    out, elapsed_ms, meta = fs.flagsparse_spmm_csr(data, idx, ptr, B, (8, 12), op="trans", return_time=True,
                                                   return_meta=True)
    assert out.shape == (12, 4)
    assert meta["op"] == "trans"
    assert meta["symbolic_ms"] >= 0.0 and meta["compute_ms"] >= 0.0
    assert elapsed_ms == pytest.approx(meta["op_total_ms"])
    assert meta["op_total_ms"] == pytest.approx(meta["symbolic_ms"] + meta["compute_ms"])


def test_spmm_non_meta_has_zero_symbolic_time(device):
    ip, ix, v = _case(8, 12, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    out, meta = fs.flagsparse_spmm_csr(data, idx, ptr, torch.randn(12, 4, device=device), (8, 12), op="non",
                                       return_meta=True)
    assert out.shape == (8, 4) and meta["op"] == "non" and meta["symbolic_ms"] == 0.0


def test_spmm_op_validation_errors(device):
    ip, ix, v = _case(8, 12, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    B_non = torch.randn(12, 4, device=device)
    B_trans = torch.randn(8, 4, device=device)
    with pytest.raises(ValueError, match="op must be one of"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_non, (8, 12), op="bad")
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_non, (8, 12), op="non", transpose=True)
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_trans, (8, 12), op="trans", transpose=False)
    with pytest.raises(ValueError, match="B.shape\\[0\\] must be n_cols=8"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_non, (8, 12), op="trans")
    with pytest.raises(ValueError, match="out shape/dtype must match result"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_trans, (8, 12), op="trans",
                               out=torch.empty((8, 4), dtype=torch.float32, device=device))
    with pytest.raises(TypeError, match="B dtype must match data dtype"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_non.double(), (8, 12))
    with pytest.raises(ValueError, match="indptr must be nondecreasing"):
        bad = ptr.clone()
        bad[3] = ptr[4] + 1
        fs.flagsparse_spmm_csr(data, idx, bad, B_non, (8, 12))
    with pytest.raises(ValueError, match="max_segments override"):
        fs.flagsparse_spmm_csr(data, idx, ptr, B_non, (8, 12), block_nnz=1, max_segments=1)


def test_spmm_out_argument_written_in_place(device):
    ip, ix, v = _case(30, 20, seed=4)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    B = torch.randn(20, 8, device=device)
    out = torch.full((30, 8), float("nan"), device=device)
    C = fs.flagsparse_spmm_csr(data, idx, ptr, B, (30, 20), out=out)
    assert C is out and not torch.isnan(out).any()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_spmm_opt_matches_oracle(device, dtype, index_dtype):
    for M, N, K in MNK:
        ip, ix, v = _case(M, K, seed=M + K)
        data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
        B = np.random.default_rng(3).standard_normal((K, N)) * 0.125
        prepared = fs.prepare_spmm_csr_opt(data, idx, ptr, (M, K))
        C, meta = fs.flagsparse_spmm_csr_opt(B=torch.tensor(B, dtype=dtype, device=device), prepared=prepared,
                                             return_meta=True)
        assert_rel(C, oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(B, dtype), (M, K)), dtype)
        assert {"long_part_count", "long_row_count", "bucket_count"} <= set(meta)
    with pytest.raises(TypeError, match="only supports float32 and float64"):
        p16 = fs.prepare_spmm_csr_opt(data.half(), idx, ptr, (M, K))
        fs.flagsparse_spmm_csr_opt(B=torch.randn(K, N, device=device).half(), prepared=p16)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_spmm_opt_alg2_matches_oracle(device, dtype, index_dtype, monkeypatch):
    # opt-alg2 = the autotuned launch; every candidate must give the oracle's result
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", "")
    cases = [(M, N, K) + _case(M, K, seed=M + K) for M, N, K in MNK]
    ip, ix, v, Bp = synthetic.power_law_rows(3000, 3000, 2.1, 24, 1024, n_dense=32, seed=5)
    cases.append((3000, 32, 3000, ip, ix, v))
    for M, N, K, ip, ix, v in cases:
        data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
        B = np.random.default_rng(4).standard_normal((K, N)) * 0.125
        Bt = torch.tensor(B, dtype=dtype, device=device)
        prepared = fs.prepare_spmm_csr_opt_alg2(data, idx, ptr, (M, K))
        C, ms, meta = fs.flagsparse_spmm_csr_opt_alg2(B=Bt, prepared=prepared, return_time=True, return_meta=True)
        ref = oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(B, dtype), (M, K))
        assert_rel(C, ref, dtype, f"alg2 {M}x{K} n{N}")
        assert {"device_name", "sm_count", "warp_size", "bucket_hits", "launch_configs", "max_row_nnz",
                "n_dense_cols", "symbolic_ms", "compute_ms", "op_total_ms"} <= set(meta)
        assert meta["n_dense_cols"] == N and meta["op_total_ms"] == ms and ms >= 0
        out = torch.full((M, N), 7.0, dtype=dtype, device=device)
        got = fs.flagsparse_spmm_csr_opt_alg2_preprocess(data, idx, ptr, Bt, (M, K), out=out)
        assert got is out and torch.equal(out, C)
    with pytest.raises(TypeError, match="only supports float32"):
        p16 = fs.prepare_spmm_csr_opt_alg2(data.half(), idx, ptr, (M, K))
        fs.flagsparse_spmm_csr_opt_alg2(B=Bt.half(), prepared=p16)
    with pytest.raises(TypeError, match="PreparedCsrSpmmOptAlg2"):
        fs.flagsparse_spmm_csr_opt_alg2(B=Bt, prepared=fs.prepare_spmm_csr_opt(data, idx, ptr, (M, K)))
    with pytest.raises(ValueError, match="does not match prepared.shape"):
        fs.flagsparse_spmm_csr_opt_alg2(B=Bt, prepared=prepared, shape=(M + 1, K))
    with pytest.raises(ValueError, match="B.shape\\[0\\] must be n_cols"):
        fs.flagsparse_spmm_csr_opt_alg2(B=Bt[1:], prepared=prepared)
    with pytest.raises(ValueError, match="out shape/dtype"):
        fs.flagsparse_spmm_csr_opt_alg2(B=Bt, prepared=prepared, out=torch.empty(M, N + 1, dtype=dtype,
                                                                                  device=device))


@pytest.mark.parametrize("alg", sorted(fs.SPMM_CSR_ALGORITHMS))
@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_spmm_route_api(device, alg, op, layout):
    ip, ix, v = _case(64, 96, seed=11)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float64)
    prepared = fs.prepare_spmm_csr_route(data, idx, ptr, (64, 96), op=op, alg=alg)
    eff = 64 if op == "trans" else 96
    Bn = np.random.default_rng(2).standard_normal((eff, 16))
    B = torch.tensor(Bn, device=device)
    if layout == "col":
        B = B.t().contiguous().t()
    C, meta = fs.flagsparse_spmm_csr_run(prepared, B, dense_layout=layout, return_meta=True, timing=True,
                                         diagnostics=True)
    assert meta["alg"] == alg and meta["output_layout"] == layout
    assert "diagnostics" in meta and meta["compute_ms"] >= 0
    assert_rel(C, oracle.spmm(ip, ix, v, Bn, (64, 96), op), torch.float64)


def test_spmm_bitwise_equal_across_op_entry_points(device):
    # trans through the CSC permutation equals non on the materialised transpose
    ip, ix, v = _case(200, 150, seed=8, density=0.05)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    from flagsparse_amd._csc import transpose_csr

    dt, it, pt, shape_t = transpose_csr(data, idx, ptr, (200, 150))
    B = torch.randn(200, 32, device=device)
    C_trans = fs.flagsparse_spmm_csr(data, idx, ptr, B, (200, 150), op="trans")
    C_non = fs.flagsparse_spmm_csr(dt, it, pt, B, shape_t)
    assert torch.equal(C_trans, C_non)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_spmm_trans_gathered_values_equal_values_via_perm(device, dtype, index_dtype):
    # A^T B with values gathered into CSC order first (the default) against the
    # kernel that loads data[perm[q]] itself; the long rows of A^T (columns 5,
    # 900, 2047 of A, 3000 nonzeros each) take the segmented path
    ip_t, ix_t, v_t = long_row_matrix(2048, 3000, [5, 900, 2047], 3000, 4, seed=12)
    At = sp.csr_matrix((v_t, ix_t, ip_t), shape=(2048, 3000))
    A = At.T.tocsr()
    A.sort_indices()
    data, idx, ptr = device_csr(A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, device, dtype,
                                index_dtype)
    h = CsrHandle(idx, ptr, A.shape)
    B = torch.randn(3000, 96, device=device).to(dtype)
    C_gather = h.spmm(data, B, transpose=True)
    C_perm = h.spmm(data, B, transpose=True, values_via_perm=True)
    assert torch.equal(C_gather, C_perm)
    ref = oracle.spmm(A.indptr.astype(np.int64), A.indices.astype(np.int64), np_dtype_round(A.data, dtype),
                      np_dtype_round(B.double().cpu().numpy(), dtype), A.shape, "trans")
    assert_rel(C_gather, ref, dtype, "trans gathered")


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16], ids=["bf16", "f16"])
@pytest.mark.parametrize("op", ["non", "trans"])
def test_spmm_half_precision_storage(device, dtype, op):
    ip, ix, v, B = synthetic.power_law_rows(5000, 5000, 2.1, 24, 4096, n_dense=128, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    eff_B = B if op == "non" else np.random.default_rng(1).standard_normal((5000, 128))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(eff_B, dtype=dtype, device=device), (5000, 5000), op=op)
    assert C.dtype == dtype
    # against the FP64 oracle on the unrounded inputs: storage + output rounding
    ref = oracle.spmm(ip, ix, v, eff_B, (5000, 5000), op)
    assert_rel(C, ref, dtype, f"{dtype} vs fp64")
