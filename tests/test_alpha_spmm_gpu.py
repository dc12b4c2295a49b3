"""The AlphaSparse-ALG1 SpMM routes on the GPU vs the C oracle.

Mirrors tests/pytest/test_alpha_spmm_alg1_accuracy.py of the reference: every
route against the FP64 product (float32 / float64, int32 / int64 indices,
(M, K, N) = (72, 64, 33) and dense widths 1 and 65 as in its heuristics
grid), prepared + meta equal to the raw call, the empty-rows / tail-columns
matrix, plus the argument contract (raw or prepared, dtype and shape errors,
``out``).  The reference's warp/factor launch-heuristic test has no
counterpart: the meta here describes the HIP launch instead.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import autotune, synthetic
from oracle import oracle

from ._helpers import assert_rel, device_csr, np_dtype_round

pytestmark = pytest.mark.gpu

ROUTES = {
    "alpha_spmm_alg1": (fs.flagsparse_alpha_spmm_alg1, fs.prepare_alpha_spmm_alg1),
    "alpha_spmm_alg1_tle": (fs.flagsparse_alpha_spmm_alg1_tle, fs.prepare_alpha_spmm_alg1_tle),
    "alpha_spmm_alg1_tle_opt": (fs.flagsparse_alpha_spmm_alg1_tle_opt, fs.prepare_alpha_spmm_alg1_tle_opt),
    "alpha_spmm_alg1_tle_opt2": (fs.flagsparse_alpha_spmm_alg1_tle_opt2, fs.prepare_alpha_spmm_alg1_tle_opt2),
}
DTYPES = [torch.float32, torch.float64]


@pytest.fixture(autouse=True)
def _no_tune_file(monkeypatch):
    # the tuned route may tune: keep its cache in memory only
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", "")
    autotune.clear_memory()


def _case(M, K, N, dtype, index_dtype, seed):
    ip, ix, v = synthetic.random_csr(M, K, 0.15, seed=seed)
    data, idx, ptr = device_csr(ip, ix, v, "cuda", dtype, index_dtype, index_dtype)
    b = np.random.default_rng(seed + 1).standard_normal((K, N))
    B = torch.tensor(b, dtype=dtype, device="cuda")
    ref = oracle.spmm(ip, ix, np_dtype_round(v, dtype), np_dtype_round(b, dtype), (M, K))
    return data, idx, ptr, B, ref


@pytest.mark.parametrize("route", list(ROUTES))
@pytest.mark.parametrize("dtype", DTYPES, ids=["float32", "float64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("M,K,N", [(72, 64, 33), (40, 48, 1), (40, 48, 65), (300, 200, 128)])
def test_routes_match_oracle(device, route, dtype, index_dtype, M, K, N):
    fn, _ = ROUTES[route]
    data, idx, ptr, B, ref = _case(M, K, N, dtype, index_dtype, seed=M + K + N)
    C, meta = fn(data, idx, ptr, B, (M, K), return_meta=True)
    assert C.shape == (M, N) and C.dtype == dtype
    assert_rel(C, ref, dtype, f"{route} {M}x{K} n{N}")
    assert meta["route"] == route and meta["n_dense_cols"] == N and meta["backend"] == "hip"
    assert meta["lanes_n"] in (1, 2, 4, 8, 16, 32, 64)
    tuned = route == "alpha_spmm_alg1_tle_opt2"
    assert meta["launch_version"] == ("hip_row_bins_tuned" if tuned else "hip_row_bins")


@pytest.mark.parametrize("route", ["alpha_spmm_alg1_tle_opt", "alpha_spmm_alg1_tle_opt2"])
@pytest.mark.parametrize("dtype", DTYPES, ids=["float32", "float64"])
def test_prepare_and_meta_match_raw(device, route, dtype):
    fn, prepare = ROUTES[route]
    build = fs.build_alpha_spmm_alg1_tle_opt_meta if route.endswith("opt") else fs.build_alpha_spmm_alg1_tle_opt2_meta
    data, idx, ptr, B, ref = _case(72, 64, 33, dtype, torch.int32, seed=5)
    prepared = prepare(data, idx, ptr, (72, 64))
    meta = build(prepared, B)
    assert meta["route"] == route
    raw = fn(data, idx, ptr, B, (72, 64))
    via_prepared = fn(B=B, prepared=prepared, meta=meta)
    assert torch.equal(raw, via_prepared)  # the same launch
    assert_rel(via_prepared, ref, dtype)
    # a meta asking for other lanes along N runs that launch, same result within tolerance
    other = dict(meta, lanes_n=1 if meta["lanes_n"] > 1 else 2)
    assert_rel(fn(B=B, prepared=prepared, meta=other), ref, dtype, "meta lanes")


@pytest.mark.parametrize("route", list(ROUTES))
@pytest.mark.parametrize("dtype", DTYPES, ids=["float32", "float64"])
def test_empty_rows_and_tail_columns(device, route, dtype):
    """The reference's 5 x 9 case: rows 1 and 3 empty, a nonzero in the last
    column, 17 dense columns (a tail past every vector width)."""
    fn, _ = ROUTES[route]
    ip = np.array([0, 2, 2, 3, 3, 4], dtype=np.int64)
    ix = np.array([0, 4, 2, 8], dtype=np.int64)
    v = np.array([1.5, -2.0, 3.0, 4.0])
    data = torch.tensor(v, dtype=dtype, device=device)
    idx = torch.tensor(ix, dtype=torch.int32, device=device)
    ptr = torch.tensor(ip, dtype=torch.int64, device=device)
    b = np.random.default_rng(9).standard_normal((9, 17))
    C = fn(data, idx, ptr, torch.tensor(b, dtype=dtype, device=device), (5, 9))
    assert_rel(C, oracle.spmm(ip, ix, v, np_dtype_round(b, dtype), (5, 9)), dtype)
    assert torch.all(C[1] == 0) and torch.all(C[3] == 0)


def test_argument_contract(device):
    data, idx, ptr, B, _ = _case(30, 20, 8, torch.float32, torch.int32, seed=3)
    prepared = fs.prepare_alpha_spmm_alg1(data, idx, ptr, (30, 20))
    for fn, _ in ROUTES.values():
        with pytest.raises(ValueError, match="Pass either raw CSR inputs or prepared, not both"):
            fn(data, idx, ptr, B, (30, 20), prepared=prepared)
        with pytest.raises(ValueError, match="required when prepared is not provided"):
            fn(data, idx, None, B, (30, 20))
        with pytest.raises(TypeError, match="PreparedAlphaSpmmAlg1"):
            fn(B=B, prepared=object())
        with pytest.raises(ValueError, match="B is required"):
            fn(prepared=prepared)
        with pytest.raises(TypeError, match="B dtype must match"):
            fn(B=B.double(), prepared=prepared)
        with pytest.raises(ValueError, match=r"B.shape\[0\] must be n_cols=20"):
            fn(B=B[1:], prepared=prepared)
        with pytest.raises(ValueError, match="out shape/dtype must match result"):
            fn(B=B, prepared=prepared, out=torch.empty(30, 9, device=device))
        with pytest.raises(ValueError, match="B must be a CUDA tensor"):
            fn(B=B.cpu(), prepared=prepared)
    with pytest.raises(TypeError, match="only support float32, float64, complex64, and complex128"):
        fs.prepare_alpha_spmm_alg1(data.half(), idx, ptr, (30, 20))
    with pytest.raises(ValueError, match="indptr"):
        fs.prepare_alpha_spmm_alg1(data, idx, ptr[:-1], (30, 20))
    cplx = fs.prepare_alpha_spmm_alg1(data.to(torch.complex64), idx, ptr, (30, 20))
    with pytest.raises(TypeError, match="has no HIP kernel"):
        fs.flagsparse_alpha_spmm_alg1(B=B.to(torch.complex64), prepared=cplx)
    with pytest.raises(ValueError, match="lanes_n must be a power of two"):
        fs.flagsparse_alpha_spmm_alg1_tle_opt(B=B, prepared=prepared, meta={"lanes_n": 3})
    with pytest.raises(TypeError, match="meta must be a dict"):
        fs.flagsparse_alpha_spmm_alg1_tle_opt(B=B, prepared=prepared, meta=[1])
    assert fs.is_alpha_spmm_alg1_tle_available() and fs.is_alpha_spmm_alg1_tle_opt_available()
    assert fs.is_alpha_spmm_alg1_tle_opt2_available() and fs.alpha_spmm_alg1_tle_unavailable_reason() == ""


def test_out_in_place_and_strided(device):
    data, idx, ptr, B, ref = _case(64, 50, 16, torch.float32, torch.int32, seed=4)
    prepared = fs.prepare_alpha_spmm_alg1_tle(data, idx, ptr, (64, 50))
    out = torch.empty(64, 16, device=device)
    got = fs.flagsparse_alpha_spmm_alg1_tle(B=B, prepared=prepared, out=out)
    assert got.data_ptr() == out.data_ptr()
    assert_rel(out, ref, torch.float32)
    wide = torch.full((64, 32), 5.0, device=device)
    got = fs.flagsparse_alpha_spmm_alg1_tle(B=B, prepared=prepared, out=wide[:, ::2])
    assert torch.equal(wide[:, ::2], out) and torch.all(wide[:, 1::2] == 5.0)
    # a column-major B is made row-major first
    Bt = B.t().contiguous().t()
    assert torch.equal(fs.flagsparse_alpha_spmm_alg1_tle(B=Bt, prepared=prepared), out)
