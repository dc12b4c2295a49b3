"""The reference-harness benchmark cases on the GPU: result layout, the HIP
kernel's agreement with the float64 golden (north-star metric), and the
rocSPARSE leg (torch.sparse.mm) agreeing too, on small unseeded-style inputs
(seeded here through torch.manual_seed)."""

import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import benchmarks

from ._helpers import REL_TOL

pytestmark = pytest.mark.gpu

SECTIONS = {"parameters", "performance", "verification", "backend_status", "samples"}


@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_benchmark_spmv_case(device, op, dtype):
    torch.manual_seed(0)
    r = fs.benchmark_spmv_case(n_rows=700, n_cols=500, nnz=6000, value_dtype=dtype, warmup=1, iters=3, op=op)
    assert SECTIONS <= set(r)
    assert r["parameters"]["op"] == op and r["performance"]["triton_ms"] > 0
    assert r["verification"]["triton_max_relative_error"] <= REL_TOL[dtype]
    assert r["verification"]["triton_match_reference"]
    y, ref = r["samples"]["triton"], r["samples"]["reference"]
    assert y.shape == ref.shape == ((500,) if op == "trans" else (700,))
    if r["samples"]["pytorch"] is not None:
        assert r["verification"]["pytorch_match_reference"]


@pytest.mark.parametrize("op", ["non", "trans"])
def test_benchmark_spmm_case(device, op):
    torch.manual_seed(1)
    r = fs.benchmark_spmm_case(n_rows=600, n_cols=900, nnz=9000, n_dense_cols=24, warmup=1, iters=3, op=op)
    assert SECTIONS <= set(r)
    v = r["verification"]
    assert v["triton_max_relative_error"] <= REL_TOL[torch.float32]
    assert v["triton_strict_allclose_match"]
    assert r["samples"]["triton"].shape == ((900, 24) if op == "trans" else (600, 24))
    assert r["backend_status"]["cusparse_unavailable_reason"]
    if r["samples"]["pytorch"] is not None:
        assert v["pytorch_max_relative_error"] <= 1e-5


def test_benchmark_sddmm_case(device):
    torch.manual_seed(2)
    r = fs.benchmark_sddmm_case(n_rows=300, n_cols=200, nnz=4000, k_dim=40, warmup=1, iters=3, alpha=0.5, beta=2.0)
    assert SECTIONS <= set(r)
    assert r["verification"]["triton_max_relative_error"] <= REL_TOL[torch.float32]
    assert r["verification"]["pytorch_max_relative_error"] <= REL_TOL[torch.float32]
    assert r["samples"]["triton"].shape == (4000,)
    with pytest.raises(TypeError, match="float32 or torch.float64"):
        fs.benchmark_sddmm_case(value_dtype=torch.bfloat16)


def test_benchmark_empty_and_graph_timer(device):
    d, i, p = benchmarks._build_random_csr(5, 4, 0, torch.float32, torch.int32, device)
    assert d.numel() == 0 and p.shape == (6,) and p.dtype == torch.int64
    x = torch.randn(1 << 12, device=device)
    y = torch.empty_like(x)
    ms = benchmarks._benchmark_cuda_graph_op(lambda: torch.mul(x, 2.0, out=y), graph_batch=4, warmup=1, repeats=3)
    assert ms > 0 and torch.equal(y, x * 2.0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_benchmark_spmm_opt_cases(device, dtype):
    torch.manual_seed(3)
    r = fs.benchmark_spmm_opt_case(n_rows=500, n_cols=400, nnz=6000, n_dense_cols=24, value_dtype=dtype, warmup=1,
                                   iters=2)
    assert SECTIONS <= set(r)
    v, p = r["verification"], r["performance"]
    assert v["base_ok"] and v["opt_ok"] and v["status"] == "PASS"
    assert p["base_ms"] > 0 and p["opt_ms"] > 0 and p["cu_ms"] is None
    assert r["samples"]["opt"].shape == (500, 24) and r["samples"]["reference"].dtype == dtype
    assert torch.equal(r["samples"]["base"], r["samples"]["opt"])  # the same kernel and plan
    torch.manual_seed(4)
    r2 = fs.benchmark_spmm_opt_alg2_case(n_rows=300, n_cols=700, nnz=5000, n_dense_cols=40, value_dtype=dtype,
                                         warmup=1, iters=2)
    v2 = r2["verification"]
    for key in ("base_vs_torch_err", "opt_vs_torch_err", "opt_alg2_vs_torch_err"):
        assert v2[key] <= 1.0, key
    assert v2["opt_alg2_vs_cusparse_err"] is None and r2["backend_status"]["cusparse_unavailable_reason"]
    assert r2["samples"]["opt_alg2_triton"].shape == (300, 40) and r2["performance"]["opt_alg2_ms"] > 0
    with pytest.raises(TypeError, match="only supports float32 and float64"):
        fs.benchmark_spmm_opt_case(value_dtype=torch.bfloat16)
    with pytest.raises(TypeError, match="only supports float32 and float64"):
        fs.benchmark_spmm_opt_alg2_case(value_dtype=torch.float16)


def test_comprehensive_spmm_test_is_the_spmm_case(device):
    torch.manual_seed(5)
    r = fs.comprehensive_spmm_test(n_rows=200, n_cols=150, nnz=1500, n_dense_cols=16, dtype=torch.float64, warmup=1,
                                   iters=2, op="trans")
    assert SECTIONS <= set(r) and r["parameters"]["op"] == "trans"
    assert r["parameters"]["value_dtype"] == str(torch.float64)
    assert r["verification"]["triton_max_relative_error"] <= REL_TOL[torch.float64]
