"""Host-side logic that runs without a GPU: op codes, validation order and
messages (the reference's validators reject CPU tensors first in the same
order), plan configuration, metrics and the synthetic generators."""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import _common, _plan, spmm_csr, spmv_csr, synthetic


@pytest.mark.parametrize(
    "op,code", [(None, 0), ("non", 0), (" TRANS ", 1), ("conj", 2), (0, 0), (1, 1), (2, 2), (2.0, 2)]
)
def test_normalize_op(op, code):
    assert _common.normalize_op(op) == code


@pytest.mark.parametrize("op", ["bad", "2", 3, -1, object()])
def test_normalize_op_rejects(op):
    with pytest.raises(ValueError, match="op must be one of: 0=non, 1=trans, 2=conj"):
        _common.normalize_op(op)


def test_transpose_flag_maps_to_trans():
    assert _common.normalize_op(None, transpose=True) == 1


def _cpu_csr():
    data = torch.ones(3)
    indices = torch.tensor([0, 1, 2], dtype=torch.int32)
    indptr = torch.tensor([0, 1, 3], dtype=torch.int32)
    return data, indices, indptr


def test_spmv_validation_order_on_cpu_tensors():
    data, indices, indptr = _cpu_csr()
    with pytest.raises(TypeError, match="must all be torch.Tensor"):
        spmv_csr._prepare_spmv_csr_matrix([1.0], indices, indptr, (2, 3))
    with pytest.raises(ValueError, match="1D tensors"):
        spmv_csr._prepare_spmv_csr_matrix(data.view(3, 1), indices, indptr, (2, 3))
    with pytest.raises(ValueError, match="indptr length must be n_rows\\+1=4, got 3"):
        spmv_csr._prepare_spmv_csr_matrix(data, indices, indptr, (3, 3))
    with pytest.raises(ValueError, match="same length"):
        spmv_csr._prepare_spmv_csr_matrix(data[:2], indices, indptr, (2, 3))
    with pytest.raises(ValueError, match="must be CUDA tensors"):
        spmv_csr._prepare_spmv_csr_matrix(data, indices, indptr, (2, 3))
    with pytest.raises(ValueError, match="index_fallback_policy"):
        spmv_csr._prepare_spmv_csr_matrix(data, indices, indptr, (2, 3), index_fallback_policy="x")


def test_spmv_public_argument_errors():
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmv_csr(op="non", transpose=True)
    with pytest.raises(ValueError, match="required when prepared is not provided"):
        fs.flagsparse_spmv_csr(x=torch.ones(3))


def test_spmm_public_argument_errors():
    data, indices, indptr = _cpu_csr()
    B = torch.ones(3, 2)
    with pytest.raises(ValueError, match="op must be one of"):
        fs.flagsparse_spmm_csr(data, indices, indptr, B, (2, 3), op="bad")
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmm_csr(data, indices, indptr, B, (2, 3), op="non", transpose=True)
    with pytest.raises(ValueError, match="block_n must be positive"):
        fs.flagsparse_spmm_csr(data, indices, indptr, B, (2, 3), block_n=0)
    with pytest.raises(ValueError, match="shape must be a 2-tuple"):
        spmm_csr._prepare_spmm_csr_matrix(data, indices, indptr, (2, 3, 4))
    with pytest.raises(ValueError, match="must be CUDA tensors"):
        fs.flagsparse_spmm_csr(data, indices, indptr, B, (2, 3))


def test_sddmm_pattern_errors_on_cpu():
    _, indices, indptr = _cpu_csr()
    with pytest.raises(ValueError, match="must be CUDA tensors"):
        fs.prepare_sddmm_csr(indices, indptr, (2, 3))
    with pytest.raises(ValueError, match="required when prepared is not provided"):
        fs.flagsparse_sddmm_csr(x=torch.ones(2, 2), y=torch.ones(3, 2))


def test_route_registry_names_and_resolution():
    names = set(fs.SPMM_CSR_ALGORITHMS)
    assert {"csr_base", "csr_base_accuracy", "alpha_alg1_tle_opt", "alpha_alg1_tle_opt2", "spmm_csr_alg1",
            "spmm_csr_alg2", "spmm_csr_alg2_accuracy", "spmm_csr_alg2_accuracy_hp"} == names
    assert fs.resolve_spmm_csr_algorithm("auto", "non", torch.float32).name == "csr_base"
    assert fs.resolve_spmm_csr_algorithm("alg2", "trans", torch.float64).name == "spmm_csr_alg2"
    with pytest.raises(ValueError, match="unsupported CSR SpMM algorithm"):
        fs.resolve_spmm_csr_algorithm("nope", "non", torch.float32)
    with pytest.raises(TypeError, match="does not support dtype"):
        fs.resolve_spmm_csr_algorithm("spmm_csr_alg1", "non", torch.float16)
    assert "csr_base" in fs.list_spmm_csr_algorithms(op="conj", dtype=torch.bfloat16)
    with pytest.raises(ValueError, match="dense_layout must be one of"):
        spmm_csr._normalize_dense_layout("diagonal")


def test_plan_configs():
    cfg = _plan.spmv_config()
    assert cfg.lanes == (4, 8, 16, 32, 64)
    assert cfg.uppers == (16, 32, 64, 128, _plan.SPMV_LONG_THRESHOLD)
    cfg = _plan.spmm_config(16, slot_scale=1)
    assert cfg.lanes == (1, 2, 4) and cfg.uppers == (4, 8, _plan.SPMM_LONG_THRESHOLD)
    cfg = _plan.spmm_config(16, slot_scale=4)
    assert cfg.lanes == (1, 2, 4) and cfg.uppers == (16, 32, _plan.SPMM_LONG_THRESHOLD)
    cfg = _plan.spmm_config(64)
    assert cfg.lanes == (1,) and cfg.uppers == (_plan.SPMM_LONG_THRESHOLD,)
    cfg = _plan.spmm_config(1, long_threshold=16)
    assert list(cfg.uppers) == sorted(set(cfg.uppers)) and cfg.uppers[-1] == 16
    assert len(_plan.spmm_config(1).lanes) + 1 <= 8  # regular bins + long bin fit the C table


@pytest.mark.parametrize("n,elem,g", [(64, 4, 16), (128, 2, 16), (128, 4, 32), (1, 4, 1), (300, 4, 64), (32, 8, 16)])
def test_lanes_for_dense(n, elem, g):
    assert _plan.lanes_for_dense(n, elem) == g


def test_error_metrics_definition():
    ref = torch.tensor([1.0, -4.0, 2.0], dtype=torch.float64)
    cand = torch.tensor([1.0, -4.5, 2.0], dtype=torch.float64)
    m = _common.error_metrics(cand, ref)
    assert m["max_abs_error"] == pytest.approx(0.5)
    assert m["max_relative_error"] == pytest.approx(0.125)
    assert m["sum_relative_error"] == pytest.approx(0.5 / 7.0)
    v = _common.validation_metrics(cand.float(), ref)
    assert v["relative_threshold"] == 1e-6 and not v["matches_threshold"]


def test_uniform_rows_generator():
    ip, ix, v, x = synthetic.uniform_rows(512, 300, 16, seed=0)
    assert np.all(np.diff(ip) == 16)
    for r in range(512):
        seg = ix[ip[r]: ip[r + 1]]
        assert np.all(np.diff(seg) > 0) and seg.min() >= 0 and seg.max() < 300
    assert np.all(np.abs(v) <= 1) and np.all(np.abs(x) <= 1)
    ip2, ix2, v2, _ = synthetic.uniform_rows(512, 300, 16, seed=0)
    assert np.array_equal(ix, ix2) and np.array_equal(v, v2)


def test_power_law_scale_hits_the_mean():
    xm = synthetic.power_law_scale(2.1, 24, 4096)
    k = np.arange(1, 4097)
    assert np.minimum(1.0, (xm / k) ** 1.1).sum() == pytest.approx(24.0, rel=1e-9)
    rng = np.random.Generator(np.random.PCG64(1))
    lens = synthetic.power_law_lengths(rng, 200000, 2.1, 24, 4096)
    assert lens.max() <= 4096 and abs(lens.mean() - 24) < 1.0


def test_power_law_graph_is_symmetric_and_normalised():
    ip, ix, v = synthetic.power_law_graph(2000, 2.3, 14, seed=3)
    import scipy.sparse as sp

    A = sp.csr_matrix((v, ix, ip), shape=(2000, 2000))
    assert abs(A - A.T).max() < 1e-15
    assert A.diagonal().max() == 0
    deg = np.diff(ip)
    r = np.repeat(np.arange(2000), deg)
    np.testing.assert_allclose(v, 1 / np.sqrt(deg[r] * deg[ix]))


# The reference's exported names on the CSR SpMV / SpMM / SDDMM path and its
# adjacent formats (src/flagsparse/__init__.py __all__), each of which this
# package provides under the same name.
REFERENCE_PATH_NAMES = (
    "flagsparse_spmv_csr", "prepare_spmv_csr", "PreparedCsrSpmv",
    "flagsparse_spmv_csc", "prepare_spmv_csc", "PreparedCscSpmv",
    "flagsparse_spmv_coo", "prepare_spmv_coo", "PreparedCoo",
    "flagsparse_spmv_coo_tocsr", "prepare_spmv_coo_tocsr", "coo_to_csr_for_spmv",
    "flagsparse_spmv_bsr", "prepare_spmv_bsr", "PreparedBsrSpmv",
    "flagsparse_spmm_csr", "flagsparse_spmm_csr_opt", "prepare_spmm_csr_opt", "PreparedCsrSpmmOpt",
    "flagsparse_spmm_csr_opt_alg1", "prepare_spmm_csr_opt_alg1", "flagsparse_spmm_csr_opt_alg1_preprocess",
    "prepare_spmm_csr_opt_alg1_preprocess", "flagsparse_spmm_csr_opt_alg2", "prepare_spmm_csr_opt_alg2",
    "flagsparse_spmm_csr_opt_alg2_preprocess", "prepare_spmm_csr_opt_alg2_preprocess", "PreparedCsrSpmmOptAlg2",
    "SPMM_CSR_ALGORITHMS", "SpmmCsrAlgorithm", "SpmmCsrAlgorithmUnavailable", "PreparedCsrSpmmRoute",
    "prepare_spmm_csr_route", "flagsparse_spmm_csr_run", "list_spmm_csr_algorithms", "resolve_spmm_csr_algorithm",
    "flagsparse_spmm_coo", "flagsparse_spmm_coo_run", "prepare_spmm_coo_route", "PreparedCooSpmmRoute",
    "SPMM_COO_ALGORITHMS", "SpmmCooAlgorithm", "SpmmCooAlgorithmUnavailable", "list_spmm_coo_algorithms",
    "resolve_spmm_coo_algorithm",
    "flagsparse_alpha_spmm_alg1", "flagsparse_alpha_spmm_alg1_tle", "flagsparse_alpha_spmm_alg1_tle_opt",
    "flagsparse_alpha_spmm_alg1_tle_opt2", "prepare_alpha_spmm_alg1", "prepare_alpha_spmm_alg1_tle",
    "prepare_alpha_spmm_alg1_tle_opt", "prepare_alpha_spmm_alg1_tle_opt2", "PreparedAlphaSpmmAlg1",
    "build_alpha_spmm_alg1_tle_opt_meta", "build_alpha_spmm_alg1_tle_opt2_meta",
    "is_alpha_spmm_alg1_tle_available", "is_alpha_spmm_alg1_tle_opt_available",
    "is_alpha_spmm_alg1_tle_opt2_available", "alpha_spmm_alg1_tle_unavailable_reason",
    "alpha_spmm_alg1_tle_opt_unavailable_reason", "alpha_spmm_alg1_tle_opt2_unavailable_reason",
    "flagsparse_sddmm_csr", "prepare_sddmm_csr", "SDDMMPrepared",
    "benchmark_spmv_case", "benchmark_spmm_case", "benchmark_sddmm_case", "benchmark_spmm_opt_case",
    "benchmark_spmm_opt_alg2_case", "comprehensive_spmm_test",
    "SUPPORTED_VALUE_DTYPES", "SUPPORTED_INDEX_DTYPES", "error_metrics", "validation_metrics",
    "load_mtx_to_csr_torch",
)


def test_reference_path_names_are_exported():
    missing = [n for n in REFERENCE_PATH_NAMES if n not in fs.__all__ or not hasattr(fs, n)]
    assert not missing, missing
    assert len(set(fs.__all__)) == len(fs.__all__)  # no duplicates in __all__


def test_availability_queries_and_registries_need_no_device():
    assert fs.is_alpha_spmm_alg1_tle_opt2_available() and fs.alpha_spmm_alg1_tle_opt_unavailable_reason() == ""
    assert fs.list_spmm_coo_algorithms(op="trans") == ("coo_rowrun", "coo_atomic", "spmm_coo_alg1")
    assert fs.resolve_spmm_coo_algorithm("rowrun", 0, torch.float64).display_name == "COORowRun"
    with pytest.raises(ValueError, match="op must be one of"):
        fs.list_spmm_coo_algorithms(op="sideways")


def test_row_len_if_constant():
    from flagsparse_amd._plan import _row_len_if_constant

    assert _row_len_if_constant(4, 0, 64, 16) == 16
    assert _row_len_if_constant(4, 40, 104, 16) == 16  # indptr starting at 40
    assert _row_len_if_constant(4, 0, 63, 16) == 0  # one row shorter
    assert _row_len_if_constant(0, 0, 0, 0) == 0
    assert _row_len_if_constant(5, 0, 0, 0) == 0  # empty rows only
