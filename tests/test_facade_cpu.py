"""The reference's import surface resolves on this build (CPU only: nothing
is launched).  The expected names are those the reference's own public-API
test checks (tests/ci/test_public_api.py:19-48) and the module paths its
tests import (tests/test_spmv.py:28, tests/test_spmm.py:42,
tests/test_sddmm.py:47, tests/test_spmm_opt_alg2.py:41-42,
tests/ci/test_sparse_common_cpu.py:28)."""

import importlib

import pytest
import torch

import flagsparse
import flagsparse_amd

EXPECTED_TOP_LEVEL = {
    "flagsparse_gather", "flagsparse_scatter", "flagsparse_spmv_csr", "flagsparse_spmv_coo",
    "flagsparse_spmv_csc", "flagsparse_spmv_bsr", "prepare_spmv_csc", "prepare_spmv_bsr", "flagsparse_spmm_csr",
    "flagsparse_spmm_coo", "flagsparse_spgemm_csr", "flagsparse_sddmm_csr", "flagsparse_spsv_csr",
    "flagsparse_spsv_coo", "flagsparse_spsv_sell", "flagsparse_spsv_analysis_sell", "flagsparse_spsv_solve_sell",
    "flagsparse_spsm_csr", "flagsparse_spsm_coo", "create_csr_matrix", "create_coo_matrix", "read_mtx_file",
}

HOT_PATH = {
    "flagsparse_spmv_csr", "prepare_spmv_csr", "PreparedCsrSpmv", "flagsparse_spmm_csr", "flagsparse_spmm_csr_opt",
    "prepare_spmm_csr_opt", "PreparedCsrSpmmOpt", "flagsparse_spmm_csr_run", "prepare_spmm_csr_route",
    "PreparedCsrSpmmRoute", "SPMM_CSR_ALGORITHMS", "list_spmm_csr_algorithms", "resolve_spmm_csr_algorithm",
    "flagsparse_spmm_csr_opt_alg2", "prepare_spmm_csr_opt_alg2", "PreparedCsrSpmmOptAlg2", "flagsparse_sddmm_csr",
    "prepare_sddmm_csr", "SDDMMPrepared", "benchmark_spmv_case", "benchmark_spmm_case", "benchmark_sddmm_case",
}


def test_public_surface_exposes_core_entry_points():
    exported = set(dir(flagsparse))
    assert EXPECTED_TOP_LEVEL <= exported
    assert EXPECTED_TOP_LEVEL <= set(flagsparse.__all__)


def test_hot_path_names_are_this_builds_operators():
    for name in HOT_PATH:
        assert getattr(flagsparse, name) is getattr(flagsparse_amd, name), name


def test_star_import_resolves_every_name():
    ns = {}
    exec("from flagsparse import *", ns)  # noqa: S102
    assert set(flagsparse.__all__) <= set(ns)


@pytest.mark.parametrize("module,names", [
    ("spmv_csr", ["flagsparse_spmv_csr", "prepare_spmv_csr", "PreparedCsrSpmv", "flagsparse_spmv_coo_tocsr"]),
    ("spmm_csr", ["flagsparse_spmm_csr", "flagsparse_spmm_csr_opt", "flagsparse_spmm_csr_run", "SPMM_CSR_ALGORITHMS",
                  "benchmark_spmm_case", "comprehensive_spmm_test", "prepare_spmm_csr_opt_alg1"]),
    ("spmm_csr_opt_alg2", ["flagsparse_spmm_csr_opt_alg2", "prepare_spmm_csr_opt_alg2", "PreparedCsrSpmmOptAlg2",
                           "benchmark_spmm_opt_alg2_case"]),
    ("sddmm_csr", ["flagsparse_sddmm_csr", "prepare_sddmm_csr", "SDDMMPrepared", "benchmark_sddmm_case"]),
    ("_common", ["SUPPORTED_INDEX_DTYPES", "SUPPORTED_VALUE_DTYPES"]),
])
def test_reference_module_paths(module, names):
    mod = importlib.import_module(f"flagsparse.sparse_operations.{module}")
    for name in names:
        assert getattr(mod, name) is getattr(flagsparse_amd, name, getattr(mod, name)), name
    assert getattr(flagsparse.sparse_operations, module) is mod


def test_common_dtypes_match_the_reference():
    from flagsparse.sparse_operations import _common

    assert _common.SUPPORTED_INDEX_DTYPES == (torch.int32, torch.int64)


def test_names_outside_the_path_raise_not_implemented():
    for name in ("flagsparse_spgemm_csr", "flagsparse_spsv_csr", "flagsparse_gather", "read_mtx_file"):
        fn = getattr(flagsparse, name)
        with pytest.raises(NotImplementedError, match="outside the CSR SpMV / SpMM / SDDMM path"):
            fn()
    with pytest.raises(AttributeError):
        getattr(flagsparse, "no_such_operator")
