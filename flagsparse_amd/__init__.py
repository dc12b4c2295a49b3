"""flagsparse_amd: MI355X-native (gfx950 HIP) CSR SpMV / SpMM / SDDMM.

A drop-in for the CSR SpMV / SpMM (op non/trans/conj) / SDDMM path of
FlagSparse: the public names below are those the reference exports for this
path (src/flagsparse/__init__.py:19-154, conf/operators.yaml:46-61, 132-148,
263-278), with the same arguments and behaviour.  Use it as
``import flagsparse_amd as flagsparse``.  The compute is hand-written HIP for
gfx950 in libflagsparse_hip.so; there is no other backend.

Beyond the reference: the differentiable SpMM (``CsrMatrix``,
``spmm_csr_autograd``) and the BSR SpMM on the matrix cores
(``csr_to_bsr``, ``quantize_bsr_fp8``, ``flagsparse_spmm_bsr``).
Adjacent to the path: COO SpMV through a GPU COO -> CSR conversion
(``flagsparse_spmv_coo_tocsr``), BSR SpMV (``flagsparse_spmv_bsr``) and
Matrix Market ingestion
(``load_mtx_to_csr_torch``).
"""

from ._common import (
    SUPPORTED_INDEX_DTYPES,
    SUPPORTED_VALUE_DTYPES,
    error_metrics,
    validation_metrics,
)
from .alpha_spmm_alg1 import (
    PreparedAlphaSpmmAlg1,
    alpha_spmm_alg1_tle_opt2_unavailable_reason,
    alpha_spmm_alg1_tle_opt_unavailable_reason,
    alpha_spmm_alg1_tle_unavailable_reason,
    build_alpha_spmm_alg1_tle_opt2_meta,
    build_alpha_spmm_alg1_tle_opt_meta,
    flagsparse_alpha_spmm_alg1,
    flagsparse_alpha_spmm_alg1_tle,
    flagsparse_alpha_spmm_alg1_tle_opt,
    flagsparse_alpha_spmm_alg1_tle_opt2,
    is_alpha_spmm_alg1_tle_available,
    is_alpha_spmm_alg1_tle_opt2_available,
    is_alpha_spmm_alg1_tle_opt_available,
    prepare_alpha_spmm_alg1,
    prepare_alpha_spmm_alg1_tle,
    prepare_alpha_spmm_alg1_tle_opt,
    prepare_alpha_spmm_alg1_tle_opt2,
)
from .benchmarks import (
    benchmark_sddmm_case,
    benchmark_spmm_case,
    benchmark_spmm_opt_alg2_case,
    benchmark_spmm_opt_case,
    benchmark_spmv_case,
    comprehensive_spmm_test,
)
from .autograd import CsrMatrix, flagsparse_spmm_csr_autograd, spmm_csr_autograd
from .bsr import csr_to_bsr, flagsparse_spmm_bsr, quantize_bsr_fp8
from .sddmm_csr import SDDMMPrepared, flagsparse_sddmm_csr, prepare_sddmm_csr
from .spmm_csr import (
    SPMM_CSR_ALGORITHMS,
    PreparedCsrSpmmOpt,
    PreparedCsrSpmmOptAlg2,
    PreparedCsrSpmmRoute,
    SpmmCsrAlgorithm,
    SpmmCsrAlgorithmUnavailable,
    flagsparse_spmm_csr,
    flagsparse_spmm_csr_opt,
    flagsparse_spmm_csr_opt_alg1,
    flagsparse_spmm_csr_opt_alg1_preprocess,
    flagsparse_spmm_csr_opt_alg2,
    flagsparse_spmm_csr_opt_alg2_preprocess,
    flagsparse_spmm_csr_run,
    list_spmm_csr_algorithms,
    prepare_spmm_csr_opt,
    prepare_spmm_csr_opt_alg1,
    prepare_spmm_csr_opt_alg1_preprocess,
    prepare_spmm_csr_opt_alg2,
    prepare_spmm_csr_opt_alg2_preprocess,
    prepare_spmm_csr_route,
    resolve_spmm_csr_algorithm,
)
from .spmm_coo import (
    SPMM_COO_ALGORITHMS,
    PreparedCooSpmmRoute,
    SpmmCooAlgorithm,
    SpmmCooAlgorithmUnavailable,
    flagsparse_spmm_coo,
    flagsparse_spmm_coo_run,
    list_spmm_coo_algorithms,
    prepare_spmm_coo_route,
    resolve_spmm_coo_algorithm,
)
from .mtx import load_mtx_to_csr_torch
from .spmv_bsr import PreparedBsrSpmv, flagsparse_spmv_bsr, prepare_spmv_bsr
from .spmv_coo import (
    PreparedCoo,
    coo_to_csr_for_spmv,
    flagsparse_spmv_coo,
    flagsparse_spmv_coo_tocsr,
    prepare_spmv_coo,
    prepare_spmv_coo_tocsr,
)
from .spmv_csr import PreparedCsrSpmv, flagsparse_spmv_csr, prepare_spmv_csr
from .spmv_csc import PreparedCscSpmv, flagsparse_spmv_csc, prepare_spmv_csc

__version__ = "0.1.0"

__all__ = [
    "comprehensive_spmm_test",
    "SPMM_COO_ALGORITHMS",
    "PreparedCooSpmmRoute",
    "SpmmCooAlgorithm",
    "SpmmCooAlgorithmUnavailable",
    "flagsparse_spmm_coo",
    "flagsparse_spmm_coo_run",
    "list_spmm_coo_algorithms",
    "prepare_spmm_coo_route",
    "resolve_spmm_coo_algorithm",
    "benchmark_spmm_opt_case",
    "benchmark_spmm_opt_alg2_case",
    "PreparedCoo",
    "flagsparse_spmv_coo",
    "prepare_spmv_coo",
    "PreparedCscSpmv",
    "flagsparse_spmv_csc",
    "prepare_spmv_csc",
    "PreparedAlphaSpmmAlg1",
    "alpha_spmm_alg1_tle_opt2_unavailable_reason",
    "alpha_spmm_alg1_tle_opt_unavailable_reason",
    "alpha_spmm_alg1_tle_unavailable_reason",
    "build_alpha_spmm_alg1_tle_opt2_meta",
    "build_alpha_spmm_alg1_tle_opt_meta",
    "flagsparse_alpha_spmm_alg1",
    "flagsparse_alpha_spmm_alg1_tle",
    "flagsparse_alpha_spmm_alg1_tle_opt",
    "flagsparse_alpha_spmm_alg1_tle_opt2",
    "is_alpha_spmm_alg1_tle_available",
    "is_alpha_spmm_alg1_tle_opt2_available",
    "is_alpha_spmm_alg1_tle_opt_available",
    "prepare_alpha_spmm_alg1",
    "prepare_alpha_spmm_alg1_tle",
    "prepare_alpha_spmm_alg1_tle_opt",
    "prepare_alpha_spmm_alg1_tle_opt2",
    "benchmark_spmv_case",
    "benchmark_spmm_case",
    "benchmark_sddmm_case",
    "SUPPORTED_INDEX_DTYPES",
    "SUPPORTED_VALUE_DTYPES",
    "error_metrics",
    "validation_metrics",
    "PreparedCsrSpmv",
    "flagsparse_spmv_csr",
    "prepare_spmv_csr",
    "SPMM_CSR_ALGORITHMS",
    "PreparedCsrSpmmOpt",
    "PreparedCsrSpmmOptAlg2",
    "PreparedCsrSpmmRoute",
    "SpmmCsrAlgorithm",
    "SpmmCsrAlgorithmUnavailable",
    "flagsparse_spmm_csr",
    "flagsparse_spmm_csr_opt",
    "flagsparse_spmm_csr_opt_alg1",
    "flagsparse_spmm_csr_opt_alg1_preprocess",
    "flagsparse_spmm_csr_opt_alg2",
    "flagsparse_spmm_csr_opt_alg2_preprocess",
    "flagsparse_spmm_csr_run",
    "list_spmm_csr_algorithms",
    "prepare_spmm_csr_opt",
    "prepare_spmm_csr_opt_alg1",
    "prepare_spmm_csr_opt_alg1_preprocess",
    "prepare_spmm_csr_opt_alg2",
    "prepare_spmm_csr_opt_alg2_preprocess",
    "prepare_spmm_csr_route",
    "resolve_spmm_csr_algorithm",
    "SDDMMPrepared",
    "flagsparse_sddmm_csr",
    "prepare_sddmm_csr",
    "CsrMatrix",
    "flagsparse_spmm_csr_autograd",
    "spmm_csr_autograd",
    "csr_to_bsr",
    "quantize_bsr_fp8",
    "flagsparse_spmm_bsr",
    "coo_to_csr_for_spmv",
    "prepare_spmv_coo_tocsr",
    "flagsparse_spmv_coo_tocsr",
    "load_mtx_to_csr_torch",
    "PreparedBsrSpmv",
    "prepare_spmv_bsr",
    "flagsparse_spmv_bsr",
]
