"""Names of the reference's public surface that lie outside the CSR SpMV /
SpMM / SDDMM path this MI355X build implements.

They exist so that ``import flagsparse`` and ``from flagsparse import *``
resolve every name the reference exports (src/flagsparse/__init__.py:19-154),
but calling one raises ``NotImplementedError`` naming the family; nothing
here computes anything.
"""

# family -> the reference's names in it (src/flagsparse/__init__.py:19-154,
# src/flagsparse/sparse_operations/__init__.py:36-137)
FAMILIES = {
    "gather / scatter": (
        "flagsparse_gather", "flagsparse_scatter", "pytorch_index_gather", "pytorch_index_scatter",
        "cusparse_spmv_gather", "cusparse_spmv_scatter", "cusparse_native_gather", "benchmark_gather_case",
        "benchmark_scatter_case", "benchmark_performance", "comprehensive_gather_test",
        "comprehensive_scatter_test",
    ),
    "SpGEMM": ("SpGEMMPrepared", "prepare_spgemm_csr", "flagsparse_spgemm_csr", "benchmark_spgemm_case"),
    "SpSV (triangular solve)": (
        "flagsparse_spsv_csr", "flagsparse_spsv_coo", "flagsparse_spsv_sell", "flagsparse_spsv_buffer_size",
        "flagsparse_spsv_buffer_size_ex", "flagsparse_spsv_analysis_csr", "flagsparse_spsv_analysis_coo",
        "flagsparse_spsv_analysis_sell", "flagsparse_spsv_analysis_ex", "flagsparse_spsv_solve_csr",
        "flagsparse_spsv_solve_coo", "flagsparse_spsv_solve_sell", "flagsparse_spsv_solve_ex",
        "flagsparse_spsv_preprocess_csr", "flagsparse_spsv_preprocess_coo", "flagsparse_spsv_create_workspace",
        "flagsparse_create_spsv_handle", "flagsparse_create_spmat_csr", "flagsparse_create_spmat_coo",
        "flagsparse_create_dnvec", "FlagSparseSpSVDescr", "FlagSparseSpSVHandle", "FlagSparseSpSVWorkspace",
        "FlagSparseSpMatDescr", "FlagSparseDnVecDescr",
    ),
    "SpSM (triangular solve, many right-hand sides)": (
        "flagsparse_spsm_csr", "flagsparse_spsm_coo", "benchmark_spsm_case", "comprehensive_spsm_test",
    ),
    "CuPy-backed sparse format containers": (
        "create_csr_matrix", "create_coo_matrix", "create_csc_matrix", "create_bsr_matrix", "create_sell_matrix",
        "create_blocked_ell_matrix", "coo_to_csr", "coo_to_csc", "coo_to_bsr", "coo_to_sell", "coo_to_blocked_ell",
        "CSRMatrix", "COOMatrix", "CSCMatrix", "BSRMatrix", "SELLMatrix", "BLOCKEDELLMatrix",
        "generate_random_sparse_matrix", "read_mtx_file",
    ),
}

NAMES = {name: family for family, names in FAMILIES.items() for name in names}

# in-scope replacements worth naming in the error message
_INSTEAD = {
    "read_mtx_file": "flagsparse.load_mtx_to_csr_torch (torch CSR tensors) covers the .mtx path",
    "create_csr_matrix": "pass data / indices / indptr tensors to flagsparse_spmv_csr / flagsparse_spmm_csr",
    "create_bsr_matrix": "flagsparse.csr_to_bsr builds the 16 x 16 BSR of the matrix-core SpMM",
    "coo_to_csr": "flagsparse.coo_to_csr_for_spmv converts COO to CSR on the GPU",
}


class Unsupported:
    """Callable placeholder of a reference name this build does not implement."""

    __slots__ = ("__name__", "family")

    def __init__(self, name, family):
        self.__name__ = name
        self.family = family

    def __call__(self, *args, **kwargs):
        hint = _INSTEAD.get(self.__name__)
        raise NotImplementedError(
            f"flagsparse.{self.__name__} ({self.family}) is outside the CSR SpMV / SpMM / SDDMM path that this "
            f"MI355X build implements" + (f"; {hint}" if hint else "")
        )

    def __repr__(self):
        return f"<unsupported flagsparse.{self.__name__} ({self.family})>"


def get(name):
    family = NAMES.get(name)
    if family is None:
        raise AttributeError(name)
    return Unsupported(name, family)
