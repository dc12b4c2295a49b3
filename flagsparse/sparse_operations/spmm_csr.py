"""``flagsparse.sparse_operations.spmm_csr`` (the reference's
src/flagsparse/sparse_operations/spmm_csr.py): the names of that module on this
build, from ``flagsparse_amd.spmm_csr``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmm_csr as _impl
from flagsparse_amd.spmm_csr import *  # noqa: F401,F403
from flagsparse_amd.benchmarks import benchmark_spmm_case, benchmark_spmm_opt_case, comprehensive_spmm_test  # noqa: F401


def __getattr__(name):
    return getattr(_impl, name)
