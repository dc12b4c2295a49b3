"""``flagsparse.sparse_operations.sddmm_csr`` (the reference's
src/flagsparse/sparse_operations/sddmm_csr.py): the names of that module on this
build, from ``flagsparse_amd.sddmm_csr``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.sddmm_csr as _impl
from flagsparse_amd.sddmm_csr import *  # noqa: F401,F403
from flagsparse_amd.benchmarks import benchmark_sddmm_case  # noqa: F401


def __getattr__(name):
    return getattr(_impl, name)
