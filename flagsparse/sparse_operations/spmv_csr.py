"""``flagsparse.sparse_operations.spmv_csr`` (the reference's
src/flagsparse/sparse_operations/spmv_csr.py): the names of that module on this
build, from ``flagsparse_amd.spmv_csr``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmv_csr as _impl
from flagsparse_amd.spmv_csr import *  # noqa: F401,F403
from flagsparse_amd.spmv_coo import flagsparse_spmv_coo_tocsr, prepare_spmv_coo_tocsr, coo_to_csr_for_spmv  # noqa: F401


def __getattr__(name):
    return getattr(_impl, name)
