"""``flagsparse.sparse_operations.spmv_coo`` (the reference's
src/flagsparse/sparse_operations/spmv_coo.py): the names of that module on this
build, from ``flagsparse_amd.spmv_coo``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmv_coo as _impl
from flagsparse_amd.spmv_coo import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
