"""``flagsparse.sparse_operations.spmv_csc`` (the reference's
src/flagsparse/sparse_operations/spmv_csc.py): the names of that module on this
build, from ``flagsparse_amd.spmv_csc``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmv_csc as _impl
from flagsparse_amd.spmv_csc import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
