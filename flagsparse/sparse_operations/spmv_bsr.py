"""``flagsparse.sparse_operations.spmv_bsr`` (the reference's
src/flagsparse/sparse_operations/spmv_bsr.py): the names of that module on this
build, from ``flagsparse_amd.spmv_bsr``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmv_bsr as _impl
from flagsparse_amd.spmv_bsr import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
