"""``flagsparse.sparse_operations.spmm_coo`` (the reference's
src/flagsparse/sparse_operations/spmm_coo.py): the names of that module on this
build, from ``flagsparse_amd.spmm_coo``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.spmm_coo as _impl
from flagsparse_amd.spmm_coo import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
