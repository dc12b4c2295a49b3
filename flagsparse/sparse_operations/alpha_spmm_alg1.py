"""``flagsparse.sparse_operations.alpha_spmm_alg1`` (the reference's
src/flagsparse/sparse_operations/alpha_spmm_alg1.py): the names of that module on this
build, from ``flagsparse_amd.alpha_spmm_alg1``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.alpha_spmm_alg1 as _impl
from flagsparse_amd.alpha_spmm_alg1 import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
