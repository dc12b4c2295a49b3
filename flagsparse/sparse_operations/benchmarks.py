"""``flagsparse.sparse_operations.benchmarks`` (the reference's
src/flagsparse/sparse_operations/benchmarks.py): the names of that module on this
build, from ``flagsparse_amd.benchmarks``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd.benchmarks as _impl
from flagsparse_amd.benchmarks import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
