"""``flagsparse.sparse_operations._common`` (the reference's
src/flagsparse/sparse_operations/_common.py): the names of that module on this
build, from ``flagsparse_amd._common``.  Any other attribute, private ones
included, is looked up there."""

import flagsparse_amd._common as _impl
from flagsparse_amd._common import *  # noqa: F401,F403


def __getattr__(name):
    return getattr(_impl, name)
