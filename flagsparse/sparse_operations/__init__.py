"""``flagsparse.sparse_operations`` on this build: the reference's operator
package (src/flagsparse/sparse_operations/__init__.py:17-137) with its module
names -- ``spmv_csr``, ``spmm_csr``, ``spmm_csr_opt_alg2``, ``sddmm_csr``,
``_common``, ... -- resolving to the MI355X implementation in
``flagsparse_amd``.  Names of the reference's other operator families
(gather / scatter, SpGEMM, SpSV, SpSM) resolve to placeholders that raise
``NotImplementedError``.
"""

import flagsparse_amd as _amd

from .. import _unsupported
from . import (  # noqa: F401
    _common,
    alpha_spmm_alg1,
    benchmarks,
    sddmm_csr,
    spmm_coo,
    spmm_csr,
    spmm_csr_opt_alg2,
    spmv_bsr,
    spmv_coo,
    spmv_csc,
    spmv_csr,
)

_IMPLEMENTED = [n for n in _amd.__all__]
# the CuPy format containers live in the reference's sparse_formats module,
# not in sparse_operations
_FORMATS = set(_unsupported.FAMILIES["CuPy-backed sparse format containers"])
_UNSUPPORTED = [n for n in _unsupported.NAMES if n not in _FORMATS]

__all__ = sorted(set(_IMPLEMENTED) | set(_UNSUPPORTED))


def __getattr__(name):
    if name in _IMPLEMENTED:
        return getattr(_amd, name)
    if name in _UNSUPPORTED:
        return _unsupported.get(name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


def __dir__():
    return sorted(set(globals()) | set(__all__))
