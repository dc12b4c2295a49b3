"""``flagsparse``: the reference's import surface on the MI355X build.

Code written against FlagSparse imports it unchanged -- ``import
flagsparse``, ``from flagsparse import flagsparse_spmm_csr``, ``import
flagsparse.sparse_operations.spmv_csr as m`` -- and reaches the gfx950 HIP
kernels of ``flagsparse_amd``.  Like the reference's facade
(src/flagsparse/__init__.py:156-306) it resolves names lazily: the operator
names of the CSR SpMV / SpMM / SDDMM path (and the formats and routes built
beside it) come from ``flagsparse_amd``; the reference's names outside that
path (gather / scatter, SpGEMM, SpSV, SpSM and the CuPy format containers)
resolve to placeholders whose call raises ``NotImplementedError``.  Names
this build adds beyond the reference (the differentiable SpMM, the BSR SpMM
on the matrix cores, ``load_mtx_to_csr_torch``) are exported as well.
"""

__version__ = "1.0.0"

from . import _unsupported  # noqa: E402


def _implemented():
    import flagsparse_amd  # noqa: PLC0415

    return set(flagsparse_amd.__all__)


__all__ = sorted(_implemented() | set(_unsupported.NAMES))


def __getattr__(name):
    if name == "sparse_operations":
        import importlib  # noqa: PLC0415

        return importlib.import_module(".sparse_operations", __name__)
    if name in _implemented():
        import flagsparse_amd  # noqa: PLC0415

        return getattr(flagsparse_amd, name)
    if name in _unsupported.NAMES:
        return _unsupported.get(name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


def __dir__():
    return sorted(set(globals()) | set(__all__))
