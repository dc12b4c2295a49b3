"""The C-ABI library loads without a GPU and exports every declared symbol."""

import ctypes
import os
import re

import pytest
import torch  # noqa: F401  (the library binds to torch's HIP runtime)

from flagsparse_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(fsp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_symbols()
    for required in ("fsp_spmv_csr", "fsp_spmm_csr", "fsp_sddmm_csr", "fsp_csr_to_csc", "fsp_csr_row_bins"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.library_path())
    missing = [n for n in _declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(_declared_symbols()) == set(_lib.SIGNATURES)


def test_abi_version_and_error_channel():
    lib = _lib.load()
    assert lib.fsp_abi_version() == _lib.ABI_VERSION
    # A rejected call (no device work) reports through fsp_last_error.
    st = lib.fsp_csr_row_bins_workspace(10, 0, None)
    assert st == -1
    assert b"bad arguments" in lib.fsp_last_error()


def test_workspace_queries_need_no_device():
    lib = _lib.load()
    n = ctypes.c_int64(0)
    assert lib.fsp_csr_row_bins_workspace(1 << 20, 6, ctypes.byref(n)) == 0
    assert n.value >= 6 * 1024 * 8
    # The CSC workspace query asks rocPRIM for its temporary size, which reads
    # device properties: without a GPU it reports a launch error instead.
    st = lib.fsp_csr_to_csc_workspace(1000, 1000, 5000, _lib.FSP_I32, ctypes.byref(n))
    if torch.cuda.is_available():
        assert st == 0 and n.value >= 5000 * 12
    else:
        assert st in (0, -3)


def test_only_torch_hip_runtime_is_mapped():
    _lib.load()
    maps = open("/proc/self/maps").read()
    hip = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(hip) == 1, hip


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.FlagSparseLibraryError, match="missing"):
        _lib.load(str(tmp_path / "nope.so"))


def test_gather_probe_lives_in_the_bench_library():
    from flagsparse_amd import _bench_lib

    assert not hasattr(ctypes.CDLL(_lib.library_path()), "fsp_gather_rows_probe")
    lib = _bench_lib.load()
    assert lib.fsp_gather_rows_probe(None, 1 << 20, 64, 256, 10, None, None) == -1  # NULL table
    assert lib.fsp_gather_rows_probe(None, 1 << 20, 64, 24, 10, None, None) == -1
    assert b"row_bytes" in lib.fsp_bench_last_error()
    assert lib.fsp_gather_rows_probe(None, 1 << 40, 1 << 32, 256, 10, None, None) == -1
    assert b"n_rows" in lib.fsp_bench_last_error()
    # a table smaller than n_rows x row_bytes is refused before any launch
    assert lib.fsp_gather_rows_probe(None, 64 * 256 - 1, 64, 256, 10, None, None) == -1
    assert b"table_bytes" in lib.fsp_bench_last_error()


def test_noalloc_spmm_rejects_oversized_partials_without_a_device():
    lib = _lib.load()
    st = _lib.FspCsrPlan()
    st.n_bins = 0
    st.unroll = 4
    st.n_long_rows = 10
    st.n_segments = _lib.NOALLOC_SCRATCH_BYTES // (64 * 4) + 1  # just past the scratch at N = 64, f32
    status = lib.fsp_spmm_csr_noalloc(_lib.FSP_F32, _lib.FSP_I32, _lib.FSP_I32, None, None, None, 0, None, None, 64,
                                      None, 64, 10, 64, 100, 10, 16, ctypes.byref(st), None)
    assert status == -4
    assert b"static scratch" in lib.fsp_last_error()
    assert lib.fsp_spmm_csr_noalloc(_lib.FSP_F32, _lib.FSP_I32, _lib.FSP_I32, None, None, None, 0, None, None, 64,
                                    None, 64, 10, 64, 100, 10, 16, None, None) == -1


def test_fused_backward_entry_rejects_bad_arguments_without_a_device():
    lib = _lib.load()
    st = _lib.FspCsrPlan()
    st.n_bins = 0
    st.unroll = 4
    args = [_lib.FSP_F32, _lib.FSP_I32, _lib.FSP_I32, None, None, None, _lib.FSP_I32, None, None, 8, None, 8, None, 8,
            None]
    # negative sizes
    assert lib.fsp_spmm_csr_trans_sddmm(*args, -1, 8, 0, 0, 16, ctypes.byref(st), None) == -1
    assert b"negative size" in lib.fsp_last_error()
    # no plan
    assert lib.fsp_spmm_csr_trans_sddmm(*args, 0, 8, 0, 0, 16, None, None) == -1
    # nonzeros without arrays
    assert lib.fsp_spmm_csr_trans_sddmm(*args, 0, 8, 5, 0, 16, ctypes.byref(st), None) == -1
    assert b"NULL" in lib.fsp_last_error()
