"""The gather probe (fsp_gather_rows_probe) through the C ABI: every row size
runs and leaves the sink untouched for an all-ones table (its lane sums are
small integers), and bad arguments are rejected with FSP_ERR_ARG."""

import pytest
import torch

from flagsparse_amd import _lib

pytestmark = pytest.mark.gpu

FSP_ERR_ARG = -1


@pytest.mark.parametrize("row_bytes", [16, 32, 64, 128, 256, 512, 1024])
def test_probe_runs_for_every_row_size(device, row_bytes):
    n_rows = 1000
    table = torch.ones(n_rows * row_bytes // 4, dtype=torch.float32, device=device)
    sink = torch.zeros(1, dtype=torch.float32, device=device)
    lib = _lib.load()
    stream = torch.cuda.current_stream(device).cuda_stream
    for n_gathers in (0, 1, 777, 100000):
        _lib.check(lib.fsp_gather_rows_probe(table.data_ptr(), n_rows, row_bytes, n_gathers, sink.data_ptr(), stream),
                   "fsp_gather_rows_probe")
    torch.cuda.synchronize()
    assert float(sink.item()) == 0.0
    assert torch.all(table == 1.0)  # read only


def test_probe_rejects_bad_arguments(device):
    table = torch.ones(4096, dtype=torch.float32, device=device)
    sink = torch.zeros(1, dtype=torch.float32, device=device)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    assert lib.fsp_gather_rows_probe(table.data_ptr(), 64, 24, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), 64, 2048, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), 0, 256, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr() + 4, 32, 256, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), 64, 256, -1, sink.data_ptr(), s) == FSP_ERR_ARG
    assert b"row_bytes" in lib.fsp_last_error() or b"n_rows" in lib.fsp_last_error()
