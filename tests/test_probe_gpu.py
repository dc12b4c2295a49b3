"""The gather probe of the measurement library (fsp_gather_rows_probe in
libflagsparse_bench.so) through its C ABI: every row size runs and leaves the
sink untouched for an all-ones table (its lane sums are small integers), and
bad arguments, a table smaller than n_rows x row_bytes included, are rejected
with -1."""

import pytest
import torch

from flagsparse_amd import _bench_lib

pytestmark = pytest.mark.gpu

FSP_ERR_ARG = -1


@pytest.mark.parametrize("row_bytes", [16, 32, 64, 128, 256, 512, 1024])
def test_probe_runs_for_every_row_size(device, row_bytes):
    n_rows = 1000
    table = torch.ones(n_rows * row_bytes // 4, dtype=torch.float32, device=device)
    sink = torch.zeros(1, dtype=torch.float32, device=device)
    lib = _bench_lib.load()
    stream = torch.cuda.current_stream(device).cuda_stream
    for n_gathers in (0, 1, 777, 100000):
        _bench_lib.check(lib.fsp_gather_rows_probe(table.data_ptr(), table.numel() * 4, n_rows, row_bytes, n_gathers,
                                                   sink.data_ptr(), stream), "fsp_gather_rows_probe")
    torch.cuda.synchronize()
    assert float(sink.item()) == 0.0
    assert torch.all(table == 1.0)  # read only


def test_probe_rejects_bad_arguments(device):
    table = torch.ones(4096, dtype=torch.float32, device=device)
    sink = torch.zeros(1, dtype=torch.float32, device=device)
    lib = _bench_lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    tb = table.numel() * 4
    assert lib.fsp_gather_rows_probe(table.data_ptr(), tb, 64, 24, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), tb, 64, 2048, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), tb, 0, 256, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr() + 4, tb - 4, 32, 256, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), tb, 64, 256, -1, sink.data_ptr(), s) == FSP_ERR_ARG
    assert lib.fsp_gather_rows_probe(table.data_ptr(), tb, 65, 256, 10, sink.data_ptr(), s) == FSP_ERR_ARG
    assert b"table_bytes" in lib.fsp_bench_last_error()
