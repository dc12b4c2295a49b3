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


@pytest.mark.parametrize("n_arrays", [1, 2])
@pytest.mark.parametrize("contiguous", [0, 1])
def test_stream_probe_reads_and_writes_nothing(device, n_arrays, contiguous):
    lib = _bench_lib.load()
    a = torch.zeros(4096 * 4 + 8, dtype=torch.float32, device=device)  # 4098 chunks: a ragged tail per grid
    b = torch.zeros_like(a)
    sink = torch.zeros(1, dtype=torch.float32, device=device)
    s = torch.cuda.current_stream(device).cuda_stream
    for wgs in (1, 3, 512):
        assert lib.fsp_stream_read_probe(a.data_ptr(), b.data_ptr(), a.numel() // 4, n_arrays, wgs, contiguous,
                                         sink.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert float(sink.item()) == 0.0 and torch.all(a == 0) and torch.all(b == 0)
    assert lib.fsp_stream_read_probe(a.data_ptr() + 4, b.data_ptr(), 8, n_arrays, 1, contiguous, sink.data_ptr(),
                                     s) == FSP_ERR_ARG
    assert lib.fsp_stream_read_probe(a.data_ptr(), b.data_ptr(), 8, 3, 1, contiguous, sink.data_ptr(), s) == FSP_ERR_ARG


@pytest.mark.parametrize("n_slabs,spin", [(1, -1), (4, 1 << 20), (4, -1)])
def test_slab_probe_matches_spmm(device, n_slabs, spin):
    """The column-slab schedule probe computes the same SpMM (N = 64) as
    flagsparse_spmm_csr, with and without its XCD phase counter."""
    import numpy as np

    import flagsparse_amd as fs
    from flagsparse_amd import synthetic

    ip, ix, v, Bn = synthetic.power_law_rows(n_rows=3000, n_cols=3000, cap=600, seed=9)
    n = 3000
    data, idx, ptr = synthetic.to_torch(ip, ix, v, device)
    B = synthetic.dense(Bn, device)
    ref = fs.flagsparse_spmm_csr(data, idx, ptr, B, (n, n))
    lens = np.diff(ip)
    assert lens.max() <= 65535
    # one virtual row per row, longest first
    order = np.argsort(-lens, kind="stable")
    vb = ip[:-1][order]
    vlen = lens[order]
    vid = np.repeat(np.arange(n), vlen)
    pos = np.repeat(vb - (np.cumsum(vlen) - vlen), vlen) + np.arange(vid.size)
    cnt = np.bincount(vid * n_slabs + ix[pos].astype(np.int64) * n_slabs // n, minlength=n * n_slabs)
    slab_end = torch.from_numpy(np.cumsum(cnt.reshape(n, n_slabs), axis=1).astype(np.uint16).view(np.int16)).to(device)
    vbeg = torch.from_numpy(vb.astype(np.int32)).to(device)
    vslot = torch.from_numpy(order.astype(np.int32)).to(device)
    C = torch.full((n, 64), float("nan"), device=device)
    counters = torch.zeros(8, dtype=torch.int32, device=device)
    flag = torch.zeros(1, dtype=torch.int32, device=device)
    wgs = torch.cuda.get_device_properties(device).multi_processor_count
    st = _bench_lib.load().fsp_slab_spmm_probe(8, vbeg.data_ptr(), slab_end.data_ptr(), vslot.data_ptr(),
                                               idx.data_ptr(), data.data_ptr(), B.data_ptr(), C.data_ptr(),
                                               C.data_ptr(), n, n, n_slabs, wgs, counters.data_ptr(), spin,
                                               flag.data_ptr(), torch.cuda.current_stream(device).cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert int(flag.item()) == 0
    assert float((C - ref).abs().max() / ref.abs().max()) <= 1e-6
