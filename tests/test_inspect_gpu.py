"""The one-readback CSR inspection (fsp_csr_inspect + row bins): every
figure against numpy on valid and broken row pointers, and the plan it
returns against build_plan's."""

import numpy as np
import pytest
import torch

from flagsparse_amd import _plan, synthetic

from ._helpers import long_row_matrix

pytestmark = pytest.mark.gpu


def _np_stats(ip, ix):
    ip = np.asarray(ip, dtype=np.int64)
    ix = np.asarray(ix, dtype=np.int64)
    d = np.diff(ip)
    return {
        "indptr_first": int(ip[0]),
        "indptr_last": int(ip[-1]),
        "n_decreasing": int((d < 0).sum()),
        "min_col": int(ix.min()) if ix.size else None,
        "max_col": int(ix.max()) if ix.size else None,
        "max_row": int(max(0, d.max())) if d.size else 0,
    }


@pytest.mark.parametrize("pt", [torch.int32, torch.int64], ids=["p32", "p64"])
@pytest.mark.parametrize("it", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_inspect_matches_numpy(device, pt, it):
    rng = np.random.default_rng(7)
    cases = []
    ip, ix, _ = synthetic.random_csr(3000, 5000, 0.01, seed=3)
    cases.append((ip, ix))
    bad = ip.copy()
    bad[10], bad[2000] = bad[11] + 5, bad[2001] + 3  # two decreasing steps at least
    cases.append((bad, ix))
    neg = ix.copy()
    neg[rng.integers(0, neg.size, 3)] = -4  # out-of-range columns
    cases.append((ip, neg))
    cases.append((ip + 1, ix))  # indptr[0] != 0
    cases.append((np.zeros(9, dtype=np.int64), np.zeros(0, dtype=np.int64)))  # no nonzeros
    cases.append((np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int64)))  # no rows
    for ip_, ix_ in cases:
        p = torch.tensor(ip_, dtype=pt, device=device)
        i = torch.tensor(ix_, dtype=it, device=device)
        stats, plan = _plan.inspect_csr(p, i, ip_.size - 1, ix_.size)
        assert stats == _np_stats(ip_, ix_)
        assert plan is None


def test_inspect_plan_equals_build_plan(device):
    ip, ix, _ = long_row_matrix(4000, 3000, long_rows=(5, 3999), long_len=2900, short_len=9, seed=2)
    p = torch.tensor(ip, dtype=torch.int32, device=device)
    i = torch.tensor(ix, dtype=torch.int32, device=device)
    for cfg in (_plan.spmv_config(), _plan.spmm_config(16), _plan.spmm_config(4, long_threshold=64, seg_len=64)):
        stats, plan = _plan.inspect_csr(p, i, 4000, ix.size, cfg)
        ref = _plan.build_plan(p, 4000, cfg)
        assert stats["max_row"] == int(np.diff(ip).max())
        assert plan.bin_counts == ref.bin_counts and plan.identity == ref.identity
        assert (plan.n_segments, plan.n_long_rows) == (ref.n_segments, ref.n_long_rows)
        for a, b in ((plan.seg_row, ref.seg_row), (plan.seg_begin, ref.seg_begin), (plan.seg_end, ref.seg_end),
                     (plan.long_rows, ref.long_rows), (plan.long_seg_ptr, ref.long_seg_ptr)):
            assert torch.equal(a.cpu(), b.cpu())
        if not plan.identity:
            assert torch.equal(plan.rows.cpu(), ref.rows.cpu())
    # an invalid row pointer gives the figures but no plan
    bad = torch.tensor(ip[::-1].copy(), dtype=torch.int32, device=device)
    stats, plan = _plan.inspect_csr(bad, i, 4000, ix.size, _plan.spmv_config())
    assert stats["n_decreasing"] > 0 and plan is None
