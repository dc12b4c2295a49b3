"""Autotuner bookkeeping and batch host logic (no GPU): shape classes, the
candidate grids, the versioned on-disk cache and the offset-table checks."""

import json

import numpy as np
import pytest
import torch

from flagsparse_amd import _plan, autotune, batch, synthetic


@pytest.fixture
def cache_file(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "library_hash", lambda: "h1")
    autotune.clear_memory()
    yield path
    autotune.clear_memory()


def test_shape_class_buckets():
    k1 = autotune.shape_class("spmm", torch.float32, torch.int32, torch.int32, 32, 7_000_000, 239_000_000, 64, 256)
    k2 = autotune.shape_class("spmm", torch.float32, torch.int32, torch.int32, 32, 7_100_000, 241_000_000, 63, 256)
    k3 = autotune.shape_class("spmm", torch.float32, torch.int32, torch.int32, 64, 7_000_000, 239_000_000, 64, 256)
    assert k1 == k2 != k3
    assert k1.startswith("spmm|float32|int32|int32|n32|") and k1.endswith("|cu256")
    assert autotune.shape_class("spmv", torch.float64, torch.int64, torch.int64, 1, 0, 0, 0, 8).count("-1") == 3


def test_candidates_and_configs():
    sv = autotune.candidates("spmv")
    assert len(sv) == len(set(sv)) == 36
    assert {p.variant for p in sv} == {0, 1}
    sm = autotune.candidates("spmm", 32, 4)
    assert len(sm) == len(set(sm)) == 48
    assert {p.xcd_groups for p in sm} == {0, 8}
    assert {p.lanes_n for p in sm} == {8, 4}
    assert {p.lanes_n for p in autotune.candidates("spmm", 8, 4)} == {2}
    p = autotune.TuneParams(unroll=8, slot_scale=2, long_threshold=512)
    cfg = p.spmm_config(8)
    assert cfg == _plan.spmm_config(8, unroll=8, long_threshold=512, seg_len=512, slot_scale=2)
    assert autotune.TuneParams().spmv_config() == _plan.spmv_config()
    assert autotune.TuneParams().max_blocks(None) == 0


def test_cache_round_trip_and_invalidation(cache_file, monkeypatch):
    p = autotune.TuneParams(unroll=8, slot_scale=1, lanes_n=4, blocks_per_cu=8)
    assert autotune.lookup("k") is None
    autotune.store("k", p, 1.25)
    assert autotune.lookup("k") == p
    data = json.loads(cache_file.read_text())
    assert data["k"]["params"]["unroll"] == 8 and data["k"]["ms"] == 1.25
    autotune.clear_memory()
    assert autotune.lookup("k") == p  # from the file
    # another library build: the entry is stale
    autotune.clear_memory()
    monkeypatch.setattr(autotune, "library_hash", lambda: "h2")
    assert autotune.lookup("k") is None


def test_cache_disabled_and_corrupt(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "library_hash", lambda: "h1")
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", "")
    autotune.clear_memory()
    autotune.store("k", autotune.TuneParams())
    assert autotune.lookup("k") == autotune.TuneParams()  # memory only
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", str(bad))
    autotune.clear_memory()
    assert autotune.lookup("k") is None
    autotune.store("k", autotune.TuneParams(unroll=2))
    assert json.loads(bad.read_text())["k"]["params"]["unroll"] == 2


def test_tune_picks_fastest_without_gpu(cache_file, monkeypatch):
    cands = [autotune.TuneParams(unroll=u) for u in (2, 4, 8)]
    fake = {2: 3.0, 4: 1.0, 8: 2.0}
    calls = []

    def run(p):
        calls.append(p)

    def timed(fn, reps=5):  # stands in for the device-event timer
        fn()
        return fake[calls[-1].unroll]

    monkeypatch.setattr(autotune, "time_launch", timed)
    best, results = autotune.tune("key", cands, run)
    assert best.unroll == 4 and len(results) == 3
    again, none = autotune.tune("key", cands, run)
    assert again == best and none is None  # cached: nothing re-timed
    assert len(calls) == 3


def test_batch_offset_checks():
    assert batch._host_offsets([0, 3, 3, 7], "x").tolist() == [0, 3, 3, 7]
    with pytest.raises(ValueError, match="start at 0"):
        batch._host_offsets([1, 3], "x")
    with pytest.raises(ValueError, match="nondecreasing"):
        batch._host_offsets([0, 3, 2], "x")
    with pytest.raises(ValueError, match="entries"):
        batch._host_offsets([0], "x")
    with pytest.raises(ValueError, match="entries"):
        batch._host_offsets([0, 1, 2], "x", n=2)


def test_batch_rejects_cpu_tensors():
    with pytest.raises(ValueError, match="CUDA"):
        batch.CsrBatch(torch.ones(2), torch.zeros(2, dtype=torch.int32), torch.tensor([0, 2], dtype=torch.int32),
                       [0, 1])


def test_csr_batch_generator_structure():
    ip, ix, v, ro, no = (t.numpy() for t in synthetic.csr_batch(n_mats=40, rows_lo=8, rows_hi=300, seed=4))
    n = np.diff(ro)
    assert n.min() >= 8 and n.max() <= 300
    assert ip.size == ro[-1] + 40 and ix.size == v.size == no[-1]
    assert v.dtype == np.float32 and np.abs(v).max() <= 1.0
    for m in range(40):
        lp = ip[ro[m] + m: ro[m + 1] + m + 1].astype(np.int64)
        assert lp[0] == 0 and lp[-1] == no[m + 1] - no[m]
        lens = np.diff(lp)
        assert lens.min() >= 4 and lens.max() <= 64
        cols = ix[no[m]: no[m + 1]]
        for r in range(n[m]):
            c = cols[lp[r]: lp[r + 1]]
            assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < n[m]
    again = synthetic.csr_batch(n_mats=40, rows_lo=8, rows_hi=300, seed=4)
    assert torch.equal(again[1], torch.from_numpy(ix))  # seeded


def test_log_uniform_rows_and_lengths_distribution():
    ip, ix, v, ro, no = synthetic.csr_batch(n_mats=2000, rows_lo=64, rows_hi=512, seed=5)
    n = np.diff(ro.numpy())
    # log-uniform sizes: mean of ln n close to the midpoint of [ln 64, ln 513]
    assert abs(np.log(n).mean() - 0.5 * (np.log(64) + np.log(513))) < 0.05
    lens = np.diff(ip.numpy().astype(np.int64))
    lens = lens[lens >= 0]  # drop the jumps back to 0 between matrices
    assert abs(lens.mean() - 34) < 0.2


def test_batch_parts_table_covers_rows_once():
    import numpy as np

    from flagsparse_amd.batch import parts_table

    rng = np.random.default_rng(0)
    for trial in range(50):
        n_mats = int(rng.integers(1, 40))
        n_rows = rng.integers(0, 300, n_mats) * (rng.random(n_mats) > 0.15)  # some empty matrices
        n_cols = rng.integers(0, 500, n_mats)
        nnz_m = np.where(n_rows > 0, rng.integers(0, 20000, n_mats), 0)
        rows = np.concatenate([[0], np.cumsum(n_rows)])
        cols = np.concatenate([[0], np.cumsum(n_cols)])
        nnz = np.concatenate([[0], np.cumsum(nnz_m)])
        part_nnz = int(rng.choice([1, 64, 1000, 1 << 15]))
        table, start = parts_table(rows, cols, nnz, part_nnz)
        assert table.shape == (start[-1], 4) and start.shape == (n_mats + 1,)
        covered = np.zeros(rows[-1], dtype=np.int64)
        for m in range(n_mats):
            own = table[start[m]:start[m + 1]]
            if n_rows[m] == 0:
                assert own.shape[0] == 0
                continue
            assert own.shape[0] == min(n_rows[m], max(1, -(-nnz_m[m] // part_nnz)))
            assert own[0, 0] == rows[m] and own[-1, 1] == rows[m + 1]
            assert (own[:, 1] > own[:, 0]).all() and (own[1:, 0] == own[:-1, 1]).all()  # contiguous, non-empty
            assert (own[:, 2] == cols[m]).all() and (own[:, 3] == cols[m + 1]).all()
            for rb, re, _, _ in own:
                covered[rb:re] += 1
        assert (covered == 1).all()
