"""Matrix Market ingestion (CPU): the numpy loader against the oracle's
restatement of the reference loader (bit-equal values, identical structure)
and against scipy.io.mmread (an independent reader)."""

import io

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
import torch

from flagsparse_amd import mtx
from oracle import oracle as O


def _text(banner, n_rows, n_cols, entries, nnz=None, comments=("% a comment",)):
    out = [banner] + list(comments) + [f"{n_rows} {n_cols} {len(entries) if nnz is None else nnz}"]
    out += [" ".join(str(t) for t in e) for e in entries]
    return "\n".join(out) + "\n"


def _both(text):
    lines = text.splitlines(keepends=True)
    return mtx.mtx_to_csr_arrays(lines), O.mtx_to_csr_loop(lines)


def _same(a, b):
    (v1, c1, p1, s1), (v2, c2, p2, s2) = a, b
    assert s1 == s2
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(c1, c2)
    assert v1.tobytes() == v2.tobytes()  # bit-equal, signed zeros included


def _random_entries(rng, n_rows, n_cols, count, dup_frac=0.3):
    r = rng.integers(1, n_rows + 1, count)
    c = rng.integers(1, n_cols + 1, count)
    k = int(count * dup_frac)
    r[:k] = r[count - k:]
    c[:k] = c[count - k:]  # repeated coordinates
    v = rng.standard_normal(count) * 10 ** rng.uniform(-8, 8, count)
    return [(int(a), int(b), repr(float(x))) for a, b, x in zip(r, c, v)]


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric", "hermitian"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_loader_matches_reference_procedure(symmetry, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    ents = _random_entries(rng, n, n, int(rng.integers(1, 400)))
    text = _text(f"%%MatrixMarket matrix coordinate real {symmetry}", n, n, ents)
    _same(*_both(text))


def test_pattern_rectangular_out_of_range_and_nnz_cut():
    ents = [(1, 1), (2, 5), (9, 1), (3, 3), (3, 3), (0, 2), (4, 7), (2, 1)]
    # 9 and 0 are outside a 4 x 6 shape; only the first 7 lines count
    text = _text("%%MatrixMarket matrix coordinate pattern general", 4, 6, ents, nnz=7)
    got, want = _both(text)
    _same(got, want)
    v, c, p, shape = got
    assert shape == (4, 6) and v.tolist() == [1.0, 1.0, 2.0]  # (1,1), (2,5), (3,3) twice
    assert c.tolist() == [0, 4, 2] and p.tolist() == [0, 1, 2, 3, 3]


def test_many_duplicates_sum_in_file_order():
    vals = [1e16, 1.0, -1e16, 1.0, 3.0, 1e-16, -0.0]
    ents = [(2, 2, repr(x)) for x in vals] + [(1, 1, "-0.0")]
    got, want = _both(_text("%%MatrixMarket matrix coordinate real general", 3, 3, ents))
    _same(got, want)
    seq = 0.0
    for x in vals:
        seq = seq + x
    assert got[0][1] == seq and str(got[0][0]) == "0.0"  # 0.0 + (-0.0) = +0.0, as the reference


def test_symmetric_mirror_interleaves_with_explicit_entries():
    # (2,1) is given explicitly and also mirrored from (1,2): order matters
    ents = [(2, 1, "0.1"), (1, 2, "1e20"), (2, 1, "-1e20"), (3, 3, "2.5")]
    got, want = _both(_text("%%MatrixMarket matrix coordinate real symmetric", 3, 3, ents))
    _same(got, want)


def test_empty_and_header_errors():
    got, want = _both("%%MatrixMarket matrix coordinate real general\n3 4 0\n")
    _same(got, want)
    assert got[2].tolist() == [0, 0, 0, 0]
    with pytest.raises(ValueError, match="Cannot parse"):
        mtx.mtx_to_csr_arrays(["% only comments\n", "\n"])


def test_against_scipy_mmread():
    rng = np.random.default_rng(7)
    for symmetry in ("general", "symmetric", "skew-symmetric"):
        n = 40
        ents = _random_entries(rng, n, n, 300, dup_frac=0.2)
        if symmetry != "general":  # scipy wants the lower triangle only
            ents = [(max(a, b), min(a, b), v) for a, b, v in ents if not (symmetry == "skew-symmetric" and a == b)]
        text = _text(f"%%MatrixMarket matrix coordinate real {symmetry}", n, n, ents)
        v, c, p, shape = mtx.mtx_to_csr_arrays(text.splitlines(True))
        ref = sp.csr_matrix(scipy.io.mmread(io.StringIO(text)))
        ref.sum_duplicates()
        ref.sort_indices()
        mine = sp.csr_matrix((v, c, p), shape=shape)
        diff = abs(mine - ref)
        assert diff.max() <= 1e-9 * max(1.0, abs(ref).max()), symmetry


def test_load_to_torch(tmp_path):
    path = tmp_path / "m.mtx"
    path.write_text(_text("%%MatrixMarket matrix coordinate integer general", 3, 3, [(1, 2, 3), (3, 1, -4)]))
    data, idx, ptr, shape = mtx.load_mtx_to_csr_torch(str(path), dtype=torch.float64, device="cpu")
    assert data.tolist() == [3.0, -4.0] and idx.dtype == torch.int64 and ptr.tolist() == [0, 1, 1, 2]
    bad = tmp_path / "bad.mtx"
    bad.write_text("% nothing\n")
    with pytest.raises(ValueError, match="Cannot parse .mtx header"):
        mtx.load_mtx_to_csr_torch(str(bad), device="cpu")
