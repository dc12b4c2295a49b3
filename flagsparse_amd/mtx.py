"""Matrix Market (.mtx) ingestion and a per-file SpMV / SpMM sweep.

``load_mtx_to_csr_torch`` has the semantics of the reference's loader of the
same name (tests/test_spmv.py:116-189, also tests/test_spmm.py): the
``%%MatrixMarket`` banner's field and symmetry; comments and blank lines
skipped anywhere; the first non-comment line gives ``n_rows n_cols [nnz]``;
only the first ``nnz`` entry lines are read; entries outside the shape are
dropped; ``pattern`` entries are 1.0; ``symmetric`` / ``hermitian`` mirror
every off-diagonal entry, ``skew-symmetric`` mirrors it negated; duplicate
(row, column) entries are summed in float64 in file order starting from 0.0
(a mirrored entry right after the entry it mirrors), then cast to the value
dtype; columns ascend inside each row; indices / indptr are int64.  An empty
file (nnz = 0) gives an empty CSR.  The reference accumulates into one
Python dict per row; this build parses with numpy and sums each (row, col)
group level by level, which keeps the file-order summation (bit-equal).

``run_one_mtx`` / ``python -m flagsparse_amd.mtx FILES...`` run the HIP
kernels on each file and report time, GB/s and the error against torch's
sparse product on the same device in float64 (the reference's
``run_one_mtx`` reports the same fields, tests/test_spmv.py:501-693).
"""

import argparse
import math
import os
import sys
import time

import numpy as np
import torch


def _parse_header_and_entries(lines):
    field, symmetry = "real", "general"
    header = None
    entries = []
    for line in lines:
        s = line.strip()
        if s.startswith("%%MatrixMarket"):
            tok = s.split()
            if len(tok) >= 5:
                field, symmetry = tok[3].lower(), tok[4].lower()
            continue
        if s.startswith("%"):
            continue
        if header is None and s:
            parts = s.split()
            header = (int(parts[0]), int(parts[1]), int(parts[2]) if len(parts) > 2 else 0)
            continue
        if s:
            entries.append(s)
    return field, symmetry, header, entries


def mtx_to_csr_arrays(lines):
    """(values float64, indices int64, indptr int64, (n_rows, n_cols)) from
    the lines of a .mtx file (host numpy; the loader's core)."""
    field, symmetry, header, entries = _parse_header_and_entries(lines)
    if header is None:
        raise ValueError("Cannot parse .mtx header")
    n_rows, n_cols, nnz = header
    if nnz == 0:
        return np.zeros(0), np.zeros(0, np.int64), np.zeros(n_rows + 1, np.int64), (n_rows, n_cols)
    entries = entries[:nnz]
    split = [e.split() for e in entries]
    r = np.array([p[0] for p in split], dtype=np.int64) - 1
    c = np.array([p[1] for p in split], dtype=np.int64) - 1
    if field == "pattern":
        v = np.ones(len(split))
    else:
        v = np.array([float(p[2]) for p in split], dtype=np.float64)
    keep = (r >= 0) & (r < n_rows) & (c >= 0) & (c < n_cols)
    seq = np.arange(r.size, dtype=np.int64) * 2  # file order; a mirror sits at 2k + 1
    rows, cols, vals, order = [r[keep]], [c[keep]], [v[keep]], [seq[keep]]
    if symmetry in ("symmetric", "hermitian", "skew-symmetric"):
        m = keep & (r != c) & (c < n_rows) & (r < n_cols)
        rows.append(c[m])
        cols.append(r[m])
        vals.append(-v[m] if symmetry == "skew-symmetric" else v[m])
        order.append(seq[m] + 1)
    rows, cols, vals, order = (np.concatenate(a) for a in (rows, cols, vals, order))
    # group by (row, col) with file order inside a group
    perm = np.lexsort((order, cols, rows))
    rows, cols, vals = rows[perm], cols[perm], vals[perm]
    if rows.size:
        new = np.ones(rows.size, dtype=bool)
        new[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
        starts = np.flatnonzero(new)
    else:
        starts = np.zeros(0, np.int64)
    sizes = np.diff(np.append(starts, rows.size))
    acc = 0.0 + vals[starts]  # the reference starts every sum from 0.0
    level = 1
    live = np.flatnonzero(sizes > level)
    while live.size:  # add the level-th duplicate of every group that has one
        acc[live] = acc[live] + vals[starts[live] + level]
        level += 1
        live = live[sizes[live] > level]
    out_rows, out_cols = rows[starts], cols[starts]
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(np.bincount(out_rows, minlength=n_rows), out=indptr[1:])
    return acc, out_cols.astype(np.int64), indptr, (n_rows, n_cols)


def load_mtx_to_csr_torch(file_path, dtype=torch.float32, device=None):
    """Load a Matrix Market file into CSR torch tensors (data, indices,
    indptr, shape) on `device` (default: cuda when available)."""
    if device is None:
        device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    with open(file_path, "r", encoding="utf-8") as f:
        lines = f.readlines()
    try:
        vals, cols, indptr, shape = mtx_to_csr_arrays(lines)
    except ValueError as exc:
        if str(exc) == "Cannot parse .mtx header":
            raise ValueError(f"Cannot parse .mtx header: {file_path}") from None
        raise
    data = torch.tensor(vals, dtype=torch.float64).to(dtype=dtype, device=device)
    return (data, torch.from_numpy(cols).to(device), torch.from_numpy(indptr).to(device), shape)


# ------------------------------------------------------------------ sweep
# max|d| / max|ref| bounds of a PASS (FP32 / FP64 / BF16 / FP16 storage)
_TOLERANCE = {torch.float32: 1e-5, torch.float64: 1e-12, torch.bfloat16: 1e-2, torch.float16: 5e-3}


def _time_ms(fn, warmup, iters):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(iters):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / iters


def run_one_mtx(mtx_path, value_dtype=torch.float32, index_dtype=torch.int32, warmup=10, iters=50, op="non",
                kind="spmv", n_dense=32):
    """Run one .mtx through the HIP SpMV (or SpMM with n_dense columns) and
    torch's sparse product on the same device (float64 reference)."""
    from ._common import normalize_op, op_transposes
    from .spmm_csr import flagsparse_spmm_csr_run, prepare_spmm_csr_route
    from .spmv_csr import flagsparse_spmv_csr, prepare_spmv_csr

    op = normalize_op(op)
    device = torch.device("cuda")
    data, indices, indptr, shape = load_mtx_to_csr_torch(mtx_path, dtype=value_dtype, device=device)
    indices = indices.to(index_dtype)
    indptr = indptr.to(index_dtype)
    n_rows, n_cols = shape
    nnz = data.numel()
    transpose = op_transposes(op)
    in_size, out_size = (n_rows, n_cols) if transpose else (n_cols, n_rows)
    g = torch.Generator(device=device)
    g.manual_seed(0)
    if kind == "spmv":
        x = torch.rand(in_size, generator=g, device=device, dtype=torch.float32).to(value_dtype) * 2 - 1
        prepared = prepare_spmv_csr(data, indices, indptr, shape, op=op)
        run = lambda: flagsparse_spmv_csr(x=x, prepared=prepared)  # noqa: E731
        dense_in = x.unsqueeze(1)
    else:
        x = torch.randn(in_size, n_dense, generator=g, device=device, dtype=torch.float32).to(value_dtype)
        prepared = prepare_spmm_csr_route(data, indices, indptr, shape, op=op)
        run = lambda: flagsparse_spmm_csr_run(prepared, x)  # noqa: E731
        dense_in = x
    y = run()
    ms = _time_ms(run, warmup, iters)

    def torch_csr(values):
        A = torch.sparse_csr_tensor(indptr.to(torch.int64), indices.to(torch.int64), values, size=shape,
                                    device=device)
        return A.to_sparse_coo().t().coalesce().to_sparse_csr() if transpose else A

    ref = (torch_csr(data.to(torch.float64)) @ dense_in.to(torch.float64)).reshape(y.shape)
    diff = float((y.to(torch.float64) - ref).abs().max()) if y.numel() else 0.0
    mag = float(ref.abs().max()) if ref.numel() else 0.0
    rel = 0.0 if diff == 0 else (diff / mag if mag else math.inf)
    try:  # torch's sparse product on the same device and dtype, as a baseline
        A_v = torch_csr(data)
        torch_ms = _time_ms(lambda: A_v @ dense_in, warmup, iters)
    except (RuntimeError, NotImplementedError):
        torch_ms = None
    s_v = data.element_size()
    s_i = indices.element_size()
    cols_out = 1 if kind == "spmv" else n_dense
    nbytes = nnz * (s_v + s_i) + (n_rows + 1) * indptr.element_size() + (in_size + out_size) * cols_out * s_v
    tol = _TOLERANCE.get(value_dtype, 1e-5)
    return {
        "path": mtx_path, "shape": shape, "nnz": nnz, "op": op, "transpose": transpose, "kind": kind,
        "hip_ms": ms, "torch_sparse_ms": torch_ms, "gbs": nbytes / (ms * 1e-3) / 1e9 if ms else None,
        "rel_err": rel, "status": "PASS" if rel <= tol else "FAIL",
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description="HIP SpMV / SpMM over Matrix Market files")
    ap.add_argument("paths", nargs="+", help=".mtx files or directories")
    ap.add_argument("--kind", choices=("spmv", "spmm"), default="spmv")
    ap.add_argument("--op", default="non")
    ap.add_argument("--dtype", choices=("float32", "float64", "bfloat16", "float16"), default="float32")
    ap.add_argument("--index-dtype", choices=("int32", "int64"), default="int32")
    ap.add_argument("--n-dense", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--iters", type=int, default=50)
    args = ap.parse_args(argv)
    files = []
    for p in args.paths:
        if os.path.isdir(p):
            files += sorted(os.path.join(p, f) for f in os.listdir(p) if f.endswith(".mtx"))
        else:
            files.append(p)
    print(f"{'file':40s} {'shape':>20s} {'nnz':>12s} {'hip ms':>9s} {'GB/s':>9s} {'torch ms':>9s} {'rel err':>10s} status")
    bad = 0
    for f in files:
        t0 = time.perf_counter()
        r = run_one_mtx(f, getattr(torch, args.dtype), getattr(torch, args.index_dtype), args.warmup, args.iters,
                        args.op, args.kind, args.n_dense)
        bad += r["status"] != "PASS"
        tms = f"{r['torch_sparse_ms']:.4f}" if r["torch_sparse_ms"] is not None else "n/a"
        print(f"{os.path.basename(f)[:40]:40s} {str(r['shape']):>20s} {r['nnz']:>12d} {r['hip_ms']:>9.4f} "
              f"{r['gbs']:>9.1f} {tms:>9s} {r['rel_err']:>10.2e} {r['status']} ({time.perf_counter() - t0:.1f}s)",
              flush=True)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
