"""Autotuner for the row-bin CSR kernels, cached per shape class.

The tuned parameters are the ones the kernels take from their plan:

* ``unroll``         nonzeros per lane (SpMV) or per slot (SpMM) per trip --
                     the nonzeros a wave has in flight (block-nnz);
* ``slot_scale``     SpMM row-bin thresholds: a row of length n gets the fewest
                     slots S with n <= S * unroll * slot_scale;
* ``lanes_n``        SpMM lanes along the dense dimension (block-n =
                     lanes_n * 16 bytes of B per row and pass);
* ``long_threshold`` rows longer than this are cut into segments of
                     ``seg_len`` nonzeros (= long_threshold);
* ``blocks_per_cu``  0 = one workgroup per row tile, else a persistent grid of
                     blocks_per_cu workgroups per compute unit, each walking a
                     contiguous run of rows;
* ``xcd_groups``     8 = cut each row bin into 8 contiguous regions, one per
                     XCD, so neighbouring rows share an L2 (0 = off);
* ``variant``        SpMV loop: 0 = software-pipelined, 1 = plain, 2 = plain
                     and 3 = pipelined with x staged in LDS (x of at most
                     64 KiB), 4 = vector chunks with x in LDS; for a batch,
                     5 = row-range parts with each matrix's column range of
                     x in LDS (``lanes_n`` lanes per row, ``part_nnz``
                     nonzeros per part).

A shape class is (kernel, value / index dtypes, dense width, rounded log2 of
the mean and maximum row length and of the row count, compute units).  The
winner of a class is kept in memory and, unless disabled, in a JSON file
(``FLAGSPARSE_AMD_TUNE_CACHE``, default ``~/.cache/flagsparse_amd/tune.json``)
whose entries carry the library's ABI version and a hash of the built
library, so a rebuilt kernel invalidates them.  Candidates are timed with
device events on the current stream (median of a few launches after one
warm-up launch).
"""

import hashlib
import json
import math
import os
import threading
from dataclasses import asdict, dataclass

import torch

from . import _lib, _plan

_CACHE_LOCK = threading.Lock()
_MEMORY = {}
_LIB_HASH = None


@dataclass(frozen=True)
class TuneParams:
    unroll: int = 4
    slot_scale: int = _plan.SLOT_SCALE
    lanes_n: int = 0  # 0 = lanes_for_dense(n_dense)
    long_threshold: int = 0  # 0 = the kernel family's default
    blocks_per_cu: int = 0
    xcd_groups: int = 0
    variant: int = 0
    part_nnz: int = 0  # batch SpMV variant 5: nonzeros per row-range part

    def spmv_config(self):
        lt = self.long_threshold or _plan.SPMV_LONG_THRESHOLD
        return _plan.spmv_config(unroll=self.unroll, long_threshold=lt, seg_len=lt)

    def spmm_config(self, lanes_n):
        lt = self.long_threshold or _plan.SPMM_LONG_THRESHOLD
        return _plan.spmm_config(lanes_n, unroll=self.unroll, long_threshold=lt, seg_len=lt,
                                 slot_scale=self.slot_scale)

    def max_blocks(self, device):
        if self.blocks_per_cu <= 0:
            return 0
        return int(self.blocks_per_cu) * compute_units(device)


def compute_units(device):
    return int(torch.cuda.get_device_properties(device).multi_processor_count)


def _log2_bucket(v):
    return int(round(math.log2(v))) if v > 0 else -1


def shape_class(kind, value_dtype, index_dtype, indptr_dtype, n_dense, n_rows, nnz, max_row, cus, extra=""):
    """Key of a shape class (a string, stable across processes); `extra`
    tags classes whose candidate sets differ (e.g. x fits in LDS)."""
    mean = nnz / n_rows if n_rows else 0.0
    parts = [
        kind,
        str(value_dtype).replace("torch.", ""),
        str(index_dtype).replace("torch.", ""),
        str(indptr_dtype).replace("torch.", ""),
        f"n{int(n_dense)}",
        f"mean{_log2_bucket(mean)}",
        f"max{_log2_bucket(max_row)}",
        f"rows{_log2_bucket(n_rows)}",
        f"cu{int(cus)}",
    ]
    if extra:
        parts.append(extra)
    return "|".join(parts)


def batch_spmv_candidates(elem_bytes=4, x_len=0, chunk_ok=False):
    """SpMV candidates of a batch: the single-matrix set plus, with chunk_ok,
    the row-range parts kernel (variant 5: each matrix's column range of x
    in LDS; lanes_n lanes per row, part_nnz nonzeros per part)."""
    out = candidates("spmv", 1, elem_bytes, x_len, chunk_ok)
    if chunk_ok:
        for lanes in (4, 8, 16):
            for part_nnz in (1 << 13, 1 << 15, 1 << 16, 1 << 17):
                out.append(TuneParams(unroll=4, lanes_n=lanes, variant=5, part_nnz=part_nnz))
    return out


def candidates(kind, n_dense=1, elem_bytes=4, x_len=0, chunk_ok=False):
    """Candidate parameter sets of a kernel family (small, fixed grids).
    For SpMV, x_len > 0 with x_len * elem_bytes <= 64 KiB adds the loops with
    x staged in LDS (variants 2 and 3: 1024-thread workgroups, 1 or 2 per CU)
    and, with chunk_ok (4-element aligned arrays), the vector-chunk loop
    (variant 4)."""
    out = []
    if kind == "spmv":
        for variant in (0, 1):
            for unroll in (2, 4, 8):
                for bpc in (0, 4, 8):
                    for xg in (0, 8):
                        out.append(TuneParams(unroll=unroll, blocks_per_cu=bpc, xcd_groups=xg, variant=variant))
        if 0 < x_len * elem_bytes <= _plan.LDSX_MAX_BYTES:
            for variant in (2, 3):
                for unroll in (2, 4, 8):
                    for bpc in (1, 2):
                        for xg in (0, 8):
                            out.append(TuneParams(unroll=unroll, blocks_per_cu=bpc, xcd_groups=xg, variant=variant))
            if chunk_ok:
                for unroll in (4, 8):
                    for bpc in (1, 2):
                        for xg in (0, 8):
                            out.append(TuneParams(unroll=unroll, blocks_per_cu=bpc, xcd_groups=xg, variant=4))
        return out
    g = _plan.lanes_for_dense(n_dense, elem_bytes)
    lanes = [g] + ([g // 2] if g >= 8 else [])
    for ln in lanes:
        for unroll in (4, 8):
            for scale in (1, 2, 4):
                for bpc in (0, 8):
                    for xg in (0, 8):
                        out.append(TuneParams(unroll=unroll, slot_scale=scale, lanes_n=ln, blocks_per_cu=bpc,
                                              xcd_groups=xg))
    return out


# ------------------------------------------------------------------- cache
def cache_path():
    p = os.environ.get("FLAGSPARSE_AMD_TUNE_CACHE")
    if p is not None:
        return p or None  # "" disables the file
    return os.path.join(os.path.expanduser("~"), ".cache", "flagsparse_amd", "tune.json")


def library_hash():
    """sha1 of the built library file (entries of another build are stale)."""
    global _LIB_HASH
    if _LIB_HASH is None:
        lib = _lib.load()
        h = hashlib.sha1()
        with open(lib._name, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        _LIB_HASH = h.hexdigest()[:16]
    return _LIB_HASH


def _version_tag():
    return f"abi{_lib.ABI_VERSION}-{library_hash()}"


def _read_file(path):
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if isinstance(data, dict) else {}


def lookup(key):
    with _CACHE_LOCK:
        hit = _MEMORY.get(key)
        if hit is not None:
            return hit
        path = cache_path()
        if path:
            entry = _read_file(path).get(key)
            if isinstance(entry, dict) and entry.get("version") == _version_tag():
                try:
                    params = TuneParams(**entry["params"])
                except (TypeError, KeyError):
                    return None
                _MEMORY[key] = params
                return params
    return None


def store(key, params, timing_ms=None):
    with _CACHE_LOCK:
        _MEMORY[key] = params
        path = cache_path()
        if not path:
            return
        data = _read_file(path)
        data[key] = {"version": _version_tag(), "params": asdict(params), "ms": timing_ms}
        try:
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            tmp = f"{path}.{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass  # an unwritable cache only costs a re-tune


def clear_memory():
    with _CACHE_LOCK:
        _MEMORY.clear()


# ------------------------------------------------------------------ tuning
def spmv_class_tag(x_len, indices, data):
    """Shape-class tag of the SpMV candidate set: x fits in LDS (the LDS
    loops), and then whether the arrays sit on 4-element boundaries (the
    vector-chunk loop)."""
    if not 0 < x_len * data.element_size() <= _plan.LDSX_MAX_BYTES:
        return ""
    return "xlds-chunk" if _plan.chunk_aligned(indices, data) else "xlds"


def tune_spmv(handle, data, allow_tuning=True):
    """Parameters of y = A x for a CsrHandle (shape class of its rows): the
    cached winner, or tuned now on this matrix (x = 1, y scratch)."""
    device = data.device
    lengths = handle.indptr[1:] - handle.indptr[:-1] if handle.n_rows else handle.indptr[:0]
    max_row = int(lengths.max().item()) if handle.n_rows else 0
    key = shape_class("spmv", data.dtype, handle.indices.dtype, handle.indptr.dtype, 1, handle.n_rows, handle.nnz,
                      max_row, compute_units(device), extra=spmv_class_tag(handle.n_cols, handle.indices, data))
    hit = lookup(key)
    if hit is not None or not allow_tuning:
        return hit or TuneParams()
    x = torch.ones(handle.n_cols, dtype=data.dtype, device=device)
    y = torch.empty(handle.n_rows, dtype=data.dtype, device=device)
    cands = candidates("spmv", 1, data.element_size(), handle.n_cols, _plan.chunk_aligned(handle.indices, data))
    for p in cands:  # plans (host syncs) before any timing
        handle.plan(p.spmv_config())

    def run(p):
        handle.spmv(data, x, out=y, config=p.spmv_config(), max_blocks=p.max_blocks(device),
                    xcd_groups=p.xcd_groups, variant=p.variant)

    best, _ = tune(key, cands, run)
    return best


def tune_spmm(handle, data, n_dense, allow_tuning=True):
    """Parameters of C = A B (B row-major, n_dense columns) for a CsrHandle:
    the cached winner of its shape class, or tuned now on this matrix with a
    B of ones (C scratch)."""
    device = data.device
    lengths = handle.indptr[1:] - handle.indptr[:-1] if handle.n_rows else handle.indptr[:0]
    max_row = int(lengths.max().item()) if handle.n_rows else 0
    key = shape_class("spmm", data.dtype, handle.indices.dtype, handle.indptr.dtype, n_dense, handle.n_rows,
                      handle.nnz, max_row, compute_units(device))
    hit = lookup(key)
    if hit is not None or not allow_tuning or handle.n_rows == 0 or n_dense == 0:
        return hit or TuneParams()
    elem = data.element_size()
    B = torch.ones(handle.n_cols, n_dense, dtype=data.dtype, device=device)
    C = torch.empty(handle.n_rows, n_dense, dtype=data.dtype, device=device)
    cands = candidates("spmm", n_dense, elem)
    default_lanes = _plan.lanes_for_dense(n_dense, elem)
    for p in cands:
        handle.plan(p.spmm_config(p.lanes_n or default_lanes))

    def run(p):
        ln = p.lanes_n or default_lanes
        handle.spmm(data, B, out=C, lanes_n=ln, config=p.spmm_config(ln), max_blocks=p.max_blocks(device),
                    xcd_groups=p.xcd_groups)

    best, _ = tune(key, cands, run)
    return best



def time_launch(fn, reps=5):
    """Median device time (ms) of fn() on the current stream after a warm-up."""
    fn()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune(key, cands, run, reps=5):
    """Time run(params) for every candidate; cache and return the fastest.
    `run` must launch only (no host syncs) so the events time the kernels."""
    hit = lookup(key)
    if hit is not None:
        return hit, None
    best, best_ms = None, float("inf")
    results = []
    for p in cands:
        ms = time_launch(lambda: run(p), reps)
        results.append((p, ms))
        if ms < best_ms:
            best, best_ms = p, ms
    store(key, best, best_ms)
    return best, results
