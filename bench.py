"""Benchmark of the CSR SpMM hot path on one MI355X.

Headline workload (one step = one SpMM over the whole matrix, inputs resident
in HBM): C = A B, A 262144 x 262144 CSR with power-law row lengths (exponent
2.1, mean 24, cap 4096, distinct uniform columns, values U[-1,1]), B dense
262144 x 64 N(0,1), float32 values, int32 indices, seed 1.

value = algorithmic bytes / time per step (GB/s), where algorithmic bytes =
nnz*(4+4) + (M+1)*4 + K*N*4 + M*N*4 (A streamed once, B read once, C written
once).  The roofline object prices the dominant kernel (spmm_bins_kernel)
against the 8 TB/s HBM peak with its duration from HIP events on the stream
the kernels run on.  The same line also carries the bf16-vs-fp32 fwd+bwd
speedup of the differentiable SpMM (200000-node power-law graph, N=128) and
the 16384 x 16384 x 16 SpMV rate, and a CPU baseline: the C oracle (one core)
on a bounded slice of the same workload, plus torch.sparse on all host cores.
Extras: the measured HBM copy rate, an HBM-bound 2^24-row SpMV, the BSR SpMM
on the matrix cores (bf16 / f16 / fp8) and the 4096-matrix batched launch.
"""

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import flagsparse_amd as fs  # noqa: E402
from flagsparse_amd import _plan, synthetic  # noqa: E402
from flagsparse_amd._matrix import CsrHandle  # noqa: E402

METRIC = "SpMV/SpMM GB/s (% HBM peak); bf16 fwd+bwd SpMM speedup vs FP32 at relerr<=1e-2"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmm_bytes(nnz, m, k, n, s_v=4, s_i=4, s_p=4, s_c=4):
    return nnz * (s_v + s_i) + (m + 1) * s_p + k * n * s_v + m * n * s_c


def time_steps(fn, steps, warmup, stream):
    """Warm up, then time `steps` calls between device syncs (wall clock,
    nothing else in the stream), then `steps` more calls each bracketed by
    HIP events on `stream` (the stream the kernels are launched on).
    Returns (wall ms of the uninstrumented steps, per-step event ms): the
    event records cost the stream ~10 us per step, so they stay out of the
    wall-clock window."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000.0
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for s, e in events:
        s.record(stream)
        fn()
        e.record(stream)
    torch.cuda.synchronize()
    ev = [s.elapsed_time(e) for s, e in events]
    return wall_ms, ev


def bench_spmm_power_law(dev, steps, warmup):
    """The headline step through the public prepared entry points: one
    flagsparse_spmm_csr_opt(B=..., prepared=..., out=C) per step (non), and
    one flagsparse_spmm_csr_run on the route prepared for op trans (the CSC
    permutation is built by the first call, outside the timed steps)."""
    log("generating 262144^2 power-law CSR (seed 1)")
    ip, ix, v, B = synthetic.power_law_rows(seed=1)
    n = 262144
    nnz = int(ix.size)
    data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, torch.float32, torch.int32)
    Bt = synthetic.dense(B, dev, torch.float32)
    prepared = fs.prepare_spmm_csr_opt(data, idx, ptr, (n, n))
    C = torch.empty((n, 64), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        fs.flagsparse_spmm_csr_opt(B=Bt, prepared=prepared, out=C)

    step()  # builds the row-bin plan (prepare-time work, outside the timed steps)
    plan = prepared._handle.plan(_plan.spmm_config(_plan.lanes_for_dense(64, 4)))
    log(f"nnz={nnz} bins={plan.bin_counts} long_rows={plan.n_long_rows} segments={plan.n_segments}")
    wall_ms, ev = time_steps(step, steps, warmup, stream)
    ms = wall_ms / steps
    nbytes = spmm_bytes(nnz, n, n, 64)
    route = fs.prepare_spmm_csr_route(data, idx, ptr, (n, n), op="trans")

    def step_t():
        fs.flagsparse_spmm_csr_run(route, Bt)

    step_t()  # builds the CSC permutation and its plan once
    wall_t, ev_t = time_steps(step_t, steps, warmup, stream)
    ms_t = wall_t / steps
    nbytes_t = nbytes + nnz * 4  # the permutation is read as well
    return {
        "nnz": nnz,
        "ms": ms,
        "event_ms_median": statistics.median(ev),
        "bytes": nbytes,
        "flops": 2.0 * nnz * 64,
        "trans_ms": ms_t,
        "trans_event_ms_median": statistics.median(ev_t),
        "trans_bytes": nbytes_t,
        "plan": {"bins": plan.bin_counts, "long_rows": plan.n_long_rows, "segments": plan.n_segments},
        "inputs": (ip, ix, v, B),
    }


def bench_gather_probe(dev, n_rows, row_bytes, n_gathers, steps, warmup):
    """Random whole-row reads from a table of B's footprint (fsp_gather_rows_probe
    of the measurement library), one per nonzero: the practical ceiling of the
    SpMM's B-row gathers, timed like the SpMM (HIP events on the launch
    stream).  Returns (ms, GB/s)."""
    from flagsparse_amd import _bench_lib  # noqa: PLC0415

    table = torch.empty(n_rows * row_bytes // 4, dtype=torch.float32, device=dev).normal_()
    sink = torch.zeros(1, dtype=torch.float32, device=dev)
    lib = _bench_lib.load()
    stream = torch.cuda.current_stream(dev)
    table_bytes = table.numel() * 4

    def run():
        _bench_lib.check(lib.fsp_gather_rows_probe(table.data_ptr(), table_bytes, n_rows, row_bytes, n_gathers,
                                                   sink.data_ptr(), stream.cuda_stream), "fsp_gather_rows_probe")

    _, ev = time_steps(run, steps, warmup, stream)
    ms = statistics.median(ev)
    return ms, n_gathers * row_bytes / (ms * 1e-3) / 1e9


def bench_slab_schedule(dev, inputs, steps, configs=((1, 8, -1), (4, 8, 1 << 20), (16, 8, 1 << 20), (16, 8, -1))):
    """The headline SpMM under the column-slab schedule (fsp_slab_spmm_probe of
    the measurement library): B swept in S slabs, the workgroups of one XCD
    passing a phase counter after each slab (spin >= 0) or not (spin < 0).
    Rows above 256 nonzeros run as 256-nonzero segments; rows are dealt to
    lane groups longest first.  Reports kernel time (HIP events, median),
    the error against flagsparse_spmm_csr and whether a bounded wait ran
    out.  Not part of the default line (a measurement of a rejected design)."""
    from flagsparse_amd import _bench_lib  # noqa: PLC0415

    ip, ix, v, Bn = inputs
    n = ip.size - 1
    data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, torch.float32, torch.int32)
    B = synthetic.dense(Bn, dev, torch.float32)
    ref = fs.flagsparse_spmm_csr(data, idx, ptr, B, (n, n))
    lens = np.diff(ip)
    seg = 256
    short = lens <= seg
    beg, end, slot, seg_row = [ip[:-1][short]], [ip[1:][short]], [np.nonzero(short)[0]], []
    n_seg = 0
    for r in np.nonzero(~short)[0]:
        k = -(-int(lens[r]) // seg)
        j = np.arange(k)
        beg.append(ip[r] + lens[r] * j // k)
        end.append(ip[r] + lens[r] * (j + 1) // k)
        slot.append(n + n_seg + j)
        seg_row.append(np.full(k, r))
        n_seg += k
    vb, ve, vs = (np.concatenate(a).astype(np.int64) for a in (beg, end, slot))
    order = np.argsort(-(ve - vb), kind="stable")  # longest first, dealt round-robin to lane groups
    vb, ve, vs = vb[order], ve[order], vs[order]
    nv = vb.size
    vlen = ve - vb
    vid = np.repeat(np.arange(nv), vlen)
    pos = np.repeat(vb - (np.cumsum(vlen) - vlen), vlen) + np.arange(vid.size)
    cols = ix[pos].astype(np.int64)
    vbeg_t = torch.from_numpy(vb.astype(np.int32)).to(dev)
    vslot_t = torch.from_numpy(vs.astype(np.int32)).to(dev)
    seg_row_t = torch.from_numpy(np.concatenate(seg_row)).to(dev) if seg_row else None
    C = torch.zeros((n, 64), device=dev)  # rows cut into segments are never written: their sum is added below
    P = torch.empty((max(1, n_seg), 64), device=dev)
    counters = torch.zeros(8, dtype=torch.int32, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    wgs = torch.cuda.get_device_properties(dev).multi_processor_count
    stream = torch.cuda.current_stream(dev)
    lib = _bench_lib.load()
    out = {"virtual_rows": int(nv), "segments": int(n_seg), "workgroups": wgs}
    for n_slabs, rpg, spin in configs:
        cnt = np.bincount(vid * n_slabs + cols * n_slabs // n, minlength=nv * n_slabs).reshape(nv, n_slabs)
        slab_end = torch.from_numpy(np.cumsum(cnt, axis=1).astype(np.uint16).view(np.int16)).to(dev)

        def run():
            st = lib.fsp_slab_spmm_probe(rpg, vbeg_t.data_ptr(), slab_end.data_ptr(), vslot_t.data_ptr(),
                                         idx.data_ptr(), data.data_ptr(), B.data_ptr(), C.data_ptr(), P.data_ptr(), n,
                                         nv, n_slabs, wgs, counters.data_ptr(), spin, flag.data_ptr(),
                                         stream.cuda_stream)
            if st != 0:
                raise RuntimeError(f"fsp_slab_spmm_probe failed with status {st}")

        flag.zero_()
        _, ev = time_steps(run, max(3, steps // 2), 1, stream)
        got = C.clone()
        if seg_row_t is not None:
            got.index_add_(0, seg_row_t, P[:n_seg])
        err = float((got - ref).abs().max() / ref.abs().max())
        out[f"slabs{n_slabs}_rpg{rpg}_{'sync' if spin >= 0 else 'nosync'}"] = {
            "ms": round(statistics.median(ev), 4), "rel_err_vs_spmm": err, "wait_ran_out": bool(flag.item())}
    return out


def graph_time_ms(fn, dev, per_graph=64, reps=15):
    """Median device time per call of `fn` replayed from a captured graph of
    `per_graph` calls: the kernels back to back, without host launch gaps."""
    s = torch.cuda.Stream(dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        fn()
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(per_graph):
            fn()
    times = []
    with torch.cuda.stream(s):
        g.replay()
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(s)
            g.replay()
            b.record(s)
            b.synchronize()
            times.append(a.elapsed_time(b) / per_graph)
    s.synchronize()
    return statistics.median(times)


def bench_spmv_uniform(dev, steps, warmup):
    """16384^2 x 16 SpMV (2.3 MB, cache-resident after the first call): host
    wall clock per call, and device time per call from graph replay (the
    launch-latency-free figure; bytes / that = cache-resident GB/s)."""
    from flagsparse_amd import autotune  # noqa: PLC0415

    ip, ix, v, x = synthetic.uniform_rows(16384, 16384, 16, seed=0)
    data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, torch.float32, torch.int32)
    xt = synthetic.dense(x, dev, torch.float32)
    handle = CsrHandle(idx, ptr, (16384, 16384))
    y = torch.empty(16384, device=dev)
    handle.spmv(data, xt, out=y)
    stream = torch.cuda.current_stream(dev)
    wall_ms, ev = time_steps(lambda: handle.spmv(data, xt, out=y), steps * 5, warmup, stream)
    nbytes = ix.size * 8 + 16385 * 4 + 16384 * 4 * 2
    graph_ms = graph_time_ms(lambda: handle.spmv(data, xt, out=y), dev)
    best = None
    for p in autotune.candidates("spmv", 1, 4, 16384):  # graph-timed: launch gaps would hide the kernels
        ms = graph_time_ms(lambda: handle.spmv(data, xt, out=y, config=p.spmv_config(), max_blocks=p.max_blocks(dev),
                                               xcd_groups=p.xcd_groups, variant=p.variant), dev, 32, 5)
        if best is None or ms < best[0]:
            best = (ms, p)
    out = {"ms": wall_ms / (steps * 5), "event_ms_median": statistics.median(ev), "bytes": nbytes,
           "graph_ms_per_call": graph_ms, "gbs_cache_resident": nbytes / (graph_ms * 1e-3) / 1e9,
           "best_graph_ms_per_call": best[0], "best_params": best[1].__dict__,
           "best_gbs_cache_resident": nbytes / (best[0] * 1e-3) / 1e9}
    out["replicas_256"] = bench_spmv_replicas(dev, data, idx, ptr, xt, nbytes)
    return out


def bench_spmv_replicas(dev, data, idx, ptr, xt, nbytes, n_rep=256):
    """The 16384^2 x 16 SpMV over 256 independent in-memory replicas (587 MB,
    beyond the 256 MB Infinity Cache), all products in one batched launch
    (graph replay): the HBM-resident rate per product."""
    from flagsparse_amd.batch import CsrBatch  # noqa: PLC0415

    b = CsrBatch.from_matrices([(data, idx, ptr, (16384, 16384))] * n_rep)
    x = xt.repeat(n_rep)
    y = torch.empty(b.total_rows, dtype=torch.float32, device=dev)
    p = b.params("spmv")
    g = b.graph_spmv(x, y)
    stream = torch.cuda.current_stream(dev)
    _, ev = time_steps(g.replay, 20, 3, stream)
    med = statistics.median(ev)
    ok = torch.equal(y[:16384], b.spmv(x, params=p)[:16384])
    del b, x, y, g
    return {"replicas": n_rep, "bytes": nbytes * n_rep, "graph_event_ms_median": round(med, 5),
            "us_per_product": round(med * 1e3 / n_rep, 4), "gbs": round(nbytes * n_rep / (med * 1e-3) / 1e9, 2),
            "tuned": p.__dict__, "replay_equals_eager": bool(ok)}


def bench_spmm_tuned(dev, steps, warmup, inputs):
    """The headline SpMM with the autotuner's parameters for its shape class."""
    from flagsparse_amd import autotune  # noqa: PLC0415

    ip, ix, v, B = inputs
    n = 262144
    data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, torch.float32, torch.int32)
    Bt = synthetic.dense(B, dev, torch.float32)
    handle = CsrHandle(idx, ptr, (n, n))
    p = autotune.tune_spmm(handle, data, 64)
    ln = p.lanes_n or _plan.lanes_for_dense(64, 4)
    cfg = p.spmm_config(ln)
    C = torch.empty((n, 64), dtype=torch.float32, device=dev)

    def step():
        handle.spmm(data, Bt, out=C, lanes_n=ln, config=cfg, max_blocks=p.max_blocks(dev), xcd_groups=p.xcd_groups)

    _, ev = time_steps(step, steps, warmup, torch.cuda.current_stream(dev))
    med = statistics.median(ev)
    return {"params": p.__dict__, "event_ms_median": round(med, 5),
            "gbs": round(spmm_bytes(int(ix.size), n, n, 64) / (med * 1e-3) / 1e9, 2)}


def bench_autograd_bf16_vs_fp32(dev, steps, warmup):
    log("generating 200000-node power-law graph (seed 3)")
    ip, ix, v = synthetic.power_law_graph(200000, 2.3, 14, seed=3)
    n = 200000
    rng = np.random.Generator(np.random.PCG64(3))
    Bn = rng.standard_normal((n, 128))
    dCn = rng.standard_normal((n, 128))
    out = {"nnz": int(ix.size)}
    results = {}
    stream = torch.cuda.current_stream(dev)
    for dtype, name in ((torch.float32, "fp32"), (torch.bfloat16, "bf16")):
        data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, dtype, torch.int32)
        mat = fs.CsrMatrix(idx, ptr, (n, n)).prepare_backward()
        vals = data.clone().requires_grad_(True)
        B = synthetic.dense(Bn, dev, dtype).requires_grad_(True)
        dC = synthetic.dense(dCn, dev, dtype)

        def step():
            vals.grad = None
            B.grad = None
            C = fs.spmm_csr_autograd(vals, B, mat)
            C.backward(dC)

        wall_ms, ev = time_steps(step, steps, warmup, stream)
        out[name + "_ms"] = wall_ms / steps
        out[name + "_event_ms_median"] = statistics.median(ev)
        step()
        results[name] = (fs.spmm_csr_autograd(vals, B, mat).detach().double(), B.grad.double(), vals.grad.double())
    # bf16 error relative to the fp32 run (max|d| / max|ref| per output)
    for i, what in enumerate(("fwd", "dB", "dvalues")):
        ref = results["fp32"][i]
        out[f"bf16_rel_err_{what}_vs_fp32"] = float((results["bf16"][i] - ref).abs().max() / ref.abs().max())
    # per-step device time (median of the HIP-event windows, each holding the
    # forward and both backward launches of one step): the host-side Python
    # overhead of the autograd call, which the wall-clock mean also carries,
    # is not part of the kernels being compared
    out["speedup_bf16_vs_fp32"] = out["fp32_event_ms_median"] / out["bf16_event_ms_median"]
    out["speedup_bf16_vs_fp32_wall"] = out["fp32_ms"] / out["bf16_ms"]
    return out


def bench_trans_forms(dev, steps, warmup, inputs):
    """The two forms of the transposed SpMM A^T B against the forward A B:
    values gathered into CSC order by fsp_gather and then the plain kernel
    over the CSC (the library's form), and the kernel that loads
    data[perm[q]] itself.  The 262144^2 power-law matrix at N = 64 and the
    200000-node graph at N = 128, f32 and bf16; the forms are checked bitwise
    equal."""
    from flagsparse_amd._csc import gather_values  # noqa: PLC0415

    stream = torch.cuda.current_stream(dev)
    out = {}
    ip, ix, v, _ = inputs
    cases = [("power_law_262144_n64", ip, ix, v, 262144, 64)]
    gip, gix, gv = synthetic.power_law_graph(200000, 2.3, 14, seed=3)
    cases.append(("graph_200000_n128", gip, gix, gv, 200000, 128))
    for name, ip_, ix_, v_, n, width in cases:
        for dtype, dname in ((torch.float32, "f32"), (torch.bfloat16, "bf16")):
            data, idx, ptr = synthetic.to_torch(ip_, ix_, v_, dev, dtype, torch.int32)
            h = CsrHandle(idx, ptr, (n, n))
            B = torch.randn(n, width, device=dev).to(dtype)
            C_g = torch.empty((n, width), dtype=dtype, device=dev)
            C_p = torch.empty((n, width), dtype=dtype, device=dev)
            perm = h.csc()[2]
            res = {}
            for key, fn in (("fwd", lambda: h.spmm(data, B, out=C_g)),
                            ("trans_values_via_perm", lambda: h.spmm(data, B, transpose=True, out=C_p,
                                                                     values_via_perm=True)),
                            ("gather_only", lambda: gather_values(data, perm)),
                            ("trans_gathered", lambda: h.spmm(data, B, transpose=True, out=C_g))):
                _, ev = time_steps(fn, steps, warmup, stream)
                res[key + "_ms"] = round(statistics.median(ev), 5)
            torch.cuda.synchronize()
            res["bitwise_equal"] = bool(torch.equal(C_g, C_p))
            res["gain"] = round(res["trans_values_via_perm_ms"] / res["trans_gathered_ms"], 3)
            out[f"{name}_{dname}"] = res
            log(f"trans forms {name} {dname}: {res}")
    return out


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "spmm_bins_kernel_power_law_262144_n64_pmc.md")
SPMV_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "spmv_vec_kernel_uniform_16m_x16_pmc.md")


def measured_traffic(path=TRAFFIC_SUMMARY):
    """HBM bytes per launch of a kernel (default: the headline SpMM) from a
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by the copy kernel's calibration factor); None when absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1].strip()))
    except (OSError, ValueError):
        return None
    return None


def bench_copy(dev, steps):
    """Measured HBM copy rate (1 GiB device-to-device, read + write bytes);
    its kernel also calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    n = 1 << 28  # 2^28 float32 = 1 GiB
    src = torch.empty(n, dtype=torch.float32, device=dev).uniform_()
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream(dev)
    _, ev = time_steps(lambda: dst.copy_(src), max(5, steps // 5), 2, stream)
    med = statistics.median(ev)
    del src, dst
    return {"bytes": 2 * 4 * n, "event_ms_median": round(med, 5), "gbs": round(2 * 4 * n / (med * 1e-3) / 1e9, 2)}


def bench_spmv_large(dev, steps, warmup):
    """HBM-bound SpMV: 2^24 rows x 16384 columns, 16 distinct uniform columns
    per row -- the 16384 x 16384 x 16 pattern with 1024x the rows, so the
    index / value stream (2.2 GB) is far beyond the caches while x (64 KB)
    stays cache-resident, as it does for the small matrix."""
    from flagsparse_amd import autotune  # noqa: PLC0415

    m, n = 1 << 24, 16384
    ptr, idx, data, x = synthetic.uniform_rows_device(m, n, 16, seed=0, device=dev)
    handle = CsrHandle(idx, ptr, (m, n))
    y = torch.empty(m, dtype=torch.float32, device=dev)
    p = autotune.tune_spmv(handle, data)  # per shape class, cached

    def step():
        handle.spmv(data, x, out=y, config=p.spmv_config(), max_blocks=p.max_blocks(dev), xcd_groups=p.xcd_groups,
                    variant=p.variant)

    step()
    stream = torch.cuda.current_stream(dev)
    _, ev = time_steps(step, steps, warmup, stream)
    med = statistics.median(ev)
    nnz = m * 16
    nbytes = nnz * 8 + (m + 1) * 4 + n * 4 + m * 4
    gbs = nbytes / (med * 1e-3) / 1e9
    probe = bench_read_probe(dev, idx, data, steps, warmup)
    probe["spmv_frac_of_best"] = round(gbs / probe["best_gbs"], 4)
    # rows of one length: the loop computes row bounds and reads no row pointer,
    # so the bytes it moves are the CSR bytes less indptr (PMC-measured below)
    computed = handle.plan(p.spmv_config()).row_len > 0
    traffic = measured_traffic(SPMV_TRAFFIC_SUMMARY) if p.variant == 4 and computed else None
    del ptr, idx, data, x, y, handle
    return {"rows": m, "cols": n, "nnz": nnz, "bytes": nbytes, "event_ms_median": round(med, 5),
            "tuned": {k: v for k, v in p.__dict__.items()}, "row_bounds": "computed" if computed else "loaded",
            "traffic": traffic, "moved_gbs": round(traffic / (med * 1e-3) / 1e9, 2) if traffic else None,
            "gbs": round(gbs, 2), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "read_probe": probe}


def bench_read_probe(dev, a, b, steps, warmup):
    """Streaming reads of the SpMV's column-id and value arrays themselves
    (fsp_stream_read_probe of the measurement library: 16-byte nontemporal
    chunks, both arrays in lock step, 1024-thread workgroups): the practical
    read rate of the index / value stream pattern, with the SpMV's
    persistent grid (two workgroups per CU, contiguous ranges) and with a
    grid-stride one; also one array alone."""
    from flagsparse_amd import _bench_lib  # noqa: PLC0415

    lib = _bench_lib.load()
    sink = torch.zeros(1, device=dev)
    stream = torch.cuda.current_stream(dev)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    chunks = a.numel() * a.element_size() // 16
    out = {"arrays_bytes": 2 * chunks * 16}
    best = 0.0
    for name, n_arrays, wgs, contiguous in (("two_arrays_2wg_per_cu_contiguous", 2, 2 * cus, 1),
                                            ("two_arrays_2wg_per_cu_strided", 2, 2 * cus, 0),
                                            ("two_arrays_8wg_per_cu_strided", 2, 8 * cus, 0),
                                            ("one_array_8wg_per_cu_strided", 1, 8 * cus, 0)):
        def run():
            st = lib.fsp_stream_read_probe(a.data_ptr(), b.data_ptr(), chunks, n_arrays, wgs, contiguous,
                                           sink.data_ptr(), stream.cuda_stream)
            if st != 0:
                raise RuntimeError(f"fsp_stream_read_probe failed with status {st}")

        _, ev = time_steps(run, steps, warmup, stream)
        ms = statistics.median(ev)
        g = n_arrays * chunks * 16 / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 5), "gbs": round(g, 1)}
        if n_arrays == 2:
            best = max(best, g)
    out["best_gbs"] = round(best, 1)
    return out


def bench_vendor(dev, steps, warmup, inputs):
    """rocSPARSE (torch.sparse.mm on a ROCm build of torch) on the same two
    workloads: the 262144^2 power-law SpMM (n = 64, f32/i32) and the 2^24-row
    x16 SpMV.  Device-event medians on the current stream, output allocation
    included as torch.sparse.mm does it per call."""
    ip, ix, v, B = inputs
    n = 262144
    data, idx, ptr = synthetic.to_torch(ip, ix, v, dev, torch.float32, torch.int32)
    Bt = synthetic.dense(B, dev, torch.float32)
    A = torch.sparse_csr_tensor(ptr, idx, data, size=(n, n))
    stream = torch.cuda.current_stream(dev)
    out = {}
    try:
        _, ev = time_steps(lambda: torch.sparse.mm(A, Bt), steps, warmup, stream)
        med = statistics.median(ev)
        nbytes = spmm_bytes(int(ix.size), n, n, 64)
        out["spmm_power_law_262144_n64"] = {"event_ms_median": round(med, 5),
                                            "gbs": round(nbytes / (med * 1e-3) / 1e9, 2)}
    except RuntimeError as exc:
        out["spmm_power_law_262144_n64"] = {"error": str(exc)[:200]}
    del A, data, idx, ptr, Bt
    m, k = 1 << 24, 16384
    ptr, idx, data, x = synthetic.uniform_rows_device(m, k, 16, seed=0, device=dev)
    A = torch.sparse_csr_tensor(ptr, idx, data, size=(m, k))
    xv = x.unsqueeze(1)
    try:
        _, ev = time_steps(lambda: torch.sparse.mm(A, xv), steps, warmup, stream)
        med = statistics.median(ev)
        nbytes = m * 16 * 8 + (m + 1) * 4 + k * 4 + m * 4
        out["spmv_uniform_16m_rows_x16"] = {"event_ms_median": round(med, 5),
                                            "gbs": round(nbytes / (med * 1e-3) / 1e9, 2)}
    except RuntimeError as exc:
        out["spmv_uniform_16m_rows_x16"] = {"error": str(exc)[:200]}
    del A, ptr, idx, data, x, xv
    out["library"] = "rocSPARSE via torch.sparse.mm (torch " + torch.__version__ + ")"
    return out


def bench_bsr(dev, steps, warmup):
    """BSR 65536 x 65536 (16 x 16 blocks, 40 per block row, seed 2) times
    B 65536 x 128 on the matrix cores: bf16, f16 and fp8 e4m3 (per-block
    absmax scales for A, one absmax scale for B).  event_ms_median is the
    kernel alone (B packed and A quantised beforehand); per_call_ms is the
    public flagsparse_spmm_bsr on the float32 B, which packs (fp8: takes the
    absmax and quantises) B on every call."""
    from flagsparse_amd import _lib, bsr  # noqa: PLC0415

    log("generating BSR 65536^2, 40 blocks per block row (seed 2)")
    bp, bc, bd = synthetic.bsr_blocks()
    B = np.random.default_rng(2).standard_normal((65536, 128))
    nnzb = int(bc.size)
    indptr = torch.from_numpy(bp).to(dev, torch.int32)
    indices = torch.from_numpy(bc).to(dev, torch.int32)
    blocks32 = torch.from_numpy(bd).to(dev, torch.float32)
    Bt = synthetic.dense(B, dev, torch.float32)
    C = torch.empty((65536, 128), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    flops = 2.0 * nnzb * 256 * 128
    peaks = {"bf16": 2500.0, "f16": 2500.0, "fp8": 5000.0}  # dense MFMA TFLOP/s
    out = {"nnzb": nnzb, "flops": flops}
    lib = _lib.load()
    for name in ("bf16", "f16", "fp8"):
        if name == "fp8":
            data, scales = bsr.quantize_bsr_fp8(blocks32)
            Bp, absmax = bsr.pack_dense_operand(Bt, torch.uint8)
            vcode, elem = _lib.FSP_F8E4M3, 1
        else:
            dt = torch.bfloat16 if name == "bf16" else torch.float16
            data, scales = blocks32.to(dt), None
            Bp, absmax = bsr.pack_dense_operand(Bt, dt)
            vcode, elem = _lib.value_code(dt), 2

        def step():
            _lib.check(lib.fsp_bsr_spmm(vcode, _lib.FSP_I32, indptr.data_ptr(), indices.data_ptr(), data.data_ptr(),
                                        None if scales is None else scales.data_ptr(), 4096, nnzb, Bp.data_ptr(),
                                        None if absmax is None else absmax.data_ptr(), 65536, 128, C.data_ptr(), 128,
                                        65536, stream.cuda_stream), "fsp_bsr_spmm")

        _, ev = time_steps(step, steps, warmup, stream)
        med = statistics.median(ev)
        # algorithmic bytes: blocks (+ scales), block columns and row pointer,
        # B once in the kernel's storage type, C (float32) once
        nbytes = nnzb * 256 * elem + nnzb * 4 + 4097 * 4 + 65536 * 128 * elem + 65536 * 128 * 4
        if name == "fp8":
            nbytes += nnzb * 4
        tflops = flops / (med * 1e-3) / 1e12

        def public_call():  # what a caller runs: B (f32) packed / quantised, then the kernel
            fs.flagsparse_spmm_bsr(data, indices, indptr, Bt, (65536, 65536), scales=scales, out=C)

        _, ev_call = time_steps(public_call, steps, warmup, stream)
        out[name] = {"event_ms_median": round(med, 5), "tflops": round(tflops, 2),
                     "mfma_util": round(tflops / peaks[name], 4), "bytes": nbytes,
                     "gbs": round(nbytes / (med * 1e-3) / 1e9, 2),
                     "hbm_frac": round(nbytes / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "per_call_ms": round(statistics.median(ev_call), 5)}
        del data, Bp
    return out


def bench_batch(dev, steps, warmup):
    """4096 square CSR matrices (rows log-uniform in [64, 8192], 4..64
    nonzeros per row, seed 4) in one launch, HIP-graph replayed: SpMV and
    SpMM (k = 32) with autotuned parameters, against the per-matrix launch
    loop (4096 launches) of the same kernels."""
    from flagsparse_amd.batch import CsrBatch  # noqa: PLC0415

    log("generating the 4096-matrix batch (seed 4)")
    ip, ix, v, ro, no = synthetic.csr_batch(seed=4, device=dev)
    b = CsrBatch(v, ix, ip, ro, None, no)
    del ip, ix, ro, no
    torch.cuda.empty_cache()
    rows, nnz = b.total_rows, b.nnz
    x = torch.rand(b.total_cols, device=dev) * 2 - 1
    y = torch.empty(rows, device=dev)
    Bm = torch.randn(b.total_cols, 32, device=dev)
    Cm = torch.empty(rows, 32, device=dev)
    stream = torch.cuda.current_stream(dev)
    out = {"matrices": b.n_mats, "rows": rows, "nnz": nnz}
    log("tuning the batch kernels")
    pv = b.params("spmv")
    pm = b.params("spmm", 32)
    out["tuned"] = {"spmv": pv.__dict__, "spmm": pm.__dict__}
    gv = b.graph_spmv(x, y)
    gm = b.graph_spmm(Bm, Cm)
    bytes_v = nnz * 8 + (rows + 1) * 4 + b.total_cols * 4 + rows * 4
    bytes_m = nnz * 8 + (rows + 1) * 4 + b.total_cols * 32 * 4 + rows * 32 * 4
    for name, g, nbytes, loop in (("spmv", gv, bytes_v, lambda: b.spmv_per_matrix(x, out=y, params=pv)),
                                  ("spmm_k32", gm, bytes_m, lambda: b.spmm_per_matrix(Bm, out=Cm, params=pm))):
        _, ev = time_steps(g.replay, steps, warmup, stream)
        med = statistics.median(ev)
        loop()  # per-matrix plans are built before the loop is timed
        wall_loop, ev_loop = time_steps(loop, max(3, steps // 10), 1, stream)
        loop_ms = wall_loop / max(3, steps // 10)
        out[name] = {"graph_event_ms_median": round(med, 5), "bytes": nbytes,
                     "gbs": round(nbytes / (med * 1e-3) / 1e9, 2),
                     "hbm_frac": round(nbytes / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "per_matrix_loop_ms": round(loop_ms, 4),
                     "speedup_vs_loop": round(loop_ms / med, 2)}
    del b, x, y, Bm, Cm, gv, gm
    torch.cuda.empty_cache()
    return out


def time_cpu(fn, warmup=3, min_reps=10, min_s=1.0, budget_s=30.0):
    """Median wall time (s) of `fn` on the host: `warmup` untimed calls, then
    at least `min_reps` timed ones, more while under `min_s` (short calls),
    stopping once `budget_s` is spent (benchmark/performance_utils.py:81-93:
    warm-ups, then the median; matrix construction outside)."""
    for _ in range(warmup):
        fn()
    reps = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        fn()
        reps.append(time.perf_counter() - t0)
        spent = time.perf_counter() - t_start
        if len(reps) >= min_reps and (spent >= min_s or len(reps) >= 10000):
            break
        if spent >= budget_s and len(reps) >= 3:
            break
    return statistics.median(reps), len(reps)


def _cpu_info():
    return {"os_cpu_count": os.cpu_count(), "torch_threads": torch.get_num_threads(), "cpu_model": _cpu_model(),
            "torch_parallel_info": torch.__config__.parallel_info(), "torch_config": torch.__config__.show()}


def cpu_baseline(inputs):
    """The reference's CPU path on the box's host cores: torch.sparse.mm on
    CPU tensors (its oracle call, ops/benchmarks.py:67-72 and
    ops/spmm_csr.py:4520-4560) over the whole headline workload, and the
    16384 x 16384 x 16 SpMV as torch.sparse.mm(A,
    x.unsqueeze(1)); plus the C oracle (one core, long double) on the
    headline workload as a side figure."""
    from oracle import oracle  # noqa: PLC0415

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    ip, ix, v, B = inputs
    n = ip.size - 1
    A = torch.sparse_csr_tensor(torch.from_numpy(ip), torch.from_numpy(ix), torch.from_numpy(v.astype(np.float32)),
                                size=(n, n))
    Bc = torch.from_numpy(B.astype(np.float32))
    med, reps = time_cpu(lambda: torch.sparse.mm(A, Bc))
    nbytes = spmm_bytes(int(ix.size), n, n, 64)
    info = _cpu_info()
    res = {"torch_sparse": {
        "value": round(nbytes / med / 1e9, 3), "unit": "GB/s", "cores": torch.get_num_threads(), "kind": "reference",
        "sample": f"torch.sparse.mm(A, B) on CPU tensors (the reference's CPU oracle path) over the whole headline "
                  f"workload: 262144^2 power-law CSR ({ix.size} nnz) x 262144 x 64 f32, int32 indices, seed 1; "
                  f"median of {reps} reps after 3 warm-ups, construction excluded",
        "ms": round(med * 1e3, 4), "reps": reps, **info}}
    del A, Bc
    # BASELINE config 1 (the reference's own CPU-runnable case)
    ip1, ix1, v1, x1 = synthetic.uniform_rows(16384, 16384, 16, seed=0)
    A1 = torch.sparse_csr_tensor(torch.from_numpy(ip1), torch.from_numpy(ix1), torch.from_numpy(v1.astype(np.float32)),
                                 size=(16384, 16384))
    xc = torch.from_numpy(x1.astype(np.float32)).unsqueeze(1)
    med1, reps1 = time_cpu(lambda: torch.sparse.mm(A1, xc))
    nbytes1 = ix1.size * 8 + 16385 * 4 + 16384 * 4 * 2
    res["c1_spmv"] = {"value": round(nbytes1 / med1 / 1e9, 3), "unit": "GB/s", "cores": torch.get_num_threads(),
                      "kind": "reference", "ms": round(med1 * 1e3, 5), "reps": reps1,
                      "sample": "torch.sparse.mm(A, x.unsqueeze(1)) on CPU tensors, 16384^2 CSR with 16 distinct "
                                "uniform columns per row, f32 / int32, seed 0 (whole workload)"}
    # the C oracle (one core, long double) on the headline workload
    t0 = time.perf_counter()
    oracle.spmm(ip, ix, v, B, (n, n))
    dt = time.perf_counter() - t0
    res["oracle"] = {"value": round(nbytes / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
                     "sample": f"C oracle (long double) SpMM over the whole headline workload, one run, {dt:.2f} s"}
    return res


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline SpMM only (profiling runs)")
    ap.add_argument("--with-copy", action="store_true",
                    help="also run the 1 GiB copy (the counter calibration kernel) under --no-extras")
    ap.add_argument("--only", default="",
                    help="comma-separated extras to run (copy, spmv_large, spmv_small, spmm_tuned, autograd, bsr, batch, "
                         "vendor, slab_schedule, trans_forms); default: all but slab_schedule and trans_forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the hot path is a single-GPU kernel); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    r = bench_spmm_power_law(dev, args.steps, args.warmup)
    value = r["bytes"] / (r["ms"] * 1e-3) / 1e9
    kern_ms = r["event_ms_median"]
    achieved = r["bytes"] / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "csr_spmm_power_law_262144x262144_n64",
            "op": "non",
            "rows": 262144,
            "cols": 262144,
            "nnz": r["nnz"],
            "dense_cols": 64,
            "row_length": "power-law exponent 2.1, mean 24, cap 4096",
            "value_dtype": "float32",
            "index_dtype": "int32",
            "seed": 1,
        },
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "gflops": round(r["flops"] / (r["ms"] * 1e-3) / 1e9, 2),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(),
            "kernel": "spmm_bins_kernel (+ spmm_long_reduce_kernel inside the event window)",
            "kernel_ms": round(kern_ms, 5),
        },
        "plan": r["plan"],
        "trans": {
            "gbs": round(r["trans_bytes"] / (r["trans_ms"] * 1e-3) / 1e9, 2),
            "ms_per_step": round(r["trans_ms"], 5),
        },
    }
    # the gather ceiling: the SpMM's B-row reads (nnz x 64 x 4 B) against the
    # same number of random 256-byte row reads from a 64 MiB table
    probe_ms, probe_gbs = bench_gather_probe(dev, 262144, 256, r["nnz"], args.steps, args.warmup)
    spmm_gather_gbs = r["nnz"] * 256 / (kern_ms * 1e-3) / 1e9
    line["roofline"]["gather_probe"] = {
        "what": "random 256-byte row reads from a 64 MiB table, one per nonzero (fsp_gather_rows_probe)",
        "ms": round(probe_ms, 5),
        "gbs": round(probe_gbs, 1),
        "spmm_gather_gbs": round(spmm_gather_gbs, 1),
        "frac": round(spmm_gather_gbs / probe_gbs, 4),
    }
    extras = set(e for e in args.only.split(",") if e) if args.only else {
        "copy", "spmv_large", "spmv_small", "spmm_tuned", "autograd", "bsr", "batch", "vendor"}
    if args.no_extras:
        extras = {"copy"} if args.with_copy else set()
    if "copy" in extras:
        line["hbm_copy_1gib"] = bench_copy(dev, args.steps)
    if "spmv_large" in extras:
        line["spmv_uniform_16m_rows_x16"] = bench_spmv_large(dev, args.steps, args.warmup)
    if "spmv_small" in extras:
        s = bench_spmv_uniform(dev, args.steps, args.warmup)
        line["spmv_uniform_16384x16"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in s.items()}
    if "spmm_tuned" in extras:
        line["spmm_power_law_tuned"] = bench_spmm_tuned(dev, args.steps, args.warmup, r["inputs"])
    if "autograd" in extras:
        a = bench_autograd_bf16_vs_fp32(dev, max(5, args.steps // 5), args.warmup)
        line["bf16_fwd_bwd_speedup_vs_fp32"] = round(a["speedup_bf16_vs_fp32"], 3)
        line["autograd_200k"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in a.items()}
    if "bsr" in extras:
        line["bsr_65536_b40_n128"] = bench_bsr(dev, args.steps, args.warmup)
    if "batch" in extras:
        line["batch_4096"] = bench_batch(dev, args.steps, args.warmup)
    if "slab_schedule" in extras:
        line["spmm_slab_schedule"] = bench_slab_schedule(dev, r["inputs"], args.steps)
    if "trans_forms" in extras:
        line["spmm_trans_forms"] = bench_trans_forms(dev, args.steps, args.warmup, r["inputs"])
    if "vendor" in extras:
        vend = bench_vendor(dev, args.steps, args.warmup, r["inputs"])
        vs = vend.get("spmm_power_law_262144_n64", {}).get("event_ms_median")
        if vs:
            vend["spmm_speedup_vs_rocsparse"] = round(vs / kern_ms, 3)
        sl = line.get("spmv_uniform_16m_rows_x16", {}).get("event_ms_median")
        vv = vend.get("spmv_uniform_16m_rows_x16", {}).get("event_ms_median")
        if sl and vv:
            vend["spmv_speedup_vs_rocsparse"] = round(vv / sl, 3)
        line["rocsparse"] = vend
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cb = cpu_baseline(r["inputs"])
        line["cpu_baseline"] = cb["torch_sparse"]
        line["cpu_baseline"]["gpu_speedup"] = round(value / cb["torch_sparse"]["value"], 2)
        line["cpu_oracle_port"] = cb["oracle"]
        c1 = dict(cb["c1_spmv"])
        small = line.get("spmv_uniform_16384x16")
        if small:
            c1["gpu_graph_us_per_call"] = round(small["graph_ms_per_call"] * 1e3, 3)
            c1["gpu_speedup"] = round(c1["ms"] / small["graph_ms_per_call"], 2)
            small["cpu_torch_sparse"] = c1
        else:
            line["cpu_c1_spmv_16384x16"] = c1
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
