"""Operator runner and FlagGems-style reports for the CSR SpMV / SpMM / SDDMM
path, so results drop into the reference's reporting.

The file formats follow the reference's run_flagsparse_pytest.py (and the
contract its tests/ci/test_runner_result_format.py pins):

* ``<op>/accuracy_result.json``  per-test outcomes (``_pytest_record``);
* ``<op>/performance.csv`` and ``<op>/performance_result.json``
  (``{op: {"result", "test_case", "details": [{"dtype", "result": [{
  "shape_detail", "latency_base", "latency", "speedup", ...}]}]}}``,
  :1385-1430);
* ``<op>/<phase>_detail.json`` (:1035-1078);
* ``summary.json`` -- ``{"timestamp", "env", "result": {op: {"customized",
  "accuracy", "performance", "labels"}}}`` with the strict FlagGems phase
  schemas (:501-570, :892-965, :1852-1866) -- plus ``summary_flat.json`` and
  ``summary.csv`` (:2451-2530).

Each operator's accuracy phase runs its GPU parity tests in a child pytest;
its performance phase runs the benchmark cases (this package's HIP kernels
vs torch.sparse, i.e. rocSPARSE) in a child Python, one operator at a time.
``result.html`` (:2360-2449) is written beside the summaries; the spreadsheet
rendering is not produced.

CLI: ``python -m flagsparse_amd.report [--ops spmv_csr,spmm_csr]
[--phase both|accuracy|performance] [--results-dir DIR] [--quick]
[--list-ops]``.
"""

import argparse
import csv
import datetime as _dt
import importlib
import importlib.metadata as _md
import json
import math
import os
import platform
import re
import shlex
import signal
import statistics
import subprocess
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

PROJECT_ROOT = Path(__file__).resolve().parent.parent

# ------------------------------------------------------------------ formats
ANSI_RE = re.compile(r"\x1B\[[0-?]*[ -/]*[@-~]")
SUMMARY_RE = re.compile(r"(\d+)\s+([A-Za-z_]+)")
PYTEST_CASE_RE = re.compile(r"^(?P<nodeid>\S+::\S+)\s+(?P<status>PASSED|FAILED|SKIPPED|ERROR|XFAIL|XPASS)\b"
                            r"(?P<message>.*)$")
STATUS_TO_FLAGGEMS = {
    "PASS": "Passed", "FAIL": "Failed", "SKIP": "Skipped", "TIMEOUT": "Timeout", "NO_TESTS": "NotFound",
    "CRASH": "Error", "NOT_CONFIGURED": "NotFound",
    "Passed": "Passed", "Failed": "Failed", "Skipped": "Skipped", "Timeout": "Timeout", "NotFound": "NotFound",
    "Error": "Error",
}
PYTEST_STATUS_TO_FLAGGEMS = {"PASSED": "Passed", "FAILED": "Failed", "SKIPPED": "Skipped", "ERROR": "Error",
                             "XFAIL": "Skipped", "XPASS": "Passed"}
# (speedup column, baseline latency column, candidate latency column), in
# the order a row's schema is looked up
PERFORMANCE_SPEEDUP_SCHEMAS = (
    ("speedup", "latency_base", "latency"),
    ("triton_speedup_vs_pytorch", "pytorch_ms", "triton_ms"),
    ("triton_speedup_vs_cusparse", "cusparse_ms", "triton_ms"),
    ("opt_vs_base", "base_ms", "opt_ms"),
    ("base_vs_alg2_speedup", "base_ms", "alg2_ms"),
    ("base_vs_alg1_speedup", "base_ms", "alg1_ms"),
)
_METRIC_COLUMNS = {"speedup", "triton_ms", "pytorch_ms", "cusparse_ms", "base_ms", "alg1_ms", "alg2_ms",
                   "max_abs_err", "max_rel_err"}
_FLAGGEMS_DTYPES = {"float16": "fp16", "torch.float16": "fp16", "half": "fp16", "float32": "fp32",
                    "torch.float32": "fp32", "float": "fp32", "float64": "fp64", "torch.float64": "fp64",
                    "double": "fp64", "bfloat16": "bf16", "torch.bfloat16": "bf16"}
_SUMMARY_CSV_HEADERS = ["operator", "gpu", "phase", "status", "configured", "passed", "failed", "skipped",
                        "errors", "total", "exit_code", "returncode", "duration", "duration_sec", "row_count",
                        "speedup", "data_file", "log_path", "stdout_log_path", "stderr_log_path", "data_path",
                        "reason", "command"]


def _to_float(value):
    if value is None:
        return None
    text = str(value).strip()
    if not text or text.upper() in {"N/A", "NA", "NONE", "NULL"}:
        return None
    if text.endswith("x"):
        text = text[:-1]
    try:
        number = float(text)
    except ValueError:
        return None
    return None if (math.isnan(number) or math.isinf(number)) else number


def _row_dtype(row):
    for key in ("dtype", "value_dtype", "data_dtype", "x_dtype", "B_dtype"):
        if row.get(key):
            return str(row[key])
    return "unknown"


def _row_shape(row, index):
    for key in ("shape", "matrix", "name", "case", "m,n,k", "M,N,K"):
        if row.get(key):
            return str(row[key])
    for keys in (("M", "N", "K"), ("rows", "cols", "nnz")):
        dims = [str(row[k]) for k in keys if row.get(k)]
        if dims:
            return "x".join(dims)
    return f"row_{index}"


def _detail_shape(row, index, seen):
    parts = [_row_shape(row, index)]
    for key in ("index_dtype", "op", "opA", "transpose", "format", "dense_cols", "n_rhs", "mode", "case_id",
                "seed"):
        if row.get(key):
            parts.append(f"{key}={row[key]}")
    name = "|".join(parts)
    if name in seen:
        name = f"{name}|row={index}"
    seen.add(name)
    return name


def _is_metric_column(key):
    k = key.lower()
    return k in _METRIC_COLUMNS or k.endswith("_ms") or "speedup" in k or "latency" in k or "err" in k


def _performance_schema(row):
    for speedup_key, base_key, latency_key in PERFORMANCE_SPEEDUP_SCHEMAS:
        if row.get(speedup_key):
            return speedup_key, base_key, latency_key
    for key in row:
        if "speedup" in key.lower():
            return key, None, None
    return "speedup", None, None


def _benchmark_json_detail(row, index):
    sk, bk, lk = _performance_schema(row)
    base = _to_float(row.get(bk)) if bk else None
    lat = _to_float(row.get(lk)) if lk else None
    sp = _to_float(row.get(sk))
    out = {"shape_detail": _row_shape(row, index), "latency_base": base or 0.0, "latency": lat or 0.0,
           "speedup": sp or 0.0}
    for key, value in row.items():
        if key not in out and value not in (None, ""):
            num = _to_float(value)
            out[key] = num if num is not None else value
    return out


def benchmark_json_from_csv(op, csv_path, *, status="passed", test_case="csv"):
    """performance_result.json content of a benchmark CSV (run_flagsparse_pytest.py:1404-1420)."""
    with Path(csv_path).open("r", encoding="utf-8", newline="") as f:
        rows = list(csv.DictReader(f))
    by_dtype = {}
    for i, row in enumerate(rows):
        by_dtype.setdefault(_row_dtype(row), []).append(_benchmark_json_detail(row, i))
    details = [{"dtype": d, "result": r} for d, r in sorted(by_dtype.items())]
    return {op: {"result": status, "test_case": test_case, "details": details}}


def write_benchmark_json_from_csv(op, csv_path, json_path):
    Path(json_path).write_text(json.dumps(benchmark_json_from_csv(op, csv_path), indent=2, sort_keys=True),
                               encoding="utf-8")


def parse_performance_json(op, path):
    """run_flagsparse_pytest.py:1460-1510."""
    try:
        raw = json.loads(Path(path).read_text(encoding="utf-8"))
    except (OSError, ValueError) as exc:
        return {"status": "Error", "reason": f"Invalid JSON in {path}: {exc}", "data": {}}
    data = raw.get(op, {})
    if not data:
        return {"status": "NotFound", "data": {}}
    result = str(data.get("result", "NotFound"))
    if result.lower() in {"failed", "skipped"}:
        return {"status": result.title(), "reason": data.get("reason", "Unknown"),
                "test_case": data.get("test_case", "Unknown"), "data": {}}
    bench = {}
    for item in data.get("details", []):
        details, speedups = {}, []
        for res in item.get("result", []):
            shape = str(res.get("shape_detail", "Unknown")).replace(" ", "")
            sp = _to_float(res.get("speedup")) or 0.0
            details[shape] = {"base": _to_float(res.get("latency_base")) or 0.0,
                              "gems": _to_float(res.get("latency")) or 0.0, "speedup": sp}
            speedups.append(sp)
        bench[str(item.get("dtype", "unknown"))] = {"result": "OK" if details else "Unknown", "details": details,
                                                    "speedup": sum(speedups) / len(speedups) if speedups else 0}
    return {"status": result.title(), "data": bench, "test_case": data.get("test_case", "Unknown")}


def _flaggems_perf_data(rows):
    grouped, totals, seen = {}, {}, {}
    for i, row in enumerate(rows):
        dtype = _row_dtype(row)
        shape = _detail_shape(row, i, seen.setdefault(dtype, set()))
        sk, bk, lk = _performance_schema(row)
        sp = _to_float(row.get(sk))
        entry = grouped.setdefault(dtype, {"result": "OK", "details": {}, "speedup": 0})
        detail = {"base": _to_float(row.get(bk)) if bk else None, "gems": _to_float(row.get(lk)) if lk else None,
                  "speedup": sp}
        for key in ("status", "matrix_status"):
            if row.get(key):
                detail[key] = row[key]
                if row[key].upper() in {"FAIL", "ERROR"}:
                    entry["result"] = "Failed"
        entry["details"][shape] = detail
        if sp is not None:
            totals.setdefault(dtype, []).append(sp)
    for dtype, values in totals.items():
        grouped[dtype]["speedup"] = statistics.mean(values)
    return grouped


def summarize_performance_csv(path):
    """Phase fields of a benchmark CSV (run_flagsparse_pytest.py:1541-1597)."""
    with Path(path).open("r", encoding="utf-8", newline="") as f:
        rows = list(csv.DictReader(f))
    bench, raw_by = {}, {}
    for i, row in enumerate(rows):
        dtype, shape = _row_dtype(row), _row_shape(row, i)
        metrics, meta = {}, {}
        for key, value in row.items():
            if value in (None, ""):
                continue
            num = _to_float(value)
            (metrics if (num is not None and _is_metric_column(key)) else meta)[key] = (
                num if (num is not None and _is_metric_column(key)) else value)
        bench.setdefault(dtype, {}).setdefault(shape, []).append({"metrics": metrics, "metadata": meta, "raw": row})
        raw_by.setdefault(dtype, {}).setdefault(shape, []).append(row)
    out = {"data_path": str(path), "csv_file": str(path), "row_count": len(rows), "records": rows,
           "records_by_dtype_shape": raw_by, "benchmark": bench, "data": _flaggems_perf_data(rows),
           "test_case": "csv"}
    by_col = {}
    for row in rows:
        for key, value in row.items():
            num = _to_float(value) if "speedup" in key.lower() else None
            if num is not None:
                by_col.setdefault(key, []).append(num)
    if by_col:
        out["speedup_by_column"] = {k: statistics.mean(v) for k, v in by_col.items()}
        for key in ("speedup", "triton_speedup_vs_pytorch", "triton_speedup_vs_cusparse"):
            if key in out["speedup_by_column"]:
                out["speedup"] = out["speedup_by_column"][key]
                break
        else:
            out["speedup"] = out["speedup_by_column"][sorted(by_col)[0]]
    return out


def _parse_test_parameters(nodeid):
    if "[" not in nodeid or not nodeid.endswith("]"):
        return {}
    text = nodeid.rsplit("[", 1)[1][:-1]
    return {f"param_{i}": v for i, v in enumerate(text.split("-"))} if text else {}


def parse_pytest_cases(text):
    """Per-test lines of `pytest -v` output (run_flagsparse_pytest.py:746-773)."""
    cases, seen = [], set()
    for line in ANSI_RE.sub("", text).splitlines():
        m = PYTEST_CASE_RE.match(line.strip())
        if not m:
            continue
        key = (m.group("nodeid"), m.group("status"))
        if key in seen:
            continue
        seen.add(key)
        file_part, _, name = key[0].partition("::")
        cases.append({"nodeid": key[0], "file": file_part, "name": name,
                      "status": PYTEST_STATUS_TO_FLAGGEMS.get(key[1], key[1]), "status_raw": key[1],
                      "parameters": _parse_test_parameters(key[0]),
                      "message": m.group("message").strip() or None})
    return cases


def pytest_counts(text):
    counts = dict.fromkeys(("passed", "failed", "skipped", "errors", "xfailed", "xpassed", "deselected"), 0)
    lines = [ln for ln in ANSI_RE.sub("", text).splitlines() if ln.strip()]
    for line in reversed(lines):
        if " in " in line and ("passed" in line or "failed" in line or "skipped" in line or "error" in line
                               or "no tests ran" in line or "deselected" in line):
            for num, word in SUMMARY_RE.findall(line):
                word = {"error": "errors", "warning": "warnings"}.get(word, word)
                if word in counts:
                    counts[word] = int(num)
            break
    return counts


def status_from_pytest_counts(counts, returncode):
    """run_flagsparse_pytest.py:776-791."""
    has = any(counts[k] for k in ("passed", "failed", "skipped", "errors", "xfailed", "xpassed"))
    if returncode == 5 and not has:
        return "NO_TESTS"
    if returncode not in (0, 5) and not has:
        return "CRASH"
    if counts["failed"] or counts["errors"] or returncode not in (0, 5):
        return "FAIL"
    if counts["passed"] or counts["xfailed"]:
        return "PASS"
    if counts["skipped"]:
        return "SKIP"
    return "NO_TESTS"


def parse_accuracy_json(path):
    """Outcome counts of accuracy_result.json (run_flagsparse_pytest.py:968-1032)."""
    try:
        raw = json.loads(Path(path).read_text(encoding="utf-8"))
    except (OSError, ValueError) as exc:
        return {"total": 0, "skipped": 0, "failed": 0, "passed": 0, "status": "Error",
                "details": {"error": f"Invalid JSON in {path}: {exc}"}, "errors": 1}
    skipped, failed, passed, issue = {}, {}, 0, False
    for case, item in raw.items():
        if not isinstance(item, dict):
            continue
        params = [case[: case.find("[")] if "[" in case else case]
        params += [f"{k}={str(v).replace(' ', '')}" for k, v in (item.get("params") or {}).items()]
        label = ":".join(params)
        result = str(item.get("result") or "").lower()
        reason = str(item.get("reason") or "Unknown")
        if result == "passed":
            passed += 1
        elif result == "skipped":
            issue = issue or "Issue" in reason
            skipped.setdefault(reason, []).append(label)
        else:
            failed.setdefault(reason, []).append(label)
    n_skip = sum(len(v) for v in skipped.values())
    n_fail = sum(len(v) for v in failed.values())
    details = {}
    if failed:
        details["failed"] = failed
    if skipped:
        details["skipped"] = skipped
    if n_fail or issue:
        status = "Failed"
    elif n_skip:
        status = "Skipped"
    elif passed:
        status = "Passed"
    else:
        status = "NotFound"
    return {"total": passed + n_skip + n_fail, "skipped": n_skip, "failed": n_fail, "passed": passed,
            "details": details, "status": status}


# --------------------------------------------------------- phase schemas
def _flaggems_status(status):
    return STATUS_TO_FLAGGEMS.get(str(status or ""), str(status or "Unknown"))


def _duration(r):
    return r.get("duration", r.get("duration_sec"))


def _exit_code(r):
    return r.get("exit_code", r.get("returncode"))


def _accuracy_details(r):
    if isinstance(r.get("details"), dict):
        return r["details"]
    details = {}
    for key, name in (("failures", "failed"), ("tests", "tests"), ("marker", "marker"), ("reason", "reason")):
        if r.get(key):
            details[name] = r[key]
    return details


def _flaggems_accuracy_result(r):
    """The strict accuracy schema (run_flagsparse_pytest.py:892-904)."""
    return {"total": int(r.get("total") or 0), "skipped": int(r.get("skipped") or 0),
            "failed": int(r.get("failed") or 0), "passed": int(r.get("passed") or 0),
            "details": _accuracy_details(r), "status": _flaggems_status(r.get("status") or "UNKNOWN"),
            "exit_code": int(_exit_code(r) or 0), "duration": float(_duration(r) or 0.0),
            "data_file": str(r.get("data_file") or "")}


def _strict_perf_data(value):
    if not isinstance(value, dict):
        return {}
    out = {}
    for dtype, dv in value.items():
        if not isinstance(dv, dict):
            continue
        details = dv.get("details") if isinstance(dv.get("details"), dict) else {}
        out[_FLAGGEMS_DTYPES.get(str(dtype), str(dtype))] = {
            "result": str(dv.get("result") or "Unknown"),
            "details": {str(s): {"base": float((m or {}).get("base") or 0.0),
                                 "gems": float((m or {}).get("gems") or 0.0),
                                 "speedup": float((m or {}).get("speedup") or 0.0)} for s, m in details.items()},
            "speedup": float(dv.get("speedup") or 0.0),
        }
    return out


def _flaggems_performance_result(r):
    """The strict performance schema (run_flagsparse_pytest.py:950-959)."""
    return {"duration": float(_duration(r) or 0.0), "exit_code": int(_exit_code(r) or 0),
            "data_file": str(r.get("data_file") or ""), "data": _strict_perf_data(r.get("data")),
            "status": _flaggems_status(r.get("status") or "UNKNOWN"), "test_case": str(r.get("test_case") or "Unknown")}


def _flaggems_phase_result(r):
    return _flaggems_performance_result(r) if r.get("phase") == "performance" else _flaggems_accuracy_result(r)


def _operator_summary(result):
    entry = {"customized": bool(result.get("customized", True))}
    for phase in ("accuracy", "performance"):
        r = result.get(phase)
        if not isinstance(r, dict):
            r = {"operator": result.get("operator"), "phase": phase, "status": "NOT_CONFIGURED", "duration": 0.0,
                 "exit_code": 0, "data_file": ""}
        entry[phase] = _flaggems_phase_result({**r, "phase": phase})
    labels = result.get("labels")
    entry["labels"] = [str(x) for x in labels] if isinstance(labels, list) else []
    return entry


def _phase_summary(r):
    if r.get("phase") == "performance":
        s = {"row_count": r.get("row_count", 0), "speedup": r.get("speedup", 0)}
        if "speedup_by_column" in r:
            s["speedup_by_column"] = r["speedup_by_column"]
        return s
    return {k: r.get(k, 0) for k in ("passed", "failed", "skipped", "errors", "xfailed", "xpassed", "total")}


def _phase_detail_for_file(r):
    out = {**_flaggems_phase_result(r), "status_raw": str(r.get("status") or "UNKNOWN"), "phase": r.get("phase"),
           "configured": r.get("configured"), "returncode": r.get("returncode"),
           "duration_sec": r.get("duration_sec"), "command": r.get("command", []), "log_path": r.get("log_path"),
           "stdout_log_path": r.get("stdout_log_path"), "stderr_log_path": r.get("stderr_log_path"),
           "data_path": r.get("data_path"), "csv_file": r.get("csv_file"), "reason": r.get("reason"),
           "summary": _phase_summary(r)}
    if r.get("phase") == "accuracy":
        out["tests"] = r.get("tests", [])
    else:
        out["benchmark"] = r.get("benchmark", {})
        out["details"] = {k: r[k] for k in ("row_count", "speedup", "speedup_by_column") if k in r}
    return out


def write_phase_result(op_dir, phase, r):
    """<phase>_detail.json beside <phase>_result.json (run_flagsparse_pytest.py:1062-1078)."""
    op_dir = Path(op_dir)
    path = op_dir / f"{phase}_result.json"
    r["result_path"] = str(path)
    if path.exists():
        r["data_file"] = str(path.relative_to(op_dir.parent)) if op_dir.name == r.get("operator") else path.name
    if phase == "performance" and r.get("csv_file") is None:
        r["csv_file"] = r.get("data_path")
    (op_dir / f"{phase}_detail.json").write_text(json.dumps(_phase_detail_for_file(r), indent=2, sort_keys=True,
                                                            default=str), encoding="utf-8")
    return path


def _phase_rows(results):
    rows = []
    for res in results:
        for phase in ("accuracy", "performance"):
            r = res.get(phase)
            if not isinstance(r, dict):
                continue
            cmd = r.get("command")
            rows.append({
                "operator": res.get("operator"), "gpu": res.get("gpu"), "phase": phase, "status": r.get("status"),
                "configured": r.get("configured"), "passed": r.get("passed", ""), "failed": r.get("failed", ""),
                "skipped": r.get("skipped", ""), "errors": r.get("errors", ""), "total": r.get("total", ""),
                "exit_code": _exit_code(r), "returncode": r.get("returncode", ""), "duration": _duration(r),
                "duration_sec": r.get("duration_sec", ""), "row_count": r.get("row_count", ""),
                "speedup": r.get("speedup", ""), "data_file": r.get("data_file", r.get("data_path")),
                "log_path": r.get("log_path", ""), "stdout_log_path": r.get("stdout_log_path", ""),
                "stderr_log_path": r.get("stderr_log_path", ""), "data_path": r.get("data_path", ""),
                "reason": r.get("reason", ""), "command": shlex.join(cmd) if cmd else "",
            })
    return rows


def _totals(rows):
    by_status, by_phase = {}, {}
    for row in rows:
        s, p = str(row.get("status") or "UNKNOWN"), str(row.get("phase") or "unknown")
        by_status[s] = by_status.get(s, 0) + 1
        by_phase.setdefault(p, {})
        by_phase[p][s] = by_phase[p].get(s, 0) + 1
    return {"by_status": by_status, "by_phase": by_phase}


# ---------------------------------------------------------------- env info
def _module_version(name):
    try:
        return {"version": _md.version(name)}
    except _md.PackageNotFoundError:
        try:
            mod = importlib.import_module(name)
        except Exception:  # noqa: BLE001 -- an absent or broken package is reported as such
            return {"version": None}
        return {"version": getattr(mod, "__version__", None)}


def collect_env_info(project_root=PROJECT_ROOT):
    env = {
        "python": {"version": platform.python_version(), "executable": sys.executable,
                   "implementation": platform.python_implementation()},
        "platform": {"system": platform.system(), "release": platform.release(), "machine": platform.machine(),
                     "platform": platform.platform()},
        "packages": {"torch": _module_version("torch"), "triton": _module_version("triton"),
                     "flagsparse": {"version": _package_version()}},
    }
    try:
        git = subprocess.run(["git", "rev-parse", "HEAD"], cwd=project_root, capture_output=True, text=True,
                             timeout=10, check=False)
        env["git"] = {"commit": git.stdout.strip() or None}
    except (OSError, subprocess.SubprocessError):
        env["git"] = {"commit": None}
    try:
        import torch  # noqa: PLC0415
    except ImportError:
        return env
    gpu = {"available": bool(torch.cuda.is_available()), "version": getattr(torch.version, "hip", None),
           "device_count": 0, "devices": []}
    if gpu["available"]:
        gpu["device_count"] = torch.cuda.device_count()
        for i in range(gpu["device_count"]):
            p = torch.cuda.get_device_properties(i)
            gpu["devices"].append({"index": i, "name": p.name, "total_memory": int(p.total_memory),
                                   "arch": getattr(p, "gcnArchName", None)})
    env["cuda"] = gpu
    return env


def _package_version():
    from . import __version__  # noqa: PLC0415

    return __version__


def _flag_gems_env_info(env):
    """The strict FlagGems env schema (run_flagsparse_pytest.py:501-570)."""
    py, plat = env.get("python") or {}, env.get("platform") or {}
    pk, gpu = env.get("packages") or {}, env.get("cuda") or {}
    try:
        osr = platform.freedesktop_os_release()
    except (AttributeError, OSError):
        osr = {}
    devices = gpu.get("devices") or []
    avail = bool(gpu.get("available"))
    return {
        "architecture": str(plat.get("machine") or ""),
        "os_name": str(osr.get("ID") or plat.get("system") or "").lower(),
        "os_release": str(osr.get("VERSION_ID") or plat.get("release") or ""),
        "python": str(py.get("version") or ""),
        "torch": {"version": str((pk.get("torch") or {}).get("version") or ""), "cuda_available": avail,
                  "device_name": str(devices[0].get("name") or "") if devices else "",
                  "device_count": int(gpu.get("device_count") or 0)},
        "flagtree": None,
        "triton": {"version": str((pk.get("triton") or {}).get("version") or ""), "has_config": False},
        "flag_gems": {"version": str((pk.get("flagsparse") or {}).get("version") or ""),
                      "vendor": "amd" if avail else "cpu", "device": "cuda" if avail else "cpu"},
    }


def write_summary(results, results_dir, env_info):
    """summary.json (FlagGems schema), summary_flat.json and summary.csv
    (run_flagsparse_pytest.py:1852-1866, 2451-2530)."""
    results_dir = Path(results_dir)
    results_dir.mkdir(parents=True, exist_ok=True)
    ordered = sorted(results, key=lambda r: str(r["operator"]))
    rows = _phase_rows(ordered)
    per_op = {str(r["operator"]): _operator_summary(r) for r in ordered}
    stamp = _dt.datetime.now()
    summary = {"timestamp": stamp.strftime("%Y-%m-%d %H:%M:%S"), "env": _flag_gems_env_info(env_info),
               "result": per_op}
    (results_dir / "summary.json").write_text(json.dumps(summary, indent=2) + "\n", encoding="utf-8")
    iso = stamp.isoformat(timespec="seconds")
    flat = {"generated_at": iso, "timestamp": iso, "env": env_info, "totals": _totals(rows), "results": ordered,
            "result": per_op}
    (results_dir / "summary_flat.json").write_text(json.dumps(flat, indent=2, sort_keys=True, default=str),
                                                   encoding="utf-8")
    with (results_dir / "summary.csv").open("w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=_SUMMARY_CSV_HEADERS)
        w.writeheader()
        for row in rows:
            w.writerow({k: row.get(k, "") for k in _SUMMARY_CSV_HEADERS})
    write_result_html(ordered, results_dir, env_info)
    return summary


# ------------------------------------------------------------- result.html
_HTML_DTYPES = ("fp16", "fp32", "fp64", "bf16")
_HTML_STYLE = """<style>
body { font-family: sans-serif; font-size: 13px; }
table { border-collapse: collapse; margin-bottom: 12px; }
th, td { border: 1px solid #bbb; padding: 3px 6px; vertical-align: top; }
th { background: #eee; }
pre { max-height: 320px; overflow: auto; background: #f7f7f7; margin: 4px 0; }
.table-stats { margin: 6px 0; }
</style>"""
# filterTable() hides rows whose accuracy / performance status does not match
# the header selects; sortTable(col, dir) orders rows by a numeric column
# (dir "reset" restores the operator order).
_HTML_SCRIPT = """<script>
function filterTable() {
  var sel = document.querySelectorAll('select.status-filter');
  var rows = document.querySelectorAll('#main-table tbody tr'); var shown = 0;
  rows.forEach(function (r) {
    var ok = true;
    sel.forEach(function (s) {
      var want = s.value; var cell = r.cells[parseInt(s.dataset.col)].querySelector('summary');
      if (want && cell && cell.textContent.trim() !== want) { ok = false; }
    });
    r.style.display = ok ? '' : 'none'; if (ok) { shown += 1; }
  });
  document.getElementById('visible-count').textContent = shown;
}
function sortTable(col, dir) {
  var body = document.querySelector('#main-table tbody');
  var rows = Array.prototype.slice.call(body.rows);
  rows.sort(function (a, b) {
    if (dir === 'reset') { return parseInt(a.cells[0].textContent) - parseInt(b.cells[0].textContent); }
    var x = parseFloat(a.cells[col].textContent), y = parseFloat(b.cells[col].textContent);
    x = isNaN(x) ? -Infinity : x; y = isNaN(y) ? -Infinity : y;
    return dir === 'asc' ? x - y : y - x;
  });
  rows.forEach(function (r) { body.appendChild(r); });
}
</script>"""


def _esc(value):
    import html  # noqa: PLC0415

    return html.escape("" if value is None else str(value))


def _html_status(r):
    return _flaggems_status(r.get("status")) if isinstance(r, dict) else "NotFound"


def _html_counts(r):
    if not isinstance(r, dict) or r.get("phase") == "performance":
        return ""
    return f"{r.get('passed', 0)}/{r.get('failed', 0)}/{r.get('skipped', 0)}"


def _html_speedups(r):
    """(mean speedup, {dtype: mean speedup}) of a performance phase."""
    if not isinstance(r, dict):
        return None, {}
    per = {}
    for row in r.get("records") or []:
        sp = _to_float(row.get("speedup"))
        if sp is None:
            continue
        per.setdefault(_FLAGGEMS_DTYPES.get(str(row.get("dtype")), str(row.get("dtype"))), []).append(sp)
    means = {k: statistics.mean(v) for k, v in per.items()}
    mean = _to_float(r.get("speedup"))
    if mean is None and means:
        mean = statistics.mean(means.values())
    return mean, means


def _html_details(phase, r, results_dir):
    """The phase's data file and log, each in a <pre> block."""
    if not isinstance(r, dict):
        return ""
    parts = []
    for key in ("data_file", "stdout_log_path"):
        rel = r.get(key)
        if not rel:
            continue
        path = Path(rel) if Path(str(rel)).is_absolute() else Path(results_dir) / str(rel)
        try:
            text = path.read_text(encoding="utf-8", errors="replace")[:20000]
        except OSError:
            text = "(not found)"
        parts.append(f"<div>{_esc(Path(str(rel)).name)}</div><pre>{_esc(text)}</pre>")
    parts.append(f"<div>{phase}.json</div><pre>{_esc(json.dumps(_flaggems_phase_result({**r, 'phase': phase}), indent=2, default=str))}</pre>")
    return "".join(parts)


def write_result_html(results, results_dir, env_info):
    """result.html: the environment table and one row per operator with its
    accuracy / performance status, counts, mean speedup and per-dtype
    speedups (the report run_flagsparse_pytest.py:2360-2449 writes)."""
    results_dir = Path(results_dir)
    ordered = sorted(results, key=lambda r: str(r["operator"]))
    fg = _flag_gems_env_info(env_info)
    env_rows = [f"<tr><td>generated</td><td><tt>{_esc(_dt.datetime.now().strftime('%Y-%m-%d %H:%M:%S'))}</tt></td></tr>"]
    for key, value in (("architecture", fg["architecture"]), ("os", f"{fg['os_name']} {fg['os_release']}"),
                       ("python", fg["python"]), ("torch", fg["torch"]["version"]),
                       ("device", fg["torch"]["device_name"]), ("device_count", fg["torch"]["device_count"]),
                       ("flagsparse_amd", fg["flag_gems"]["version"]),
                       ("git", (env_info.get("git") or {}).get("commit"))):
        env_rows.append(f"<tr><td>{_esc(key)}</td><td><tt>{_esc(value)}</tt></td></tr>")
    statuses = ("Passed", "Failed", "Skipped", "Error", "Timeout", "NotFound")

    def status_select(col):
        opts = "".join(f'<option value="{s}">{s}</option>' for s in statuses)
        return (f'<select class="status-filter" data-col="{col}" onchange="filterTable()">'
                f'<option value="">All</option>{opts}</select>')

    rows = []
    for i, res in enumerate(ordered, start=1):
        op = str(res["operator"])
        acc, perf = res.get("accuracy"), res.get("performance")
        mean, by_dtype = _html_speedups(perf)
        dtype_cells = "".join(
            f"<td>{'' if by_dtype.get(d) is None else f'{by_dtype[d]:.3f}'}</td>" for d in _HTML_DTYPES)
        note = next((str(r.get("reason")) for r in (acc, perf) if isinstance(r, dict) and r.get("reason")), "")
        rows.append(
            f"<tr><td>{i}</td><td>{_esc(op)}</td>"
            f"<td><details><summary>{_esc(_html_status(acc))}</summary>{_html_details('accuracy', acc, results_dir)}"
            f"</details></td><td>{_esc(_html_counts(acc))}</td>"
            f"<td><details><summary>{_esc(_html_status(perf))}</summary>"
            f"{_html_details('performance', perf, results_dir)}</details></td>"
            f"<td>{'' if mean is None else f'{mean:.3f}'}</td>{dtype_cells}<td>{_esc(note)}</td></tr>")
    sort_buttons = "".join(f"<button onclick=\"sortTable(5, '{d}')\">{label}</button>"
                           for d, label in (("asc", "&#9650;"), ("desc", "&#9660;"), ("reset", "&#8634;")))
    dtype_heads = "".join(f"<th>{d}</th>" for d in _HTML_DTYPES)
    text = f"""<!DOCTYPE html>
<html>
<head>
<meta charset="utf-8">
<title>FlagSparse Test Report</title>
{_HTML_STYLE}
</head>
<body>
<h3>Test Environment</h3>
<table><thead><tr><th>Env</th><th>Setting</th></tr></thead>
<tbody>{"".join(env_rows)}</tbody></table>
<h3>Test Result</h3>
<div class="table-stats">Total: <span id="total-count">{len(ordered)}</span> | Showing: <span id="visible-count">{len(ordered)}</span></div>
<table id="main-table">
<thead><tr><th>No.</th><th>ID</th><th>AccRes<br>{status_select(2)}</th><th>AccStat(pass/fail/skip)</th>
<th>PerfRes<br>{status_select(4)}</th><th>OPAverageSpeedUp<br>{sort_buttons}</th>{dtype_heads}<th>Note</th></tr></thead>
<tbody>{"".join(rows)}</tbody></table>
{_HTML_SCRIPT}
</body>
</html>
"""
    (results_dir / "result.html").write_text(text, encoding="utf-8")


def _should_fail(results, strict=False):
    failing = {"FAIL", "CRASH", "TIMEOUT", "NO_TESTS", "Failed", "Error", "Timeout", "NotFound"}
    if strict:
        failing |= {"NOT_CONFIGURED", "NotConfigured"}
    return any(row.get("status") in failing for row in _phase_rows(results))


# ------------------------------------------------------ operator catalog
@dataclass(frozen=True)
class OperatorSpec:
    tests: tuple  # pytest paths relative to the project root
    keyword: str = ""  # pytest -k expression
    perf: str = ""  # performance case family ("" = none)
    labels: tuple = field(default=("flagsparse", "sparse", "hip", "public-api"))


OPERATORS = {
    "spmv_csr": OperatorSpec(("tests/test_spmv_gpu.py",), perf="spmv_csr"),
    "spmv_coo_tocsr": OperatorSpec(("tests/test_coo_gpu.py",)),
    "spmv_bsr": OperatorSpec(("tests/test_spmv_bsr_gpu.py",)),
    "spmm_csr": OperatorSpec(("tests/test_spmm_gpu.py", "tests/test_autograd_gpu.py"), keyword="not opt",
                             perf="spmm_csr"),
    "spmm_csr_opt": OperatorSpec(("tests/test_spmm_gpu.py",), keyword="opt and not alg2", perf="spmm_csr_opt"),
    "spmm_csr_opt_alg1": OperatorSpec(("tests/test_spmm_gpu.py",), keyword="opt and not alg2",
                                      perf="spmm_csr_opt"),
    "spmm_csr_opt_alg2": OperatorSpec(("tests/test_spmm_gpu.py",), keyword="alg2", perf="spmm_csr_opt_alg2"),
    "sddmm_csr": OperatorSpec(("tests/test_sddmm_gpu.py",), perf="sddmm_csr"),
    "spmm_bsr": OperatorSpec(("tests/test_bsr_gpu.py",), labels=("flagsparse", "sparse", "hip", "bsr", "mfma")),
    "csr_batch": OperatorSpec(("tests/test_batch_gpu.py",), labels=("flagsparse", "sparse", "hip", "batch")),
}

# (n_rows, n_cols, nnz[, dense width or K]) of the performance sweeps
_SPMV_SHAPES = [(4096, 4096, 65536), (16384, 16384, 262144), (65536, 65536, 1048576), (262144, 262144, 4194304)]
_SPMM_SHAPES = [(4096, 4096, 65536, 32), (16384, 16384, 262144, 64), (65536, 65536, 1048576, 64)]
_SDDMM_SHAPES = [(1024, 1024, 16384, 64), (16384, 16384, 262144, 64), (65536, 65536, 1048576, 128)]


def _perf_rows(op, quick, warmup, iters):
    """Benchmark rows (one dict per case) of a performance family."""
    import torch  # noqa: PLC0415

    from . import benchmarks  # noqa: PLC0415

    dtypes = [torch.float32] if quick else [torch.float32, torch.float64]
    take = (lambda s: s[:2]) if quick else (lambda s: s)
    rows = []
    name = lambda d: str(d).replace("torch.", "")  # noqa: E731
    if op == "spmv_csr":
        for dt in dtypes:
            for m, n, nnz in take(_SPMV_SHAPES):
                for opn in ("non", "trans"):
                    r = benchmarks.benchmark_spmv_case(m, n, nnz, dt, torch.int32, warmup, iters, op=opn)
                    rows.append(_row(name(dt), f"{m}x{n}x{nnz}", opn, r, r["verification"]["triton_max_relative_error"]))
    elif op in ("spmm_csr", "spmm_csr_opt", "spmm_csr_opt_alg2"):
        for dt in dtypes:
            for m, n, nnz, k in take(_SPMM_SHAPES):
                r = benchmarks.benchmark_spmm_case(m, n, nnz, k, dt, torch.int32, warmup, iters)
                row = _row(name(dt), f"{m}x{n}x{nnz}", "non", r, r["verification"]["triton_max_relative_error"])
                row["dense_cols"] = k
                if op != "spmm_csr":
                    row.update(_opt_timing(op, m, n, nnz, k, dt, warmup, iters))
                rows.append(row)
    elif op == "sddmm_csr":
        for dt in dtypes:
            for m, n, nnz, k in take(_SDDMM_SHAPES):
                r = benchmarks.benchmark_sddmm_case(m, n, nnz, k, dt, warmup, iters)
                row = _row(name(dt), f"{m}x{n}x{nnz}", "non", r, r["verification"]["triton_max_relative_error"],
                           baseline="torch gather formulation")
                row["max_abs_err"] = r["verification"]["triton_max_abs_error"]
                row["k_dim"] = k
                rows.append(row)
    return rows


# max|cand - golden| / max|golden| bounds a row must meet (float64 golden)
_REL_TOL = {"float32": 1e-5, "float64": 1e-12, "bfloat16": 1e-2, "float16": 5e-3}


def _row(dtype, shape, opn, r, rel, baseline="rocSPARSE (torch.sparse)"):
    """One CSV row; status PASS when the kernel meets the relative-error bound
    against the float64 golden (elementwise allclose between two rounded
    results is reported, not judged)."""
    perf = r["performance"]
    v = r["verification"]
    row = {"dtype": dtype, "shape": shape, "op": opn, "triton_ms": perf["triton_ms"], "pytorch_ms": perf["pytorch_ms"],
           "triton_speedup_vs_pytorch": perf["triton_speedup_vs_pytorch"], "baseline": baseline}
    if rel is not None:
        row["max_rel_err"] = rel
    row["allclose"] = v.get("triton_match_reference", v.get("triton_match_pytorch"))
    row["status"] = "PASS" if rel is not None and rel <= _REL_TOL.get(dtype, 1e-5) else "FAIL"
    return row


def _opt_timing(op, m, n, nnz, k, dt, warmup, iters):
    """base_ms (flagsparse_spmm_csr) beside the opt / opt-alg2 entry point."""
    import torch  # noqa: PLC0415

    from . import benchmarks, spmm_csr  # noqa: PLC0415

    dev = torch.device("cuda")
    data, idx, ptr = benchmarks._build_random_csr(m, n, nnz, dt, torch.int32, dev)
    B = benchmarks._build_random_dense((n, k), dt, dev)
    _, base_ms = benchmarks._benchmark_cuda_op(lambda: spmm_csr.flagsparse_spmm_csr(data, idx, ptr, B, (m, n)),
                                               warmup, iters)
    if op == "spmm_csr_opt":
        prep = spmm_csr.prepare_spmm_csr_opt(data, idx, ptr, (m, n))
        _, ms = benchmarks._benchmark_cuda_op(lambda: spmm_csr.flagsparse_spmm_csr_opt(B=B, prepared=prep), warmup,
                                              iters)
        return {"base_ms": base_ms, "opt_ms": ms, "opt_vs_base": base_ms / ms if ms > 0 else None}
    prep = spmm_csr.prepare_spmm_csr_opt_alg2(data, idx, ptr, (m, n))
    spmm_csr.flagsparse_spmm_csr_opt_alg2(B=B, prepared=prep)  # tunes once for the shape class
    _, ms = benchmarks._benchmark_cuda_op(lambda: spmm_csr.flagsparse_spmm_csr_opt_alg2(B=B, prepared=prep), warmup,
                                          iters)
    return {"base_ms": base_ms, "alg2_ms": ms, "base_vs_alg2_speedup": base_ms / ms if ms > 0 else None}


def run_perf_child(op, csv_path, quick, warmup, iters):
    """Body of the performance child process: the CSV of one operator."""
    import torch  # noqa: PLC0415

    if not torch.cuda.is_available():
        print("CUDA is not available; performance cases skipped", flush=True)
        return 0
    rows = _perf_rows(op, quick, warmup, iters)
    if not rows:
        return 0
    keys = []
    for row in rows:
        keys += [k for k in row if k not in keys]
    Path(csv_path).parent.mkdir(parents=True, exist_ok=True)
    with Path(csv_path).open("w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=keys)
        w.writeheader()
        for row in rows:
            w.writerow(row)
    print(f"{op}: {len(rows)} performance rows -> {csv_path}", flush=True)
    return 0


# ------------------------------------------------------------------ runner
def _child_env():
    env = os.environ.copy()
    env["PYTHONUNBUFFERED"] = "1"
    env["PYTHONPATH"] = os.pathsep.join([str(PROJECT_ROOT)] + ([env["PYTHONPATH"]] if env.get("PYTHONPATH") else []))
    return env


def _run(cmd, timeout):
    """(returncode, stdout, stderr, seconds, timed_out); a timed-out child's
    process group (started here) is terminated."""
    t0 = time.perf_counter()
    proc = subprocess.Popen(cmd, cwd=PROJECT_ROOT, env=_child_env(), stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True, start_new_session=True)
    try:
        out, err = proc.communicate(timeout=timeout)
        return proc.returncode, out, err, time.perf_counter() - t0, False
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGTERM)
        try:
            out, err = proc.communicate(timeout=10)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            out, err = proc.communicate()
        return -100, out or "", err or "", time.perf_counter() - t0, True


def _phase_base(op, phase, cmd, rc, out, err, dur, op_dir):
    (op_dir / f"{phase}_stdout.log").write_text(out, encoding="utf-8")
    (op_dir / f"{phase}_stderr.log").write_text(err, encoding="utf-8")
    rel = lambda p: str(p.relative_to(op_dir.parent))  # noqa: E731
    return {"operator": op, "phase": phase, "configured": True, "returncode": rc, "exit_code": rc,
            "duration_sec": dur, "duration": dur, "command": cmd,
            "log_path": rel(op_dir / f"{phase}_stdout.log"), "stdout_log_path": rel(op_dir / f"{phase}_stdout.log"),
            "stderr_log_path": rel(op_dir / f"{phase}_stderr.log")}


def run_accuracy(op, spec, op_dir, timeout=1800, extra=()):
    result_path = op_dir / "accuracy_result.json"
    if result_path.exists():
        result_path.unlink()
    cmd = [sys.executable, "-m", "pytest", *spec.tests, "-m", "gpu", "-v", "-p", "flagsparse_amd._pytest_record",
           "--record", "json", "--output", str(result_path)]
    if spec.keyword:
        cmd += ["-k", spec.keyword]
    cmd += list(extra)
    rc, out, err, dur, timed_out = _run(cmd, timeout)
    r = _phase_base(op, "accuracy", cmd, rc, out, err, dur, op_dir)
    counts = pytest_counts(out)
    r.update({k: counts[k] for k in ("passed", "failed", "skipped", "errors", "xfailed", "xpassed")})
    r["total"] = sum(counts[k] for k in ("passed", "failed", "skipped", "errors", "xfailed", "xpassed"))
    r["tests"] = parse_pytest_cases(out)
    r["failures"] = [ln.strip() for ln in out.splitlines() if ln.startswith("FAILED ") or ln.startswith("ERROR ")]
    r["status"] = "TIMEOUT" if timed_out else status_from_pytest_counts(counts, rc)
    if result_path.exists():
        parsed = parse_accuracy_json(result_path)
        r["details"] = parsed["details"]
    write_phase_result(op_dir, "accuracy", r)
    return r


def run_performance(op, spec, op_dir, quick=False, warmup=5, iters=20, timeout=1800):
    if not spec.perf:
        return {"operator": op, "phase": "performance", "configured": False, "status": "NOT_CONFIGURED",
                "reason": "no performance cases for this operator", "duration": 0.0, "exit_code": 0}
    csv_path = op_dir / "performance.csv"
    json_path = op_dir / "performance_result.json"
    for p in (csv_path, json_path):
        if p.exists():
            p.unlink()
    cmd = [sys.executable, "-m", "flagsparse_amd.report", "--run-perf", spec.perf, "--csv", str(csv_path),
           "--warmup", str(warmup), "--iters", str(iters)] + (["--quick"] if quick else [])
    rc, out, err, dur, timed_out = _run(cmd, timeout)
    r = _phase_base(op, "performance", cmd, rc, out, err, dur, op_dir)
    if timed_out:
        r["status"] = "TIMEOUT"
    elif rc != 0:
        r["status"] = "FAIL"
    elif "CUDA is not available" in out + err:
        r["status"] = "SKIP"
    elif not csv_path.exists():
        r["status"] = "NO_TESTS"
    else:
        r["status"] = "PASS"
    if csv_path.exists():
        r["data_path"] = str(csv_path.relative_to(op_dir.parent))
        r.update({k: v for k, v in summarize_performance_csv(csv_path).items()
                  if k not in ("records", "data_path", "csv_file")})
        write_benchmark_json_from_csv(op, csv_path, json_path)
    write_phase_result(op_dir, "performance", r)
    return r


def run(ops, results_dir, phase="both", quick=False, warmup=5, iters=20, timeout=1800, gpu=None):
    results_dir = Path(results_dir)
    results = []
    for op in ops:
        spec = OPERATORS[op]
        op_dir = results_dir / op
        op_dir.mkdir(parents=True, exist_ok=True)
        res = {"operator": op, "gpu": gpu if gpu is not None else 0, "customized": True, "labels": list(spec.labels)}
        if phase in ("both", "accuracy"):
            res["accuracy"] = run_accuracy(op, spec, op_dir, timeout)
            print(f"[{op}] accuracy {res['accuracy']['status']} ({res['accuracy'].get('passed', 0)} passed)",
                  flush=True)
        if phase in ("both", "performance"):
            res["performance"] = run_performance(op, spec, op_dir, quick, warmup, iters, timeout)
            print(f"[{op}] performance {res['performance']['status']}", flush=True)
        results.append(res)
    write_summary(results, results_dir, collect_env_info())
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--ops", default="", help="comma-separated operator ids (default: all)")
    ap.add_argument("--phase", default="both", choices=["both", "accuracy", "performance"])
    ap.add_argument("--results-dir", default="results")
    ap.add_argument("--quick", action="store_true", help="fewer shapes and dtypes")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--timeout", type=int, default=1800, help="seconds per phase child")
    ap.add_argument("--gpu", type=int, default=None, help="HIP device index for the children")
    ap.add_argument("--strict", action="store_true", help="exit 1 also for unconfigured phases")
    ap.add_argument("--list-ops", action="store_true")
    ap.add_argument("--run-perf", default="", help=argparse.SUPPRESS)
    ap.add_argument("--csv", default="", help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.run_perf:
        return run_perf_child(args.run_perf, args.csv, args.quick, args.warmup, args.iters)
    if args.list_ops:
        for op, spec in OPERATORS.items():
            print(f"{op}\taccuracy={','.join(spec.tests)}{' -k ' + spec.keyword if spec.keyword else ''}"
                  f"\tperformance={spec.perf or '-'}")
        return 0
    ops = [o for o in args.ops.split(",") if o] if args.ops else list(OPERATORS)
    unknown = [o for o in ops if o not in OPERATORS]
    if unknown:
        ap.error(f"unknown operator(s): {', '.join(unknown)}; known: {', '.join(OPERATORS)}")
    if args.gpu is not None:
        os.environ["HIP_VISIBLE_DEVICES"] = str(args.gpu)
    results = run(ops, args.results_dir, args.phase, args.quick, args.warmup, args.iters, args.timeout, args.gpu)
    return 1 if _should_fail(results, args.strict) else 0


if __name__ == "__main__":
    sys.exit(main())
